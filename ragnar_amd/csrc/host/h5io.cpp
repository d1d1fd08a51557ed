// HDF5 input/output of the `ragnar` module: H5read1DArray_* / H5write1DArray_*
// (io/h5.cpp:16-118) and the Tristan-v2 particle reader TristanV2_{1,2,3}D
// (plugins/tristan-v2.cpp:51-258), on the HDF5 C API.
//
// Reads stream straight into device buffers: each dataset is read as
// hyperslab chunks into two pinned staging buffers that alternate, and each
// chunk's host->device copy is queued on a copy stream while the next chunk is
// read from the file.
#include <pybind11/pybind11.h>

#include "containers.hpp"

namespace py = pybind11;
using namespace pybind11::literals;

#if defined(RAGNAR_USE_HDF5)

#include <hdf5.h>

#include <algorithm>
#include <array>
#include <cstring>
#include <stdexcept>
#include <string>
#include <sys/stat.h>

namespace rgnr {
namespace {

struct H5Handle {
  hid_t id = -1;
  herr_t (*close)(hid_t) = nullptr;
  H5Handle(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  H5Handle(const H5Handle&) = delete;
  H5Handle& operator=(const H5Handle&) = delete;
  ~H5Handle() {
    if (id >= 0 && close) close(id);
  }
};

void quiet_hdf5() {
  static bool done = false;
  if (!done) {
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    done = true;
  }
}

template <class T>
hid_t native_type();
template <>
hid_t native_type<int>() { return H5T_NATIVE_INT; }
template <>
hid_t native_type<float>() { return H5T_NATIVE_FLOAT; }
template <>
hid_t native_type<double>() { return H5T_NATIVE_DOUBLE; }

hid_t open_file_ro(const std::string& fname) {
  quiet_hdf5();
  const hid_t f = H5Fopen(fname.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) throw std::runtime_error("Unable to open file " + fname);
  return f;
}

// Extent of a 1-D dataset ("Dataset is not 1D" otherwise).
hsize_t dataset_extent(hid_t dset) {
  H5Handle space(H5Dget_space(dset), H5Sclose);
  if (H5Sget_simple_extent_ndims(space.id) != 1) throw std::runtime_error("Dataset is not 1D");
  hsize_t dims[1] = {0};
  H5Sget_simple_extent_dims(space.id, dims, nullptr);
  return dims[0];
}

// Pinned double buffer + copy stream for chunked file -> device transfers.
class Stager {
 public:
  static constexpr size_t kChunkBytes = size_t(16) << 20;
  Stager() {
    for (auto& b : buf_) check(rgn_host_malloc(&b, kChunkBytes));
    for (auto& e : ev_) check(rgn_event_create(&e));
    check(rgn_stream_create(&stream_));
  }
  ~Stager() {
    rgn_sync(stream_);
    for (auto& e : ev_) rgn_event_destroy(e);
    for (auto& b : buf_) rgn_host_free(b);
    rgn_stream_destroy(stream_);
  }
  // Reads `count` elements of `dset` (hyperslab start/stride) into d_dst.
  template <class T>
  void read(hid_t dset, hsize_t start, hsize_t count, hsize_t stride, T* d_dst) {
    const hsize_t per_chunk = kChunkBytes / sizeof(T);
    H5Handle fspace(H5Dget_space(dset), H5Sclose);
    for (hsize_t done = 0, k = 0; done < count; done += per_chunk, ++k) {
      const hsize_t n = std::min(per_chunk, count - done);
      const int slot = static_cast<int>(k & 1);
      if (k >= 2) check(rgn_event_sync(ev_[slot]));  // its previous copy has finished
      const hsize_t s0[1] = {start + done * stride}, st[1] = {stride}, cnt[1] = {n};
      if (H5Sselect_hyperslab(fspace.id, H5S_SELECT_SET, s0, st, cnt, nullptr) < 0) {
        throw std::runtime_error("hyperslab selection out of range");
      }
      H5Handle mspace(H5Screate_simple(1, cnt, nullptr), H5Sclose);
      if (H5Dread(dset, native_type<T>(), mspace.id, fspace.id, H5P_DEFAULT, buf_[slot]) < 0) {
        throw std::runtime_error("HDF5 read failed");
      }
      check(rgn_memcpy_h2d_async(d_dst + done, buf_[slot], n * sizeof(T), stream_));
      check(rgn_event_record(ev_[slot], stream_));
    }
    check(rgn_sync(stream_));
  }

 private:
  std::array<void*, 2> buf_{};
  std::array<void*, 2> ev_{};
  rgn_stream_t stream_ = nullptr;
};

// io/h5.cpp:16-48
template <class T>
Array1D<T> Read1DArray(const std::string& filename, const std::string& dsetname, std::size_t size,
                       std::size_t stride) {
  H5Handle file(open_file_ro(filename), H5Fclose);
  H5Handle dset(H5Dopen2(file.id, dsetname.c_str(), H5P_DEFAULT), H5Dclose);
  if (dset.id < 0) throw std::runtime_error("Unable to open dataset " + dsetname);
  const hsize_t n0 = dataset_extent(dset.id);
  if (stride == 0) {
    throw std::runtime_error("Stride must be greater than 0");
  } else if (size == 0) {
    size = n0;
  } else if (n0 / stride > size) {
    throw std::runtime_error("Number of read quantity exceeds allocated space");
  }
  if (verbose()) py::print("Reading", dsetname, "from", filename, "...", "end"_a = "", "flush"_a = true);
  Array1D<T> out = Array1D<T>::device(size);
  if (size && (size - 1) * stride >= n0) {
    if (verbose()) py::print("Error reading", dsetname, "from", filename, ":", "selection out of range");
    throw std::runtime_error("hyperslab selection out of range");
  }
  if (size) {
    Stager st;
    st.read<T>(dset.id, 0, size, stride, out.dptr());
  }
  out.mark_device_written();
  if (verbose()) py::print(": OK", "flush"_a = true);
  return out;
}

// io/h5.cpp:50-68 (HighFive ReadWrite | Create: open, or create when missing)
template <class T>
void Write1DArray(const std::string& filename, const std::string& dsetname, const Array1D<T>& array) {
  quiet_hdf5();
  struct stat sb;
  const bool exists = ::stat(filename.c_str(), &sb) == 0;
  const hid_t fid = exists ? H5Fopen(filename.c_str(), H5F_ACC_RDWR, H5P_DEFAULT)
                           : H5Fcreate(filename.c_str(), H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT);
  if (fid < 0) throw std::runtime_error("Unable to open or create file " + filename);
  H5Handle file(fid, H5Fclose);
  if (verbose()) py::print("Writing", dsetname, "to", filename, "...", "end"_a = "", "flush"_a = true);
  const auto& h = array.host();
  const hsize_t dims[1] = {h.size()};
  H5Handle space(H5Screate_simple(1, dims, nullptr), H5Sclose);
  H5Handle dset(H5Dcreate2(file.id, dsetname.c_str(), native_type<T>(), space.id, H5P_DEFAULT,
                           H5P_DEFAULT, H5P_DEFAULT),
                H5Dclose);
  if (dset.id < 0 || H5Dwrite(dset.id, native_type<T>(), H5S_ALL, H5S_ALL, H5P_DEFAULT, h.data()) < 0) {
    if (verbose()) py::print("Error writing", dsetname, "to", filename);
    throw std::runtime_error("Unable to write dataset " + dsetname);
  }
  if (verbose()) py::print(": OK", "flush"_a = true);
}

// plugins/tristan-v2.hpp:16-43
template <dim_t D>
class TristanV2 {
  std::string m_path;
  std::size_t m_step = 0;
  bool is_path_set{false};
  bool is_step_set{false};

 public:
  std::string label() const { return "Tristan V2"; }
  void setPath(const std::string& path) {
    m_path = path;
    is_path_set = true;
  }
  void setStep(std::size_t st) {
    m_step = st;
    is_step_set = true;
  }
  std::string getPath() const {
    if (!is_path_set) throw std::runtime_error("Path not set");
    return m_path;
  }
  std::size_t getStep() const {
    if (!is_step_set) throw std::runtime_error("Step not set");
    return m_step;
  }

  // tristan-v2.cpp:51-74: hyperslab checks; size 0 means dims / stride.
  static std::size_t read_quantity(Stager& st, hid_t file, const std::string& dsetname,
                                   std::size_t start, std::size_t size, std::size_t stride,
                                   real_t* d_dst, std::size_t capacity) {
    H5Handle dset(H5Dopen2(file, dsetname.c_str(), H5P_DEFAULT), H5Dclose);
    if (dset.id < 0) throw std::runtime_error("Unable to open dataset " + dsetname);
    const hsize_t n0 = dataset_extent(dset.id);
    if (start >= n0) throw std::runtime_error("Start index out of bounds");
    if (start + size > n0) throw std::runtime_error("Size exceeds dataset dimensions");
    if (size == 0) size = n0 / stride;
    if (size > capacity) throw std::runtime_error("Number of particles mismatch");
    if (size && start + (size - 1) * stride >= n0) {
      throw std::runtime_error("hyperslab selection out of range");
    }
    st.read<real_t>(dset.id, start, size, stride, d_dst);
    return size;
  }

  // tristan-v2.cpp:95-188
  Particles<D> readParticles(const std::string& label, unsigned short sp, std::size_t start,
                             std::size_t size, std::size_t stride, bool ignore_coordinates) const {
    if (stride == 0) {
      throw std::runtime_error("Stride must be greater than 0");
    } else if (stride != 1 and size != 0) {
      throw std::runtime_error("Size must be determined automatically (0) when stride != 1");
    }
    Particles<D> prtls{label};
    const auto step = std::to_string(getStep());
    const auto sp_str = std::to_string(sp);
    const std::string fname = getPath() + "/output/prtl/prtl.tot." +
                              std::string(step.length() < 5 ? 5 - step.length() : 0, '0') + step;
    if (verbose()) py::print("Reading particles #", sp, "from", fname, "...", "flush"_a = true);
    H5Handle file(open_file_ro(fname), H5Fclose);
    const std::array<std::string, 3> comps_coord{"x", "y", "z"};
    const std::array<std::string, 3> comps_vel{"u", "v", "w"};
    prtls.setIgnoreCoords(ignore_coordinates);
    std::size_t ntotal;
    {
      H5Handle dx(H5Dopen2(file.id, ("x_" + sp_str).c_str(), H5P_DEFAULT), H5Dclose);
      if (dx.id < 0) throw std::runtime_error("Unable to open dataset x_" + sp_str);
      ntotal = dataset_extent(dx.id);
    }
    if (start + size >= ntotal) throw std::runtime_error("start + size >= total number of particles");
    const std::size_t nparticles = (size == 0) ? (ntotal / stride) : size;
    prtls.allocate(nparticles);
    if (verbose()) {
      py::print(" found", ToHumanReadable(static_cast<double>(ntotal), false), "particles, reading",
                ToHumanReadable(static_cast<double>(nparticles), false), "starting from", start,
                "flush"_a = true);
    }
    auto report_ok = [&](std::size_t nread, const std::string& quantity) {
      if (nread != nparticles) throw std::runtime_error("Number of particles mismatch");
      if (verbose()) py::print(" ", quantity, ": OK", "flush"_a = true);
    };
    Stager st;
    if (not ignore_coordinates) {
      for (dim_t d = 0; d < D; ++d) {
        auto& dst = prtls.X[d];
        report_ok(read_quantity(st, file.id, comps_coord[d] + "_" + sp_str, start, size, stride,
                                dst.dptr(), nparticles),
                  comps_coord[d]);
        dst.mark_device_written();
      }
    }
    for (int d = 0; d < 3; ++d) {
      report_ok(read_quantity(st, file.id, comps_vel[d] + "_" + sp_str, start, size, stride,
                              prtls.U[d].dptr(), nparticles),
                comps_vel[d]);
      report_ok(read_quantity(st, file.id, "e" + comps_coord[d] + "_" + sp_str, start, size,
                              stride, prtls.E[d].dptr(), nparticles),
                "e" + comps_coord[d]);
      report_ok(read_quantity(st, file.id, "b" + comps_coord[d] + "_" + sp_str, start, size,
                              stride, prtls.B[d].dptr(), nparticles),
                "b" + comps_coord[d]);
      prtls.U[d].mark_device_written();
      prtls.E[d].mark_device_written();
      prtls.B[d].mark_device_written();
    }
    prtls.setNactive(nparticles);
    return prtls;
  }
};

template <class T>
void define_h5(py::module& m, const char* suffix) {
  m.def((std::string("H5read1DArray_") + suffix).c_str(), &Read1DArray<T>, "filename"_a,
        "dsetname"_a, "size"_a = 0, "stride"_a = 1, "Read a 1D array from an HDF5 file");
  m.def((std::string("H5write1DArray_") + suffix).c_str(), &Write1DArray<T>, "filename"_a,
        "dsetname"_a, "array"_a, "Write a 1D array to an HDF5 file");
}

template <dim_t D>
void define_tristan(py::module& m) {
  const std::string name = "TristanV2_" + std::to_string(D) + "D";
  py::class_<TristanV2<D>>(m, name.c_str())
    .def(py::init<>())
    .def("label", &TristanV2<D>::label)
    .def("setPath", &TristanV2<D>::setPath)
    .def("setStep", &TristanV2<D>::setStep)
    .def("getPath", &TristanV2<D>::getPath)
    .def("getStep", &TristanV2<D>::getStep)
    .def("readParticles", &TristanV2<D>::readParticles, "label"_a, "sp"_a, "start"_a = 0,
         "size"_a = 0, "stride"_a = 1, "ignore_coordinates"_a = false,
         "Read particles of species sp from output/prtl/prtl.tot.<step>")
    .doc() = "Plugin for Tristan v2 particle output";
}

}  // namespace

void pyDefineH5(py::module& m) {
  // pyinterface.cpp:68-82; the reference names use typeid(T).name(): i, f, d
  define_h5<int>(m, "i");
  define_h5<float>(m, "f");
  define_h5<double>(m, "d");
  define_tristan<1>(m);
  define_tristan<2>(m);
  define_tristan<3>(m);
  m.attr("HAS_HDF5") = true;
}

}  // namespace rgnr

#else

namespace rgnr {
void pyDefineH5(py::module& m) { m.attr("HAS_HDF5") = false; }
}  // namespace rgnr

#endif
