// The `ragnar` Python module: the reference's pybind11 API (pyinterface.cpp:26-91)
// over the gfx950 kernels of libragnar_hip.so.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <stdexcept>
#include <string>

#include "containers.hpp"

namespace py = pybind11;
using namespace pybind11::literals;
using namespace rgnr;

namespace rgnr {
void pyDefineH5(py::module& m);  // h5io.cpp (present when built with HDF5)
}

namespace {

template <class T>
void define_array(py::module& m, const char* name) {
  // array.cpp:64-110; the reference names these Array1D_<typeid(T).name()>
  py::class_<Array1D<T>>(m, name)
    .def(py::init<>())
    .def(py::init<const py::array_t<T, py::array::forcecast>&>(), "arr"_a)
    .def("head", &Array1D<T>::head, "n"_a = 10, "start"_a = 0,
         "Print n elements starting from index start")
    .def("__repr__", &Array1D<T>::repr)
    .def("as_array", &Array1D<T>::as_array, "Copy the array into a numpy array")
    .def("extent", &Array1D<T>::extent, "d"_a = 0, "Extent along dimension d")
    .doc() = "A one-dimensional array in device memory";
}

template <bool LG>
void define_tabulated(py::module& m) {
  // tabulation.cpp:119-171
  const std::string name = std::string("TabulatedFunction") + (LG ? "_log" : "");
  py::class_<TabulatedFunction<LG>>(m, name.c_str())
    .def(py::init<const Array1D<real_t>&, const Array1D<real_t>&, real_t>(), "x"_a, "y"_a,
         "yfill"_a = 0.0)
    .def(py::init([](const py::array_t<real_t, py::array::forcecast>& x,
                     const py::array_t<real_t, py::array::forcecast>& y, real_t yfill) {
           // tabulation.cpp:43-61: n = x.size(), then x.at(i) and y.at(i) for i < n,
           // so pybind11's index checks raise IndexError for a shorter y or an
           // array of more than one dimension, and a longer y is cut to n
           const size_t n = static_cast<size_t>(x.size());
           std::vector<real_t> xs(n), ys(n);
           for (size_t i = 0; i < n; ++i) {
             xs[i] = x.at(static_cast<py::ssize_t>(i));
             ys[i] = y.at(static_cast<py::ssize_t>(i));
           }
           return TabulatedFunction<LG>(Array1D<real_t>::from_host(xs.data(), n),
                                        Array1D<real_t>::from_host(ys.data(), n), yfill);
         }),
         "x"_a, "y"_a, "yfill"_a = 0.0)
    .def("xArr", &TabulatedFunction<LG>::xArr)
    .def("yArr", &TabulatedFunction<LG>::yArr)
    .def("nPoints", &TabulatedFunction<LG>::nPoints)
    .def("yFill", &TabulatedFunction<LG>::yFill)
    .def("xMin", &TabulatedFunction<LG>::xMin)
    .def("xMax", &TabulatedFunction<LG>::xMax)
    .doc() = "A tabulated function x -> y (yfill outside the x range)";
}

template <dim_t D>
void define_particles(py::module& m) {
  // particles.cpp:385-540
  const std::string name = "Particles_" + std::to_string(D) + "D";
  py::class_<Particles<D>>(m, name.c_str())
    .def(py::init<const std::string&>())
    .def("__repr__", &Particles<D>::repr)
    .def("__len__", &Particles<D>::nactive)
    .def("fromArrays", &Particles<D>::fromArrays, "arrays"_a, "append"_a = false,
         "Fill from a dict of arrays with keys X1..X3, U1..U3, E1..E3, B1..B3 (and W, "
         "per-particle weights); missing components are zero (W: one)")
    .def("setNactive", &Particles<D>::setNactive)
    .def("setIgnoreCoords", &Particles<D>::setIgnoreCoords)
    .def("allocate", &Particles<D>::allocate)
    .def("printHead", &Particles<D>::printHead, "number"_a = 5, "start"_a = 0)
    .def("is_allocated", &Particles<D>::is_allocated)
    .def("nactive", &Particles<D>::nactive)
    .def("nalloc", &Particles<D>::nalloc)
    .def("label", &Particles<D>::label)
    .def("energyDistribution", &Particles<D>::energyDistribution, "energy_bins"_a,
         "fourvel"_a = true, "Energy distribution dN/dE of |u| (fourvel) or sqrt(1 + u^2)")
    .def("X", &Particles<D>::Xarr, "d"_a)
    .def("U", &Particles<D>::Uarr, "d"_a)
    .def("E", &Particles<D>::Earr, "d"_a)
    .def("B", &Particles<D>::Barr, "d"_a)
    .def("W", &Particles<D>::Warr)
    .def("has_weights", &Particles<D>::has_weights)
    .def("_device_pointers",
         [](const Particles<D>& p) {
           // (U1, U2, U3, E1, E2, E3, B1, B2, B3, W) device addresses, 0 when absent
           py::list out;
           for (const auto* q : {&p.U, &p.E, &p.B}) {
             for (int d = 0; d < 3; ++d) {
               out.append(reinterpret_cast<std::uintptr_t>(q->size() > static_cast<size_t>(d)
                                                             ? (*q)[d].dptr()
                                                             : nullptr));
             }
           }
           out.append(reinterpret_cast<std::uintptr_t>(p.has_weights() ? p.W.dptr() : nullptr));
           return out;
         })
    .doc() = "Particle data (device-resident structure of arrays)";
}

template <dim_t D>
void define_sync(py::module& m) {
  // synchrotron.cpp:181-215
  const std::string name = "SynchrotronSpectrum_" + std::to_string(D) + "D";
  m.def(name.c_str(), &SynchrotronSpectrum<D>, "prtls"_a, "bins_e_syn"_a, "B0"_a, "g_syn"_a,
        "e_syn_at_g_syn"_a,
        "Synchrotron spectrum E^2 dN/dE of a particle population (bins in units of mc^2)");
}

}  // namespace

PYBIND11_MODULE(ragnar, m) {
  m.doc() = "ragnar: radiative post-processing of PIC data (MI355X kernels)";

  // pyinterface.cpp:30-49
  m.def(
    "Initialize",
    []() {
      if (rgn_is_initialized()) {
        py::print("ragnar is already initialized");
      } else {
        check(rgn_init(0));
      }
    },
    "Initialize the device runtime");
  m.def(
    "Finalize",
    []() {
      if (!rgn_is_initialized()) {
        py::print("ragnar is not initialized");
      } else {
        check(rgn_finalize());
      }
    },
    "Finalize the device runtime");

  // global.cpp:9-16
  py::class_<EnergyUnits>(m, "EnergyUnits")
    .def_readonly_static("eV", &EnergyUnits::eV)
    .def_readonly_static("MeV", &EnergyUnits::MeV)
    .def_readonly_static("GeV", &EnergyUnits::GeV)
    .def_readonly_static("mec2", &EnergyUnits::mec2)
    .def_readonly_static("mpc2", &EnergyUnits::mpc2);

  // snippets.cpp:170-209
  m.def("Linspace", &Linspace, "start"_a, "stop"_a, "num"_a, "num linearly spaced values");
  m.def("Logspace", &Logspace, "start"_a, "stop"_a, "num"_a, "num logarithmically spaced values");
  define_tabulated<true>(m);
  define_tabulated<false>(m);

  define_array<int>(m, "Array1D_i");
  define_array<float>(m, "Array1D_f");
  define_array<double>(m, "Array1D_d");

  // bins.cpp:33-110
  py::class_<Bins, Array1D<real_t>>(m, "Bins")
    .def(py::init<const Array1D<real_t>&, const std::string&>(), "arr"_a, "unit"_a = "")
    .def(py::init<const std::string&>(), "unit"_a = "")
    .def(py::init<const py::array_t<real_t, py::array::forcecast>&, const std::string&>(), "arr"_a,
         "unit"_a = "")
    .def_readwrite("log_spaced", &Bins::log_spaced)
    .def_readwrite("unit", &Bins::unit)
    .doc() = "Histogram bins with a unit and a log-spacing flag";
  m.def("Linbins", &Linbins, "start"_a, "stop"_a, "num"_a, "unit"_a = "");
  m.def("Logbins", &Logbins, "start"_a, "stop"_a, "num"_a, "unit"_a = "");

  define_particles<1>(m);
  define_particles<2>(m);
  define_particles<3>(m);

  pyDefineH5(m);

  // distributions.cpp:168-346
  py::class_<PlawGenerator>(m, "PlawGenerator")
    .def(py::init<real_t, real_t, real_t>(), "p"_a, "emin"_a = 0.0, "emax"_a = 0.0)
    .def_readonly("p", &PlawGenerator::p)
    .def_readonly("emin", &PlawGenerator::emin)
    .def_readonly("emax", &PlawGenerator::emax)
    .def("compute", &PlawGenerator::compute, "energy_bins"_a);
  py::class_<BrokenPlawGenerator>(m, "BrokenPlawGenerator")
    .def(py::init<real_t, real_t, real_t, real_t, real_t>(), "e_break"_a, "p1"_a, "p2"_a,
         "emin"_a = 0.0, "emax"_a = 0.0)
    .def_readonly("e_break", &BrokenPlawGenerator::e_break)
    .def_readonly("p1", &BrokenPlawGenerator::p1)
    .def_readonly("p2", &BrokenPlawGenerator::p2)
    .def_readonly("emin", &BrokenPlawGenerator::emin)
    .def_readonly("emax", &BrokenPlawGenerator::emax)
    .def("compute", &BrokenPlawGenerator::compute, "energy_bins"_a);
  py::class_<DeltaGenerator>(m, "DeltaGenerator")
    .def(py::init<real_t, real_t>(), "energy0"_a, "denergy"_a)
    .def_readonly("energy0", &DeltaGenerator::energy0)
    .def_readonly("denergy", &DeltaGenerator::denergy)
    .def("compute", &DeltaGenerator::compute, "energy_bins"_a);
  py::class_<TabulatedDistribution>(m, "TabulatedDistribution")
    .def(py::init<const Bins&, const Array1D<real_t>&>(), "bins_energy"_a, "f"_a)
    .def(py::init<const Bins&, const PlawGenerator&>(), "bins_energy"_a, "generator"_a)
    .def(py::init<const Bins&, const BrokenPlawGenerator&>(), "bins_energy"_a, "generator"_a)
    .def(py::init<const Bins&, const DeltaGenerator&>(), "bins_energy"_a, "generator"_a)
    .def("extent", &TabulatedDistribution::extent)
    .def("log_spaced", &TabulatedDistribution::log_spaced)
    .def("EnergyBins", &TabulatedDistribution::EnergyBins)
    .def("F", &TabulatedDistribution::F);

  // synchrotron.cpp:147-215
  m.def("Ffunc_integrand", &Ffunc_integrand, "x"_a);
  m.def("SynchrotronSpectrumFromDist", &SynchrotronSpectrumFromDist, "dist_prtls"_a,
        "bins_e_syn"_a, "g_syn"_a, "e_syn_at_g_syn"_a,
        "Synchrotron spectrum E^2 dN/dE of a tabulated distribution");
  define_sync<1>(m);
  define_sync<2>(m);
  define_sync<3>(m);

  // ic.cpp:48-74
  m.def("ICSpectrum", &ICSpectrum, "dist_prtls"_a, "dist_soft_photons"_a, "bins_e_ic"_a,
        "Inverse-Compton spectrum E^2 dN/dE of a particle distribution scattering soft photons");

  m.def("_default_ftable_nodes", []() {
    const rgn_ftable* t = default_ftable();
    const int n = rgn_ftable_size(t);
    py::array_t<float> x(n), y(n);
    check(rgn_ftable_nodes(t, x.mutable_data(), y.mutable_data()));
    return py::make_tuple(x, y);
  });
}
