// Host-side mirror of the reference's containers (src/containers/*), holding
// device-resident data allocated through the C ABI of libragnar_hip.so.  The
// Python module `ragnar` (pyragnar.cpp) binds these with the reference's names.
#pragma once

#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <cstddef>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "ragnar_hip.h"

namespace py = pybind11;

namespace rgnr {

using real_t = float;
using dim_t = unsigned short;

// Converts a C-ABI status into the reference's exception types (pybind11 maps
// std::runtime_error -> RuntimeError, std::range_error -> ValueError).
void check(int status);

// global.h:21-27
struct EnergyUnits {
  inline static const std::string eV = "eV";
  inline static const std::string MeV = "MeV";
  inline static const std::string GeV = "GeV";
  inline static const std::string mec2 = "mec2";
  inline static const std::string mpc2 = "mpc2";
};

bool verbose();
std::string ToHumanReadable(double value, bool use_suffixes);

// Ref-counted device storage with a lazily refreshed host mirror.  Copies of an
// Array1D alias the same storage, as Kokkos::View copies do (array.hpp:15-42).
template <class T>
struct Storage {
  T* d = nullptr;
  size_t n = 0;
  std::vector<T> host;
  bool host_valid = false;
  Storage() = default;
  Storage(const Storage&) = delete;
  Storage& operator=(const Storage&) = delete;
  ~Storage() {
    if (d) rgn_free(d);
  }
};

template <class T>
struct Array1D {
  std::shared_ptr<Storage<T>> s = std::make_shared<Storage<T>>();

  Array1D() = default;
  explicit Array1D(const py::array_t<T, py::array::forcecast>& arr);  // array.hpp:21-30

  static Array1D from_host(const T* p, size_t n);
  static Array1D device(size_t n);  // zero-filled

  T* dptr() const { return s->d; }
  size_t size() const { return s->n; }
  const std::vector<T>& host() const;  // D->H on first use after a device write
  void mark_device_written() const { s->host_valid = false; }

  void head(std::size_t n = 10, std::size_t start = 0) const;  // array.cpp:18-32
  std::string repr() const;                                     // array.cpp:34-44
  py::array_t<T> as_array() const;                              // array.cpp:46-51
  std::size_t extent(unsigned short d = 0) const { return d == 0 ? s->n : 1; }
};

// bins.hpp:17-30
struct Bins : Array1D<real_t> {
  bool log_spaced{false};
  std::string unit;
  Bins(const Array1D<real_t>& arr, const std::string& unit = "") : Array1D<real_t>{arr}, unit{unit} {}
  explicit Bins(const std::string& unit = "") : unit{unit} {}
  Bins(const py::array_t<real_t, py::array::forcecast>& arr, const std::string& unit = "")
    : Array1D<real_t>{arr}, unit{unit} {}
};

Array1D<real_t> Linspace(real_t start, real_t stop, std::size_t num);  // snippets.cpp:21-37
Array1D<real_t> Logspace(real_t start, real_t stop, std::size_t num);  // snippets.cpp:39-62
Bins Linbins(real_t start, real_t stop, std::size_t num, const std::string& unit = "");
Bins Logbins(real_t start, real_t stop, std::size_t num, const std::string& unit = "");

// tabulation.hpp:55-116
template <bool LG>
struct TabulatedFunction {
  Array1D<real_t> m_x, m_y;
  real_t m_yfill;
  std::size_t m_n;
  real_t m_xmin, m_xmax;
  TabulatedFunction(const Array1D<real_t>& x, const Array1D<real_t>& y, real_t yfill = 0.0);
  const Array1D<real_t>& xArr() const { return m_x; }
  const Array1D<real_t>& yArr() const { return m_y; }
  std::size_t nPoints() const { return m_n; }
  real_t yFill() const { return m_yfill; }
  real_t xMin() const { return m_xmin; }
  real_t xMax() const { return m_xmax; }
};

// distributions.hpp:15-61 (host-evaluated generators)
struct DistGenerator {
  virtual ~DistGenerator() = default;
  virtual real_t f(real_t energy) const = 0;
  Array1D<real_t> compute(const Bins& energy_bins) const;  // distributions.cpp:19-35
 protected:
  real_t m_norm{0};
};

struct PlawGenerator : DistGenerator {
  const real_t p, emin, emax;
  PlawGenerator(real_t p, real_t emin = 0.0, real_t emax = 0.0);
  real_t f(real_t energy) const override;
};

struct BrokenPlawGenerator : DistGenerator {
  const real_t e_break, emin, emax, p1, p2;
  BrokenPlawGenerator(real_t e_break, real_t p1, real_t p2, real_t emin = 0.0, real_t emax = 0.0);
  real_t f(real_t energy) const override;
};

struct DeltaGenerator : DistGenerator {
  const real_t energy0, denergy;
  DeltaGenerator(real_t energy0, real_t denergy);
  real_t f(real_t energy) const override;
};

// distributions.hpp:63-78
struct TabulatedDistribution {
  Bins m_e_bins;
  Array1D<real_t> m_f;
  TabulatedDistribution(const Bins& e_bins, const Array1D<real_t>& f);
  TabulatedDistribution(const Bins& e_bins, const DistGenerator& generator);
  const Bins& EnergyBins() const { return m_e_bins; }
  const Array1D<real_t>& F() const { return m_f; }
  std::size_t extent() const { return m_f.size(); }
  bool log_spaced() const { return m_e_bins.log_spaced; }
};

// particles.hpp:18-72, stored as device structure-of-arrays.  W (per-particle
// weights, default 1) has no reference counterpart.
template <dim_t D>
struct Particles {
  std::vector<Array1D<real_t>> X, U, E, B;  // D, 3, 3, 3 components
  Array1D<real_t> W;                       // empty unless weights were given

  explicit Particles(const std::string& label) : m_label{label} {}

  void fromArrays(const std::map<std::string, py::array_t<real_t, py::array::forcecast>>& arrays,
                  bool append = false);
  void allocate(std::size_t nalloc);
  void reallocate(std::size_t nalloc);
  void setNactive(std::size_t nactive);
  void setIgnoreCoords(bool ignore) { m_coords_ignored = ignore; }
  bool is_allocated() const { return m_is_allocated; }
  std::size_t nalloc() const { return m_nalloc; }
  std::size_t nactive() const { return m_nactive; }
  const std::string& label() const { return m_label; }
  bool has_weights() const { return W.size() > 0; }
  void printHead(std::size_t number = 5, std::size_t start = 0) const;
  std::string repr() const;
  TabulatedDistribution energyDistribution(const Bins& energy_bins, bool fourvel = true) const;
  Array1D<real_t> Xarr(std::size_t d) const;
  Array1D<real_t> Uarr(std::size_t d) const;
  Array1D<real_t> Earr(std::size_t d) const;
  Array1D<real_t> Barr(std::size_t d) const;
  Array1D<real_t> Warr() const;

 private:
  bool m_is_allocated{false};
  bool m_coords_ignored{false};
  std::size_t m_nactive{0};
  std::size_t m_nalloc{0};
  const std::string m_label;
};

// physics
real_t Ffunc_integrand(real_t x);
Array1D<real_t> SynchrotronSpectrumFromDist(const TabulatedDistribution& dist, const Bins& bins,
                                            real_t g_syn, real_t e_syn_at_g_syn);
// ic.cpp:15-46
Array1D<real_t> ICSpectrum(const TabulatedDistribution& dist_prtls,
                           const TabulatedDistribution& dist_soft_photons, const Bins& bins_e_ic);
template <dim_t D>
Array1D<real_t> SynchrotronSpectrum(const Particles<D>& prtls, const Bins& bins, real_t B0,
                                    real_t g_syn, real_t e_syn_at_g_syn);

// The reference F table (TabulateFfunc defaults), built once per process.
const rgn_ftable* default_ftable();

}  // namespace rgnr
