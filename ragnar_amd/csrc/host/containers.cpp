// Host-side containers and physics drivers of the `ragnar` module.  All device
// memory and every kernel go through the C ABI (ragnar_hip.h).
#include "containers.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <stdexcept>

using namespace pybind11::literals;

namespace rgnr {

void check(int status) {
  if (status == RGN_OK) return;
  const std::string msg = rgn_last_error();
  switch (status) {
    case RGN_ERR_RANGE:
      throw std::range_error(msg);
    case RGN_ERR_NOMEM:
      throw std::runtime_error("out of device memory: " + msg);
    default:
      throw std::runtime_error(msg);
  }
}

bool verbose() {
  static const bool v = [] {
    const char* s = std::getenv("RAGNAR_VERBOSE");
    return !(s && std::strcmp(s, "0") == 0);
  }();
  return v;
}

// snippets.cpp:64-111 (the branch with use_suffixes == false is the one the
// reference calls, USE_POW10 = false).  A zero value is printed as 0.00·10^0;
// the reference loops forever on it.
std::string ToHumanReadable(double value, bool use_suffixes) {
  const bool negative = value < 0;
  double v = negative ? -value : value;
  char buf[64];
  if (use_suffixes) {
    static const char* suffixes[9] = {"p", "n", "μ", "m", "", "k", "M", "G", "T"};
    int sidx = 4;
    while (v != 0.0 && (v < 0.01 || v >= 1000)) {
      if (v < 0.01) {
        if (sidx == 0) break;
        v *= 1000;
        --sidx;
      } else {
        if (sidx == 8) break;
        v /= 1000;
        ++sidx;
      }
    }
    std::snprintf(buf, sizeof buf, "%.2f", v);
    return (negative ? "-" : "") + std::string(buf) + " " + suffixes[sidx];
  }
  int pow = 0;
  while (v != 0.0 && (v < 0.1 || v >= 10)) {
    if (v < 0.1) {
      v *= 10;
      --pow;
    } else {
      v /= 10;
      ++pow;
    }
  }
  std::snprintf(buf, sizeof buf, "%.2f", v);
  return (negative ? "-" : "") + std::string(buf) + "·10^" + std::to_string(pow);
}

// ---------------------------------------------------------------------------
// Array1D
// ---------------------------------------------------------------------------
template <class T>
Array1D<T>::Array1D(const py::array_t<T, py::array::forcecast>& arr) {
  const auto c = py::array_t<T, py::array::c_style | py::array::forcecast>::ensure(arr);
  if (!c) throw std::runtime_error("cannot convert the input to a contiguous array");
  *this = from_host(c.data(), static_cast<size_t>(c.size()));
}

template <class T>
Array1D<T> Array1D<T>::from_host(const T* p, size_t n) {
  Array1D<T> a;
  a.s->n = n;
  a.s->host.assign(p, p + n);
  a.s->host_valid = true;
  if (n) {
    void* d = nullptr;
    check(rgn_malloc(&d, n * sizeof(T)));
    a.s->d = static_cast<T*>(d);
    check(rgn_memcpy_h2d(d, p, n * sizeof(T)));
  }
  return a;
}

template <class T>
Array1D<T> Array1D<T>::device(size_t n) {
  Array1D<T> a;
  a.s->n = n;
  if (n) {
    void* d = nullptr;
    check(rgn_malloc(&d, n * sizeof(T)));
    a.s->d = static_cast<T*>(d);
    check(rgn_memset(d, 0, n * sizeof(T)));
  }
  a.s->host.assign(n, T(0));
  a.s->host_valid = true;
  return a;
}

template <class T>
const std::vector<T>& Array1D<T>::host() const {
  if (!s->host_valid) {
    s->host.resize(s->n);
    if (s->n) check(rgn_memcpy_d2h(s->host.data(), s->d, s->n * sizeof(T)));
    s->host_valid = true;
  }
  return s->host;
}

template <class T>
void Array1D<T>::head(std::size_t n, std::size_t start) const {
  if (start + n > s->n) throw std::range_error("Array::head: n > data.extent(0)");
  n = std::min(n, s->n - start);
  const auto& h = host();
  for (std::size_t i = 0; i < n; ++i) py::print(h[start + i], " ", "end"_a = "");
  py::print();
}

template <class T>
std::string Array1D<T>::repr() const {
  return "1D Array [ size: " + std::to_string(s->n) + " ]";
}

template <class T>
py::array_t<T> Array1D<T>::as_array() const {
  const auto& h = host();
  py::array_t<T> out(static_cast<py::ssize_t>(h.size()));
  if (!h.empty()) std::memcpy(out.mutable_data(), h.data(), h.size() * sizeof(T));
  return out;
}

template struct Array1D<int>;
template struct Array1D<float>;
template struct Array1D<double>;

// ---------------------------------------------------------------------------
// Grids (snippets.cpp:21-62, bins.cpp:15-31)
// ---------------------------------------------------------------------------
Array1D<real_t> Linspace(real_t start, real_t stop, std::size_t num) {
  if (start >= stop) throw std::runtime_error("Linspace start must be < stop");
  std::vector<real_t> v(num);
  for (std::size_t i = 0; i < num; ++i) {
    v[i] = (num == 1) ? start
                      : start + static_cast<real_t>(i) * (stop - start) / static_cast<real_t>(num - 1);
  }
  return Array1D<real_t>::from_host(v.data(), num);
}

Array1D<real_t> Logspace(real_t start, real_t stop, std::size_t num) {
  if (start <= 0.0 or stop <= 0.0) {
    throw std::runtime_error("Logspace start and stop must be strictly positive");
  }
  if (start >= stop) throw std::runtime_error("Logspace start must be < stop");
  std::vector<real_t> v(num);
  for (std::size_t i = 0; i < num; ++i) {
    if (num == 1) {
      v[i] = start;
    } else {
      // float exponent; pow(10, float) promotes to double (Kokkos math::pow)
      const real_t e = std::log10(start) +
                       static_cast<real_t>(i) * (std::log10(stop / start)) / static_cast<real_t>(num - 1);
      v[i] = static_cast<real_t>(std::pow(10.0, static_cast<double>(e)));
    }
  }
  return Array1D<real_t>::from_host(v.data(), num);
}

Bins Linbins(real_t start, real_t stop, std::size_t num, const std::string& unit) {
  Bins b(Linspace(start, stop, num), unit);
  b.log_spaced = false;
  return b;
}

Bins Logbins(real_t start, real_t stop, std::size_t num, const std::string& unit) {
  Bins b(Logspace(start, stop, num), unit);
  b.log_spaced = true;
  return b;
}

// ---------------------------------------------------------------------------
// TabulatedFunction (tabulation.cpp:19-117)
// ---------------------------------------------------------------------------
template <bool LG>
TabulatedFunction<LG>::TabulatedFunction(const Array1D<real_t>& x, const Array1D<real_t>& y,
                                         real_t yfill)
  : m_x{x}, m_y{y}, m_yfill{yfill}, m_n{x.size()} {
  const auto& h = m_x.host();
  m_xmin = h.empty() ? 0.0f : h[0];
  m_xmax = m_xmin;
  for (real_t v : h) {
    m_xmin = v < m_xmin ? v : m_xmin;
    m_xmax = v > m_xmax ? v : m_xmax;
  }
  if (m_y.size() != m_n) throw std::range_error("y.size != x.size in TabulatedFunction");
  if (m_xmin >= m_xmax) throw std::range_error("xmin >= xmax in TabulatedFunction");
  if (LG && m_xmin <= 0.0) throw std::range_error("xmin <= 0.0 in Logspace TabulatedFunction");
}

template struct TabulatedFunction<true>;
template struct TabulatedFunction<false>;

// ---------------------------------------------------------------------------
// Generators (distributions.cpp:19-166)
// ---------------------------------------------------------------------------
Array1D<real_t> DistGenerator::compute(const Bins& energy_bins) const {
  const auto& e = energy_bins.host();
  std::vector<real_t> f(e.size());
  for (size_t i = 0; i < e.size(); ++i) f[i] = this->f(e[i]);
  return Array1D<real_t>::from_host(f.data(), f.size());
}

PlawGenerator::PlawGenerator(real_t p, real_t emin, real_t emax) : p{p}, emin{emin}, emax{emax} {
  if (emin < 0.0) {
    throw std::runtime_error("emin < 0.0");
  } else if (emin > emax and emax > 0.0) {
    throw std::runtime_error("emin > emax");
  } else if (emin == 0.0 and p <= -1.0) {
    throw std::runtime_error("p <= -1 and emin = 0.0 : normalization diverges");
  } else if (emax == 0.0 and p >= -1.0) {
    throw std::runtime_error("p >= -1 and emax = 0.0 (infinity) : normalization diverges");
  } else if (emax == 0.0) {
    m_norm = -std::pow(emin, p + 1) / (p + 1);
  } else if (p != -1.0) {
    m_norm = (std::pow(emax, p + 1) - std::pow(emin, p + 1)) / (p + 1);
  } else {
    m_norm = std::log(emax / emin);
  }
}

real_t PlawGenerator::f(real_t energy) const {
  if (energy < emin or (emax > 0.0 and energy >= emax)) return 0.0;
  return std::pow(energy, p) / m_norm;
}

BrokenPlawGenerator::BrokenPlawGenerator(real_t e_break, real_t p1, real_t p2, real_t emin,
                                         real_t emax)
  : e_break{e_break}, emin{emin}, emax{emax}, p1{p1}, p2{p2} {
  if (e_break <= 0.0) throw std::runtime_error("e_break <= 0.0");
  if (emin < 0.0) throw std::runtime_error("emin < 0.0");
  if (emin > emax and emax > 0.0) throw std::runtime_error("emin > emax");
  real_t pre_break, post_break;
  if (emin == 0.0 and p1 <= -1.0) {
    throw std::runtime_error("p1 <= -1 and emin = 0.0 : normalization diverges");
  } else if (e_break <= emin) {
    pre_break = 0.0;
  } else {
    pre_break = (1 - std::pow(emin / e_break, p1 + 1)) / (p1 + 1);
  }
  if (emax == 0.0 and p2 >= -1.0) {
    throw std::runtime_error("p2 >= -1 and emax = 0.0 (infinity) : normalization diverges");
  } else if (emax == 0.0) {
    post_break = 1 / (-p2 - 1);
  } else if (p2 == -1.0) {
    post_break = std::log(emax / e_break);
  } else {
    post_break = (std::pow(emax / e_break, p2 + 1) - 1) / (p2 + 1);
  }
  m_norm = (pre_break + post_break) * e_break;
}

real_t BrokenPlawGenerator::f(real_t energy) const {
  if (energy < emin or (emax > 0.0 and energy >= emax)) return 0.0;
  if (energy < e_break) return std::pow(energy / e_break, p1) / m_norm;
  return std::pow(energy / e_break, p2) / m_norm;
}

DeltaGenerator::DeltaGenerator(real_t energy0, real_t denergy) : energy0{energy0}, denergy{denergy} {
  m_norm = static_cast<real_t>(1.0 / denergy);
}

real_t DeltaGenerator::f(real_t energy) const {
  if (std::abs(energy - energy0) < denergy * 0.5) return static_cast<real_t>(1.0 * m_norm);
  return 0.0;
}

TabulatedDistribution::TabulatedDistribution(const Bins& e_bins, const Array1D<real_t>& f)
  : m_e_bins{e_bins}, m_f{f} {
  if (m_e_bins.size() != m_f.size()) throw std::runtime_error("e_bins.extent() != f.extent()");
}

TabulatedDistribution::TabulatedDistribution(const Bins& e_bins, const DistGenerator& generator)
  : m_e_bins{e_bins}, m_f{generator.compute(e_bins)} {}

// ---------------------------------------------------------------------------
// Particles (particles.cpp:25-383)
// ---------------------------------------------------------------------------
namespace {
Array1D<real_t> grown_copy(const Array1D<real_t>& old, std::size_t old_n, std::size_t new_n) {
  Array1D<real_t> a = Array1D<real_t>::device(new_n);
  if (old_n && old.dptr()) check(rgn_memcpy_d2d(a.dptr(), old.dptr(), old_n * sizeof(real_t)));
  a.mark_device_written();
  return a;
}

void fill_ones(real_t* d, std::size_t n) {
  if (!n) return;
  std::vector<real_t> ones(n, 1.0f);
  check(rgn_memcpy_h2d(d, ones.data(), n * sizeof(real_t)));
}
}  // namespace

template <dim_t D>
void Particles<D>::allocate(std::size_t nalloc) {
  if (is_allocated()) throw std::runtime_error("Particles already allocated");
  X.clear();
  if (not m_coords_ignored) {
    for (dim_t d = 0; d < D; ++d) X.push_back(Array1D<real_t>::device(nalloc));
  }
  U.clear();
  E.clear();
  B.clear();
  for (int d = 0; d < 3; ++d) {
    U.push_back(Array1D<real_t>::device(nalloc));
    E.push_back(Array1D<real_t>::device(nalloc));
    B.push_back(Array1D<real_t>::device(nalloc));
  }
  m_is_allocated = true;
  m_nalloc = nalloc;
}

template <dim_t D>
void Particles<D>::reallocate(std::size_t nalloc) {
  if (!is_allocated()) {
    throw std::runtime_error("Particles not allocated, if you want to allocate, call `allocate` instead");
  }
  if (nalloc <= m_nalloc) {
    throw std::runtime_error("New allocation size must be greater than the current one");
  }
  for (auto& a : X) a = grown_copy(a, m_nalloc, nalloc);
  for (auto* q : {&U, &E, &B}) {
    for (auto& a : *q) a = grown_copy(a, m_nalloc, nalloc);
  }
  if (has_weights()) {
    W = grown_copy(W, m_nalloc, nalloc);
    fill_ones(W.dptr() + m_nalloc, nalloc - m_nalloc);
  }
  m_nalloc = nalloc;
}

template <dim_t D>
void Particles<D>::setNactive(std::size_t nactive) {
  if (!is_allocated()) throw std::runtime_error("Particles not allocated");
  m_nactive = nactive;
}

template <dim_t D>
void Particles<D>::fromArrays(
  const std::map<std::string, py::array_t<real_t, py::array::forcecast>>& arrays, bool append) {
  if ((not append) and is_allocated()) {
    throw std::runtime_error("Particles already allocated, if you want to append, specify `append = True`");
  }
  std::size_t nprtls = 0;
  for (const auto& [name, arr] : arrays) {
    const auto n = static_cast<std::size_t>(arr.ndim() ? arr.shape(0) : 1);
    if (nprtls == 0) {
      nprtls = n;
    } else if (nprtls != n) {
      throw std::runtime_error("Inconsistent number of particles");
    }
  }
  if (nprtls == 0) throw std::runtime_error("No particles provided");
  std::size_t start = 0;
  if (is_allocated()) {
    start = m_nactive;
    if (start + nprtls > m_nalloc) reallocate(m_nactive + nprtls);
  } else {
    allocate(nprtls);
  }
  const std::size_t bytes = nprtls * sizeof(real_t);
  auto put = [&](Array1D<real_t>& dst, const std::string& key) {
    auto it = arrays.find(key);
    if (it != arrays.end()) {
      const auto c =
        py::array_t<real_t, py::array::c_style | py::array::forcecast>::ensure(it->second);
      check(rgn_memcpy_h2d(dst.dptr() + start, c.data(), bytes));
    } else {
      check(rgn_memset(dst.dptr() + start, 0, bytes));
    }
    dst.mark_device_written();
  };
  bool has_coords = not m_coords_ignored;
  for (dim_t d = 0; d < D; ++d) {
    if (arrays.count("X" + std::to_string(d + 1))) has_coords = true;
  }
  setIgnoreCoords(not has_coords);
  if (has_coords) {
    if (X.empty()) {
      for (dim_t d = 0; d < D; ++d) X.push_back(Array1D<real_t>::device(m_nalloc));
    }
    for (dim_t d = 0; d < D; ++d) put(X[d], "X" + std::to_string(d + 1));
  }
  for (int d = 0; d < 3; ++d) {
    put(U[d], "U" + std::to_string(d + 1));
    put(E[d], "E" + std::to_string(d + 1));
    put(B[d], "B" + std::to_string(d + 1));
  }
  auto wit = arrays.find("W");
  if (wit != arrays.end() && !has_weights()) {
    W = Array1D<real_t>::device(m_nalloc);
    fill_ones(W.dptr(), m_nalloc);
  }
  if (has_weights()) {
    if (wit != arrays.end()) {
      put(W, "W");
    } else {
      fill_ones(W.dptr() + start, nprtls);
      W.mark_device_written();
    }
  }
  setNactive(start + nprtls);
}

template <dim_t D>
void Particles<D>::printHead(std::size_t number, std::size_t start) const {
  py::print("Particles:", m_label);
  if (!is_allocated()) {
    py::print(" [ not allocated ]");
    return;
  }
  if (start + number > m_nalloc) {
    throw std::runtime_error("Number of particles to print exceeds allocated space");
  }
  py::print(" [", m_nactive, "/", m_nalloc, "]");
  py::print(" showing", start, "to", start + number);
  auto slice = [&](const Array1D<real_t>& a) {
    std::vector<real_t> h(number);
    if (number) check(rgn_memcpy_d2h(h.data(), a.dptr() + start, number * sizeof(real_t)));
    return h;
  };
  if (not m_coords_ignored && !X.empty()) {
    for (dim_t d = 0; d < D; ++d) {
      const auto h = slice(X[d]);
      py::print(" X_" + std::to_string(d + 1), ":", "end"_a = "");
      if (start > 0) py::print("...", "end"_a = "");
      for (std::size_t i = 0; i < number; ++i) py::print(h[i], "end"_a = "");
      if (m_nactive > number) {
        py::print("...");
      } else {
        py::print();
      }
    }
  }
  const char* names[3] = {"U", "E", "B"};
  const std::vector<Array1D<real_t>>* qs[3] = {&U, &E, &B};
  for (int q = 0; q < 3; ++q) {
    for (int d = 0; d < 3; ++d) {
      const auto h = slice((*qs[q])[d]);
      py::print(" ", std::string(names[q]) + "_" + std::to_string(d + 1), ":", "end"_a = "");
      if (start > 0) py::print("...", "end"_a = "");
      for (std::size_t i = 0; i < number; ++i) py::print(h[i], " ", "end"_a = "");
      if (m_nalloc > start + number) {
        py::print("...");
      } else {
        py::print();
      }
    }
  }
}

template <dim_t D>
std::string Particles<D>::repr() const {
  return "Particles<" + std::to_string(D) + "D> (" + m_label + ") : " +
         (is_allocated() ? ToHumanReadable(static_cast<double>(nactive()), false) : "not allocated");
}

namespace {
Array1D<real_t> component_copy(const std::vector<Array1D<real_t>>& v, std::size_t comp,
                               std::size_t n) {
  if (comp >= v.size()) throw std::out_of_range("Invalid component");
  Array1D<real_t> a = Array1D<real_t>::device(n);
  if (n) check(rgn_memcpy_d2d(a.dptr(), v[comp].dptr(), n * sizeof(real_t)));
  a.mark_device_written();
  return a;
}
}  // namespace

template <dim_t D>
Array1D<real_t> Particles<D>::Xarr(std::size_t d) const {
  if (m_coords_ignored) throw std::runtime_error("Particle coordinates ignored");
  if (d - 1 >= D) throw std::out_of_range("Invalid component");
  return component_copy(X, d - 1, m_nactive);
}

template <dim_t D>
Array1D<real_t> Particles<D>::Uarr(std::size_t d) const {
  if (d - 1 >= 3) throw std::out_of_range("Invalid component");
  return component_copy(U, d - 1, m_nactive);
}

template <dim_t D>
Array1D<real_t> Particles<D>::Earr(std::size_t d) const {
  if (d - 1 >= 3) throw std::out_of_range("Invalid component");
  return component_copy(E, d - 1, m_nactive);
}

template <dim_t D>
Array1D<real_t> Particles<D>::Barr(std::size_t d) const {
  if (d - 1 >= 3) throw std::out_of_range("Invalid component");
  return component_copy(B, d - 1, m_nactive);
}

template <dim_t D>
Array1D<real_t> Particles<D>::Warr() const {
  if (!has_weights()) {
    Array1D<real_t> a = Array1D<real_t>::device(m_nactive);
    fill_ones(a.dptr(), m_nactive);
    a.mark_device_written();
    return a;
  }
  return component_copy(std::vector<Array1D<real_t>>{W}, 0, m_nactive);
}

template <dim_t D>
TabulatedDistribution Particles<D>::energyDistribution(const Bins& energy_bins, bool fourvel) const {
  if (verbose()) {
    py::print("Computing energy distribution for", label(), "...", "end"_a = "", "flush"_a = true);
  }
  const std::size_t nb = energy_bins.size();
  Array1D<real_t> hist = Array1D<real_t>::device(nb);
  if (nb && m_nactive && is_allocated()) {
    rgn_hist_plan* plan = nullptr;
    check(rgn_hist_plan_create(energy_bins.host().data(), static_cast<int>(nb),
                               energy_bins.log_spaced ? 1 : 0, &plan));
    const float* u[3] = {U[0].dptr(), U[1].dptr(), U[2].dptr()};
    const int st = rgn_energy_hist(plan, u, m_nactive, fourvel ? 1 : 0, hist.dptr(), nullptr, nullptr);
    const int st2 = (st == RGN_OK) ? rgn_sync(nullptr) : st;
    rgn_hist_plan_destroy(plan);
    check(st2);
    hist.mark_device_written();
  }
  if (verbose()) py::print(": OK", "flush"_a = true);
  return TabulatedDistribution{energy_bins, hist};
}

template struct Particles<1>;
template struct Particles<2>;
template struct Particles<3>;

// ---------------------------------------------------------------------------
// Physics drivers (synchrotron.cpp:28-145)
// ---------------------------------------------------------------------------
const rgn_ftable* default_ftable() {
  static std::mutex mu;
  static rgn_ftable* table = nullptr;
  std::lock_guard<std::mutex> lk(mu);
  if (!table) {
    // TabulateFfunc() defaults, synchrotron.hpp:26-29
    check(rgn_ftable_build(RGN_FTABLE_REFERENCE, 200, static_cast<real_t>(1e-6),
                           static_cast<real_t>(100), &table));
  }
  return table;
}

real_t Ffunc_integrand(real_t x) {
  float out = 0.0f;
  check(rgn_ffunc_integrand(x, &out));
  return out;
}

namespace {
constexpr int kBinsPerPlan = 512;

struct PlanGuard {
  rgn_sync_plan* p = nullptr;
  ~PlanGuard() {
    if (p) rgn_sync_plan_destroy(p);
  }
};
}  // namespace

Array1D<real_t> SynchrotronSpectrumFromDist(const TabulatedDistribution& dist, const Bins& bins,
                                            real_t g_syn, real_t e_syn_at_g_syn) {
  if (verbose()) py::print("Computing synchrotron spectrum from a distribution", "flush"_a = true);
  const rgn_ftable* table = default_ftable();
  const std::size_t nbins = bins.size();
  Array1D<real_t> out = Array1D<real_t>::device(nbins);
  if (verbose()) {
    py::print(" Launching", ToHumanReadable(static_cast<double>(dist.extent() * nbins), false),
              "threads", "end"_a = "", "flush"_a = true);
  }
  const auto& hb = bins.host();
  for (std::size_t c0 = 0; c0 < nbins; c0 += kBinsPerPlan) {
    const int nc = static_cast<int>(std::min<std::size_t>(kBinsPerPlan, nbins - c0));
    PlanGuard plan;
    check(rgn_sync_plan_create(table, hb.data() + c0, nc, 1.0f, g_syn, e_syn_at_g_syn, &plan.p));
    check(rgn_sync_fromdist(plan.p, dist.EnergyBins().dptr(), dist.F().dptr(),
                            static_cast<int>(dist.extent()), dist.log_spaced() ? 1 : 0,
                            out.dptr() + c0, nullptr));
    check(rgn_sync(nullptr));
  }
  out.mark_device_written();
  if (verbose()) py::print(": OK", "flush"_a = true);
  return out;
}

// ic.cpp:15-46.  The progress lines come first and the unit checks of the
// reference's kernel constructor (ic.hpp:52-59) after them, as in the reference.
Array1D<real_t> ICSpectrum(const TabulatedDistribution& dist_prtls,
                           const TabulatedDistribution& dist_soft_photons, const Bins& bins_e_ic) {
  if (verbose()) py::print("Computing IC spectrum ...", "flush"_a = true);
  const std::size_t ng = dist_prtls.extent(), nes = dist_soft_photons.extent(), nic = bins_e_ic.size();
  if (verbose()) {
    py::print(" Launching", ToHumanReadable(static_cast<double>(ng * nes * nic), false), "threads",
              "end"_a = "", "flush"_a = true);
  }
  if (dist_soft_photons.EnergyBins().unit != EnergyUnits::mec2) {
    throw std::runtime_error("Soft photons energy bins must be in units of mec^2");
  }
  if (bins_e_ic.unit != EnergyUnits::mec2) {
    throw std::runtime_error("E_ic must be in units of mec^2");
  }
  constexpr std::size_t kMaxBins = static_cast<std::size_t>(std::numeric_limits<int>::max());
  if (ng > kMaxBins || nes > kMaxBins || nic > kMaxBins) throw std::range_error("too many bins");
  Array1D<real_t> out = Array1D<real_t>::device(nic);
  check(rgn_ic_spectrum(dist_prtls.EnergyBins().dptr(), dist_prtls.F().dptr(), static_cast<int>(ng),
                        dist_prtls.log_spaced() ? 1 : 0, dist_soft_photons.EnergyBins().dptr(),
                        dist_soft_photons.F().dptr(), static_cast<int>(nes), bins_e_ic.dptr(),
                        static_cast<int>(nic), out.dptr(), nullptr));
  check(rgn_sync(nullptr));
  out.mark_device_written();
  if (verbose()) py::print(": OK", "flush"_a = true);
  return out;
}

template <dim_t D>
Array1D<real_t> SynchrotronSpectrum(const Particles<D>& prtls, const Bins& bins, real_t B0,
                                    real_t g_syn, real_t e_syn_at_g_syn) {
  if (verbose()) py::print("Computing synchrotron spectrum for", prtls.label(), "flush"_a = true);
  const rgn_ftable* table = default_ftable();
  const std::size_t nbins = bins.size();
  if (verbose()) {
    py::print(" Launching", ToHumanReadable(static_cast<double>(prtls.nactive() * nbins), false),
              "threads", "end"_a = "", "flush"_a = true);
  }
  // synchrotron.hpp:139-142
  if (bins.unit != EnergyUnits::mec2 and bins.unit != EnergyUnits::mpc2) {
    throw std::runtime_error("bins_e_syn must be in units of mc^2");
  }
  Array1D<real_t> out = Array1D<real_t>::device(nbins);
  const std::size_t n = prtls.is_allocated() ? prtls.nactive() : 0;
  if (n && nbins) {
    const float* u[3] = {prtls.U[0].dptr(), prtls.U[1].dptr(), prtls.U[2].dptr()};
    const float* e[3] = {prtls.E[0].dptr(), prtls.E[1].dptr(), prtls.E[2].dptr()};
    const float* b[3] = {prtls.B[0].dptr(), prtls.B[1].dptr(), prtls.B[2].dptr()};
    const float* w = prtls.has_weights() ? prtls.W.dptr() : nullptr;
    const auto& hb = bins.host();
    for (std::size_t c0 = 0; c0 < nbins; c0 += kBinsPerPlan) {
      const int nc = static_cast<int>(std::min<std::size_t>(kBinsPerPlan, nbins - c0));
      PlanGuard plan;
      check(rgn_sync_plan_create(table, hb.data() + c0, nc, B0, g_syn, e_syn_at_g_syn, &plan.p));
      check(rgn_sync_spectrum(plan.p, u, e, b, w, n, out.dptr() + c0, nullptr));
      check(rgn_sync(nullptr));
    }
    out.mark_device_written();
  }
  if (verbose()) py::print(": OK", "flush"_a = true);
  return out;
}

template Array1D<real_t> SynchrotronSpectrum<1>(const Particles<1>&, const Bins&, real_t, real_t, real_t);
template Array1D<real_t> SynchrotronSpectrum<2>(const Particles<2>&, const Bins&, real_t, real_t, real_t);
template Array1D<real_t> SynchrotronSpectrum<3>(const Particles<3>&, const Bins&, real_t, real_t, real_t);

}  // namespace rgnr
