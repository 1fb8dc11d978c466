// Host build of the kernel arithmetic in cp3_math.hpp: the exact per-point
// decomposition the HIP kernels execute, run on the CPU one point at a time.
// It is a measurement and test aid (numerics of the rewrite against the
// oracle without a GPU, and operation counting), not a fallback: the
// propagator library never calls it.
#include <cstdio>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "cp3_host.hpp"
#include "cp3_math.hpp"

namespace {

void put_err(char* err, int errlen, const char* msg) {
  if (err && errlen > 0) std::snprintf(err, (size_t)errlen, "%s", msg);
}

// One trajectory through the table row of cosine c, as a kernel lane runs it.
template <class P>
void atmos_point(const cp3::SetConst& sc, const cp3::TrajTable& tt, int c,
                 double E, double* out9) {
  using D = typename P::D;
  using F = typename P::F;
  const D e = (D)E;
  const D inv = cp3::m_recip(e);
  const int M = tt.M[(size_t)c];
  cp3::Traj<P> t;
  cp3::traj_init<P>(t, M == 0);
  for (int i = 0; i < M; ++i) {
    cp3::Slab<P> s;
    cp3::solve_slab<P>(sc, e, inv, (D)tt.rye[(size_t)tt.rye_cs * c + (size_t)i], s);
    const D L = (D)tt.seg[(size_t)c * tt.S + (size_t)i];
    if (i < M - 1)
      cp3::traj_middle<P>(t, s, L);
    else
      cp3::traj_inner<P>(t, s, L);
  }
  cp3::Slab<P> vac;
  cp3::solve_slab<P>(sc, e, inv, D(0), vac);
  F prob[9];
  cp3::traj_finish<P>(t, vac, (D)tt.Lvac[(size_t)c], sc, prob);
  for (int k = 0; k < 9; ++k) out9[k] = (double)prob[k];
}

// Production-height bins of one point, read from host arrays: hmw of the
// cosine ([nb][2], km) and the bin probabilities in the reference layout
// [type][flavour][E][cos][bin] at this (type, E, cos).
struct HostBins {
  int n;
  const double* hmw;
  const double* w;
  int64_t flav;
  template <class D>
  void shift(int b, D& hm, D& hw) const {
    hm = (D)hmw[2 * b];
    hw = (D)hmw[2 * b + 1];
  }
  double weight(int a, int b) const { return w[(int64_t)a * flav + b]; }
};

// The same point with production-height averaging.
template <class P>
void atmos_point_avg(const cp3::SetConst& sc, const cp3::TrajTable& tt, int c,
                     double E, const HostBins& bins, double* out9) {
  using D = typename P::D;
  using F = typename P::F;
  const D e = (D)E;
  const D inv = cp3::m_recip(e);
  const int M = tt.M[(size_t)c];
  cp3::Traj<P> t;
  cp3::traj_init<P>(t, M == 0);
  for (int i = 0; i < M; ++i) {
    cp3::Slab<P> s;
    cp3::solve_slab<P>(sc, e, inv, (D)tt.rye[(size_t)tt.rye_cs * c + (size_t)i], s);
    const D L = (D)tt.seg[(size_t)c * tt.S + (size_t)i];
    if (i < M - 1)
      cp3::traj_middle<P>(t, s, L);
    else
      cp3::traj_inner<P>(t, s, L);
  }
  cp3::Slab<P> vac;
  cp3::solve_slab<P>(sc, e, inv, D(0), vac);
  F prob[9];
  cp3::traj_finish_avg<P>(t, vac, (D)tt.Lvac[(size_t)c], sc, bins, prob);
  for (int k = 0; k < 9; ++k) out9[k] = (double)prob[k];
}

template <class P>
void run_atmos_avg(const cp3::SetConst& sc, const cp3::TrajTable& tt, const double* ev,
                   int nE, int nutype, int nb, const double* prob,
                   const std::vector<double>& hmw, double* out) {
  const int64_t nC = tt.nC;
  const int64_t flav = (int64_t)nE * nC * nb;
#pragma omp parallel for schedule(dynamic)
  for (int c = 0; c < tt.nC; ++c)
    for (int e = 0; e < nE; ++e) {
      HostBins b{nb, hmw.data() + (size_t)c * nb * 2,
                 prob + (int64_t)nutype * 3 * flav + ((int64_t)e * nC + c) * nb, flav};
      atmos_point_avg<P>(sc, tt, c, ev[e], b, out + ((size_t)c * nE + e) * 9);
    }
}

template <class P>
void run_atmos(const cp3::SetConst& sc, const cp3::TrajTable& tt,
               const double* ev, int nE, double* out) {
#pragma omp parallel for schedule(dynamic)
  for (int c = 0; c < tt.nC; ++c)
    for (int e = 0; e < nE; ++e)
      atmos_point<P>(sc, tt, c, ev[e], out + ((size_t)c * nE + e) * 9);
}

template <class P>
void run_beam(const cp3::SetConst& sc, const double* ev, int nE, double rho,
              double L, double* out) {
  using D = typename P::D;
  using F = typename P::F;
#pragma omp parallel for
  for (int e = 0; e < nE; ++e) {
    F prob[9];
    cp3::beam_point<P>(sc, (D)ev[e], (D)(rho * cp3::kBeamYe), (D)L, prob);
    for (int k = 0; k < 9; ++k) out[(size_t)e * 9 + k] = (double)prob[k];
  }
}

// ---------------------------------------------------------------- counting

// Operation counters: add/sub/mul/div/sqrt = 1 FLOP each (a fused
// multiply-add appears as its mul and add, 2 FLOPs); transcendental calls are
// counted separately and not as FLOPs.
struct OpCount {
  double f64 = 0, f32 = 0;
  double sincos64 = 0, sincos32 = 0, acos64 = 0, acos32 = 0;
  double sqrt64 = 0, sqrt32 = 0, rint64 = 0;
};
static thread_local OpCount g_ops;

template <class T>
inline void count_flop() {
  if (sizeof(T) == 8)
    g_ops.f64 += 1;
  else
    g_ops.f32 += 1;
}

template <class T>
struct Cnt {
  T v;
  Cnt() = default;
  constexpr Cnt(double x) : v((T)x) {}
  template <class U>
  explicit Cnt(const Cnt<U>& o) : v((T)o.v) {}
  explicit operator double() const { return (double)v; }
  explicit operator float() const { return (float)v; }
  Cnt operator-() const { return Cnt((double)-v); }
  Cnt& operator+=(const Cnt& o) { count_flop<T>(); v += o.v; return *this; }
  Cnt& operator-=(const Cnt& o) { count_flop<T>(); v -= o.v; return *this; }
  Cnt& operator*=(const Cnt& o) { count_flop<T>(); v *= o.v; return *this; }
};
template <class T> Cnt<T> mk(T x) { Cnt<T> c; c.v = x; return c; }
template <class T> Cnt<T> operator+(Cnt<T> a, Cnt<T> b) { count_flop<T>(); return mk<T>(a.v + b.v); }
template <class T> Cnt<T> operator-(Cnt<T> a, Cnt<T> b) { count_flop<T>(); return mk<T>(a.v - b.v); }
template <class T> Cnt<T> operator*(Cnt<T> a, Cnt<T> b) { count_flop<T>(); return mk<T>(a.v * b.v); }
template <class T> Cnt<T> operator/(Cnt<T> a, Cnt<T> b) { count_flop<T>(); return mk<T>(a.v / b.v); }
template <class T> bool operator<(Cnt<T> a, Cnt<T> b) { return a.v < b.v; }
template <class T> bool operator>(Cnt<T> a, Cnt<T> b) { return a.v > b.v; }
template <class T> bool operator<=(Cnt<T> a, Cnt<T> b) { return a.v <= b.v; }
template <class T> bool operator>=(Cnt<T> a, Cnt<T> b) { return a.v >= b.v; }

// Counted overloads, found by argument-dependent lookup from the templates
// of cp3_math.hpp.
inline Cnt<double> m_sqrt(Cnt<double> x) { g_ops.sqrt64 += 1; g_ops.f64 += 1; return mk<double>(std::sqrt(x.v)); }
inline Cnt<float> m_sqrt(Cnt<float> x) { g_ops.sqrt32 += 1; g_ops.f32 += 1; return mk<float>(std::sqrt(x.v)); }
inline Cnt<double> m_rsqrt(Cnt<double> x) { g_ops.sqrt64 += 1; g_ops.f64 += 2; return mk<double>(1.0 / std::sqrt(x.v)); }
inline Cnt<float> m_rsqrt(Cnt<float> x) { g_ops.sqrt32 += 1; g_ops.f32 += 2; return mk<float>(1.0f / std::sqrt(x.v)); }
inline Cnt<double> m_recip(Cnt<double> x) { g_ops.f64 += 1; return mk<double>(1.0 / x.v); }
inline Cnt<float> m_recip(Cnt<float> x) { g_ops.f32 += 1; return mk<float>(1.0f / x.v); }
inline Cnt<float> m_acos(Cnt<float> x) { g_ops.acos32 += 1; return mk<float>(std::acos(x.v)); }
inline Cnt<double> m_fabs(Cnt<double> x) { return mk<double>(std::fabs(x.v)); }
inline Cnt<float> m_fabs(Cnt<float> x) { return mk<float>(std::fabs(x.v)); }
inline Cnt<double> m_rint(Cnt<double> x) { g_ops.rint64 += 1; return mk<double>(std::rint(x.v)); }
inline Cnt<float> m_rint(Cnt<float> x) { return mk<float>(std::rint(x.v)); }
// The layer phase polynomial e^{2 pi i f} - 1 is counted as one sincos call.
template <class T>
inline void m_cispi_m1(Cnt<T> f, Cnt<T>* wr, Cnt<T>* wi) {
  if (sizeof(T) == 8)
    g_ops.sincos64 += 1;
  else
    g_ops.sincos32 += 1;
  T r, i;
  cp3::m_cispi_m1(f.v, &r, &i);
  wr->v = r;
  wi->v = i;
}
inline void m_sincos(Cnt<float> x, Cnt<float>* s, Cnt<float>* c) {
  g_ops.sincos32 += 1;
  s->v = std::sin(x.v);
  c->v = std::cos(x.v);
}
// The FP64 third angle's FP32 seed: an FP32 acos, one FP32 multiply and an
// FP32 sincos; the Newton steps after it are counted as they run.
inline void m_third_angle_seed(Cnt<double> x, Cnt<double>* c, Cnt<double>* s) {
  g_ops.acos32 += 1;
  g_ops.sincos32 += 1;
  g_ops.f32 += 1;
  cp3::m_third_angle_seed(x.v, &c->v, &s->v);
}
// The FP32 form's third angle: acos and sincos, counted by their overloads.
inline void third_angle(Cnt<float> x, Cnt<float>* c, Cnt<float>* s) {
  m_sincos(m_acos(x) * Cnt<float>(1.0 / 3.0), s, c);
}

}  // namespace

namespace cp3 {
struct CountF64 {
  using D = Cnt<double>;
  using F = Cnt<double>;
  static constexpr bool kRateAcc = PolicyF64::kRateAcc;
};
struct CountF32 {
  using D = Cnt<float>;
  using F = Cnt<float>;
  static constexpr bool kRateAcc = PolicyF32::kRateAcc;
};
struct CountMixed {
  using D = Cnt<double>;
  using F = Cnt<float>;
  static constexpr bool kRateAcc = PolicyMixed::kRateAcc;
};

}  // namespace cp3

namespace {

// Counts one call of each kernel stage; out[stage*9 + k] with stages
// 0 solve_slab, 1 traj_middle, 2 traj_inner, 3 traj_finish (with traj_init),
// 4 beam_point, and k = f64, f32, sincos64, sincos32, acos64, acos32,
// sqrt64, sqrt32, rint64.
template <class P>
void stage_ops(double* out) {
  using D = typename P::D;
  using F = typename P::F;
  const cp3::SetConst sc = cp3::make_set(0.58725, 0.14887, 0.83146, -1.6, 1, 7.53e-5,
                                         2.494e-3, 0);
  auto take = [&](int stage) {
    const OpCount& o = g_ops;
    const double v[9] = {o.f64, o.f32, o.sincos64, o.sincos32, o.acos64,
                         o.acos32, o.sqrt64, o.sqrt32, o.rint64};
    for (int k = 0; k < 9; ++k) out[stage * 9 + k] = v[k];
    g_ops = OpCount();
  };
  g_ops = OpCount();
  const D E = D(2.0), inv = D(0.5);
  cp3::Slab<P> s;
  cp3::solve_slab<P>(sc, E, inv, D(4.0 * 0.494), s);
  take(0);
  cp3::Traj<P> t;
  cp3::traj_init<P>(t, false);
  g_ops = OpCount();
  cp3::traj_middle<P>(t, s, D(700.0));
  take(1);
  cp3::traj_inner<P>(t, s, D(1500.0));
  take(2);
  cp3::Slab<P> vac;
  cp3::solve_slab<P>(sc, E, inv, D(0.0), vac);
  g_ops = OpCount();
  F prob[9];
  cp3::traj_finish<P>(t, vac, D(80.0), sc, prob);
  take(3);
  cp3::beam_point<P>(sc, E, D(1.3), D(1300.0), prob);
  take(4);
}

}  // namespace

extern "C" {

// Operation counts of the kernel stages (see stage_ops), 45 doubles.
void cp3m_stage_ops(int precision, double* out) {
  if (precision == 0)
    stage_ops<cp3::CountF64>(out);
  else if (precision == 1)
    stage_ops<cp3::CountF32>(out);
  else
    stage_ops<cp3::CountMixed>(out);
}

// The layer phase function w = e^{2 pi i f} - 1 (m_cispi_m1) at n points,
// in FP64 (precision 0) or FP32 (otherwise).
void cp3m_cispi_m1(int precision, const double* f, int n, double* wr,
                   double* wi) {
  for (int k = 0; k < n; ++k) {
    if (precision == 0) {
      cp3::m_cispi_m1(f[k], &wr[k], &wi[k]);
    } else {
      float r, i;
      cp3::m_cispi_m1((float)f[k], &r, &i);
      wr[k] = r;
      wi[k] = i;
    }
  }
}

// cos and sin of the cubic's third angle acos(x) / 3 (third_angle) at n
// points, in FP64 (precision 0: FP32 seed and Newton steps for the cube root)
// or FP32 (otherwise: acos and sincos).
void cp3m_third_angle(int precision, const double* x, int n, double* c, double* s) {
  for (int k = 0; k < n; ++k) {
    if (precision == 0) {
      cp3::third_angle(x[k], &c[k], &s[k]);
    } else {
      float cf, sf;
      cp3::third_angle((float)x[k], &cf, &sf);
      c[k] = cf;
      s[k] = sf;
    }
  }
}

// A density file as the product loads it (setDensityFromFile): rows stored
// descending, 5 values per row (radius, rho or a, b, c, Ye; b = c = 0 for a
// constant model); *poly = 1 for a polynomial model.  Returns -1 with the
// message in err on an unsupported or unreadable file, -2 if cap is short.
int cp3m_load_earth_file(const char* path, double* rows, int cap, int* nrows, int* poly,
                         char* err, int errlen) {
  try {
    cp3::EarthModel em;
    cp3::earth_load_file(em, path);
    const size_t n = em.radii.size();
    if ((int)n > cap) return -2;
    for (size_t i = 0; i < n; ++i) {
      double* r = rows + 5 * i;
      r[0] = em.radii[i];
      r[1] = em.poly ? em.as[i] : em.rhos[i];
      r[2] = em.poly ? em.bs[i] : 0.0;
      r[3] = em.poly ? em.cs[i] : 0.0;
      r[4] = em.yps[i];
    }
    *nrows = (int)n;
    *poly = em.poly ? 1 : 0;
  } catch (const std::exception& ex) {
    put_err(err, errlen, ex.what());
    return -1;
  }
  return 0;
}

// Shell counts per cosine (the M of the trajectory table) for a model.
int cp3m_max_layers(const double* cosv, int nC, const double* radii,
                    const double* rhos, const double* yps, int nrows,
                    int32_t* M_out) {
  try {
    cp3::EarthModel em;
    cp3::earth_set(em, std::vector<double>(radii, radii + nrows),
                   std::vector<double>(rhos, rhos + nrows),
                   std::vector<double>(yps, yps + nrows));
    for (int c = 0; c < nC; ++c) M_out[c] = cp3::earth_max_layer(em, cosv[c]);
  } catch (...) {
    return -1;
  }
  return 0;
}

// Atmospheric probabilities [cos][E][9] through the kernel decomposition.
// precision: 0 = FP64, 1 = FP32, 2 = mixed.
int cp3m_atmos(int precision, int nutype, const double* cosv, int nC,
               const double* ev, int nE, const double* radii,
               const double* rhos, const double* yps, int nrows, double th12,
               double th13, double th23, double dcp, int kNuType, double dm12,
               double dm23, double h_km, double* out, char* err, int errlen) {
  try {
    cp3::EarthModel em;
    cp3::earth_set(em, std::vector<double>(radii, radii + nrows),
                   std::vector<double>(rhos, rhos + nrows),
                   std::vector<double>(yps, yps + nrows));
    cp3::TrajTable tt;
    cp3::build_traj_table(tt, em, cosv, nC, h_km * 100000.0);
    const cp3::SetConst sc =
        cp3::make_set(th12, th13, th23, dcp, kNuType, dm12, dm23, nutype);
    if (precision == 0)
      run_atmos<cp3::PolicyF64>(sc, tt, ev, nE, out);
    else if (precision == 1)
      run_atmos<cp3::PolicyF32>(sc, tt, ev, nE, out);
    else
      run_atmos<cp3::PolicyMixed>(sc, tt, ev, nE, out);
  } catch (const std::exception& ex) {
    put_err(err, errlen, ex.what());
    return -1;
  }
  return 0;
}

// Atmospheric probabilities with production-height averaging: nb bins,
// prob[2][3][nE][nC][nb] (setProductionHeightList's layout) and the nb + 1
// bin edges in km.
int cp3m_atmos_pha(int precision, int nutype, const double* cosv, int nC,
                   const double* ev, int nE, const double* radii, const double* rhos,
                   const double* yps, int nrows, double th12, double th13, double th23,
                   double dcp, int kNuType, double dm12, double dm23, double h_km,
                   int nb, const double* prob, const double* edges_km, double* out,
                   char* err, int errlen) {
  try {
    cp3::EarthModel em;
    cp3::earth_set(em, std::vector<double>(radii, radii + nrows),
                   std::vector<double>(rhos, rhos + nrows),
                   std::vector<double>(yps, yps + nrows));
    cp3::TrajTable tt;
    cp3::build_traj_table(tt, em, cosv, nC, h_km * 100000.0);
    std::vector<double> hmw;
    cp3::build_height_bins(hmw, cosv, nC, h_km * 100000.0,
                           std::vector<double>(edges_km, edges_km + nb + 1));
    const cp3::SetConst sc =
        cp3::make_set(th12, th13, th23, dcp, kNuType, dm12, dm23, nutype);
    if (precision == 0)
      run_atmos_avg<cp3::PolicyF64>(sc, tt, ev, nE, nutype, nb, prob, hmw, out);
    else if (precision == 1)
      run_atmos_avg<cp3::PolicyF32>(sc, tt, ev, nE, nutype, nb, prob, hmw, out);
    else
      run_atmos_avg<cp3::PolicyMixed>(sc, tt, ev, nE, nutype, nb, prob, hmw, out);
  } catch (const std::exception& ex) {
    put_err(err, errlen, ex.what());
    return -1;
  }
  return 0;
}

// The same through a polynomial model (setDensity(radii, a, b, c, yps)).
int cp3m_atmos_poly(int precision, int nutype, const double* cosv, int nC,
                    const double* ev, int nE, const double* radii, const double* pa,
                    const double* pb, const double* pc, const double* yps, int nrows,
                    double th12, double th13, double th23, double dcp, int kNuType,
                    double dm12, double dm23, double h_km, double* out, char* err,
                    int errlen) {
  try {
    auto v = [&](const double* x) { return std::vector<double>(x, x + nrows); };
    cp3::EarthModel em;
    cp3::earth_set_poly(em, v(radii), v(pa), v(pb), v(pc), v(yps));
    cp3::TrajTable tt;
    cp3::build_traj_table(tt, em, cosv, nC, h_km * 100000.0);
    const cp3::SetConst sc =
        cp3::make_set(th12, th13, th23, dcp, kNuType, dm12, dm23, nutype);
    if (precision == 0)
      run_atmos<cp3::PolicyF64>(sc, tt, ev, nE, out);
    else if (precision == 1)
      run_atmos<cp3::PolicyF32>(sc, tt, ev, nE, out);
    else
      run_atmos<cp3::PolicyMixed>(sc, tt, ev, nE, out);
  } catch (const std::exception& ex) {
    put_err(err, errlen, ex.what());
    return -1;
  }
  return 0;
}

// The per-(cosine, layer) densities of a polynomial model's trajectory table,
// out[c * nrows + i] (0 past the cosine's innermost layer).
int cp3m_poly_rye(const double* cosv, int nC, const double* radii, const double* pa,
                  const double* pb, const double* pc, const double* yps, int nrows,
                  double h_km, double* out) {
  try {
    auto v = [&](const double* x) { return std::vector<double>(x, x + nrows); };
    cp3::EarthModel em;
    cp3::earth_set_poly(em, v(radii), v(pa), v(pb), v(pc), v(yps));
    cp3::TrajTable tt;
    cp3::build_traj_table(tt, em, cosv, nC, h_km * 100000.0);
    if (tt.rye_cs != tt.S) return -1;
    for (size_t k = 0; k < tt.rye.size(); ++k) out[k] = tt.rye[k];
  } catch (...) {
    return -1;
  }
  return 0;
}

// Beam probabilities [E][9].
int cp3m_beam(int precision, int nutype, const double* ev, int nE, double rho,
              double L_km, double th12, double th13, double th23, double dcp,
              int kNuType, double dm12, double dm23, double* out, char* err,
              int errlen) {
  try {
    const cp3::SetConst sc =
        cp3::make_set(th12, th13, th23, dcp, kNuType, dm12, dm23, nutype);
    if (precision == 0)
      run_beam<cp3::PolicyF64>(sc, ev, nE, rho, L_km, out);
    else if (precision == 1)
      run_beam<cp3::PolicyF32>(sc, ev, nE, rho, L_km, out);
    else
      run_beam<cp3::PolicyMixed>(sc, ev, nE, rho, L_km, out);
  } catch (const std::exception& ex) {
    put_err(err, errlen, ex.what());
    return -1;
  }
  return 0;
}

}  // extern "C"
