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
  const D inv = D(1) / e;
  const int M = tt.M[(size_t)c];
  cp3::Traj<P> t;
  cp3::traj_init<P>(t, M == 0);
  for (int i = 0; i < M; ++i) {
    cp3::Slab<P> s;
    cp3::solve_slab<P>(sc, e, inv, (D)tt.rye[(size_t)i], s);
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

}  // namespace

extern "C" {

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
