// Three-flavour Barger propagation arithmetic shared by the HIP kernels and
// the host build of the same code (libcp3model.so, used for numerics tests and
// operation counting).  Everything here is templated on a precision policy P
// that fixes two scalar types:
//   P::D - eigen-solve, phase rates, phases and their reduction;
//   P::F - eigenvectors, the per-layer amplitude updates and the output.
//
// Method (Barger et al., as in physics.hpp:267-629 and 789-1150):
//   * per (energy, shell): the matter eigenvalues from the trigonometric
//     solution of the characteristic cubic with the reference's coefficients
//     (physics.hpp:283-323), ordered by continuity to vacuum
//     (physics.hpp:198-252, 327-330), and the Sylvester projectors
//     prod_{m != k}(H - l_m) / prod_{m != k}(l_k - l_m) (physics.hpp:394-427);
//   * per (cosine, shell): the layer matrix sum_k C_k e^{-i 2.534 l_k L / E}
//     (physics.hpp:439-450, math.hpp:57-75) and the ordered product with the
//     mirror for the outgoing half of the path (physics.hpp:956-990);
//   * P(a->b) = |A[b][a]|^2 (physics.hpp:1102-1144).
//
// Exact rewrites used to cut the arithmetic (all algebraic identities):
//   1. With the PDG matrix U = W O G^+ (W = R23 diag(1,1,e^{i d}),
//      O = R13 R12, G = diag(1,1,e^{i d})), the flavour Hamiltonian is
//      W H' W^+ with H' = O diag(DM(k,0)) O^T - fac e_0 e_0^T real symmetric.
//      Layer matrices in that frame are complex symmetric, so the mirrored
//      product T_1 ... T_{M-1} equals (T_{M-1} ... T_1)^T and the path is
//      A' = Y^T T_M Y T_0 with one 3x3 product chain Y instead of two.
//   2. The projectors of the real symmetric H' are rank one, v_k v_k^T, so a
//      layer acts as y -> y + sum_k (e^{i psi_k} - 1) v_k (v_k . y) after
//      removing the common phase e^{i phi_ref} (a global phase of the layer,
//      which no |A|^2 sees).  The reference index is one of the two closest
//      eigenvalues so that an ill-conditioned eigenvector of a near-degenerate
//      pair only ever multiplies a small e^{i psi} - 1.
//   3. Y is unitary, so only two columns are carried; the third is
//      det(Y) conj(y_0 x y_1), with det(Y) = e^{i sum psi} accumulated.
//   4. cos(t0 -+ 2pi/3) = -cos(t0)/2 +- (sqrt(3)/2) sin(t0): one (cos t0,
//      sin t0) pair for the three roots.
// Phases: every layer phase is carried in revolutions, reduced to
// [-1/2, 1/2] in P::D, and e^{i psi} - 1 is evaluated in P::F from minimax
// polynomials of sin and cos of the half angle (no quadrant logic).
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define CP3_HD __host__ __device__ __forceinline__
#else
#define CP3_HD inline
#endif

#include <math.h>
#include <stdint.h>

namespace cp3 {

// constants.hpp:28-40 and physics.hpp:444.
constexpr double kGf2 = 1.52588e-4;
constexpr double kLoE = 2.534;
constexpr double kREarthKm = 6371.0;
constexpr double kKm2Cm = 1.0e5;
constexpr double kREarthCm = kREarthKm * kKm2Cm;
constexpr double kBeamYe = 0.5;
constexpr double kSqrt3Half = 0.86602540378443864676;
constexpr double kInvTwoPi = 0.15915494309189533577;

// Per (parameter set, neutrino type) constants, built on the host from the
// reference's setter arguments.  Plain old data: copied to the device as is.
struct SetConst {
  double g0;          // |U_e1|^2 as physics.hpp:294 forms it
  double q1, q2;      // 1 - |U_e2|^2, 1 - |U_e3|^2 as physics.hpp:286-291
  double dm01, dm02;  // DM(0,1), DM(0,2) (propagator.hpp:200-208)
  double dm00;        // DM(0,0)
  double hv[6];       // O diag(DM(k,0)) O^T, upper triangle 00 01 02 11 12 22
  double sgn;         // sign of the matter term: -1 neutrino, +1 antineutrino
  double c23, s23;    // W = R23 diag(1, 1, e^{i d})
  double cd, sd;      // cos d, sin d (d already negated for kNuType < 0)
  // The cubic's roots in the reference's vacuum order (physics.hpp:232-246,
  // 327-330): eigenvalue k is amp (ra[k] cos t0 + rb[k] sin t0) + shift, with
  // (ra, rb) = (1, 0), (-1/2, sqrt(3)/2) or (-1/2, -sqrt(3)/2) for roots 0-2.
  double ra[3], rb[3];
  // The parts of H'^2 that do not depend on the matter term (H' differs from
  // O diag(DM(k,0)) O^T only in H'_00): hh00 + H'_00^2, hh01 + H'_00 hv[1],
  // hh02 + H'_00 hv[2], and hh11, hh12, hh22 as they are.
  double hh[6];
  int32_t order[3];   // vacuum ordering (physics.hpp:232-246), host side
  int32_t pad;
};

// ---------------------------------------------------------------- policies

// kRateAcc: the determinant phase advances once per layer by the summed rate
// times the length (one FP64 multiply-add, where FP64 instructions issue at
// about half the FP32 rate) instead of summing the two reduced phases.
struct PolicyF64 {
  using D = double;
  using F = double;
  static constexpr bool kRateAcc = false;
};
struct PolicyF32 {
  using D = float;
  using F = float;
  static constexpr bool kRateAcc = false;
};
struct PolicyMixed {
  using D = double;
  using F = float;
  static constexpr bool kRateAcc = true;
};

// Scalar wrappers, one overload per type, so the host build can count
// operations by precision.
#if defined(__HIP_DEVICE_COMPILE__)
// FP64 square root for the eigen-solve, whose arguments are 0 or normal: the
// arithmetic of the compiled sqrt(double) (v_rsq_f64, then the coupled
// refinement of s = sqrt(x) and h = 1/(2 sqrt(x)) with two residual
// corrections), in the same order, without its rescaling of arguments below
// 2^-767 and its special-case selects (7 of its 17 instructions).  Over 4M
// sampled normals on gfx950 it gives the compiled sqrt's bits for x >= 2^-900
// and is within 1 ulp of them below.  rsq is taken of max(x, DBL_MIN), so
// x = 0 gives s = 0 * y = 0 throughout; subnormal x are outside its domain.
CP3_HD double m_sqrt(double x) {
  const double y = __builtin_amdgcn_rsq(fmax(x, 2.2250738585072014e-308));
  double s = x * y;
  double h = y * 0.5;
  const double r = fma(-h, s, 0.5);
  s = fma(s, r, s);
  h = fma(h, r, h);
  double e = fma(-s, s, x);
  s = fma(e, h, s);
  e = fma(-s, s, x);
  return fma(e, h, s);
}
// 1/x for a normal x > 0 (an energy in GeV): v_rcp_f64 and two Newton steps,
// within an ulp of the division (5 instructions instead of the 11 of the
// IEEE division sequence).
CP3_HD double m_recip(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}
#else
CP3_HD double m_sqrt(double x) { return sqrt(x); }
CP3_HD double m_recip(double x) { return 1.0 / x; }
#endif
CP3_HD float m_sqrt(float x) { return sqrtf(x); }
CP3_HD float m_recip(float x) { return 1.0f / x; }
CP3_HD float m_acos(float x) { return acosf(x); }
CP3_HD double m_fabs(double x) { return fabs(x); }
CP3_HD float m_fabs(float x) { return fabsf(x); }
CP3_HD double m_rint(double x) { return rint(x); }
CP3_HD float m_rint(float x) { return rintf(x); }

#if defined(__HIP_DEVICE_COMPILE__)
CP3_HD double m_rsqrt(double x) { return rsqrt(x); }
CP3_HD float m_rsqrt(float x) { return rsqrtf(x); }
#else
CP3_HD double m_rsqrt(double x) { return 1.0 / sqrt(x); }
CP3_HD float m_rsqrt(float x) { return 1.0f / sqrtf(x); }
#endif
#if defined(__HIP_DEVICE_COMPILE__)
// FP32: reduce in revolutions and use the hardware v_sin_f32 / v_cos_f32.
CP3_HD void m_sincos(float x, float* s, float* c) {
  float f = x * 0.15915494309189535f;
  f = f - rintf(f);
  *s = __builtin_amdgcn_sinf(f);
  *c = __builtin_amdgcn_cosf(f);
}
#else
CP3_HD void m_sincos(float x, float* s, float* c) {
  *s = sinf(x);
  *c = cosf(x);
}
#endif

// w = e^{2 pi i f} - 1 for |f| <= 1/2, with no argument reduction and no
// quadrant: a = sqrt(2) sin(pi f) and b = sqrt(2) cos(pi f) from minimax
// polynomials in f (degrees 9 and 10 in FP32, 15 and 16 in FP64; the
// coefficients and their errors, at most 2.1e-7 and 4.5e-16 with rounding,
// come from fit_cispi.py), then w = (cos 2 pi f - 1, sin 2 pi f)
// = (-2 sin^2, 2 sin cos)(pi f) = (-a^2, a b), free of cancellation for
// small f.
// The two polynomials: pa = sqrt(2) sin(pi f) / f and b = sqrt(2) cos(pi f).
CP3_HD void m_half_sc(float f, float* pa_out, float* b_out) {
  const float z = f * f;
  float pa = 0.10920575261116028f;
  pa = pa * z + -0.845763623714447f;
  pa = pa * z + 3.6062889099121094f;
  pa = pa * z + -7.308241367340088f;
  pa = pa * z + 4.442883014678955f;
  float b = -0.034502167254686356f;
  b = b * z + 0.3322518765926361f;
  b = b * z + -1.8882750272750854f;
  b = b * z + 5.73988151550293f;
  b = b * z + -6.9788641929626465f;
  b = b * z + 1.4142135381698608f;
  *pa_out = pa;
  *b_out = b;
}
CP3_HD void m_half_sc(double f, double* pa_out, double* b_out) {
  const double z = f * f;
  double pa = -2.9819668671076778e-05;
  pa = fma(pa, z, 0.0006589415933150775);
  pa = fma(pa, z, -0.010423245350225985);
  pa = fma(pa, z, 0.11617181136154904);
  pa = fma(pa, z, -0.8474880236479588);
  pa = fma(pa, z, 3.606476571424355);
  pa = fma(pa, z, -7.308249499994612);
  pa = fma(pa, z, 4.442882938158362);
  double b = 5.8684089087724815e-06;
  b = fma(b, z, -0.00014787947669297268);
  b = fma(b, z, 0.0027288046656365815);
  b = fma(b, z, -0.03649645205784972);
  b = fma(b, z, 0.33280776877138185);
  b = fma(b, z, -1.8883467170302788);
  b = fma(b, z, 5.739885734946355);
  b = fma(b, z, -6.978864199638876);
  b = fma(b, z, 1.4142135623730951);
  *pa_out = pa;
  *b_out = b;
}
template <class T>
CP3_HD void m_cispi_m1(T f, T* wr, T* wi) {
  T pa, b;
  m_half_sc(f, &pa, &b);
  const T a = pa * f;
  *wr = -(a * a);
  *wi = a * b;
}

// sin(2 pi q) / (2 pi q) for |q| <= 1/2, from the same polynomials with no
// division: sin(2 pi q) = a b = pa q b, so the quotient is pa b / (2 pi)
// (exactly 1 at q = 0 up to the polynomials' rounding).
template <class T>
CP3_HD T m_sinc2pi_half(T q) {
  T pa, b;
  m_half_sc(q, &pa, &b);
  return pa * b * T(kInvTwoPi);
}

// Layer phases in revolutions: om is in revolutions per km, the phase
// q = om * L is reduced in D to f = q - rint(q) in [-1/2, 1/2], summed into
// acc (det Y = e^{2 pi i acc}), and e^{2 pi i f} - 1 is evaluated in F.
template <class D, class F>
struct PhaseRev {
  static constexpr double kOmScale = kInvTwoPi;
  CP3_HD static void eval(D om, D L, D& acc, F& wr, F& wi) {
    const D q = om * L;
    const D f = q - m_rint(q);
    acc += f;
    // an FP32 sum is kept in [-1/2, 1/2] so that it keeps its resolution
    if (sizeof(D) == 4) acc -= m_rint(acc);
    m_cispi_m1(F(f), &wr, &wi);
  }
  // the same without accumulating the phase
  CP3_HD static void eval(D om, D L, F& wr, F& wi) {
    const D q = om * L;
    m_cispi_m1(F(q - m_rint(q)), &wr, &wi);
  }
  // e^{2 pi i acc}
  CP3_HD static void unit(D acc, F& cr, F& ci) {
    F wr;
    m_cispi_m1(F(acc - m_rint(acc)), &wr, &ci);
    cr = wr + F(1);
  }
};

template <class P>
struct Phase : PhaseRev<typename P::D, typename P::F> {};

// ------------------------------------------------------- per (E, shell)

// Eigen-solution of one homogeneous slab: two unit eigenvectors and their
// phase rates relative to the reference eigenvalue (revolutions per km).
template <class P>
struct Slab {
  typename P::F va[3], vb[3];
  typename P::D oma, omb;
};

// Unit eigenvector of the real symmetric matrix H for the eigenvalue lk from
// the Sylvester numerator P = (H - lp)(H - lq) = H^2 - (lp + lq) H + lp lq,
// whose columns are all multiples of v_k (physics.hpp:394-427 builds the same
// product): the column j with the largest P_jj dk = dk^2 v_j^2,
// dk = (lk - lp)(lk - lq), scaled by 1/sqrt(P_jj dk).  H and H2 = H^2 as
// upper triangles 00 01 02 11 12 22.
// kFNorm (beam path only): an F eigenvector of the mixed form is normalised
// in F; the column itself, whose entries cancel, is still formed in D.
template <class D, class F, bool kFNorm = false>
CP3_HD void eigvec(const D H[6], const D H2[6], D lk, D lp, D lq, F v[3]) {
  const D sp = lp + lq, pp = lp * lq;
  const D p00 = (H2[0] - sp * H[0]) + pp;
  const D p11 = (H2[3] - sp * H[3]) + pp;
  const D p22 = (H2[5] - sp * H[5]) + pp;
  const D p01 = H2[1] - sp * H[1];
  const D p02 = H2[2] - sp * H[2];
  const D p12 = H2[4] - sp * H[4];
  const D dk = (lk - lp) * (lk - lq);
  const D s0 = p00 * dk, s1 = p11 * dk, s2 = p22 * dk;
  const bool j0 = s0 >= s1 && s0 >= s2;
  const bool j1 = !j0 && s1 >= s2;
  const D c0 = j0 ? p00 : (j1 ? p01 : p02);
  const D c1 = j0 ? p01 : (j1 ? p11 : p12);
  const D c2 = j0 ? p02 : (j1 ? p12 : p22);
  const D sj = j0 ? s0 : (j1 ? s1 : s2);
  if (kFNorm && sizeof(F) < sizeof(D)) {
    const F inv = m_rsqrt((F)sj);
    v[0] = (F)c0 * inv;
    v[1] = (F)c1 * inv;
    v[2] = (F)c2 * inv;
  } else {
    const D inv = m_rsqrt(sj);
    v[0] = (F)(c0 * inv);
    v[1] = (F)(c1 * inv);
    v[2] = (F)(c2 * inv);
  }
}

// cos and sin of the cubic's third angle acos(x) / 3 (physics.hpp:310).
// FP32: acos and sincos.  FP64: e^{i acos(x) / 3} is the principal cube root
// of z = x + i sqrt(1 - x^2), found by Newton's method for u^3 = z,
//   u <- u + (z / u^2 - u) / 3,  z / u^2 = z conj(u)^2 / |u|^4,
// from a seed of FP32 accuracy.  1 / |u|^4 is taken as (2 - |u|^2)^2, which
// differs from it by O((|u|^2 - 1)^2) and keeps the convergence quadratic
// (at u = u* (1 + e) the step leaves u* + O(e^2)), so two steps reach FP64
// rounding (within one ulp of the exact values over [-1, 1]).  On gfx950
// this took a slab solve from 439 VALU instructions, with an FP64 acos
// (about 80, with a divergent branch) and sincos (about 65), to 357.
// 1 - x^2 is formed as (1 - x)(1 + x), accurate to a few ulp up to |x| = 1.
CP3_HD void m_third_angle_seed(double x, double* c, double* s) {
  const float t = acosf((float)x) * (1.0f / 3.0f);
#if defined(__HIP_DEVICE_COMPILE__)
  const float r = t * 0.15915494309189535f;  // revolutions
  *c = (double)__builtin_amdgcn_cosf(r);
  *s = (double)__builtin_amdgcn_sinf(r);
#else
  *c = (double)cosf(t);
  *s = (double)sinf(t);
#endif
}

CP3_HD void third_angle(float x, float* c, float* s) {
  m_sincos(m_acos(x) * (1.0f / 3.0f), s, c);
}

template <class D>
CP3_HD void third_angle(D x, D* c, D* s) {
  const D w = m_sqrt((D(1) - x) * (D(1) + x));
  D uc, us;
  m_third_angle_seed(x, &uc, &us);
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const D c2 = uc * uc, s2 = us * us, cs = uc * us;
    const D t = D(2) - (c2 + s2);
    const D q = t * t;
    const D a = c2 - s2;  // conj(u)^2 = a - 2i cs
    const D pr = x * a + D(2) * (w * cs);
    const D pi = w * a - D(2) * (x * cs);
    uc = uc + (pr * q - uc) * D(1.0 / 3.0);
    us = us + (pi * q - us) * D(1.0 / 3.0);
  }
  *c = uc;
  *s = us;
}

// Matter eigenvalues in the reference's order: the trigonometric roots of
// the cubic with physics.hpp:283-323's coefficients, in ORDER's sequence
// through the per-set root coefficients ra, rb.
template <class D>
CP3_HD void matter_eigs(const SetConst& sc, D fac, D lam[3]) {
  const D d01 = D(sc.dm01), d02 = D(sc.dm02);
  const D alpha = (fac + d01) + d02;
  const D beta = d01 * d02 + fac * (d01 * D(sc.q1) + d02 * D(sc.q2));
  const D gamma = ((fac * d01) * d02) * D(sc.g0);
  D tmp = alpha * alpha - D(3) * beta;
  tmp = tmp < D(0) ? D(0) : tmp;
  const D st = m_sqrt(tmp);
  // the quotient by a reciprocal and a product (no IEEE division sequence);
  // a denominator clamped at the smallest normal keeps a triple root (tmp
  // rounded to 0) finite: a nonzero numerator then clamps to +-1 below, as
  // its quotient by 0 would, and a zero one gives 0
  const D den = D(2) * (tmp * st);
  D arg = (D(2) * alpha * alpha * alpha - D(9) * alpha * beta + D(27) * gamma) *
          m_recip(den > D(2.2250738585072014e-308) ? den : D(2.2250738585072014e-308));
  if (m_fabs(arg) > D(1)) arg = arg > D(0) ? D(1) : D(-1);
  D sn, cs;
  third_angle(arg, &cs, &sn);
  const D amp = D(-2.0 / 3.0) * st;
  const D shift = D(sc.dm00) - alpha * D(1.0 / 3.0);
#pragma unroll
  for (int k = 0; k < 3; ++k)
    lam[k] = amp * (D(sc.ra[k]) * cs + D(sc.rb[k]) * sn) + shift;
}

// Solves one slab of electron density rhoYe at energy E (invE = 1/E);
// kFNorm as for eigvec.
template <class P, bool kFNorm = false>
CP3_HD void solve_slab(const SetConst& sc, typename P::D E,
                       typename P::D invE, typename P::D rhoYe, Slab<P>& s) {
  using D = typename P::D;
  const D fac = D(sc.sgn) * ((D(kGf2) * E) * rhoYe);
  D lam[3];
  matter_eigs<D>(sc, fac, lam);
  const D h0 = D(sc.hv[0]) - fac;
  const D H[6] = {h0, D(sc.hv[1]), D(sc.hv[2]), D(sc.hv[3]), D(sc.hv[4]),
                  D(sc.hv[5])};
  const D H2[6] = {h0 * h0 + D(sc.hh[0]), h0 * D(sc.hv[1]) + D(sc.hh[1]),
                   h0 * D(sc.hv[2]) + D(sc.hh[2]), D(sc.hh[3]), D(sc.hh[4]),
                   D(sc.hh[5])};
  // Reference index: one end of the closest pair; `near` is the other end,
  // `far` the remaining eigenvalue.
  const D g01 = m_fabs(lam[0] - lam[1]);
  const D g12 = m_fabs(lam[1] - lam[2]);
  const D g02 = m_fabs(lam[0] - lam[2]);
  const bool p01 = g01 <= g12 && g01 <= g02;
  const bool p12 = !p01 && g12 <= g02;
  const D lr = p01 ? lam[0] : lam[2];
  const D ln = (p01 || p12) ? lam[1] : lam[0];
  const D lf = p01 ? lam[2] : (p12 ? lam[0] : lam[1]);
  eigvec<D, typename P::F, kFNorm>(H, H2, lf, lr, ln, s.va);
  eigvec<D, typename P::F, kFNorm>(H, H2, ln, lr, lf, s.vb);
  const D k = D(-kLoE * Phase<P>::kOmScale) * invE;
  s.oma = k * (lf - lr);
  s.omb = k * (ln - lr);
}

// ---------------------------------------------------- complex helpers

template <class F>
struct C3 {  // complex 3-vector
  F re[3], im[3];
};

// y -> y + w v (v . y) with w complex and v real.
template <class F>
CP3_HD void rank1(const F v[3], F wr, F wi, C3<F>& y) {
  const F tr = v[0] * y.re[0] + v[1] * y.re[1] + v[2] * y.re[2];
  const F ti = v[0] * y.im[0] + v[1] * y.im[1] + v[2] * y.im[2];
  const F ur = wr * tr - wi * ti;
  const F ui = wr * ti + wi * tr;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    y.re[m] += v[m] * ur;
    y.im[m] += v[m] * ui;
  }
}

// ------------------------------------------------ per (cosine, shell)

// Per-trajectory state.  Before the innermost shell: the first two columns
// of Y (y0, y1) and acc = sum of the layer phases (det Y = e^{i acc}).  After
// it: Z = Y^T T_M Y (complex symmetric) packed as y0 = (Z00, Z01, Z02),
// y1 = (Z11, Z12, Z22).
template <class P>
struct Traj {
  C3<typename P::F> y0, y1;
  typename P::D acc;
};

template <class P>
CP3_HD void traj_init(Traj<P>& t, bool no_shells) {
  using F = typename P::F;
  // Y = I (y0 = e0, y1 = e1) or, without shells, Z = I packed.  Written as
  // value selects: conditional stores to different fields would be merged
  // into a store through a selected address, which keeps the state out of
  // registers.
  const F one = F(1), zero = F(0);
  t.y0.re[0] = one;
  t.y0.re[1] = zero;
  t.y0.re[2] = zero;
  t.y1.re[0] = no_shells ? one : zero;
  t.y1.re[1] = no_shells ? zero : one;
  t.y1.re[2] = no_shells ? one : zero;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    t.y0.im[m] = zero;
    t.y1.im[m] = zero;
  }
  t.acc = typename P::D(0);
}

// A middle shell crossed twice (physics.hpp:971-978): Y -> T Y.
template <class P>
CP3_HD void traj_middle(Traj<P>& t, const Slab<P>& s, typename P::D L) {
  using F = typename P::F;
  F wr, wi;
  if (P::kRateAcc) {
    // det Y advances by (oma + omb) L per layer, kept unreduced in D
    Phase<P>::eval(s.oma, L, wr, wi);
    rank1(s.va, wr, wi, t.y0);
    rank1(s.va, wr, wi, t.y1);
    Phase<P>::eval(s.omb, L, wr, wi);
    rank1(s.vb, wr, wi, t.y0);
    rank1(s.vb, wr, wi, t.y1);
    t.acc = t.acc + (s.oma + s.omb) * L;
    return;
  }
  Phase<P>::eval(s.oma, L, t.acc, wr, wi);
  rank1(s.va, wr, wi, t.y0);
  rank1(s.va, wr, wi, t.y1);
  Phase<P>::eval(s.omb, L, t.acc, wr, wi);
  rank1(s.vb, wr, wi, t.y0);
  rank1(s.vb, wr, wi, t.y1);
}

// Bilinear (unconjugated) dot product p . q of complex 3-vectors, as chains
// of single products (every step one fused multiply-add).
template <class F>
CP3_HD void dot_cc(const C3<F>& p, const C3<F>& q, F& r, F& i) {
  r = p.re[0] * q.re[0];
  i = p.re[0] * q.im[0];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    if (m > 0) {
      r = r + p.re[m] * q.re[m];
      i = i + p.re[m] * q.im[m];
    }
    r = r - p.im[m] * q.im[m];
    i = i + p.im[m] * q.re[m];
  }
}

// The innermost shell, crossed once along its full chord
// (physics.hpp:979-983, 785): Z = Y^T T_M Y.  The chord's transfer is the
// square of the half chord's, T_M = T_h T_h, and T_h is symmetric (real
// eigenvectors in this frame), so Z = (T_h Y)^T (T_h Y): one more layer
// update over half the chord, as for a middle shell, then Z = Y^T Y of the
// updated columns (the third, y2 = det(Y) conj(y0 x y1), from the other two).
// This takes about 40 fewer instructions per trajectory than
// Z = Y^T Y + sum_k w_k u_k u_k^T over the full chord, with the same
// property that an error in an ill-conditioned eigenvector of a
// near-degenerate pair only enters through its small phase factor.
template <class P>
CP3_HD void traj_inner(Traj<P>& t, const Slab<P>& s, typename P::D L) {
  using F = typename P::F;
  using D = typename P::D;
  traj_middle<P>(t, s, D(0.5) * L);
  // third column y2 = det(Y) conj(y0 x y1)
  F er, ei;
  Phase<P>::unit(t.acc, er, ei);
  C3<F> y2;
  const C3<F>& a = t.y0;
  const C3<F>& b = t.y1;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const int p = (m + 1) % 3, q = (m + 2) % 3;
    F xr = a.re[p] * b.re[q];
    xr = xr - a.im[p] * b.im[q];
    xr = xr - a.re[q] * b.re[p];
    xr = xr + a.im[q] * b.im[p];
    F xi = a.re[p] * b.im[q];
    xi = xi + a.im[p] * b.re[q];
    xi = xi - a.re[q] * b.im[p];
    xi = xi - a.im[q] * b.re[p];
    // e * conj(x)
    y2.re[m] = er * xr + ei * xi;
    y2.im[m] = ei * xr - er * xi;
  }
  // packed upper triangle 00 01 02 11 12 22
  F zr[6], zi[6];
  dot_cc(t.y0, t.y0, zr[0], zi[0]);
  dot_cc(t.y0, t.y1, zr[1], zi[1]);
  dot_cc(t.y0, y2, zr[2], zi[2]);
  dot_cc(t.y1, t.y1, zr[3], zi[3]);
  dot_cc(t.y1, y2, zr[4], zi[4]);
  dot_cc(y2, y2, zr[5], zi[5]);
  t.y0.re[0] = zr[0]; t.y0.im[0] = zi[0];
  t.y0.re[1] = zr[1]; t.y0.im[1] = zi[1];
  t.y0.re[2] = zr[2]; t.y0.im[2] = zi[2];
  t.y1.re[0] = zr[3]; t.y1.im[0] = zi[3];
  t.y1.re[1] = zr[4]; t.y1.im[1] = zi[4];
  t.y1.re[2] = zr[5]; t.y1.im[2] = zi[5];
}

// Applies the atmosphere (vacuum) segment, rotates back to the flavour basis
// and forms the nine probabilities prob[a*3+b] = P(a->b) = |A[b][a]|^2.
template <class P>
CP3_HD void traj_finish(const Traj<P>& t, const Slab<P>& vac,
                        typename P::D Lvac, const SetConst& sc,
                        typename P::F prob[9]) {
  using F = typename P::F;
  // A = Z T_0 = Z + sum_k w_k (Z v_k) v_k^T
  F ar[3][3], ai[3][3];
  const F zr[3][3] = {{t.y0.re[0], t.y0.re[1], t.y0.re[2]},
                      {t.y0.re[1], t.y1.re[0], t.y1.re[1]},
                      {t.y0.re[2], t.y1.re[1], t.y1.re[2]}};
  const F zi[3][3] = {{t.y0.im[0], t.y0.im[1], t.y0.im[2]},
                      {t.y0.im[1], t.y1.im[0], t.y1.im[1]},
                      {t.y0.im[2], t.y1.im[1], t.y1.im[2]}};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      ar[i][j] = zr[i][j];
      ai[i][j] = zi[i][j];
    }
  for (int k = 0; k < 2; ++k) {
    const F* v = k == 0 ? vac.va : vac.vb;
    F wr, wi;
    Phase<P>::eval(k == 0 ? vac.oma : vac.omb, Lvac, wr, wi);
    for (int i = 0; i < 3; ++i) {
      const F tr = zr[i][0] * v[0] + zr[i][1] * v[1] + zr[i][2] * v[2];
      const F ti = zi[i][0] * v[0] + zi[i][1] * v[1] + zi[i][2] * v[2];
      const F sr = wr * tr - wi * ti;
      const F si = wr * ti + wi * tr;
      for (int j = 0; j < 3; ++j) {
        ar[i][j] += sr * v[j];
        ai[i][j] += si * v[j];
      }
    }
  }
  // G A G^+ with G = diag(1, 1, e^{i d})
  const F cd = F(sc.cd), sd = F(sc.sd);
  for (int j = 0; j < 2; ++j) {
    const F r = ar[2][j], i = ai[2][j];
    ar[2][j] = cd * r - sd * i;
    ai[2][j] = cd * i + sd * r;
    const F r2 = ar[j][2], i2 = ai[j][2];
    ar[j][2] = cd * r2 + sd * i2;
    ai[j][2] = cd * i2 - sd * r2;
  }
  // R23 (.) R23^T
  const F c = F(sc.c23), s = F(sc.s23);
  for (int j = 0; j < 3; ++j) {
    const F r1 = ar[1][j], i1 = ai[1][j], r2 = ar[2][j], i2 = ai[2][j];
    ar[1][j] = c * r1 + s * r2;
    ai[1][j] = c * i1 + s * i2;
    ar[2][j] = c * r2 - s * r1;
    ai[2][j] = c * i2 - s * i1;
  }
  for (int i = 0; i < 3; ++i) {
    const F r1 = ar[i][1], i1 = ai[i][1], r2 = ar[i][2], i2 = ai[i][2];
    ar[i][1] = c * r1 + s * r2;
    ai[i][1] = c * i1 + s * i2;
    ar[i][2] = c * r2 - s * r1;
    ai[i][2] = c * i2 - s * i1;
  }
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      prob[a * 3 + b] = ar[b][a] * ar[b][a] + ai[b][a] * ai[b][a];
}

// Production-height averaging (physics.hpp:997-1067, 1106-1130).  The
// atmosphere segment splits the amplitude by vacuum eigenstate:
// A = sum_k u_k v_k^T with u_k = e^{i psi_k} Z v_k (psi_0 = 0 for the
// reference eigenvector, v_0 = v_a x v_b), so lengthening the path by s turns
// component k by e^{2 pi i om_k s} (om_k in revolutions per km).  A bin of
// path-length shifts [s0, s1] averages the interference of components j < k
// into sinc(pi d hw) e^{2 pi i d hm} with d = om_j - om_k, hm = (s0 + s1) / 2
// and hw = s1 - s0 (zero rates when the atmosphere segment has zero length,
// physics.hpp:960-963), and the bins add up weighted by the bin probability
// of the initial flavour a:
//   P(a->b) = sum_k |q_k|^2 + 2 sum_{j<k} Re(q_j conj(q_k) X_jk(a)),
//   q_k = U_k[b] V_k[a], U_k = R23 G u_k, V_k = R23 conj(G) v_k.
// With every X equal to 1 this is |A[b][a]|^2.  The sinc is sin(x)/x for
// |x| >= 1e-6 and 1 - x^2/6 + x^4/120 below (math.hpp:49-55, Epsilon() of
// constants.hpp:49).  `bins` gives the bin count `n`, shift(bin, hm, hw) in km
// and weight(a, bin).
template <class P, class Bins>
CP3_HD void traj_finish_avg(const Traj<P>& t, const Slab<P>& vac,
                            typename P::D Lvac, const SetConst& sc,
                            const Bins& bins, typename P::F prob[9]) {
  using D = typename P::D;
  using F = typename P::F;
  const F zr[3][3] = {{t.y0.re[0], t.y0.re[1], t.y0.re[2]},
                      {t.y0.re[1], t.y1.re[0], t.y1.re[1]},
                      {t.y0.re[2], t.y1.re[1], t.y1.re[2]}};
  const F zi[3][3] = {{t.y0.im[0], t.y0.im[1], t.y0.im[2]},
                      {t.y0.im[1], t.y1.im[0], t.y1.im[1]},
                      {t.y0.im[2], t.y1.im[1], t.y1.im[2]}};
  F v[3][3];
  for (int m = 0; m < 3; ++m) {
    v[1][m] = vac.va[m];
    v[2][m] = vac.vb[m];
  }
  for (int m = 0; m < 3; ++m) {
    const int p = (m + 1) % 3, q = (m + 2) % 3;
    v[0][m] = vac.va[p] * vac.vb[q] - vac.va[q] * vac.vb[p];
  }
  // u_k = c_k Z v_k, then U_k = R23 G u_k; V_k = R23 conj(G) v_k
  const F cd = F(sc.cd), sd = F(sc.sd), c = F(sc.c23), s = F(sc.s23);
  F ur[3][3], ui[3][3], vr[3][3], vi[3][3];
  for (int k = 0; k < 3; ++k) {
    F cr = F(1), ci = F(0);
    if (k > 0) {
      F wr, wi;
      Phase<P>::eval(k == 1 ? vac.oma : vac.omb, Lvac, wr, wi);
      cr = cr + wr;
      ci = wi;
    }
    F tr[3], ti[3];
    for (int i = 0; i < 3; ++i) {
      const F xr = zr[i][0] * v[k][0] + zr[i][1] * v[k][1] + zr[i][2] * v[k][2];
      const F xi = zi[i][0] * v[k][0] + zi[i][1] * v[k][1] + zi[i][2] * v[k][2];
      tr[i] = cr * xr - ci * xi;
      ti[i] = cr * xi + ci * xr;
    }
    const F g2r = cd * tr[2] - sd * ti[2], g2i = cd * ti[2] + sd * tr[2];
    ur[k][0] = tr[0];
    ui[k][0] = ti[0];
    ur[k][1] = c * tr[1] + s * g2r;
    ui[k][1] = c * ti[1] + s * g2i;
    ur[k][2] = c * g2r - s * tr[1];
    ui[k][2] = c * g2i - s * ti[1];
    const F h2r = v[k][2] * cd, h2i = -(v[k][2] * sd);
    vr[k][0] = v[k][0];
    vi[k][0] = F(0);
    vr[k][1] = c * v[k][1] + s * h2r;
    vi[k][1] = s * h2i;
    vr[k][2] = c * h2r - s * v[k][1];
    vi[k][2] = c * h2i;
  }
  // pairs (0, 1), (0, 2), (1, 2)
  const bool moving = Lvac != D(0);
  const D om1 = moving ? vac.oma : D(0), om2 = moving ? vac.omb : D(0);
  const D d[3] = {-om1, -om2, om1 - om2};
  F xr[3][3], xi[3][3];  // [pair][initial flavour]
  for (int p = 0; p < 3; ++p)
    for (int a = 0; a < 3; ++a) {
      xr[p][a] = F(0);
      xi[p][a] = F(0);
    }
  for (int bin = 0; bin < bins.n; ++bin) {
    D hm, hw;
    bins.shift(bin, hm, hw);
    // e^{2 pi i d hm} per pair; the third is the first two's quotient
    F er[3], ei[3];
    Phase<P>::unit(d[0] * hm, er[0], ei[0]);
    Phase<P>::unit(d[1] * hm, er[1], ei[1]);
    er[2] = er[1] * er[0] + ei[1] * ei[0];
    ei[2] = ei[1] * er[0] - er[1] * ei[0];
    F fr[3], fi[3];
    for (int p = 0; p < 3; ++p) {
      // sinc(x), x = 2 pi q: a widths' phase is below half a revolution in
      // practice (polynomial quotient); wider bins divide the reduced sine
      const D q = D(0.5) * d[p] * hw;
      F sinc;
      if (m_fabs(q) <= D(0.5)) {
        const F x = F(D(2.0 * 3.14159265358979323846) * q);
        sinc = m_fabs(x) >= F(1e-6) ? m_sinc2pi_half(F(q))
                                    : F(1) - x * x / F(6) + x * x * x * x / F(120);
      } else {
        F wr, sn;
        m_cispi_m1(F(q - m_rint(q)), &wr, &sn);
        sinc = sn / F(D(2.0 * 3.14159265358979323846) * q);
      }
      fr[p] = sinc * er[p];
      fi[p] = sinc * ei[p];
    }
    for (int a = 0; a < 3; ++a) {
      const F w = F(bins.weight(a, bin));
      for (int p = 0; p < 3; ++p) {
        xr[p][a] = xr[p][a] + w * fr[p];
        xi[p][a] = xi[p][a] + w * fi[p];
      }
    }
  }
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      F qr[3], qi[3];
      F sum = F(0);
      for (int k = 0; k < 3; ++k) {
        qr[k] = ur[k][b] * vr[k][a] - ui[k][b] * vi[k][a];
        qi[k] = ur[k][b] * vi[k][a] + ui[k][b] * vr[k][a];
        sum = sum + (qr[k] * qr[k] + qi[k] * qi[k]);
      }
      for (int p = 0; p < 3; ++p) {
        const int j = p < 2 ? 0 : 1, k = p == 0 ? 1 : 2;
        const F pr = qr[j] * qr[k] + qi[j] * qi[k];
        const F pi = qi[j] * qr[k] - qr[j] * qi[k];
        sum = sum + F(2) * (pr * xr[p][a] - pi * xi[p][a]);
      }
      prob[a * 3 + b] = sum;
    }
}

// A single constant-density layer of length L (the beam path,
// physics.hpp:1153-1260).  In the real frame the layer matrix is
// T = I + sum_k w_k v_k v_k^T over the two non-reference eigenvectors, and
// the flavour-basis amplitude is B = W T W^+ with W = R23 G, so
//   B = I + sum_k w_k u_k u_k^+,  u_k = W v_k = (v0, c v1 + s e^{id} v2,
//                                               -s v1 + c e^{id} v2),
// and P(a->b) = |B[b][a]|^2.  The Hermitian products u_k[b] conj(u_k[a])
// serve both B[b][a] and B[a][b] (the latter takes their conjugates), so a
// pair of entries costs one set of products.
template <class P>
CP3_HD void beam_point(const SetConst& sc, typename P::D E,
                       typename P::D rhoYe, typename P::D L,
                       typename P::F prob[9]) {
  using D = typename P::D;
  using F = typename P::F;
  Slab<P> s;
  // the mixed form normalises the eigenvectors in F here: one layer adds no
  // error of its own (1.2e-6 from FP64 on 0.1-20 GeV at 1300 km), where
  // chained layers add up
  solve_slab<P, true>(sc, E, m_recip(E), rhoYe, s);
  const F c = F(sc.c23), sn = F(sc.s23), cd = F(sc.cd), sd = F(sc.sd);
  F wr[2], wi[2], ur[2][3], ui[2][3];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const F* v = k == 0 ? s.va : s.vb;
    Phase<P>::eval(k == 0 ? s.oma : s.omb, L, wr[k], wi[k]);
    const F g2r = cd * v[2], g2i = sd * v[2];
    ur[k][0] = v[0];
    ui[k][0] = F(0);
    ur[k][1] = c * v[1] + sn * g2r;
    ui[k][1] = sn * g2i;
    ur[k][2] = c * g2r - sn * v[1];
    ui[k][2] = c * g2i;
  }
  // diagonal: B[m][m] = 1 + sum_k w_k |u_k[m]|^2
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    F br = F(1), bi = F(0);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const F d = ur[k][m] * ur[k][m] + ui[k][m] * ui[k][m];
      br = br + wr[k] * d;
      bi = bi + wi[k] * d;
    }
    prob[m * 3 + m] = br * br + bi * bi;
  }
  // pairs (b, a) = (1, 0), (2, 0), (2, 1): h_k = u_k[b] conj(u_k[a]),
  // B[b][a] = sum_k w_k h_k, B[a][b] = sum_k w_k conj(h_k)
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    const int b = p == 0 ? 1 : 2, a = p == 2 ? 1 : 0;
    F x = F(0), y = F(0), z = F(0), q = F(0);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const F hr = ur[k][b] * ur[k][a] + ui[k][b] * ui[k][a];
      const F hi = ui[k][b] * ur[k][a] - ur[k][b] * ui[k][a];
      x = x + wr[k] * hr;
      y = y + wi[k] * hi;
      z = z + wr[k] * hi;
      q = q + wi[k] * hr;
    }
    const F bar = x - y, bai = z + q;  // B[b][a]
    const F abr = x + y, abi = q - z;  // B[a][b]
    prob[a * 3 + b] = bar * bar + bai * bai;
    prob[b * 3 + a] = abr * abr + abi * abi;
  }
}

}  // namespace cp3
