// ORACLE - test infrastructure only.
//
// A CPU restatement, in plain double-precision C++, of the reference's
// three-flavour Barger propagator as run by AtmosCpuPropagator<double> and
// BeamCpuPropagator<double>.  It exists to check the HIP path: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only
// as the checker.  It is never on the product path.
//
// Each function follows the reference operation for operation (the same
// formulas, evaluation order and branches) so that it agrees with the compiled
// reference to the last bit or nearly so; the reference lines are cited on each
// function.  Differences on purpose:
//   * no MaxNLayers cap (constants.hpp:46): shell scratch is sized per call;
//   * the innermost-shell lookup of radii[i+1] (physics.hpp:777) reads 0 when a
//     model is passed without its 0-km row, instead of reading past the end;
//     the value is unused on that branch (physics.hpp:780-786);
//   * unsupported density files raise an error instead of std::terminate
//     (atmoscpupropagator.hpp:338-342);
//   * polynomial models need b and c lists as long as the radii (the
//     reference checks only a's size, atmoscpupropagator.hpp:190-200, and
//     would read past a shorter list), and the cos = -1 branch of the
//     polynomial average reads radius 0 past a model without its 0-km row;
//   * production-height averaging's sinc (math.hpp:49-55) is the intended
//     sin(A)/A for |A| >= 1e-6 unless `int_abs_sinc` asks for the compiled
//     reference's form, whose abs() truncates to an integer.
//
// Pins (tests/): golden vectors made by the compiled reference
// (tests/golden/), the reference library itself where it is built
// (oracle/_ref/), closed forms (vacuum oscillation, U.U^+ = I, eigenvalues vs
// numpy.linalg.eigvalsh, one layer vs scipy.linalg.expm, chord geometry sums)
// and invariants (unitarity, nu-bar <-> (delta, A) -> (-delta, -A), T-symmetry
// at delta = 0); production-height averaging against zero-width bins (the
// plain propagation from that height), Gauss-Legendre means over path-length
// shifts and the incoherent limit.

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

struct cx {
  double re;
  double im;
};

// constants.hpp:28-40 and the hard-coded L/E factor of physics.hpp:444, 524.
constexpr double kGf2 = 1.52588e-4;  // 2*sqrt(2)*G_F*N_A*(hbar c)^3 [eV^2/GeV/(g/cm^3)]
constexpr double kKm2Cm = 1.0e5;
constexpr double kREarthKm = 6371.0;
constexpr double kREarthCm = kREarthKm * kKm2Cm;
constexpr double kBeamYe = 0.5;
constexpr double kLoE = 2.534;

// The state Propagator<double> carries into the physics core:
// Mix_U, dm (propagator.hpp:279-280) and the vacuum ordering computed by
// prepare_getMfast (physics.hpp:198-252).
struct Osc {
  cx U[3][3];
  double DM[3][3];
  int order[3];
};

// propagator.hpp:148-183: PDG parametrisation, delta negated when kNuType < 0.
void set_pmns(Osc& o, double t12, double t13, double t23, double dcp,
              int kNuType) {
  if (kNuType < 0) dcp *= -1;
  const double s12 = std::sin(t12), s13 = std::sin(t13), s23 = std::sin(t23);
  const double c12 = std::cos(t12), c13 = std::cos(t13), c23 = std::cos(t23);
  const double sd = std::sin(dcp), cd = std::cos(dcp);
  o.U[0][0] = {c12 * c13, 0.0};
  o.U[0][1] = {s12 * c13, 0.0};
  o.U[0][2] = {s13 * cd, -s13 * sd};
  o.U[1][0] = {-s12 * c23 - c12 * s23 * s13 * cd, -c12 * s23 * s13 * sd};
  o.U[1][1] = {c12 * c23 - s12 * s23 * s13 * cd, -s12 * s23 * s13 * sd};
  o.U[1][2] = {s23 * c13, 0.0};
  o.U[2][0] = {s12 * s23 - c12 * c23 * s13 * cd, -c12 * c23 * s13 * sd};
  o.U[2][1] = {-c12 * s23 - s12 * c23 * s13 * cd, -s12 * c23 * s13 * sd};
  o.U[2][2] = {c23 * c13, 0.0};
}

// propagator.hpp:188-209: vacuum masses (0, dm12, dm12+dm23) with the 5e-9
// degeneracy breaker, DM[i][j] = m_i - m_j.
void set_masses(Osc& o, double dm12sq, double dm23sq) {
  double m[3] = {0.0, dm12sq, dm12sq + dm23sq};
  const double brk = 5.0e-9;
  if (dm12sq == 0.0) m[0] -= brk;
  if (dm23sq == 0.0) m[2] += brk;
  o.DM[0][0] = o.DM[1][1] = o.DM[2][2] = 0.0;
  o.DM[0][1] = m[0] - m[1];
  o.DM[1][0] = -o.DM[0][1];
  o.DM[0][2] = m[0] - m[2];
  o.DM[2][0] = -o.DM[0][2];
  o.DM[1][2] = m[1] - m[2];
  o.DM[2][1] = -o.DM[1][2];
}

// Trigonometric roots of Barger et al. eq. (21) for the coefficients
// (alpha, beta, gamma); shared shape of physics.hpp:213-229 and 296-323.
// `clamp_tmp` reproduces getMfast's max(tmp, 0) (physics.hpp:297), which the
// vacuum version does not apply (physics.hpp:214).
static void cubic_roots(double alpha, double beta, double gamma, double DM00,
                        bool clamp_tmp, double root[3]) {
  double tmp = alpha * alpha - 3.0 * beta;
  if (clamp_tmp && tmp < 0) tmp = 0;
  double arg = (2.0 * alpha * alpha * alpha - 9.0 * alpha * beta + 27.0 * gamma) /
               (2.0 * std::sqrt(tmp * tmp * tmp));
  if (std::fabs(arg) > 1.0) arg = arg / std::fabs(arg);
  const double th0 = std::acos(arg) / 3.0;
  const double th1 = th0 - (2.0 * M_PI / 3.0);
  const double th2 = th0 + (2.0 * M_PI / 3.0);
  const double amp = -(2.0 / 3.0) * std::sqrt(tmp);
  root[0] = amp;
  root[1] = amp;
  root[2] = amp;
  root[0] *= std::cos(th0);
  root[1] *= std::cos(th1);
  root[2] *= std::cos(th2);
  const double shift = DM00 - alpha / 3.0;
  root[0] += shift;
  root[1] += shift;
  root[2] += shift;
}

// physics.hpp:198-252: order[i] = index of the vacuum root closest to DM(i,0)
// (first minimum wins).
void set_order(Osc& o) {
  const double a = o.DM[0][1] + o.DM[0][2];
  const double b = o.DM[0][1] * o.DM[0][2];
  double r[3];
  cubic_roots(a, b, 0.0, o.DM[0][0], false, r);
  for (int i = 0; i < 3; ++i) {
    double best = std::fabs(o.DM[i][0] - r[0]);
    int k = 0;
    for (int j = 1; j < 3; ++j) {
      const double d = std::fabs(o.DM[i][0] - r[j]);
      if (d < best) {
        k = j;
        best = d;
      }
    }
    o.order[i] = k;
  }
}

// physics.hpp:276-281, 352-357: matter term, negative for neutrinos.
static inline double matter_fac(double E, double rhoYe, bool anti) {
  return anti ? kGf2 * E * rhoYe : -kGf2 * E * rhoYe;
}

// physics.hpp:267-340 (getMfast): matter eigenvalues ordered by continuity to
// vacuum; dMM[i][j] = l_i - l_j, dMV[i][j] = l_i - DM(j,0).
void matter_masses(const Osc& o, double E, double rhoYe, bool anti,
                   double dMM[3][3], double dMV[3][3]) {
  const double fac = matter_fac(E, rhoYe, anti);
  const cx* U0 = o.U[0];
  const double d01 = o.DM[0][1], d02 = o.DM[0][2];
  const double alpha = fac + d01 + d02;
  const double beta =
      d01 * d02 +
      fac * (d01 * (1.0 - U0[1].re * U0[1].re - U0[1].im * U0[1].im) +
             d02 * (1.0 - U0[2].re * U0[2].re - U0[2].im * U0[2].im));
  const double gamma =
      fac * d01 * d02 * (U0[0].re * U0[0].re + U0[0].im * U0[0].im);
  double r[3], lam[3];
  cubic_roots(alpha, beta, gamma, o.DM[0][0], true, r);
  for (int i = 0; i < 3; ++i) lam[i] = r[o.order[i]];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      dMM[i][j] = lam[i] - lam[j];
      dMV[i][j] = lam[i] - o.DM[j][0];
    }
}

// physics.hpp:345-430 (get_product): Sylvester projectors of the mass-basis
// matrix, P[i][j][k] = prod_{m != k} (M - l_m)[i][j] / prod_{m != k}(l_k - l_m).
void projectors(const Osc& o, double E, double rhoYe, bool anti,
                const double dMV[3][3], const double dMM[3][3],
                cx P[3][3][3]) {
  const double fac = matter_fac(E, rhoYe, anti);
  cx H[3][3][3];  // H[n][m][j] = (M - l_j)[n][m]
  for (int n = 0; n < 3; ++n)
    for (int m = 0; m < 3; ++m) {
      const cx& a = o.U[0][n];
      const cx& b = o.U[0][m];
      const double re = -fac * (a.re * b.re + a.im * b.im);
      const double im = -fac * (a.re * b.im - a.im * b.re);
      for (int j = 0; j < 3; ++j) H[n][m][j] = {re, im};
    }
  for (int j = 0; j < 3; ++j)
    for (int n = 0; n < 3; ++n) H[n][n][j].re -= dMV[j][n];

  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) P[i][j][k] = {0.0, 0.0};
  // projector k uses the two factors (M - l_a)(M - l_b) with {a,b} = the
  // other two indices in cyclic order (1,2), (2,0), (0,1).
  static const int fa[3] = {1, 2, 0};
  static const int fb[3] = {2, 0, 1};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      for (int m = 0; m < 3; ++m)
        for (int k = 0; k < 3; ++k) {
          const cx& x = H[i][m][fa[k]];
          const cx& y = H[m][j][fb[k]];
          P[i][j][k].re += x.re * y.re - x.im * y.im;
          P[i][j][k].im += x.re * y.im + x.im * y.re;
        }
      const double d0 = dMM[0][1] * dMM[0][2];
      const double d1 = dMM[1][2] * dMM[1][0];
      const double d2 = dMM[2][0] * dMM[2][1];
      P[i][j][0].re /= d0;
      P[i][j][0].im /= d0;
      P[i][j][1].re /= d1;
      P[i][j][1].im /= d1;
      P[i][j][2].re /= d2;
      P[i][j][2].im /= d2;
    }
}

// physics.hpp:439-450 (getArg): phases -2.534 * l_k * L / E (phase offset 0).
void phases(double L, double E, const double dMV[3][3], double arg[3]) {
  for (int k = 0; k < 3; ++k) arg[k] = -kLoE * dMV[k][0] * L / E;
}

// physics.hpp:459-514 (getC): C_k = U P_k U^+ in the flavour basis.
void flavour_projectors(const Osc& o, const cx P[3][3][3], cx C[3][3][3]) {
  for (int k = 0; k < 3; ++k)
    for (int a = 0; a < 3; ++a) {
      double rr[3] = {0., 0., 0.}, ri[3] = {0., 0., 0.};
      double ir[3] = {0., 0., 0.}, ii[3] = {0., 0., 0.};
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          rr[j] += o.U[a][i].re * P[i][j][k].re;
          ri[j] += o.U[a][i].re * P[i][j][k].im;
          ir[j] += o.U[a][i].im * P[i][j][k].re;
          ii[j] += o.U[a][i].im * P[i][j][k].im;
        }
      for (int b = 0; b < 3; ++b) {
        double sre = 0., sim = 0.;
        for (int j = 0; j < 3; ++j) {
          const cx& u = o.U[b][j];
          sre += rr[j] * u.re;
          sre += ri[j] * u.im;
          sre += ir[j] * u.im;
          sre -= ii[j] * u.re;
          sim += ii[j] * u.im;
          sim += ir[j] * u.re;
          sim += ri[j] * u.re;
          sim -= rr[j] * u.im;
        }
        C[k][a][b] = {sre, sim};
      }
    }
}

// physics.hpp:620-629 (get_transition_matrix_expansion).
void layer_expansion(const Osc& o, double E, double rhoYe, double L_km,
                     bool anti, cx C[3][3][3], double arg[3]) {
  double dMM[3][3], dMV[3][3];
  cx P[3][3][3];
  matter_masses(o, E, rhoYe, anti, dMM, dMV);
  phases(L_km, E, dMV, arg);
  projectors(o, E, rhoYe, anti, dMV, dMM, P);
  flavour_projectors(o, P, C);
}

// math.hpp:57-75 (multiply_phase_matrix, host branch): B += A * e^{i phase}.
static inline void add_phased(double phase, const cx A[3][3], cx B[3][3]) {
  const double s = std::sin(phase);
  const double c = std::cos(phase);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      B[i][j].re += c * A[i][j].re - s * A[i][j].im;
      B[i][j].im += c * A[i][j].im + s * A[i][j].re;
    }
}

// math.hpp:82-96 (multiply_complex_matrix): C += A * B.
static inline void mul_acc(const cx A[3][3], const cx B[3][3], cx C[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) {
        C[i][j].re += A[i][k].re * B[k][j].re - A[i][k].im * B[k][j].im;
        C[i][j].im += A[i][k].im * B[k][j].re + A[i][k].re * B[k][j].im;
      }
}

static inline void zero(cx A[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) A[i][j] = {0.0, 0.0};
}

// C = A * B into a fresh matrix (clear + multiply, physics.hpp:972-974).
static inline void mul(const cx A[3][3], const cx B[3][3], cx C[3][3]) {
  zero(C);
  mul_acc(A, B, C);
}

// physics.hpp:745-787 (getTraversedDistanceOfLayer), in cm.  `radii` is
// descending with n entries; radii[n] reads as 0 (see header).
double segment_cm(const double* radii, int n, int layer, int max_layer,
                  double path_cm, double earth_cm, double cz) {
  if (cz >= 0) return path_cm;
  if (layer == 0) return path_cm - earth_cm;
  const int i = (layer <= max_layer) ? layer - 1 : 2 * max_layer - layer - 1;
  const double r2min = kREarthKm * kREarthKm * (1 - cz * cz);
  const double ri = radii[i];
  const double rn = (i + 1 < n) ? radii[i + 1] : 0.0;
  double cross_this = 2.0 * std::sqrt(ri * ri - r2min);
  if (std::sqrt(r2min) > ri) cross_this = 0;
  double cross_next = 2.0 * std::sqrt(rn * rn - r2min);
  if (std::sqrt(r2min) > rn) cross_next = 0;
  if (i < max_layer - 1) return 0.5 * (cross_this - cross_next) * kKm2Cm;
  return cross_this * kKm2Cm;
}

// physics.hpp:634-646 (getDensityOfLayer): rho*Ye of the shell, 0 in vacuum.
double layer_rho_ye(const double* rhos, const double* yps, int layer,
                    int max_layer) {
  if (layer == 0) return 0.0;
  const int i = (layer <= max_layer) ? layer - 1 : 2 * max_layer - layer - 1;
  return rhos[i] * yps[i];
}

// An Earth model as AtmosCpuPropagator<double> holds it after setDensity
// (atmoscpupropagator.hpp:116-170): rows descending, coslimit per row.  With
// `poly` (setDensity(radii, a, b, c, yps), :181-245) the density of row i is
// rho(R) = as[i] R^2 + bs[i] R + cs[i], R in km, averaged per trajectory.
struct Earth {
  std::vector<double> radii, rhos, yps, coslimit;
  std::vector<double> as, bs, cs;
  bool poly = false;
};

// physics.hpp:653-739 (getDensityOfLayerPoly): rho*Ye averaged over the
// layer's segment of dist_cm, the segment parametrised from the shell's outer
// boundary inward by t (km) with R(t)^2 = t^2 - 2 d t + R_i^2, d the half chord
// at R_i.  The integrals of R^2, R and 1 over t in [0, t2] are the closed
// forms of the reference, in its operation order.
double layer_rho_ye_poly(const Earth& e, double cz, int layer, int max_layer,
                         double dist_cm) {
  if (cz >= 0) return 0.0;
  if (layer == 0) return 0.0;
  const int i = (layer <= max_layer) ? layer - 1 : 2 * max_layer - layer - 1;
  const size_t n = e.radii.size();
  const double Ri = e.radii[(size_t)i];
  const double Rn = ((size_t)i + 1 < n) ? e.radii[(size_t)i + 1] : 0.0;
  const double A = e.as[(size_t)i], B = e.bs[(size_t)i], Cc = e.cs[(size_t)i];
  const double rmin2 = std::pow(kREarthKm, 2) * (1 - std::pow(cz, 2));
  const double rmax2 = std::pow(Ri, 2);
  const double half = std::sqrt(rmax2 - rmin2);
  if (rmin2 == 0) {
    // straight through the centre: the radial average over [Rn, Ri]
    const double hi = A * std::pow(Ri, 3) / 3. + B * std::pow(Ri, 2) / 2. + Cc * Ri;
    const double lo = A * std::pow(Rn, 3) / 3. + B * std::pow(Rn, 2) / 2. + Cc * Rn;
    double rho = hi - lo;
    rho /= (Ri - Rn);
    return rho * e.yps[(size_t)i];
  }
  const double qa = 1.0, qb = -2 * half, qc = rmax2;  // R^2 = qa t^2 + qb t + qc
  const double t2 = dist_cm / kKm2Cm;
  const double t1 = 0.;
  auto sq_int = [&](double t) {
    return qa * std::pow(t, 3) / 3. + qb * std::pow(t, 2) / 2. + qc * t;
  };
  auto lin_a = [&](double t) {
    return (qb + 2. * qa * t) * std::sqrt(qc + t * (qb + qa * t)) / (4. * qa);
  };
  auto lin_b = [&](double t) {
    return (qb * qb - 4. * qa * qc) *
           std::log(qb + 2. * qa * t + 2. * std::sqrt(qa) * std::sqrt(qc + t * (qb + qa * t))) /
           (8. * std::pow(qa, 3. / 2.));
  };
  const double sq = sq_int(t2) - sq_int(t1);
  const double lin = (lin_a(t2) - lin_b(t2)) - (lin_a(t1) - lin_b(t1));
  const double con = t2 - t1;
  const double rho = (A * sq + B * lin + Cc * con) / (t2 - t1);
  return rho * e.yps[(size_t)i];
}

// atmoscpupropagator.hpp:116-170.
void set_density(Earth& e, const std::vector<double>& r,
                 const std::vector<double>& rho,
                 const std::vector<double>& ye) {
  if (rho.size() != r.size())
    throw std::runtime_error("setDensity : rhos.size() != radii.size()");
  if (rho.size() != ye.size())
    throw std::runtime_error("setDensity : rhos.size() != yps.size()");
  if (rho.empty() || r.empty() || ye.empty())
    throw std::runtime_error("setDensity : vectors must not be empty");
  bool flip = false;
  if (r.size() >= 2) {
    const int sign = (r[1] - r[0] > 0 ? 1 : -1);
    for (size_t i = 1; i < r.size(); ++i)
      if ((r[i] - r[i - 1]) * sign < 0)
        throw std::runtime_error("radii order messed up");
    flip = (sign == 1);
  }
  e.radii = r;
  e.rhos = rho;
  e.yps = ye;
  if (flip) {
    std::vector<double> a(e.radii.rbegin(), e.radii.rend());
    std::vector<double> b(e.rhos.rbegin(), e.rhos.rend());
    std::vector<double> c(e.yps.rbegin(), e.yps.rend());
    e.radii.swap(a);
    e.rhos.swap(b);
    e.yps.swap(c);
  }
  e.coslimit.clear();
  for (size_t i = 0; i < e.radii.size(); ++i) {
    double x = -1 * std::sqrt(1 - (e.radii[i] * e.radii[i] /
                                   (kREarthKm * kREarthKm)));
    if (i == 0) x = 0;
    e.coslimit.push_back(x);
  }
}

// atmoscpupropagator.hpp:431-444 without the MaxNLayers cap: number of shell
// rows whose cosine limit lies strictly above c.
// atmoscpupropagator.hpp:181-245: the polynomial model, flipped to
// descending rows like the constant one (rhos untouched).
void set_density_poly(Earth& e, const std::vector<double>& r,
                      const std::vector<double>& a, const std::vector<double>& b,
                      const std::vector<double>& c, const std::vector<double>& ye) {
  if (a.size() != r.size())
    throw std::runtime_error("setDensity : a.size() != radii.size()");
  if (a.size() != ye.size())
    throw std::runtime_error("setDensity : a.size() != yps.size()");
  if (a.empty() || b.empty() || c.empty() || r.empty() || ye.empty())
    throw std::runtime_error("setDensity : vectors must not be empty");
  if (b.size() != r.size() || c.size() != r.size())
    throw std::runtime_error("setDensity : b and c must have radii.size() entries");
  bool flip = false;
  if (r.size() >= 2) {
    const int sign = (r[1] - r[0] > 0 ? 1 : -1);
    for (size_t i = 1; i < r.size(); ++i)
      if ((r[i] - r[i - 1]) * sign < 0)
        throw std::runtime_error("radii order messed up");
    flip = (sign == 1);
  }
  e.radii = r;
  e.as = a;
  e.bs = b;
  e.cs = c;
  e.yps = ye;
  if (flip)
    for (std::vector<double>* v : {&e.radii, &e.yps, &e.as, &e.bs, &e.cs})
      std::reverse(v->begin(), v->end());
  e.coslimit.clear();
  for (size_t i = 0; i < e.radii.size(); ++i) {
    double x = -1 * std::sqrt(1 - (e.radii[i] * e.radii[i] /
                                   (kREarthKm * kREarthKm)));
    if (i == 0) x = 0;
    e.coslimit.push_back(x);
  }
  e.poly = true;
}

int max_layer(const Earth& e, double c) {
  int n = 0;
  for (double lim : e.coslimit)
    if (c < lim) ++n;
  return n;
}

// Splits a density-file line on single or repeated spaces the way
// atmoscpupropagator.hpp:265-281 / 313-322 does, returning the field count.
static std::vector<std::string> split_fields(std::string line) {
  std::vector<std::string> f;
  while (line.find_first_of(" ") != std::string::npos) {
    size_t pos = line.find_first_of(" ");
    f.push_back(line.substr(0, pos));
    while (pos + 1 < line.size() && line[pos] == line[pos + 1]) ++pos;
    line = line.substr(pos + 1);
  }
  f.push_back(line);
  return f;
}

// atmoscpupropagator.hpp:250-344: 3 columns (radius, rho, Ye) or 5 columns
// (radius, a, b, c, Ye; fields converted with atof as :313-327 does).
void set_density_file(Earth& e, const std::string& path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("could not open density file " + path);
  std::string line;
  int ncol = 0;
  while (std::getline(in, line)) {
    if (line.empty() || line[0] == '#') continue;
    const int n = (int)split_fields(line).size();
    if (ncol == 0) ncol = n;
  }
  if (ncol != 3 && ncol != 5)
    throw std::runtime_error("unsupported Earth model: " +
                             std::to_string(ncol) + " entries per line");
  in.clear();
  in.seekg(0);
  if (ncol == 5) {
    std::vector<double> r, a, b, c, y;
    while (std::getline(in, line)) {
      if (line.empty() || line[0] == '#') continue;
      const std::vector<std::string> f = split_fields(line);
      if (f.size() < 5) throw std::runtime_error("short line in " + path);
      r.push_back(std::atof(f[0].c_str()));
      a.push_back(std::atof(f[1].c_str()));
      b.push_back(std::atof(f[2].c_str()));
      c.push_back(std::atof(f[3].c_str()));
      y.push_back(std::atof(f[4].c_str()));
    }
    set_density_poly(e, r, a, b, c, y);
    return;
  }
  std::vector<double> r, d, y;
  double a, b, c;
  while (in >> a >> b >> c) {
    r.push_back(a);
    d.push_back(b);
    y.push_back(c);
  }
  set_density(e, r, d, y);
}

// Production-height averaging input (physics.hpp:803-807): nb bins, bin
// probabilities prob[type][flavour][E][cos][bin] (the layout indexed at
// physics.hpp:1042-1043) and nb + 1 bin edges in km (multiplied by km2cm at
// physics.hpp:1013).  `int_abs_sinc` reproduces the compiled reference's
// defined_sinc (math.hpp:49-55): there g++ resolves the unqualified abs() to
// int abs(int), so the series branch is taken for every |A| < 1 instead of
// |A| < 1e-6.  Without it the sinc is the intended one.
struct Pha {
  int nb;
  const double* prob;
  const double* edges_km;
  bool int_abs_sinc;
};

// math.hpp:49-55 (defined_sinc), Epsilon() = 1e-6 (constants.hpp:49).
static inline double sinc(double A, bool int_abs) {
  const double mag = int_abs ? (double)std::abs((int)A) : std::fabs(A);
  if (mag >= 1e-6) return std::sin(A) / A;
  return 1.0 - A * A / 6. + A * A * A * A / 120.;
}

// physics.hpp:789-1150 (calculate_atmos; constant or polynomial densities,
// :915-920; production-height averaging when pha is given, :997-1067).  out is
// [cos][E][9], channel a*3+b = P(a->b).
void calculate_atmos(const Osc& o, const Earth& earth, bool anti,
                     const double* cosv, int nC, const double* ev, int nE,
                     double h_cm, double* out, const Pha* pha = nullptr) {
  const int nrows = (int)earth.radii.size();
  std::vector<int> maxl((size_t)nC);
  for (int c = 0; c < nC; ++c) maxl[(size_t)c] = max_layer(earth, cosv[c]);

#pragma omp parallel for schedule(dynamic)
  for (int ic = 0; ic < nC; ++ic) {
    const double cz = cosv[ic];
    const double earth_cm = -2.0 * cz * kREarthCm;
    const int M = maxl[(size_t)ic];
    cx C_atm[3][3][3];
    for (int ie = 0; ie < nE; ++ie) {
      const double E = ev[ie];
      cx fin[3][3], ctm[3][3], T[3][3], tmp[3][3];
      double darg0[3] = {0., 0., 0.};
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) ctm[i][j] = {i == j ? 1.0 : 0.0, 0.0};
      const double path_cm =
          std::sqrt((kREarthCm + h_cm) * (kREarthCm + h_cm) -
                    (kREarthCm * kREarthCm) * (1 - cz * cz)) -
          kREarthCm * cz;
      for (int layer = 0; layer <= M; ++layer) {
        const double dist = segment_cm(earth.radii.data(), nrows, layer, M,
                                       path_cm, earth_cm, cz);
        const double rye =
            earth.poly ? layer_rho_ye_poly(earth, cz, layer, M, dist)
                       : layer_rho_ye(earth.rhos.data(), earth.yps.data(), layer, M);
        cx C[3][3][3];
        double arg[3];
        layer_expansion(o, E, rye, dist / kKm2Cm, anti, C, arg);
        zero(T);
        for (int k = 0; k < 3; ++k) add_phased(arg[k], C[k], T);
        if (layer == 0) {
          std::memcpy(fin, T, sizeof(fin));
          std::memcpy(C_atm, C, sizeof(C_atm));
          // physics.hpp:958-967: phase per cm of the atmosphere segment
          for (int k = 0; k < 3; ++k) darg0[k] = dist == 0. ? 0. : arg[k] / dist;
        } else if (layer < M) {
          mul(T, fin, tmp);
          std::memcpy(fin, tmp, sizeof(fin));
          mul(ctm, T, tmp);
          std::memcpy(ctm, tmp, sizeof(ctm));
        } else {
          mul(T, fin, tmp);
          std::memcpy(fin, tmp, sizeof(fin));
        }
      }
      mul(ctm, fin, tmp);
      std::memcpy(fin, tmp, sizeof(fin));

      // physics.hpp:997-1067: length-shift factors S[k][j][flavour before],
      // 1 (off the diagonal) without averaging.
      cx S[3][3][3];
      if (pha) {
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j)
            for (int f = 0; f < 3; ++f) S[i][j][f] = {i == j ? 1.0 : 0.0, 0.};
        std::vector<double> shift((size_t)pha->nb + 1);
        for (int b = 0; b < pha->nb + 1; ++b) {
          const double hb = kKm2Cm * pha->edges_km[b];
          const double pl = (std::sqrt((kREarthCm + hb) * (kREarthCm + hb) -
                                       (kREarthCm * kREarthCm) * (1 - cz * cz)) -
                             kREarthCm * cz);
          shift[(size_t)b] = pl - path_cm;
        }
        const size_t nb = (size_t)pha->nb;
        for (int bin = 0; bin < pha->nb; ++bin) {
          const double h0 = shift[(size_t)bin], h1 = shift[(size_t)bin + 1];
          const double hm = (h1 + h0) / 2.;
          const double hw = (h1 - h0);
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < i; ++j) {
              const double dd = darg0[i] - darg0[j];
              const double A = 0.5 * dd * hw;
              const cx f = {sinc(A, pha->int_abs_sinc) * std::cos(dd * hm),
                            sinc(A, pha->int_abs_sinc) * std::sin(dd * hm)};
              for (int fl = 0; fl < 3; ++fl) {
                const size_t idx = (size_t)(anti ? 1 : 0) * 3 * (size_t)nE * nC * nb +
                                   (size_t)fl * (size_t)nE * nC * nb +
                                   (size_t)ie * nC * nb + (size_t)ic * nb + (size_t)bin;
                const double w = pha->prob[idx];
                S[i][j][fl].re += w * f.re;
                S[i][j][fl].im += w * f.im;
                S[j][i][fl].re += w * f.re;
                S[j][i][fl].im -= w * f.im;
              }
            }
        }
      } else {
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < i; ++j)
            for (int f = 0; f < 3; ++f) S[i][j][f] = {1.0, 0.};
      }

      // physics.hpp:1102-1130: P(a->b) assembled from the expansion of the
      // atmosphere layer.
      cx prod[3][3][3];
      for (int k = 0; k < 3; ++k) mul(fin, C_atm[k], prod[k]);
      double prob[3][3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
      for (int k = 0; k < 3; ++k) {
        for (int b = 0; b < 3; ++b)
          for (int a = 0; a < 3; ++a)
            prob[a][b] += prod[k][a][b].re * prod[k][a][b].re +
                          prod[k][a][b].im * prod[k][a][b].im;
        for (int j = 0; j < k; ++j)
          for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 3; ++a)
              prob[a][b] += 2. * prod[j][a][b].re * prod[k][a][b].re * S[k][j][b].re +
                            2. * prod[j][a][b].im * prod[k][a][b].im * S[k][j][b].re +
                            2. * prod[j][a][b].im * prod[k][a][b].re * S[k][j][b].im -
                            2. * prod[j][a][b].re * prod[k][a][b].im * S[k][j][b].im;
      }
      double* dst = out + ((size_t)ic * nE + ie) * 9;
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) dst[a * 3 + b] = prob[b][a];
    }
  }
}

// physics.hpp:516-598 (getA): beam amplitude U X U^+ with
// X = sum_k P_k e^{i arg_k}, contracted through the AXFAC factors of
// physics.hpp:157-174.
void beam_amplitude(const Osc& o, double E, double rhoYe, double L_km,
                    bool anti, cx A[3][3]) {
  double dMM[3][3], dMV[3][3];
  cx P[3][3][3];
  matter_masses(o, E, rhoYe, anti, dMM, dMV);
  projectors(o, E, rhoYe, anti, dMV, dMM, P);
  cx X[3][3];
  zero(X);
  for (int k = 0; k < 3; ++k) {
    const double arg = -kLoE * dMV[k][0] * L_km / E;
    const double s = std::sin(arg);
    const double c = std::cos(arg);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        X[i][j].re += c * P[i][j][k].re - s * P[i][j][k].im;
        X[i][j].im += c * P[i][j][k].im + s * P[i][j][k].re;
      }
  }
  for (int n = 0; n < 3; ++n)
    for (int m = 0; m < 3; ++m) {
      double are = 0, aim = 0;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          const cx& u = o.U[n][i];
          const cx& v = o.U[m][j];
          const double f0 = u.re * v.re + u.im * v.im;
          const double f1 = u.re * v.im - u.im * v.re;
          const double f2 = u.im * v.im + u.re * v.re;
          const double f3 = u.im * v.re - u.re * v.im;
          are += f0 * X[i][j].re + f1 * X[i][j].im;
          aim += f2 * X[i][j].im + f3 * X[i][j].re;
        }
      A[n][m] = {are, aim};
    }
}

// physics.hpp:1153-1260 (calculate_beam).  out is [E][9].
void calculate_beam(const Osc& o, bool anti, const double* ev, int nE,
                    double rho, double L_km, double* out) {
#pragma omp parallel for
  for (int ie = 0; ie < nE; ++ie) {
    cx A[3][3];
    beam_amplitude(o, ev[ie], rho * kBeamYe, L_km, anti, A);
    double* dst = out + (size_t)ie * 9;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        dst[a * 3 + b] = 0. + (A[b][a].re * A[b][a].re + A[b][a].im * A[b][a].im);
  }
}

static void make_osc(Osc& o, double t12, double t13, double t23, double dcp,
                     int kNuType, double dm12, double dm23) {
  set_pmns(o, t12, t13, t23, dcp, kNuType);
  set_masses(o, dm12, dm23);
  set_order(o);
}

static void put_err(char* err, int errlen, const char* msg) {
  if (err && errlen > 0) std::snprintf(err, (size_t)errlen, "%s", msg);
}

static void set_threads(int threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
}

}  // namespace orc

extern "C" {

// Atmospheric probabilities, [cos][E][9].  Argument meaning follows the
// reference call sequence (example/mainatmos.cpp:121-148).
int orc_atmos_f64(int nutype, const double* cosv, int nC, const double* ev,
                  int nE, const double* radii, const double* rhos,
                  const double* yps, int nrows, double th12, double th13,
                  double th23, double dcp, int kNuType, double dm12,
                  double dm23, double h_km, int threads, double* out,
                  char* err, int errlen) {
  try {
    orc::Osc o;
    orc::make_osc(o, th12, th13, th23, dcp, kNuType, dm12, dm23);
    orc::Earth e;
    orc::set_density(e, std::vector<double>(radii, radii + nrows),
                     std::vector<double>(rhos, rhos + nrows),
                     std::vector<double>(yps, yps + nrows));
    orc::set_threads(threads);
    orc::calculate_atmos(o, e, nutype != 0, cosv, nC, ev, nE,
                         h_km * 100000.0, out);
  } catch (const std::exception& ex) {
    orc::put_err(err, errlen, ex.what());
    return -1;
  }
  return 0;
}

// The same with production-height averaging (physics.hpp:997-1067): nb bins,
// prob[2][3][nE][nC][nb] (neutrino type, initial flavour, energy, cosine, bin)
// and edges_km[nb + 1]; int_abs_sinc selects the compiled reference's sinc.
int orc_atmos_pha_f64(int nutype, const double* cosv, int nC, const double* ev,
                      int nE, const double* radii, const double* rhos,
                      const double* yps, int nrows, double th12, double th13,
                      double th23, double dcp, int kNuType, double dm12,
                      double dm23, double h_km, int nb, const double* prob,
                      const double* edges_km, int int_abs_sinc, int threads,
                      double* out, char* err, int errlen) {
  try {
    if (nb < 0 || nb > 20)
      throw std::runtime_error("invalid number of production height bins");
    orc::Osc o;
    orc::make_osc(o, th12, th13, th23, dcp, kNuType, dm12, dm23);
    orc::Earth e;
    orc::set_density(e, std::vector<double>(radii, radii + nrows),
                     std::vector<double>(rhos, rhos + nrows),
                     std::vector<double>(yps, yps + nrows));
    orc::set_threads(threads);
    const orc::Pha pha{nb, prob, edges_km, int_abs_sinc != 0};
    orc::calculate_atmos(o, e, nutype != 0, cosv, nC, ev, nE,
                         h_km * 100000.0, out, &pha);
  } catch (const std::exception& ex) {
    orc::put_err(err, errlen, ex.what());
    return -1;
  }
  return 0;
}

// Production-height averaging through a polynomial Earth model.
int orc_atmos_poly_pha_f64(int nutype, const double* cosv, int nC, const double* ev,
                           int nE, const double* radii, const double* pa,
                           const double* pb, const double* pc, const double* yps,
                           int nrows, double th12, double th13, double th23,
                           double dcp, int kNuType, double dm12, double dm23,
                           double h_km, int nb, const double* prob,
                           const double* edges_km, int int_abs_sinc, int threads,
                           double* out, char* err, int errlen) {
  try {
    if (nb < 0 || nb > 20)
      throw std::runtime_error("invalid number of production height bins");
    orc::Osc o;
    orc::make_osc(o, th12, th13, th23, dcp, kNuType, dm12, dm23);
    orc::Earth e;
    auto v = [&](const double* x) { return std::vector<double>(x, x + nrows); };
    orc::set_density_poly(e, v(radii), v(pa), v(pb), v(pc), v(yps));
    orc::set_threads(threads);
    const orc::Pha pha{nb, prob, edges_km, int_abs_sinc != 0};
    orc::calculate_atmos(o, e, nutype != 0, cosv, nC, ev, nE,
                         h_km * 100000.0, out, &pha);
  } catch (const std::exception& ex) {
    orc::put_err(err, errlen, ex.what());
    return -1;
  }
  return 0;
}

// The same with a polynomial Earth model (setDensity(radii, a, b, c, yps)).
int orc_atmos_poly_f64(int nutype, const double* cosv, int nC, const double* ev,
                       int nE, const double* radii, const double* pa,
                       const double* pb, const double* pc, const double* yps,
                       int nrows, double th12, double th13, double th23,
                       double dcp, int kNuType, double dm12, double dm23,
                       double h_km, int threads, double* out, char* err,
                       int errlen) {
  try {
    orc::Osc o;
    orc::make_osc(o, th12, th13, th23, dcp, kNuType, dm12, dm23);
    orc::Earth e;
    auto v = [&](const double* x) { return std::vector<double>(x, x + nrows); };
    orc::set_density_poly(e, v(radii), v(pa), v(pb), v(pc), v(yps));
    orc::set_threads(threads);
    orc::calculate_atmos(o, e, nutype != 0, cosv, nC, ev, nE,
                         h_km * 100000.0, out);
  } catch (const std::exception& ex) {
    orc::put_err(err, errlen, ex.what());
    return -1;
  }
  return 0;
}

// Probe: rho*Ye of every (cosine, layer 1..M) of a polynomial model, as the
// propagation uses it; out[c * nrows + l - 1], 0 past M.
int orc_poly_rye(const double* cosv, int nC, const double* radii, const double* pa,
                 const double* pb, const double* pc, const double* yps, int nrows,
                 double h_km, double* out, char* err, int errlen) {
  try {
    orc::Earth e;
    auto v = [&](const double* x) { return std::vector<double>(x, x + nrows); };
    orc::set_density_poly(e, v(radii), v(pa), v(pb), v(pc), v(yps));
    const double h_cm = h_km * 100000.0;
    for (int ic = 0; ic < nC; ++ic) {
      const double cz = cosv[ic];
      const int M = orc::max_layer(e, cz);
      const double earth_cm = -2.0 * cz * orc::kREarthCm;
      const double path_cm =
          std::sqrt((orc::kREarthCm + h_cm) * (orc::kREarthCm + h_cm) -
                    (orc::kREarthCm * orc::kREarthCm) * (1 - cz * cz)) -
          orc::kREarthCm * cz;
      for (int l = 1; l <= nrows; ++l) {
        double r = 0.0;
        if (l <= M) {
          const double dist = orc::segment_cm(e.radii.data(), nrows, l, M, path_cm,
                                              earth_cm, cz);
          r = orc::layer_rho_ye_poly(e, cz, l, M, dist);
        }
        out[(size_t)ic * nrows + (size_t)(l - 1)] = r;
      }
    }
  } catch (const std::exception& ex) {
    orc::put_err(err, errlen, ex.what());
    return -1;
  }
  return 0;
}

int orc_atmos_file_f64(int nutype, const double* cosv, int nC,
                       const double* ev, int nE, const char* path, double th12,
                       double th13, double th23, double dcp, int kNuType,
                       double dm12, double dm23, double h_km, int threads,
                       double* out, char* err, int errlen) {
  try {
    orc::Osc o;
    orc::make_osc(o, th12, th13, th23, dcp, kNuType, dm12, dm23);
    orc::Earth e;
    orc::set_density_file(e, path);
    orc::set_threads(threads);
    orc::calculate_atmos(o, e, nutype != 0, cosv, nC, ev, nE,
                         h_km * 100000.0, out);
  } catch (const std::exception& ex) {
    orc::put_err(err, errlen, ex.what());
    return -1;
  }
  return 0;
}

// Beam probabilities, [E][9]; rho in g/cm^3 (Ye = 0.5 applied), L in km.
int orc_beam_f64(int nutype, const double* ev, int nE, double rho,
                 double L_km, double th12, double th13, double th23,
                 double dcp, int kNuType, double dm12, double dm23,
                 int threads, double* out, char* err, int errlen) {
  try {
    orc::Osc o;
    orc::make_osc(o, th12, th13, th23, dcp, kNuType, dm12, dm23);
    orc::set_threads(threads);
    orc::calculate_beam(o, nutype != 0, ev, nE, rho, L_km, out);
  } catch (const std::exception& ex) {
    orc::put_err(err, errlen, ex.what());
    return -1;
  }
  return 0;
}

// --- single-function probes for the closed-form tests -------------------

// U as [3][3] complex (18 doubles, re/im interleaved).
void orc_pmns(double th12, double th13, double th23, double dcp, int kNuType,
              double* U18) {
  orc::Osc o;
  orc::set_pmns(o, th12, th13, th23, dcp, kNuType);
  std::memcpy(U18, o.U, sizeof(o.U));
}

// Ordered matter eigenvalues l_i (mass-basis matrix M, eV^2) and the vacuum
// ordering.
void orc_matter_eigs(double th12, double th13, double th23, double dcp,
                     int kNuType, double dm12, double dm23, double E,
                     double rhoYe, int nutype, double* lam3, int* order3) {
  orc::Osc o;
  orc::make_osc(o, th12, th13, th23, dcp, kNuType, dm12, dm23);
  double dMM[3][3], dMV[3][3];
  orc::matter_masses(o, E, rhoYe, nutype != 0, dMM, dMV);
  for (int i = 0; i < 3; ++i) lam3[i] = dMV[i][0] + o.DM[0][0];
  for (int i = 0; i < 3; ++i) order3[i] = o.order[i];
}

// One layer's transition matrix T = sum_k C_k e^{i arg_k} (flavour basis,
// [after][before], 18 doubles) and the mass-basis projectors P (54 doubles,
// [i][j][k]).
void orc_layer(double th12, double th13, double th23, double dcp, int kNuType,
               double dm12, double dm23, double E, double rhoYe, double L_km,
               int nutype, double* T18, double* P54) {
  orc::Osc o;
  orc::make_osc(o, th12, th13, th23, dcp, kNuType, dm12, dm23);
  orc::cx C[3][3][3], T[3][3], P[3][3][3];
  double arg[3], dMM[3][3], dMV[3][3];
  orc::layer_expansion(o, E, rhoYe, L_km, nutype != 0, C, arg);
  orc::zero(T);
  for (int k = 0; k < 3; ++k) orc::add_phased(arg[k], C[k], T);
  std::memcpy(T18, T, sizeof(T));
  orc::matter_masses(o, E, rhoYe, nutype != 0, dMM, dMV);
  orc::projectors(o, E, rhoYe, nutype != 0, dMV, dMM, P);
  std::memcpy(P54, P, sizeof(P));
}

// Geometry of one trajectory: max layer index M and the 2M+1 segment lengths
// (km) walked by the reference layer loop (layer 0 = atmosphere, then shells
// inward to the innermost full chord, listed once each as the loop visits
// them, i.e. M+1 values in seg_km[0..M]).  Returns M, or -1 on bad input.
int orc_segments(double cz, const double* radii, const double* rhos,
                 const double* yps, int nrows, double h_km, double* seg_km,
                 double* rho_ye, int cap) {
  try {
    orc::Earth e;
    orc::set_density(e, std::vector<double>(radii, radii + nrows),
                     std::vector<double>(rhos, rhos + nrows),
                     std::vector<double>(yps, yps + nrows));
    const int M = orc::max_layer(e, cz);
    if (M + 1 > cap) return -1;
    const double h_cm = h_km * 100000.0;
    const double path_cm =
        std::sqrt((orc::kREarthCm + h_cm) * (orc::kREarthCm + h_cm) -
                  (orc::kREarthCm * orc::kREarthCm) * (1 - cz * cz)) -
        orc::kREarthCm * cz;
    const double earth_cm = -2.0 * cz * orc::kREarthCm;
    for (int l = 0; l <= M; ++l) {
      seg_km[l] = orc::segment_cm(e.radii.data(), nrows, l, M, path_cm,
                                  earth_cm, cz) /
                  orc::kKm2Cm;
      rho_ye[l] = orc::layer_rho_ye(e.rhos.data(), e.yps.data(), l, M);
    }
    return M;
  } catch (...) {
    return -1;
  }
}

}  // extern "C"
