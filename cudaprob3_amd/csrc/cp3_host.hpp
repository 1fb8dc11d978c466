// Host-side preparation for the propagation kernels: per parameter-set
// constants, the Earth model as the reference's AtmosCpuPropagator holds it,
// and the per-grid trajectory table (segment lengths per cosine and shell).
// Geometry is evaluated in double with the reference's formulas and order, so
// the segment lengths the kernels see are the ones the reference computes per
// point (physics.hpp:745-787, 897-898).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "cp3_math.hpp"

namespace cp3 {

// PMNS matrix as propagator.hpp:148-183 builds it ([flavour][mass], re/im).
struct Pmns {
  double re[3][3], im[3][3];
};

inline Pmns make_pmns(double t12, double t13, double t23, double dcp,
                      int kNuType) {
  if (kNuType < 0) dcp *= -1;
  const double s12 = std::sin(t12), s13 = std::sin(t13), s23 = std::sin(t23);
  const double c12 = std::cos(t12), c13 = std::cos(t13), c23 = std::cos(t23);
  const double sd = std::sin(dcp), cd = std::cos(dcp);
  Pmns u;
  u.re[0][0] = c12 * c13;                        u.im[0][0] = 0.0;
  u.re[0][1] = s12 * c13;                        u.im[0][1] = 0.0;
  u.re[0][2] = s13 * cd;                         u.im[0][2] = -s13 * sd;
  u.re[1][0] = -s12 * c23 - c12 * s23 * s13 * cd; u.im[1][0] = -c12 * s23 * s13 * sd;
  u.re[1][1] = c12 * c23 - s12 * s23 * s13 * cd;  u.im[1][1] = -s12 * s23 * s13 * sd;
  u.re[1][2] = s23 * c13;                        u.im[1][2] = 0.0;
  u.re[2][0] = s12 * s23 - c12 * c23 * s13 * cd;  u.im[2][0] = -c12 * c23 * s13 * sd;
  u.re[2][1] = -c12 * s23 - s12 * c23 * s13 * cd; u.im[2][1] = -s12 * c23 * s13 * sd;
  u.re[2][2] = c23 * c13;                        u.im[2][2] = 0.0;
  return u;
}

// Mass-difference matrix DM[i][j] = m_i - m_j with the 5e-9 degeneracy
// breaker (propagator.hpp:188-209).
inline void make_dm(double dm12sq, double dm23sq, double DM[3][3]) {
  double m[3] = {0.0, dm12sq, dm12sq + dm23sq};
  const double brk = 5.0e-9;
  if (dm12sq == 0.0) m[0] -= brk;
  if (dm23sq == 0.0) m[2] += brk;
  for (int i = 0; i < 3; ++i) DM[i][i] = 0.0;
  DM[0][1] = m[0] - m[1];
  DM[1][0] = -DM[0][1];
  DM[0][2] = m[0] - m[2];
  DM[2][0] = -DM[0][2];
  DM[1][2] = m[1] - m[2];
  DM[2][1] = -DM[1][2];
}

// Vacuum ordering (physics.hpp:198-252): for each vacuum mass DM(i,0), the
// index of the closest root of the vacuum cubic, first minimum winning.
inline void vacuum_order(const double DM[3][3], int32_t order[3]) {
  const double a = DM[0][1] + DM[0][2];
  const double b = DM[0][1] * DM[0][2];
  const double g = 0.0;
  const double tmp = a * a - 3.0 * b;
  double arg = (2.0 * a * a * a - 9.0 * a * b + 27.0 * g) /
               (2.0 * std::sqrt(tmp * tmp * tmp));
  if (std::fabs(arg) > 1.0) arg = arg / std::fabs(arg);
  const double t0 = std::acos(arg) / 3.0;
  const double t1 = t0 - (2.0 * M_PI / 3.0);
  const double t2 = t0 + (2.0 * M_PI / 3.0);
  double r[3];
  r[0] = r[1] = r[2] = -(2.0 / 3.0) * std::sqrt(tmp);
  r[0] *= std::cos(t0);
  r[1] *= std::cos(t1);
  r[2] *= std::cos(t2);
  const double sh = DM[0][0] - a / 3.0;
  for (int k = 0; k < 3; ++k) r[k] += sh;
  for (int i = 0; i < 3; ++i) {
    double best = std::fabs(DM[i][0] - r[0]);
    int k = 0;
    for (int j = 1; j < 3; ++j) {
      const double d = std::fabs(DM[i][0] - r[j]);
      if (d < best) {
        k = j;
        best = d;
      }
    }
    order[i] = k;
  }
}

// All per (parameter set, neutrino type) constants the kernels read.
// Arguments as the reference setters take them: setMNSMatrix(t12, t13, t23,
// dcp, kNuType), setNeutrinoMasses(dm12sq, dm23sq) and
// calculateProbabilities(nutype) with nutype 0 = neutrino, 1 = antineutrino.
inline SetConst make_set(double t12, double t13, double t23, double dcp,
                         int kNuType, double dm12sq, double dm23sq,
                         int nutype) {
  SetConst s{};
  const Pmns u = make_pmns(t12, t13, t23, dcp, kNuType);
  double DM[3][3];
  make_dm(dm12sq, dm23sq, DM);
  s.g0 = u.re[0][0] * u.re[0][0] + u.im[0][0] * u.im[0][0];
  s.q1 = 1.0 - u.re[0][1] * u.re[0][1] - u.im[0][1] * u.im[0][1];
  s.q2 = 1.0 - u.re[0][2] * u.re[0][2] - u.im[0][2] * u.im[0][2];
  s.dm01 = DM[0][1];
  s.dm02 = DM[0][2];
  s.dm00 = DM[0][0];
  vacuum_order(DM, s.order);
  // O = R13 R12, real frame mass matrix O diag(DM(k,0)) O^T.
  const double s12 = std::sin(t12), c12 = std::cos(t12);
  const double s13 = std::sin(t13), c13 = std::cos(t13);
  const double O[3][3] = {{c13 * c12, c13 * s12, s13},
                          {-s12, c12, 0.0},
                          {-s13 * c12, -s13 * s12, c13}};
  const double d[3] = {DM[0][0], DM[1][0], DM[2][0]};
  const int ii[6] = {0, 0, 0, 1, 1, 2};
  const int jj[6] = {0, 1, 2, 1, 2, 2};
  for (int n = 0; n < 6; ++n) {
    double acc = 0.0;
    for (int k = 0; k < 3; ++k) acc += O[ii[n]][k] * d[k] * O[jj[n]][k];
    s.hv[n] = acc;
  }
  // roots in vacuum order: root r = amp (A_r cos t0 + B_r sin t0) + shift
  const double A[3] = {1.0, -0.5, -0.5};
  const double B[3] = {0.0, kSqrt3Half, -kSqrt3Half};
  for (int k = 0; k < 3; ++k) {
    s.ra[k] = A[s.order[k]];
    s.rb[k] = B[s.order[k]];
  }
  // H'^2 without the matter term's contributions
  const double* h = s.hv;
  s.hh[0] = h[1] * h[1] + h[2] * h[2];
  s.hh[1] = h[1] * h[3] + h[2] * h[4];
  s.hh[2] = h[1] * h[4] + h[2] * h[5];
  s.hh[3] = h[1] * h[1] + h[3] * h[3] + h[4] * h[4];
  s.hh[4] = h[1] * h[2] + h[3] * h[4] + h[4] * h[5];
  s.hh[5] = h[2] * h[2] + h[4] * h[4] + h[5] * h[5];
  s.sgn = nutype ? 1.0 : -1.0;
  s.c23 = std::cos(t23);
  s.s23 = std::sin(t23);
  const double dd = kNuType < 0 ? -dcp : dcp;
  s.cd = std::cos(dd);
  s.sd = std::sin(dd);
  return s;
}

// Earth model after setDensity (atmoscpupropagator.hpp:116-170): rows
// descending by radius, coslimit per row.  A polynomial model
// (setDensity(radii, a, b, c, yps), :181-245) gives row i the density
// as[i] R^2 + bs[i] R + cs[i] at radius R km instead of rhos[i].
struct EarthModel {
  std::vector<double> radii, rhos, yps, coslimit;
  std::vector<double> as, bs, cs;
  bool set = false, poly = false;
};

inline void earth_coslimits(EarthModel& e) {
  e.coslimit.clear();
  for (size_t i = 0; i < e.radii.size(); ++i) {
    double x = -1 * std::sqrt(1 - (e.radii[i] * e.radii[i] /
                                   (kREarthKm * kREarthKm)));
    if (i == 0) x = 0;
    e.coslimit.push_back(x);
  }
}

inline void earth_set(EarthModel& e, const std::vector<double>& r,
                      const std::vector<double>& rho,
                      const std::vector<double>& ye) {
  if (rho.size() != r.size())
    throw std::invalid_argument("setDensity : rhos.size() != radii.size()");
  if (rho.size() != ye.size())
    throw std::invalid_argument("setDensity : rhos.size() != yps.size()");
  if (rho.empty() || r.empty() || ye.empty())
    throw std::invalid_argument("setDensity : vectors must not be empty");
  bool flip = false;
  if (r.size() >= 2) {
    const int sign = (r[1] - r[0] > 0 ? 1 : -1);
    for (size_t i = 1; i < r.size(); ++i)
      if ((r[i] - r[i - 1]) * sign < 0)
        throw std::invalid_argument("radii order messed up");
    flip = (sign == 1);
  }
  e.radii = r;
  e.rhos = rho;
  e.yps = ye;
  if (flip) {
    std::reverse(e.radii.begin(), e.radii.end());
    std::reverse(e.rhos.begin(), e.rhos.end());
    std::reverse(e.yps.begin(), e.yps.end());
  }
  earth_coslimits(e);
  e.set = true;
  e.poly = false;
}

// setDensity(radii, a, b, c, yps) (atmoscpupropagator.hpp:181-245): the same
// checks and flip, rhos untouched.  b and c must be as long as the radii (the
// reference checks only a and would read past shorter lists).
inline void earth_set_poly(EarthModel& e, const std::vector<double>& r,
                           const std::vector<double>& a,
                           const std::vector<double>& b,
                           const std::vector<double>& c,
                           const std::vector<double>& ye) {
  if (a.size() != r.size())
    throw std::invalid_argument("setDensity : a.size() != radii.size()");
  if (a.size() != ye.size())
    throw std::invalid_argument("setDensity : a.size() != yps.size()");
  if (a.empty() || b.empty() || c.empty() || r.empty() || ye.empty())
    throw std::invalid_argument("setDensity : vectors must not be empty");
  if (b.size() != r.size() || c.size() != r.size())
    throw std::invalid_argument("setDensity : b.size() and c.size() must equal radii.size()");
  bool flip = false;
  if (r.size() >= 2) {
    const int sign = (r[1] - r[0] > 0 ? 1 : -1);
    for (size_t i = 1; i < r.size(); ++i)
      if ((r[i] - r[i - 1]) * sign < 0)
        throw std::invalid_argument("radii order messed up");
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
  earth_coslimits(e);
  e.set = true;
  e.poly = true;
}

// setChemicalComposition (atmoscpupropagator.hpp:481-490): the new list is
// taken in the model's stored (descending) row order.
inline void earth_set_yps(EarthModel& e, const std::vector<double>& ye) {
  if (ye.size() != e.yps.size())
    throw std::invalid_argument(
        "setChemicalComposition : size of input list not equal to "
        "expectation");
  e.yps = ye;
}

// Number of shell rows whose cosine limit lies strictly above c
// (atmoscpupropagator.hpp:434).
inline int earth_max_layer(const EarthModel& e, double c) {
  int n = 0;
  for (double lim : e.coslimit)
    if (c < lim) ++n;
  return n;
}

// Splits a density-file line on single or repeated spaces the way
// atmoscpupropagator.hpp:265-281 counts fields.
inline int density_file_fields(std::string line) {
  int n = 0;
  while (line.find_first_of(" ") != std::string::npos) {
    size_t pos = line.find_first_of(" ");
    while (pos + 1 < line.size() && line[pos] == line[pos + 1]) ++pos;
    line = line.substr(pos + 1);
    ++n;
  }
  return n + 1;
}

// The fields of a density-file line, split as density_file_fields counts them
// (atmoscpupropagator.hpp:313-322).
inline std::vector<std::string> density_file_split(std::string line) {
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

// setDensityFromFile (atmoscpupropagator.hpp:250-344): 3 columns (radius km,
// density g/cm^3, Ye) or 5 columns (radius, a, b, c, Ye: a polynomial model,
// fields converted with atof as :313-327).  Other formats raise
// std::invalid_argument (the reference ends the process there, :338-342).
inline void earth_load_file(EarthModel& e, const std::string& path) {
  std::ifstream in(path);
  if (!in) throw std::invalid_argument("could not open density file " + path);
  std::string line;
  int ncol = 0;
  while (std::getline(in, line)) {
    if (line.empty() || line[0] == '#') continue;
    const int n = density_file_fields(line);
    if (ncol == 0) ncol = n;
  }
  if (ncol != 3 && ncol != 5)
    throw std::invalid_argument("unsupported Earth model in " + path + ": " +
                                std::to_string(ncol) +
                                " entries per line (3 or 5 expected)");
  in.clear();
  in.seekg(0);
  if (ncol == 5) {
    std::vector<double> r, a, b, c, ye;
    while (std::getline(in, line)) {
      if (line.empty() || line[0] == '#') continue;
      const std::vector<std::string> f = density_file_split(line);
      if (f.size() < 5) throw std::invalid_argument("short line in density file " + path);
      r.push_back(std::atof(f[0].c_str()));
      a.push_back(std::atof(f[1].c_str()));
      b.push_back(std::atof(f[2].c_str()));
      c.push_back(std::atof(f[3].c_str()));
      ye.push_back(std::atof(f[4].c_str()));
    }
    earth_set_poly(e, r, a, b, c, ye);
    return;
  }
  std::vector<double> r, rho, ye;
  double a, b, c;
  while (in >> a >> b >> c) {
    r.push_back(a);
    rho.push_back(b);
    ye.push_back(c);
  }
  earth_set(e, r, rho, ye);
}

// Trajectory table of one (Earth model, cosine list, production height):
//   M[c]        innermost layer index (= number of shells crossed);
//   Lvac[c]     atmosphere segment, km (physics.hpp:757-759);
//   seg[c*S+i]  km in shell i for i < M[c]: half chord difference for
//               i < M[c]-1, full chord for the innermost (physics.hpp:780-786);
//   rye[i]      rho*Ye of shell i (physics.hpp:634-646); for a polynomial
//               model rye[c*S+i], the density averaged over that segment
//               (physics.hpp:653-739), and rye_cs = S (0 otherwise): row
//               stride of rye per cosine;
//   perm        cosine indices ordered by decreasing M (stable), so blocks of
//               neighbouring entries cross similar numbers of shells.
struct TrajTable {
  int nC = 0, S = 0, rye_cs = 0;
  std::vector<int32_t> M, perm;
  std::vector<double> Lvac, seg, rye;
};

inline double segment_cm(const std::vector<double>& radii, int layer,
                         int max_layer, double path_cm, double earth_cm,
                         double cz) {
  if (cz >= 0) return path_cm;
  if (layer == 0) return path_cm - earth_cm;
  const int i = (layer <= max_layer) ? layer - 1 : 2 * max_layer - layer - 1;
  const double r2min = kREarthKm * kREarthKm * (1 - cz * cz);
  const double ri = radii[(size_t)i];
  // A model passed without its 0-km row would read past the end here in the
  // reference (physics.hpp:777); the value is unused on that branch.
  const double rn = ((size_t)i + 1 < radii.size()) ? radii[(size_t)i + 1] : 0.0;
  double cross_this = 2.0 * std::sqrt(ri * ri - r2min);
  if (std::sqrt(r2min) > ri) cross_this = 0;
  double cross_next = 2.0 * std::sqrt(rn * rn - r2min);
  if (std::sqrt(r2min) > rn) cross_next = 0;
  if (i < max_layer - 1) return 0.5 * (cross_this - cross_next) * kKm2Cm;
  return cross_this * kKm2Cm;
}

// rho*Ye of layer `layer` of a trajectory at cosine cz through a polynomial
// model, averaged over its segment of dist_cm (physics.hpp:653-739): the
// segment runs from the shell's outer boundary inward, R(t)^2 = t^2 - 2 d t +
// R_i^2 with d the half chord at R_i, and the integrals of R^2, R and 1 over
// t in [0, dist] take the reference's closed forms in its operation order
// (the results must agree to the last bit).  Straight through the centre
// (cz = -1) it is the radial mean over the shell.
inline double poly_layer_rye(const EarthModel& e, double cz, int layer, int max_layer,
                             double dist_cm) {
  if (cz >= 0 || layer == 0) return 0.0;
  const int i = (layer <= max_layer) ? layer - 1 : 2 * max_layer - layer - 1;
  const size_t k = (size_t)i;
  const double Ri = e.radii[k];
  // a model without its 0-km row reads radius 0 below its last row
  const double Rn = (k + 1 < e.radii.size()) ? e.radii[k + 1] : 0.0;
  const double pa = e.as[k], pb = e.bs[k], pc = e.cs[k];
  const double r2min = std::pow(kREarthKm, 2) * (1 - std::pow(cz, 2));
  const double r2max = std::pow(Ri, 2);
  const double d = std::sqrt(r2max - r2min);
  if (r2min == 0) {
    const double up = pa * std::pow(Ri, 3) / 3. + pb * std::pow(Ri, 2) / 2. + pc * Ri;
    const double dn = pa * std::pow(Rn, 3) / 3. + pb * std::pow(Rn, 2) / 2. + pc * Rn;
    double rho = up - dn;
    rho /= (Ri - Rn);
    return rho * e.yps[k];
  }
  // R^2 = u t^2 + v t + w
  const double u = 1.0, v = -2 * d, w = r2max;
  const double t1 = 0., t2 = dist_cm / kKm2Cm;
  const auto r2_int = [&](double t) {
    return u * std::pow(t, 3) / 3. + v * std::pow(t, 2) / 2. + w * t;
  };
  const auto r_int_a = [&](double t) {
    return (v + 2. * u * t) * std::sqrt(w + t * (v + u * t)) / (4. * u);
  };
  const auto r_int_b = [&](double t) {
    return (v * v - 4. * u * w) *
           std::log(v + 2. * u * t + 2. * std::sqrt(u) * std::sqrt(w + t * (v + u * t))) /
           (8. * std::pow(u, 3. / 2.));
  };
  const double i2 = r2_int(t2) - r2_int(t1);
  const double i1 = (r_int_a(t2) - r_int_b(t2)) - (r_int_a(t1) - r_int_b(t1));
  const double i0 = t2 - t1;
  const double rho = (pa * i2 + pb * i1 + pc * i0) / (t2 - t1);
  // A grazing trajectory (cz a little below 0, about -1e-9 and up) crosses the
  // outer shell along a segment so short that these closed forms reach 0/0 or
  // log(0), and the reference's density is NaN.  The mean over a vanishing
  // segment is the density at the shell's outer radius.
  if (!std::isfinite(rho)) return (pa * r2max + pb * Ri + pc) * e.yps[k];
  return rho * e.yps[k];
}

// Production-height bins of a cosine list (physics.hpp:1009-1033): the path
// length from each of the nb + 1 edges (km) minus the one from h_cm, in cm
// and in the reference's operation order, then per bin the mid shift and the
// width, hmw[(c * nb + b) * 2 + {0, 1}], in km.
inline void build_height_bins(std::vector<double>& hmw, const double* cosv, int nC,
                              double h_cm, const std::vector<double>& edges_km) {
  const int nb = (int)edges_km.size() - 1;
  hmw.assign((size_t)nC * (size_t)(nb > 0 ? nb : 0) * 2, 0.0);
  std::vector<double> shift(edges_km.size());
  for (int c = 0; c < nC; ++c) {
    const double cz = cosv[c];
    const double path_cm = std::sqrt((kREarthCm + h_cm) * (kREarthCm + h_cm) -
                                     (kREarthCm * kREarthCm) * (1 - cz * cz)) -
                           kREarthCm * cz;
    for (size_t b = 0; b < edges_km.size(); ++b) {
      const double hb = kKm2Cm * edges_km[b];
      const double pl = (std::sqrt((kREarthCm + hb) * (kREarthCm + hb) -
                                   (kREarthCm * kREarthCm) * (1 - cz * cz)) -
                         kREarthCm * cz);
      shift[b] = pl - path_cm;
    }
    for (int b = 0; b < nb; ++b) {
      const double h0 = shift[(size_t)b], h1 = shift[(size_t)b + 1];
      double* d = &hmw[((size_t)c * nb + (size_t)b) * 2];
      d[0] = (h1 + h0) / 2. / kKm2Cm;
      d[1] = (h1 - h0) / kKm2Cm;
    }
  }
}

// The bin probabilities of setProductionHeightList, [type][flavour][E][cos][bin]
// (physics.hpp:1042-1043), reordered to [type][flavour][bin][cos][E] so that
// the lanes of a wave (consecutive energies) read consecutive entries.
template <class T>
void reorder_height_prob(std::vector<T>& out, const double* prob, int nC, int nE, int nb) {
  const size_t plane = (size_t)nC * (size_t)nE;
  out.resize((size_t)6 * (size_t)nb * plane);
  for (int tf = 0; tf < 6; ++tf)
    for (int e = 0; e < nE; ++e)
      for (int c = 0; c < nC; ++c)
        for (int b = 0; b < nb; ++b)
          out[((size_t)tf * nb + (size_t)b) * plane + (size_t)c * nE + (size_t)e] =
              (T)prob[(((size_t)tf * nE + (size_t)e) * nC + (size_t)c) * nb + (size_t)b];
}

inline void build_traj_table(TrajTable& t, const EarthModel& e,
                             const double* cosv, int nC, double h_cm) {
  const int S = (int)e.radii.size();
  t.nC = nC;
  t.S = S;
  t.M.assign((size_t)nC, 0);
  t.Lvac.assign((size_t)nC, 0.0);
  t.seg.assign((size_t)nC * (size_t)S, 0.0);
  t.rye_cs = e.poly ? S : 0;
  if (e.poly) {
    t.rye.assign((size_t)nC * (size_t)S, 0.0);
  } else {
    t.rye.assign((size_t)S, 0.0);
    for (int i = 0; i < S; ++i) t.rye[(size_t)i] = e.rhos[(size_t)i] * e.yps[(size_t)i];
  }
  for (int c = 0; c < nC; ++c) {
    const double cz = cosv[c];
    const int M = earth_max_layer(e, cz);
    t.M[(size_t)c] = M;
    const double earth_cm = -2.0 * cz * kREarthCm;
    const double path_cm = std::sqrt((kREarthCm + h_cm) * (kREarthCm + h_cm) -
                                     (kREarthCm * kREarthCm) * (1 - cz * cz)) -
                           kREarthCm * cz;
    t.Lvac[(size_t)c] = segment_cm(e.radii, 0, M, path_cm, earth_cm, cz) / kKm2Cm;
    for (int l = 1; l <= M; ++l) {
      const double dist_cm = segment_cm(e.radii, l, M, path_cm, earth_cm, cz);
      t.seg[(size_t)c * S + (size_t)(l - 1)] = dist_cm / kKm2Cm;
      if (e.poly)
        t.rye[(size_t)c * S + (size_t)(l - 1)] = poly_layer_rye(e, cz, l, M, dist_cm);
    }
  }
  t.perm.resize((size_t)nC);
  std::iota(t.perm.begin(), t.perm.end(), 0);
  std::stable_sort(t.perm.begin(), t.perm.end(), [&](int a, int b) {
    return t.M[(size_t)a] > t.M[(size_t)b];
  });
}

}  // namespace cp3
