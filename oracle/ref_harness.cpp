// Test infrastructure only: a C-ABI harness around the reference's own CPU
// propagators (AtmosCpuPropagator<T>, BeamCpuPropagator<T>), compiled from the
// reference headers where they lie (see oracle/Makefile).  It is used to pin the
// oracle restatement and to produce golden vectors; it is never on the product
// path.
//
// Two builds are made from this file:
//   libcp3ref.so       - the reference exactly as shipped (MaxNLayers() == 10,
//                        constants.hpp:46);
//   libcp3ref_deep.so  - identical except that Constants<T>::MaxNLayers() is
//                        64, so models with more than 10 shells (12 and 59 shell
//                        Earth models) can run.  MaxNLayers only sizes a
//                        per-point stack array (physics.hpp:842-850) and the
//                        check in setMaxlayers (atmoscpupropagator.hpp:436-440);
//                        the arithmetic is unchanged.  The other constants are
//                        restated with the values of constants.hpp:28-49.
//
// Output layout of every entry point: the reference CPU layout [cos][E][9]
// (physics.hpp:1142-1144), channel a*3+b = P(a->b) (types.hpp:23-27), read back
// through the public getProbability() accessor.

#include "constants.hpp"

#ifdef CP3REF_DEEP
namespace cudaprob3linear {
template <>
struct Constants<double> {
  static constexpr double tworttwoGf() { return 1.52588e-4; }
  static constexpr double km2cm() { return 1.0e5; }
  static constexpr double REarth() { return 6371.0; }
  static constexpr double REarthcm() { return REarth() * km2cm(); }
  static constexpr double density_convert() { return 0.5; }
  static constexpr int MaxProdHeightBins() { return 20; }
  static constexpr int MaxNLayers() { return 64; }
  static constexpr double Epsilon() { return 1e-6; }
};
template <>
struct Constants<float> {
  static constexpr float tworttwoGf() { return 1.52588e-4; }
  static constexpr float km2cm() { return 1.0e5; }
  static constexpr float REarth() { return 6371.0; }
  static constexpr float REarthcm() { return REarth() * km2cm(); }
  static constexpr float density_convert() { return 0.5; }
  static constexpr int MaxProdHeightBins() { return 20; }
  static constexpr int MaxNLayers() { return 64; }
  static constexpr float Epsilon() { return 1e-6; }
};
}  // namespace cudaprob3linear
#endif

#include "atmoscpupropagator.hpp"
#include "beamcpupropagator.hpp"

#include <cstdio>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

namespace {

void put_err(char* err, int errlen, const char* msg) {
  if (err && errlen > 0) {
    std::snprintf(err, (size_t)errlen, "%s", msg);
  }
}

template <typename T>
std::vector<T> to_vec(const double* p, int n) {
  std::vector<T> v((size_t)(n > 0 ? n : 0));
  for (int i = 0; i < n; ++i) v[(size_t)i] = (T)p[i];
  return v;
}

// Runs the reference atmospheric CPU propagator in the call order of
// example/mainatmos.cpp:121-148 (energies, cosines, MNS, masses, density,
// production height, calculate).  `path` selects setDensityFromFile instead of
// setDensity(radii, rhos, yps); `poly` (a, b, c per row) selects
// setDensity(radii, a, b, c, yps).
template <typename T, typename OUT>
int run_atmos(int nutype, const double* cosv, int nC, const double* ev, int nE,
              const double* radii, const double* rhos, const double* yps,
              int nrows, const char* path, const double* const* poly, double th12,
              double th13,
              double th23, double dcp, int kNuType, double dm12, double dm23,
              double h_km, int threads, OUT* out, char* err, int errlen) {
  using namespace cudaprob3linear;
  try {
    AtmosCpuPropagator<T> p(nC, nE, threads);
    p.setEnergyList(to_vec<T>(ev, nE));
    p.setCosineList(to_vec<T>(cosv, nC));
    p.setMNSMatrix((T)th12, (T)th13, (T)th23, (T)dcp, kNuType);
    p.setNeutrinoMasses((T)dm12, (T)dm23);
    if (path) {
      p.setDensityFromFile(std::string(path));
    } else if (poly) {
      p.setDensity(to_vec<T>(radii, nrows), to_vec<T>(poly[0], nrows),
                   to_vec<T>(poly[1], nrows), to_vec<T>(poly[2], nrows),
                   to_vec<T>(yps, nrows));
    } else {
      p.setDensity(to_vec<T>(radii, nrows), to_vec<T>(rhos, nrows),
                   to_vec<T>(yps, nrows));
    }
    p.setProductionHeight((T)h_km);
    p.calculateProbabilities(nutype ? Antineutrino : Neutrino);
    for (int c = 0; c < nC; ++c)
      for (int e = 0; e < nE; ++e)
        for (int t = 0; t < 9; ++t)
          out[((size_t)c * nE + e) * 9 + t] =
              (OUT)p.getProbability(c, e, (ProbType)t);
  } catch (const std::exception& ex) {
    put_err(err, errlen, ex.what());
    return -1;
  }
  return 0;
}

// The reference class never turns production-height averaging on: the flag
// defaults to false (atmoscpupropagator.hpp:532) and
// SetNumberOfProductionHeightBinsForAveraging clears it for every n >= 1
// (:97-99).  This subclass sets the protected flag after that call, so that
// the reference's own setProductionHeightList (:395-427), calculateProbabilities
// (:445-477) and averaging arithmetic (physics.hpp:997-1067, 1106-1130) run.
template <typename T>
struct AveragingPropagator : cudaprob3linear::AtmosCpuPropagator<T> {
  AveragingPropagator(int nC, int nE, int threads)
      : cudaprob3linear::AtmosCpuPropagator<T>(nC, nE, threads) {}
  void enableAveraging() { this->useProductionHeightAveraging = true; }
};

// `poly` (a, b, c per row) selects setDensity(radii, a, b, c, yps).
int run_atmos_pha(int nutype, const double* cosv, int nC, const double* ev, int nE,
                  const double* radii, const double* rhos, const double* yps,
                  int nrows, const double* const* poly, double th12, double th13,
                  double th23, double dcp, int kNuType, double dm12, double dm23,
                  double h_km, int nb, const double* prob, const double* edges_km,
                  int threads, double* out, char* err, int errlen) {
  using namespace cudaprob3linear;
  try {
    AveragingPropagator<double> p(nC, nE, threads);
    p.setEnergyList(to_vec<double>(ev, nE));
    p.setCosineList(to_vec<double>(cosv, nC));
    p.setMNSMatrix(th12, th13, th23, dcp, kNuType);
    p.setNeutrinoMasses(dm12, dm23);
    if (poly)
      p.setDensity(to_vec<double>(radii, nrows), to_vec<double>(poly[0], nrows),
                   to_vec<double>(poly[1], nrows), to_vec<double>(poly[2], nrows),
                   to_vec<double>(yps, nrows));
    else
      p.setDensity(to_vec<double>(radii, nrows), to_vec<double>(rhos, nrows),
                   to_vec<double>(yps, nrows));
    p.setProductionHeight(h_km);
    p.SetNumberOfProductionHeightBinsForAveraging(nb);
    p.enableAveraging();
    p.setProductionHeightList(to_vec<double>(prob, nb * 2 * 3 * nE * nC),
                              to_vec<double>(edges_km, nb + 1));
    p.calculateProbabilities(nutype ? Antineutrino : Neutrino);
    for (int c = 0; c < nC; ++c)
      for (int e = 0; e < nE; ++e)
        for (int t = 0; t < 9; ++t)
          out[((size_t)c * nE + e) * 9 + t] = p.getProbability(c, e, (ProbType)t);
  } catch (const std::exception& ex) {
    put_err(err, errlen, ex.what());
    return -1;
  }
  return 0;
}

// Runs the reference beam CPU propagator in the call order of
// example/mainbeam.cpp:113-133.
template <typename T, typename OUT>
int run_beam(int nutype, const double* ev, int nE, double rho, double L_km,
             double th12, double th13, double th23, double dcp, int kNuType,
             double dm12, double dm23, int threads, OUT* out, char* err,
             int errlen) {
  using namespace cudaprob3linear;
  try {
    BeamCpuPropagator<T> p(nE, threads);
    p.setEnergyList(to_vec<T>(ev, nE));
    p.setMNSMatrix((T)th12, (T)th13, (T)th23, (T)dcp, kNuType);
    p.setNeutrinoMasses((T)dm12, (T)dm23);
    p.setDensity((T)rho);
    p.setPathLength((T)L_km);
    p.calculateProbabilities(nutype ? Antineutrino : Neutrino);
    for (int e = 0; e < nE; ++e)
      for (int t = 0; t < 9; ++t)
        out[(size_t)e * 9 + t] = (OUT)p.getProbability(e, (ProbType)t);
  } catch (const std::exception& ex) {
    put_err(err, errlen, ex.what());
    return -1;
  }
  return 0;
}

}  // namespace

extern "C" {

int ref_max_layers(void) {
  return cudaprob3linear::Constants<double>::MaxNLayers();
}

int ref_atmos_f64(int nutype, const double* cosv, int nC, const double* ev,
                  int nE, const double* radii, const double* rhos,
                  const double* yps, int nrows, double th12, double th13,
                  double th23, double dcp, int kNuType, double dm12,
                  double dm23, double h_km, int threads, double* out,
                  char* err, int errlen) {
  return run_atmos<double>(nutype, cosv, nC, ev, nE, radii, rhos, yps, nrows,
                           nullptr, nullptr, th12, th13, th23, dcp, kNuType, dm12,
                           dm23, h_km, threads, out, err, errlen);
}

int ref_atmos_f32(int nutype, const double* cosv, int nC, const double* ev,
                  int nE, const double* radii, const double* rhos,
                  const double* yps, int nrows, double th12, double th13,
                  double th23, double dcp, int kNuType, double dm12,
                  double dm23, double h_km, int threads, float* out, char* err,
                  int errlen) {
  return run_atmos<float>(nutype, cosv, nC, ev, nE, radii, rhos, yps, nrows,
                          nullptr, nullptr, th12, th13, th23, dcp, kNuType, dm12,
                          dm23, h_km, threads, out, err, errlen);
}

int ref_atmos_file_f64(int nutype, const double* cosv, int nC,
                       const double* ev, int nE, const char* path, double th12,
                       double th13, double th23, double dcp, int kNuType,
                       double dm12, double dm23, double h_km, int threads,
                       double* out, char* err, int errlen) {
  return run_atmos<double>(nutype, cosv, nC, ev, nE, nullptr, nullptr, nullptr,
                           0, path, nullptr, th12, th13, th23, dcp, kNuType, dm12,
                           dm23, h_km, threads, out, err, errlen);
}

int ref_atmos_poly_f64(int nutype, const double* cosv, int nC, const double* ev,
                       int nE, const double* radii, const double* pa,
                       const double* pb, const double* pc, const double* yps,
                       int nrows, double th12, double th13, double th23,
                       double dcp, int kNuType, double dm12, double dm23,
                       double h_km, int threads, double* out, char* err,
                       int errlen) {
  const double* const poly[3] = {pa, pb, pc};
  return run_atmos<double>(nutype, cosv, nC, ev, nE, radii, nullptr, yps, nrows,
                           nullptr, poly, th12, th13, th23, dcp, kNuType, dm12,
                           dm23, h_km, threads, out, err, errlen);
}

// Atmospheric probabilities with production-height averaging on: nb bins,
// prob[2][3][nE][nC][nb] and edges_km[nb + 1] as setProductionHeightList takes
// them.
int ref_atmos_pha_f64(int nutype, const double* cosv, int nC, const double* ev,
                      int nE, const double* radii, const double* rhos,
                      const double* yps, int nrows, double th12, double th13,
                      double th23, double dcp, int kNuType, double dm12,
                      double dm23, double h_km, int nb, const double* prob,
                      const double* edges_km, int threads, double* out, char* err,
                      int errlen) {
  return run_atmos_pha(nutype, cosv, nC, ev, nE, radii, rhos, yps, nrows, nullptr, th12,
                       th13, th23, dcp, kNuType, dm12, dm23, h_km, nb, prob, edges_km,
                       threads, out, err, errlen);
}

// The same through a polynomial Earth model.
int ref_atmos_poly_pha_f64(int nutype, const double* cosv, int nC, const double* ev,
                           int nE, const double* radii, const double* pa,
                           const double* pb, const double* pc, const double* yps,
                           int nrows, double th12, double th13, double th23,
                           double dcp, int kNuType, double dm12, double dm23,
                           double h_km, int nb, const double* prob,
                           const double* edges_km, int threads, double* out,
                           char* err, int errlen) {
  const double* const poly[3] = {pa, pb, pc};
  return run_atmos_pha(nutype, cosv, nC, ev, nE, radii, nullptr, yps, nrows, poly, th12,
                       th13, th23, dcp, kNuType, dm12, dm23, h_km, nb, prob, edges_km,
                       threads, out, err, errlen);
}

int ref_beam_f64(int nutype, const double* ev, int nE, double rho, double L_km,
                 double th12, double th13, double th23, double dcp,
                 int kNuType, double dm12, double dm23, int threads,
                 double* out, char* err, int errlen) {
  return run_beam<double>(nutype, ev, nE, rho, L_km, th12, th13, th23, dcp,
                          kNuType, dm12, dm23, threads, out, err, errlen);
}

int ref_beam_f32(int nutype, const double* ev, int nE, double rho, double L_km,
                 double th12, double th13, double th23, double dcp,
                 int kNuType, double dm12, double dm23, int threads,
                 float* out, char* err, int errlen) {
  return run_beam<float>(nutype, ev, nE, rho, L_km, th12, th13, th23, dcp,
                         kNuType, dm12, dm23, threads, out, err, errlen);
}

}  // extern "C"
