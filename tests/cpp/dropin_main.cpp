// Drives HipPropagatorSingle<double> / HipBeamPropagatorSingle<double> the way
// example/mainatmos.cpp:116-160 and example/mainbeam.cpp:112-139 drive the
// reference GPU propagators, and writes every probability to a binary file
// for the parity test (tests/test_gpu_dropin.py).
//
//   dropin_main atmos <out.bin> <precision 0|1|2> <nutype 0|1>
//   dropin_main poly  <out.bin> <precision 0|1|2> <nutype 0|1>
//   dropin_main avg   <out.bin> <precision 0|1|2> <nutype 0|1>
//   dropin_main beam  <out.bin> <precision 0|1|2> <nutype 0|1>
//   dropin_main virtual <out.bin> <precision 0|1|2> <nutype 0|1>
//   dropin_main file  <out.bin> <precision 0|1|2> <nutype 0|1> <model.dat>
//   dropin_main errors
//
// atmos: the 4-shell model, 100 x 100 grid (cos -1 + i/100, E 0.1-100 GeV
// log), nominal parameters, h = 22 km; output [cos][E][9] doubles.
// poly: the same grid through setDensity(radii, a, b, c, yps) with the
// quadratic profiles of cudaprob3_amd.workloads.prem4_poly.
// avg: the atmos case with production-height averaging over 5 bins (edges 5,
// 10, 15, 22, 30, 50 km), bin probabilities height_prob() below.
// beam: 1300 km, 2.6 g/cm^3, 500 energies over 0.1-20 GeV; output [E][9].
// virtual: the atmos and beam cases through the reference's names
// (cudaprob3linear.hpp) and base-class pointers only, the concrete class
// picked at run time as a fitter would: std::unique_ptr<Propagator<double>>
// for the calls of the reference's base (propagator.hpp:148-252) and
// AtmosPropagator<double>& for the atmospheric setters; output the atmos
// block [cos][E][9] then the beam block [E][9].
// file: setDensityFromFile(model.dat) on a 20 x 20 grid (cos -1 + i/20,
// E 1-100 GeV log), nominal parameters, h = 22 km; output [cos][E][9].
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "cudaprob3_amd/cudaprob3linear.hpp"
#include "cudaprob3_amd/hip_propagator.hpp"

using namespace cp3hip;

static std::vector<double> logspace(double lo, double hi, int n) {
  std::vector<double> v((size_t)n);
  for (int i = 0; i < n; ++i) v[(size_t)i] = lo * std::pow(hi / lo, (double)i / (n - 1));
  return v;
}

static const double kTh12 = std::asin(std::sqrt(0.307));
static const double kTh13 = std::asin(std::sqrt(0.0220));
static const double kTh23 = std::asin(std::sqrt(0.546));
static const double kDcp = -1.6, kDm12 = 7.53e-5, kDm23 = 2.494e-3;

// Bin probabilities [nutype][flavour][E][cos][bin] for the avg case:
// proportional to 1 + (7t + 5f + 3e + c + 11b) mod 13, normalised per
// (t, f, e, c).
static std::vector<double> height_prob(int nC, int nE, int nb) {
  std::vector<double> v((size_t)2 * 3 * nE * nC * nb);
  for (int t = 0; t < 2; ++t)
    for (int f = 0; f < 3; ++f)
      for (int e = 0; e < nE; ++e)
        for (int c = 0; c < nC; ++c) {
          double* w = &v[((((size_t)t * 3 + f) * nE + e) * nC + c) * nb];
          double sum = 0;
          for (int b = 0; b < nb; ++b) {
            w[b] = 1.0 + (double)((7 * t + 5 * f + 3 * e + c + 11 * b) % 13);
            sum += w[b];
          }
          for (int b = 0; b < nb; ++b) w[b] /= sum;
        }
  return v;
}

static int run_atmos(const char* path, int precision, int nutype, bool poly = false,
                     bool avg = false) {
  const int nC = 100, nE = 100;
  std::vector<double> cz((size_t)nC);
  for (int i = 0; i < nC; ++i) cz[(size_t)i] = -1.0 + i / 100.0;
  const std::vector<double> ev = logspace(0.1, 100.0, nE);
  HipPropagatorSingle<double> p(0, nC, nE, (Precision)precision);
  p.setEnergyList(ev);
  p.setCosineList(cz);
  p.setMNSMatrix(kTh12, kTh13, kTh23, kDcp, nutype ? -1 : 1);
  p.setNeutrinoMasses(kDm12, kDm23);
  if (poly) {
    const double R2 = 6371.0 * 6371.0, R = 6371.0;
    p.setDensity({0.0, 1220.0, 3480.0, 5701.0, 6371.0},
                 {-8.8381 / R2, -8.8381 / R2, -3.6426 / R2, 5.5283 / R2, 0.0 / R2},
                 {0.0 / R, 0.0 / R, -1.2638 / R, -6.4761 / R, -1.4836 / R},
                 {13.0885, 13.0885, 12.5815, 7.9565, 5.3197},
                 {0.466, 0.466, 0.466, 0.494, 0.494});
  } else {
    p.setDensity({0.0, 1220.0, 3480.0, 5701.0, 6371.0}, {13.0, 13.0, 11.3, 5.0, 3.3},
                 {0.5, 0.5, 0.5, 0.5, 0.5});
  }
  p.setProductionHeight(22.0);
  p.setMaxlayers();
  if (avg) {
    p.SetNumberOfProductionHeightBinsForAveraging(5);
    p.setProductionHeightAveraging(true);
    p.setProductionHeightList(height_prob(nC, nE, 5), {5.0, 10.0, 15.0, 22.0, 30.0, 50.0});
  }
  if (poly || avg) {
    p.calculateProbabilities(nutype ? Antineutrino : Neutrino);
  } else {
    // the CUDA class's asynchronous names (atmoscudapropagator.cuh:114-120)
    p.calculateAtmosProbabilitiesAsync(nutype ? Antineutrino : Neutrino);
    p.waitForCompletion();
    p.getResultFromDevice();
  }
  std::vector<double> out((size_t)nC * nE * 9);
  for (int c = 0; c < nC; ++c)
    for (int e = 0; e < nE; ++e)
      for (int t = 0; t < 9; ++t)
        out[((size_t)c * nE + e) * 9 + t] = p.getProbability(c, e, (ProbType)t);
  // getProbabilityArr: energy outer, cosine inner
  std::vector<double> arr((size_t)nC * nE);
  p.getProbabilityArr(arr.data(), m_e);
  for (int e = 0; e < nE; ++e)
    for (int c = 0; c < nC; ++c)
      if (arr[(size_t)e * nC + c] != out[((size_t)c * nE + e) * 9 + m_e]) {
        std::fprintf(stderr, "getProbabilityArr order mismatch\n");
        return 3;
      }
  FILE* f = std::fopen(path, "wb");
  if (!f) return 2;
  std::fwrite(out.data(), sizeof(double), out.size(), f);
  std::fclose(f);
  std::printf("atmos ok: %d x %d, kernel %.3f ms\n", nC, nE, p.lastKernelMs());
  return 0;
}

static int run_beam(const char* path, int precision, int nutype) {
  const int nE = 500;
  const std::vector<double> ev = logspace(0.1, 20.0, nE);
  HipBeamPropagatorSingle<double> p(0, nE, 1, (Precision)precision);
  p.setEnergyList(ev);
  p.setMNSMatrix(kTh12, kTh13, kTh23, kDcp, nutype ? -1 : 1);
  p.setNeutrinoMasses(kDm12, kDm23);
  p.setDensity(2.6);
  p.setPathLength(1300.0);
  p.calculateBeamProbabilitiesAsync(nutype ? Antineutrino : Neutrino);
  p.waitForCompletion();
  p.getResultFromDevice();
  std::vector<double> out((size_t)nE * 9);
  for (int e = 0; e < nE; ++e)
    for (int t = 0; t < 9; ++t) out[(size_t)e * 9 + t] = p.getProbability(e, (ProbType)t);
  FILE* f = std::fopen(path, "wb");
  if (!f) return 2;
  std::fwrite(out.data(), sizeof(double), out.size(), f);
  std::fclose(f);
  std::printf("beam ok: %d energies\n", nE);
  return 0;
}

// A fitter's factory: the concrete class behind a base pointer, chosen at
// run time.
static std::unique_ptr<cudaprob3linear::Propagator<double>> make_propagator(
    bool beam, int n_cosines, int n_energies, int precision) {
  if (beam)
    return std::unique_ptr<cudaprob3linear::Propagator<double>>(
        new cudaprob3linear::BeamCudaPropagatorSingle(0, n_energies, 1, (Precision)precision));
  return std::unique_ptr<cudaprob3linear::Propagator<double>>(
      new cudaprob3linear::AtmosCudaPropagatorSingle(0, n_cosines, n_energies,
                                                     (Precision)precision));
}

// Every channel through the base's getProbabilityArr (energy outer, cosine
// inner), re-ordered to [cos][E][9].
static void all_channels(cudaprob3linear::Propagator<double>& p, int nC, int nE,
                         double* out) {
  std::vector<double> arr((size_t)nC * nE);
  for (int t = 0; t < 9; ++t) {
    p.getProbabilityArr(arr.data(), (cudaprob3linear::ProbType)t);
    for (int e = 0; e < nE; ++e)
      for (int c = 0; c < nC; ++c) out[((size_t)c * nE + e) * 9 + t] = arr[(size_t)e * nC + c];
  }
}

static int run_virtual(const char* path, int precision, int nutype) {
  const cudaprob3linear::NeutrinoType type =
      nutype ? cudaprob3linear::Antineutrino : cudaprob3linear::Neutrino;
  const int nC = 100, nE = 100, nB = 500;
  std::vector<double> cz((size_t)nC);
  for (int i = 0; i < nC; ++i) cz[(size_t)i] = -1.0 + i / 100.0;
  std::vector<double> out((size_t)nC * nE * 9 + (size_t)nB * 9);

  std::unique_ptr<cudaprob3linear::Propagator<double>> p =
      make_propagator(false, nC, nE, precision);
  cp3hip::AtmosPropagator<double>* atm = dynamic_cast<cp3hip::AtmosPropagator<double>*>(p.get());
  if (!atm) return 5;
  p->setEnergyList(logspace(0.1, 100.0, nE));
  atm->setCosineList(cz);
  p->setMNSMatrix(kTh12, kTh13, kTh23, kDcp, nutype ? -1 : 1);
  p->setNeutrinoMasses(kDm12, kDm23);
  atm->setDensity({0.0, 1220.0, 3480.0, 5701.0, 6371.0}, {13.0, 13.0, 11.3, 5.0, 3.3},
                  {0.3, 0.3, 0.3, 0.3, 0.3});
  // the composition through the base: Ye back to 0.5 on every row
  p->setChemicalComposition({0.5, 0.5, 0.5, 0.5, 0.5});
  atm->setProductionHeight(22.0);
  p->calculateProbabilities(type);
  all_channels(*p, nC, nE, out.data());
  for (int c = 0; c < nC; c += 7)
    for (int e = 0; e < nE; e += 11)
      if (atm->getProbability(c, e, cudaprob3linear::m_e) != out[((size_t)c * nE + e) * 9 + 3]) {
        std::fprintf(stderr, "getProbability through the interface differs\n");
        return 3;
      }

  p = make_propagator(true, 0, nB, precision);
  cp3hip::BeamPropagator<double>* bm = dynamic_cast<cp3hip::BeamPropagator<double>*>(p.get());
  if (!bm) return 5;
  p->setEnergyList(logspace(0.1, 20.0, nB));
  p->setMNSMatrix(kTh12, kTh13, kTh23, kDcp, nutype ? -1 : 1);
  p->setNeutrinoMasses(kDm12, kDm23);
  p->setDensity(2.6);
  p->setPathLength(1300.0);
  p->calculateProbabilities(type);
  double* bout = out.data() + (size_t)nC * nE * 9;
  all_channels(*p, 1, nB, bout);
  if (bm->getProbability(nB - 1, cudaprob3linear::e_m) != bout[(size_t)(nB - 1) * 9 + 1]) {
    std::fprintf(stderr, "beam getProbability through the interface differs\n");
    return 3;
  }
  FILE* f = std::fopen(path, "wb");
  if (!f) return 2;
  std::fwrite(out.data(), sizeof(double), out.size(), f);
  std::fclose(f);
  std::printf("virtual ok\n");
  return 0;
}

static int run_file(const char* path, int precision, int nutype, const char* model) {
  const int nC = 20, nE = 20;
  std::vector<double> cz((size_t)nC);
  for (int i = 0; i < nC; ++i) cz[(size_t)i] = -1.0 + i / 20.0;
  HipPropagatorSingle<double> p(0, nC, nE, (Precision)precision);
  p.setEnergyList(logspace(1.0, 100.0, nE));
  p.setCosineList(cz);
  p.setMNSMatrix(kTh12, kTh13, kTh23, kDcp, nutype ? -1 : 1);
  p.setNeutrinoMasses(kDm12, kDm23);
  p.setDensityFromFile(model);
  p.setProductionHeight(22.0);
  p.calculateProbabilities(nutype ? Antineutrino : Neutrino);
  std::vector<double> out((size_t)nC * nE * 9);
  for (int c = 0; c < nC; ++c)
    for (int e = 0; e < nE; ++e)
      for (int t = 0; t < 9; ++t)
        out[((size_t)c * nE + e) * 9 + t] = p.getProbability(c, e, (ProbType)t);
  FILE* f = std::fopen(path, "wb");
  if (!f) return 2;
  std::fwrite(out.data(), sizeof(double), out.size(), f);
  std::fclose(f);
  std::printf("file ok: %d layer boundaries\n", p.getNlayerBoundaries());
  return 0;
}

template <class Fn>
static bool throws(Fn fn, const char* needle) {
  try {
    fn();
  } catch (const std::runtime_error& e) {
    if (std::strstr(e.what(), needle)) return true;
    std::fprintf(stderr, "unexpected message: %s\n", e.what());
    return false;
  }
  return false;
}

// The reference's exceptions (atmoscpupropagator.hpp:123-142, 369-370,
// 387-388; propagator.hpp:214-215; atmoscudapropagator.cu:223-224, 267-268).
static int run_errors() {
  int bad = 0;
  HipPropagatorSingle<double> p(0, 3, 4);
  bad += !throws([&] { p.setEnergyList({1.0, 2.0}); }, "number of energy nodes");
  bad += !throws([&] { p.setCosineList({-1.0}); }, "number of cosine nodes");
  bad += !throws([&] { p.setProductionHeight(22.0); }, "must set cosine list");
  bad += !throws([&] { p.setDensity({0.0, 3000.0, 1000.0}, {1, 2, 3}, {0.5, 0.5, 0.5}); },
                 "radii order");
  bad += !throws([&] { p.setDensity({0.0, 1.0}, {1.0}, {0.5, 0.5}); }, "rhos.size() != radii.size()");
  bad += !throws([&] { p.setDensity(std::vector<double>{}, {}, {}); }, "must not be empty");
  bad += !throws([&] { p.setDensity({0.0, 1.0}, {1.0}, {0.0, 0.0}, {1.0, 1.0}, {0.5, 0.5}); },
                 "a.size() != radii.size()");
  bad += !throws([&] { p.setDensityFromFile("/nonexistent/model.dat"); }, "could not open");
  p.setCosineList({-1.0, -0.5, 0.2});
  p.setEnergyList({1.0, 2.0, 3.0, 4.0});
  bad += !throws([&] { p.calculateProbabilities(Neutrino); }, "production height was not set");
  {
    // a model is needed before the first calculation (checked before any
    // upload or launch)
    HipPropagatorSingle<double> nomodel(0, 3, 4);
    nomodel.setCosineList({-1.0, -0.5, 0.2});
    nomodel.setEnergyList({1.0, 2.0, 3.0, 4.0});
    nomodel.setMNSMatrix(kTh12, kTh13, kTh23, kDcp, 1);
    nomodel.setNeutrinoMasses(kDm12, kDm23);
    nomodel.setProductionHeight(15.0);
    bad += !throws([&] { nomodel.calculateProbabilities(Neutrino); }, "density model was not set");
  }
  p.setMNSMatrix(kTh12, kTh13, kTh23, kDcp, 1);
  p.setNeutrinoMasses(kDm12, kDm23);
  p.setDensity({0.0, 3480.0, 6371.0}, {11.3, 11.3, 4.0}, {0.5, 0.5, 0.5});
  p.setProductionHeight(15.0);
  p.calculateProbabilities(Neutrino);
  bad += !throws([&] { p.getProbability(3, 0, e_e); }, "Invalid indices");
  bad += !throws([&] { p.getProbability(0, 4, e_e); }, "Invalid indices");
  bad += !throws([&] { p.setChemicalComposition({0.5}); }, "Size of input list");
  bad += !throws([&] { p.setProductionHeightList({1.0}, {1.0, 2.0}); }, "not expecting");
  // production-height averaging, which the reference class never turns on
  bad += !throws([&] { p.SetNumberOfProductionHeightBinsForAveraging(21); },
                 "invalid number of production height bins");
  std::vector<double> plain(3 * 4 * 9);
  for (int c = 0; c < 3; ++c)
    for (int e = 0; e < 4; ++e)
      for (int t = 0; t < 9; ++t) plain[(c * 4 + e) * 9 + t] = p.getProbability(c, e, (ProbType)t);
  p.SetNumberOfProductionHeightBinsForAveraging(1);
  p.setProductionHeightAveraging(true);
  bad += !throws([&] { p.calculateProbabilities(Neutrino); }, "production height array was not set");
  bad += !throws([&] { p.setProductionHeightList({1.0}, {15.0, 15.0}); },
                 "Prob array is not the expected size");
  bad += !throws([&] { p.setProductionHeightList(std::vector<double>(72, 1.0), {15.0}); },
                 "ProductionHeightBins array is not expected size");
  // one zero-width bin at the production height: the unaveraged result
  p.setProductionHeightList(std::vector<double>(72, 1.0), {15.0, 15.0});
  p.calculateProbabilities(Neutrino);
  double dev = 0;
  for (int c = 0; c < 3; ++c)
    for (int e = 0; e < 4; ++e)
      for (int t = 0; t < 9; ++t)
        dev = std::fmax(dev, std::fabs(p.getProbability(c, e, (ProbType)t) -
                                       plain[(c * 4 + e) * 9 + t]));
  if (dev > 1e-12) {
    std::fprintf(stderr, "zero-width bin moved results by %.3g\n", dev);
    ++bad;
  }
  // n >= 1 turns averaging off again, as in the reference
  p.SetNumberOfProductionHeightBinsForAveraging(2);
  bad += !throws([&] { p.setProductionHeightList(std::vector<double>(144, 0.5), {1.0, 2.0, 3.0}); },
                 "not expecting");
  HipPropagatorSingle<double> q(std::move(p));
  bad += !throws([&] { p.calculateProbabilities(Neutrino); }, "moved from");
  q.calculateProbabilities(Antineutrino);
  double s = 0;
  for (int t = 0; t < 3; ++t) s += q.getProbability(0, 0, (ProbType)t);
  if (std::fabs(s - 1.0) > 1e-12) {
    std::fprintf(stderr, "unitarity %.17g\n", s);
    ++bad;
  }
  bad += !throws([] { HipPropagatorSingle<double> r(1 << 20, 2, 2); }, "not available");
  HipBeamPropagatorSingle<double> beam(0, 4, 1);
  beam.setChemicalComposition({});
  bad += !throws([&] { beam.setChemicalComposition({0.5}); }, "Size of input list");
  std::printf("errors: %d unexpected\n", bad);
  return bad ? 1 : 0;
}

int main(int argc, char** argv) {
  try {
    if (argc >= 2 && std::string(argv[1]) == "errors") return run_errors();
    if (argc < 5) {
      std::fprintf(stderr,
                   "usage: %s atmos|poly|avg|beam|virtual out.bin precision nutype | errors\n",
                   argv[0]);
      return 2;
    }
    const std::string mode = argv[1];
    const int prec = std::atoi(argv[3]), nt = std::atoi(argv[4]);
    if (mode == "atmos") return run_atmos(argv[2], prec, nt);
    if (mode == "poly") return run_atmos(argv[2], prec, nt, true);
    if (mode == "avg") return run_atmos(argv[2], prec, nt, false, true);
    if (mode == "beam") return run_beam(argv[2], prec, nt);
    if (mode == "virtual") return run_virtual(argv[2], prec, nt);
    if (mode == "file" && argc >= 6) return run_file(argv[2], prec, nt, argv[5]);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "exception: %s\n", e.what());
    return 4;
  }
  return 2;
}
