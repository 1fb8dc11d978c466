// C++ drop-in propagators for MI355X over the C ABI of libcp3hip.so
// (include/cp3_hip.h).
//
//   Propagator<FLOAT_T>               the abstract base with the virtual set of
//                                     the reference's Propagator<FLOAT_T>
//                                     (propagator.hpp:148-252);
//   AtmosPropagator<FLOAT_T>          the atmospheric interface: the virtual
//                                     members AtmosCpuPropagator<FLOAT_T> adds
//                                     (atmoscpupropagator.hpp:85-513);
//   BeamPropagator<FLOAT_T>           the beam interface: the virtual members
//                                     BeamCpuPropagator<FLOAT_T> adds
//                                     (beamcpupropagator.hpp:84-140);
//   HipPropagatorSingle<FLOAT_T>      : AtmosPropagator<FLOAT_T>, replaces
//                                     AtmosCudaPropagatorSingle
//                                     (atmoscudapropagator.cuh:42-145);
//   HipBeamPropagatorSingle<FLOAT_T>  : BeamPropagator<FLOAT_T>, replaces
//                                     BeamCudaPropagatorSingle
//                                     (beamcudapropagator.cuh:41-121).
//
// A caller that holds a Propagator<double> (or AtmosPropagator<double>)
// pointer and picks the concrete class at run time can take either class;
// include/cudaprob3_amd/cudaprob3linear.hpp maps the reference's names onto
// these.
//
// FLOAT_T = double computes in FP64 and FLOAT_T = float in FP32 by default;
// Precision::Mixed (FP64 eigen-solve and phases, FP32 amplitude work) can be
// chosen for either.  Misuse throws std::runtime_error with the reference's
// messages; device failures throw too (the reference exits).  Each object owns
// its device buffers and stream; no state is process-global, so objects may be
// used from different threads concurrently (one thread per object).
#ifndef CUDAPROB3_AMD_HIP_PROPAGATOR_HPP
#define CUDAPROB3_AMD_HIP_PROPAGATOR_HPP

#include <cstdint>
#include <iostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../cp3_hip.h"

namespace cp3hip {

// types.hpp:23-29
enum ProbType : int {
  e_e = 0, e_m = 1, e_t = 2,
  m_e = 3, m_m = 4, m_t = 5,
  t_e = 6, t_m = 7, t_t = 8
};
enum NeutrinoType { Neutrino, Antineutrino };

enum class Precision : int { FP64 = CP3_FP64, FP32 = CP3_FP32, Mixed = CP3_MIXED };

template <class FLOAT_T>
constexpr Precision default_precision() {
  return std::is_same<FLOAT_T, float>::value ? Precision::FP32 : Precision::FP64;
}

inline void check(int rc) {
  if (rc != CP3_OK) throw std::runtime_error(cp3_last_error());
}

namespace detail {
template <class FLOAT_T>
std::vector<double> to_double(const std::vector<FLOAT_T>& v) {
  return std::vector<double>(v.begin(), v.end());
}
}  // namespace detail

// Propagator<FLOAT_T> (propagator.hpp:45-252): the members the reference
// declares virtual, all pure here (the reference gives the first three a
// host-side body that every concrete class here replaces with a device
// upload).
template <class FLOAT_T>
class Propagator {
 public:
  virtual ~Propagator() {}
  // propagator.hpp:148-183; delta is negated when kNuType < 0
  virtual void setMNSMatrix(FLOAT_T theta12, FLOAT_T theta13, FLOAT_T theta23,
                            FLOAT_T dCP, int kNuType) = 0;
  // propagator.hpp:188-209; dm23sq is Delta m^2_32
  virtual void setNeutrinoMasses(FLOAT_T dm12sq, FLOAT_T dm23sq) = 0;
  // propagator.hpp:213-218
  virtual void setEnergyList(const std::vector<FLOAT_T>& list) = 0;
  // propagator.hpp:223
  virtual void setChemicalComposition(const std::vector<FLOAT_T>& list) = 0;
  // propagator.hpp:227
  virtual void calculateProbabilities(NeutrinoType type) = 0;
  // propagator.hpp:248
  virtual void getProbabilityArr(FLOAT_T* probArr, ProbType t) = 0;
  // propagator.hpp:250
  virtual void setPathLength(FLOAT_T path_length) = 0;
  // propagator.hpp:252
  virtual void setDensity(FLOAT_T rho) = 0;
};

// The virtual members AtmosCpuPropagator<FLOAT_T> adds to the base
// (atmoscpupropagator.hpp:85-513).
template <class FLOAT_T>
class AtmosPropagator : public Propagator<FLOAT_T> {
 public:
  using Propagator<FLOAT_T>::setDensity;
  // atmoscpupropagator.hpp:85
  virtual int getNlayerBoundaries() = 0;
  // atmoscpupropagator.hpp:88-108
  virtual void SetNumberOfProductionHeightBinsForAveraging(int n) = 0;
  // atmoscpupropagator.hpp:116-170
  virtual void setDensity(const std::vector<FLOAT_T>& radii, const std::vector<FLOAT_T>& rhos,
                          const std::vector<FLOAT_T>& yps) = 0;
  // atmoscpupropagator.hpp:181-245
  virtual void setDensity(const std::vector<FLOAT_T>& radii, const std::vector<FLOAT_T>& a,
                          const std::vector<FLOAT_T>& b, const std::vector<FLOAT_T>& c,
                          const std::vector<FLOAT_T>& yps) = 0;
  // atmoscpupropagator.hpp:250-344
  virtual void setDensityFromFile(const std::string& filename) = 0;
  // atmoscpupropagator.hpp:368-381
  virtual void setCosineList(const std::vector<FLOAT_T>& list) = 0;
  // atmoscpupropagator.hpp:386-393
  virtual void setProductionHeight(FLOAT_T heightKM) = 0;
  // atmoscpupropagator.hpp:395-427
  virtual void setProductionHeightList(const std::vector<FLOAT_T>& list_prob,
                                       const std::vector<FLOAT_T>& list_bins) = 0;
  // atmoscpupropagator.hpp:431-444
  virtual void setMaxlayers() = 0;
  // atmoscpupropagator.hpp:492-500
  virtual FLOAT_T getProbability(int index_cosine, int index_energy, ProbType t) = 0;
};

// The virtual members BeamCpuPropagator<FLOAT_T> adds to the base
// (beamcpupropagator.hpp:123-130).
template <class FLOAT_T>
class BeamPropagator : public Propagator<FLOAT_T> {
 public:
  virtual FLOAT_T getProbability(int index_energy, ProbType t) = 0;
};

template <class FLOAT_T>
class HipPropagatorSingle : public AtmosPropagator<FLOAT_T> {
  static_assert(std::is_floating_point<FLOAT_T>::value, "FLOAT_T must be float or double");

 public:
  // AtmosCudaPropagatorSingle(int id, int n_cosines, int n_energies)
  // (atmoscudapropagator.cu:30).
  HipPropagatorSingle(int id, int n_cosines, int n_energies,
                      Precision precision = default_precision<FLOAT_T>())
      : n_cosines_(n_cosines), n_energies_(n_energies) {
    check(cp3_atmos_create(id, n_cosines, n_energies, (int)precision, &h_));
  }
  // AtmosCudaPropagatorSingle(int n_cosines, int n_energies): device 0
  // (atmoscudapropagator.cu:68).
  HipPropagatorSingle(int n_cosines, int n_energies)
      : HipPropagatorSingle(0, n_cosines, n_energies) {}
  ~HipPropagatorSingle() {
    if (h_) cp3_atmos_destroy(h_);
  }
  HipPropagatorSingle(const HipPropagatorSingle&) = delete;
  HipPropagatorSingle& operator=(const HipPropagatorSingle&) = delete;
  HipPropagatorSingle(HipPropagatorSingle&& o) noexcept
      : h_(o.h_), n_cosines_(o.n_cosines_), n_energies_(o.n_energies_),
        avg_(o.avg_) {
    o.h_ = nullptr;
  }
  HipPropagatorSingle& operator=(HipPropagatorSingle&& o) noexcept {
    if (this != &o) {
      if (h_) cp3_atmos_destroy(h_);
      h_ = o.h_;
      n_cosines_ = o.n_cosines_;
      n_energies_ = o.n_energies_;
      avg_ = o.avg_;
      o.h_ = nullptr;
    }
    return *this;
  }

  // propagator.hpp:148-183
  void setMNSMatrix(FLOAT_T theta12, FLOAT_T theta13, FLOAT_T theta23,
                    FLOAT_T dCP, int kNuType) override {
    check(cp3_atmos_set_mns_matrix(live(), theta12, theta13, theta23, dCP, kNuType));
  }
  // propagator.hpp:188-209
  void setNeutrinoMasses(FLOAT_T dm12sq, FLOAT_T dm23sq) override {
    check(cp3_atmos_set_neutrino_masses(live(), dm12sq, dm23sq));
  }
  // propagator.hpp:213-218
  void setEnergyList(const std::vector<FLOAT_T>& list) override {
    const std::vector<double> v = detail::to_double(list);
    check(cp3_atmos_set_energy_list(live(), v.data(), (int)v.size()));
  }
  // atmoscpupropagator.hpp:368-381
  void setCosineList(const std::vector<FLOAT_T>& list) override {
    const std::vector<double> v = detail::to_double(list);
    check(cp3_atmos_set_cosine_list(live(), v.data(), (int)v.size()));
  }
  // atmoscpupropagator.hpp:116-170
  void setDensity(const std::vector<FLOAT_T>& radii,
                  const std::vector<FLOAT_T>& rhos,
                  const std::vector<FLOAT_T>& yps) override {
    if (rhos.size() != radii.size())
      throw std::runtime_error("setDensity : rhos.size() != radii.size()");
    if (rhos.size() != yps.size())
      throw std::runtime_error("setDensity : rhos.size() != yps.size()");
    const std::vector<double> r = detail::to_double(radii), d = detail::to_double(rhos),
                              y = detail::to_double(yps);
    check(cp3_atmos_set_density(live(), r.data(), d.data(), y.data(), (int)r.size()));
  }
  // atmoscpupropagator.hpp:181-245: densities a R^2 + b R + c per shell,
  // averaged per trajectory segment.
  void setDensity(const std::vector<FLOAT_T>& radii, const std::vector<FLOAT_T>& a,
                  const std::vector<FLOAT_T>& b, const std::vector<FLOAT_T>& c,
                  const std::vector<FLOAT_T>& yps) override {
    if (a.size() != radii.size())
      throw std::runtime_error("setDensity : a.size() != radii.size()");
    if (a.size() != yps.size())
      throw std::runtime_error("setDensity : a.size() != yps.size()");
    const std::vector<double> r = detail::to_double(radii), va = detail::to_double(a),
                              vb = detail::to_double(b), vc = detail::to_double(c),
                              y = detail::to_double(yps);
    if (vb.size() != r.size() || vc.size() != r.size())
      throw std::runtime_error("setDensity : b.size() and c.size() must equal radii.size()");
    check(cp3_atmos_set_density_poly(live(), r.data(), va.data(), vb.data(), vc.data(),
                                     y.data(), (int)r.size()));
  }
  // atmoscpupropagator.hpp:347-357: a placeholder in the reference too.
  void setDensity(FLOAT_T rho) override {
    (void)rho;
    std::cout << "DUMMY FUNCTION: ATMOS class uses setDensity(radii, rhos, yps) "
                 "or setDensityFromFile(filename)" << std::endl;
  }
  // atmoscpupropagator.hpp:250-344 (3- and 5-column files)
  void setDensityFromFile(const std::string& filename) override {
    check(cp3_atmos_set_density_from_file(live(), filename.c_str()));
  }
  // atmoscpupropagator.hpp:481-490
  void setChemicalComposition(const std::vector<FLOAT_T>& list) override {
    const std::vector<double> v = detail::to_double(list);
    check(cp3_atmos_set_chemical_composition(live(), v.data(), (int)v.size()));
  }
  // atmoscpupropagator.hpp:386-393
  void setProductionHeight(FLOAT_T heightKM) override {
    check(cp3_atmos_set_production_height(live(), heightKM));
  }
  // atmoscpupropagator.hpp:88-108: at most 20 bins; like the reference, any
  // n >= 1 turns averaging off (setProductionHeightAveraging turns it on).
  void SetNumberOfProductionHeightBinsForAveraging(int n) override {
    check(cp3_atmos_set_production_height_bins(live(), n));
    if (n >= 1) avg_ = false;
    if (avg_)
      std::cout << "Set " << n << " Production height bins" << std::endl;
    else
      std::cout << "Using fixed production height" << std::endl;
  }
  // No counterpart in the reference class, which keeps this flag protected
  // (atmoscpupropagator.hpp:532) and never sets it: turns the averaging of
  // physics.hpp:997-1067 on or off; a list set before is dropped.
  void setProductionHeightAveraging(bool on) {
    check(cp3_atmos_set_production_height_averaging(live(), on ? 1 : 0));
    avg_ = on;
  }
  // atmoscpupropagator.hpp:395-427: bin probabilities
  // [nutype][flavour][energy][cosine][bin] and the bin edges in km; throws
  // unless averaging is on.
  void setProductionHeightList(const std::vector<FLOAT_T>& list_prob,
                               const std::vector<FLOAT_T>& list_bins) override {
    const std::vector<double> p = detail::to_double(list_prob), b = detail::to_double(list_bins);
    check(cp3_atmos_set_production_height_list(live(), p.data(), (int64_t)p.size(), b.data(),
                                               (int)b.size()));
  }
  // atmoscpupropagator.hpp:361-364: a placeholder in the reference too.
  void setPathLength(FLOAT_T) override {
    std::cout << "DUMMY FUNCTION - ATMOS class calculates PATH LENGTH" << std::endl;
  }
  // atmoscpupropagator.hpp:85
  int getNlayerBoundaries() override {
    int n = 0;
    check(cp3_atmos_get_nlayer_boundaries(live(), &n));
    return n;
  }

  // atmoscudapropagator.cu:202-205
  void calculateProbabilities(NeutrinoType type) override {
    check(cp3_atmos_calculate(live(), type == Antineutrino ? 1 : 0));
  }
  void calculateProbabilitiesAsync(NeutrinoType type) {
    check(cp3_atmos_calculate_async(live(), type == Antineutrino ? 1 : 0));
  }
  // atmoscudapropagator.cuh:114-120 (the CUDA class's names)
  void calculateAtmosProbabilitiesAsync(NeutrinoType type) { calculateProbabilitiesAsync(type); }
  void waitForCompletion() { check(cp3_atmos_wait(live())); }
  // atmoscudapropagator.cu:319-325: waits, then copies the whole result to
  // host memory (the first read of a channel copies that channel).
  void getResultFromDevice() {
    double v = 0;
    for (int t = 0; t < 9; ++t) check(cp3_atmos_get_probability(live(), 0, 0, t, &v));
  }
  // atmoscpupropagator.hpp:431-444: the shell counts are rebuilt with the
  // trajectory table whenever the cosines or the model change.
  void setMaxlayers() override {}

  // atmoscudapropagator.cu:222-235
  FLOAT_T getProbability(int index_cosine, int index_energy, ProbType t) override {
    double v = 0;
    check(cp3_atmos_get_probability(live(), index_cosine, index_energy, (int)t, &v));
    return (FLOAT_T)v;
  }
  // atmoscudapropagator.cu:238-257: energy outer, cosine inner.
  void getProbabilityArr(FLOAT_T* probArr, ProbType t) override {
    std::vector<double> tmp((size_t)n_cosines_ * (size_t)n_energies_);
    check(cp3_atmos_get_probability_arr(live(), tmp.data(), (int)t));
    for (size_t i = 0; i < tmp.size(); ++i) probArr[i] = (FLOAT_T)tmp[i];
  }

  // Device result [9][nC][nE] (double for FP64, float otherwise) and stream.
  const void* deviceResult(int* elem_bytes = nullptr) {
    const void* p = nullptr;
    check(cp3_atmos_result_device(live(), &p, elem_bytes));
    return p;
  }
  float lastKernelMs() {
    float ms = 0;
    check(cp3_atmos_last_kernel_ms(live(), &ms));
    return ms;
  }
  cp3_atmos* handle() { return h_; }

 private:
  cp3_atmos* live() {
    if (!h_)
      throw std::runtime_error(
          "CudaPropagatorSingle::calculateProbabilities. Object has been moved from.");
    return h_;
  }
  cp3_atmos* h_ = nullptr;
  int n_cosines_ = 0, n_energies_ = 0;
  bool avg_ = false;
};

template <class FLOAT_T>
class HipBeamPropagatorSingle : public BeamPropagator<FLOAT_T> {
  static_assert(std::is_floating_point<FLOAT_T>::value, "FLOAT_T must be float or double");

 public:
  // BeamCudaPropagatorSingle(int id, int n_energies, int n_threads)
  // (beamcudapropagator.cu:29).
  HipBeamPropagatorSingle(int id, int n_energies, int n_threads,
                          Precision precision = default_precision<FLOAT_T>())
      : n_energies_(n_energies) {
    (void)n_threads;
    check(cp3_beam_create(id, n_energies, (int)precision, &h_));
  }
  // BeamCudaPropagatorSingle(int n_energies, int n_threads): device 0.
  HipBeamPropagatorSingle(int n_energies, int n_threads)
      : HipBeamPropagatorSingle(0, n_energies, n_threads) {}
  ~HipBeamPropagatorSingle() {
    if (h_) cp3_beam_destroy(h_);
  }
  HipBeamPropagatorSingle(const HipBeamPropagatorSingle&) = delete;
  HipBeamPropagatorSingle& operator=(const HipBeamPropagatorSingle&) = delete;
  HipBeamPropagatorSingle(HipBeamPropagatorSingle&& o) noexcept
      : h_(o.h_), n_energies_(o.n_energies_) {
    o.h_ = nullptr;
  }
  HipBeamPropagatorSingle& operator=(HipBeamPropagatorSingle&& o) noexcept {
    if (this != &o) {
      if (h_) cp3_beam_destroy(h_);
      h_ = o.h_;
      n_energies_ = o.n_energies_;
      o.h_ = nullptr;
    }
    return *this;
  }

  void setMNSMatrix(FLOAT_T theta12, FLOAT_T theta13, FLOAT_T theta23,
                    FLOAT_T dCP, int kNuType) override {
    check(cp3_beam_set_mns_matrix(live(), theta12, theta13, theta23, dCP, kNuType));
  }
  void setNeutrinoMasses(FLOAT_T dm12sq, FLOAT_T dm23sq) override {
    check(cp3_beam_set_neutrino_masses(live(), dm12sq, dm23sq));
  }
  void setEnergyList(const std::vector<FLOAT_T>& list) override {
    const std::vector<double> v = detail::to_double(list);
    check(cp3_beam_set_energy_list(live(), v.data(), (int)v.size()));
  }
  // beamcpupropagator.hpp:84-91
  void setDensity(FLOAT_T rho) override { check(cp3_beam_set_density(live(), rho)); }
  // beamcpupropagator.hpp:112-120: a beam holds no per-row composition (its
  // Ye is 0.5, physics.hpp:1219), so only an empty list is accepted.
  void setChemicalComposition(const std::vector<FLOAT_T>& list) override {
    if (!list.empty())
      throw std::runtime_error(
          "BeamCpuPropagator::setChemicalComposition. Size of input list not equal to expectation");
  }
  void setPathLength(FLOAT_T path_length) override {
    check(cp3_beam_set_path_length(live(), path_length));
  }
  void calculateProbabilities(NeutrinoType type) override {
    check(cp3_beam_calculate(live(), type == Antineutrino ? 1 : 0));
  }
  void calculateProbabilitiesAsync(NeutrinoType type) {
    check(cp3_beam_calculate_async(live(), type == Antineutrino ? 1 : 0));
  }
  // beamcudapropagator.cuh:95-101 (the CUDA class's names)
  void calculateBeamProbabilitiesAsync(NeutrinoType type) { calculateProbabilitiesAsync(type); }
  void waitForCompletion() { check(cp3_beam_wait(live())); }
  void getResultFromDevice() {
    double v = 0;
    for (int t = 0; t < 9; ++t) check(cp3_beam_get_probability(live(), 0, t, &v));
  }
  // beamcudapropagator.cu:146-159
  FLOAT_T getProbability(int index_energy, ProbType t) override {
    double v = 0;
    check(cp3_beam_get_probability(live(), index_energy, (int)t, &v));
    return (FLOAT_T)v;
  }
  // beamcudapropagator.cu:162-178
  void getProbabilityArr(FLOAT_T* probArr, ProbType t) override {
    std::vector<double> tmp((size_t)n_energies_);
    check(cp3_beam_get_probability_arr(live(), tmp.data(), (int)t));
    for (size_t i = 0; i < tmp.size(); ++i) probArr[i] = (FLOAT_T)tmp[i];
  }
  const void* deviceResult(int* elem_bytes = nullptr) {
    const void* p = nullptr;
    check(cp3_beam_result_device(live(), &p, elem_bytes));
    return p;
  }
  float lastKernelMs() {
    float ms = 0;
    check(cp3_beam_last_kernel_ms(live(), &ms));
    return ms;
  }

 private:
  cp3_beam* live() {
    if (!h_)
      throw std::runtime_error(
          "BeamCpuPropagator::calculateProbabilities. Object has been moved from.");
    return h_;
  }
  cp3_beam* h_ = nullptr;
  int n_energies_ = 0;
};

}  // namespace cp3hip

#endif  // CUDAPROB3_AMD_HIP_PROPAGATOR_HPP
