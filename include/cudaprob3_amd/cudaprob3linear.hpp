// The reference's names for the MI355X drop-in classes, so that caller code
// written against CUDAProb3++'s GPU propagators compiles unchanged after
// swapping its include:
//
//   #include "atmoscudapropagator.cuh"        // reference
//   #include "cudaprob3_amd/cudaprob3linear.hpp"  // this library
//
// Mapped names (reference declaration -> this library):
//   cudaprob3linear::ProbType, e_e .. t_t        (types.hpp:23-27)
//       -> cp3hip::ProbType
//   cudaprob3linear::NeutrinoType, Neutrino, Antineutrino (types.hpp:29)
//       -> cp3hip::NeutrinoType
//   cudaprob3linear::Propagator<FLOAT_T>         (propagator.hpp:45-252)
//       -> cp3hip::Propagator<FLOAT_T>
//   cudaprob3linear::AtmosCudaPropagatorSingle   (atmoscudapropagator.cuh:42)
//       -> cp3hip::HipPropagatorSingle<double>
//   cudaprob3linear::BeamCudaPropagatorSingle    (beamcudapropagator.cuh:41)
//       -> cp3hip::HipBeamPropagatorSingle<double>
//
// The reference's CPU classes (AtmosCpuPropagator<FLOAT_T>,
// BeamCpuPropagator<FLOAT_T>) are not provided.  They are concrete in the
// reference, and both GPU classes derive from them, so code that holds a GPU
// object through an AtmosCpuPropagator<double> or BeamCpuPropagator<double>
// pointer names the interfaces cp3hip::AtmosPropagator<double> or
// cp3hip::BeamPropagator<double> instead (the same virtual members).
#ifndef CUDAPROB3_AMD_CUDAPROB3LINEAR_HPP
#define CUDAPROB3_AMD_CUDAPROB3LINEAR_HPP

#include "hip_propagator.hpp"

namespace cudaprob3linear {

using cp3hip::ProbType;
using cp3hip::e_e;
using cp3hip::e_m;
using cp3hip::e_t;
using cp3hip::m_e;
using cp3hip::m_m;
using cp3hip::m_t;
using cp3hip::t_e;
using cp3hip::t_m;
using cp3hip::t_t;

using cp3hip::NeutrinoType;
using cp3hip::Neutrino;
using cp3hip::Antineutrino;

template <class FLOAT_T>
using Propagator = cp3hip::Propagator<FLOAT_T>;

using AtmosCudaPropagatorSingle = cp3hip::HipPropagatorSingle<double>;
using BeamCudaPropagatorSingle = cp3hip::HipBeamPropagatorSingle<double>;

}  // namespace cudaprob3linear

#endif  // CUDAPROB3_AMD_CUDAPROB3LINEAR_HPP
