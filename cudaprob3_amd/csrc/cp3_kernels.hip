// HIP kernels for gfx950: atmospheric and beam propagation.
//
// Atmospheric work unit = one wave of 64 lanes: 64 consecutive energies of one
// job times a block of NC cosines.  Each lane owns one energy and carries NC
// trajectories.  The shell loop is the outer loop: each (energy, shell)
// eigen-solution is computed once and applied to every trajectory of the
// block that crosses the shell.  Everything that depends only on the cosine
// or the shell (segment lengths, densities, shell counts) is wave-uniform and
// read with scalar loads; the 9 probabilities are stored as 9 coalesced rows
// of the [9][nC][nE] output.
#include <hip/hip_runtime.h>

#include "cp3_launch.hpp"
#include "cp3_math.hpp"

namespace cp3 {

template <class P>
struct OutT {
  using type = float;
};
template <>
struct OutT<PolicyF64> {
  using type = double;
};

template <class P, int NC>
__device__ __forceinline__ void atmos_unit(const AtmosJob& job, int64_t local) {
  using D = typename P::D;
  using F = typename P::F;
  using O = typename OutT<P>::type;
  const GridDev& g = job.g;
  const int ntile = (g.nE + kWave - 1) / kWave;
  const int cb = (int)(local / ntile);
  const int et = (int)(local - (int64_t)cb * ntile);
  const int e = et * kWave + (int)threadIdx.x;
  const bool e_ok = e < g.nE;
  const D E = (D)g.ev[e_ok ? e : g.nE - 1];
  const D inv = D(1) / E;

  int cidx[NC];
  int Mc[NC];
  bool c_ok[NC];
  int Mmax = 0;
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    const int k = cb * NC + j;
    c_ok[j] = k < g.nC;
    cidx[j] = g.perm[c_ok[j] ? k : g.nC - 1];
    Mc[j] = g.M[cidx[j]];
    Mmax = Mc[j] > Mmax ? Mc[j] : Mmax;
  }

  Traj<P> t[NC];
#pragma unroll
  for (int j = 0; j < NC; ++j) traj_init<P>(t[j], Mc[j] == 0);

  for (int i = 0; i < Mmax; ++i) {
    Slab<P> s;
    solve_slab<P>(job.sc, E, inv, (D)g.rye[i], s);
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      if (i < Mc[j] - 1) {
        traj_middle<P>(t[j], s, (D)g.seg[(int64_t)cidx[j] * g.S + i]);
      } else if (i == Mc[j] - 1) {
        traj_inner<P>(t[j], s, (D)g.seg[(int64_t)cidx[j] * g.S + i]);
      }
    }
  }

  Slab<P> vac;
  solve_slab<P>(job.sc, E, inv, D(0), vac);
  O* out = (O*)job.out;
  const int64_t plane = (int64_t)g.nC * g.nE;
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    F prob[9];
    traj_finish<P>(t[j], vac, (D)g.Lvac[cidx[j]], job.sc, prob);
    if (e_ok && c_ok[j]) {
      O* dst = out + (int64_t)cidx[j] * g.nE + e;
#pragma unroll
      for (int ch = 0; ch < 9; ++ch) dst[ch * plane] = (O)prob[ch];
    }
  }
}

template <class P, int NC>
__global__ void __launch_bounds__(kWave) atmos_single_kernel(AtmosJob job) {
  atmos_unit<P, NC>(job, (int64_t)blockIdx.x);
}

// Finds the job owning work unit u: jobs[j].unit_begin <= u < jobs[j+1]...
__device__ __forceinline__ int find_job_atmos(const AtmosJob* jobs, int njobs,
                                              int64_t u) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].unit_begin <= u)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Persistent form: a fixed grid of waves walks the flattened work list of
// every job of the batch with a static stride.
template <class P, int NC>
__global__ void __launch_bounds__(kWave) atmos_batch_kernel(
    const AtmosJob* __restrict__ jobs, int njobs, int64_t nunits) {
  for (int64_t u = blockIdx.x; u < nunits; u += gridDim.x) {
    const int j = find_job_atmos(jobs, njobs, u);
    const AtmosJob& job = jobs[j];
    atmos_unit<P, NC>(job, u - job.unit_begin);
  }
}

template <class P>
__device__ __forceinline__ void beam_unit(const BeamJob& job, int64_t local) {
  using D = typename P::D;
  using F = typename P::F;
  using O = typename OutT<P>::type;
  const int e = (int)local * kWave + (int)threadIdx.x;
  const bool e_ok = e < job.nE;
  const D E = (D)job.ev[e_ok ? e : job.nE - 1];
  F prob[9];
  beam_point<P>(job.sc, E, (D)job.rye, (D)job.L, prob);
  if (e_ok) {
    O* out = (O*)job.out;
#pragma unroll
    for (int ch = 0; ch < 9; ++ch) out[(int64_t)ch * job.nE + e] = (O)prob[ch];
  }
}

template <class P>
__global__ void __launch_bounds__(kWave) beam_single_kernel(BeamJob job) {
  beam_unit<P>(job, (int64_t)blockIdx.x);
}

template <class P>
__global__ void __launch_bounds__(kWave) beam_batch_kernel(
    const BeamJob* __restrict__ jobs, int njobs, int64_t nunits) {
  for (int64_t u = blockIdx.x; u < nunits; u += gridDim.x) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (jobs[mid].unit_begin <= u)
        lo = mid;
      else
        hi = mid - 1;
    }
    const BeamJob& job = jobs[lo];
    beam_unit<P>(job, u - job.unit_begin);
  }
}

// ------------------------------------------------------------- launchers

template <class P>
static hipError_t atmos_single_nc(int nc, const AtmosJob& job,
                                  hipStream_t st) {
  const int64_t n = atmos_units(job.g, nc);
  if (n <= 0) return hipSuccess;
  if (n > 0x7fffffff) return hipErrorInvalidConfiguration;
  const dim3 grid((unsigned)n), block(kWave);
  switch (nc) {
    case 1: atmos_single_kernel<P, 1><<<grid, block, 0, st>>>(job); break;
    case 2: atmos_single_kernel<P, 2><<<grid, block, 0, st>>>(job); break;
    case 4: atmos_single_kernel<P, 4><<<grid, block, 0, st>>>(job); break;
    case 8: atmos_single_kernel<P, 8><<<grid, block, 0, st>>>(job); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_atmos_single(int precision, int nc, const AtmosJob& job,
                               hipStream_t st) {
  switch (precision) {
    case kFP64: return atmos_single_nc<PolicyF64>(nc, job, st);
    case kFP32: return atmos_single_nc<PolicyF32>(nc, job, st);
    case kMixed: return atmos_single_nc<PolicyMixed>(nc, job, st);
    default: return hipErrorInvalidValue;
  }
}

template <class P>
static hipError_t atmos_batch_nc(int nc, const AtmosJob* jobs, int njobs,
                                 int64_t nunits, int grid, hipStream_t st) {
  if (nunits <= 0 || njobs <= 0) return hipSuccess;
  const dim3 g((unsigned)grid), b(kWave);
  switch (nc) {
    case 1: atmos_batch_kernel<P, 1><<<g, b, 0, st>>>(jobs, njobs, nunits); break;
    case 2: atmos_batch_kernel<P, 2><<<g, b, 0, st>>>(jobs, njobs, nunits); break;
    case 4: atmos_batch_kernel<P, 4><<<g, b, 0, st>>>(jobs, njobs, nunits); break;
    case 8: atmos_batch_kernel<P, 8><<<g, b, 0, st>>>(jobs, njobs, nunits); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_atmos_batch(int precision, int nc, const AtmosJob* jobs,
                              int njobs, int64_t nunits, int grid,
                              hipStream_t st) {
  if (grid <= 0) return hipErrorInvalidConfiguration;
  switch (precision) {
    case kFP64: return atmos_batch_nc<PolicyF64>(nc, jobs, njobs, nunits, grid, st);
    case kFP32: return atmos_batch_nc<PolicyF32>(nc, jobs, njobs, nunits, grid, st);
    case kMixed: return atmos_batch_nc<PolicyMixed>(nc, jobs, njobs, nunits, grid, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_beam_single(int precision, const BeamJob& job,
                              hipStream_t st) {
  const int64_t n = beam_units(job.nE);
  if (n <= 0) return hipSuccess;
  const dim3 grid((unsigned)n), block(kWave);
  switch (precision) {
    case kFP64: beam_single_kernel<PolicyF64><<<grid, block, 0, st>>>(job); break;
    case kFP32: beam_single_kernel<PolicyF32><<<grid, block, 0, st>>>(job); break;
    case kMixed: beam_single_kernel<PolicyMixed><<<grid, block, 0, st>>>(job); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_beam_batch(int precision, const BeamJob* jobs, int njobs,
                             int64_t nunits, int grid, hipStream_t st) {
  if (nunits <= 0 || njobs <= 0) return hipSuccess;
  if (grid <= 0) return hipErrorInvalidConfiguration;
  const dim3 g((unsigned)grid), b(kWave);
  switch (precision) {
    case kFP64: beam_batch_kernel<PolicyF64><<<g, b, 0, st>>>(jobs, njobs, nunits); break;
    case kFP32: beam_batch_kernel<PolicyF32><<<g, b, 0, st>>>(jobs, njobs, nunits); break;
    case kMixed: beam_batch_kernel<PolicyMixed><<<g, b, 0, st>>>(jobs, njobs, nunits); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <class P>
static hipError_t atmos_occ(int nc, int* n) {
  switch (nc) {
    case 1: return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, atmos_batch_kernel<P, 1>, kWave, 0);
    case 2: return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, atmos_batch_kernel<P, 2>, kWave, 0);
    case 4: return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, atmos_batch_kernel<P, 4>, kWave, 0);
    case 8: return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, atmos_batch_kernel<P, 8>, kWave, 0);
    default: return hipErrorInvalidValue;
  }
}

hipError_t atmos_batch_occupancy(int precision, int nc, int* n) {
  switch (precision) {
    case kFP64: return atmos_occ<PolicyF64>(nc, n);
    case kFP32: return atmos_occ<PolicyF32>(nc, n);
    case kMixed: return atmos_occ<PolicyMixed>(nc, n);
    default: return hipErrorInvalidValue;
  }
}

hipError_t beam_batch_occupancy(int precision, int* n) {
  switch (precision) {
    case kFP64: return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, beam_batch_kernel<PolicyF64>, kWave, 0);
    case kFP32: return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, beam_batch_kernel<PolicyF32>, kWave, 0);
    case kMixed: return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, beam_batch_kernel<PolicyMixed>, kWave, 0);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace cp3
