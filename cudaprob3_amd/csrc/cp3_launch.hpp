// Device-side job descriptors and the kernel launchers (cp3_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cp3_math.hpp"

namespace cp3 {

// Output precision forms.
enum Precision : int { kFP64 = 0, kFP32 = 1, kMixed = 2 };

inline int out_elem_bytes(int precision) { return precision == kFP64 ? 8 : 4; }

// One resident atmospheric grid: energies, Earth model and cosine list
// reduced to the trajectory table (cp3_host.hpp, TrajTable).
struct GridDev {
  const double* ev;     // [nE] GeV
  const int32_t* perm;  // [nC] cosine indices by decreasing shell count
  const int32_t* M;     // [nC] shells crossed
  const double* Lvac;   // [nC] atmosphere segment, km
  const double* seg;    // [nC * S] km per shell
  const double* rye;    // [S] rho * Ye per shell, or [nC * S] (rye_cs = S)
  // row stride of rye per cosine: 0, or S for a polynomial model's
  // per-trajectory densities (shared solve mode with cosine block 1 and one
  // wave per workgroup only: those solve per cosine)
  int32_t nC, nE, S, rye_cs;
};

// Production-height averaging of a single-grid job: the bin probabilities of
// the job's neutrino type, [3][nb][nC][nE] (initial flavour, bin, cosine,
// energy; element type of the output), and per cosine the bins' mid
// path-length shifts and widths in km, [nC][nb][2].  Read only by the
// averaging kernels.
struct PhaDev {
  const void* w;
  const double* hmw;
  int32_t nb, pad;
};

// One atmospheric evaluation: a parameter set and neutrino type on a grid.
// Output layout [9][nC][nE] (channel a*3+b = P(a->b)), element type double
// for kFP64 and float otherwise.
struct AtmosJob {
  SetConst sc;
  GridDev g;
  void* out;
  int64_t unit_begin;  // first work unit of this job in a batch
  // 64-energy tiles this job's own work units cover: all of them
  // (energy_tiles), or only the full ones when the last nE % 64 energies
  // run in a packed tail tile (TailGroup)
  int32_t etiles, pad;
  // Slab-table mode (batch engine): the pre-solved eigen-solutions of this
  // job, Slab<P> entries [g.S + 1][energy_tiles(g) * 64] (row g.S: the
  // vacuum).
  void* slabs;
  PhaDev pha;
};

// Packed tail tiles: the last `tail` = nE % 64 energies of up to `per` =
// 64 / tail jobs that share one grid, side by side in the lanes of one wave
// (lane l: job jobs[t * per + l / tail], energy nE - tail + l % tail), so the
// partial last tile of each job does not leave 64 - tail lanes idle.
// Cosine-dependent data stay wave-uniform; only the parameter set and the
// output array differ between lanes.
struct TailGroup {
  GridDev g;
  const int32_t* jobs;  // [njobs] indices into the AtmosJob table
  int32_t njobs, tail, per, pad;
  int64_t unit_begin;   // first work unit of this group in the tail launch
};

// One beam evaluation: a parameter set and neutrino type on an energy list
// through one constant-density layer.  Output layout [9][nE].
struct BeamJob {
  SetConst sc;
  const double* ev;
  double rye;  // rho * 0.5 (physics.hpp:1219)
  double L;    // km
  void* out;
  int64_t unit_begin;
  int32_t nE, pad;
};

constexpr int kWave = 64;

// Cosine blocks of one grid (w waves x nc cosines each) and 64-energy tiles.
CP3_HD int64_t cosine_blocks(const GridDev& g, int nc, int w) {
  return (g.nC + nc * w - 1) / (nc * w);
}
CP3_HD int32_t energy_tiles(const GridDev& g) { return (g.nE + kWave - 1) / kWave; }
// A job's slab table: rows (shells, then the vacuum) of energy_tiles * 64
// entries; the pre-pass solves its (S + 1) x nE used entries in waves of 64.
CP3_HD int64_t slab_entries(const GridDev& g) {
  return (int64_t)(g.S + 1) * energy_tiles(g) * kWave;
}
CP3_HD int64_t slab_waves(const GridDev& g) {
  return ((int64_t)(g.S + 1) * g.nE + kWave - 1) / kWave;
}

// Work units of one job over all its energy tiles: (cosine block) x (tile).
inline int64_t atmos_units(const GridDev& g, int nc, int w = 1) {
  return cosine_blocks(g, nc, w) * energy_tiles(g);
}
// Work units of a tail group: (cosine block) x (packed tile).
inline int64_t tail_units(const TailGroup& t, int nc, int w) {
  return cosine_blocks(t.g, nc, w) * ((t.njobs + t.per - 1) / t.per);
}
inline int64_t beam_units(int nE) { return (nE + kWave - 1) / kWave; }

// Supported cosine-block sizes and waves per workgroup.
inline bool valid_nc(int nc) { return nc == 1 || nc == 2 || nc == 4 || nc == 8; }
inline bool valid_group(int w) { return w == 1 || w == 2 || w == 4; }

hipError_t launch_atmos_single(int precision, int nc, int w, const AtmosJob& job,
                               hipStream_t stream);
hipError_t launch_atmos_batch(int precision, int nc, int w, const AtmosJob* d_jobs,
                              int njobs, int64_t nunits, int grid,
                              hipStream_t stream);
// Slab-table mode: bytes per table entry (sizeof(Slab<P>)), the pre-pass
// over every job's table (max_waves 64-entry tiles per job, max_waves the
// largest slab_waves, a few tiles per one-wave workgroup; at most 65535
// jobs), and the propagation launches (one wave per workgroup) reading the
// tables.
size_t slab_entry_bytes(int precision);
// The single grid's pre-pass (job by value), which also stores the job
// descriptor at d_job for the propagation launch that follows.
hipError_t launch_slab_table_single(int precision, const AtmosJob& job, AtmosJob* d_job,
                                   hipStream_t stream);
hipError_t launch_slab_table(int precision, const AtmosJob* d_jobs, int njobs,
                             int64_t max_waves, hipStream_t stream);
hipError_t launch_atmos_table(int precision, int nc, const AtmosJob* d_jobs, int njobs,
                              int64_t nunits, int grid, hipStream_t stream);
hipError_t launch_tail_table(int precision, int nc, const AtmosJob* d_jobs,
                             const TailGroup* d_groups, int ngroups, int64_t nunits,
                             hipStream_t stream);
hipError_t atmos_table_occupancy(int precision, int nc, int* blocks_per_cu);

// Production-height averaging (job.pha) on a single grid, one cosine per
// wave: the in-kernel solve (one workgroup per unit) or, after
// launch_slab_table_single, the slab-table form reading the job at d_job
// (grid a multiple of 8).
hipError_t launch_atmos_single_avg(int precision, const AtmosJob& job,
                                   hipStream_t stream);
hipError_t launch_atmos_table_avg(int precision, const AtmosJob* d_job, int64_t nunits,
                                  int grid, hipStream_t stream);

// The packed tail tiles of a batch: one workgroup per unit.
hipError_t launch_atmos_tail(int precision, int nc, int w, const AtmosJob* d_jobs,
                             const TailGroup* d_groups, int ngroups, int64_t nunits,
                             hipStream_t stream);
hipError_t launch_beam_single(int precision, const BeamJob& job,
                              hipStream_t stream);
// The batched beam kernel on a (128-energy groups of the longest job) x (job)
// grid; max_job_units counts the longest job's 64-energy units (beam_units).
// No persistent loop: it would hold the solve's constants in VGPRs.
hipError_t launch_beam_batch(int precision, const BeamJob* d_jobs, int njobs,
                             int64_t max_job_units, hipStream_t stream);

// Resident 64-lane workgroups per CU for the batched kernels (occupancy).
hipError_t atmos_batch_occupancy(int precision, int nc, int w, int* blocks_per_cu);
hipError_t beam_batch_occupancy(int precision, int* blocks_per_cu);

}  // namespace cp3
