// HIP kernels for gfx950: atmospheric and beam propagation.
//
// Atmospheric work unit = one workgroup of W waves: 64 consecutive energies of
// one job times W blocks of NC cosines (cosines ordered by decreasing shell
// count, so a unit's trajectories cross similar depths).  Each lane owns one
// energy and carries NC trajectories.  The shell loop is the outer loop: each
// (energy, shell) eigen-solution is computed once per unit and applied to
// every trajectory that crosses the shell.  With W > 1 the waves take turns
// solving the shells of a chunk of W shells and publish the solutions in LDS
// (double-buffered, one barrier per chunk), so a solve serves W * NC
// trajectories.  Everything that depends only on the cosine or the shell
// (segment lengths, densities, shell counts) is wave-uniform and read with
// scalar loads; the 9 probabilities are stored as 9 coalesced rows of the
// [9][nC][nE] output.
#include <hip/hip_runtime.h>

#include <algorithm>

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

// Reads of a grid's tables (perm, M, seg, rye, Lvac) at wave-uniform indices.
// No kernel writes them, so they are read through the constant address
// space: scalar loads into SGPRs, with the index arithmetic on the scalar unit
// instead of per-lane 64-bit address math.
template <class T>
__device__ __forceinline__ T uload(const T* p, int64_t i) {
  return ((const __attribute__((address_space(4))) T*)p)[i];
}

// Per-lane accesses through the global address space.  Pointers taken from
// job descriptors are generic, and a generic (flat) access counts against the
// scalar-memory/LDS wait counter as well as the vector-memory one: every wait
// for a scalar load (a segment length, a shell count) would also wait for the
// slab-row prefetch and the output stores in flight.  A global access counts
// against the vector-memory counter only.  No pointer so accessed points into
// LDS or private memory (they are device buffers, the static scratch array or
// a caller's device-accessible output), which is all a global access needs.
template <class T>
__device__ __forceinline__ T gload(const T* p, int64_t i) {
  return ((const __attribute__((address_space(1))) T*)p)[i];
}
template <class T>
__device__ __forceinline__ void gstore(T* p, int64_t i, T v) {
  ((__attribute__((address_space(1))) T*)p)[i] = v;
}

// Shell i as a middle shell of all NC trajectories of a wave (i < Mmin - 1,
// Mmin the smallest shell count among them): branch-free, so the NC
// independent update chains interleave.
template <class P, int NC>
__device__ __forceinline__ void middle_all(Traj<P> (&t)[NC], const Slab<P>& s,
                                           const GridDev& g, const int (&cidx)[NC],
                                           int i) {
  using D = typename P::D;
#pragma unroll
  for (int j = 0; j < NC; ++j)
    traj_middle<P>(t[j], s, (D)uload(g.seg, (int64_t)cidx[j] * g.S + i));
}

// Shell i for trajectories that may end at or before it: a middle shell of
// those with more shells, the innermost of those with exactly i + 1.
template <class P, int NC>
__device__ __forceinline__ void apply_shell(Traj<P> (&t)[NC], const Slab<P>& s,
                                            const GridDev& g, const int (&cidx)[NC],
                                            const int (&Mc)[NC], int i) {
  using D = typename P::D;
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    const D L = (D)uload(g.seg, (int64_t)cidx[j] * g.S + i);
    if (i < Mc[j] - 1) {
      traj_middle<P>(t[j], s, L);
    } else if (i == Mc[j] - 1) {
      traj_inner<P>(t[j], s, L);
    }
  }
}

// Shell i as a middle shell of the trajectories with more than i + 1 shells
// (the others have reached their innermost shell, taken later).
template <class P, int NC>
__device__ __forceinline__ void middle_some(Traj<P> (&t)[NC], const Slab<P>& s,
                                            const GridDev& g, const int (&cidx)[NC],
                                            const int (&Mc)[NC], int i) {
  using D = typename P::D;
#pragma unroll
  for (int j = 0; j < NC; ++j)
    if (i < Mc[j] - 1) traj_middle<P>(t[j], s, (D)uload(g.seg, (int64_t)cidx[j] * g.S + i));
}

// LDS staging of one chunk of W slabs, two buffers; struct of arrays with one
// entry per lane so every access is a conflict-free contiguous row.
template <class P, int W>
struct SlabStage {
  typename P::F v[2][W][6][kWave];
  typename P::D om[2][W][2][kWave];
};

template <class P, int W>
__device__ __forceinline__ void stage_put(SlabStage<P, W>& st, int buf, int w,
                                          int lane, const Slab<P>& s) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    st.v[buf][w][k][lane] = s.va[k];
    st.v[buf][w][3 + k][lane] = s.vb[k];
  }
  st.om[buf][w][0][lane] = s.oma;
  st.om[buf][w][1][lane] = s.omb;
}

template <class P, int W>
__device__ __forceinline__ void stage_get(const SlabStage<P, W>& st, int buf,
                                          int w, int lane, Slab<P>& s) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    s.va[k] = st.v[buf][w][k][lane];
    s.vb[k] = st.v[buf][w][3 + k][lane];
  }
  s.oma = st.om[buf][w][0][lane];
  s.omb = st.om[buf][w][1][lane];
}

// Dummy staging type for W == 1 (no LDS).
struct NoStage {};

// Wave `wave`'s share of chunk c: the solve of shell c * W + wave, if the
// unit crosses it, published in LDS buffer buf.  Shell Mwg stands for the
// vacuum (the atmosphere segment), solved once and shared like the others.
template <class P, int W>
__device__ __forceinline__ void stage_chunk(SlabStage<P, W>& st, int buf, int c,
                                            int wave, int Mwg, const SetConst& sc,
                                            typename P::D E, typename P::D inv,
                                            const GridDev& g, int lane) {
  using D = typename P::D;
  const int is = c * W + wave;
  if (is <= Mwg) {
    const D rye = (D)uload(g.rye, is < Mwg ? is : 0);
    Slab<P> s;
    solve_slab<P>(sc, E, inv, is < Mwg ? rye : D(0), s);
    stage_put<P, W>(st, buf, wave, lane, s);
  }
}

// The NC cosines of a wave (perm positions base .. base + NC - 1): table
// rows, shell counts (0 past the end of the list) and their extremes.
template <int NC>
struct CosSet {
  int cidx[NC];
  int Mc[NC];
  bool ok[NC];
  int Mmin, Mmax;
};

template <int NC>
__device__ __forceinline__ void cos_set(const GridDev& g, int base, CosSet<NC>& cs) {
  cs.Mmax = 0;
  cs.Mmin = 1 << 30;
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    const int k = base + j;
    cs.ok[j] = k < g.nC;
    cs.cidx[j] = uload(g.perm, cs.ok[j] ? k : g.nC - 1);
    cs.Mc[j] = cs.ok[j] ? uload(g.M, cs.cidx[j]) : 0;
    cs.Mmax = cs.Mc[j] > cs.Mmax ? cs.Mc[j] : cs.Mmax;
    cs.Mmin = cs.Mc[j] < cs.Mmin ? cs.Mc[j] : cs.Mmin;
  }
}

// Production-height bins of one lane's trajectory (PhaDev): the cosine's
// shifts through scalar loads, the lane's bin probabilities as coalesced rows.
template <class P>
struct DevBins {
  using D = typename P::D;
  using O = typename OutT<P>::type;
  int n;
  const double* hmw;  // this cosine's [nb][2]
  const O* w;         // this (cosine, energy) entry of [3][nb][nC][nE]
  int64_t plane;
  __device__ DevBins(const PhaDev& p, const GridDev& g, int c, int e)
      : n(p.nb),
        hmw(p.hmw + (int64_t)c * p.nb * 2),
        w((const O*)p.w + (int64_t)c * g.nE + e),
        plane((int64_t)g.nC * g.nE) {}
  __device__ void shift(int bin, D& hm, D& hw) const {
    hm = (D)uload(hmw, 2 * bin);
    hw = (D)uload(hmw, 2 * bin + 1);
  }
  __device__ O weight(int a, int bin) const { return gload(w, ((int64_t)a * n + bin) * plane); }
};

// The atmosphere segment and the nine probabilities of trajectory j of a
// cosine set, stored as rows of the [9][nC][nE] output (AVG: averaged over
// the production-height bins of pha).
template <class P, int NC, bool AVG = false>
__device__ __forceinline__ void finish_store_one(const Traj<P>& t, int j, const Slab<P>& vac,
                                                 const GridDev& g, const CosSet<NC>& cs,
                                                 const SetConst& sc, void* outp, int e,
                                                 bool e_ok, const PhaDev& pha) {
  using D = typename P::D;
  using F = typename P::F;
  using O = typename OutT<P>::type;
  O* out = (O*)outp;
  const int64_t plane = (int64_t)g.nC * g.nE;
  F prob[9];
  if constexpr (AVG) {
    const DevBins<P> bins(pha, g, cs.cidx[j], e_ok ? e : g.nE - 1);
    traj_finish_avg<P>(t, vac, (D)uload(g.Lvac, cs.cidx[j]), sc, bins, prob);
  } else {
    (void)pha;
    traj_finish<P>(t, vac, (D)uload(g.Lvac, cs.cidx[j]), sc, prob);
  }
  if (e_ok && cs.ok[j]) {
    O* dst = out + (int64_t)cs.cidx[j] * g.nE + e;
#pragma unroll
    for (int ch = 0; ch < 9; ++ch) gstore(dst, ch * plane, (O)prob[ch]);
  }
}

// The same for all NC trajectories.
template <class P, int NC, bool AVG = false>
__device__ __forceinline__ void finish_store(const Traj<P> (&t)[NC], const Slab<P>& vac,
                                             const GridDev& g, const CosSet<NC>& cs,
                                             const SetConst& sc, void* outp, int e,
                                             bool e_ok, const PhaDev& pha) {
#pragma unroll
  for (int j = 0; j < NC; ++j)
    finish_store_one<P, NC, AVG>(t[j], j, vac, g, cs, sc, outp, e, e_ok, pha);
}

// The trajectories of cosine block sb of grid g for one energy per lane:
// lane energy index e (stored only if e_ok), the lane's parameter set sc and
// output array outp ([9][nC][nE]).  sc and outp are wave-uniform for a job's
// own work units and per lane in packed tail tiles.
template <class P, int NC, int W, bool AVG = false, class Stage>
__device__ __forceinline__ void atmos_block(const GridDev& g, int sb, int e, bool e_ok,
                                            const SetConst& sc, void* outp,
                                            const PhaDev& pha, Stage& st) {
  using D = typename P::D;
  const int lane = (int)threadIdx.x % kWave;
  const int wave = W == 1 ? 0 : __builtin_amdgcn_readfirstlane((int)threadIdx.x / kWave);
  const D E = (D)gload(g.ev, e_ok ? e : g.nE - 1);
  const D inv = m_recip(E);

  const int base = sb * (W * NC);
  CosSet<NC> cs;
  cos_set<NC>(g, base + wave * NC, cs);
  const int(&cidx)[NC] = cs.cidx;
  const int(&Mc)[NC] = cs.Mc;
  const int Mmin = cs.Mmin, Mmax = cs.Mmax;

  Traj<P> t[NC];
#pragma unroll
  for (int j = 0; j < NC; ++j) traj_init<P>(t[j], Mc[j] == 0);

  // Shells below Mmin - 1 run in a loop of their own: joining the
  // branch-free path with the general one inside a single loop costs
  // register copies of the whole trajectory state at every shell.
  Slab<P> vac;
  if constexpr (W == 1) {
    (void)st;
    // per-cosine densities (rye_cs = S) come with NC == 1
    const double* rye = g.rye + (int64_t)g.rye_cs * cidx[0];
    int i = 0;
    for (; i < Mmin - 1; ++i) {
      Slab<P> s;
      solve_slab<P>(sc, E, inv, (D)uload(rye, i), s);
      middle_all<P, NC>(t, s, g, cidx, i);
    }
    for (; i < Mmax; ++i) {
      Slab<P> s;
      solve_slab<P>(sc, E, inv, (D)uload(rye, i), s);
      apply_shell<P, NC>(t, s, g, cidx, Mc, i);
    }
    solve_slab<P>(sc, E, inv, D(0), vac);
  } else {
    // deepest trajectory of the unit: its first cosine (decreasing order)
    const int Mwg = uload(g.M, uload(g.perm, base));
    // Mwg shells and the vacuum
    const int nchunk = Mwg / W + 1;
    // chunks of W shells that are middle shells of every trajectory of this
    // wave; nfast < nchunk since Mmin <= Mwg.  Every wave passes nchunk
    // barriers whatever its own split.
    const int nfast = Mmin > 1 ? (Mmin - 1) / W : 0;
    int buf = 0;
    int c = 0;
    for (; c < nfast; ++c) {
      stage_chunk<P, W>(st, buf, c, wave, Mwg, sc, E, inv, g, lane);
      __syncthreads();
#pragma unroll
      for (int w = 0; w < W; ++w) {
        Slab<P> s;
        stage_get<P, W>(st, buf, w, lane, s);
        middle_all<P, NC>(t, s, g, cidx, c * W + w);
      }
      buf ^= 1;
    }
    for (; c < nchunk; ++c) {
      stage_chunk<P, W>(st, buf, c, wave, Mwg, sc, E, inv, g, lane);
      __syncthreads();
#pragma unroll
      for (int w = 0; w < W; ++w) {
        const int i = c * W + w;
        if (i < Mmax) {
          Slab<P> s;
          stage_get<P, W>(st, buf, w, lane, s);
          apply_shell<P, NC>(t, s, g, cidx, Mc, i);
        }
      }
      buf ^= 1;
    }
    // the vacuum slab, staged as shell Mwg of chunk Mwg / W
    stage_get<P, W>(st, (Mwg / W) & 1, Mwg % W, lane, vac);
    // the next unit of a persistent workgroup restarts at buffer 0: every
    // wave must be done reading this unit's last chunk first
    __syncthreads();
  }

  finish_store<P, NC, AVG>(t, vac, g, cs, sc, outp, e, e_ok, pha);
}

// A work unit of one job: cosine block x one of its own 64-energy tiles.
template <class P, int NC, int W, bool AVG = false, class Stage>
__device__ __forceinline__ void atmos_unit(const AtmosJob& job, int64_t local,
                                           Stage& st) {
  const int lane = (int)threadIdx.x % kWave;
  const int sb = (int)(local / job.etiles);
  const int et = (int)(local - (int64_t)sb * job.etiles);
  const int e = et * kWave + lane;
  atmos_block<P, NC, W, AVG>(job.g, sb, e, e < job.g.nE, job.sc, job.out, job.pha, st);
}

// A work unit of a tail group: cosine block x packed tile t, whose lanes
// hold the last tg.tail energies of tg.per jobs.
template <class P, int NC, int W, class Stage>
__device__ __forceinline__ void atmos_tail_unit(const AtmosJob* jobs,
                                                const TailGroup& tg, int64_t local,
                                                Stage& st) {
  const int lane = (int)threadIdx.x % kWave;
  const int ntt = (tg.njobs + tg.per - 1) / tg.per;
  const int sb = (int)(local / ntt);
  const int t = (int)(local - (int64_t)sb * ntt);
  const int q = lane / tg.tail;
  const int slot = t * tg.per + q;
  const bool ok = q < tg.per && slot < tg.njobs;
  const int e = tg.g.nE - tg.tail + (lane - q * tg.tail);
  const AtmosJob& job = jobs[gload(tg.jobs, ok ? slot : 0)];
  atmos_block<P, NC, W>(tg.g, sb, e, ok, job.sc, job.out, job.pha, st);
}

// Minimum waves per SIMD asked of each atmospheric variant, i.e. a cap of
// 512 / waves VGPRs per lane (0: the compiler's own allocation).  Chosen from
// sweeps of every (precision, cosine block, group, grid) on the 1024-set,
// 59-shell batch on MI355X: a third wave hides the LDS, scalar-load and
// barrier latency of the mixed NC = 1 and 4 variants (NC = 4, group 4:
// 12.5 -> 10.6 ms), two waves instead of one help FP64 NC = 4; elsewhere
// the cap costs more in spills than it gains.
template <class P, int NC>
struct MinWaves {
  static constexpr int value = 0;
};
template <int NC>
struct MinWaves<PolicyMixed, NC> {
  static constexpr int value = (NC == 1 || NC == 4) ? 3 : 0;
};
template <int NC>
struct MinWaves<PolicyF64, NC> {
  static constexpr int value = NC == 1 ? 3 : (NC == 4 ? 2 : 0);
};
template <int NC>
struct MinWaves<PolicyF32, NC> {
  static constexpr int value = NC == 8 ? 3 : 0;
};

template <class P, int W>
struct StageOf {
  using type = SlabStage<P, W>;
};
template <class P>
struct StageOf<P, 1> {
  using type = NoStage;
};

template <class P, int NC, int W>
__global__ void __launch_bounds__(kWave * W, (MinWaves<P, NC>::value))
atmos_single_kernel(AtmosJob job) {
  __shared__ typename StageOf<P, W>::type st;
  atmos_unit<P, NC, W>(job, (int64_t)blockIdx.x, st);
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

// Batched form over the flattened work list of every job: one workgroup per
// work unit, or a persistent grid walking the list with a static stride.
template <class P, int NC, int W>
__global__ void __launch_bounds__(kWave * W, (MinWaves<P, NC>::value)) atmos_batch_kernel(
    const AtmosJob* __restrict__ jobs, int njobs, int64_t nunits) {
  __shared__ typename StageOf<P, W>::type st;
  for (int64_t u = blockIdx.x; u < nunits; u += gridDim.x) {
    const int j = find_job_atmos(jobs, njobs, u);
    const AtmosJob& job = jobs[j];
    atmos_unit<P, NC, W>(job, u - job.unit_begin, st);
  }
}

// The packed tail tiles of a batch (TailGroup), one workgroup per unit.
template <class P, int NC, int W>
__global__ void __launch_bounds__(kWave * W, (MinWaves<P, NC>::value)) atmos_tail_kernel(
    const AtmosJob* __restrict__ jobs, const TailGroup* __restrict__ groups, int ngroups,
    int64_t nunits) {
  __shared__ typename StageOf<P, W>::type st;
  for (int64_t u = blockIdx.x; u < nunits; u += gridDim.x) {
    int lo = 0, hi = ngroups - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (groups[mid].unit_begin <= u)
        lo = mid;
      else
        hi = mid - 1;
    }
    const TailGroup& tg = groups[lo];
    atmos_tail_unit<P, NC, W>(jobs, tg, u - tg.unit_begin, st);
  }
}

// ------------------------------------------------------------ slab table
//
// The batch engine can solve every (job, shell, energy) slab once, in a
// launch of its own, into a table in HBM; the propagation kernels then read
// each slab with coalesced row loads instead of solving it in every work
// unit (a work unit shares a solve among its own trajectories, the table
// among every cosine of the grid).  Table of a job: Slab<P> rows
// [g.S + 1][energy_tiles(g) * 64], row g.S the vacuum, stored by tile as
// TileIO lays it out.

// The 64 entries of one (row, energy tile) of a slab table, word-major: the
// first 16 bytes of every lane's entry, then the next 16 bytes, and a last
// part of 8 bytes likewise, so that each load or store instruction of a
// wave touches one contiguous block (1 KiB, or 512 bytes for the last part)
// instead of every line of the tile's span.
template <class T>
struct TileIO {
  static_assert(sizeof(T) % 8 == 0, "slab entries are whole 8-byte words");
  static constexpr int kQuads = (int)(sizeof(T) / 16);
  static constexpr bool kHalf = sizeof(T) % 16 != 0;
  static constexpr int64_t kTileBytes = (int64_t)sizeof(T) * kWave;
  // clang vector types (not HIP's vector classes) so that the accesses can
  // go through gload / gstore
  typedef uint32_t W4 __attribute__((ext_vector_type(4)));
  typedef uint32_t W2 __attribute__((ext_vector_type(2)));
  __device__ static T load(const char* tile, int lane) {
    uint32_t w[sizeof(T) / 4];
#pragma unroll
    for (int k = 0; k < kQuads; ++k) {
      const W4 v = gload((const W4*)(tile + k * 16 * kWave), lane);
      w[4 * k] = v.x;
      w[4 * k + 1] = v.y;
      w[4 * k + 2] = v.z;
      w[4 * k + 3] = v.w;
    }
    if constexpr (kHalf) {
      const W2 v = gload((const W2*)(tile + kQuads * 16 * kWave), lane);
      w[4 * kQuads] = v.x;
      w[4 * kQuads + 1] = v.y;
    }
    T t;
    __builtin_memcpy(&t, w, sizeof(T));
    return t;
  }
  __device__ static void store(char* tile, int lane, const T& t) {
    uint32_t w[sizeof(T) / 4];
    __builtin_memcpy(w, &t, sizeof(T));
#pragma unroll
    for (int k = 0; k < kQuads; ++k) {
      const W4 v = {w[4 * k], w[4 * k + 1], w[4 * k + 2], w[4 * k + 3]};
      gstore((W4*)(tile + k * 16 * kWave), lane, v);
    }
    if constexpr (kHalf) {
      const W2 v = {w[4 * kQuads], w[4 * kQuads + 1]};
      gstore((W2*)(tile + kQuads * 16 * kWave), lane, v);
    }
  }
};

// The tile block of row i and energy e of a job's table (lane e % 64 of it).
template <class P>
__device__ __forceinline__ const char* slab_tile(const void* slabs, const GridDev& g, int i,
                                                 int e) {
  return (const char*)slabs +
         ((int64_t)i * energy_tiles(g) + e / kWave) * TileIO<Slab<P>>::kTileBytes;
}

// Pre-pass: the (S + 1) x nE used entries of each job's table, 64
// consecutive (row, energy) pairs per wave, row-major, so that no lane solves
// a padding column.  K consecutive such waves w0 .. w0 + K - 1 of job `job`
// in one wave: every tile's indices and loads first, so that the later
// tiles' loads are in flight during the first solve, and the set constants
// copied once (the tile stores would otherwise make the compiler reload them
// from the job for every tile); then the solves and stores in tile order.
// A lane past the job's last wave or last row solves a valid pair and does
// not store.
template <class P, int K>
__device__ __forceinline__ void slab_tiles(const AtmosJob& job, int64_t w0) {
  using D = typename P::D;
  const int lane = (int)threadIdx.x;
  const GridDev& g = job.g;
  const int nE = g.nE, S = g.S;
  const int64_t nw = slab_waves(g);
  char* const slabs = (char*)job.slabs;
  const int64_t tiles = energy_tiles(g);
  int ii[K], ee[K];
  bool ok[K];
  D E[K], rye[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int64_t w = w0 + k;
    // the wave's first pair (row i0, energy e0), wave-uniform
    const int64_t first = (w < nw ? w : 0) * kWave;
    const int i0 = (int)(first / nE);
    int i = i0;
    int e = (int)(first - (int64_t)i0 * nE) + lane;
    while (e >= nE) {  // at most ceil(64 / nE) steps
      e -= nE;
      ++i;
    }
    ok[k] = w < nw && i <= S;
    ii[k] = i;
    ee[k] = e;
    E[k] = (D)gload(g.ev, e);
    // S >= 1 (an Earth model has rows), so row 0 of rye exists
    rye[k] = (D)gload(g.rye, i < S ? i : 0);
  }
  const SetConst sc = job.sc;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    Slab<P> sl;
    solve_slab<P>(sc, E[k], m_recip(E[k]), ii[k] < S ? rye[k] : D(0), sl);
    // loads and stores share one counter (vmcnt), so a wait for the later
    // tiles' loads placed after this tile's stores would also wait for the
    // stores; the loads were issued before this solve and have arrived, so
    // wait for them here (vmcnt(0), expcnt and lgkmcnt left alone)
    if (k == 0 && K > 1) __builtin_amdgcn_s_waitcnt(0x0F70);
    if (ok[k])
      TileIO<Slab<P>>::store(
          slabs + ((int64_t)ii[k] * tiles + ee[k] / kWave) * TileIO<Slab<P>>::kTileBytes,
          ee[k] % kWave, sl);
  }
}

// The batch's pre-pass on a 2-D grid (y: job, x: groups of kSlabTiles of the
// job's waves) so that a workgroup finds its job without a search; tiles past
// a smaller job's count are skipped at once.  A one-tile wave spends most of
// its life waiting (its job's loads, its stores).  With every tile's loads
// issued before the first solve, three tiles per wave take the 1024-set,
// 59-shell batch's pre-pass to 0.414 ms on MI355X, against 0.441 ms with two
// and 0.417 ms with four (one tile at a time, loads after the previous
// tile's stores: 0.485 ms with one tile, 0.448 with two).
constexpr int kSlabTiles = 3;
template <class P>
__global__ void __launch_bounds__(kWave) slab_table_kernel(const AtmosJob* __restrict__ jobs) {
  const AtmosJob& job = jobs[blockIdx.y];
  slab_tiles<P, kSlabTiles>(job, (int64_t)blockIdx.x * kSlabTiles);
}

// The single grid's pre-pass, its job passed as a kernel argument; workgroup
// 0 also stores the job descriptor for the propagation kernel that follows
// in stream order (no host staging buffer and no extra launch).
template <class P>
__global__ void __launch_bounds__(kWave) slab_table_single_kernel(AtmosJob job,
                                                                  AtmosJob* dst) {
  static_assert(sizeof(AtmosJob) % 4 == 0, "AtmosJob is copied as 32-bit words");
  if (blockIdx.x == 0) {
    constexpr int kWords = (int)(sizeof(AtmosJob) / 4);
    const uint32_t* src = reinterpret_cast<const uint32_t*>(&job);
    uint32_t* out = reinterpret_cast<uint32_t*>(dst);
    for (int i = (int)threadIdx.x; i < kWords; i += kWave) out[i] = src[i];
  }
  slab_tiles<P, 1>(job, (int64_t)blockIdx.x);
}

// The NC trajectories of cosine block sb (NC cosines) for one energy per
// lane, reading slab row i of the lane's table column: entry `lane` of the
// tile block col + i * row (row: bytes per table row), each row loaded one
// shell ahead of its use.
template <class P, int NC, bool AVG = false>
__device__ __forceinline__ void table_block(const GridDev& g, int sb, int e, bool e_ok,
                                            const char* col, int lane, int64_t row,
                                            const SetConst& sc, void* outp,
                                            const PhaDev& pha) {
  using IO = TileIO<Slab<P>>;
  CosSet<NC> cs;
  cos_set<NC>(g, sb * NC, cs);
  Traj<P> t[NC];
#pragma unroll
  for (int j = 0; j < NC; ++j) traj_init<P>(t[j], cs.Mc[j] == 0);
  // rows i + 1 <= Mmax <= g.S all exist (row g.S: the vacuum)
  Slab<P> s = IO::load(col, lane);
  int i = 0;
  for (; i < cs.Mmin - 1; ++i) {
    const Slab<P> nx = IO::load(col + (i + 1) * row, lane);
    middle_all<P, NC>(t, s, g, cs.cidx, i);
    s = nx;
  }
  for (; i < cs.Mmax - 1; ++i) {
    const Slab<P> nx = IO::load(col + (i + 1) * row, lane);
    middle_some<P, NC>(t, s, g, cs.cidx, cs.Mc, i);
    s = nx;
  }
  // then each trajectory's innermost shell (its own row) and its finish: the
  // state of the finished ones is dead, and the shell loop above never holds
  // the innermost shell's temporaries
  const Slab<P> vac = IO::load(col + g.S * row, lane);
  if constexpr (NC == 4 && !AVG) {
    // two trajectories at a time, in one basic block (both innermost shells,
    // then both finishes, then both stores) so that their chains interleave:
    // on the 1024-set, 59-shell batch on MI355X the FP64 step is 0.8% faster
    // and the mixed packed tails 1.7%; with block 8 the pairs spill and the
    // mixed kernel is 1% slower, so block 8 keeps one at a time
#pragma unroll
    for (int j = 0; j < NC; j += 2) {
      using D = typename P::D;
      using F = typename P::F;
      using O = typename OutT<P>::type;
      if (cs.Mc[j] >= 1 && cs.Mc[j + 1] >= 1) {
        const int ia = cs.Mc[j] - 1, ib = cs.Mc[j + 1] - 1;
        const Slab<P> sa = IO::load(col + ia * row, lane);
        const Slab<P> sb2 = IO::load(col + ib * row, lane);
        traj_inner<P>(t[j], sa, (D)uload(g.seg, (int64_t)cs.cidx[j] * g.S + ia));
        traj_inner<P>(t[j + 1], sb2, (D)uload(g.seg, (int64_t)cs.cidx[j + 1] * g.S + ib));
      } else {
#pragma unroll
        for (int k = j; k < j + 2; ++k)
          if (cs.Mc[k] >= 1) {
            const int im = cs.Mc[k] - 1;
            const Slab<P> si = IO::load(col + im * row, lane);
            traj_inner<P>(t[k], si, (D)uload(g.seg, (int64_t)cs.cidx[k] * g.S + im));
          }
      }
      F pa[9], pb[9];
      traj_finish<P>(t[j], vac, (D)uload(g.Lvac, cs.cidx[j]), sc, pa);
      traj_finish<P>(t[j + 1], vac, (D)uload(g.Lvac, cs.cidx[j + 1]), sc, pb);
      O* out = (O*)outp;
      const int64_t plane = (int64_t)g.nC * g.nE;
      if (e_ok && cs.ok[j]) {
        O* dst = out + (int64_t)cs.cidx[j] * g.nE + e;
#pragma unroll
        for (int ch = 0; ch < 9; ++ch) gstore(dst, ch * plane, (O)pa[ch]);
      }
      if (e_ok && cs.ok[j + 1]) {
        O* dst = out + (int64_t)cs.cidx[j + 1] * g.nE + e;
#pragma unroll
        for (int ch = 0; ch < 9; ++ch) gstore(dst, ch * plane, (O)pb[ch]);
      }
    }
  } else {
    // one trajectory at a time
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      if (cs.Mc[j] >= 1) {
        using D = typename P::D;
        const int im = cs.Mc[j] - 1;
        const Slab<P> si = IO::load(col + im * row, lane);
        traj_inner<P>(t[j], si, (D)uload(g.seg, (int64_t)cs.cidx[j] * g.S + im));
      }
      finish_store_one<P, NC, AVG>(t[j], j, vac, g, cs, sc, outp, e, e_ok, pha);
    }
  }
}

// A work unit of one job in table mode: energy tile x cosine block, the
// cosine blocks of a tile consecutive so that they share its table rows.
template <class P, int NC, bool AVG = false>
__device__ __forceinline__ void table_unit(const AtmosJob& job, int64_t local) {
  const int lane = (int)threadIdx.x;
  const int ncb = (int)cosine_blocks(job.g, NC, 1);
  const int et = (int)(local / ncb);
  const int sb = (int)(local - (int64_t)et * ncb);
  const int e = et * kWave + lane;
  const int64_t row = energy_tiles(job.g) * TileIO<Slab<P>>::kTileBytes;
  table_block<P, NC, AVG>(job.g, sb, e, e < job.g.nE, slab_tile<P>(job.slabs, job.g, 0, e),
                          lane, row, job.sc, job.out, job.pha);
}

// A packed tail unit in table mode: packed tile x cosine block.
template <class P, int NC>
__device__ __forceinline__ void tail_table_unit(const AtmosJob* jobs, const TailGroup& tg,
                                                int64_t local) {
  const int lane = (int)threadIdx.x;
  const int ncb = (int)cosine_blocks(tg.g, NC, 1);
  const int t = (int)(local / ncb);
  const int sb = (int)(local - (int64_t)t * ncb);
  const int q = lane / tg.tail;
  const int slot = t * tg.per + q;
  const bool ok = q < tg.per && slot < tg.njobs;
  const int e = tg.g.nE - tg.tail + (lane - q * tg.tail);
  const AtmosJob& job = jobs[gload(tg.jobs, ok ? slot : 0)];
  const int64_t row = energy_tiles(tg.g) * TileIO<Slab<P>>::kTileBytes;
  table_block<P, NC>(tg.g, sb, e, ok, slab_tile<P>(job.slabs, tg.g, 0, e), e % kWave, row,
                     job.sc, job.out, job.pha);
}

// Minimum waves per SIMD of the table-mode kernels (0: natural allocation).
// The mixed kernel with 8 cosines per lane is held to 3 waves (168 VGPRs,
// a few spilled): left alone the scheduler takes 235 VGPRs and 2 waves,
// 13% slower on the 1024-set, 59-shell batch on MI355X.
template <class P, int NC>
struct MinWavesTable {
  static constexpr int value = 0;
};
template <int NC>
struct MinWavesTable<PolicyMixed, NC> {
  static constexpr int value = NC == 8 ? 3 : 0;
};


// Table-mode propagation.  Workgroups are dealt to the 8 XCDs round-robin;
// block b takes unit (b % 8) * per + b / 8, so each XCD walks a contiguous
// eighth of the unit list and the table rows its units share stay in its
// own L2 (the grid is a multiple of 8, so a block keeps its XCD share).
template <class P, int NC>
__global__ void __launch_bounds__(kWave, (MinWavesTable<P, NC>::value)) atmos_table_kernel(
    const AtmosJob* __restrict__ jobs, int njobs, int64_t nunits) {
  const int64_t per = (nunits + 7) / 8;
  for (int64_t b = blockIdx.x; b < per * 8; b += gridDim.x) {
    const int64_t u = (b % 8) * per + b / 8;
    if (u >= nunits) continue;
    const int j = find_job_atmos(jobs, njobs, u);
    const AtmosJob& job = jobs[j];
    table_unit<P, NC>(job, u - job.unit_begin);
  }
}

// Table-mode packed tails, with the same XCD-local walk of the unit list.
template <class P, int NC>
__global__ void __launch_bounds__(kWave, (MinWavesTable<P, NC>::value)) tail_table_kernel(
    const AtmosJob* __restrict__ jobs, const TailGroup* __restrict__ groups, int ngroups,
    int64_t nunits) {
  const int64_t per = (nunits + 7) / 8;
  for (int64_t b = blockIdx.x; b < per * 8; b += gridDim.x) {
    const int64_t u = (b % 8) * per + b / 8;
    if (u >= nunits) continue;
    int lo = 0, hi = ngroups - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (groups[mid].unit_begin <= u)
        lo = mid;
      else
        hi = mid - 1;
    }
    const TailGroup& tg = groups[lo];
    tail_table_unit<P, NC>(jobs, tg, u - tg.unit_begin);
  }
}

// Production-height averaging on a single grid, one cosine per wave: the
// in-kernel solve, one workgroup per unit ...
template <class P>
__global__ void __launch_bounds__(kWave) atmos_single_avg_kernel(AtmosJob job) {
  NoStage st;
  atmos_unit<P, 1, 1, true>(job, (int64_t)blockIdx.x, st);
}

// ... and the slab-table form, with the XCD-local walk of atmos_table_kernel.
template <class P>
__global__ void __launch_bounds__(kWave) atmos_table_avg_kernel(
    const AtmosJob* __restrict__ job, int64_t nunits) {
  const int64_t per = (nunits + 7) / 8;
  for (int64_t b = blockIdx.x; b < per * 8; b += gridDim.x) {
    const int64_t u = (b % 8) * per + b / 8;
    if (u >= nunits) continue;
    table_unit<P, 1, true>(*job, u);
  }
}

template <class P>
__device__ __forceinline__ void beam_unit(const BeamJob& job, int64_t local, int lane) {
  using D = typename P::D;
  using F = typename P::F;
  using O = typename OutT<P>::type;
  const int e = (int)local * kWave + lane;
  const bool e_ok = e < job.nE;
  const D E = (D)gload(job.ev, e_ok ? e : job.nE - 1);
  F prob[9];
  beam_point<P>(job.sc, E, (D)job.rye, (D)job.L, prob);
  if (e_ok) {
    O* out = (O*)job.out;
#pragma unroll
    for (int ch = 0; ch < 9; ++ch) gstore(out, (int64_t)ch * job.nE + e, (O)prob[ch]);
  }
}

template <class P>
__global__ void __launch_bounds__(kWave) beam_single_kernel(BeamJob job) {
  beam_unit<P>(job, (int64_t)blockIdx.x, (int)threadIdx.x);
}

// kBeamPerLane energies of one job per lane (e, e + 64, ...: independent
// solves in flight), so a wave covers kBeamPerLane * 64 consecutive energies,
// and kBeamWaves waves per workgroup.  Without a loop the solve's constants
// are not held in registers across units.  The grid is 2-D (y: job, x: groups
// of kBeamWaves waves), so a wave finds its job without a search; waves past
// a shorter job's energies exit at once.  The kernel is VALU-bound: on the
// 4096-set x 1e5-energy batch (MI355X) the same arithmetic without the
// stores takes 73-84% of the time of the kernel with two energies per lane
// (FP64 to FP32 forms), the stores alone 60% (FP32 form, 5.6 TB/s); four
// energies per lane overlap the two better (FP32 4.41 -> 3.86 ms, mixed
// 6.50 -> 6.22 ms, FP64 8.36 -> 8.12 ms).
constexpr int kBeamWaves = 4;
constexpr int kBeamPerLane = 4;
template <class P>
__global__ void __launch_bounds__(kWave * kBeamWaves) beam_batch_kernel(
    const BeamJob* __restrict__ jobs, int job0) {
  using D = typename P::D;
  using F = typename P::F;
  using O = typename OutT<P>::type;
  const BeamJob& job = jobs[job0 + (int)blockIdx.y];
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kWave);
  const int64_t local = (int64_t)blockIdx.x * kBeamWaves + wave;
  if (local * kBeamPerLane * kWave >= job.nE) return;
  const int lane = (int)threadIdx.x % kWave;
  O* out = (O*)job.out;
  F p[kBeamPerLane][9];
  int e[kBeamPerLane];
#pragma unroll
  for (int k = 0; k < kBeamPerLane; ++k) {
    e[k] = (int)local * kBeamPerLane * kWave + k * kWave + lane;
    const D E = (D)gload(job.ev, e[k] < job.nE ? e[k] : job.nE - 1);
    beam_point<P>(job.sc, E, (D)job.rye, (D)job.L, p[k]);
  }
#pragma unroll
  for (int k = 0; k < kBeamPerLane; ++k)
    if (e[k] < job.nE) {
#pragma unroll
      for (int ch = 0; ch < 9; ++ch) gstore(out, (int64_t)ch * job.nE + e[k], (O)p[k][ch]);
    }
}

// ------------------------------------------------------------- launchers

#define CP3_NCW_CASES(X)                                                   \
  X(1, 1) X(2, 1) X(4, 1) X(8, 1) X(1, 2) X(2, 2) X(4, 2) X(8, 2) X(1, 4) \
  X(2, 4) X(4, 4) X(8, 4)

template <class P>
static hipError_t atmos_single_p(int nc, int w, const AtmosJob& job,
                                 hipStream_t st) {
  const int64_t n = atmos_units(job.g, nc, w);
  if (n <= 0) return hipSuccess;
  if (n > 0x7fffffff) return hipErrorInvalidConfiguration;
  const dim3 grid((unsigned)n);
#define X(NC_, W_)                                                        \
  if (nc == NC_ && w == W_) {                                             \
    atmos_single_kernel<P, NC_, W_><<<grid, dim3(kWave * W_), 0, st>>>(job); \
    return hipGetLastError();                                             \
  }
  CP3_NCW_CASES(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_atmos_single(int precision, int nc, int w, const AtmosJob& job,
                               hipStream_t st) {
  switch (precision) {
    case kFP64: return atmos_single_p<PolicyF64>(nc, w, job, st);
    case kFP32: return atmos_single_p<PolicyF32>(nc, w, job, st);
    case kMixed: return atmos_single_p<PolicyMixed>(nc, w, job, st);
    default: return hipErrorInvalidValue;
  }
}

template <class P>
static hipError_t atmos_batch_p(int nc, int w, const AtmosJob* jobs, int njobs,
                                int64_t nunits, int grid, hipStream_t st) {
  if (nunits <= 0 || njobs <= 0) return hipSuccess;
  const dim3 g((unsigned)grid);
#define X(NC_, W_)                                                             \
  if (nc == NC_ && w == W_) {                                                  \
    atmos_batch_kernel<P, NC_, W_><<<g, dim3(kWave * W_), 0, st>>>(jobs, njobs, \
                                                                   nunits);    \
    return hipGetLastError();                                                  \
  }
  CP3_NCW_CASES(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_atmos_batch(int precision, int nc, int w, const AtmosJob* jobs,
                              int njobs, int64_t nunits, int grid,
                              hipStream_t st) {
  if (grid <= 0) return hipErrorInvalidConfiguration;
  switch (precision) {
    case kFP64: return atmos_batch_p<PolicyF64>(nc, w, jobs, njobs, nunits, grid, st);
    case kFP32: return atmos_batch_p<PolicyF32>(nc, w, jobs, njobs, nunits, grid, st);
    case kMixed: return atmos_batch_p<PolicyMixed>(nc, w, jobs, njobs, nunits, grid, st);
    default: return hipErrorInvalidValue;
  }
}

template <class P>
static hipError_t atmos_tail_p(int nc, int w, const AtmosJob* jobs,
                               const TailGroup* groups, int ngroups, int64_t nunits,
                               hipStream_t st) {
  if (nunits <= 0 || ngroups <= 0) return hipSuccess;
  if (nunits > 0x7fffffff) return hipErrorInvalidConfiguration;
  const dim3 g((unsigned)nunits);
#define X(NC_, W_)                                                            \
  if (nc == NC_ && w == W_) {                                                 \
    atmos_tail_kernel<P, NC_, W_><<<g, dim3(kWave * W_), 0, st>>>(jobs, groups, \
                                                                  ngroups, nunits); \
    return hipGetLastError();                                                 \
  }
  CP3_NCW_CASES(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_atmos_tail(int precision, int nc, int w, const AtmosJob* jobs,
                             const TailGroup* groups, int ngroups, int64_t nunits,
                             hipStream_t st) {
  switch (precision) {
    case kFP64: return atmos_tail_p<PolicyF64>(nc, w, jobs, groups, ngroups, nunits, st);
    case kFP32: return atmos_tail_p<PolicyF32>(nc, w, jobs, groups, ngroups, nunits, st);
    case kMixed: return atmos_tail_p<PolicyMixed>(nc, w, jobs, groups, ngroups, nunits, st);
    default: return hipErrorInvalidValue;
  }
}

#define CP3_NC_CASES(X) X(1) X(2) X(4) X(8)

size_t slab_entry_bytes(int precision) {
  switch (precision) {
    case kFP64: return sizeof(Slab<PolicyF64>);
    case kFP32: return sizeof(Slab<PolicyF32>);
    case kMixed: return sizeof(Slab<PolicyMixed>);
    default: return 0;
  }
}

hipError_t launch_slab_table_single(int precision, const AtmosJob& job, AtmosJob* d_job,
                                   hipStream_t st) {
  const int64_t waves = slab_waves(job.g);
  if (waves <= 0 || waves > 0x7fffffff) return hipErrorInvalidConfiguration;
  const dim3 g((unsigned)waves), b(kWave);
  switch (precision) {
    case kFP64: slab_table_single_kernel<PolicyF64><<<g, b, 0, st>>>(job, d_job); break;
    case kFP32: slab_table_single_kernel<PolicyF32><<<g, b, 0, st>>>(job, d_job); break;
    case kMixed: slab_table_single_kernel<PolicyMixed><<<g, b, 0, st>>>(job, d_job); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_slab_table(int precision, const AtmosJob* jobs, int njobs,
                             int64_t max_waves, hipStream_t st) {
  if (max_waves <= 0 || njobs <= 0) return hipSuccess;
  if (max_waves > 0x7fffffff || njobs > 65535) return hipErrorInvalidConfiguration;
  const dim3 g((unsigned)((max_waves + kSlabTiles - 1) / kSlabTiles), (unsigned)njobs),
      b(kWave);
  switch (precision) {
    case kFP64: slab_table_kernel<PolicyF64><<<g, b, 0, st>>>(jobs); break;
    case kFP32: slab_table_kernel<PolicyF32><<<g, b, 0, st>>>(jobs); break;
    case kMixed: slab_table_kernel<PolicyMixed><<<g, b, 0, st>>>(jobs); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_atmos_single_avg(int precision, const AtmosJob& job, hipStream_t st) {
  const int64_t n = atmos_units(job.g, 1, 1);
  if (n <= 0) return hipSuccess;
  if (n > 0x7fffffff) return hipErrorInvalidConfiguration;
  const dim3 g((unsigned)n), b(kWave);
  switch (precision) {
    case kFP64: atmos_single_avg_kernel<PolicyF64><<<g, b, 0, st>>>(job); break;
    case kFP32: atmos_single_avg_kernel<PolicyF32><<<g, b, 0, st>>>(job); break;
    case kMixed: atmos_single_avg_kernel<PolicyMixed><<<g, b, 0, st>>>(job); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_atmos_table_avg(int precision, const AtmosJob* d_job, int64_t nunits,
                                  int grid, hipStream_t st) {
  if (nunits <= 0) return hipSuccess;
  if (grid <= 0 || grid % 8) return hipErrorInvalidConfiguration;
  const dim3 g((unsigned)grid), b(kWave);
  switch (precision) {
    case kFP64: atmos_table_avg_kernel<PolicyF64><<<g, b, 0, st>>>(d_job, nunits); break;
    case kFP32: atmos_table_avg_kernel<PolicyF32><<<g, b, 0, st>>>(d_job, nunits); break;
    case kMixed: atmos_table_avg_kernel<PolicyMixed><<<g, b, 0, st>>>(d_job, nunits); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <class P>
static hipError_t atmos_table_p(int nc, const AtmosJob* jobs, int njobs, int64_t nunits,
                                int grid, hipStream_t st) {
  if (nunits <= 0 || njobs <= 0) return hipSuccess;
  if (grid <= 0 || grid % 8) return hipErrorInvalidConfiguration;
#define X(NC_)                                                                  \
  if (nc == NC_) {                                                              \
    atmos_table_kernel<P, NC_><<<dim3((unsigned)grid), dim3(kWave), 0, st>>>(    \
        jobs, njobs, nunits);                                                   \
    return hipGetLastError();                                                   \
  }
  CP3_NC_CASES(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_atmos_table(int precision, int nc, const AtmosJob* jobs, int njobs,
                              int64_t nunits, int grid, hipStream_t st) {
  switch (precision) {
    case kFP64: return atmos_table_p<PolicyF64>(nc, jobs, njobs, nunits, grid, st);
    case kFP32: return atmos_table_p<PolicyF32>(nc, jobs, njobs, nunits, grid, st);
    case kMixed: return atmos_table_p<PolicyMixed>(nc, jobs, njobs, nunits, grid, st);
    default: return hipErrorInvalidValue;
  }
}

template <class P>
static hipError_t tail_table_p(int nc, const AtmosJob* jobs, const TailGroup* groups,
                               int ngroups, int64_t nunits, hipStream_t st) {
  if (nunits <= 0 || ngroups <= 0) return hipSuccess;
  const int64_t grid = (nunits + 7) / 8 * 8;
  if (grid > 0x7fffffff) return hipErrorInvalidConfiguration;
#define X(NC_)                                                                   \
  if (nc == NC_) {                                                               \
    tail_table_kernel<P, NC_><<<dim3((unsigned)grid), dim3(kWave), 0, st>>>(      \
        jobs, groups, ngroups, nunits);                                          \
    return hipGetLastError();                                                    \
  }
  CP3_NC_CASES(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_tail_table(int precision, int nc, const AtmosJob* jobs,
                             const TailGroup* groups, int ngroups, int64_t nunits,
                             hipStream_t st) {
  switch (precision) {
    case kFP64: return tail_table_p<PolicyF64>(nc, jobs, groups, ngroups, nunits, st);
    case kFP32: return tail_table_p<PolicyF32>(nc, jobs, groups, ngroups, nunits, st);
    case kMixed: return tail_table_p<PolicyMixed>(nc, jobs, groups, ngroups, nunits, st);
    default: return hipErrorInvalidValue;
  }
}

template <class P>
static hipError_t table_occ(int nc, int* n) {
#define X(NC_)                                                                \
  if (nc == NC_)                                                              \
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, atmos_table_kernel<P, NC_>, \
                                                        kWave, 0);
  CP3_NC_CASES(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t atmos_table_occupancy(int precision, int nc, int* n) {
  switch (precision) {
    case kFP64: return table_occ<PolicyF64>(nc, n);
    case kFP32: return table_occ<PolicyF32>(nc, n);
    case kMixed: return table_occ<PolicyMixed>(nc, n);
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
                             int64_t max_job_units, hipStream_t st) {
  if (max_job_units <= 0 || njobs <= 0) return hipSuccess;
  if (max_job_units > 0x7fffffff) return hipErrorInvalidConfiguration;
  // at most 65535 jobs per launch (grid y)
  for (int j0 = 0; j0 < njobs; j0 += 65535) {
    // waves of kBeamPerLane * 64 energies
    const int64_t waves = (max_job_units + kBeamPerLane - 1) / kBeamPerLane;
    const dim3 g((unsigned)((waves + kBeamWaves - 1) / kBeamWaves),
                 (unsigned)std::min(njobs - j0, 65535)),
        b(kWave * kBeamWaves);
    switch (precision) {
      case kFP64: beam_batch_kernel<PolicyF64><<<g, b, 0, st>>>(jobs, j0); break;
      case kFP32: beam_batch_kernel<PolicyF32><<<g, b, 0, st>>>(jobs, j0); break;
      case kMixed: beam_batch_kernel<PolicyMixed><<<g, b, 0, st>>>(jobs, j0); break;
      default: return hipErrorInvalidValue;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template <class P>
static hipError_t atmos_occ(int nc, int w, int* n) {
#define X(NC_, W_)                                                     \
  if (nc == NC_ && w == W_)                                            \
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(               \
        n, atmos_batch_kernel<P, NC_, W_>, kWave * W_, 0);
  CP3_NCW_CASES(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t atmos_batch_occupancy(int precision, int nc, int w, int* n) {
  switch (precision) {
    case kFP64: return atmos_occ<PolicyF64>(nc, w, n);
    case kFP32: return atmos_occ<PolicyF32>(nc, w, n);
    case kMixed: return atmos_occ<PolicyMixed>(nc, w, n);
    default: return hipErrorInvalidValue;
  }
}

hipError_t beam_batch_occupancy(int precision, int* n) {
  switch (precision) {
    case kFP64: return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, beam_batch_kernel<PolicyF64>, kWave * kBeamWaves, 0);
    case kFP32: return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, beam_batch_kernel<PolicyF32>, kWave * kBeamWaves, 0);
    case kMixed: return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, beam_batch_kernel<PolicyMixed>, kWave * kBeamWaves, 0);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace cp3
