// Batched engine (include/cp3_hip.h, cp3_batch_*): resident grids, a job list
// of (parameter set, neutrino type, grid), one persistent launch per kind over
// the flattened work list of all jobs, optional hipGraph replay, and an
// autotuner for the cosine block and the persistent grid size.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/cp3_hip.h"
#include "cp3_capi_util.hpp"
#include "cp3_grid.hpp"
#include "cp3_host.hpp"
#include "cp3_launch.hpp"

using cp3::capi::DevBuf;
using cp3::capi::DeviceScope;
using cp3::capi::Error;
using cp3::capi::guard;
using cp3::capi::hip_check;
using cp3::capi::PinnedBuf;

namespace {

constexpr size_t kAlign = 256;

size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

}  // namespace

using cp3::capi::AtmosGrid;
using cp3::capi::BeamGrid;

struct cp3_batch {
  int dev = 0, precision = CP3_FP64;
  int cus = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // recorded at the start of a run's launches and after each phase:
  // slab-table pre-pass, atmospheric propagation, packed tails, beam (in the
  // enqueued work itself, so that a replayed graph records them too)
  hipEvent_t evp[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  bool ran = false;
  std::vector<std::unique_ptr<AtmosGrid>> agrids;
  std::vector<std::unique_ptr<BeamGrid>> bgrids;

  std::vector<cp3_job> jobs;
  std::vector<int> slot;  // job -> index in its kind's array
  std::vector<cp3::AtmosJob> ajobs;
  std::vector<cp3::BeamJob> bjobs;
  DevBuf d_ajobs, d_bjobs, d_out;
  // packed tail tiles (cp3::TailGroup) and their job index lists
  std::vector<cp3::TailGroup> tgroups;
  std::vector<int32_t> tjobs;
  std::vector<size_t> tjob_off;
  DevBuf d_tgroups, d_tjobs;
  PinnedBuf staging;
  std::vector<size_t> out_off, out_len;
  int64_t a_units = 0, t_units = 0, b_units = 0, evals = 0;
  size_t out_total = 0;
  // slab tables (CP3_SOLVE_TABLE): per atmospheric job its byte offset in
  // d_slabs, and the pre-pass grid width (the largest job's waves)
  std::vector<size_t> slab_off;
  size_t slab_total = 0;
  int64_t s_waves = 0;
  DevBuf d_slabs;

  int tune_nc = 0, tune_w = 0, tune_waves = 0;  // 0 = automatic
  int tune_mode = CP3_SOLVE_AUTO;
  // nc, w: cosine block and group waves of the propagation; tnc: the packed
  // tails' cosine block
  int nc = 1, w = 1, tnc = 1, mode = CP3_SOLVE_SHARED, a_grid = 0;
  int64_t b_max_units = 0;  // units of the longest beam job

  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  ~cp3_batch() {
    DeviceScope ds(dev);
    drop_graph();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    for (hipEvent_t e : evp)
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }

  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }

  int elem() const { return cp3::out_elem_bytes(precision); }

  // Builds the device job tables from `jobs` (constants only when
  // `params_only`) and uploads them through pinned staging on the stream.
  void build_tables(bool params_only) {
    const size_t na = ajobs.size(), nb = bjobs.size();
    if (!params_only && mode == CP3_SOLVE_TABLE) d_slabs.alloc(slab_total, dev);
    for (size_t j = 0; j < jobs.size(); ++j) {
      const cp3_job& in = jobs[j];
      const cp3::SetConst sc =
          cp3::make_set(in.theta12, in.theta13, in.theta23, in.dcp,
                        in.knutype < 0 ? -1 : 1, in.dm12sq, in.dm23sq, in.nutype);
      if (in.kind == CP3_JOB_ATMOS) {
        cp3::AtmosJob& a = ajobs[(size_t)slot[j]];
        a.sc = sc;
        if (!params_only) {
          a.g = agrids[(size_t)in.grid]->dev();
          a.out = (char*)d_out.p + out_off[j];
          a.slabs = mode == CP3_SOLVE_TABLE
                        ? (void*)((char*)d_slabs.p + slab_off[(size_t)slot[j]])
                        : nullptr;
        }
      } else {
        cp3::BeamJob& b = bjobs[(size_t)slot[j]];
        b.sc = sc;
        if (!params_only) {
          const BeamGrid& g = *bgrids[(size_t)in.grid];
          b.ev = (const double*)g.ev.p;
          b.nE = g.nE;
          b.rye = g.rye;
          b.L = g.L;
          b.out = (char*)d_out.p + out_off[j];
        }
      }
    }
    const size_t abytes = na * sizeof(cp3::AtmosJob);
    const size_t bbytes = nb * sizeof(cp3::BeamJob);
    const size_t gbytes = params_only ? 0 : tgroups.size() * sizeof(cp3::TailGroup);
    const size_t jbytes = params_only ? 0 : tjobs.size() * sizeof(int32_t);
    DeviceScope ds(dev);
    // staging may still feed an earlier copy on this stream
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    staging.alloc(abytes + bbytes + gbytes + jbytes);
    char* const st = (char*)staging.p;
    if (!params_only) {
      d_ajobs.alloc(abytes, dev);
      d_bjobs.alloc(bbytes, dev);
      d_tgroups.alloc(gbytes, dev);
      d_tjobs.alloc(jbytes, dev);
      for (size_t t = 0; t < tgroups.size(); ++t)
        tgroups[t].jobs = (const int32_t*)d_tjobs.p + tjob_off[t];
    }
    if (abytes) std::memcpy(st, ajobs.data(), abytes);
    if (bbytes) std::memcpy(st + abytes, bjobs.data(), bbytes);
    if (gbytes) std::memcpy(st + abytes + bbytes, tgroups.data(), gbytes);
    if (jbytes) std::memcpy(st + abytes + bbytes + gbytes, tjobs.data(), jbytes);
    const struct {
      void* dst;
      size_t off, n;
    } ups[4] = {{d_ajobs.p, 0, abytes},
                {d_bjobs.p, abytes, bbytes},
                {d_tgroups.p, abytes + bbytes, gbytes},
                {d_tjobs.p, abytes + bbytes + gbytes, jbytes}};
    for (const auto& u : ups)
      if (u.n)
        hip_check(hipMemcpyAsync(u.dst, st + u.off, u.n, hipMemcpyHostToDevice, stream),
                  "job table upload");
  }

  // Slab-table layout: job a's table at slab_off[a], [S + 1][tiles * 64]
  // entries.
  void plan_slabs() {
    const size_t entry = cp3::slab_entry_bytes(precision);
    slab_off.assign(ajobs.size(), 0);
    size_t off = 0;
    int64_t wv = 0;
    for (size_t j = 0; j < jobs.size(); ++j) {
      if (jobs[j].kind != CP3_JOB_ATMOS) continue;
      const size_t a = (size_t)slot[j];
      const cp3::GridDev g = agrids[(size_t)jobs[j].grid]->dev();
      slab_off[a] = off;
      off = align_up(off + (size_t)cp3::slab_entries(g) * entry);
      wv = std::max(wv, cp3::slab_waves(g));
    }
    slab_total = off;
    s_waves = wv;
  }

  // Table mode unless the tables would take more than kTableMaxBytes (or the
  // pre-pass grid would have more than 65535 rows of jobs).
  static constexpr size_t kTableMaxBytes = (size_t)16 << 30;
  int auto_mode() const {
    return slab_total > 0 && slab_total <= kTableMaxBytes && ajobs.size() <= 65535
               ? CP3_SOLVE_TABLE
               : CP3_SOLVE_SHARED;
  }

  // Atmospheric jobs sharing a grid whose nE % 64 leaves room for two or
  // more tails in a wave run their last energies in packed tail tiles; every
  // other job covers all its energy tiles itself.
  void plan_tails() {
    tgroups.clear();
    tjobs.clear();
    tjob_off.clear();
    std::vector<std::vector<int32_t>> by_grid(agrids.size());
    for (size_t j = 0; j < jobs.size(); ++j)
      if (jobs[j].kind == CP3_JOB_ATMOS) by_grid[(size_t)jobs[j].grid].push_back(slot[j]);
    for (size_t gi = 0; gi < agrids.size(); ++gi) {
      const cp3::GridDev g = agrids[gi]->dev();
      const int tail = g.nE % cp3::kWave;
      const int per = tail ? cp3::kWave / tail : 0;
      const bool pack = per >= 2 && by_grid[gi].size() >= 2;
      for (int32_t a : by_grid[gi])
        ajobs[(size_t)a].etiles = pack ? g.nE / cp3::kWave : cp3::energy_tiles(g);
      if (!pack) continue;
      cp3::TailGroup t{};
      t.g = g;
      t.njobs = (int32_t)by_grid[gi].size();
      t.tail = tail;
      t.per = per;
      tgroups.push_back(t);
      tjob_off.push_back(tjobs.size());
      tjobs.insert(tjobs.end(), by_grid[gi].begin(), by_grid[gi].end());
    }
  }

  void check_job(const cp3_job& j) const {
    if (j.kind == CP3_JOB_ATMOS) {
      if (j.grid < 0 || (size_t)j.grid >= agrids.size())
        throw Error(CP3_ERR_INVALID, "unknown atmospheric grid id " + std::to_string(j.grid));
    } else if (j.kind == CP3_JOB_BEAM) {
      if (j.grid < 0 || (size_t)j.grid >= bgrids.size())
        throw Error(CP3_ERR_INVALID, "unknown beam grid id " + std::to_string(j.grid));
    } else {
      throw Error(CP3_ERR_INVALID, "job kind must be CP3_JOB_ATMOS or CP3_JOB_BEAM");
    }
    if (j.nutype != 0 && j.nutype != 1)
      throw Error(CP3_ERR_INVALID, "neutrino type must be 0 or 1");
  }

  void set_jobs(const cp3_job* in, int n) {
    // validate before touching any state: a rejected list leaves the batch
    // as it was
    for (int j = 0; j < n; ++j) check_job(in[j]);
    jobs.assign(in, in + n);
    slot.assign(jobs.size(), 0);
    ajobs.clear();
    bjobs.clear();
    out_off.assign(jobs.size(), 0);
    out_len.assign(jobs.size(), 0);
    size_t off = 0;
    evals = 0;
    for (size_t j = 0; j < jobs.size(); ++j) {
      size_t pts;
      if (jobs[j].kind == CP3_JOB_ATMOS) {
        const AtmosGrid& g = *agrids[(size_t)jobs[j].grid];
        pts = (size_t)g.nC * (size_t)g.nE;
        slot[j] = (int)ajobs.size();
        ajobs.push_back(cp3::AtmosJob{});
      } else {
        pts = (size_t)bgrids[(size_t)jobs[j].grid]->nE;
        slot[j] = (int)bjobs.size();
        bjobs.push_back(cp3::BeamJob{});
      }
      out_off[j] = off;
      out_len[j] = pts * 9 * (size_t)elem();
      off = align_up(off + out_len[j]);
      evals += (int64_t)pts;
    }
    out_total = off;
    d_out.alloc(out_total, dev);
    plan_tails();
    plan_slabs();
    choose_tuning();
    build_tables(false);
    drop_graph();
  }

  void update_jobs(const cp3_job* in, int n) {
    if ((size_t)n != jobs.size())
      throw Error(CP3_ERR_INVALID, "update_jobs: job count differs from set_jobs");
    for (int j = 0; j < n; ++j) {
      check_job(in[j]);
      if (in[j].kind != jobs[(size_t)j].kind || in[j].grid != jobs[(size_t)j].grid)
        throw Error(CP3_ERR_INVALID, "update_jobs: kinds and grids must not change");
    }
    jobs.assign(in, in + n);
    build_tables(true);
  }

  // Unit prefix sums for the chosen cosine block, and grid sizes.
  void choose_tuning() {
    mode = tune_mode ? tune_mode : auto_mode();
    if (mode == CP3_SOLVE_TABLE && (slab_total == 0 || ajobs.size() > 65535))
      mode = CP3_SOLVE_SHARED;
    nc = tune_nc ? tune_nc : (mode == CP3_SOLVE_TABLE ? 4 : auto_nc());
    // the table replaces the waves' shared solves: one wave per workgroup
    w = mode == CP3_SOLVE_TABLE ? 1 : (tune_w ? tune_w : 4);
    // Table-mode tails take at most 4 cosines per lane: on the 1024-set
    // 59-shell batch (MI355X) the tail kernel takes 0.350 ms at block 4 and
    // 0.360 ms at block 8 (0.40 ms at blocks 8 and 2 before block-4 kernels
    // finished their trajectories in pairs), while the propagation kernel is
    // fastest at block 8.
    tnc = mode == CP3_SOLVE_TABLE ? std::min(nc, 4) : nc;
    int64_t u = 0;
    for (size_t j = 0; j < jobs.size(); ++j) {
      if (jobs[j].kind != CP3_JOB_ATMOS) continue;
      cp3::AtmosJob& a = ajobs[(size_t)slot[j]];
      a.unit_begin = u;
      u += cp3::cosine_blocks(agrids[(size_t)jobs[j].grid]->dev(), nc, w) * a.etiles;
    }
    a_units = u;
    u = 0;
    for (cp3::TailGroup& t : tgroups) {
      t.unit_begin = u;
      u += cp3::tail_units(t, tnc, w);
    }
    t_units = u;
    u = 0;
    b_max_units = 0;
    for (size_t j = 0; j < jobs.size(); ++j) {
      if (jobs[j].kind != CP3_JOB_BEAM) continue;
      cp3::BeamJob& b = bjobs[(size_t)slot[j]];
      b.unit_begin = u;
      const int64_t n = cp3::beam_units(bgrids[(size_t)jobs[j].grid]->nE);
      u += n;
      b_max_units = std::max(b_max_units, n);
    }
    b_units = u;
    a_grid = grid_for(a_units, tune_waves, true);
    if (mode == CP3_SOLVE_TABLE && a_grid > 0) {
      // a multiple of 8 (one share per XCD), at most one block per unit
      const int64_t cap = (a_units + 7) / 8 * 8;
      a_grid = (int)std::min<int64_t>(cap, ((int64_t)a_grid + 7) / 8 * 8);
    }
  }

  int auto_nc() const {
    // the largest block that still leaves several waves per SIMD
    const int cap = precision == CP3_FP64 ? 4 : 8;
    for (int c = cap; c > 1; c /= 2) {
      int64_t u = 0;
      for (size_t j = 0; j < jobs.size(); ++j)
        if (jobs[j].kind == CP3_JOB_ATMOS)
          u += cp3::atmos_units(agrids[(size_t)jobs[j].grid]->dev(), c, 4);
      if (u * 4 >= (int64_t)cus * 4 * 8) return c;
    }
    return 1;
  }

  int grid_for(int64_t units, int waves, bool atmos) const {
    if (units <= 0) return 0;
    if (waves < 0 || units <= (int64_t)cus) return (int)std::min<int64_t>(units, 0x7fffffff);
    int occ = 0;
    if (atmos && mode == CP3_SOLVE_TABLE)
      hip_check(cp3::atmos_table_occupancy(precision, nc, &occ), "occupancy");
    else if (atmos)
      hip_check(cp3::atmos_batch_occupancy(precision, nc, w, &occ), "occupancy");
    else
      hip_check(cp3::beam_batch_occupancy(precision, &occ), "occupancy");
    if (occ <= 0) occ = 1;
    const int per_cu = waves > 0 ? std::min(waves, occ) : occ;
    return (int)std::min<int64_t>(units, (int64_t)cus * per_cu);
  }

  void mark(int k) { hip_check(hipEventRecord(evp[k], stream), "hipEventRecord"); }

  void enqueue() {
    const cp3::AtmosJob* aj = (const cp3::AtmosJob*)d_ajobs.p;
    const int na = (int)ajobs.size();
    const bool table = mode == CP3_SOLVE_TABLE;
    mark(0);
    if (table && s_waves > 0)
      hip_check(cp3::launch_slab_table(precision, aj, na, s_waves, stream),
                "slab table launch");
    mark(1);
    if (a_units > 0)
      hip_check(table ? cp3::launch_atmos_table(precision, nc, aj, na, a_units, a_grid, stream)
                      : cp3::launch_atmos_batch(precision, nc, w, aj, na, a_units, a_grid,
                                                stream),
                "batched atmospheric launch");
    mark(2);
    if (t_units > 0) {
      const cp3::TailGroup* tg = (const cp3::TailGroup*)d_tgroups.p;
      const int ng = (int)tgroups.size();
      hip_check(table ? cp3::launch_tail_table(precision, tnc, aj, tg, ng, t_units, stream)
                      : cp3::launch_atmos_tail(precision, nc, w, aj, tg, ng, t_units, stream),
                "packed tail launch");
    }
    mark(3);
    if (b_units > 0)
      hip_check(cp3::launch_beam_batch(precision, (const cp3::BeamJob*)d_bjobs.p,
                                       (int)bjobs.size(), b_max_units, stream),
                "batched beam launch");
    mark(4);
  }

  // Milliseconds of the last run's phases (about 0 for a phase with no
  // launch: consecutive markers can read a few microseconds apart either
  // way, so a negative difference reads as 0).
  void phase_ms(float out[4]) {
    if (!ran) throw Error(CP3_ERR_STATE, "the batch was not run");
    DeviceScope ds(dev);
    hip_check(hipEventSynchronize(ev1), "hipEventSynchronize");
    for (int k = 0; k < 4; ++k) {
      hip_check(hipEventElapsedTime(&out[k], evp[k], evp[k + 1]), "hipEventElapsedTime");
      if (out[k] < 0.f) out[k] = 0.f;
    }
  }

  void run() {
    if (jobs.empty()) throw Error(CP3_ERR_STATE, "no jobs set");
    DeviceScope ds(dev);
    hip_check(hipEventRecord(ev0, stream), "hipEventRecord");
    if (use_graph) {
      if (!exec) {
        hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal),
                  "hipStreamBeginCapture");
        try {
          enqueue();
        } catch (...) {
          hipGraph_t g = nullptr;
          (void)hipStreamEndCapture(stream, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        hip_check(hipStreamEndCapture(stream, &graph), "hipStreamEndCapture");
        hip_check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0),
                  "hipGraphInstantiate");
      }
      hip_check(hipGraphLaunch(exec, stream), "hipGraphLaunch");
    } else {
      enqueue();
    }
    hip_check(hipEventRecord(ev1, stream), "hipEventRecord");
    ran = true;
  }

  void wait() {
    DeviceScope ds(dev);
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
  }

  float last_ms() {
    if (!ran) throw Error(CP3_ERR_STATE, "the batch was not run");
    DeviceScope ds(dev);
    hip_check(hipEventSynchronize(ev1), "hipEventSynchronize");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, ev0, ev1), "hipEventElapsedTime");
    return ms;
  }

  void autotune(int reps, int* best_nc, int* best_w, int* best_waves) {
    if (jobs.empty()) throw Error(CP3_ERR_STATE, "no jobs set");
    if (reps < 1) reps = 1;
    const bool had_graph = use_graph;
    use_graph = false;
    drop_graph();
    float best = 1e30f;
    int bn = 1, bw = 1, bwv = 0, bm = CP3_SOLVE_SHARED;
    const bool any_atmos = !ajobs.empty();
    const int ncs[4] = {1, 2, 4, 8};
    const int gws[3] = {1, 2, 4};
    const int waves[2] = {0, -1};
    // the modes tried: a fixed one, or both (table mode only if it fits)
    int modes[2] = {CP3_SOLVE_SHARED, CP3_SOLVE_TABLE}, nm = 2;
    if (tune_mode != CP3_SOLVE_AUTO) {
      modes[0] = tune_mode;
      nm = 1;
    } else if (auto_mode() != CP3_SOLVE_TABLE) {
      nm = 1;
    }
    if (!any_atmos) nm = 1;
    const int fixed_mode = tune_mode;
    for (int mi = 0; mi < nm; ++mi)
      for (int ci = 0; ci < (any_atmos ? 4 : 1); ++ci)
        for (int gi = 0; gi < (any_atmos && modes[mi] == CP3_SOLVE_SHARED ? 3 : 1); ++gi)
          for (int wi = 0; wi < 2; ++wi) {
            tune_mode = any_atmos ? modes[mi] : fixed_mode;
            tune_nc = any_atmos ? ncs[ci] : 0;
            tune_w = any_atmos && modes[mi] == CP3_SOLVE_SHARED ? gws[gi] : 0;
            tune_waves = waves[wi];
            choose_tuning();
            build_tables(false);
            run();  // warm-up
            wait();
            float tot = 0.f;
            for (int r = 0; r < reps; ++r) {
              run();
              tot += last_ms();
            }
            if (tot / reps < best) {
              best = tot / reps;
              bn = tune_nc;
              bw = tune_w;
              bwv = tune_waves;
              bm = tune_mode;
            }
          }
    tune_nc = bn;
    tune_w = bw;
    tune_waves = bwv;
    tune_mode = fixed_mode == CP3_SOLVE_AUTO ? bm : fixed_mode;
    choose_tuning();
    build_tables(false);
    use_graph = had_graph;
    if (best_nc) *best_nc = nc;
    if (best_w) *best_w = w;
    if (best_waves) *best_waves = tune_waves;
  }
};

namespace {

cp3_batch* need(cp3_batch* b) {
  if (!b) throw Error(CP3_ERR_INVALID, "null batch handle");
  return b;
}

void need_ptr(const void* p, const char* what) {
  if (!p) throw Error(CP3_ERR_INVALID, std::string("null pointer: ") + what);
}

}  // namespace

extern "C" {

int cp3_batch_create(int device, int precision, cp3_batch** out) {
  return guard([&] {
    need_ptr(out, "out");
    *out = nullptr;
    if (precision != CP3_FP64 && precision != CP3_FP32 && precision != CP3_MIXED)
      throw Error(CP3_ERR_INVALID, "unknown precision");
    int n = 0;
    hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (device < 0 || device >= n)
      throw Error(CP3_ERR_INVALID,
                  "The requested GPU Id " + std::to_string(device) + " is not available.");
    std::unique_ptr<cp3_batch> b(new cp3_batch());
    b->dev = device;
    b->precision = precision;
    DeviceScope ds(device);
    hip_check(hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, device),
              "hipDeviceGetAttribute");
    hip_check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipEventCreate(&b->ev0), "hipEventCreate");
    hip_check(hipEventCreate(&b->ev1), "hipEventCreate");
    for (hipEvent_t& e : b->evp) hip_check(hipEventCreate(&e), "hipEventCreate");
    *out = b.release();
  });
}

int cp3_batch_destroy(cp3_batch* b) {
  return guard([&] { delete b; });
}

int cp3_batch_add_atmos_grid(cp3_batch* b, const double* cosv, int nC,
                             const double* ev, int nE, const double* radii,
                             const double* rhos, const double* yps, int nrows,
                             double h_km, int* grid_id) {
  return guard([&] {
    need(b);
    need_ptr(grid_id, "grid_id");
    if (nC <= 0 || nE <= 0) throw Error(CP3_ERR_INVALID, "grid sizes must be positive");
    need_ptr(cosv, "cosines");
    need_ptr(ev, "energies");
    if (nrows < 0) throw Error(CP3_ERR_INVALID, "negative number of model rows");
    if (nrows > 0) {
      need_ptr(radii, "radii");
      need_ptr(rhos, "rhos");
      need_ptr(yps, "yps");
    }
    std::unique_ptr<AtmosGrid> g =
        cp3::capi::make_atmos_grid(cosv, nC, ev, nE, radii, rhos, yps, nrows, h_km, b->dev);
    *grid_id = (int)b->agrids.size();
    b->agrids.push_back(std::move(g));
  });
}

int cp3_batch_add_beam_grid(cp3_batch* b, const double* ev, int nE, double rho,
                            double L, int* grid_id) {
  return guard([&] {
    need(b);
    need_ptr(grid_id, "grid_id");
    if (nE <= 0) throw Error(CP3_ERR_INVALID, "number of energies must be positive");
    need_ptr(ev, "energies");
    std::unique_ptr<BeamGrid> g = cp3::capi::make_beam_grid(ev, nE, rho, L, b->dev);
    *grid_id = (int)b->bgrids.size();
    b->bgrids.push_back(std::move(g));
  });
}

int cp3_batch_set_jobs(cp3_batch* b, const cp3_job* jobs, int n) {
  return guard([&] {
    need(b);
    if (n <= 0) throw Error(CP3_ERR_INVALID, "job list must not be empty");
    need_ptr(jobs, "jobs");
    b->set_jobs(jobs, n);
  });
}

int cp3_batch_update_jobs(cp3_batch* b, const cp3_job* jobs, int n) {
  return guard([&] {
    need(b);
    need_ptr(jobs, "jobs");
    b->update_jobs(jobs, n);
  });
}

int cp3_batch_set_tuning(cp3_batch* b, int nc, int w, int waves) {
  return guard([&] {
    need(b);
    if (nc != 0 && !cp3::valid_nc(nc))
      throw Error(CP3_ERR_INVALID, "cosine block must be 0, 1, 2, 4 or 8");
    if (w != 0 && !cp3::valid_group(w))
      throw Error(CP3_ERR_INVALID, "group waves must be 0, 1, 2 or 4");
    b->tune_nc = nc;
    b->tune_w = w;
    b->tune_waves = waves;
    if (!b->jobs.empty()) {
      b->choose_tuning();
      b->build_tables(false);
      b->drop_graph();
    }
  });
}

int cp3_batch_set_solve_mode(cp3_batch* b, int mode) {
  return guard([&] {
    need(b);
    if (mode != CP3_SOLVE_AUTO && mode != CP3_SOLVE_SHARED && mode != CP3_SOLVE_TABLE)
      throw Error(CP3_ERR_INVALID, "solve mode must be CP3_SOLVE_AUTO, _SHARED or _TABLE");
    b->tune_mode = mode;
    if (!b->jobs.empty()) {
      b->choose_tuning();
      b->build_tables(false);
      b->drop_graph();
    }
  });
}

int cp3_batch_get_tuning(cp3_batch* b, int* nc, int* w, int* waves, int* mode) {
  return guard([&] {
    need(b);
    if (b->jobs.empty()) throw Error(CP3_ERR_STATE, "no jobs set");
    if (nc) *nc = b->nc;
    if (w) *w = b->w;
    if (waves) *waves = b->tune_waves;
    if (mode) *mode = b->mode;
  });
}

int cp3_batch_autotune(cp3_batch* b, int reps, int* nc, int* w, int* waves) {
  return guard([&] { need(b)->autotune(reps, nc, w, waves); });
}

int cp3_batch_use_graph(cp3_batch* b, int enable) {
  return guard([&] {
    need(b);
    b->use_graph = enable != 0;
    if (!b->use_graph) b->drop_graph();
  });
}

int cp3_batch_run(cp3_batch* b) {
  return guard([&] { need(b)->run(); });
}

int cp3_batch_wait(cp3_batch* b) {
  return guard([&] { need(b)->wait(); });
}

int cp3_batch_last_ms(cp3_batch* b, float* ms) {
  return guard([&] {
    need(b);
    need_ptr(ms, "ms");
    *ms = b->last_ms();
  });
}

int cp3_batch_phase_ms(cp3_batch* b, float* ms) {
  return guard([&] {
    need(b);
    need_ptr(ms, "ms");
    b->phase_ms(ms);
  });
}

int cp3_batch_result_device(cp3_batch* b, int job, const void** dptr, int* elem) {
  return guard([&] {
    need(b);
    need_ptr(dptr, "dptr");
    if (job < 0 || (size_t)job >= b->jobs.size())
      throw Error(CP3_ERR_RANGE, "job index out of range");
    *dptr = (const char*)b->d_out.p + b->out_off[(size_t)job];
    if (elem) *elem = b->elem();
  });
}

int cp3_batch_copy_result(cp3_batch* b, int job, double* dst) {
  return guard([&] {
    need(b);
    need_ptr(dst, "dst");
    if (job < 0 || (size_t)job >= b->jobs.size())
      throw Error(CP3_ERR_RANGE, "job index out of range");
    const size_t bytes = b->out_len[(size_t)job];
    DeviceScope ds(b->dev);
    hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    if (b->precision == CP3_FP64) {
      hip_check(hipMemcpy(dst, (const char*)b->d_out.p + b->out_off[(size_t)job], bytes,
                          hipMemcpyDeviceToHost),
                "result copy");
    } else {
      std::vector<float> tmp(bytes / sizeof(float));
      hip_check(hipMemcpy(tmp.data(), (const char*)b->d_out.p + b->out_off[(size_t)job],
                          bytes, hipMemcpyDeviceToHost),
                "result copy");
      for (size_t i = 0; i < tmp.size(); ++i) dst[i] = (double)tmp[i];
    }
  });
}

int cp3_batch_stats(cp3_batch* b, int64_t* units, int64_t* evals, int64_t* out_bytes) {
  return guard([&] {
    need(b);
    if (units) *units = b->a_units + b->b_units;
    if (evals) *evals = b->evals;
    if (out_bytes) *out_bytes = (int64_t)b->out_total;
  });
}

int cp3_batch_get_stream(cp3_batch* b, void** s) {
  return guard([&] {
    need(b);
    need_ptr(s, "stream");
    *s = (void*)b->stream;
  });
}

}  // extern "C"
