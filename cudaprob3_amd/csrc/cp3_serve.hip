// Request server (include/cp3_hip.h, cp3_server_*): continuous batching of
// single-evaluation requests on one device.
//
// Clients submit jobs (a parameter set and neutrino type on a resident grid)
// from any thread.  A scheduler thread keeps a FIFO of pending requests and
// K launch slots (one stream each).  Whenever a slot is free it admits the
// requests at the head of the queue whose outputs fit in the device-memory
// arena (a first-fit allocator over a region of the configured cap), builds
// their job table and launches the batched kernels on that slot's stream;
// requests that arrive while other batches run are admitted into the next
// free slot, so their work overlaps on the device.  A finished request's
// output stays resident until the client fetches or releases it, which
// returns its memory to the arena.
//
// Every tuning of the kernels gives bitwise identical results, so a request's
// probabilities do not depend on which other requests share its batch.
#include <hip/hip_runtime.h>
#include <sys/prctl.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/cp3_hip.h"
#include "cp3_capi_util.hpp"
#include "cp3_grid.hpp"
#include "cp3_host.hpp"
#include "cp3_launch.hpp"

using cp3::capi::AtmosGrid;
using cp3::capi::BeamGrid;
using cp3::capi::DevBuf;
using cp3::capi::DeviceScope;
using cp3::capi::Error;
using cp3::capi::guard;
using cp3::capi::hip_check;
using cp3::capi::PinnedBuf;

namespace {

using Clock = std::chrono::steady_clock;

// First-fit allocator of offsets in [0, cap) with a coalescing free list.
class ArenaAlloc {
 public:
  explicit ArenaAlloc(size_t cap) : cap_(cap) {
    if (cap) free_[0] = cap;
  }
  static size_t round(size_t n) { return (n + 255) / 256 * 256; }
  bool alloc(size_t n, size_t* off) {
    n = round(n);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      *off = it->first;
      const size_t rest = it->second - n, start = it->first + n;
      free_.erase(it);
      if (rest) free_[start] = rest;
      used_ += n;
      return true;
    }
    return false;
  }
  void release(size_t off, size_t n) {
    n = round(n);
    used_ -= n;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t used() const { return used_; }
  size_t cap() const { return cap_; }

 private:
  std::map<size_t, size_t> free_;
  size_t cap_, used_ = 0;
};

enum ReqState { kPending, kRunning, kDone, kFailed };

struct Request {
  cp3_job job{};
  // bytes: the result at off; tbytes: the slab table right after it (at
  // off + round(bytes)) while the request's batch runs, 0 otherwise
  size_t bytes = 0, off = 0, tbytes = 0;
  int64_t evals = 0;
  Clock::time_point t_submit, t_launch, t_done;
  ReqState state = kPending;
  bool has_mem = false;
  // set while cp3_server_fetch copies the result out of the arena: release
  // and a second fetch are refused until the copy is done
  bool fetching = false;
  std::string err;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  std::vector<int64_t> reqs;
  // the admitted batch's launch shape (set under the server lock by admit,
  // read by enqueue without it)
  bool table = false;
  int nc = 1, w = 1, precision = CP3_MIXED;
  int64_t au = 0, bu = 0, bmax = 0, max_waves = 0;
  std::vector<cp3::AtmosJob> ajobs;
  std::vector<cp3::BeamJob> bjobs;
  DevBuf d_ajobs, d_bjobs;
  PinnedBuf staging;
};

double us_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::micro>(b - a).count();
}

}  // namespace

struct cp3_server {
  int dev = 0, precision = CP3_MIXED;
  // cosine block (0: automatic, 1 in table mode and 2 in shared mode: a
  // request's grid then spreads over the most work units) and group waves
  int nc = 0, w = 4;
  // atmospheric batches pre-solve their slabs into a table (CP3_SOLVE_TABLE,
  // the default) or solve inside the propagation kernel (CP3_SOLVE_SHARED)
  int solve_mode = CP3_SOLVE_TABLE;
  int max_batch = 256;
  DevBuf arena;
  std::unique_ptr<ArenaAlloc> alloc;
  std::vector<std::unique_ptr<AtmosGrid>> agrids;
  std::vector<std::unique_ptr<BeamGrid>> bgrids;
  std::vector<std::unique_ptr<Slot>> slots;

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<int64_t> pending;
  std::unordered_map<int64_t, Request> reqs;
  int64_t next_id = 1;
  bool stop = false;
  std::thread worker;
  int64_t n_done = 0, evals_done = 0, n_batches = 0, peak_inflight = 0;
  size_t peak_mem = 0;

  ~cp3_server() { shutdown(); }

  int elem() const { return cp3::out_elem_bytes(precision); }

  void shutdown() {
    {
      std::lock_guard<std::mutex> lk(mu);
      stop = true;
    }
    cv_work.notify_all();
    if (worker.joinable()) worker.join();
    DeviceScope ds(dev);
    for (auto& sp : slots) {
      Slot& s = *sp;
      if (s.stream) (void)hipStreamSynchronize(s.stream);
      if (s.done) (void)hipEventDestroy(s.done);
      if (s.stream) (void)hipStreamDestroy(s.stream);
      s.done = nullptr;
      s.stream = nullptr;
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
    if (j.nutype != 0 && j.nutype != 1) throw Error(CP3_ERR_INVALID, "neutrino type must be 0 or 1");
  }

  // Slab-table bytes of an atmospheric job's grid.
  size_t table_bytes(const cp3_job& j) const {
    if (j.kind != CP3_JOB_ATMOS) return 0;
    return (size_t)cp3::slab_entries(agrids[(size_t)j.grid]->dev()) *
           cp3::slab_entry_bytes(precision);
  }

  size_t points(const cp3_job& j) const {
    return j.kind == CP3_JOB_ATMOS ? agrids[(size_t)j.grid]->points()
                                   : bgrids[(size_t)j.grid]->points();
  }

  // Admission, called with mu held: takes the requests at the head of the
  // queue whose output (and slab table) fit in the arena, in arrival order,
  // and builds their job tables in the slot.  Returns false when the head of
  // the queue does not fit (it waits for memory; admission stays FIFO).
  bool admit(Slot& s) {
    std::vector<int64_t> take;
    const bool table = solve_mode == CP3_SOLVE_TABLE;
    const int nc = this->nc ? this->nc : (table ? 1 : 2);
    while (!pending.empty() && (int)take.size() < max_batch) {
      Request& r = reqs[pending.front()];
      const size_t tb = table ? table_bytes(r.job) : 0;
      size_t off = 0;
      if (!alloc->alloc(ArenaAlloc::round(r.bytes) + ArenaAlloc::round(tb), &off)) break;
      r.off = off;
      r.tbytes = tb;
      r.has_mem = true;
      take.push_back(pending.front());
      pending.pop_front();
    }
    if (take.empty()) return false;
    peak_mem = std::max(peak_mem, alloc->used());
    s.ajobs.clear();
    s.bjobs.clear();
    s.au = s.bu = s.bmax = s.max_waves = 0;
    s.table = table;
    s.nc = nc;
    s.w = w;
    s.precision = precision;
    for (int64_t id : take) {
      const Request& r = reqs[id];
      const cp3_job& in = r.job;
      const cp3::SetConst sc = cp3::make_set(in.theta12, in.theta13, in.theta23, in.dcp,
                                             in.knutype < 0 ? -1 : 1, in.dm12sq, in.dm23sq,
                                             in.nutype);
      void* out = (char*)arena.p + r.off;
      if (in.kind == CP3_JOB_ATMOS) {
        cp3::AtmosJob a{};
        a.sc = sc;
        a.g = agrids[(size_t)in.grid]->dev();
        a.out = out;
        a.unit_begin = s.au;
        a.etiles = cp3::energy_tiles(a.g);
        if (table) {
          a.slabs = (char*)out + ArenaAlloc::round(r.bytes);
          s.au += cp3::atmos_units(a.g, nc, 1);
          s.max_waves = std::max(s.max_waves, cp3::slab_waves(a.g));
        } else {
          s.au += cp3::atmos_units(a.g, nc, w);
        }
        s.ajobs.push_back(a);
      } else {
        const BeamGrid& g = *bgrids[(size_t)in.grid];
        cp3::BeamJob b{};
        b.sc = sc;
        b.ev = (const double*)g.ev.p;
        b.nE = g.nE;
        b.rye = g.rye;
        b.L = g.L;
        b.out = out;
        b.unit_begin = s.bu;
        s.bu += cp3::beam_units(g.nE);
        s.bmax = std::max<int64_t>(s.bmax, cp3::beam_units(g.nE));
        s.bjobs.push_back(b);
      }
    }
    s.reqs = std::move(take);
    s.busy = true;
    return true;
  }

  // The HIP work of an admitted batch, called without mu: job tables to the
  // device, the kernels and the completion event, all on the slot's stream.
  // Only the scheduler thread touches a slot, and the requests of an admitted
  // batch are neither pending nor finished, so no client call reads or frees
  // them meanwhile.
  void enqueue(Slot& s) {
    // A batch of one request (the usual case at moderate rates) passes its
    // job by value to the kernels instead of copying a job table to the
    // device first: the same per-point arithmetic, so the same bits.  In
    // table mode the pre-pass stores the descriptor for the propagation
    // kernel, as the single-grid propagator does.
    if (s.ajobs.size() + s.bjobs.size() == 1) {
      if (!s.ajobs.empty()) {
        const cp3::AtmosJob& a = s.ajobs[0];
        cp3::AtmosJob* dj = (cp3::AtmosJob*)s.d_ajobs.p;
        if (s.table) {
          hip_check(cp3::launch_slab_table_single(s.precision, a, dj, s.stream),
                    "slab table launch");
          hip_check(cp3::launch_atmos_table(s.precision, s.nc, dj, 1, s.au,
                                            (int)std::min<int64_t>((s.au + 7) / 8 * 8,
                                                                   0x7ffffff8),
                                            s.stream),
                    "atmospheric launch");
        } else {
          hip_check(cp3::launch_atmos_single(s.precision, s.nc, s.w, a, s.stream),
                    "atmospheric launch");
        }
      } else {
        hip_check(cp3::launch_beam_single(s.precision, s.bjobs[0], s.stream), "beam launch");
      }
      hip_check(hipEventRecord(s.done, s.stream), "hipEventRecord");
      return;
    }
    const size_t abytes = s.ajobs.size() * sizeof(cp3::AtmosJob);
    const size_t bbytes = s.bjobs.size() * sizeof(cp3::BeamJob);
    std::memcpy(s.staging.p, s.ajobs.data(), abytes);
    std::memcpy((char*)s.staging.p + abytes, s.bjobs.data(), bbytes);
    if (abytes)
      hip_check(hipMemcpyAsync(s.d_ajobs.p, s.staging.p, abytes, hipMemcpyHostToDevice,
                               s.stream),
                "job table upload");
    if (bbytes)
      hip_check(hipMemcpyAsync(s.d_bjobs.p, (char*)s.staging.p + abytes, bbytes,
                               hipMemcpyHostToDevice, s.stream),
                "job table upload");
    const cp3::AtmosJob* aj = (const cp3::AtmosJob*)s.d_ajobs.p;
    const int na = (int)s.ajobs.size();
    if (s.au > 0 && s.table) {
      hip_check(cp3::launch_slab_table(s.precision, aj, na, s.max_waves, s.stream),
                "slab table launch");
      // one workgroup per unit, a multiple of 8 (one share per XCD)
      hip_check(cp3::launch_atmos_table(s.precision, s.nc, aj, na, s.au,
                                        (int)std::min<int64_t>((s.au + 7) / 8 * 8, 0x7ffffff8),
                                        s.stream),
                "atmospheric launch");
    } else if (s.au > 0) {
      hip_check(cp3::launch_atmos_batch(s.precision, s.nc, s.w, aj, na, s.au,
                                        (int)std::min<int64_t>(s.au, 0x7fffffff), s.stream),
                "atmospheric launch");
    }
    if (s.bu > 0)
      hip_check(cp3::launch_beam_batch(s.precision, (const cp3::BeamJob*)s.d_bjobs.p,
                                       (int)s.bjobs.size(), s.bmax, s.stream),
                "beam launch");
    hip_check(hipEventRecord(s.done, s.stream), "hipEventRecord");
  }

  // Called with mu held after enqueue: the batch's requests run (or, if the
  // enqueue failed, fail and give their memory back).
  void started(Slot& s, const std::string& error) {
    if (!error.empty()) {
      for (int64_t id : s.reqs) fail(id, error);
      s.reqs.clear();
      s.busy = false;
      cv_done.notify_all();
      return;
    }
    const Clock::time_point now = Clock::now();
    for (int64_t id : s.reqs) {
      Request& r = reqs[id];
      r.state = kRunning;
      r.t_launch = now;
    }
    ++n_batches;
    int64_t inflight = 0;
    for (const auto& o : slots) inflight += o->busy ? 1 : 0;
    peak_inflight = std::max(peak_inflight, inflight);
  }

  void fail(int64_t id, const std::string& why) {
    Request& r = reqs[id];
    r.state = kFailed;
    r.err = why;
    if (r.has_mem) {
      alloc->release(r.off, ArenaAlloc::round(r.bytes) + ArenaAlloc::round(r.tbytes));
      r.tbytes = 0;
      r.has_mem = false;
    }
  }

  // The slab table of a finished request goes back to the arena; its result
  // stays until fetched or released.
  void drop_table(Request& r) {
    if (r.has_mem && r.tbytes) {
      alloc->release(r.off + ArenaAlloc::round(r.bytes), r.tbytes);
      r.tbytes = 0;
    }
  }

  void retire(Slot& s, hipError_t status) {
    const Clock::time_point now = Clock::now();
    for (int64_t id : s.reqs) {
      auto it = reqs.find(id);
      if (it == reqs.end()) continue;
      if (status == hipSuccess) {
        drop_table(it->second);
        it->second.state = kDone;
        it->second.t_done = now;
        ++n_done;
        evals_done += it->second.evals;
      } else {
        fail(id, std::string("kernel failed: ") + hipGetErrorString(status));
      }
    }
    s.reqs.clear();
    s.busy = false;
  }

  // Poll interval while batches are in flight, and how long the scheduler
  // keeps polling after its last launch or completion before it blocks on
  // the submit signal.  A blocked thread takes tens of microseconds to wake,
  // and over a hundred on a loaded host; the polling thread sees a
  // submission within kPollUs.  With Poisson arrivals at rate r, a fraction
  // e^{-r T} of requests find the scheduler blocked after an idle window T:
  // 13.5% at 2000 req/s for T = 1 ms, which set the p99 latency; for
  // T = 20 ms it is nil above 500 req/s.  Polling costs part of one host
  // core while requests keep coming.
  static constexpr int kPollUs = 2;
  static constexpr int kIdleSpinUs = 20000;

  void loop() {
    DeviceScope ds(dev);
    // The poll below sleeps kPollUs between event queries while batches are
    // in flight; the default 50 us timer slack of a Linux thread would
    // stretch each sleep to about 60 us, and a request's completion (or a new
    // submission) is noticed only at the next poll.
    (void)prctl(PR_SET_TIMERSLACK, 1000UL, 0UL, 0UL, 0UL);
    // Slot state (busy, done, the job tables) is touched by this thread
    // only; the lock guards the queue, the requests and the arena, and is
    // released for every HIP call, so clients submit, wait and fetch while
    // batches are queried and launched.
    std::vector<std::pair<Slot*, hipError_t>> finished;
    std::unique_lock<std::mutex> lk(mu);
    Clock::time_point last = Clock::now();
    for (;;) {
      bool progress = false, busy = false;
      finished.clear();
      lk.unlock();
      for (auto& sp : slots) {
        Slot& s = *sp;
        if (!s.busy) continue;
        const hipError_t q = hipEventQuery(s.done);
        if (q == hipErrorNotReady) {
          busy = true;
          continue;
        }
        finished.emplace_back(&s, q);
      }
      lk.lock();
      for (auto& f : finished) retire(*f.first, f.second);
      if (!finished.empty()) {
        progress = true;
        cv_done.notify_all();
      }
      for (auto& sp : slots) {
        Slot& s = *sp;
        if (pending.empty()) break;
        if (s.busy) continue;
        if (!admit(s)) break;
        lk.unlock();
        std::string error;
        try {
          enqueue(s);
        } catch (const std::exception& e) {
          // work enqueued before the failure may still use the requests'
          // memory
          (void)hipStreamSynchronize(s.stream);
          error = e.what();
        }
        lk.lock();
        started(s, error);
        busy = busy || s.busy;
        progress = true;
      }
      if (stop && !busy) break;
      if (progress) {
        last = Clock::now();
        continue;
      }
      if (busy || Clock::now() - last < std::chrono::microseconds(kIdleSpinUs)) {
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(kPollUs));
        lk.lock();
      } else {
        cv_work.wait(lk);
      }
    }
    for (int64_t id : pending) fail(id, "server stopped");
    pending.clear();
    cv_done.notify_all();
  }
};

namespace {

cp3_server* need(cp3_server* s) {
  if (!s) throw Error(CP3_ERR_INVALID, "null server handle");
  return s;
}

void need_ptr(const void* p, const char* what) {
  if (!p) throw Error(CP3_ERR_INVALID, std::string("null pointer: ") + what);
}

}  // namespace

extern "C" {

int cp3_server_create(int device, int precision, size_t mem_cap_bytes, int n_slots,
                      int max_batch_jobs, cp3_server** out) {
  return guard([&] {
    need_ptr(out, "out");
    *out = nullptr;
    if (precision != CP3_FP64 && precision != CP3_FP32 && precision != CP3_MIXED)
      throw Error(CP3_ERR_INVALID, "unknown precision");
    if (mem_cap_bytes == 0) throw Error(CP3_ERR_INVALID, "memory cap must be positive");
    if (n_slots < 1 || n_slots > 16) throw Error(CP3_ERR_INVALID, "slots must be in 1..16");
    if (max_batch_jobs < 1) throw Error(CP3_ERR_INVALID, "max batch must be positive");
    int n = 0;
    hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (device < 0 || device >= n)
      throw Error(CP3_ERR_INVALID,
                  "The requested GPU Id " + std::to_string(device) + " is not available.");
    std::unique_ptr<cp3_server> s(new cp3_server());
    s->dev = device;
    s->precision = precision;
    s->max_batch = max_batch_jobs;
    DeviceScope ds(device);
    s->arena.alloc(mem_cap_bytes, device);
    s->alloc.reset(new ArenaAlloc(mem_cap_bytes));
    for (int k = 0; k < n_slots; ++k) s->slots.emplace_back(new Slot());
    for (auto& slp : s->slots) {
      Slot& sl = *slp;
      hip_check(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking), "hipStreamCreate");
      hip_check(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming), "hipEventCreate");
      sl.d_ajobs.alloc((size_t)max_batch_jobs * sizeof(cp3::AtmosJob), device);
      sl.d_bjobs.alloc((size_t)max_batch_jobs * sizeof(cp3::BeamJob), device);
      sl.staging.alloc((size_t)max_batch_jobs *
                       (sizeof(cp3::AtmosJob) + sizeof(cp3::BeamJob)));
    }
    cp3_server* raw = s.get();
    s->worker = std::thread([raw] { raw->loop(); });
    *out = s.release();
  });
}

int cp3_server_destroy(cp3_server* s) {
  return guard([&] { delete s; });
}

int cp3_server_set_tuning(cp3_server* s, int nc, int w) {
  return guard([&] {
    need(s);
    if (!cp3::valid_nc(nc)) throw Error(CP3_ERR_INVALID, "cosine block must be 1, 2, 4 or 8");
    if (!cp3::valid_group(w)) throw Error(CP3_ERR_INVALID, "group waves must be 1, 2 or 4");
    std::lock_guard<std::mutex> lk(s->mu);
    s->nc = nc;
    s->w = w;
  });
}

int cp3_server_set_solve_mode(cp3_server* s, int mode) {
  return guard([&] {
    need(s);
    if (mode != CP3_SOLVE_AUTO && mode != CP3_SOLVE_SHARED && mode != CP3_SOLVE_TABLE)
      throw Error(CP3_ERR_INVALID, "solve mode must be CP3_SOLVE_AUTO, _SHARED or _TABLE");
    std::lock_guard<std::mutex> lk(s->mu);
    s->solve_mode = mode == CP3_SOLVE_SHARED ? CP3_SOLVE_SHARED : CP3_SOLVE_TABLE;
  });
}

int cp3_server_add_atmos_grid(cp3_server* s, const double* cosv, int nC, const double* ev,
                              int nE, const double* radii, const double* rhos,
                              const double* yps, int nrows, double h_km, int* grid_id) {
  return guard([&] {
    need(s);
    need_ptr(grid_id, "grid_id");
    std::unique_ptr<AtmosGrid> g =
        cp3::capi::make_atmos_grid(cosv, nC, ev, nE, radii, rhos, yps, nrows, h_km, s->dev);
    std::lock_guard<std::mutex> lk(s->mu);
    *grid_id = (int)s->agrids.size();
    s->agrids.push_back(std::move(g));
  });
}

int cp3_server_add_beam_grid(cp3_server* s, const double* ev, int nE, double rho,
                             double L, int* grid_id) {
  return guard([&] {
    need(s);
    need_ptr(grid_id, "grid_id");
    std::unique_ptr<BeamGrid> g = cp3::capi::make_beam_grid(ev, nE, rho, L, s->dev);
    std::lock_guard<std::mutex> lk(s->mu);
    *grid_id = (int)s->bgrids.size();
    s->bgrids.push_back(std::move(g));
  });
}

int cp3_server_submit(cp3_server* s, const cp3_job* job, int64_t* request_id) {
  return guard([&] {
    need(s);
    need_ptr(job, "job");
    need_ptr(request_id, "request_id");
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) throw Error(CP3_ERR_STATE, "server is stopping");
    s->check_job(*job);
    Request r;
    r.job = *job;
    r.evals = (int64_t)s->points(*job);
    r.bytes = (size_t)r.evals * 9 * (size_t)s->elem();
    // the result and, for an atmospheric request, its slab table while it runs
    const size_t need = ArenaAlloc::round(r.bytes) + ArenaAlloc::round(s->table_bytes(*job));
    if (need > s->alloc->cap())
      throw Error(CP3_ERR_CAPACITY, "request output and table (" + std::to_string(need) +
                                        " B) exceed the server's memory cap");
    r.t_submit = Clock::now();
    const int64_t id = s->next_id++;
    s->reqs.emplace(id, std::move(r));
    s->pending.push_back(id);
    *request_id = id;
    s->cv_work.notify_one();
  });
}

int cp3_server_wait(cp3_server* s, int64_t id, int timeout_ms) {
  return guard([&] {
    need(s);
    std::unique_lock<std::mutex> lk(s->mu);
    auto ready = [&] {
      auto it = s->reqs.find(id);
      return it == s->reqs.end() || it->second.state == kDone || it->second.state == kFailed;
    };
    if (timeout_ms < 0) {
      s->cv_done.wait(lk, ready);
    } else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) {
      throw Error(CP3_ERR_STATE, "timed out waiting for request " + std::to_string(id));
    }
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) throw Error(CP3_ERR_RANGE, "unknown request " + std::to_string(id));
    if (it->second.state == kFailed) throw Error(CP3_ERR_DEVICE, it->second.err);
  });
}

int cp3_server_fetch(cp3_server* s, int64_t id, double* dst) {
  return guard([&] {
    need(s);
    need_ptr(dst, "dst");
    size_t off, bytes;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      auto it = s->reqs.find(id);
      if (it == s->reqs.end()) throw Error(CP3_ERR_RANGE, "unknown request " + std::to_string(id));
      if (it->second.state != kDone) throw Error(CP3_ERR_STATE, "request is not finished");
      if (it->second.fetching)
        throw Error(CP3_ERR_STATE, "request " + std::to_string(id) + " is being fetched");
      // pinned: its memory stays allocated until the copy below is done
      it->second.fetching = true;
      off = it->second.off;
      bytes = it->second.bytes;
    }
    try {
      DeviceScope ds(s->dev);
      if (s->precision == CP3_FP64) {
        hip_check(hipMemcpy(dst, (const char*)s->arena.p + off, bytes, hipMemcpyDeviceToHost),
                  "result copy");
      } else {
        std::vector<float> tmp(bytes / sizeof(float));
        hip_check(hipMemcpy(tmp.data(), (const char*)s->arena.p + off, bytes,
                            hipMemcpyDeviceToHost),
                  "result copy");
        for (size_t i = 0; i < tmp.size(); ++i) dst[i] = (double)tmp[i];
      }
    } catch (...) {
      std::lock_guard<std::mutex> lk(s->mu);
      auto it = s->reqs.find(id);
      if (it != s->reqs.end()) it->second.fetching = false;
      throw;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it != s->reqs.end()) {
      if (it->second.has_mem) s->alloc->release(it->second.off, it->second.bytes);
      s->reqs.erase(it);
    }
    s->cv_work.notify_one();
  });
}

int cp3_server_result_device(cp3_server* s, int64_t id, const void** dptr, int* elem) {
  return guard([&] {
    need(s);
    need_ptr(dptr, "dptr");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) throw Error(CP3_ERR_RANGE, "unknown request " + std::to_string(id));
    if (it->second.state != kDone) throw Error(CP3_ERR_STATE, "request is not finished");
    *dptr = (const char*)s->arena.p + it->second.off;
    if (elem) *elem = s->elem();
  });
}

int cp3_server_release(cp3_server* s, int64_t id) {
  return guard([&] {
    need(s);
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) throw Error(CP3_ERR_RANGE, "unknown request " + std::to_string(id));
    if (it->second.state == kPending || it->second.state == kRunning)
      throw Error(CP3_ERR_STATE, "request is not finished");
    if (it->second.fetching)
      throw Error(CP3_ERR_STATE, "request " + std::to_string(id) + " is being fetched");
    if (it->second.has_mem) s->alloc->release(it->second.off, it->second.bytes);
    s->reqs.erase(it);
    s->cv_work.notify_one();
  });
}

int cp3_server_latency(cp3_server* s, int64_t id, double* queue_us, double* total_us) {
  return guard([&] {
    need(s);
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) throw Error(CP3_ERR_RANGE, "unknown request " + std::to_string(id));
    if (it->second.state != kDone) throw Error(CP3_ERR_STATE, "request is not finished");
    if (queue_us) *queue_us = us_between(it->second.t_submit, it->second.t_launch);
    if (total_us) *total_us = us_between(it->second.t_submit, it->second.t_done);
  });
}

int cp3_server_stats(cp3_server* s, int64_t* completed, int64_t* evals, int64_t* batches,
                     int64_t* mem_in_use, int64_t* mem_peak, int64_t* peak_inflight) {
  return guard([&] {
    need(s);
    std::lock_guard<std::mutex> lk(s->mu);
    if (completed) *completed = s->n_done;
    if (evals) *evals = s->evals_done;
    if (batches) *batches = s->n_batches;
    if (mem_in_use) *mem_in_use = (int64_t)s->alloc->used();
    if (mem_peak) *mem_peak = (int64_t)s->peak_mem;
    if (peak_inflight) *peak_inflight = s->peak_inflight;
  });
}

}  // extern "C"
