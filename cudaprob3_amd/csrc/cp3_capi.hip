// C ABI implementation (include/cp3_hip.h): the single-grid atmospheric and
// beam propagators.  Host logic mirrors AtmosCpuPropagator / BeamCpuPropagator
// (validation, call order, Earth model handling) and AtmosCudaPropagatorSingle
// / BeamCudaPropagatorSingle (device residency, stream, lazy result copy);
// all state is per object.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <exception>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cp3_hip.h"
#include "cp3_capi_util.hpp"
#include "cp3_host.hpp"
#include "cp3_launch.hpp"

using cp3::capi::DevBuf;
using cp3::capi::DeviceScope;
using cp3::capi::Error;
using cp3::capi::guard;
using cp3::capi::hip_check;
using cp3::capi::PinnedBuf;

namespace {

constexpr int kVersion = 10000;  // 1.0.0

void check_precision(int precision) {
  if (precision != CP3_FP64 && precision != CP3_FP32 && precision != CP3_MIXED)
    throw Error(CP3_ERR_INVALID, "unknown precision " + std::to_string(precision));
}

void check_nutype(int nutype) {
  if (nutype != 0 && nutype != 1)
    throw Error(CP3_ERR_INVALID, "neutrino type must be 0 (neutrino) or 1 (antineutrino)");
}

void check_channel(int t) {
  if (t < 0 || t > 8) throw Error(CP3_ERR_RANGE, "probability type must be in 0..8");
}

void open_device(int device) {
  int n = 0;
  hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
  if (n == 0) throw Error(CP3_ERR_DEVICE, "No GPU found");
  if (device < 0 || device >= n)
    throw Error(CP3_ERR_INVALID,
                "The requested GPU Id " + std::to_string(device) + " is not available.");
}

// Oscillation parameters as the reference setters store them.
struct OscArgs {
  double th12 = 0, th13 = 0, th23 = 0, dcp = 0;
  int knu = 1;
  double dm12 = 0, dm23 = 0;
  cp3::SetConst make(int nutype) const {
    return cp3::make_set(th12, th13, th23, dcp, knu, dm12, dm23, nutype);
  }
};

// Largest cosine block per precision that keeps the state in registers at a
// useful occupancy, while leaving at least 4096 waves; a single grid then
// shares each eigen-solve between two waves.
void auto_tuning(int precision, const cp3::GridDev& g, int* nc, int* w) {
  const int cap = precision == CP3_FP64 ? 2 : (precision == CP3_FP32 ? 8 : 4);
  *w = 2;
  for (int c = cap; c > 1; c /= 2)
    if (cp3::atmos_units(g, c, *w) * *w >= 4096) {
      *nc = c;
      return;
    }
  *nc = 1;
  if (cp3::atmos_units(g, 1, *w) * *w < 4096) *w = 1;
}

// Slab-table mode: the largest cosine block up to a cap that still leaves
// 2048 work units (8 per CU), else 1.  The cap is 4, or on models of 32 or
// more shells 2 in mixed and 1 in FP32 precision (2000 x 2000 grids on
// MI355X: 12 shells, mixed block 4 / 2 / 1 = 0.077 / 0.080 / 0.083 ms;
// 59 shells, mixed 0.172 / 0.153 / 0.174 ms, FP32 0.168 / 0.151 / 0.147 ms;
// FP64 blocks 4 and 1 within 1% of each other on both).
int table_nc(int precision, const cp3::GridDev& g) {
  int cap = 4;
  if (g.S >= 32) cap = precision == CP3_MIXED ? 2 : (precision == CP3_FP32 ? 1 : 4);
  for (int c = cap; c > 1; c /= 2)
    if (cp3::cosine_blocks(g, c, 1) * cp3::energy_tiles(g) >= 2048) return c;
  return 1;
}

// Scratch of cp3_atmos_calculate_async_static: the job descriptor (first
// kScratchJobBytes) and the slab table of one calculation, in a static device
// array instead of memory the object allocates.  Each device holds its own
// copy until the process ends; host code reaches it only through the address
// hipGetSymbolAddress gives with that device current (the array's name is a
// host-side stand-in, not a device address).
constexpr size_t kScratchBytes = (size_t)16 << 20;
constexpr size_t kScratchJobBytes = 512;
static_assert(sizeof(cp3::AtmosJob) <= kScratchJobBytes, "job descriptor exceeds its slot");
constexpr int kScratchDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

struct ScratchTable {
  std::mutex mu;
  bool resolved[kScratchDevices] = {};
  void* ptr[kScratchDevices] = {};
  hipError_t err[kScratchDevices] = {};
};

ScratchTable& scratch_table() {
  static ScratchTable t;
  return t;
}

// Resolves `device`'s copy of the array once, with that device current.
void resolve_scratch(int device) {
  if (device < 0 || device >= kScratchDevices) return;
  ScratchTable& t = scratch_table();
  std::lock_guard<std::mutex> lk(t.mu);
  if (t.resolved[device]) return;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  t.err[device] = e;
  t.ptr[device] = e == hipSuccess ? ptr : nullptr;
  t.resolved[device] = true;
  if (e != hipSuccess) (void)hipGetLastError();
}

// The device address of `device`'s copy (resolved by cp3_atmos_create).
void* scratch_ptr(int device) {
  if (device < 0 || device >= kScratchDevices)
    throw Error(CP3_ERR_DEVICE, "static scratch supports devices 0.." +
                                    std::to_string(kScratchDevices - 1));
  ScratchTable& t = scratch_table();
  std::lock_guard<std::mutex> lk(t.mu);
  if (!t.resolved[device])
    throw Error(CP3_ERR_STATE, "static scratch of device " + std::to_string(device) +
                                   " was not resolved");
  if (!t.ptr[device])
    throw Error(CP3_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") +
                                    hipGetErrorString(t.err[device]));
  return t.ptr[device];
}

}  // namespace

struct cp3_atmos {
  int dev = 0, nC = 0, nE = 0, precision = CP3_FP64, nc_force = 0, w_force = 0;
  int solve_mode = CP3_SOLVE_AUTO;
  // slab-table mode: the table and the job descriptor on the device
  DevBuf d_slabs, d_job;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // launched: a calculation wrote d_out; stream_used: some calculation
  // (into d_out or a caller's buffer) may still read the device tables
  bool launched = false, stream_used = false;
  std::vector<double> energies, cosines;
  bool cos_set = false, height_set = false;
  double h_cm = 0.0;
  OscArgs osc;
  cp3::EarthModel earth;
  int n_layers = 0;
  DevBuf d_ev, d_perm, d_M, d_Lvac, d_seg, d_rye, d_out;
  PinnedBuf h_out;
  int S = 0, rye_cs = 0;
  bool table_dirty = true, energies_dirty = true;
  // bit t: channel t of the last calculation is in h_out
  uint32_t on_host = 0;
  // production-height averaging (atmoscpupropagator.hpp:88-108, 395-427,
  // 532-535): the flag, the bin count, the bin edges (km) and, on the device,
  // the reordered bin probabilities of both neutrino types and the per-cosine
  // bin shifts (rebuilt with the trajectory table)
  bool avg = false, avg_list = false;
  int n_bins = 0;
  std::vector<double> edges_km;
  DevBuf d_pha_w, d_pha_hmw;

  ~cp3_atmos() {
    DeviceScope ds(dev);
    // a calculation in flight still reads the tables freed below
    if (stream) (void)hipStreamSynchronize(stream);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }

  size_t out_bytes() const {
    return (size_t)9 * (size_t)nC * (size_t)nE * (size_t)cp3::out_elem_bytes(precision);
  }

  void upload_tables() {
    // the uploads are synchronous copies outside this object's stream, which
    // an asynchronous calculation may still be reading
    if (stream_used && (energies_dirty || table_dirty))
      hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    if (energies_dirty) {
      d_ev.upload(energies.data(), energies.size() * sizeof(double), dev);
      energies_dirty = false;
    }
    if (table_dirty) {
      cp3::TrajTable tt;
      cp3::build_traj_table(tt, earth, cosines.data(), nC, h_cm);
      S = tt.S;
      rye_cs = tt.rye_cs;
      d_perm.upload(tt.perm.data(), tt.perm.size() * sizeof(int32_t), dev);
      d_M.upload(tt.M.data(), tt.M.size() * sizeof(int32_t), dev);
      d_Lvac.upload(tt.Lvac.data(), tt.Lvac.size() * sizeof(double), dev);
      d_seg.upload(tt.seg.data(), tt.seg.size() * sizeof(double), dev);
      d_rye.upload(tt.rye.data(), tt.rye.size() * sizeof(double), dev);
      if (avg_list) upload_bins();
      table_dirty = false;
    }
  }

  void upload_bins() {
    std::vector<double> hmw;
    cp3::build_height_bins(hmw, cosines.data(), nC, h_cm, edges_km);
    if (!hmw.empty()) d_pha_hmw.upload(hmw.data(), hmw.size() * sizeof(double), dev);
  }

  cp3::GridDev grid() const {
    cp3::GridDev g{};
    g.ev = (const double*)d_ev.p;
    g.perm = (const int32_t*)d_perm.p;
    g.M = (const int32_t*)d_M.p;
    g.Lvac = (const double*)d_Lvac.p;
    g.seg = (const double*)d_seg.p;
    g.rye = (const double*)d_rye.p;
    g.nC = nC;
    g.nE = nE;
    g.S = S;
    g.rye_cs = rye_cs;
    return g;
  }

  // The call-order rules of calculateProbabilities.
  void check_ready() const {
    if (!height_set)
      throw Error(CP3_ERR_STATE,
                  "CudaPropagatorSingle::calculateProbabilities. production height was not set");
    // Without a model the trajectory table has no shells and no density row,
    // which every kernel reads (the reference CPU class would propagate
    // through the atmosphere segment alone, its GPU class through empty
    // device arrays).
    if (!earth.set)
      throw Error(CP3_ERR_STATE,
                  "CudaPropagatorSingle::calculateProbabilities. density model was not set "
                  "(setDensity or setDensityFromFile)");
    if (avg && !avg_list)
      throw Error(CP3_ERR_STATE,
                  "CpuPropagator::calculateProbabilities. production height array was not set, "
                  "but been requested to use production height averaging");
  }

  // Whether a calculation runs the slab-table pre-pass (and so needs a
  // table and a device job descriptor): not for the shared solve mode, nor
  // for a polynomial model, whose densities differ per cosine (the shared
  // solve with one cosine per wave and one wave per workgroup).
  bool uses_table() const { return !(solve_mode == CP3_SOLVE_SHARED || rye_cs != 0); }
  size_t table_bytes() const {
    return (size_t)cp3::slab_entries(grid()) * cp3::slab_entry_bytes(precision);
  }

  void launch(int nutype) {
    check_ready();
    DeviceScope ds(dev);
    upload_tables();
    void* slabs = nullptr;
    cp3::AtmosJob* dj = nullptr;
    if (uses_table()) {
      d_slabs.alloc(table_bytes(), dev);
      d_job.alloc(sizeof(cp3::AtmosJob), dev);
      slabs = d_slabs.p;
      dj = (cp3::AtmosJob*)d_job.p;
    }
    on_host = 0;
    enqueue(nutype, d_out.p, slabs, dj, true);
    launched = true;
  }

  // The kernels of one calculation on the object's stream, writing `out`
  // ([9][nC][nE]); in table mode the pre-pass fills `slabs` and stores the
  // job descriptor at `dj` for the propagation kernel.  `timed`: HIP events
  // around the kernels.  Allocates nothing.
  void enqueue(int nutype, void* out, void* slabs, cp3::AtmosJob* dj, bool timed) {
    cp3::AtmosJob job{};
    job.sc = osc.make(nutype);
    job.g = grid();
    job.out = out;
    job.unit_begin = 0;
    job.etiles = cp3::energy_tiles(job.g);
    const bool table = uses_table();
    if (timed) hip_check(hipEventRecord(ev0, stream), "hipEventRecord");
    if (avg) {
      // the bin probabilities of this neutrino type, one cosine per wave
      const int64_t plane = (int64_t)nC * nE;
      job.pha.w = (const char*)d_pha_w.p +
                  (size_t)nutype * 3 * n_bins * plane * cp3::out_elem_bytes(precision);
      job.pha.hmw = (const double*)d_pha_hmw.p;
      job.pha.nb = n_bins;
      if (!table) {
        hip_check(cp3::launch_atmos_single_avg(precision, job, stream),
                  "atmospheric kernel launch");
      } else {
        job.slabs = slabs;
        const int64_t units = cp3::cosine_blocks(job.g, 1, 1) * job.etiles;
        hip_check(cp3::launch_slab_table_single(precision, job, dj, stream),
                  "slab table launch");
        hip_check(cp3::launch_atmos_table_avg(precision, dj, units,
                                              (int)((units + 7) / 8 * 8), stream),
                  "atmospheric kernel launch");
      }
    } else if (!table) {
      int nc = 1, w = 1;
      auto_tuning(precision, job.g, &nc, &w);
      if (nc_force) nc = nc_force;
      if (w_force) w = w_force;
      if (rye_cs != 0) nc = w = 1;
      hip_check(cp3::launch_atmos_single(precision, nc, w, job, stream),
                "atmospheric kernel launch");
    } else {
      // the grid's slab table (a pre-pass) read by the propagation kernel
      const int nc = nc_force ? nc_force : table_nc(precision, job.g);
      job.slabs = slabs;
      const int64_t units = cp3::cosine_blocks(job.g, nc, 1) * job.etiles;
      hip_check(cp3::launch_slab_table_single(precision, job, dj, stream),
                "slab table launch");
      hip_check(cp3::launch_atmos_table(precision, nc, dj, 1, units,
                                        (int)((units + 7) / 8 * 8), stream),
                "atmospheric kernel launch");
    }
    if (timed) hip_check(hipEventRecord(ev1, stream), "hipEventRecord");
    stream_used = true;
  }

  // cp3_atmos_calculate_async_static: the same kernels into the caller's
  // device buffer, with the table and job descriptor in the static scratch.
  void launch_static(int nutype, void* out) {
    check_ready();
    if (energies_dirty || table_dirty)
      throw Error(CP3_ERR_STATE,
                  "cp3_atmos_calculate_async_static: the energies, cosines, Earth model or "
                  "production height changed since the last calculation; a "
                  "cp3_atmos_calculate(_async) call uploads them first");
    DeviceScope ds(dev);
    void* slabs = nullptr;
    cp3::AtmosJob* dj = nullptr;
    if (uses_table()) {
      const size_t need = kScratchJobBytes + table_bytes();
      if (need > kScratchBytes)
        throw Error(CP3_ERR_INVALID,
                    "cp3_atmos_calculate_async_static: this grid's slab table needs " +
                        std::to_string(need) + " bytes of scratch, more than the " +
                        std::to_string(kScratchBytes) +
                        " bytes of the static array (cp3_atmos_calculate_async has no limit)");
      char* base = (char*)scratch_ptr(dev);
      dj = (cp3::AtmosJob*)base;
      slabs = base + kScratchJobBytes;
    }
    enqueue(nutype, out, slabs, dj, false);
  }

  void wait() {
    DeviceScope ds(dev);
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
  }

  // Channel t ([nC][nE], contiguous in the [9][nC][nE] result) to pinned
  // host memory on its first access after a calculation.  The reference
  // copies all nine channels on the first access to any
  // (atmoscudapropagator.cu:222-235); copying by channel gives the same
  // values and moves a ninth of the bytes per channel read.
  void fetch(int t) {
    if (on_host & (1u << t)) return;
    DeviceScope ds(dev);
    const size_t cb = out_bytes() / 9;
    hip_check(hipMemcpyAsync((char*)h_out.p + (size_t)t * cb, (const char*)d_out.p + (size_t)t * cb,
                             cb, hipMemcpyDeviceToHost, stream),
              "result copy");
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    on_host |= 1u << t;
  }

  double value(size_t idx) const {
    if (precision == CP3_FP64) return ((const double*)h_out.p)[idx];
    return (double)((const float*)h_out.p)[idx];
  }
};

struct cp3_beam {
  int dev = 0, nE = 0, precision = CP3_FP64;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool launched = false;
  std::vector<double> energies;
  double rho = 0.0, L = 0.0;
  OscArgs osc;
  DevBuf d_ev, d_out;
  PinnedBuf h_out;
  bool energies_dirty = true;
  // bit t: channel t of the last calculation is in h_out
  uint32_t on_host = 0;

  ~cp3_beam() {
    DeviceScope ds(dev);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }
  size_t out_bytes() const {
    return (size_t)9 * (size_t)nE * (size_t)cp3::out_elem_bytes(precision);
  }
  void launch(int nutype) {
    DeviceScope ds(dev);
    if (energies_dirty) {
      // an asynchronous calculation may still read the old list
      if (launched) hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
      d_ev.upload(energies.data(), energies.size() * sizeof(double), dev);
      energies_dirty = false;
    }
    cp3::BeamJob job{};
    job.sc = osc.make(nutype);
    job.ev = (const double*)d_ev.p;
    job.nE = nE;
    job.rye = rho * cp3::kBeamYe;
    job.L = L;
    job.out = d_out.p;
    on_host = 0;
    hip_check(hipEventRecord(ev0, stream), "hipEventRecord");
    hip_check(cp3::launch_beam_single(precision, job, stream), "beam kernel launch");
    hip_check(hipEventRecord(ev1, stream), "hipEventRecord");
    launched = true;
  }
  void wait() {
    DeviceScope ds(dev);
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
  }
  // Channel t ([nE]) to pinned host memory on its first access after a
  // calculation (the reference copies all nine, beamcudapropagator.cu:146-158).
  void fetch(int t) {
    if (on_host & (1u << t)) return;
    DeviceScope ds(dev);
    const size_t cb = out_bytes() / 9;
    hip_check(hipMemcpyAsync((char*)h_out.p + (size_t)t * cb, (const char*)d_out.p + (size_t)t * cb,
                             cb, hipMemcpyDeviceToHost, stream),
              "result copy");
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    on_host |= 1u << t;
  }
  double value(size_t idx) const {
    if (precision == CP3_FP64) return ((const double*)h_out.p)[idx];
    return (double)((const float*)h_out.p)[idx];
  }
};

namespace {

template <class T>
T* need(T* p) {
  if (!p) throw Error(CP3_ERR_INVALID, "null propagator handle");
  return p;
}

void need_ptr(const void* p, const char* what) {
  if (!p) throw Error(CP3_ERR_INVALID, std::string("null pointer: ") + what);
}

float event_ms(hipEvent_t a, hipEvent_t b, int dev) {
  DeviceScope ds(dev);
  hip_check(hipEventSynchronize(b), "hipEventSynchronize");
  float ms = 0.f;
  hip_check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
  return ms;
}

// out[ie * nC + ic] = src[ic * nE + ie] (the reference's [E][cos] order of
// getProbabilityArr, atmoscudapropagator.cu:237-254) in 64 x 64 blocks, so
// that both sides stay in cache, with the element type fixed per call: on a
// 2000 x 2000 grid 3.0 ms a channel in FP64 and 2.0 ms from FP32 results on
// the MI355X host, against 3.4 and 2.7 ms element by element (about one host
// thread's memory bandwidth; the copy from the device takes 0.57 ms).
template <class T>
void transpose_to(double* out, const T* src, int nC, int nE) {
  constexpr int B = 64;
  for (int e0 = 0; e0 < nE; e0 += B) {
    const int e1 = e0 + B < nE ? e0 + B : nE;
    for (int c0 = 0; c0 < nC; c0 += B) {
      const int c1 = c0 + B < nC ? c0 + B : nC;
      for (int ie = e0; ie < e1; ++ie) {
        double* o = out + (size_t)ie * nC;
        for (int ic = c0; ic < c1; ++ic) o[ic] = (double)src[(size_t)ic * nE + ie];
      }
    }
  }
}

}  // namespace

extern "C" {

const char* cp3_last_error(void) { return cp3::capi::last_error().c_str(); }

int cp3_version(void) { return kVersion; }

int cp3_device_count(int* n) {
  return guard([&] {
    need_ptr(n, "n");
    int c = 0;
    hip_check(hipGetDeviceCount(&c), "hipGetDeviceCount");
    *n = c;
  });
}

// ------------------------------------------------------------- atmospheric

int cp3_atmos_create(int device, int n_cosines, int n_energies, int precision,
                     cp3_atmos** out) {
  return guard([&] {
    need_ptr(out, "out");
    *out = nullptr;
    if (n_cosines <= 0 || n_energies <= 0)
      throw Error(CP3_ERR_INVALID, "number of cosines and energies must be positive");
    check_precision(precision);
    open_device(device);
    std::unique_ptr<cp3_atmos> p(new cp3_atmos());
    p->dev = device;
    p->nC = n_cosines;
    p->nE = n_energies;
    p->precision = precision;
    p->energies.assign((size_t)n_energies, 0.0);
    p->cosines.assign((size_t)n_cosines, 0.0);
    DeviceScope ds(device);
    hip_check(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipEventCreate(&p->ev0), "hipEventCreate");
    hip_check(hipEventCreate(&p->ev1), "hipEventCreate");
    p->d_out.alloc(p->out_bytes(), device);
    p->h_out.alloc(p->out_bytes());
    // this device's copy of the static scratch (cp3_atmos_calculate_async_static)
    resolve_scratch(device);
    *out = p.release();
  });
}

int cp3_atmos_destroy(cp3_atmos* p) {
  return guard([&] { delete p; });
}

int cp3_atmos_set_energy_list(cp3_atmos* p, const double* e, int n) {
  return guard([&] {
    need(p);
    if (n != p->nE)
      throw Error(CP3_ERR_INVALID,
                  "Propagator::setEnergyList. Propagator was not created for this number of energy nodes");
    need_ptr(e, "energies");
    p->energies.assign(e, e + n);
    p->energies_dirty = true;
  });
}

int cp3_atmos_set_cosine_list(cp3_atmos* p, const double* c, int n) {
  return guard([&] {
    need(p);
    if (n != p->nC)
      throw Error(CP3_ERR_INVALID,
                  "Propagator::setCosineList. Propagator was not created for this number of cosine nodes");
    need_ptr(c, "cosines");
    p->cosines.assign(c, c + n);
    p->cos_set = true;
    p->table_dirty = true;
  });
}

int cp3_atmos_set_mns_matrix(cp3_atmos* p, double t12, double t13, double t23,
                             double dcp, int knu) {
  return guard([&] {
    need(p);
    p->osc.th12 = t12;
    p->osc.th13 = t13;
    p->osc.th23 = t23;
    p->osc.dcp = dcp;
    p->osc.knu = knu;
  });
}

int cp3_atmos_set_neutrino_masses(cp3_atmos* p, double dm12, double dm23) {
  return guard([&] {
    need(p);
    p->osc.dm12 = dm12;
    p->osc.dm23 = dm23;
  });
}

int cp3_atmos_set_density(cp3_atmos* p, const double* r, const double* rho,
                          const double* ye, int n) {
  return guard([&] {
    need(p);
    if (n < 0) throw Error(CP3_ERR_INVALID, "setDensity : negative size");
    if (n > 0) {
      need_ptr(r, "radii");
      need_ptr(rho, "rhos");
      need_ptr(ye, "yps");
    }
    std::vector<double> vr, vd, vy;
    if (n > 0) {
      vr.assign(r, r + n);
      vd.assign(rho, rho + n);
      vy.assign(ye, ye + n);
    }
    cp3::earth_set(p->earth, vr, vd, vy);
    p->n_layers = (int)p->earth.radii.size();
    p->table_dirty = true;
  });
}

int cp3_atmos_set_density_poly(cp3_atmos* p, const double* r, const double* a,
                               const double* b, const double* c, const double* ye,
                               int n) {
  return guard([&] {
    need(p);
    if (n < 0) throw Error(CP3_ERR_INVALID, "setDensity : negative size");
    std::vector<double> vr, va, vb, vc, vy;
    if (n > 0) {
      need_ptr(r, "radii");
      need_ptr(a, "a");
      need_ptr(b, "b");
      need_ptr(c, "c");
      need_ptr(ye, "yps");
      vr.assign(r, r + n);
      va.assign(a, a + n);
      vb.assign(b, b + n);
      vc.assign(c, c + n);
      vy.assign(ye, ye + n);
    }
    cp3::earth_set_poly(p->earth, vr, va, vb, vc, vy);
    p->n_layers = (int)p->earth.radii.size();
    p->table_dirty = true;
  });
}

int cp3_atmos_set_density_from_file(cp3_atmos* p, const char* path) {
  return guard([&] {
    need(p);
    need_ptr(path, "path");
    cp3::earth_load_file(p->earth, path);
    p->n_layers = (int)p->earth.radii.size();
    p->table_dirty = true;
  });
}

int cp3_atmos_set_chemical_composition(cp3_atmos* p, const double* ye, int n) {
  return guard([&] {
    need(p);
    if (n < 0 || (size_t)n != p->earth.yps.size())
      throw Error(CP3_ERR_INVALID,
                  "CpuPropagator::setChemicalComposition. Size of input list not equal to expectation.");
    if (n > 0) need_ptr(ye, "yps");
    cp3::earth_set_yps(p->earth, std::vector<double>(ye, ye + n));
    p->table_dirty = true;
  });
}

int cp3_atmos_set_production_height(cp3_atmos* p, double km) {
  return guard([&] {
    need(p);
    if (!p->cos_set)
      throw Error(CP3_ERR_STATE, "must set cosine list before production height");
    p->h_cm = km * 100000.0;
    p->height_set = true;
    p->table_dirty = true;
  });
}

int cp3_atmos_set_production_height_bins(cp3_atmos* p, int n_bins) {
  return guard([&] {
    need(p);
    if (n_bins > 20)
      throw Error(CP3_ERR_INVALID,
                  "SetNumberOfProductionHeightBinsForAveraging : invalid number of production "
                  "height bins");
    if (n_bins < 0)
      throw Error(CP3_ERR_INVALID, "negative number of production height bins");
    p->n_bins = n_bins;
    if (n_bins >= 1) p->avg = false;
  });
}

int cp3_atmos_set_production_height_averaging(cp3_atmos* p, int on) {
  return guard([&] {
    need(p);
    p->avg = on != 0;
    p->avg_list = false;
  });
}

int cp3_atmos_set_production_height_list(cp3_atmos* p, const double* prob, int64_t n_prob,
                                         const double* edges_km, int n_edges) {
  return guard([&] {
    need(p);
    if (!p->avg)
      throw Error(CP3_ERR_STATE,
                  "Propagator::setProductionHeightList. Trying to set Production Height "
                  "information but propagator is not expecting to use it");
    const int64_t want = (int64_t)p->n_bins * 2 * 3 * p->nE * p->nC;
    if (n_prob != want)
      throw Error(CP3_ERR_INVALID,
                  "Propagator::setProductionHeightList. Prob array is not the expected size");
    if (n_edges - 1 != p->n_bins)
      throw Error(CP3_ERR_INVALID,
                  "Propagator::setProductionHeightList. ProductionHeightBins array is not "
                  "expected size");
    if (n_prob > 0) need_ptr(prob, "prob");
    need_ptr(edges_km, "edges_km");
    DeviceScope ds(p->dev);
    // the previous calculation may still read the old lists
    hip_check(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
    if (n_prob > 0) {
      if (p->precision == CP3_FP64) {
        std::vector<double> w;
        cp3::reorder_height_prob(w, prob, p->nC, p->nE, p->n_bins);
        p->d_pha_w.upload(w.data(), w.size() * sizeof(double), p->dev);
      } else {
        std::vector<float> w;
        cp3::reorder_height_prob(w, prob, p->nC, p->nE, p->n_bins);
        p->d_pha_w.upload(w.data(), w.size() * sizeof(float), p->dev);
      }
    }
    p->edges_km.assign(edges_km, edges_km + n_edges);
    p->avg_list = true;
    // the bin shifts follow the cosines and the production height
    p->table_dirty = true;
  });
}

int cp3_atmos_get_nlayer_boundaries(cp3_atmos* p, int* n) {
  return guard([&] {
    need(p);
    need_ptr(n, "n");
    *n = p->n_layers;
  });
}

int cp3_atmos_calculate(cp3_atmos* p, int nutype) {
  return guard([&] {
    need(p);
    check_nutype(nutype);
    p->launch(nutype);
    p->wait();
  });
}

int cp3_atmos_calculate_async(cp3_atmos* p, int nutype) {
  return guard([&] {
    need(p);
    check_nutype(nutype);
    p->launch(nutype);
  });
}

int cp3_atmos_calculate_async_static(cp3_atmos* p, int nutype, void* d_out) {
  return guard([&] {
    need(p);
    check_nutype(nutype);
    need_ptr(d_out, "d_out");
    p->launch_static(nutype, d_out);
  });
}

int cp3_atmos_wait(cp3_atmos* p) {
  return guard([&] { need(p)->wait(); });
}

int cp3_atmos_get_probability(cp3_atmos* p, int ic, int ie, int t, double* out) {
  return guard([&] {
    need(p);
    need_ptr(out, "out");
    if (ic < 0 || ie < 0 || ic >= p->nC || ie >= p->nE)
      throw Error(CP3_ERR_RANGE, "CudaPropagatorSingle::getProbability. Invalid indices");
    check_channel(t);
    if (!p->launched) throw Error(CP3_ERR_STATE, "calculateProbabilities was not called");
    p->fetch(t);
    *out = p->value(((size_t)t * p->nC + (size_t)ic) * p->nE + (size_t)ie);
  });
}

int cp3_atmos_get_probability_arr(cp3_atmos* p, double* out, int t) {
  return guard([&] {
    need(p);
    need_ptr(out, "out");
    check_channel(t);
    if (!p->launched) throw Error(CP3_ERR_STATE, "calculateProbabilities was not called");
    p->fetch(t);
    const size_t base = (size_t)t * p->nC * p->nE;
    if (p->precision == CP3_FP64)
      transpose_to(out, (const double*)p->h_out.p + base, p->nC, p->nE);
    else
      transpose_to(out, (const float*)p->h_out.p + base, p->nC, p->nE);
  });
}

int cp3_atmos_result_device(cp3_atmos* p, const void** dptr, int* elem) {
  return guard([&] {
    need(p);
    need_ptr(dptr, "dptr");
    *dptr = p->d_out.p;
    if (elem) *elem = cp3::out_elem_bytes(p->precision);
  });
}

int cp3_atmos_get_stream(cp3_atmos* p, void** s) {
  return guard([&] {
    need(p);
    need_ptr(s, "stream");
    *s = (void*)p->stream;
  });
}

int cp3_atmos_last_kernel_ms(cp3_atmos* p, float* ms) {
  return guard([&] {
    need(p);
    need_ptr(ms, "ms");
    if (!p->launched) throw Error(CP3_ERR_STATE, "no kernel was launched");
    *ms = event_ms(p->ev0, p->ev1, p->dev);
  });
}

int cp3_atmos_set_tuning(cp3_atmos* p, int nc, int w) {
  return guard([&] {
    need(p);
    if (nc != 0 && !cp3::valid_nc(nc))
      throw Error(CP3_ERR_INVALID, "cosine block must be 0, 1, 2, 4 or 8");
    if (w != 0 && !cp3::valid_group(w))
      throw Error(CP3_ERR_INVALID, "group waves must be 0, 1, 2 or 4");
    p->nc_force = nc;
    p->w_force = w;
  });
}

int cp3_atmos_set_solve_mode(cp3_atmos* p, int mode) {
  return guard([&] {
    need(p);
    if (mode != CP3_SOLVE_AUTO && mode != CP3_SOLVE_SHARED && mode != CP3_SOLVE_TABLE)
      throw Error(CP3_ERR_INVALID, "solve mode must be CP3_SOLVE_AUTO, _SHARED or _TABLE");
    p->solve_mode = mode;
  });
}

// -------------------------------------------------------------------- beam

int cp3_beam_create(int device, int n_energies, int precision, cp3_beam** out) {
  return guard([&] {
    need_ptr(out, "out");
    *out = nullptr;
    if (n_energies <= 0) throw Error(CP3_ERR_INVALID, "number of energies must be positive");
    check_precision(precision);
    open_device(device);
    std::unique_ptr<cp3_beam> p(new cp3_beam());
    p->dev = device;
    p->nE = n_energies;
    p->precision = precision;
    p->energies.assign((size_t)n_energies, 0.0);
    DeviceScope ds(device);
    hip_check(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipEventCreate(&p->ev0), "hipEventCreate");
    hip_check(hipEventCreate(&p->ev1), "hipEventCreate");
    p->d_out.alloc(p->out_bytes(), device);
    p->h_out.alloc(p->out_bytes());
    *out = p.release();
  });
}

int cp3_beam_destroy(cp3_beam* p) {
  return guard([&] { delete p; });
}

int cp3_beam_set_energy_list(cp3_beam* p, const double* e, int n) {
  return guard([&] {
    need(p);
    if (n != p->nE)
      throw Error(CP3_ERR_INVALID,
                  "Propagator::setEnergyList. Propagator was not created for this number of energy nodes");
    need_ptr(e, "energies");
    p->energies.assign(e, e + n);
    p->energies_dirty = true;
  });
}

int cp3_beam_set_mns_matrix(cp3_beam* p, double t12, double t13, double t23,
                            double dcp, int knu) {
  return guard([&] {
    need(p);
    p->osc.th12 = t12;
    p->osc.th13 = t13;
    p->osc.th23 = t23;
    p->osc.dcp = dcp;
    p->osc.knu = knu;
  });
}

int cp3_beam_set_neutrino_masses(cp3_beam* p, double dm12, double dm23) {
  return guard([&] {
    need(p);
    p->osc.dm12 = dm12;
    p->osc.dm23 = dm23;
  });
}

int cp3_beam_set_density(cp3_beam* p, double rho) {
  return guard([&] { need(p)->rho = rho; });
}

int cp3_beam_set_path_length(cp3_beam* p, double L) {
  return guard([&] { need(p)->L = L; });
}

int cp3_beam_calculate(cp3_beam* p, int nutype) {
  return guard([&] {
    need(p);
    check_nutype(nutype);
    p->launch(nutype);
    p->wait();
  });
}

int cp3_beam_calculate_async(cp3_beam* p, int nutype) {
  return guard([&] {
    need(p);
    check_nutype(nutype);
    p->launch(nutype);
  });
}

int cp3_beam_wait(cp3_beam* p) {
  return guard([&] { need(p)->wait(); });
}

int cp3_beam_get_probability(cp3_beam* p, int ie, int t, double* out) {
  return guard([&] {
    need(p);
    need_ptr(out, "out");
    if (ie < 0 || ie >= p->nE)
      throw Error(CP3_ERR_RANGE, "CudaPropagatorSingle::getProbability. Invalid indices");
    check_channel(t);
    if (!p->launched) throw Error(CP3_ERR_STATE, "calculateProbabilities was not called");
    p->fetch(t);
    *out = p->value((size_t)t * p->nE + (size_t)ie);
  });
}

int cp3_beam_get_probability_arr(cp3_beam* p, double* out, int t) {
  return guard([&] {
    need(p);
    need_ptr(out, "out");
    check_channel(t);
    if (!p->launched) throw Error(CP3_ERR_STATE, "calculateProbabilities was not called");
    p->fetch(t);
    for (int ie = 0; ie < p->nE; ++ie) out[ie] = p->value((size_t)t * p->nE + (size_t)ie);
  });
}

int cp3_beam_result_device(cp3_beam* p, const void** dptr, int* elem) {
  return guard([&] {
    need(p);
    need_ptr(dptr, "dptr");
    *dptr = p->d_out.p;
    if (elem) *elem = cp3::out_elem_bytes(p->precision);
  });
}

int cp3_beam_get_stream(cp3_beam* p, void** s) {
  return guard([&] {
    need(p);
    need_ptr(s, "stream");
    *s = (void*)p->stream;
  });
}

int cp3_beam_last_kernel_ms(cp3_beam* p, float* ms) {
  return guard([&] {
    need(p);
    need_ptr(ms, "ms");
    if (!p->launched) throw Error(CP3_ERR_STATE, "no kernel was launched");
    *ms = event_ms(p->ev0, p->ev1, p->dev);
  });
}

}  // extern "C"
