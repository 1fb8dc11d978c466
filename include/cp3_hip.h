/*
 * C ABI of the MI355X three-flavour oscillation propagator (libcp3hip.so).
 *
 * This is the only interface the HIP library exports.  It replaces the
 * device-facing part of the reference's GPU propagators:
 *   AtmosCudaPropagatorSingle  (atmoscudapropagator.cuh:42-145,
 *                               atmoscudapropagator.cu:30-325)
 *   BeamCudaPropagatorSingle   (beamcudapropagator.cuh:41-121,
 *                               beamcudapropagator.cu:29-214)
 * and their physics core (physics.hpp:789-1394).  The C++ classes in
 * include/cudaprob3_amd/hip_propagator.hpp wrap it with the reference's
 * method names and exceptions.
 *
 * Conventions
 *   - Every call returns an int status: CP3_OK (0) or a negative CP3_ERR_*.
 *     Nothing throws or exits across this boundary; cp3_last_error() gives
 *     the message of the calling thread's last failure.
 *   - Units are the reference setters': GeV, km, g/cm^3, radians, eV^2.
 *   - Channel t = a*3+b is P(a->b) with e, mu, tau = 0, 1, 2
 *     (types.hpp:23-27).  Neutrino type 0 = neutrino, 1 = antineutrino
 *     (types.hpp:29).
 *   - Device results are laid out [9][nC][nE] (atmospheric) or [9][nE] (beam),
 *     as the reference kernels write them (physics.hpp:1139-1140, 1248-1249);
 *     elements are double for CP3_FP64 and float for CP3_FP32 / CP3_MIXED.
 *   - One object is used by one host thread at a time; objects share no state,
 *     so different objects may run concurrently on their own streams.
 */
#ifndef CP3_HIP_H
#define CP3_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Precision forms. */
#define CP3_FP64 0  /* FP64 throughout                                   */
#define CP3_FP32 1  /* FP32 throughout                                   */
#define CP3_MIXED 2 /* FP64 eigen-solve and phases, FP32 amplitude work  */

/* Status codes. */
#define CP3_OK 0
#define CP3_ERR_INVALID (-1) /* bad argument; the reference throws runtime_error */
#define CP3_ERR_STATE (-2)   /* call order; the reference throws runtime_error */
#define CP3_ERR_RANGE (-3)   /* index out of range (getProbability)             */
#define CP3_ERR_DEVICE (-4)  /* HIP runtime failure; the reference exits       */
#define CP3_ERR_NOMEM (-5)   /* device or pinned allocation failed             */
#define CP3_ERR_CAPACITY (-6) /* admission refused: memory cap reached         */

typedef struct cp3_atmos cp3_atmos;
typedef struct cp3_beam cp3_beam;

/* Message of the calling thread's last failing call ("" if none). */
const char* cp3_last_error(void);
/* Library version, major*10000 + minor*100 + patch. */
int cp3_version(void);
/* Number of visible HIP devices. */
int cp3_device_count(int* n);

/* ---------------------------------------------------------------------------
 * Atmospheric propagator (one grid, one parameter set per calculate).
 * ------------------------------------------------------------------------- */

/* AtmosCudaPropagatorSingle(id, n_cosines, n_energies)
 * (atmoscudapropagator.cu:30-62).  precision: CP3_FP64 / CP3_FP32 /
 * CP3_MIXED.  Allocates the device result buffer and a stream.  An empty grid
 * (n_cosines or n_energies <= 0) is CP3_ERR_INVALID here; the reference
 * accepts it and then launches an empty grid, a CUDA error it ends the
 * process on. */
int cp3_atmos_create(int device, int n_cosines, int n_energies, int precision,
                     cp3_atmos** out);
/* ~AtmosCudaPropagatorSingle (atmoscudapropagator.cu:71-76). */
int cp3_atmos_destroy(cp3_atmos* p);

/* Propagator::setEnergyList (propagator.hpp:213-218): n must equal
 * n_energies. */
int cp3_atmos_set_energy_list(cp3_atmos* p, const double* energies, int n);
/* AtmosCpuPropagator::setCosineList (atmoscpupropagator.hpp:368-381): n must
 * equal n_cosines. */
int cp3_atmos_set_cosine_list(cp3_atmos* p, const double* cosines, int n);
/* Propagator::setMNSMatrix (propagator.hpp:148-183): delta is negated when
 * kNuType < 0. */
int cp3_atmos_set_mns_matrix(cp3_atmos* p, double theta12, double theta13,
                             double theta23, double dcp, int kNuType);
/* Propagator::setNeutrinoMasses (propagator.hpp:188-209); dm23sq is
 * Delta m^2_32. */
int cp3_atmos_set_neutrino_masses(cp3_atmos* p, double dm12sq, double dm23sq);
/* AtmosCpuPropagator::setDensity(radii, rhos, yps)
 * (atmoscpupropagator.hpp:116-170): rows sorted either way, the density on a
 * row belongs to the shell whose outer radius is that row.  No shell cap. */
int cp3_atmos_set_density(cp3_atmos* p, const double* radii_km,
                          const double* rhos, const double* yps, int n);
/* AtmosCpuPropagator::setDensity(radii, a, b, c, yps)
 * (atmoscpupropagator.hpp:181-245): the shell whose outer radius is row i has
 * density a[i] R^2 + b[i] R + c[i] (g/cm^3, R in km), averaged along each
 * trajectory segment (physics.hpp:653-739).  b and c need n entries too (the
 * reference checks only a).  Such a model solves per trajectory and layer,
 * in the shared solve mode with one cosine per lane whatever the tuning and
 * solve mode asked for. */
int cp3_atmos_set_density_poly(cp3_atmos* p, const double* radii_km, const double* a,
                               const double* b, const double* c, const double* yps,
                               int n);
/* AtmosCpuPropagator::setDensityFromFile (atmoscpupropagator.hpp:250-344):
 * 3-column (radius, density, Ye) or 5-column (radius, a, b, c, Ye: a
 * polynomial model as above) files; other formats return CP3_ERR_INVALID. */
int cp3_atmos_set_density_from_file(cp3_atmos* p, const char* path);
/* setChemicalComposition (atmoscpupropagator.hpp:481-490): n must equal the
 * number of model rows; values in the stored (outermost first) order. */
int cp3_atmos_set_chemical_composition(cp3_atmos* p, const double* yps, int n);
/* setProductionHeight (atmoscpupropagator.hpp:386-393): km; the cosine list
 * must be set first. */
int cp3_atmos_set_production_height(cp3_atmos* p, double height_km);
/* getNlayerBoundaries (atmoscpupropagator.hpp:85): number of model rows. */
int cp3_atmos_get_nlayer_boundaries(cp3_atmos* p, int* n);

/* SetNumberOfProductionHeightBinsForAveraging (atmoscpupropagator.hpp:88-108):
 * 0..20 bins (CP3_ERR_INVALID above 20).  Like the reference, any n_bins >= 1
 * turns production-height averaging off. */
int cp3_atmos_set_production_height_bins(cp3_atmos* p, int n_bins);
/* No counterpart in the reference class: turns production-height averaging on
 * (on != 0) or off.  The reference class keeps this flag protected
 * (atmoscpupropagator.hpp:532) and never sets it, while its physics core
 * takes it as an argument (physics.hpp:804) and implements the averaging
 * (physics.hpp:997-1067).  Either way a list set before is dropped. */
int cp3_atmos_set_production_height_averaging(cp3_atmos* p, int on);
/* setProductionHeightList (atmoscpupropagator.hpp:395-427): n_prob =
 * n_bins * 2 * 3 * n_energies * n_cosines bin probabilities ordered
 * [neutrino type][initial flavour][energy][cosine][bin] and n_edges =
 * n_bins + 1 bin edges in km.  CP3_ERR_STATE unless averaging is on,
 * CP3_ERR_INVALID on other sizes.  calculate then returns CP3_ERR_STATE while
 * averaging is on without a list.  Each (cosine, energy) point averages over
 * the path-length shifts of the bins, weighted per initial flavour; the sinc
 * of a bin's width is sin(x)/x for |x| >= 1e-6 (the reference's CPU build
 * truncates |x| to an integer in that test, math.hpp:50, and takes its series
 * for every |x| < 1). */
int cp3_atmos_set_production_height_list(cp3_atmos* p, const double* prob,
                                         int64_t n_prob, const double* edges_km,
                                         int n_edges);

/* calculateProbabilities (atmoscudapropagator.cu:202-205, 266-316):
 * launches and waits. */
int cp3_atmos_calculate(cp3_atmos* p, int nutype);
/* The launch alone (calculateAtmosProbabilitiesAsync, .cu:266-310) and the
 * wait (waitForCompletion, .cu:313-316). */
int cp3_atmos_calculate_async(cp3_atmos* p, int nutype);
int cp3_atmos_wait(cp3_atmos* p);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under hipGraph stream capture, tight serving loops, harnesses
 * that forbid hipMalloc): cp3_atmos_calculate_async with nothing allocated.
 * The kernels are the same and so are the results (bitwise); they are
 * written to the caller's device buffer d_out ([9][nC][nE], double for
 * CP3_FP64, float otherwise), not to the object's own result, which
 * getProbability keeps reading.  The slab table and job descriptor of the
 * default solve mode live in a static device array of 16 MiB per GPU
 * (512 B for the descriptor, the rest for the table: (shells + 1) x
 * (nE rounded up to 64) x 40 B in mixed and FP32 precision, 64 B in FP64);
 * a grid whose table exceeds it returns CP3_ERR_INVALID, before any launch.
 * The object's device inputs must be current: after a setter that changed
 * the energies, cosines, Earth model or production height, one
 * cp3_atmos_calculate(_async) uploads them, and until then this returns
 * CP3_ERR_STATE.  Launches on the object's stream (cp3_atmos_get_stream)
 * and records no timing events; cp3_atmos_wait waits.  Every call on one
 * GPU shares that GPU's array: calls that overlap on one GPU (different
 * objects, or streams not ordered by the caller) race and must be
 * serialized. */
int cp3_atmos_calculate_async_static(cp3_atmos* p, int nutype, void* d_out);

/* getProbability(ic, ie, t) (atmoscudapropagator.cu:222-235): the first
 * access to channel t after a calculate copies that channel ([nC][nE]) to
 * pinned host memory (the reference copies all nine at once; the values are
 * the same). */
int cp3_atmos_get_probability(cp3_atmos* p, int index_cosine,
                              int index_energy, int t, double* out);
/* getProbabilityArr (atmoscudapropagator.cu:238-257): n_energies*n_cosines
 * values of channel t, energy outer, cosine inner. */
int cp3_atmos_get_probability_arr(cp3_atmos* p, double* out, int t);
/* Device pointer of the [9][nC][nE] result and its element size (8 or 4). */
int cp3_atmos_result_device(cp3_atmos* p, const void** dptr, int* elem_bytes);
/* HIP stream (hipStream_t) the object launches on. */
int cp3_atmos_get_stream(cp3_atmos* p, void** stream);
/* Milliseconds of the last kernel, measured with HIP events around it. */
int cp3_atmos_last_kernel_ms(cp3_atmos* p, float* ms);
/* Kernel tuning: trajectories per lane (cosine block 1, 2, 4 or 8) and waves
 * per workgroup sharing each eigen-solve through LDS (1, 2 or 4; shared
 * solve mode only); 0 = automatic.  Results do not depend on it (bitwise). */
int cp3_atmos_set_tuning(cp3_atmos* p, int cosine_block, int group_waves);
/* Where the per-(energy, shell) eigen-solutions come from, as for the batch
 * (CP3_SOLVE_SHARED, CP3_SOLVE_TABLE; CP3_SOLVE_AUTO = the table, which for
 * one grid takes (shells + 1) x energies x 40 B in mixed precision).  No
 * counterpart in the reference; results do not depend on it (bitwise). */
int cp3_atmos_set_solve_mode(cp3_atmos* p, int mode);

/* ---------------------------------------------------------------------------
 * Beam propagator (one energy list, one parameter set per calculate).
 * ------------------------------------------------------------------------- */

/* BeamCudaPropagatorSingle(id, n_energies, n_threads)
 * (beamcudapropagator.cu:29-59); the thread count has no GPU meaning. */
int cp3_beam_create(int device, int n_energies, int precision, cp3_beam** out);
int cp3_beam_destroy(cp3_beam* p);
int cp3_beam_set_energy_list(cp3_beam* p, const double* energies, int n);
int cp3_beam_set_mns_matrix(cp3_beam* p, double theta12, double theta13,
                            double theta23, double dcp, int kNuType);
int cp3_beam_set_neutrino_masses(cp3_beam* p, double dm12sq, double dm23sq);
/* BeamCpuPropagator::setDensity (beamcpupropagator.hpp:84-86): g/cm^3;
 * Ye = 0.5 is applied in the calculation (physics.hpp:1219). */
int cp3_beam_set_density(cp3_beam* p, double rho);
/* BeamCpuPropagator::setPathLength (beamcpupropagator.hpp:89-91): km. */
int cp3_beam_set_path_length(cp3_beam* p, double length_km);
int cp3_beam_calculate(cp3_beam* p, int nutype);
int cp3_beam_calculate_async(cp3_beam* p, int nutype);
int cp3_beam_wait(cp3_beam* p);
/* getProbability(ie, t) (beamcudapropagator.cu:146-159): the first access
 * to channel t after a calculate copies that channel to pinned host memory. */
int cp3_beam_get_probability(cp3_beam* p, int index_energy, int t,
                             double* out);
/* getProbabilityArr (beamcudapropagator.cu:162-178). */
int cp3_beam_get_probability_arr(cp3_beam* p, double* out, int t);
int cp3_beam_result_device(cp3_beam* p, const void** dptr, int* elem_bytes);
int cp3_beam_get_stream(cp3_beam* p, void** stream);
int cp3_beam_last_kernel_ms(cp3_beam* p, float* ms);


/* ---------------------------------------------------------------------------
 * Batched engine: many parameter sets (MCMC walkers), each on a resident grid
 * of its own size, neutrino and antineutrino, in one persistent launch per
 * kind (atmospheric, beam), optionally replayed from a captured hipGraph.
 * No reference counterpart: the reference runs one parameter set per launch
 * through process-global constant memory (physics.hpp:115-125, 1307-1352).
 * ------------------------------------------------------------------------- */

typedef struct cp3_batch cp3_batch;

#define CP3_JOB_ATMOS 0
#define CP3_JOB_BEAM 1

/* One evaluation: a parameter set (setter units) for one neutrino type on a
 * grid made by cp3_batch_add_*_grid.  kNuType as setMNSMatrix takes it. */
typedef struct cp3_job {
  int32_t kind;    /* CP3_JOB_ATMOS or CP3_JOB_BEAM */
  int32_t grid;    /* grid id of that kind */
  int32_t nutype;  /* 0 neutrino, 1 antineutrino */
  int32_t knutype; /* sign of kNuType: delta is negated when < 0 */
  double theta12, theta13, theta23, dcp;
  double dm12sq, dm23sq;
} cp3_job;

int cp3_batch_create(int device, int precision, cp3_batch** out);
int cp3_batch_destroy(cp3_batch* b);
/* Resident atmospheric grid: Earth model rows (either order, as setDensity),
 * cosines, energies and production height, reduced once to the trajectory
 * table.  Returns its id. */
int cp3_batch_add_atmos_grid(cp3_batch* b, const double* cosines, int n_cosines,
                             const double* energies, int n_energies,
                             const double* radii_km, const double* rhos,
                             const double* yps, int n_rows, double height_km,
                             int* grid_id);
/* Resident beam grid: energies, density (g/cm^3; Ye = 0.5 applied) and path
 * length (km). */
int cp3_batch_add_beam_grid(cp3_batch* b, const double* energies,
                            int n_energies, double rho, double length_km,
                            int* grid_id);
/* Sets the job list: allocates the resident outputs, builds the work list
 * and uploads the per-job constants. */
int cp3_batch_set_jobs(cp3_batch* b, const cp3_job* jobs, int n_jobs);
/* New parameter values for the same job list (same kinds and grids): only
 * the per-job constants are rebuilt and uploaded. */
int cp3_batch_update_jobs(cp3_batch* b, const cp3_job* jobs, int n_jobs);
/* Tuning: cosine block (1, 2, 4, 8), waves per workgroup sharing each
 * eigen-solve (1, 2, 4) and resident workgroups per CU of the persistent grid
 * (-1: one workgroup per work unit instead); 0 = automatic.  Results do not
 * depend on it (bitwise). */
int cp3_batch_set_tuning(cp3_batch* b, int cosine_block, int group_waves,
                         int workgroups_per_cu);
/* Where the per-(energy, shell) eigen-solutions come from:
 *   CP3_SOLVE_SHARED - solved inside the propagation kernel, each solve
 *                      shared by the cosine block x group waves of a work
 *                      unit through LDS;
 *   CP3_SOLVE_TABLE  - solved once per (job, shell, energy) by a pre-pass
 *                      into a table in device memory (about
 *                      (shells + 1) x energies x 40 bytes per job, mixed
 *                      precision), read by every cosine; group_waves is 1;
 *   CP3_SOLVE_AUTO   - the table when the tables fit in 16 GiB (default).
 * Results do not depend on the mode (bitwise). */
#define CP3_SOLVE_AUTO 0
#define CP3_SOLVE_SHARED 1
#define CP3_SOLVE_TABLE 2
int cp3_batch_set_solve_mode(cp3_batch* b, int mode);
/* The tuning in effect for the current job list (any pointer may be NULL). */
int cp3_batch_get_tuning(cp3_batch* b, int* cosine_block, int* group_waves,
                         int* workgroups_per_cu, int* solve_mode);
/* Times every candidate tuning on the current job list (reps launches each),
 * over both solve modes unless one was set, and keeps the fastest; reports
 * it (any pointer may be NULL; cp3_batch_get_tuning gives the mode). */
int cp3_batch_autotune(cp3_batch* b, int reps, int* cosine_block,
                       int* group_waves, int* workgroups_per_cu);
/* Capture the launch into a hipGraph on the next run and replay it after. */
int cp3_batch_use_graph(cp3_batch* b, int enable);
/* Enqueue the whole batch on the batch's stream; cp3_batch_wait waits. */
int cp3_batch_run(cp3_batch* b);
int cp3_batch_wait(cp3_batch* b);
/* Milliseconds between events around the last run (kernels only). */
int cp3_batch_last_ms(cp3_batch* b, float* ms);
/* Milliseconds of the last run's phases, from events recorded between the
 * launches on the batch's stream: ms[0] slab-table pre-pass, ms[1]
 * atmospheric propagation, ms[2] packed tail tiles, ms[3] beam (about 0,
 * within a few microseconds, for a phase that launched nothing). */
int cp3_batch_phase_ms(cp3_batch* b, float* ms);
/* Device result of one job ([9][nC][nE] or [9][nE]) and its element size. */
int cp3_batch_result_device(cp3_batch* b, int job, const void** dptr,
                            int* elem_bytes);
/* Copies one job's result to host as double, layout as on the device. */
int cp3_batch_copy_result(cp3_batch* b, int job, double* host_dst);
/* Work units, evaluations and output bytes of the current job list. */
int cp3_batch_stats(cp3_batch* b, int64_t* units, int64_t* evals,
                    int64_t* out_bytes);
int cp3_batch_get_stream(cp3_batch* b, void** stream);


/* ---------------------------------------------------------------------------
 * Request server: continuous batching of single evaluations (fit or MCMC
 * requests) on one device.  No reference counterpart.
 *
 * A scheduler thread admits pending requests, in arrival order, into free
 * launch slots (one stream each) as long as their outputs fit in a device
 * arena of mem_cap_bytes; batches on different slots overlap on the device.
 * Outputs stay resident until fetched or released.  A request's result does
 * not depend on the other requests of its batch (bitwise).  All functions
 * are thread-safe.
 * ------------------------------------------------------------------------- */

typedef struct cp3_server cp3_server;

int cp3_server_create(int device, int precision, size_t mem_cap_bytes,
                      int n_slots, int max_batch_jobs, cp3_server** out);
/* Drains admitted work, fails the rest, frees everything. */
int cp3_server_destroy(cp3_server* s);
/* Kernel tuning for every batch (cosine block, group waves).  By default the
 * cosine block is 1 in table mode and 2 in shared mode, with 4 group waves. */
int cp3_server_set_tuning(cp3_server* s, int cosine_block, int group_waves);
/* Solve mode of the atmospheric batches: CP3_SOLVE_TABLE (and AUTO, the
 * default) pre-solves each request's slabs into a table that lives in the
 * memory cap while the request runs ((shells + 1) x energies x 40 B in mixed
 * precision); CP3_SOLVE_SHARED solves inside the propagation kernel.  Results
 * do not depend on it (bitwise).  No counterpart in the reference. */
int cp3_server_set_solve_mode(cp3_server* s, int mode);
int cp3_server_add_atmos_grid(cp3_server* s, const double* cosines,
                              int n_cosines, const double* energies,
                              int n_energies, const double* radii_km,
                              const double* rhos, const double* yps,
                              int n_rows, double height_km, int* grid_id);
int cp3_server_add_beam_grid(cp3_server* s, const double* energies,
                             int n_energies, double rho, double length_km,
                             int* grid_id);
/* Queues one job; CP3_ERR_CAPACITY if its output can never fit the cap. */
int cp3_server_submit(cp3_server* s, const cp3_job* job, int64_t* request_id);
/* Waits until the request finished (timeout_ms < 0: no limit). */
int cp3_server_wait(cp3_server* s, int64_t request_id, int timeout_ms);
/* Copies a finished request's result to host as double (device layout) and
 * releases it. */
int cp3_server_fetch(cp3_server* s, int64_t request_id, double* host_dst);
/* Device pointer of a finished request's result (valid until released). */
int cp3_server_result_device(cp3_server* s, int64_t request_id,
                             const void** dptr, int* elem_bytes);
/* Releases a finished request's output memory without copying. */
int cp3_server_release(cp3_server* s, int64_t request_id);
/* Microseconds from submit to launch and from submit to completion. */
int cp3_server_latency(cp3_server* s, int64_t request_id, double* queue_us,
                       double* total_us);
int cp3_server_stats(cp3_server* s, int64_t* completed, int64_t* evals,
                     int64_t* batches, int64_t* mem_in_use, int64_t* mem_peak,
                     int64_t* peak_inflight_batches);

#ifdef __cplusplus
}
#endif

#endif /* CP3_HIP_H */
