/*
 * phx.h — C ABI of the MI355X-native PHOENIX raymarch (libphx.so).
 *
 * Plain C types only: no C++ or torch types cross this boundary.  Every entry
 * point returns 0 on success or a negative PHX_E* code; phx_last_error() holds
 * the message of the last failure on the calling thread (exceptions never
 * cross the ABI).  Device work runs on one GPU (device 0, or the device set
 * with phx_set_device).
 *
 * Reference interfaces each entry point replaces are cited as file:line into
 * nohgz/phoenix.
 */
#ifndef PHX_H
#define PHX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PHX_FREQS_MAX 32  /* frequencies per request (the CPU value of PHX_MAX_FREQS, src/core/numerics/macros.hpp:110-112; named apart so that phx.h and the reference headers share a translation unit) */
#define PHX_MAX_LEVELS 16 /* AMR levels (src/core/model/bhac.hpp:154-161) */

/* error codes */
#define PHX_OK 0
#define PHX_E_ARG -1     /* invalid argument / configuration */
#define PHX_E_IO -2      /* file missing, short read, bad format */
#define PHX_E_DEVICE -3  /* HIP runtime error */
#define PHX_E_RANGE -4   /* size beyond a supported maximum */

/* arithmetic of a render (the reference's PHX_ACCURACY, CMakeLists.txt:16-22, as a run-time choice) */
#define PHX_FP32 32  /* scalar = float, compensated (hi, lo) geodesic state */
#define PHX_FP64 64  /* scalar = double */
#define PHX_MIXED 96 /* FP32 bulk; FP64 near the horizon / poles and for the AMR inversion; Kahan Stokes sums */

/* One pixel at one frequency: core::pixel::FreqOut (src/core/render/pixel.hpp:26-29).
 * res[(i * img_height + j) * nfreqs + f]; I_nu in cgs already scaled by nu^3. */
typedef struct phx_freqout {
  double IQUV[4];
  double tau, tauF;
} phx_freqout;

/* What a render reports: device::Info (src/device/kernel.cuh:23-30), with CU/VGPR
 * names and the launch's own HIP-event time. */
typedef struct phx_info {
  char device[256];
  int32_t cus, blocks, block;          /* compute units; persistent grid x workgroup size */
  int32_t vgprs, sgprs, scratch_bytes; /* per-lane registers and private-segment bytes of the march kernel */
  int32_t lds_bytes, waves_per_cu;     /* static LDS per workgroup; occupancy the launch assumed */
  double upload_s, kernel_s, decode_s; /* table upload, march kernels (HIP events), device decode */
  int64_t steps, nplasma, bad;         /* sums over all rays (src/core/render/pixel.hpp:34-37); PHX_MIXED
                                          leaves out the samples in field-free leaves (all cut or dropped) */
  int64_t rays;                        /* rays x frequencies rendered */
  int64_t fp64_steps;                  /* PHX_MIXED: geodesic steps taken in FP64 (near the horizon / poles) */
  int32_t launches;                    /* march launches: frequency chunks, PHX_CHUNK chunks, two-phase record
                                          chunks, reruns included */
  int32_t reruns;                      /* two-phase launches run again after a ray found the record pool full */
} phx_info;

/* device::Progress (src/device/kernel.cuh:32): called on the caller's thread after each chunk */
typedef void (*phx_progress_fn)(int32_t jobs_done_in_chunk, void* user);

/* The run description: core::bhac::Model (src/core/model/bhac.hpp:131-173) with
 * UserParams (:62-129), GridParams (:50-55) and PlasmaUnits (:38-45) flattened,
 * every scalar held as double (a FP32 render rounds them exactly as the
 * reference's `scalar(x)` does) and explicit widths.  Filled by phx_load_config +
 * phx_load_grid + phx_load_dump, or by the caller. */
typedef struct phx_model {
  /* [camera] */
  double inclination, fov_x, fov_y, r_cam, t_init, pa, phi_cam;
  int32_t img_width, img_height, nfreqs, pad0;
  double frequencies[PHX_FREQS_MAX];
  /* [black_hole] */
  double mass, distance;
  /* [integration] */
  double stepsize, horizon_margin;
  int32_t max_steps, max_order;
  /* [plasma] */
  double m_unit, r_low, r_high, beta_c, sigma_cut, thetae_min, thetae_max, rt_outer;
  /* [grmhd] */
  int32_t nspin, bpol, nlevels, sfc;
  /* grid file and dump header; xprobmin/max of x2, x3 in radians (x 2 pi, bhac_io.hpp:200-203) */
  int32_t nxlone[3], nx[3], ng[3];
  int32_t nleafs, ndim, nwini, nws, cells;
  double xprobmin[3], xprobmax[3], hslope_grid;
  /* derived scalars (src/host/io/bhac_io.hpp:183-251) */
  double a, hslope, gam, Ne, Bu, rhigh, rlow, ibc, r_out, cut_in, cos_pa, sin_pa;
  double L, T_unit, rho_unit, U_unit, B_unit, Ne_unit;
  double freq[PHX_FREQS_MAX], hconst[PHX_FREQS_MAX], jscale[PHX_FREQS_MAX];
  double lvl_w[PHX_MAX_LEVELS][3], x_lo[4], x_hi[4];
  int32_t level_dims[PHX_MAX_LEVELS][3];
  /* host tables, owned by phx_load_dump (freed by phx_free_model) */
  double* lb;                          /* [nleafs][3] block lower corners */
  double* dxc;                         /* [nleafs][3] cell sizes */
  double* prim;                        /* [nleafs][cells][8] primitives, or NULL: the device decodes the dump */
  int32_t* level_map[PHX_MAX_LEVELS];  /* leaf per (level, i, j, k), -1 = go finer */
  int32_t* leaf_level;                 /* [nleafs] 1-based AMR level */
  int32_t* leaf_index;                 /* [nleafs][3] block index on its level */
} phx_model;

/* [output]/[visual]/[grmhd] keys that are not part of the model (host::config::Params, config.hpp:35-50) */
typedef struct phx_config {
  char outdir[1024];
  char image[16];      /* hdf5 | text | both | none */
  int32_t block_size;  /* HDF5 pixels per block side */
  char dump_file[1024];
  char grid_file[1024];
  int32_t gpu_decode, show_header;
} phx_config;

/* ---------------------------------------------------------------- host layer */

const char* phx_last_error(void);
const char* phx_version(void);

/* host::config::load_config (src/host/io/config.hpp:229-381): TOML subset, section.key=value
 * overrides, frequency forms, unknown / missing / duplicate key errors.  Fills the [camera],
 * [black_hole], [integration], [plasma], [grmhd] fields of *model. */
int phx_load_config(const char* path, const char* const* overrides, int32_t noverrides, phx_config* cfg,
                    phx_model* model);

/* host::bhac_io::load_bhac_grid (src/host/io/bhac_io.hpp:143-174) */
int phx_load_grid(const char* grid_file, phx_model* model);

/* host::bhac_io::load (src/host/io/bhac_io.hpp:331-499) without the host decode: header,
 * finalize_model units, forest walk, block geometry, level map.  model->prim stays NULL and
 * the device decodes the dump (the reference's grmhd.gpu_decode = true path). */
int phx_load_dump(phx_model* model, const char* dump_file);

/* Re-derives the per-request scalars (units, plasma, camera rotation, cut_in, r_out, per-frequency
 * hconst/jscale) from the user fields after a caller changed them: the per-request half of
 * host::bhac_io::finalize_model (src/host/io/bhac_io.hpp:224-250).  Grid fields are untouched. */
int phx_model_derive(phx_model* model);

/* host::bhac_io::free_model (src/host/io/bhac_io.hpp:314-322) */
void phx_free_model(phx_model* model);

/* host::jobs::schedule (src/host/render/jobs.hpp:18-46): i * img_height + j, Morton-ordered if morton */
int phx_schedule(const phx_model* model, int32_t morton, int32_t* jobs, int32_t njobs);

/* host::output::compute_spec (src/host/io/output.hpp:43-53): flux [Jy] of frequency fi, non-finite pixels skipped */
int phx_compute_spec(const phx_model* model, const phx_freqout* res, int32_t fi, double S[4]);

/* host::output::compute_tau (src/host/io/output.hpp:58-71): mean tau, tauF over the finite pixels */
int phx_compute_tau(const phx_model* model, const phx_freqout* res, int32_t fi, double T[2]);
/* host::output::write_image (output.hpp:77-94): text image of frequency fi, one row per pixel
 * (i outer): x, y [arcsec], I, Q, U, V [Jy / pixel], tau, tauF */
int phx_write_image_text(const phx_model* model, const phx_freqout* res, int32_t fi, const char* path);
/* host::output::write_spectrum (output.hpp:99-111): nu, I, Q, U, V [Jy], mean tau, mean tauF per frequency */
int phx_write_spectrum(const phx_model* model, const phx_freqout* res, const char* path);
/* host::output::write_image_hdf5 (output.hpp:144-167): RAPTOR-layout HDF5 image (datasets I%e, tau%e,
 * Q%e, U%e, V%e, tauF%e per frequency, alpha, beta; [nblocks][block_size^2] doubles, Jy/sr).  The HDF5
 * library (>= 1.10) is loaded at run time (PHX_HDF5_LIB, libhdf5.so, /opt/conda/lib/libhdf5.so). */
int phx_write_image_hdf5(const phx_model* model, const phx_freqout* res, int32_t block_size, const char* path);
/* write_output of src/main.cpp:52-86: outdir/img_data.h5 and/or outdir/image[_f].dat per output.image
 * (text when HDF5 fails), then outdir/spectrum.dat; the messages the reference prints go to log */
int phx_write_output(const phx_model* model, const phx_config* cfg, const phx_freqout* res, char* log,
                     int32_t log_len);

/* ---------------------------------------------------------------- device path */

int phx_set_device(int32_t device);

/* device::warmup_start (src/device/kernel.cu:160-162): creates the HIP context off the caller's thread */
int phx_dev_warmup(void);

/* device::render (src/device/kernel.cu:284-360; declared src/device/kernel.cuh:57-66).
 * Uploads the tables, decodes the dump on the device when model->prim is NULL, marches the
 * jobs (pixel indices i * img_height + j, any order) and copies res back.  Device memory is
 * per call.  progress (may be NULL) is called after each chunk of PHX_CHUNK jobs. */
int phx_dev_render(const phx_model* model, const char* dump_file, int32_t precision, const int32_t* jobs,
                   int32_t njobs, phx_freqout* res, phx_info* info, phx_progress_fn progress, void* user);

/* ---------------------------------------------------------------- resident snapshots
 * No counterpart in the reference (its device memory is per call, kernel.cu:131-155).
 * A snapshot keeps the decoded cell table, the block table and the flattened AMR map in HBM
 * across renders; requests then differ only in camera / plasma / frequency parameters. */
typedef struct phx_snapshot phx_snapshot;

/* cell tables to keep: bit set (PHX_FP32 and PHX_MIXED renders use the FP32 cell table, PHX_FP64
 * the FP64 one; the block tables of both precisions are always resident) */
#define PHX_TABLE_FP32 1
#define PHX_TABLE_FP64 2
int phx_snapshot_open(const phx_model* model, const char* dump_file, int32_t tables, phx_snapshot** out);
int phx_snapshot_close(phx_snapshot* snap);
/* bytes of HBM the snapshot holds */
int64_t phx_snapshot_bytes(const phx_snapshot* snap);

/* One render of `request` (camera / plasma / integration / frequency fields; the grid fields must
 * describe the snapshot's dump) on a resident snapshot.  res: host array of
 * img_width*img_height*nfreqs entries, or NULL to keep the result on the device (read it with
 * phx_snapshot_result).  jobs NULL = all pixels in Morton order. */
int phx_snapshot_render(phx_snapshot* snap, const phx_model* request, int32_t precision, const int32_t* jobs,
                        int32_t njobs, phx_freqout* res, phx_info* info);
/* copies the device result of the last phx_snapshot_render into res */
int phx_snapshot_result(phx_snapshot* snap, phx_freqout* res, int64_t nres);
/* The flux of every frequency of the last phx_snapshot_render, reduced on the device:
 * spec[f][4] = I, Q, U, V in Jy, the reference's compute_spec (src/host/io/output.hpp:43-53:
 * 1e23 dOm times the sum of IQUV over the pixels whose I + Q + U + V is finite) summed in a fixed
 * order on the device (equal to the host's sum to double rounding), so that a caller who needs
 * only the spectrum copies 4 doubles per frequency instead of the image. */
int phx_snapshot_flux(phx_snapshot* snap, double* spec, int32_t nfreqs);

/* Batch render: the n requests (complete models, e.g. phx_model_derive'd copies of one run) are
 * marched by one persistent launch over a single job queue holding every pixel of every request
 * and frequency chunk, so images of uneven size, inclination, frequency, m_unit and R_high share
 * the GPU.  Requests may differ only in camera, image size, frequencies, m_unit and R_high.
 * res[i] (res or res[i] may be NULL) receives request i's img_width*img_height*nfreqs results.
 * Each request's jobs fill whole waves of its own rays, and requests with 2 or more frequencies
 * render in two phases (a launch of each kind when both are present), so every image equals its
 * phx_snapshot_render bit for bit under the default launch settings.  Requests that share the
 * camera and image (parameter-sweep points) march their rays once.
 * The batched form of device::render (src/device/kernel.cu:284-360), which marches one image. */
int phx_snapshot_render_batch(phx_snapshot* snap, const phx_model* const* requests, int32_t n, int32_t precision,
                              phx_freqout* const* res, phx_info* info);

/* Allocation-free render.  No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Marches every pixel of `request` (img_width x img_height x nfreqs) on the resident snapshot and
 * writes the results to d_res, a device buffer the caller owns (res[(i * img_height + j) * nfreqs
 * + f], as phx_snapshot_render), then waits for the march.  The per-render buffers (jobs, ray
 * start records, tasks, cameras, counters) live in a fixed 104 MiB array in device memory that
 * each GPU the library runs on holds until the process ends; the call allocates nothing.
 * Limits, refused with PHX_E_RANGE: img_width * img_height * ceil(nfreqs / 4) at most 1,048,576
 * jobs in PHX_FP32 and PHX_MIXED, 545,259 in PHX_FP64; and a request whose march kernel would use
 * per-lane scratch memory (which the HIP runtime allocates at launch): at present more than one
 * frequency in PHX_MIXED or more than two in PHX_FP32 / PHX_FP64.  Calls on one GPU share the
 * array, so the library serializes them.  Samples are processed in place (phx_tune.queue 0). */
int phx_snapshot_render_noalloc(phx_snapshot* snap, const phx_model* request, int32_t precision,
                                phx_freqout* d_res, phx_info* info);

/* Captured render (HIP graph).  No counterpart in the reference, which launches its kernel once per
 * chunk of jobs from the host (kernel.cu:332-344).  phx_graph_create records the allocation-free
 * render of `request` (setup, counter resets, camera rays, side arguments, march: kernels only) into
 * a HIP graph on the snapshot's stream and instantiates it; results go to d_res, a device buffer
 * the caller owns (img_width * img_height * nfreqs entries, as phx_snapshot_render_noalloc).
 * phx_graph_launch stores the parameter point's tasks (one small kernel), replays the graph with
 * one hipGraphLaunch and waits for it: for `request` NULL the parameter point it holds, else a
 * new parameter-sweep point of the same camera, which may change
 * m_unit, R_high and the frequency values (units re-derived with phx_model_derive) but not the
 * camera, the image size, the number of frequencies or the launch-wide integration and cut
 * parameters (PHX_E_ARG).  The limits of phx_snapshot_render_noalloc apply, and replays are
 * serialized with its calls on the GPU (they share its device array).  Destroy before closing the
 * snapshot. */
typedef struct phx_graph phx_graph;
int phx_graph_create(phx_snapshot* snap, const phx_model* request, int32_t precision, phx_freqout* d_res,
                     phx_graph** out);
int phx_graph_launch(phx_graph* graph, const phx_model* request, phx_info* info);
int phx_graph_destroy(phx_graph* graph);

/* launch configuration of batch renders */
typedef struct phx_tune {
  int32_t queue;         /* -1 default (in place), 0 samples processed in place, 1 per-wave LDS sample queue */
  int32_t blocks_per_cu; /* 0 = occupancy limit */
  int32_t order;         /* 0 request after request, 1 requests interleaved one wave (64 jobs) at a time */
  int32_t fp64_wave;     /* PHX_MIXED: 1 = a lane's FP64 step takes its whole wave to FP64 */
  int32_t refill;        /* idle lanes a wave gathers before they take new jobs;
                            64 (or 0, the default) = the whole wave, 1 = each lane at once */
  int32_t escape;        /* 0 default by precision (PHX_MIXED 2, PHX_FP32 / PHX_FP64 1); 1 every ray marches to
                            r_out / cut_in / max_steps as the reference does; 2 a ray moving outward beyond
                            rt_outer ends there (it cannot emit again: the same results, fewer steps; taken
                            only when rt_outer lies outside the photon orbits) */
  int32_t block;         /* workgroup size of the in-place march: 64 or 256 (0 = 256); the queue uses 256
                            (the reference's compile-time GPU_BLOCK, CMakeLists.txt:25-28) */
  int32_t fetch;         /* steps per fetch: a wave with rays in flight takes jobs for its idle lanes only every
                            `fetch` passes (0 or 1 = every pass) */
} phx_tune;
/* Autotuner: times every configuration (LDS queue or in place, refill threshold, steps per fetch,
 * workgroup size, job order, whole-wave FP64 promotion) on every sample_stride-th wave of the
 * batch's jobs, keeps the fastest for the snapshot's batch renders, returns it and its sampled
 * rays*freq/s.
 * Replaces the reference's compile-time GPU_BLOCK / GPU_MINBLOCKS (kernel.cu:37-42). */
int phx_autotune(phx_snapshot* snap, const phx_model* const* requests, int32_t n, int32_t precision,
                 int32_t sample_stride, phx_tune* best, double* rate);
int phx_snapshot_set_tune(phx_snapshot* snap, const phx_tune* tune);

/* ---------------------------------------------------------------- serving loop
 * No counterpart in the reference (one image per run, src/main.cpp:143-230).  A worker thread
 * serves render requests on resident snapshots with continuous batching: requests are admitted
 * in arrival order while their result buffers fit hbm_budget; every launch marches the next
 * quantum_jobs jobs (pixels x frequency chunks) across all admitted requests of a snapshot, so
 * requests arriving while others are on the GPU join the next launch; each snapshot keeps two
 * launches queued, on two streams, so one runs while the next is built and the next fills the
 * compute units the running one's last rays leave idle.
 * Result buffers come from one pool of hbm_budget bytes allocated at start (no allocation per
 * request).  Latency = submit -> result copied to the caller's buffer.
 * The requests of one snapshot must share the launch-wide parameters of a batch render. */
typedef struct phx_server phx_server;
typedef struct phx_server_opts {
  int32_t precision;    /* PHX_FP32 (default), PHX_FP64 or PHX_MIXED */
  int32_t quantum_jobs; /* jobs per launch; 0 = 262144, at least 64 (launches take whole waves of 64 jobs) */
  int64_t hbm_budget;   /* bytes of admitted result buffers (the result pool); 0 = 8 GiB */
  int32_t max_active;   /* requests in flight at most; 0 = no limit but the budget */
  int32_t order;        /* jobs per launch taken 0: oldest request first, 1: least attained service first */
} phx_server_opts;
typedef struct phx_serve_stats {
  int64_t submitted, completed, failed, active, waiting;
  int64_t rays;     /* rays x frequencies of completed requests */
  int64_t launches; /* march launches */
  int64_t peak_bytes;
  double busy_s;    /* summed launch time (HIP events; overlapping snapshots counted once per tick) */
  double wall_s;    /* since start or the last reset */
  double p50_s, p99_s, mean_s, max_s; /* latency of completed requests */
} phx_serve_stats;
int phx_server_start(phx_snapshot* const* snaps, int32_t nsnaps, const phx_server_opts* opts, phx_server** out);
/* queues a request on snapshot `snap`; res (caller-owned, img_width*img_height*nfreqs entries,
 * may be NULL) receives the result when the request completes */
int phx_server_submit(phx_server* srv, int32_t snap, const phx_model* request, phx_freqout* res, int64_t* ticket);
/* the same, and spec (caller-owned, nfreqs x 4 doubles, may be NULL) receives the request's flux
 * per frequency, reduced on the device as phx_snapshot_flux does; with res NULL only the spectrum
 * crosses to the host */
int phx_server_submit_spec(phx_server* srv, int32_t snap, const phx_model* request, phx_freqout* res, double* spec,
                           int64_t* ticket);
/* waits up to timeout_s (< 0: forever) for a ticket; *done = 1 once its result is in res (the
 * ticket is then retired) */
int phx_server_wait(phx_server* srv, int64_t ticket, double timeout_s, int32_t* done, double* latency_s);
int phx_server_stats(phx_server* srv, phx_serve_stats* stats);
int phx_server_reset_stats(phx_server* srv);
/* stops the worker once its queued launches have finished; requests not completed fail with
 * "server stopped" (a phx_server_wait on them returns PHX_E_DEVICE), and the call returns after
 * every thread waiting in phx_server_wait has left it */
int phx_server_stop(phx_server* srv);

/* decoded primitives of a snapshot, [nleafs][cells][8] as double (for checks) */
int phx_snapshot_prims(phx_snapshot* snap, int32_t table, double* out, int64_t n);

#ifdef __cplusplus
}
#endif
#endif /* PHX_H */
