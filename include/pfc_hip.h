/*
 * pfc_hip.h - C ABI of the MI355X PFC spectral stepper (libpfc_hip.so).
 *
 * The reference has no linkable API: its hot path is file-static C++ in
 * src/experiment/{misfit_uniform_time_seed_fftw_early,cvd,elastic}.cpp driven
 * through the FFTW C API (fftw_plan_dft_3d / fftw_execute, misfit.cpp:438-452).
 * Each entry point below replaces one reference routine or group of routines;
 * the citation names the reference lines it stands in for.  Plain pointers and
 * sizes only; host buffers are caller-owned, device buffers library-owned.
 * Every function returns 0 on success and a negative PFC_E* code on failure;
 * pfc_last_error() then describes the failure (thread-local).
 */
#ifndef PFC_HIP_H
#define PFC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PFC_MODEL_MISFIT 0
#define PFC_MODEL_CVD 1

#define PFC_PREC_FP64 0 /* plain FP64 drop-in path */
#define PFC_PREC_FP32 1 /* FP32 FFT/pointwise, FP64 k=0 modes and reductions */

#define PFC_OK 0
#define PFC_E_ARG (-1)
#define PFC_E_UNSUPPORTED (-2)
#define PFC_E_HIP (-3)
#define PFC_E_NOMEM (-4)
#define PFC_E_CAP (-5)

/* CLI-visible run parameters (misfit.cpp:17-35,125-143; cvd.cpp:22-53).  All
 * physics values are float32, as the reference parses them with strtof
 * (pfc_common.h:197-205).  The model's remaining constants (nu, w, c0, alpha,
 * mobilities, lattice wave numbers) are the reference's file-static values. */
typedef struct pfc_params {
    int model;      /* PFC_MODEL_MISFIT or PFC_MODEL_CVD */
    int L, M, N;    /* grid; the reference fixes it at compile time (pfc_common.h:138-146).
                       L, M in {32,64,128,256,384,512,768,1024}; N in {1,16,32,64,128} */
    int precision;  /* PFC_PREC_FP64 or PFC_PREC_FP32 */
    int narrow;     /* 1: reference float narrowing (Rhat, inc, energy terms) */
    float u0, con0, sig, dt, dx, axx;
    int grainx, grainy, grainz;
    unsigned int seed;
} pfc_params;

typedef struct pfc_ctx pfc_ctx;

/* Defaults of the reference model (misfit.cpp:116-160 / cvd.cpp:22-53). */
int pfc_params_default(pfc_params* p, int model);

/* Reference initial fields on the host, splitmix64 RNG included:
 * INITIAL_dis (misfit.cpp:610-682) or INITIAL_C2 (cvd.cpp:587-642).
 * phi, con: L*M*N doubles, index n+N*(m+M*l). */
int pfc_init_fields(const pfc_params* p, double* phi, double* con);

/* Allocate a batch of `batch` independent instances sharing grid, model and
 * precision (replaces the reference's 26 fftw_malloc arrays and 14 plans,
 * misfit.cpp:399-466).  batch = 1 is the single-run drop-in. */
int pfc_create(const pfc_params* p, int batch, pfc_ctx** out);
int pfc_destroy(pfc_ctx* ctx);

/* Load real fields of instance `inst` (host doubles, L*M*N each). */
int pfc_set_fields(pfc_ctx* ctx, int inst, const double* phi, const double* con);

/* Step mask: an inactive instance keeps its state while pfc_step and
 * pfc_step_energy advance the others (all instances start active).  Serves
 * the continuous-batching request loop, where slots of one batch hold jobs
 * admitted at different times; the reference runs one job per process
 * (MainWindow.cpp queue, misfit.cpp:311-590), so there is no counterpart. */
int pfc_set_active(pfc_ctx* ctx, int inst, int active);

/* Seed of instance `inst` for the Misfit noise stream (Gauss,
 * misfit.cpp:951-956); defaults to the creation parameters' seed.  Loading
 * fields with pfc_set_fields restarts the instance's step index at 0. */
int pfc_set_seed(pfc_ctx* ctx, int inst, unsigned int seed);

/* Advance every instance by nsteps reference time steps (the `for t` body,
 * misfit.cpp:477-498 / cvd.cpp:400-419).  Enqueues and returns; see pfc_sync. */
int pfc_step(pfc_ctx* ctx, int nsteps);
int pfc_sync(pfc_ctx* ctx);

/* Real fields of instance `inst` (UPDATECOEF output, misfit.cpp:686-697). */
int pfc_get_fields(pfc_ctx* ctx, int inst, double* phi, double* con);
int pfc_get_fields_f32(pfc_ctx* ctx, int inst, float* phi, float* con);

/* ENERGYC2 of the current state (misfit.cpp:1043-1092): mean energy and mean
 * density, FP64 reductions.  e_out: 2 doubles per instance, all instances. */
int pfc_energy(pfc_ctx* ctx, double* e_out);

/* nsteps steps with ENERGYC2 after every step (run_cvd, cvd.cpp:442), fused
 * into the step kernels.  trace: nsteps*batch*2 doubles, step-major. */
int pfc_step_energy(pfc_ctx* ctx, int nsteps, double* trace);
/* The same in two halves (no counterpart in the reference, which steps one
 * run synchronously): _enqueue queues the steps and the trace's copy to a
 * host buffer of the context and returns; _result waits for the context and
 * copies the trace of the last enqueued call (same nsteps) to `trace`. */
int pfc_step_energy_enqueue(pfc_ctx* ctx, int nsteps);
int pfc_step_energy_result(pfc_ctx* ctx, int nsteps, double* trace);

/* *done = 1 when all work queued on the context has finished, else 0; does
 * not wait (no counterpart in the reference: lets a caller step several
 * contexts and handle whichever finishes first). */
int pfc_query(pfc_ctx* ctx, int* done);

/* Phimax of the current state of instance `inst`.
 * Misfit phimax_atmposi (misfit.cpp:3261-3464): x,y,z,phi doubles, con float;
 * entries in the reference order, fallback point included.  N must be 1 or
 * the result is the 2-D extraction repeated per n, as the reference. */
int pfc_phimax_misfit(pfc_ctx* ctx, int inst, double* x, double* y, double* z, double* phi,
                      float* con, int cap, int* n);
/* CVD atompositionCFGC2 maxima (cvd.cpp:1076-1169): fractional coordinates
 * pc_x=(m+1)dx/Ly, pc_y=(n+1)dx/Lz, pc_z=l dx/Lx and phi as floats, plus the
 * threshold phic and the float max. */
int pfc_phimax_cvd(pfc_ctx* ctx, int inst, float* pc_x, float* pc_y, float* pc_z, float* pc_phi,
                   int cap, int* n, float* phic, float* phimax);

/* Elastic post-processing on float32 fields read from VTI (no context needed).
 * phimax_atmposi of elastic.cpp:2802-2921: x,y,z,phi doubles. */
int pfc_phimax_elastic(int L, int M, int N, float dx, const float* phi, double* x, double* y, double* z,
                       double* phimax, int cap, int* n);
/* No counterpart in the reference: pfc_phimax_elastic for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops).  phi (L*M*N floats), x, y, z, phimax (cap doubles each) and n (one
 * int) are DEVICE pointers; the work is enqueued on `stream` (a hipStream_t;
 * NULL = the default stream) and the call returns without waiting.  *n
 * receives the number of atoms found, which may exceed cap (then only the
 * first cap are written).  The library's scratch for it is a fixed device
 * array, one per GPU for the life of the process, sized for fields of at most
 * PFC_NOALLOC_MAX_POINTS points (larger fields: PFC_E_UNSUPPORTED); calls that
 * overlap on one GPU share it and race.  Same results as pfc_phimax_elastic. */
#define PFC_NOALLOC_MAX_POINTS (1L << 24)
int pfc_phimax_elastic_noalloc(int L, int M, int N, float dx, const float* phi, double* x, double* y, double* z,
                               double* phimax, int cap, int* n, void* stream);

/* elastic3D (elastic.cpp:3586-3729): |strain| and neighbour count per atom. */
int pfc_strain(int L, int M, int N, float dx, float benchel, int n, const double* x, const double* y,
               const double* z, double* strain, unsigned int* nb);

/* cryst_order (elastic.cpp:3013-3389): per-atom bond-order parameters of the
 * atoms from pfc_phimax_elastic, neighbours within 1.2 (minimum image), at
 * most 16 stored.  q4 and q6 are the reference's two columns (both from
 * l = 6; q6 accumulated onto q4's sum); nb (may be NULL) the neighbour count. */
int pfc_cryst_order(int L, int M, int N, float dx, int n, const double* x, const double* y, const double* z,
                    double* q4, double* q6, unsigned int* nb);

/* GR (cvd.cpp:1277-1312): pair histogram of the n CVD atom fractions that
 * pfc_phimax_cvd returns (2 per pair, bins of a/25, a = 1, pairs closer than
 * a/sqrt(2)/2 left out).  *maxbin = int(L*dx/delr); hist (may be NULL, then
 * only *maxbin is set) holds *maxbin counts.  The reference calls it on every
 * CVD checkpoint (cvd.cpp:1253). */
int pfc_cvd_gr(int L, int M, int N, float dx, int n, const float* pc_x, const float* pc_y, const float* pc_z,
               unsigned int* hist, int* maxbin);
/* Q6_GLO (cvd.cpp:1347-1510): per-entry neighbour count and local Q6
 * (neighbours at a*sqrt(3)/2*[0.9,1.2], at most 24 used), the global Q6 and
 * the 30-bin histogram (width 0.05) of the local values.  The reference calls
 * it on CVD checkpoints with kd % 20 == 0 (cvd.cpp:1254). */
int pfc_cvd_q6(int L, int M, int N, float dx, int n, const float* pc_x, const float* pc_y, const float* pc_z,
               unsigned int* nb, float* ql, float* qglobal, unsigned long long* histq);

/* Columns per workgroup of the plain-step column pass, as chosen at
 * pfc_create by the library's autotuner (no counterpart in the reference:
 * the tile width has no effect on results; PFC_AUTOTUNE=0 keeps the default). */
int pfc_col_tiles(pfc_ctx* ctx, int* columns);
/* Force tile-width candidate 0 (default), 1 (twice) or 2 (half as many
 * columns) for the plain-step column pass (no counterpart in the reference;
 * for comparisons; results do not depend on it). */
int pfc_set_col_tiles(pfc_ctx* ctx, int candidate);

/* Contexts stepped as one graph (no counterpart in the reference, which runs
 * one job per process; for the grid classes of an uneven sweep): the graph
 * holds `chunk` steps of every context as parallel branches, so one launch
 * advances all of them.  pfc_group_step runs nchunks launches, ordered after
 * work already queued on the member contexts and before work queued on them
 * afterwards.  Kernel timing must be off for the members. */
typedef struct pfc_group pfc_group;
int pfc_group_create(pfc_ctx** ctxs, int n, int chunk, pfc_group** out);
int pfc_group_step(pfc_group* group, int nchunks);
int pfc_group_destroy(pfc_group* group);

/* Capture steps in HIP graphs of `chunk` steps (0 disables). */
int pfc_set_graph(pfc_ctx* ctx, int chunk);

/* Kernel timing with HIP events on the context's stream: kernel 0 = row pass,
 * 1 = column pass, 2 = depth passes (N > 1).  While enabled, every launch of
 * that kernel carries a start/stop event pair that takes the kernel's own
 * start and completion times (hipExtLaunchKernel) and steps launch without
 * graphs; pfc_kernel_time reports the mean duration since enabling. */
int pfc_time_kernel(pfc_ctx* ctx, int kernel, int enable);
int pfc_kernel_time(pfc_ctx* ctx, double* mean_ms, int* count);

/* Algorithmic HBM bytes one launch of kernel `kernel` moves (all instances). */
int pfc_kernel_bytes(pfc_ctx* ctx, int kernel, double* bytes);

/* Measurement aid, no counterpart in the reference: two launches of a plain
 * device copy of `nbytes` (a multiple of 4096; reads and writes nbytes each),
 * the second timed.  pattern 0: linear; 1: 64-byte pieces, column-wise over
 * 4096-byte rows (the column pass's access granularity); 2: as 1 with the
 * column pass's XCD-aware tile order.  Calibrates hardware byte counters
 * against kernels of known traffic. */
int pfc_copy_probe(size_t nbytes, int pattern, double* ms);

/* Free and total device memory of the current GPU (hipMemGetInfo). */
int pfc_device_memory(size_t* free_bytes, size_t* total_bytes);

const char* pfc_last_error(void);

/* ------------------------------------------------------------------------
 * Request loop with continuous batching.  The reference runs one job per
 * pfc-exp-cli process, queued one at a time by the GUI (MainWindow.cpp:2862-
 * 2976,3453-3521, misfit.cpp:311-590); these entry points have no counterpart
 * there.  A job is one run (what one CLI invocation computes); jobs of one
 * class (model, precision, grid, sig/dt/dx/axx, narrowing) share a batch
 * context, are admitted into free slots as they arrive while the others keep
 * stepping, and leave when done.  Results are bit-identical to running the
 * job alone in a one-instance context of the same precision.
 * ------------------------------------------------------------------------ */
#define PFC_JOB_QUEUED 0
#define PFC_JOB_RUNNING 1
#define PFC_JOB_DONE 2
#define PFC_JOB_CANCELLED 3
#define PFC_JOB_FAILED 4

typedef struct pfc_server pfc_server;

typedef struct pfc_server_opts {
    double hbm_budget; /* bytes of device memory for batch contexts; <= 0: 80% of free memory */
    int max_slots;     /* instances per batch context (default 16) */
    int graph_chunk;   /* steps per captured HIP graph (default 25; 0: eager launches) */
    int quantum;       /* most steps between two admission rounds (default 500) */
} pfc_server_opts;

typedef struct pfc_job {
    pfc_params p;       /* run parameters; p.seed seeds the initial fields and the noise */
    int steps;          /* time steps (> 0) */
    int mod;            /* checkpoint interval for outdir output (0: final state only) */
    int keep_fields;    /* 1: keep the final fields for pfc_server_result */
    const char* outdir; /* NULL: no files; else the CLI's output files are written there
                           (run_config.txt, energy.txt, phi/con VTI + PVTI, Phimax_<kd>.txt,
                           checkpoint_timestamps.txt) by a writer thread */
} pfc_job;

typedef struct pfc_job_status {
    int state;              /* PFC_JOB_* */
    int steps_done;
    double t_submit, t_start, t_end; /* seconds on the server clock: arrival, admission into a slot,
                                        final files written */
    double energy, density;          /* ENERGYC2 of the final state (DONE) */
    double t_stepped;                /* the last step done (before the final files), or -1 */
    double create_s;                 /* seconds of context creation the job's admission waited for */
} pfc_job_status;

typedef struct pfc_server_info {
    double now;             /* server clock, seconds */
    long long jobs_done;
    int jobs_unfinished;    /* queued, running or awaiting file writes */
    double point_steps;     /* L*M*N*steps summed over finished jobs */
    int contexts;
    double bytes_in_use, budget;
} pfc_server_info;

int pfc_server_create(const pfc_server_opts* opts, pfc_server** out); /* opts may be NULL */
int pfc_server_submit(pfc_server* s, const pfc_job* job, long long* id);
int pfc_server_poll(pfc_server* s, long long id, pfc_job_status* st);
/* Final fields (L*M*N doubles each) of a DONE job submitted with keep_fields. */
int pfc_server_result(pfc_server* s, long long id, double* phi, double* con);
/* Failure text of a FAILED job. */
int pfc_server_error(pfc_server* s, long long id, char* buf, int cap);
/* Asks to cancel a job: a queued job ends CANCELLED when the scheduler next
 * admits jobs, a running one at its next event (checkpoint, quantum or end);
 * a job that completes first stays DONE. */
int pfc_server_cancel(pfc_server* s, long long id);
/* Wait until no job is queued, running or awaiting writes (timeout_s < 0: forever). */
int pfc_server_drain(pfc_server* s, double timeout_s);
int pfc_server_stats(pfc_server* s, pfc_server_info* st);
/* Cancels what is unfinished, stops the threads, frees the contexts. */
int pfc_server_destroy(pfc_server* s);

#ifdef __cplusplus
}
#endif
#endif
