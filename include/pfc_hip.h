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
    int L, M, N;    /* grid; the reference fixes it at compile time (pfc_common.h:138-146) */
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
/* elastic3D (elastic.cpp:3586-3729): |strain| and neighbour count per atom. */
int pfc_strain(int L, int M, int N, float dx, float benchel, int n, const double* x, const double* y,
               const double* z, double* strain, unsigned int* nb);

/* Capture steps in HIP graphs of `chunk` steps (0 disables). */
int pfc_set_graph(pfc_ctx* ctx, int chunk);

/* Kernel timing with HIP events on the context's stream: kernel 0 = row pass,
 * 1 = column pass.  While enabled, every launch of that kernel is bracketed by
 * an event pair; pfc_kernel_time reports the mean duration since enabling. */
int pfc_time_kernel(pfc_ctx* ctx, int kernel, int enable);
int pfc_kernel_time(pfc_ctx* ctx, double* mean_ms, int* count);

/* Algorithmic HBM bytes one launch of kernel `kernel` moves (all instances). */
int pfc_kernel_bytes(pfc_ctx* ctx, int kernel, double* bytes);

const char* pfc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
