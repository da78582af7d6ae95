/*
 * pfc_oracle.h - CPU FP64 restatement of the pfc-exp-cli hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product (libpfc_hip.so, pfc-exp-cli) never
 * links, loads or calls anything here.
 *
 * It restates, function by function, the reference sources (paths relative to
 * the reference repository root; misfit.cpp means
 * src/experiment/misfit_uniform_time_seed_fftw_early.cpp):
 *   pfc_common.h:165-181 (splitmix64), :237-243 (ks, min3)
 *   misfit.cpp:116-179 (parameters), :455-462 (k tables), :610-682 (init),
 *              :686-697, :700-750, :773-874, :877-925, :951-1014, :1043-1092,
 *              :1588-1810 (VTI), :3261-3556 (Phimax)
 *   cvd.cpp:26-67, :205-213, :389-454, :587-642, :865-886, :979-1040,
 *           :1067-1274 (atompositionCFGC2)
 *   elastic.cpp:30-74, :2562-2731, :2802-3011, :3586-3776
 *
 * Precision: fields and FFTs are double; every site where the reference narrows
 * to float (Rhat1..4, the CH nonlinear temporaries, per-point energy terms, all
 * model parameters, k tables and C2 kernels) is honoured when `narrow` is set.
 */
#ifndef PFC_ORACLE_H
#define PFC_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { PFO_MISFIT = 0, PFO_CVD = 1 };

typedef struct pfo_model {
    int model;               /* PFO_MISFIT or PFO_CVD */
    int L, M, N;             /* grid (reference: compile-time PFC_GRID_L/M/N) */
    float u0, sig, con0, dt, dx, axx, b0;
    int grainx, grainy, grainz;
    int dnx, dny, dnz;
    unsigned int seed;
    int steps, mod;
    float a_lattice, b_lattice, gaa1, gaa2, gab1, gab2, niu, ksi;
    float ka0, ka1, ka2, sigMa1, sigMa2, kb0, kb1, kb2, sigMb1, sigMb2;
    float c0, mn, mc, alpha, w;
    int narrow;              /* 1: reference float narrowing (default) */
} pfo_model;

/* Defaults exactly as the reference's file-static initialisers. */
void pfo_defaults(pfo_model* p, int model, int L, int M, int N);
/* RecomputeDerivedParams (misfit.cpp:162-179, cvd.cpp:205-213). */
void pfo_derive(pfo_model* p);

/* splitmix64 / u01 (pfc_common.h:165-181). */
uint64_t pfo_splitmix64(uint64_t* state);
double pfo_rng_u01(uint64_t* state);
/* The --axx noise stream of one step (misfit.cpp:951-981): n Marsaglia polar
 * values from the sequential splitmix64 stream of (seed, timestep); returns
 * the uniform pairs drawn.  Pinned by the moments of N(0,1) and the pi/4
 * acceptance rate; the (seed, timestep) mixing is parity unpinned (the
 * reference holds no noise vectors). */
uint64_t pfo_gauss(unsigned int seed, int timestep, size_t n, double* out);
/* Grain angle theta (degrees*pi/180 as float) for grain (i,j,k) (misfit.cpp:636-644). */
float pfo_grain_theta(const pfo_model* p, int iii, int jjj, int kkk);

/* k tables (float) and |k|^2. */
void pfo_ktables(const pfo_model* p, float* kl, float* km, float* kn);
/* C2A/C2B kernels on the full grid (float, index n+N*(m+M*l)); kc out. */
void pfo_c2(const pfo_model* p, float* c2a, float* c2b, float* kc1, float* kc2);

/* Initial fields (real parts; the reference sets Im(con)=1e-10, which never
 * reaches a real part). phi, con: L*M*N doubles. */
void pfo_init_fields(const pfo_model* p, double* phi, double* con);

/* Complex FFT, 3-D (L,M,N) c2c, in place on interleaved doubles, unnormalised.
 * sign = -1 (FFTW_FORWARD) or +1 (FFTW_BACKWARD). */
void pfo_fft3d(double* z, int L, int M, int N, int sign);
void pfo_fft1d(double* z, int n, int sign);

/* Pointwise terms (misfit.cpp:773-874), exposed for closed-form tests. */
double pfo_in1(const pfo_model* p, double phi);
double pfo_in2(const pfo_model* p, double c);
double pfo_in3(double c, double a, double b);
double pfo_inc(const pfo_model* p, double phi, double c, double a, double b);
/* Semi-implicit update of one mode (misfit.cpp:983-1014 with b0=0 /
 * cvd.cpp:865-886).  z = {re,im}; returns updated value in z. */
void pfo_update_phi(const pfo_model* p, float ks, double* zphi, const double* outn);
void pfo_update_con(const pfo_model* p, float ks, double* zcon, const double* outc);

/* Simulation object: full spectral-step restatement. */
typedef struct pfo_sim pfo_sim;
pfo_sim* pfo_create(const pfo_model* p);
void pfo_destroy(pfo_sim* s);
void pfo_set_fields(pfo_sim* s, const double* phi, const double* con);
void pfo_get_fields(const pfo_sim* s, double* phi, double* con);
/* One reference time step t (t = step index, used by the noise RNG). */
void pfo_step(pfo_sim* s, int t);
/* ENERGYC2: mean energy and mean density of the current state. */
void pfo_energy(pfo_sim* s, double* e_mean, double* rho_mean);
/* Run n steps; faithful=1 also performs the reference's diagnostic FFTs
 * (second UPDATEINFCMIX, UPDATEINC2GRA) every step. */
void pfo_run(pfo_sim* s, int t0, int n, int faithful);

/* Phimax / strain.  Atom arrays are caller-owned, capacity `cap`. */
typedef struct pfo_atoms {
    int n, cap;
    double *x, *y, *z, *phi;
    float *con;
} pfo_atoms;

/* Misfit phimax_atmposi (misfit.cpp:3261-3556); returns number of entries
 * (after fallback).  phi/con are real fields (doubles). */
int pfo_phimax_misfit(int L, int M, int N, float dx, float con0,
                      const double* phi, const double* con, pfo_atoms* out);
/* CVD atompositionCFGC2 (cvd.cpp:1067-1249): pc_x/pc_y/pc_z/pc_phi floats. */
int pfo_phimax_cvd(int L, int M, int N, float dx, float u0, const double* phi,
                   float* pc_x, float* pc_y, float* pc_z, float* pc_phi, int cap,
                   float* phic_out, float* phimax_out);
/* Elastic phimax_atmposi on float32 fields (elastic.cpp:2802-2921). */
int pfo_phimax_elastic(int L, int M, int N, float dx, const float* phi, pfo_atoms* out);
/* elastic3D (elastic.cpp:3586-3729): strain magnitude and neighbour count. */
void pfo_strain(int L, int M, int N, float dx, float benchel, int n,
                const double* x, const double* y, const double* z,
                double* strain, unsigned* nb);

/* VTI / PVTI (misfit.cpp:1588-1810); reader (elastic.cpp:2562-2645). */
int pfo_write_vti(const char* path, const char* name, int L, int M, int N, const double* field);
int pfo_write_pvti(const char* path, const char* name, const char* prefix, int kd, int L, int M, int N);
int pfo_read_vti(const char* path, int L, int M, int N, float* field);

/* Text writers (fixed, 4 decimals). */
int pfo_write_phimax_misfit(const char* path, const pfo_atoms* a);
int pfo_write_phimax_elastic(const char* path, const pfo_atoms* a);
int pfo_write_strain(const char* path, const pfo_atoms* a, const double* strain, const unsigned* nb);
/* cryst_order (elastic.cpp:3013-3389): per-atom q4, q6 and neighbour counts. */
void pfo_cryst_order(int L, int M, int N, float dx, int n, const double* x, const double* y, const double* z,
                     double* q4, double* q6, unsigned* nb);
float pfo_ylm_r(int l, int m, float theta, float fai);
float pfo_ylm_c(int l, int m, float theta, float fai);
int pfo_write_q4q6(const char* path, const pfo_atoms* a, const double* q4, const double* q6);

/* CVD GR (cvd.cpp:1277-1344) over the gathered pc_* fractions: number of
 * bins, and the pair histogram (2 per pair, minimum image). */
int pfo_cvd_gr_bins(int L, float dx);
void pfo_cvd_gr(int L, int M, int N, float dx, int n, const float* x, const float* y, const float* z,
                unsigned* hist);
/* hist_<kd>.txt and Grq6_<kd>.txt (cvd.cpp:1314-1339). */
int pfo_write_cvd_gr(int kd, int L, int M, int N, float dx, int n, const float* x, const unsigned* hist);
/* CVD Q6_GLO (cvd.cpp:1347-1527): local Q6 per entry, the global Q6 and the
 * 30-bin histogram of the local values. */
void pfo_cvd_q6(int L, int M, int N, float dx, int n, const float* x, const float* y, const float* z,
                unsigned* nb, float* ql, float* qglobal, unsigned long long* histq);
/* Ql_<kd>.txt and HistQ_<kd>.txt (cvd.cpp:1449-1510). */
int pfo_write_cvd_q6(int kd, int n, const unsigned* nb, const float* ql, float qglobal,
                     const unsigned long long* histq);

#ifdef __cplusplus
}
#endif
#endif
