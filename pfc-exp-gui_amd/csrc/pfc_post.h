// pfc_post.h - launchers for the Phimax / strain post-processing kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace pfc {

struct PostResult {
    int n = 0;          // entries found (may exceed the caller's capacity)
    float phic = 0.f;   // CVD threshold
    float phimax = 0.f; // CVD float max
};

// Misfit phimax_atmposi on device real fields phi/con ([L][M], prec reals).
hipError_t post_phimax_misfit(int prec, int L, int M, float dx, float con0, const void* phi, const void* con,
                              double* x, double* y, double* z, double* ph, float* co, int cap, PostResult* r,
                              hipStream_t s);
// CVD atompositionCFGC2 maxima on device real field phi.
hipError_t post_phimax_cvd(int prec, int L, int M, int N, float dx, float u0, const void* phi, float* px, float* py,
                           float* pz, float* pphi, int cap, PostResult* r, hipStream_t s);
// Elastic phimax_atmposi on a host float field.
hipError_t post_phimax_elastic(int L, int M, int N, float dx, const float* phi, double* x, double* y, double* z,
                               double* ph, int cap, PostResult* r);
// Elastic phimax_atmposi on a device float field into device outputs, on
// stream s, without allocating: scratch is a file-scope device array sized
// for at most NOALLOC_MAX_POINTS points; *n_out (device) receives the count.
constexpr long NOALLOC_MAX_POINTS = 1L << 24;
hipError_t post_phimax_elastic_noalloc(int L, int M, int N, float dx, const float* phi, double* x, double* y,
                                       double* z, double* ph, int cap, int* n_out, hipStream_t s);
// elastic3D on host atom positions.
hipError_t post_strain(int L, int M, int N, float dx, float benchel, int n, const double* x, const double* y,
                       const double* z, double* strain, unsigned* nb);

// cryst_order on host atom positions: per-atom q4, q6, neighbour counts.
hipError_t post_cryst(int L, int M, int N, float dx, int n, const double* x, const double* y, const double* z,
                      double* q4, double* q6, unsigned* nb);

// CVD GR: bins (int(L*dx/delr)) and the pair histogram of n atom fractions.
int cvd_gr_bins(int L, float dx);
hipError_t post_cvd_gr(int L, int M, int N, float dx, int n, const float* x, const float* y, const float* z,
                       unsigned* hist);
// CVD Q6_GLO: neighbour counts, local Q6, global Q6, 30-bin histogram.
hipError_t post_cvd_q6(int L, int M, int N, float dx, int n, const float* x, const float* y, const float* z,
                       unsigned* nb, float* ql, float* qglobal, unsigned long long* histq);

// Plain device copy of n 16-byte words (counter calibration); pattern 0:
// linear, 1: 64-byte pieces of a [n/256][256] array copied column-wise,
// 2: the same with the column pass's XCD-aware tile order.
hipError_t launch_copy16(const void* src, void* dst, size_t n, int pattern, hipStream_t s);

}  // namespace pfc
