// pfc_kernels.h - host-side launchers for the spectral step kernels.
//
// Device data layout (one batch of B instances of one grid L x M x N):
//   S   [B][L][N][K]  row-spectral state: r2c along m of (dpsi, dcon), K = M/2
//                     (the real Nyquist value X[M/2] is stored as Im X[0]),
//                     one element = {psi_hat_m, con_hat_m} (two complex values);
//                     the k=0 mode of both fields is held apart, as FP64 means.
//   AB  [B][L][N][K]  row-spectral C2A*psi_hat and C2B*psi_hat of the same state
//                     (the reference's c2aden/c2bden, misfit.cpp:801-833).
//   IN  [B][L][N][K]  row-spectral nonlinear terms {inn, inc} (misfit.cpp:835-874).
// Real-space planes (inputs, outputs, noise) use the reference order
// n + N*(m + M*l).  A 2-D step (N = 1) is row pass (S, AB -> IN) + column
// pass (S, IN -> S, AB).  A 3-D step adds a depth pass (FFT along n) on each
// side of the column pass: depth+ (S, IN in place), column, depth- (S, AB in
// place); between the depth passes the arrays hold (m, n)-spectral values
// with the packed k_m = 0 / M/2 column split along n (slot j < N/2: k_m = 0,
// k_n = j; slot N - j: k_m = M/2, k_n = j; slots 0 and N/2 stay packed).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>

namespace pfc {

enum { PREC_FP64 = 0, PREC_FP32 = 1 };

// Scalars the kernels need, with the reference's float products pre-formed.
struct PhysArgs {
    float niu, ksi, w, c0, alpha;
    float mn_dt;    // dynamic_mn*dt            (float, misfit.cpp:1002)
    float mc_dt;    // dynamic_mc*dt            (float, misfit.cpp:1007)
    float mc_dt_a;  // dynamic_mc*dt*para_alpha (float, misfit.cpp:1009)
    int narrow;     // reference float narrowing of Rhat/inc/energy terms
    float db;       // dt*b0 (float, misfit.cpp:1002); 0 = no noise term
    float lnc0;     // ln(c0), ln(1 - c0): the FP32 nonlinear term's log offsets
    float ln1c0;
};

struct DevBufs {
    void* S;
    void* AB;
    void* IN;
    const double* means;  // [B][4]: mean psi, mean con, C2A(0)*mean psi, C2B(0)*mean psi
    const void* twH;      // exp(-2 pi i t/H), H = M/2
    const void* twM;      // exp(-2 pi i t/M)
    const void* twL;      // exp(-2 pi i t/L)
    const void* twN;      // exp(-2 pi i t/N) (N > 1)
    const float* kl;      // [L]  reference krefl
    const float* km;      // [M]  reference krefm
    const float* kn;      // [N]  reference krefn
    const float* c2;      // [(L/2+1)][(M/2+1)][(N/2+1)][2] folded {C2A, C2B}
    double* erow;         // [B][row_blocks]  sum of pointwise energy terms
    double* ecol;         // [B][col_blocks]     sum w(k)|con_hat|^2
    const void* NZ;       // [B][L][K] complex: row-spectral noise (nullptr: no noise)
    long long* tcount;    // [B] step counter per instance (advanced by the column pass)
    const int* active;    // [B] 0 = instance idle (step kernels skip it); nullptr = all
    // Per-mode coefficients in the column layout [L][N][K] (launch_coef;
    // nullptr: the column pass evaluates them from kl/km/kn/c2), cC = {C2A, C2B}.
    const void* cA;
    const void* cC;
    int col_ts;  // plain-step column tile width chosen by the autotuner (0 default, 1 wider, 2 narrower)
    int L, M, N, K, B;
    // Energy-traced stepping: with etrace set, one workgroup per instance of
    // the (energy) column pass first reduces the previous state's partial
    // sums (erow of this step's row pass, eprev of the previous column pass,
    // which had the same grid) into etrace[inst*2 + {0,1}].
    const double* eprev = nullptr;
    double* etrace = nullptr;
    int nrow = 0;  // row-pass workgroups per instance (erow entries)
};

enum { ROW_WRITE_IN = 1, ROW_ENERGY = 2 };
enum { COL_INIT = 0, COL_STEP = 1, COL_ENERGY_ONLY = 2 };
enum { COL_ENERGY = 1 };

// Field-split passes (pfc_fs.hip): one group of threads per field, at every
// batch size.  They run the FP64 row pass on power-of-two rows and the FP32
// row pass on 1024-point rows (one row per workgroup), and the FP64 plain and
// energy-traced column steps on power-of-two columns (one column per
// workgroup).  FP32 columns keep the four-field passes: a field-split FP32
// column moves 8-byte pieces per lane (one cache line per lane and row), and
// measured slower (one 1024^2 instance 20.4 -> 27.5 us per column pass).
#ifndef PFC_FS_ROW32
#define PFC_FS_ROW32 512  // FP32 row half-lengths H on the field-split row pass (bit mask of powers of two)
#endif
constexpr bool pow2_len(int n) { return n > 0 && (n & (n - 1)) == 0; }
constexpr bool fs_row(int prec, int H) {  // H = M/2
    return prec == PREC_FP64 ? pow2_len(H) && H >= 16 && H <= 512 : pow2_len(H) && H <= 512 && (PFC_FS_ROW32 & H) != 0;
}
constexpr bool fs_col(int prec, int L) { return prec == PREC_FP64 && pow2_len(L) && L >= 32 && L <= 1024; }

// Whether the plain-step column kernels of this precision and length read
// per-mode coefficient tables: FP32, every length; FP64 where the
// field-split column pass runs (it loads them at its start, so their
// latency hides behind the state's).  (The four-field FP64 column kernels
// measured no gain from them: their columns are not held up by the C2
// gather.)
constexpr bool col_coef(bool f64, int L) { return !f64 ? L > 0 : fs_col(PREC_FP64, L); }

// Fill cA/cC from kl/km/kn/c2 and the step constants: FP32 {mk, mkc,
// 1/(1 + mk), 1/(1 + den)}, FP64 {mk, mkc, den, 0} (floats, as the reference
// forms them), and {C2A, C2B}.
hipError_t launch_coef(int prec, const DevBufs& b, const PhysArgs& p, hipStream_t s);

bool size_supported(int n);
bool depth_supported(int n);                    // N > 1 lengths of the depth pass
int row_blocks(int prec, int M, int rows);      // row-pass workgroups per instance (rows = L*N)
int col_blocks(int prec, int L, int M, int N);  // column-pass workgroups per instance (default tiles)
int col_tiles(int prec, int L, int ts);         // columns per workgroup of tile-width candidate ts

// t0/t1 (optional): events that take the kernel's own start and end times.
hipError_t launch_row_step(int prec, const DevBufs& b, const PhysArgs& p, int flags, hipStream_t s,
                           hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr);
hipError_t launch_col(int prec, const DevBufs& b, const PhysArgs& p, int mode, int flags, hipStream_t s,
                      hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr);
hipError_t launch_row_in(int prec, const DevBufs& b, const void* inphi, const void* incon, hipStream_t s);
hipError_t launch_row_out(int prec, const DevBufs& b, void* outphi, void* outcon, hipStream_t s);
// Depth pass (N > 1): FFT along n of one or two field pairs, dir -1 forward
// (then the k_m = 0 / M/2 column is split along n), +1 inverse (joined first,
// result scaled by 1/N).  in1/out1 may be null (one pair).  With one_field
// set, in0/out0 hold single complex values (the noise spectrum) instead.
hipError_t launch_depth(int prec, const DevBufs& b, int dir, const void* in0, void* out0, const void* in1,
                        void* out1, bool one_field, hipStream_t s, hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr);
// One real plane per instance [B][L][M][N] -> row spectrum [B][L][N][K] (single field).
hipError_t launch_row_r2c1(int prec, const DevBufs& b, const void* in, void* out, hipStream_t s);
// e_out[inst*2 + {0,1}] = {mean energy, mean density}; alpha_over_P = alpha/P.
hipError_t launch_energy_reduce(const DevBufs& b, int nrow, int ncol, double alpha, double* e_out,
                                hipStream_t s);

// Misfit noise (misfit.cpp:951-981): real noise planes [B][L][M][N] for the
// current step of each instance, then NZ = r2c along m of those planes.
// seeds: [B] run seeds; scratch sized by noise_scratch_bytes(P, B).
size_t noise_scratch_bytes(long P, int B);
hipError_t launch_noise(int prec, const DevBufs& b, const unsigned* seeds, void* planes, void* nz, void* scratch,
                        int* err_flag, hipStream_t s);

// The field-split passes (fs_row, fs_col above).
hipError_t launch_row_fs(int prec, const DevBufs& b, const PhysArgs& p, int flags, hipStream_t s,
                         hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr);
hipError_t launch_col_fs(int prec, const DevBufs& b, const PhysArgs& p, bool want_e, hipStream_t s,
                         hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr);

// Named kernels (for timing the dominant one).
enum { KID_ROW = 0, KID_COL = 1, KID_DEPTH = 2 };

}  // namespace pfc
