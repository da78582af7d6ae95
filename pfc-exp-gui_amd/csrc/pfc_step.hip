// pfc_step.hip - fused spectral step kernels for the Misfit / CVD binary PFC
// models on gfx950.
//
// The reference step (misfit.cpp:477-498, cvd.cpp:400-419) performs eight
// state-changing c2c FFTs over full complex-double arrays plus five
// elementwise sweeps.  Here a step is two kernels over half spectra:
//
//   row pass    per row l: c2r along m of S and AB -> psi, con, a, b in
//               registers -> in1+in2+in3 and inc (UPDATEINN3, UPDATEINFMIX,
//               UPDATEINFCMIX second half, UPDATEINC2MIX, UPDATEINC2MIXC)
//               -> r2c along m -> IN.
//   column pass per column tile: FFT along l of S and IN -> semi-implicit
//               update (MAINEQUATIONC2_NOI with b0 = 0 / MAINEQUATIONC2MIX)
//               -> inverse FFT along l of the new state -> S, and of
//               C2A/C2B times the new psi_hat -> AB (UPDATEINFCMIX first half).
//
// Both passes stream each array through HBM once; all transforms and the
// pointwise algebra happen in LDS and VGPRs.
#include <type_traits>

#include "pfc_fft.hpp"
#include "pfc_kernels.h"
#include "pfc_step_dev.hpp"

namespace pfc {

// ---------------------------------------------------------------------------
// Launch shapes
// ---------------------------------------------------------------------------
// Launch shapes (elements per thread E, rows per workgroup RPW, columns per
// workgroup T) per precision and length, and the code-shape switches below.
// The defaults are the fastest of the variants timed on MI355X for the sweep
// grids (16 instances of 128^2..512^2, FP32) and the single-instance FP64
// path; each PFC_* macro overrides one of them for timing experiments.
//
// Row pass
#ifndef PFC_ROW_THREADS
#define PFC_ROW_THREADS 128  // threads per row-pass workgroup (RPW = this / G)
#endif
#ifndef PFC_ROW_E
#define PFC_ROW_E 4  // elements per thread of power-of-two rows
#endif
#ifndef PFC_ROW_WAVE
#define PFC_ROW_WAVE 1  // row exchanges with wavefront fences when a row is one wave
#endif
// Column pass
#ifndef PFC_COL_T_DIV
#define PFC_COL_T_DIV 1  // divide / multiply the default columns per workgroup
#endif
#ifndef PFC_COL_T_MUL
#define PFC_COL_T_MUL 1
#endif
#ifndef PFC_COL_T_F64_SMALL
#define PFC_COL_T_F64_SMALL 1  // FP64 columns per workgroup, L <= 256
#endif
#ifndef PFC_COL_T_F64_MID
#define PFC_COL_T_F64_MID 1  // FP64 columns per workgroup, L = 384, 512 (fills the GPU from one instance)
#endif
#ifndef PFC_COL_SPLIT
#define PFC_COL_SPLIT -1  // exchanges in 16-byte chunks: -1 where faster (FP32 3*2^k, fixed plan), 0/1
#endif
#ifndef PFC_COL_WAVE
#define PFC_COL_WAVE -1  // one column per wavefront: -1 where faster (FP32 L <= 128), 0/1
#endif
#ifndef PFC_XCD_REMAP
#define PFC_XCD_REMAP 1  // workgroup x takes tile (x % 8)*(nx/8) + x/8: one XCD's tiles are adjacent
#endif
#ifndef PFC_ROW_TWO
#define PFC_ROW_TWO 1  // row pass: the c2r of psi, con runs while a, b still load (twiddles in LDS)
#endif
#ifndef PFC_TWO_F64
#define PFC_TWO_F64 0  // 1: the two-field row and column forms on FP64 too (they spill there: one FP64
                       // 512^2 step 28.4 -> 30.5 us)
#endif
#ifndef PFC_COL_TWO
#define PFC_COL_TWO 1  // FP32 plain column step: transforms of two fields at a time (loads/stores overlap them)
#endif
#ifndef PFC_COL_TWO_LDS16
#define PFC_COL_TWO_LDS16 1  // the two-field step's exchange buffers hold 16 bytes per slot (else 32)
#endif
#ifndef PFC_COL_COEF_LAZY
#define PFC_COL_COEF_LAZY 1  // plain steps read each mode's coefficients where the update uses them (no prefetched array)
#endif
#ifndef PFC_COL_COEF_PF
#define PFC_COL_COEF_PF 2  // with LAZY, the first modes' coefficients loaded before the last forward transform
#endif
#ifndef PFC_COL_COEF_LATE
#define PFC_COL_COEF_LATE 1  // (prefetched form) the two-field step loads the mode coefficients after the forward transforms
#endif
#ifndef PFC_COEF
#define PFC_COEF 1  // plain steps read per-mode coefficient tables (k_coef) where col_coef holds
#endif
// Both passes
#ifndef PFC_GEN
#define PFC_GEN 1  // 3*2^k lengths on the general radix-6/8 plan (<= 8 elements per thread)
#endif
#ifndef PFC_GEN_WAVE
#define PFC_GEN_WAVE 0  // with the general plan, one column per wavefront
#endif
#ifndef PFC_LDS_MOD
#define PFC_LDS_MOD 2  // 4-field exchange buffers as 16-byte chunk arrays, staggered by this many bank groups
#endif
#ifndef PFC_WPE
#define PFC_WPE 1  // waves per SIMD the register allocator targets (amdgpu_waves_per_eu)
#endif
#ifndef PFC_WPE12
#define PFC_WPE12 1  // the same for the fixed 12-element plan
#endif
#ifndef PFC_COL_WPE_LEAN
#define PFC_COL_WPE_LEAN 1  // the same for the plain-step column kernel
#endif
constexpr int col_e(bool f64, int LL) {
    return LL % 3 == 0 ? 12 : (f64 ? (LL >= 1024 ? 8 : 4) : (LL == 512 ? 8 : 4));
}
// FP64 columns up to 512 default to one per workgroup, so that a single
// instance (the drop-in path) launches enough workgroups to fill the GPU;
// the 32-byte row pieces' lines stay in one L2 through the XCD tile order.
// Where the field-split column pass runs the plain steps (fs_col: FP64
// powers of two) the other modes also take one column per workgroup, so both
// leave the same number of energy partial sums.
constexpr int col_t(bool f64, int LL) {
    return f64 ? (LL <= 256 ? PFC_COL_T_F64_SMALL : (LL <= 512 ? PFC_COL_T_F64_MID : 1))
               : (LL <= 128 ? 8 : (LL <= 768 ? 4 : 2));
}
static_assert(col_t(true, 1024) == 1 && col_t(true, 512) == 1 && col_t(true, 256) == 1,
              "field-split lengths: one column per workgroup");

// 3*2^k plan (pfc_fft.hpp): 1 (radix 3, then 4s; half the registers) where
// it measured faster: rows of 384 points (H = 192), -10% FP32 and -34% FP64
// per row pass of 384^2; FP64 rows of 768 (-10%); columns of 768 (-2% FP32,
// -8% FP64 at 768x256); FP64 columns of 384 (-14%).  FP32 columns of 384
// (+3%) and FP32 rows of 768 (+6..12%) keep plan 0.
constexpr int row_plan(bool f64, int H) { return (H == 192 || (f64 && H == 384)) ? 1 : 0; }
constexpr int col_plan(bool f64, int LL) { return ((f64 && LL == 384) || LL == 768) ? 1 : 0; }

// Row pass: RPW rows of length M = 2H per workgroup, G = H/E threads per row.
// LDS: one padded exchange buffer of 4-field elements per row.
template <typename R, int H>
struct RowCfg {
    static constexpr bool GEN = PFC_GEN && H % 3 == 0;
    static constexpr int PLAN = row_plan(sizeof(R) == 8, H);
    using SD = Sides<H, elems_for(H, PFC_ROW_E), GEN, PLAN>;
    static constexpr int E = SD::EM, G = SD::G;  // register elements, threads per row
    // Side A: the row spectrum (loads of S/AB, stores of IN); side B: real space.
    static constexpr int EA = SD::EA, GA = SD::GA, EB = SD::EB, GB = SD::GB;
    static_assert(GA == G, "side A uses every thread of the row");
    static constexpr int cap = sizeof(R) == 8 ? 4 : 8;
    static constexpr int want = PFC_ROW_THREADS / G;
    static constexpr int RPW = want < 1 ? 1 : (want > cap ? cap : want);
    static constexpr int NT = RPW * G;
    static constexpr int K = H;  // row spectrum with X[H] packed into Im X[0]
    static constexpr int PB = padded(H);                       // elements per row buffer
    static constexpr int CH = PFC_LDS_MOD ? (int)(sizeof(Vec<R, 4>) / 16) : 1;
    static constexpr int PS = PFC_LDS_MOD ? soa_stride(H, CH, PFC_LDS_MOD) : PB;
    static constexpr size_t RB = (size_t)CH * PS * (sizeof(Vec<R, 4>) / CH);  // bytes per row buffer
    static constexpr size_t LDS = (size_t)RPW * RB;
    static constexpr int WPE = (!GEN && E == 12) ? PFC_WPE12 : PFC_WPE;  // waves per SIMD the register budget targets
    // A row's G threads are consecutive; when they fit one wavefront the
    // row transforms exchange without workgroup barriers.
    static constexpr bool WAVE = PFC_ROW_WAVE != 0 && G <= 64 && 64 % G == 0;
};

// Column pass: T columns of length LL per workgroup, G = LL/E threads per
// column; consecutive threads take consecutive columns (coalesced T*16 or
// T*32 byte segments per row).  TS selects the tile width among the
// autotuner's candidates: 0 the default, 1 twice, 2 half as many columns
// (a candidate that would not fit the LDS or 1024 threads is the default).
template <typename R, int LL, int TS = 0>
struct ColCfg {
    static constexpr bool F64 = sizeof(R) == 8;
    static constexpr bool GEN = PFC_GEN && LL % 3 == 0;
    static constexpr int PLAN = col_plan(F64, LL);
    using SD = Sides<LL, col_e(F64, LL), GEN, PLAN>;
    static constexpr int E = SD::EM, G = SD::G;  // register elements, threads per column
    // Side A: the loaded/stored row spectra along l; side B: the full spectrum.
    static constexpr int EA = SD::EA, GA = SD::GA, EB = SD::EB, GB = SD::GB;
    static_assert(GA == G, "side A uses every thread of the column");
    static constexpr int T0 = col_t(F64, LL);
    static constexpr int T1 = T0 / PFC_COL_T_DIV > 0 ? T0 / PFC_COL_T_DIV : 1;
    // Wider tiles (PFC_COL_T_MUL) only while the exchange buffers fit in LDS.
    static constexpr int TD = (size_t)T1 * PFC_COL_T_MUL * padded(LL) * sizeof(Vec<R, 4>) <= 160 * 1024
                                  ? T1 * PFC_COL_T_MUL : T1;
    static constexpr int TW = TS == 1 ? 2 * TD : (TS == 2 ? (TD / 2 > 0 ? TD / 2 : 1) : TD);
    static constexpr bool TW_OK = (size_t)TW * padded(LL) * sizeof(Vec<R, 4>) <= 150 * 1024 && TW * G <= 1024;
    static constexpr int T = TW_OK ? TW : TD;
    static constexpr int NT = T * G;
    static constexpr int PB = padded(LL);
    static constexpr int CH = PFC_LDS_MOD ? (int)(sizeof(Vec<R, 4>) / 16) : 1;
    static constexpr int PS = PFC_LDS_MOD ? soa_stride(LL, CH, PFC_LDS_MOD) : PB;
    // SPLIT: the exchanges move the 4-field Vec in 16-byte chunks through a
    // buffer of 16 bytes per slot (fewer LDS bytes per wave, more waves per CU).
    static constexpr bool SPLIT = GEN ? false : (PFC_COL_SPLIT < 0 ? (!F64 && LL % 3 == 0) : PFC_COL_SPLIT != 0);
    static constexpr int PS1 = soa_stride(LL, 1, PFC_LDS_MOD ? PFC_LDS_MOD : 1);
    static constexpr size_t CB = SPLIT ? (size_t)PS1 * 16
                                       : (size_t)CH * PS * (sizeof(Vec<R, 4>) / CH);  // bytes per column buffer
    // WAVE: each column's G threads sit in one wavefront (thread = col*G + t),
    // so the transform's exchanges need no workgroup barrier and the waves of
    // a workgroup run out of step with each other.
    static constexpr bool WAVE = (GEN ? PFC_GEN_WAVE != 0
                                      : (PFC_COL_WAVE < 0 ? (!F64 && (LL % 3 == 0 || LL <= 128)) : PFC_COL_WAVE != 0)) &&
                                 G <= 64 && 64 % G == 0;
    using Lds = std::conditional_t<SPLIT, LdsSplit<R, 4, WAVE>, LdsBuf<R, 4, CH, PS, WAVE>>;
    // T column buffers; the packed column's mirror exchange (4-field Vec and
    // noise value per slot) reuses the space from the start.
    static constexpr size_t MIRROR = (size_t)PB * (sizeof(Vec<R, 4>) + sizeof(Cx<R>));
    static constexpr size_t LDS = (size_t)T * CB > MIRROR ? (size_t)T * CB : MIRROR;
    // The two-field column step's buffers: one two-field Vec per slot (16
    // bytes FP32, two 16-byte chunks FP64), columns staggered by PFC_LDS_MOD
    // bank groups like the chunked buffers, and the packed column's mirror
    // exchange from the start.
    static constexpr int CH2 = (int)(sizeof(Vec<R, 2>) / 16);
    static constexpr int PS2 = soa_stride(LL, CH2, PFC_LDS_MOD ? PFC_LDS_MOD : 1);
    static constexpr size_t CB2 = PFC_COL_TWO_LDS16 ? (size_t)CH2 * PS2 * 16 : CB;
    static constexpr size_t LDS2 = (size_t)T * CB2 > MIRROR ? (size_t)T * CB2 : MIRROR;
    static constexpr int WPE = (!GEN && E == 12) ? PFC_WPE12 : PFC_WPE;
    static constexpr int WPE_LEAN = (!GEN && E == 12) ? PFC_WPE12 : PFC_COL_WPE_LEAN;
};

// The two-field column step (k_col, TWO): FP32 lengths on the power-of-two
// plans whose default tile, with its twiddle table, keeps two workgroups per
// CU.  Decided per length, never per tile width: the two forms round
// differently (the compiler fuses a few multiply-adds differently), and every
// tile width must give the same bits.
template <typename R, int LL>
constexpr bool col_two_len() {
    using C0 = ColCfg<R, LL, 0>;
    return PFC_COL_TWO && (sizeof(R) == 4 || PFC_TWO_F64) && !C0::SPLIT && !C0::GEN &&
           C0::LDS2 + (size_t)LL * sizeof(Cx<R>) <= 80 * 1024;
}

// ---------------------------------------------------------------------------
// Row real-to-complex store
// ---------------------------------------------------------------------------

// Packed row spectra: a real row of length M = 2H has real X[0] and X[H];
// they are stored as one element X[0] + i X[H], so a row holds H elements.
// r2c of w[q] (= x[2j] + i x[2j+1] per field, j = t + G*q), stored packed
// to dst[0..H-1].  `lds` is this row's padded buffer viewed as Vec<R, NC>.
// w arrives in the row configuration C's side-B mapping (real space) and the
// spectrum leaves in its side-A mapping.
template <typename R, typename C, int NC>
__device__ __forceinline__ void row_r2c_store(Vec<R, NC> (&w)[C::E], Vec<R, NC>* lds, int t,
                                              const Cx<R>* __restrict__ twH, const Cx<R>* __restrict__ twM,
                                              Buf dst, uint32_t doff) {
    constexpr int H = C::K, EA = C::EA, GA = C::GA;
    fft_ab<R, NC, H, C::E, -1, false, C::GEN, C::PLAN>(w, LdsBuf<R, NC, 1, 0, C::WAVE>{(char*)lds}, t, twH);
    xsync<C::WAVE>();
#pragma unroll
    for (int q = 0; q < EA; q++) lds[pidx(t + GA * q)] = w[q];
    xsync<C::WAVE>();
#pragma unroll
    for (int q = 0; q < EA; q++) {
        const int k = t + GA * q;
        Vec<R, NC> X;
        if (k == 0) {
            // X[0] = Re W0 + Im W0, X[H] = Re W0 - Im W0, both real.
#pragma unroll
            for (int f = 0; f < NC; f++) X.c[f] = Cx<R>{w[q].c[f].x + w[q].c[f].y, w[q].c[f].x - w[q].c[f].y};
        } else {
            const Vec<R, NC> wc = lds[pidx(H - k)];
            const Cx<R> tw = twM[k];
#pragma unroll
            for (int f = 0; f < NC; f++) X.c[f] = r2c_merge(w[q].c[f], wc.c[f], tw);
        }
        bstore(X, dst, (doff + k) * (uint32_t)sizeof(X), 0u);
    }
}


// ---------------------------------------------------------------------------
// Row pass: S, AB -> IN (+ per-row energy partials)
// ---------------------------------------------------------------------------
template <typename R, int H>
__global__ void __launch_bounds__((RowCfg<R, H>::NT)) __attribute__((amdgpu_waves_per_eu(RowCfg<R, H>::WPE)))
k_row_step(DevBufs b, PhysArgs p, int flags) {
    using C = RowCfg<R, H>;
    constexpr int E = C::E, G = C::G, K = C::K;
    __shared__ __attribute__((aligned(16))) char smem[C::LDS];
    __shared__ double red[(C::NT + 63) / 64 + 1];
    const int tid = threadIdx.x, rl = tid / G, t = tid % G;
    const int inst = blockIdx.y;
    if (b.active && !b.active[inst]) return;  // idle instance (uniform per workgroup)
    // Rows of this workgroup from a uniform base; this thread's row at roff.
    const size_t bbase = ((size_t)inst * b.L * b.N + (size_t)blockIdx.x * C::RPW) * K;
    const uint32_t roff = (uint32_t)(rl * K);
    constexpr size_t rows_bytes = (size_t)C::RPW * K * sizeof(Vec<R, 2>);
    const Buf S = mkbuf((const Vec<R, 2>*)b.S + bbase, rows_bytes);
    const Buf AB = mkbuf((const Vec<R, 2>*)b.AB + bbase, rows_bytes);
    const Cx<R>* twH = (const Cx<R>*)b.twH;
    const Cx<R>* twM = (const Cx<R>*)b.twM;

    // One 4-field c2r: psi, con (from S) and a, b (from AB).
    constexpr int EA = C::EA, EB = C::EB, GB = C::GB;
    const bool actB = GB == G || t < GB;  // holds side-B elements
    Vec<R, 4> z[E];
    // TWO (FP32, power-of-two rows): the same c2r two fields at a time, so
    // psi and con are transformed while a and b still arrive.  Loads complete
    // in issue order, so the split twiddles and the transform's twiddle table
    // (staged in LDS) are loaded first, then S, then AB.
    constexpr bool TWO = PFC_ROW_TWO && (sizeof(R) == 4 || PFC_TWO_F64) && !C::GEN;
    __shared__ __attribute__((aligned(16))) Cx<R> tws[TWO ? H : 1];
    if constexpr (TWO) {
        constexpr int NTW = (H + C::NT - 1) / C::NT;
        constexpr uint32_t es = sizeof(Vec<R, 2>);
        Cx<R> twv[NTW], wm[EA];
        Vec<R, 2> s0[EA], s1[EA], a0[EA], a1[EA];
#pragma unroll
        for (int k2 = 0; k2 < NTW; k2++) {
            const int i = tid + k2 * C::NT;
            if (i < H) twv[k2] = twH[i];
        }
#pragma unroll
        for (int q = 0; q < EA; q++) wm[q] = twM[t + G * q];
        __asm__ volatile("" ::: "memory");
#pragma unroll
        for (int q = 0; q < EA; q++) {
            const int k = t + G * q, kc = (H - k) % H;
            s0[q] = bload<Vec<R, 2>>(S, (roff + k) * es, 0u);
            s1[q] = bload<Vec<R, 2>>(S, (roff + kc) * es, 0u);
        }
        __asm__ volatile("" ::: "memory");
#pragma unroll
        for (int q = 0; q < EA; q++) {
            const int k = t + G * q, kc = (H - k) % H;
            a0[q] = bload<Vec<R, 2>>(AB, (roff + k) * es, 0u);
            a1[q] = bload<Vec<R, 2>>(AB, (roff + kc) * es, 0u);
        }
        __asm__ volatile("" ::: "memory");
#pragma unroll
        for (int k2 = 0; k2 < NTW; k2++) {
            const int i = tid + k2 * C::NT;
            if (i < H) tws[i] = twv[k2];
        }
        __syncthreads();
        constexpr int CH2 = (int)(sizeof(Vec<R, 2>) / 16);
        const LdsBuf<R, 2, CH2, PFC_LDS_MOD ? soa_stride(H, CH2, PFC_LDS_MOD) : padded(H), C::WAVE> lds2{
            smem + rl * C::RB};
        Vec<R, 2> zs[E], za[E];
#pragma unroll
        for (int q = 0; q < EA; q++) {
            const int k = t + G * q;
            zs[q].c[0] = c2r_split_packed(s0[q].c[0], s1[q].c[0], k, wm[q]);
            zs[q].c[1] = c2r_split_packed(s0[q].c[1], s1[q].c[1], k, wm[q]);
        }
        fft_ab<R, 2, H, E, +1, true, C::GEN, C::PLAN>(zs, lds2, t, tws);
#pragma unroll
        for (int q = 0; q < EA; q++) {
            const int k = t + G * q;
            za[q].c[0] = c2r_split_packed(a0[q].c[0], a1[q].c[0], k, wm[q]);
            za[q].c[1] = c2r_split_packed(a0[q].c[1], a1[q].c[1], k, wm[q]);
        }
        fft_ab<R, 2, H, E, +1, true, C::GEN, C::PLAN>(za, lds2, t, tws);
#pragma unroll
        for (int q = 0; q < E; q++) {
            z[q].c[0] = zs[q].c[0];
            z[q].c[1] = zs[q].c[1];
            z[q].c[2] = za[q].c[0];
            z[q].c[3] = za[q].c[1];
        }
    }
#pragma unroll
    for (int q = 0; q < (TWO ? 0 : EA); q++) {
        const int k = t + G * q;
        const int kc = (H - k) % H;
        constexpr uint32_t es = sizeof(Vec<R, 2>);
        const Vec<R, 2> s0 = bload<Vec<R, 2>>(S, (roff + k) * es, 0u), s1 = bload<Vec<R, 2>>(S, (roff + kc) * es, 0u);
        const Vec<R, 2> a0 = bload<Vec<R, 2>>(AB, (roff + k) * es, 0u), a1 = bload<Vec<R, 2>>(AB, (roff + kc) * es, 0u);
        const Cx<R> w = twM[k];
        z[q].c[0] = c2r_split_packed(s0.c[0], s1.c[0], k, w);
        z[q].c[1] = c2r_split_packed(s0.c[1], s1.c[1], k, w);
        z[q].c[2] = c2r_split_packed(a0.c[0], a1.c[0], k, w);
        z[q].c[3] = c2r_split_packed(a0.c[1], a1.c[1], k, w);
    }
    if constexpr (!TWO)
        fft_ab<R, 4, H, E, +1, true, C::GEN, C::PLAN>(z, LdsBuf<R, 4, C::CH, C::PS, C::WAVE>{smem + rl * C::RB}, t,
                                                       twH);

    const double* mu = b.means + 4 * inst;
    const R mph = (R)mu[0], mco = (R)mu[1], ma = (R)mu[2], mb = (R)mu[3];
    const R invM = R(1) / R(2 * H);
    Vec<R, 2> w[E];
    double esum = 0.0;
    const bool want_e = (flags & ROW_ENERGY) != 0;
#pragma unroll
    for (int q = 0; q < EB; q++) {
        const R ph[2] = {mph + z[q].c[0].x * invM, mph + z[q].c[0].y * invM};
        const R co[2] = {mco + z[q].c[1].x * invM, mco + z[q].c[1].y * invM};
        const R aa[2] = {ma + z[q].c[2].x * invM, ma + z[q].c[2].y * invM};
        const R bb[2] = {mb + z[q].c[3].x * invM, mb + z[q].c[3].y * invM};
        R inn[2], inc[2], in3[2], in2[2];
#pragma unroll
        for (int s2 = 0; s2 < 2; s2++) {
            nonlinear(p, ph[s2], co[s2], aa[s2], bb[s2], inn[s2], inc[s2], in3[s2], in2[s2]);
            if (want_e && actB) esum += energy_point(p, ph[s2], co[s2], in3[s2], in2[s2]);
        }
        w[q].c[0] = Cx<R>{inn[0], inn[1]};
        w[q].c[1] = Cx<R>{inc[0], inc[1]};
    }
    if (flags & ROW_WRITE_IN)
        row_r2c_store<R, C, 2>(w, (Vec<R, 2>*)(smem + rl * C::RB), t, TWO ? tws : twH, twM,
                                  mkbuf((Vec<R, 2>*)b.IN + bbase, rows_bytes), roff);
    if (want_e) {
        const double se = block_sum<C::NT>(esum, red);

        if (tid == 0) {
            const size_t slot = (size_t)inst * gridDim.x + blockIdx.x;
            b.erow[slot] = se;
        }
    }
}

// Real planes (dpsi, dcon) -> S.
template <typename R, int H>
__global__ void __launch_bounds__((RowCfg<R, H>::NT)) k_row_in(DevBufs b, const R* __restrict__ inphi,
                                                               const R* __restrict__ incon) {
    using C = RowCfg<R, H>;
    constexpr int E = C::E, G = C::G, K = C::K;
    constexpr int M = 2 * H;
    __shared__ __attribute__((aligned(16))) char smem[C::LDS];
    const int tid = threadIdx.x, rl = tid / G, t = tid % G;
    const int inst = blockIdx.y, row = blockIdx.x * C::RPW + rl;
    const size_t rbase = ((size_t)inst * b.L * b.N + row) * K;
    const size_t ps = b.N;  // real point (l, m, n) of row l*N + n at pbase + m*N
    const size_t pbase = ((size_t)inst * b.L + row / b.N) * M * ps + row % b.N;
    Vec<R, 2> w[E];
    if (C::GB == G || t < C::GB) {
#pragma unroll
        for (int q = 0; q < C::EB; q++) {
            const int j = t + C::GB * q;
            w[q].c[0] = Cx<R>{inphi[pbase + 2 * j * ps], inphi[pbase + (2 * j + 1) * ps]};
            w[q].c[1] = Cx<R>{incon[pbase + 2 * j * ps], incon[pbase + (2 * j + 1) * ps]};
        }
    }
    row_r2c_store<R, C, 2>(w, (Vec<R, 2>*)smem + rl * C::PB, t, (const Cx<R>*)b.twH, (const Cx<R>*)b.twM,
                           mkbuf((Vec<R, 2>*)b.S + rbase, K * sizeof(Vec<R, 2>)), 0u);
}

// One real plane per instance -> its packed row spectrum (the noise field).
template <typename R, int H>
__global__ void __launch_bounds__((RowCfg<R, H>::NT)) k_row_r2c1(DevBufs b, const R* __restrict__ in,
                                                                 Vec<R, 1>* __restrict__ out) {
    using C = RowCfg<R, H>;
    constexpr int E = C::E, G = C::G, K = C::K;
    constexpr int M = 2 * H;
    __shared__ __attribute__((aligned(16))) char smem[C::LDS];
    const int tid = threadIdx.x, rl = tid / G, t = tid % G;
    const int inst = blockIdx.y, row = blockIdx.x * C::RPW + rl;
    if (b.active && !b.active[inst]) return;
    const size_t rbase = ((size_t)inst * b.L * b.N + row) * K;
    const size_t ps = b.N;
    const size_t pbase = ((size_t)inst * b.L + row / b.N) * M * ps + row % b.N;
    Vec<R, 1> w[E];
    if (C::GB == G || t < C::GB) {
#pragma unroll
        for (int q = 0; q < C::EB; q++) {
            const int j = t + C::GB * q;
            w[q].c[0] = Cx<R>{in[pbase + 2 * j * ps], in[pbase + (2 * j + 1) * ps]};
        }
    }
    row_r2c_store<R, C, 1>(w, (Vec<R, 1>*)smem + rl * C::PB, t, (const Cx<R>*)b.twH, (const Cx<R>*)b.twM,
                           mkbuf(out + rbase, K * sizeof(Vec<R, 1>)), 0u);
}

// S -> real planes (psi, con) including the means.
template <typename R, int H>
__global__ void __launch_bounds__((RowCfg<R, H>::NT)) k_row_out(DevBufs b, R* __restrict__ outphi,
                                                                R* __restrict__ outcon) {
    using C = RowCfg<R, H>;
    constexpr int E = C::E, G = C::G, K = C::K;
    constexpr int M = 2 * H;
    __shared__ __attribute__((aligned(16))) char smem[C::LDS];
    const int tid = threadIdx.x, rl = tid / G, t = tid % G;
    const int inst = blockIdx.y, row = blockIdx.x * C::RPW + rl;
    const size_t rbase = ((size_t)inst * b.L * b.N + row) * K;
    const size_t ps = b.N;
    const size_t pbase = ((size_t)inst * b.L + row / b.N) * M * ps + row % b.N;
    const Vec<R, 2>* S = (const Vec<R, 2>*)b.S + rbase;
    const Cx<R>* twM = (const Cx<R>*)b.twM;
    Vec<R, 2> z[E];
#pragma unroll
    for (int q = 0; q < C::EA; q++) {
        const int k = t + G * q;
        const int kc = (H - k) % H;
        const Vec<R, 2> s0 = S[k], s1 = S[kc];
        const Cx<R> w = twM[k];
        z[q].c[0] = c2r_split_packed(s0.c[0], s1.c[0], k, w);
        z[q].c[1] = c2r_split_packed(s0.c[1], s1.c[1], k, w);
    }
    fft_ab<R, 2, H, E, +1, true, C::GEN, C::PLAN>(z, LdsBuf<R, 2, 1, 0>{smem + rl * C::PB * sizeof(Vec<R, 2>)}, t,
                                          (const Cx<R>*)b.twH);
    const double* mu = b.means + 4 * inst;
    const R mph = (R)mu[0], mco = (R)mu[1];
    const R invM = R(1) / R(M);
    if (!(C::GB == G || t < C::GB)) return;
#pragma unroll
    for (int q = 0; q < C::EB; q++) {
        const int j = t + C::GB * q;
        outphi[pbase + 2 * j * ps] = mph + z[q].c[0].x * invM;
        outphi[pbase + (2 * j + 1) * ps] = mph + z[q].c[0].y * invM;
        outcon[pbase + 2 * j * ps] = mco + z[q].c[1].x * invM;
        outcon[pbase + (2 * j + 1) * ps] = mco + z[q].c[1].y * invM;
    }
}

// ---------------------------------------------------------------------------
// Column pass
// ---------------------------------------------------------------------------

// Per-mode work of one column tile after the forward transform; x[q]
// becomes {psi', con', C2A psi', C2B psi'} in place.  The packed column needs
// each mode's mirror -k: it is exchanged through `xch` (PB Vecs, then PB noise
// values), which no other thread may be using (the call starts and ends the
// exchange with barriers when the tile is packed).
// COEF (plain step with coefficient tables): cf/cc hold the tabulated
// coefficients of the column's modes (the packed column's: its k_m = 0 modes;
// its k_m = M/2 modes are read from the table's tail).
// x and nz hold the column's full-spectrum elements t + G*q (q < E) in
// threads t < G (GF: the column's thread count; the others hold none).
template <typename R, int LL, int EM, int E, int G, int GF, bool COEF = false, bool LAZY = false, int PF = 0>
__device__ __forceinline__ void col_tile_modes(const DevBufs& b, const PhysArgs& p, int mode, bool want_e,
                                               bool use_nz, Vec<R, 4> (&x)[EM], const Cx<R> (&nz)[EM], int t, int j,
                                               int kmi, int kni, bool packed_tile, bool packed, char* xch,
                                               double& g, const float4* cf = nullptr, const float2* cc = nullptr,
                                               const float* klp = nullptr, const float4* lca = nullptr,
                                               const float2* lcc = nullptr, uint32_t lcs = 0) {
    const bool act = G == GF || t < G;
    Vec<R, 4>* mx4 = (Vec<R, 4>*)xch;
    Cx<R>* nzx = (Cx<R>*)(mx4 + padded(LL));
    if (packed_tile) {
        __syncthreads();
        if (packed && act) {
#pragma unroll
            for (int q = 0; q < E; q++) {
                const int i = pidx(t + G * q);
                mx4[i] = x[q];
                if (use_nz) nzx[i] = nz[q];
            }
        }
        __syncthreads();
    }
    if (!act) return;
#pragma unroll
    for (int q = 0; q < E; q++) {
        const int i = t + G * q;
        Cx<R> ph = x[q].c[0], co = x[q].c[1], a, bb;
        if (COEF && !packed) {
            if constexpr (COEF) {
                Vec<R, 2> sv, ov;
                sv.c[0] = ph; sv.c[1] = co;
                ov.c[0] = x[q].c[2]; ov.c[1] = x[q].c[3];
                const float4 cfq = (LAZY && q >= PF) ? lca[(size_t)i * lcs] : cf[q];
                const float2 ccq = (LAZY && q >= PF) ? lcc[(size_t)i * lcs] : cc[q];
                if constexpr (sizeof(R) == 4) update_mode_coef(cfq, sv, ov, Cx<float>{0.0f, 0.0f});
                else update_mode_coef64(p, cfq, sv, ov, Cx<double>{0.0, 0.0});
                ph = sv.c[0]; co = sv.c[1];
                if (i == 0 && kmi == 0 && kni == 0) {
                    ph = Cx<R>{R(0), R(0)};
                    co = Cx<R>{R(0), R(0)};
                }
                if (want_e) g += klp ? mode_energy_kl(b, LL, i, klp[q], kmi, kni, co) : mode_energy(b, LL, i, kmi, kni, co);
                a = Cx<R>{ph.x * (R)ccq.x, ph.y * (R)ccq.x};
                bb = Cx<R>{ph.x * (R)ccq.y, ph.y * (R)ccq.y};
            }
        } else if (!packed) {
            col_mode(b, p, mode, want_e, use_nz, LL, i, kmi, kni, ph, co, x[q].c[2], x[q].c[3], nz[q], a, bb, g);
        } else if (COEF) {
            if constexpr (COEF) {
                // k_m = 0 from the column's own entries (cf, cc), k_m = M/2 from the table's tail
                const Vec<R, 4> xm = mx4[pidx((LL - i) % LL)];
                const size_t te = coef_packed_tail(b, j) + i;
                const float4 ch = ((const float4*)b.cA)[te];
                const float2 cch = ((const float2*)b.cC)[te];
                packed_modes_coef(b, p, want_e, LL, i, j, ph, co, x[q].c[2], x[q].c[3], xm.c[0], xm.c[1], xm.c[2],
                                  xm.c[3], (LAZY && q >= PF) ? lca[(size_t)i * lcs] : cf[q],
                                  (LAZY && q >= PF) ? lcc[(size_t)i * lcs] : cc[q], ch, cch, a, bb, g);
            }
        } else {
            const int im = pidx((LL - i) % LL);
            const Vec<R, 4> xm = mx4[im];
            const Cx<R> nzm = use_nz ? nzx[im] : Cx<R>{R(0), R(0)};
            Cx<R> ph0, phh, co0, coh, on0, onh, oc0, och, nz0, nzh, a0, ah, b0, bh;
            hsplit(x[q].c[0], xm.c[0], ph0, phh);
            hsplit(x[q].c[1], xm.c[1], co0, coh);
            hsplit(x[q].c[2], xm.c[2], on0, onh);
            hsplit(x[q].c[3], xm.c[3], oc0, och);
            hsplit(nz[q], nzm, nz0, nzh);
            col_mode(b, p, mode, want_e, use_nz, LL, i, 0, j, ph0, co0, on0, oc0, nz0, a0, b0, g);
            col_mode(b, p, mode, want_e, use_nz, LL, i, b.M / 2, j, phh, coh, onh, och, nzh, ah, bh, g);
            ph = hjoin(ph0, phh);
            co = hjoin(co0, coh);
            a = hjoin(a0, ah);
            bb = hjoin(b0, bh);
        }
        x[q].c[0] = ph;
        x[q].c[1] = co;
        x[q].c[2] = a;
        x[q].c[3] = bb;
    }
}


// KIND 0: every mode and flag (init, energy only, noise, energy).
// KIND 1 (COL_LEAN): the plain step (mode STEP, no noise, no energy) with the
// other paths compiled out, which frees registers for occupancy.
// KIND 2 (COL_LEAN_E): the plain step with its energy partials (energy-traced
// steps without noise).  Kinds 1 and 2 read coefficient tables where
// col_coef holds.
enum { COL_GENERAL = 0, COL_LEAN = 1, COL_LEAN_E = 2 };
template <typename R, int LL, int KIND, int TS>
__global__ void __launch_bounds__((ColCfg<R, LL, TS>::NT))
__attribute__((amdgpu_waves_per_eu(KIND != COL_GENERAL ? ColCfg<R, LL, TS>::WPE_LEAN : ColCfg<R, LL, TS>::WPE)))
k_col(DevBufs b, PhysArgs p, int mode, int flags) {
    constexpr bool LEAN = KIND != COL_GENERAL;
    using C = ColCfg<R, LL, TS>;
    constexpr int E = C::E, G = C::G, T = C::T, NT = C::NT, EA = C::EA;
    constexpr bool COEF = PFC_COEF && LEAN && col_coef(sizeof(R) == 8, LL);
    constexpr bool TWO = LEAN && col_two_len<R, LL>();
    __shared__ __attribute__((aligned(16))) char smem[TWO ? C::LDS2 : C::LDS];
    __shared__ double red[(NT + 63) / 64 + 1];
    const int tid = threadIdx.x;
    const int col = C::WAVE ? tid / G : tid % T, t = C::WAVE ? tid % G : tid / T;
    const int inst = blockIdx.y;
    // Energy-traced steps: the last workgroup of each instance also reduces
    // the previous state's partial sums, after issuing its own tile's loads
    // (TWO) so that their latency covers the reduction's.
    const bool reduce_here = KIND != COL_LEAN && b.etrace && blockIdx.x == gridDim.x - 1;
    if (b.active && !b.active[inst]) {  // idle instance (uniform per workgroup)
        if constexpr (KIND != COL_LEAN) {
            if (reduce_here) energy_reduce_block<NT>(b, p, inst, gridDim.x, red);
        }
        return;
    }
    const int K = b.K;
    const size_t CS = (size_t)b.N * K;       // stride between consecutive l
    const int bx = (PFC_XCD_REMAP && gridDim.x % 8 == 0) ? (blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8
                                                         : blockIdx.x;
    const int tile0 = bx * T;                // first column of the tile; T divides K
    const int j = tile0 / K;                 // n slot of the whole tile
    const int kc = tile0 - j * K + col;      // packed m column (kc = 0 holds k_m = 0 and M/2)
    // Element i of this column: uniform tile base + 32-bit offset col + i*CS.
    const size_t tbase = (size_t)inst * LL * CS + tile0;
    constexpr uint32_t es = sizeof(Vec<R, 2>);
    const uint32_t toff = (uint32_t)(col + t * CS) * es;     // bytes
    const size_t tbytes = ((size_t)LL * CS - tile0) * es;    // rest of the instance
    const Buf S = mkbuf((const Vec<R, 2>*)b.S + tbase, tbytes);
    const Cx<R>* twL = (const Cx<R>*)b.twL;
    const typename C::Lds lds{smem + col * C::CB};
    const int Nh = b.N / 2;
    // Columns kc = 0 of slots j = 0 and N/2 carry two Hermitian-in-l sequences
    // (k_m = 0 and M/2); uniform per workgroup.
    const bool packed_tile = tile0 - j * K == 0 && (j == 0 || j == Nh);
    const bool packed = packed_tile && col == 0;
    // Mode of a plain column; the kc = 0 column of the other slots holds
    // k_m = 0 (j < N/2) or k_m = M/2 with k_n = N - j (j > N/2).
    int kmi = kc, kni = j;
    if (kc == 0 && !packed_tile) {
        kmi = j < Nh ? 0 : b.M / 2;
        kni = j < Nh ? j : b.N - j;
    }
    if (LEAN) mode = COL_STEP;
    const bool step = mode == COL_STEP;
    const bool use_nz = !LEAN && step && b.NZ != nullptr;

    // Plain FP32 steps take the modes' coefficients from cA/cC (same layout
    // as one instance of S; loaded with the state, or after the two-field
    // forward transforms, so their latency overlaps work), in the mapping the
    // per-mode work sees (after the forward transform): rows t + GB*q,
    // q < EB, in threads t < GB.
    constexpr int EBc = C::EB, GBc = C::GB;

    // TWO (plain FP32 step, col_two_len): the transforms take two fields at a
    // time in the same per-field arithmetic, so {psi, con} is transformed
    // while the nonlinear terms are still arriving and the new state's stores
    // drain while C2A/C2B x psi is transformed back.  Global accesses complete
    // in issue order (one counter for loads and stores), so nothing the
    // transforms wait for may be issued after the data: the twiddles are
    // loaded first and read from LDS, then the state, the nonlinear terms and
    // the coefficients (after the forward transforms, which do not need them),
    // in that order.  With 16-byte exchange slots the default tiles need half
    // the LDS.  16x512^2: 42.7 -> 40.9 us per launch, the four-class sweep
    // 0.1064 -> 0.1005 ms per step, four 1024^2 instances 81.6 -> 77.7 us per
    // step; one 1024^2 instance +2%; the 3*2^k plans measured slower.
    constexpr size_t TWB = (size_t)LL * sizeof(Cx<R>);
    static_assert(!TWO || C::LDS2 + TWB <= 160 * 1024, "every tile width of a two-field length fits the LDS");
    __shared__ __attribute__((aligned(16))) Cx<R> tws[TWO ? LL : 1];
    const LdsBuf<R, 2, C::CH2, C::PS2, C::WAVE> lds2{smem + col * C::CB2};
    constexpr int NTW = (LL + NT - 1) / NT;
    Cx<R> twv[NTW];
    Vec<R, 2> xs[E], xi[E];
    if constexpr (TWO) {
#pragma unroll
        for (int k = 0; k < NTW; k++) {
            const int i = tid + k * NT;
            if (i < LL) twv[k] = twL[i];
        }
        __asm__ volatile("" ::: "memory");
#pragma unroll
        for (int q = 0; q < EA; q++) xs[q] = bload<Vec<R, 2>>(S, toff, (uint32_t)(G * q * CS) * es);
        __asm__ volatile("" ::: "memory");
        const Buf IN = mkbuf((const Vec<R, 2>*)b.IN + tbase, tbytes);
#pragma unroll
        for (int q = 0; q < EA; q++) xi[q] = bload<Vec<R, 2>>(IN, toff, (uint32_t)(G * q * CS) * es);
        __asm__ volatile("" ::: "memory");
    }
    if constexpr (KIND != COL_LEAN) {
        if (reduce_here) energy_reduce_block<NT>(b, p, inst, gridDim.x, red);
    }
    // Energy-traced plain steps: kl of this thread's modes, loaded here so
    // that the per-mode energy terms do not wait for them after the update.
    constexpr bool KLP = COEF && KIND == COL_LEAN_E;
    float klv[KLP ? C::EB : 1];
    if constexpr (KLP) {
#pragma unroll
        for (int q = 0; q < C::EB; q++) klv[q] = (C::GB == G || t < C::GB) ? b.kl[t + C::GB * q] : 0.0f;
    }
    // LAZY: each mode's coefficients are read where the update uses them
    // (L2 hits: one table serves the batch) except the first PF modes', which
    // are loaded before the last forward transform, so no full array of them
    // stays live across the transforms: 16 x 512^2 column kernel 190 -> 160
    // VGPRs, three workgroups per CU instead of two, 34.5 -> 33.8 us alone
    // (bit-identical).  Not at 1024 (one 1024^2 instance's energy-traced
    // column pass 22.5 -> 23.0 us with every mode read lazily).
    constexpr bool LAZY = COEF && PFC_COL_COEF_LAZY && LL != 1024;
    constexpr int PF = LAZY ? (PFC_COL_COEF_PF < EBc ? PFC_COL_COEF_PF : EBc) : EBc;  // modes prefetched
    float4 cf[COEF && PF > 0 ? PF : 1];
    float2 cc[COEF && PF > 0 ? PF : 1];
    auto load_coef = [&]() {
        if constexpr (COEF && PF > 0) {
            if (GBc == G || t < GBc) {
                const Buf CA = mkbuf((const float4*)b.cA + tile0, 0), CC = mkbuf((const float2*)b.cC + tile0, 0);
                const uint32_t ca_off = (uint32_t)(col + t * CS) * 16u, cc_off = (uint32_t)(col + t * CS) * 8u;
#pragma unroll
                for (int q = 0; q < PF; q++) {
                    cf[q] = bload<float4>(CA, ca_off, (uint32_t)(GBc * q * CS) * 16u);
                    cc[q] = bload<float2>(CC, cc_off, (uint32_t)(GBc * q * CS) * 8u);
                }
            }
        }
    };
    constexpr bool COEF_LATE = TWO && (PFC_COL_COEF_LATE || LAZY);
    if constexpr (!COEF_LATE) load_coef();

    // Forward transform of {psi, con, inn, inc} along l.
    Vec<R, 4> x[E];
    if constexpr (TWO) {
#pragma unroll
        for (int k = 0; k < NTW; k++) {
            const int i = tid + k * NT;
            if (i < LL) tws[i] = twv[k];
        }
        __syncthreads();
        fft_ab<R, 2, LL, E, -1, true, C::GEN, C::PLAN>(xs, lds2, t, tws);
        if constexpr (LAZY) load_coef();  // the first modes' coefficients arrive during the next transform
        fft_ab<R, 2, LL, E, -1, true, C::GEN, C::PLAN>(xi, lds2, t, tws);
        if constexpr (COEF_LATE && !LAZY) load_coef();
#pragma unroll
        for (int q = 0; q < E; q++) {
            x[q].c[0] = xs[q].c[0];
            x[q].c[1] = xs[q].c[1];
            x[q].c[2] = xi[q].c[0];
            x[q].c[3] = xi[q].c[1];
        }
    } else {
#pragma unroll
        for (int q = 0; q < EA; q++) {
            const uint32_t qo = (uint32_t)(G * q * CS) * es;  // uniform
            const Vec<R, 2> sv = bload<Vec<R, 2>>(S, toff, qo);
            x[q].c[0] = sv.c[0];
            x[q].c[1] = sv.c[1];
            if (step) {
                const Vec<R, 2> iv = bload<Vec<R, 2>>(mkbuf((const Vec<R, 2>*)b.IN + tbase, tbytes), toff, qo);
                x[q].c[2] = iv.c[0];
                x[q].c[3] = iv.c[1];
            } else {
                x[q].c[2] = x[q].c[3] = Cx<R>{R(0), R(0)};
            }
        }
        fft_ab<R, 4, LL, E, -1, true, C::GEN, C::PLAN>(x, lds, t, twL);
    }
    Cx<R> nz[E];
#pragma unroll
    for (int q = 0; q < E; q++) nz[q] = Cx<R>{R(0), R(0)};
    if (use_nz) {
        Vec<R, 1> n1[E];
#pragma unroll
        for (int q = 0; q < EA; q++)
            n1[q] = bload<Vec<R, 1>>(mkbuf((const Vec<R, 1>*)b.NZ + tbase, tbytes / 2), toff / 2,
                                     (uint32_t)(G * q * CS) * (es / 2));
        fft_ab<R, 1, LL, E, -1, true, C::GEN, C::PLAN>(n1, LdsBuf<R, 1, 1, 0>{smem + col * C::CB}, t, twL);
#pragma unroll
        for (int q = 0; q < E; q++) nz[q] = n1[q].c[0];
    }
    const bool want_e = KIND == COL_LEAN_E || (KIND == COL_GENERAL && (flags & COL_ENERGY) != 0);
    double g = 0.0;
    col_tile_modes<R, LL, E, C::EB, C::GB, G, COEF, LAZY, PF>(
        b, p, mode, want_e, use_nz, x, nz, t, j, kmi, kni, packed_tile, packed, smem, g, cf, cc, KLP ? klv : nullptr,
        LAZY ? (const float4*)b.cA + tile0 + col : nullptr, LAZY ? (const float2*)b.cC + tile0 + col : nullptr,
        (uint32_t)CS);
    if (step && b.tcount && blockIdx.x == 0 && tid == 0) b.tcount[inst] += 1;
    if (want_e) {
        const double gs = block_sum<NT>(g, red);
        if (tid == 0) b.ecol[(size_t)inst * gridDim.x + blockIdx.x] = gs;
    }
    if (!LEAN && mode == COL_ENERGY_ONLY) return;

    // Inverse transform of the new state and of C2A/C2B times the new psi.
    const R invL = R(1) / R(LL);
    const Buf AB = mkbuf((const Vec<R, 2>*)b.AB + tbase, tbytes);
    if constexpr (TWO) {
        Vec<R, 2> ys[E], ya[E];
#pragma unroll
        for (int q = 0; q < E; q++) {
            ys[q].c[0] = x[q].c[0];
            ys[q].c[1] = x[q].c[1];
            ya[q].c[0] = x[q].c[2];
            ya[q].c[1] = x[q].c[3];
        }
        fft_ab<R, 2, LL, E, +1, false, C::GEN, C::PLAN>(ys, lds2, t, tws);
#pragma unroll
        for (int q = 0; q < EA; q++) {
            Vec<R, 2> sv;
            sv.c[0] = Cx<R>{ys[q].c[0].x * invL, ys[q].c[0].y * invL};
            sv.c[1] = Cx<R>{ys[q].c[1].x * invL, ys[q].c[1].y * invL};
            bstore(sv, S, toff, (uint32_t)(G * q * CS) * es);
        }
        fft_ab<R, 2, LL, E, +1, false, C::GEN, C::PLAN>(ya, lds2, t, tws);
#pragma unroll
        for (int q = 0; q < EA; q++) {
            Vec<R, 2> av;
            av.c[0] = Cx<R>{ya[q].c[0].x * invL, ya[q].c[0].y * invL};
            av.c[1] = Cx<R>{ya[q].c[1].x * invL, ya[q].c[1].y * invL};
            bstore(av, AB, toff, (uint32_t)(G * q * CS) * es);
        }
        return;
    }
    fft_ab<R, 4, LL, E, +1, false, C::GEN, C::PLAN>(x, lds, t, twL);
#pragma unroll
    for (int q = 0; q < EA; q++) {
        const uint32_t qo = (uint32_t)(G * q * CS) * es;
        Vec<R, 2> sv, av;
        sv.c[0] = Cx<R>{x[q].c[0].x * invL, x[q].c[0].y * invL};
        sv.c[1] = Cx<R>{x[q].c[1].x * invL, x[q].c[1].y * invL};
        av.c[0] = Cx<R>{x[q].c[2].x * invL, x[q].c[2].y * invL};
        av.c[1] = Cx<R>{x[q].c[3].x * invL, x[q].c[3].y * invL};
        bstore(sv, S, toff, qo);
        bstore(av, AB, toff, qo);
    }
}

// Per-mode coefficients of the plain step in the column layout: entry
// (l, n slot j, packed column kc) holds the mode k_col gives that column
// (kc = 0 of slots other than 0 and N/2 carries k_m = 0 or M/2 with k_n
// remapped; the packed columns of slots 0 and N/2 carry k_m = 0 and M/2 at
// k_n = j: their entries hold the k_m = 0 mode), then a tail of 2 x L
// entries with the packed columns' k_m = M/2 modes (slot 0, then slot N/2;
// coef_packed_tail).  Same float expressions as update_mode, so the
// coefficient path is bit-identical.
__global__ void __launch_bounds__(256) k_coef(DevBufs b, PhysArgs p, int f64) {
    const long long n = (long long)b.L * b.N * b.K;
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= n + 2LL * b.L) return;
    const int K = b.K, Nh = b.N / 2, Mh = b.M / 2;
    int i, kmi, kni;
    if (e < n) {
        const int kc = (int)(e % K);
        const int j = (int)((e / K) % b.N);
        i = (int)(e / ((long long)K * b.N));
        kmi = kc;
        kni = j;
        if (kc == 0) {
            if (j == 0 || j == Nh) {
                kmi = 0;
            } else {
                kmi = j < Nh ? 0 : Mh;
                kni = j < Nh ? j : b.N - j;
            }
        }
    } else {
        const int s = (int)((e - n) / b.L);
        i = (int)((e - n) % b.L);
        kmi = Mh;
        kni = s == 0 ? 0 : Nh;
    }
    const float ks = ksq(b.kl[i], b.km[kmi], b.kn[kni]);
    const int il = i <= b.L / 2 ? i : b.L - i;
    const int jn = kni <= Nh ? kni : b.N - kni;
    const size_t ci = 2 * (((size_t)il * (Mh + 1) + kmi) * (Nh + 1) + jn);
    ((float4*)const_cast<void*>(b.cA))[e] = f64 ? mode_coef64(p, ks) : mode_coef(p, ks);
    ((float2*)const_cast<void*>(b.cC))[e] = make_float2(b.c2[ci], b.c2[ci + 1]);
}

hipError_t launch_coef(int prec, const DevBufs& b, const PhysArgs& p, hipStream_t s) {
    if (!b.cA || !b.cC) return hipErrorInvalidValue;
    const long long n = (long long)b.L * b.N * b.K + 2LL * b.L;
    hipLaunchKernelGGL(k_coef, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, b, p, prec == PREC_FP64 ? 1 : 0);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Depth pass (3-D grids): FFT along n of the row spectra
// ---------------------------------------------------------------------------
// T packed m columns of one l row per workgroup, G = N/E threads per column.
template <typename R, int NN>
struct DepCfg {
    static constexpr int E = NN >= 64 ? 8 : 4;
    static constexpr int G = NN / E;
    static constexpr int T = 16;  // 16 consecutive m columns: 256 or 512 byte segments per n
    static constexpr int NT = T * G;
    static constexpr int PB = padded(NN);
    static constexpr size_t LDS = (size_t)T * PB * sizeof(Vec<R, 4>);
};

// NC = 1: one complex field (in0/out0 are Cx<R>); NC = 2: one field pair
// (Vec<R, 2>); NC = 4: two field pairs (in1 null reads zeros, out1 null skips).
// Along n the kc = 0 column holds Z = X0 + i XH with X0, XH real; the forward
// pass replaces slot j in (0, N/2) by X0_hat(j) and slot N - j by XH_hat(j)
// (slots 0 and N/2 keep the packed pair), the inverse pass undoes it.
template <typename R, int NN, int NC, int DIR>
__global__ void __launch_bounds__((DepCfg<R, NN>::NT)) k_depth(DevBufs b, const void* in0, void* out0,
                                                               const void* in1, void* out1) {
    using C = DepCfg<R, NN>;
    constexpr int E = C::E, G = C::G, T = C::T;
    __shared__ __attribute__((aligned(16))) char smem[C::LDS];
    const int tid = threadIdx.x, col = tid % T, t = tid / T;
    const int inst = blockIdx.y;
    if (b.active && !b.active[inst]) return;  // idle instance (uniform per workgroup)
    const int K = b.K;
    const int tpr = K / T;  // tiles per l row; T divides K (checked at launch)
    const int l = blockIdx.x / tpr;
    const int kc0 = (blockIdx.x - l * tpr) * T;
    const size_t base = ((size_t)inst * b.L + l) * NN * K + kc0;  // slot n of column col at base + col + n*K
    const uint32_t toff = (uint32_t)(col + t * K);            // elements
    const size_t nbytes_el = (size_t)NN * K - kc0;            // elements from base to the end of this l
    Vec<R, NC>* lds = (Vec<R, NC>*)smem + col * C::PB;
    const bool split_col = kc0 == 0 && col == 0;

    Vec<R, NC> x[E];
#pragma unroll
    for (int q = 0; q < E; q++) {
        const uint32_t e = (uint32_t)(G * q * K);  // uniform, elements
        if constexpr (NC == 1) {
            constexpr uint32_t es = sizeof(Cx<R>);
            x[q].c[0] = bload<Cx<R>>(mkbuf((const Cx<R>*)in0 + base, nbytes_el * es), toff * es, e * es);
        } else {
            constexpr uint32_t es = sizeof(Vec<R, 2>);
            const Vec<R, 2> v = bload<Vec<R, 2>>(mkbuf((const Vec<R, 2>*)in0 + base, nbytes_el * es), toff * es, e * es);
            x[q].c[0] = v.c[0];
            x[q].c[1] = v.c[1];
            if constexpr (NC == 4) {
                Vec<R, 2> w;
                if (in1) w = bload<Vec<R, 2>>(mkbuf((const Vec<R, 2>*)in1 + base, nbytes_el * es), toff * es, e * es);
                else w.c[0] = w.c[1] = Cx<R>{R(0), R(0)};
                x[q].c[2] = w.c[0];
                x[q].c[3] = w.c[1];
            }
        }
    }
    if constexpr (DIR > 0) {
        // Join: slot j < N/2 becomes X0_hat(j) + i XH_hat(j), slot N - j
        // conj(X0_hat(j)) + i conj(XH_hat(j)); the mirror slot comes via LDS.
        if (kc0 == 0) {  // uniform per workgroup
            if (split_col) {
#pragma unroll
                for (int q = 0; q < E; q++) lds[pidx(t + G * q)] = x[q];
            }
            __syncthreads();
            if (split_col) {
#pragma unroll
                for (int q = 0; q < E; q++) {
                    const int jj = t + G * q;
                    if (jj == 0 || jj == NN / 2) continue;
                    const Vec<R, NC> xm = lds[pidx(NN - jj)];
#pragma unroll
                    for (int f = 0; f < NC; f++) {
                        if (jj < NN / 2) {
                            x[q].c[f] = hjoin(x[q].c[f], xm.c[f]);
                        } else {
                            const Cx<R> x0c{xm.c[f].x, -xm.c[f].y}, xhc{x[q].c[f].x, -x[q].c[f].y};
                            x[q].c[f] = hjoin(x0c, xhc);
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
    fft_reg<R, NC, NN, E, DIR>(x, lds, t, (const Cx<R>*)b.twN);
    if constexpr (DIR < 0) {
        if (kc0 == 0) {  // uniform per workgroup
            __syncthreads();
            if (split_col) {
#pragma unroll
                for (int q = 0; q < E; q++) lds[pidx(t + G * q)] = x[q];
            }
            __syncthreads();
            if (split_col) {
#pragma unroll
                for (int q = 0; q < E; q++) {
                    const int jj = t + G * q;
                    if (jj == 0 || jj == NN / 2) continue;
                    const Vec<R, NC> xm = lds[pidx(NN - jj)];
#pragma unroll
                    for (int f = 0; f < NC; f++) {
                        Cx<R> x0, xh;
                        if (jj < NN / 2) {
                            hsplit(x[q].c[f], xm.c[f], x0, xh);
                            x[q].c[f] = x0;
                        } else {
                            hsplit(xm.c[f], x[q].c[f], x0, xh);
                            x[q].c[f] = xh;
                        }
                    }
                }
            }
        }
    }
    const R sc = DIR > 0 ? R(1) / R(NN) : R(1);
#pragma unroll
    for (int q = 0; q < E; q++) {
        const uint32_t e = (uint32_t)(G * q * K);
        if constexpr (DIR > 0) {
#pragma unroll
            for (int f = 0; f < NC; f++) x[q].c[f] = Cx<R>{x[q].c[f].x * sc, x[q].c[f].y * sc};
        }
        if constexpr (NC == 1) {
            constexpr uint32_t es = sizeof(Cx<R>);
            bstore(x[q].c[0], mkbuf((Cx<R>*)out0 + base, nbytes_el * es), toff * es, e * es);
        } else {
            Vec<R, 2> v;
            v.c[0] = x[q].c[0];
            v.c[1] = x[q].c[1];
            constexpr uint32_t es = sizeof(Vec<R, 2>);
            bstore(v, mkbuf((Vec<R, 2>*)out0 + base, nbytes_el * es), toff * es, e * es);
            if constexpr (NC == 4) {
                if (out1) {
                    Vec<R, 2> w;
                    w.c[0] = x[q].c[2];
                    w.c[1] = x[q].c[3];
                    bstore(w, mkbuf((Vec<R, 2>*)out1 + base, nbytes_el * es), toff * es, e * es);
                }
            }
        }
    }
}

// Per-instance energy: e_out = {(sum e + alpha/P sum_k w|c|^2)/P, mean psi}, P = L*M*N.
// One workgroup per instance: strided FP64 partial sums, then the wavefront
// and the four waves combined in a fixed order (the result does not depend
// on the batch size or on timing).
constexpr int ENERGY_NT = 256;
__global__ __launch_bounds__(ENERGY_NT) void k_energy_reduce(DevBufs b, int nrow, int ncol, double alpha,
                                                             double* e_out) {
    const int inst = blockIdx.x;
    const int t = threadIdx.x;
    const double* er = b.erow + (size_t)inst * nrow;
    const double* ec = b.ecol + (size_t)inst * ncol;
    double se = 0.0, sg = 0.0;
    for (int i = t; i < nrow; i += ENERGY_NT) se += er[i];
    for (int i = t; i < ncol; i += ENERGY_NT) sg += ec[i];
    for (int o = 32; o > 0; o >>= 1) {
        se += __shfl_xor(se, o);
        sg += __shfl_xor(sg, o);
    }
    __shared__ double part[2][ENERGY_NT / 64];
    if ((t & 63) == 0) {
        part[0][t >> 6] = se;
        part[1][t >> 6] = sg;
    }
    __syncthreads();
    if (t != 0) return;
    se = (part[0][0] + part[0][1]) + (part[0][2] + part[0][3]);
    sg = (part[1][0] + part[1][1]) + (part[1][2] + part[1][3]);
    // Density: the k=0 mode, held in FP64 and left unchanged by every step
    // (k = 0 makes the update an identity, misfit.cpp:1002-1010); the sum of
    // the reconstructed deviations is zero up to transform round-off.
    energy_pair(b, alpha, inst, se, sg, e_out);
}

// ---------------------------------------------------------------------------
// Host dispatch
// ---------------------------------------------------------------------------
bool size_supported(int n) {
    switch (n) {
        case 32: case 64: case 128: case 256: case 384: case 512: case 768: case 1024: return true;
        default: return false;
    }
}

bool depth_supported(int n) {
    switch (n) {
        case 16: case 32: case 64: case 128: return true;
        default: return false;
    }
}

template <typename R>
static int row_blocks_t(int M, int L) {
    int rpw = 1;
    switch (M / 2) {
#define RB(H) case H: rpw = RowCfg<R, H>::RPW; break;
        RB(16) RB(32) RB(64) RB(128) RB(192) RB(256) RB(384) RB(512)
#undef RB
        default: return -1;
    }
    if (L % rpw) return -1;
    return L / rpw;
}

template <typename R, int TS>
static int col_tiles_t(int L) {
    switch (L) {
#define CB(LL) case LL: return ColCfg<R, LL, TS>::T;
        CB(32) CB(64) CB(128) CB(256) CB(384) CB(512) CB(768) CB(1024)
#undef CB
        default: return -1;
    }
}

template <typename R>
static int col_blocks_t(int L, int M, int N, int ts = 0) {
    const int T = ts == 1 ? col_tiles_t<R, 1>(L) : (ts == 2 ? col_tiles_t<R, 2>(L) : col_tiles_t<R, 0>(L));
    if (T <= 0) return -1;
    const int K = M / 2;
    if (K % T) return -1;  // k_col handles whole tiles of T columns only
    return N * (K / T);
}

int col_tiles(int prec, int L, int ts) {
    if (prec == PREC_FP64)
        return ts == 1 ? col_tiles_t<double, 1>(L) : (ts == 2 ? col_tiles_t<double, 2>(L) : col_tiles_t<double, 0>(L));
    return ts == 1 ? col_tiles_t<float, 1>(L) : (ts == 2 ? col_tiles_t<float, 2>(L) : col_tiles_t<float, 0>(L));
}

int row_blocks(int prec, int M, int rows) {
    if (fs_row(prec, M / 2)) return rows;  // the field-split row pass: one row per workgroup
    return prec == PREC_FP64 ? row_blocks_t<double>(M, rows) : row_blocks_t<float>(M, rows);
}
int col_blocks(int prec, int L, int M, int N) {
    return prec == PREC_FP64 ? col_blocks_t<double>(L, M, N) : col_blocks_t<float>(L, M, N);
}

template <typename R>
static hipError_t row_step_t(const DevBufs& b, const PhysArgs& p, int flags, hipStream_t s, hipEvent_t t0,
                             hipEvent_t t1) {
    constexpr int prec = sizeof(R) == 8 ? PREC_FP64 : PREC_FP32;
    if (fs_row(prec, b.M / 2)) return launch_row_fs(prec, b, p, flags, s, t0, t1);
    const int nb = row_blocks_t<R>(b.M, b.L * b.N);
    if (nb <= 0) return hipErrorInvalidValue;
    dim3 grid(nb, b.B);
    switch (b.M / 2) {
#define RS(H) case H: PFC_LAUNCH((k_row_step<R, H>), grid, dim3(RowCfg<R, H>::NT), s, t0, t1, b, p, flags); break;
        RS(16) RS(32) RS(64) RS(128) RS(192) RS(256) RS(384) RS(512)
#undef RS
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename R, int LL>
static void launch_col_e(dim3 grid, hipStream_t s, hipEvent_t t0, hipEvent_t t1, const DevBufs& b,
                         const PhysArgs& p, int mode, int flags) {
    PFC_LAUNCH((k_col<R, LL, COL_LEAN_E, 0>), grid, dim3(ColCfg<R, LL, 0>::NT), s, t0, t1, b, p, mode, flags);
}

template <typename R>
static hipError_t col_t(const DevBufs& b, const PhysArgs& p, int mode, int flags, hipStream_t s, hipEvent_t t0,
                        hipEvent_t t1) {
    // The specialised kernels of a length whose plain step reads coefficient
    // tables need them (every context allocates them where col_coef holds).
    const bool tables_ok = !col_coef(sizeof(R) == 8, b.L) || b.cA != nullptr;
    const bool lean = mode == COL_STEP && !b.NZ && !(flags & COL_ENERGY) && tables_ok;
    // Energy-traced steps without noise: the plain-step kernel with the
    // energy partials (default tiles: the partial count is col_blocks).
    const bool lean_e = mode == COL_STEP && !b.NZ && (flags & COL_ENERGY) && tables_ok;
    // The plain step runs the tile width the autotuner chose; the other
    // modes (and a width whose tiles do not divide K) the default.
    constexpr int prec = sizeof(R) == 8 ? PREC_FP64 : PREC_FP32;
    if ((lean || lean_e) && fs_col(prec, b.L)) {
        // one column per workgroup: the default tiles' partial count (col_t)
        if (col_blocks_t<R>(b.L, b.M, b.N) != b.N * b.K) return hipErrorInvalidValue;
        return launch_col_fs(prec, b, p, lean_e, s, t0, t1);
    }
    int ts = lean ? b.col_ts : 0;
    int nb = col_blocks_t<R>(b.L, b.M, b.N, ts);
    if (nb <= 0 && ts) nb = col_blocks_t<R>(b.L, b.M, b.N, ts = 0);
    if (nb <= 0) return hipErrorInvalidValue;
    dim3 grid(nb, b.B);
    switch (b.L) {
#define CS(LL)                                                                                                  \
    case LL:                                                                                                    \
        if (lean_e) launch_col_e<R, LL>(grid, s, t0, t1, b, p, mode, flags);                                      \
        else if (!lean) PFC_LAUNCH((k_col<R, LL, COL_GENERAL, 0>), grid, dim3(ColCfg<R, LL, 0>::NT), s, t0, t1, b, p, mode, flags); \
        else if (ts == 1) PFC_LAUNCH((k_col<R, LL, COL_LEAN, 1>), grid, dim3(ColCfg<R, LL, 1>::NT), s, t0, t1, b, p, mode, flags); \
        else if (ts == 2) PFC_LAUNCH((k_col<R, LL, COL_LEAN, 2>), grid, dim3(ColCfg<R, LL, 2>::NT), s, t0, t1, b, p, mode, flags); \
        else PFC_LAUNCH((k_col<R, LL, COL_LEAN, 0>), grid, dim3(ColCfg<R, LL, 0>::NT), s, t0, t1, b, p, mode, flags); \
        break;
        CS(32) CS(64) CS(128) CS(256) CS(384) CS(512) CS(768) CS(1024)
#undef CS
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename R>
static hipError_t row_in_t(const DevBufs& b, const void* inphi, const void* incon, hipStream_t s) {
    const int nb = row_blocks_t<R>(b.M, b.L * b.N);
    if (nb <= 0) return hipErrorInvalidValue;
    dim3 grid(nb, b.B);
    switch (b.M / 2) {
#define RI(H) case H: hipLaunchKernelGGL((k_row_in<R, H>), grid, dim3(RowCfg<R, H>::NT), 0, s, b, (const R*)inphi, (const R*)incon); break;
        RI(16) RI(32) RI(64) RI(128) RI(192) RI(256) RI(384) RI(512)
#undef RI
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename R>
static hipError_t row_out_t(const DevBufs& b, void* outphi, void* outcon, hipStream_t s) {
    const int nb = row_blocks_t<R>(b.M, b.L * b.N);
    if (nb <= 0) return hipErrorInvalidValue;
    dim3 grid(nb, b.B);
    switch (b.M / 2) {
#define RO(H) case H: hipLaunchKernelGGL((k_row_out<R, H>), grid, dim3(RowCfg<R, H>::NT), 0, s, b, (R*)outphi, (R*)outcon); break;
        RO(16) RO(32) RO(64) RO(128) RO(192) RO(256) RO(384) RO(512)
#undef RO
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename R>
static hipError_t row_r2c1_t(const DevBufs& b, const void* in, void* out, hipStream_t s) {
    const int nb = row_blocks_t<R>(b.M, b.L * b.N);
    if (nb <= 0) return hipErrorInvalidValue;
    dim3 grid(nb, b.B);
    switch (b.M / 2) {
#define R1(H) case H: hipLaunchKernelGGL((k_row_r2c1<R, H>), grid, dim3(RowCfg<R, H>::NT), 0, s, b, (const R*)in, (Vec<R, 1>*)out); break;
        R1(16) R1(32) R1(64) R1(128) R1(192) R1(256) R1(384) R1(512)
#undef R1
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename R, int NC, int DIR>
static hipError_t depth_t(const DevBufs& b, const void* in0, void* out0, const void* in1, void* out1, hipStream_t s,
                          hipEvent_t t0, hipEvent_t t1) {
    constexpr int T = 16;
    if (b.K % T) return hipErrorInvalidValue;
    dim3 grid(b.L * (b.K / T), b.B);
    switch (b.N) {
#define DP(NN) case NN: static_assert(DepCfg<R, NN>::T == T, "tile width"); \
        PFC_LAUNCH((k_depth<R, NN, NC, DIR>), grid, dim3(DepCfg<R, NN>::NT), s, t0, t1, b, in0, out0, in1, out1); break;
        DP(16) DP(32) DP(64) DP(128)
#undef DP
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename R>
static hipError_t depth_r(const DevBufs& b, int dir, const void* in0, void* out0, const void* in1, void* out1,
                          bool one_field, hipStream_t s, hipEvent_t t0, hipEvent_t t1) {
    if (dir < 0) {
        if (one_field) return depth_t<R, 1, -1>(b, in0, out0, nullptr, nullptr, s, t0, t1);
        if (!in1 && !out1) return depth_t<R, 2, -1>(b, in0, out0, nullptr, nullptr, s, t0, t1);
        return depth_t<R, 4, -1>(b, in0, out0, in1, out1, s, t0, t1);
    }
    if (one_field || !in1 || !out1) return hipErrorInvalidValue;
    return depth_t<R, 4, +1>(b, in0, out0, in1, out1, s, t0, t1);
}

hipError_t launch_depth(int prec, const DevBufs& b, int dir, const void* in0, void* out0, const void* in1,
                        void* out1, bool one_field, hipStream_t s, hipEvent_t t0, hipEvent_t t1) {
    return prec == PREC_FP64 ? depth_r<double>(b, dir, in0, out0, in1, out1, one_field, s, t0, t1)
                             : depth_r<float>(b, dir, in0, out0, in1, out1, one_field, s, t0, t1);
}

hipError_t launch_row_r2c1(int prec, const DevBufs& b, const void* in, void* out, hipStream_t s) {
    return prec == PREC_FP64 ? row_r2c1_t<double>(b, in, out, s) : row_r2c1_t<float>(b, in, out, s);
}

hipError_t launch_row_step(int prec, const DevBufs& b, const PhysArgs& p, int flags, hipStream_t s, hipEvent_t t0,
                           hipEvent_t t1) {
    return prec == PREC_FP64 ? row_step_t<double>(b, p, flags, s, t0, t1) : row_step_t<float>(b, p, flags, s, t0, t1);
}
hipError_t launch_col(int prec, const DevBufs& b, const PhysArgs& p, int mode, int flags, hipStream_t s,
                      hipEvent_t t0, hipEvent_t t1) {
    return prec == PREC_FP64 ? col_t<double>(b, p, mode, flags, s, t0, t1)
                             : col_t<float>(b, p, mode, flags, s, t0, t1);
}
hipError_t launch_row_in(int prec, const DevBufs& b, const void* inphi, const void* incon, hipStream_t s) {
    return prec == PREC_FP64 ? row_in_t<double>(b, inphi, incon, s) : row_in_t<float>(b, inphi, incon, s);
}
hipError_t launch_row_out(int prec, const DevBufs& b, void* outphi, void* outcon, hipStream_t s) {
    return prec == PREC_FP64 ? row_out_t<double>(b, outphi, outcon, s) : row_out_t<float>(b, outphi, outcon, s);
}
hipError_t launch_energy_reduce(const DevBufs& b, int nrow, int ncol, double alpha, double* e_out, hipStream_t s) {
    hipLaunchKernelGGL(k_energy_reduce, dim3(b.B), dim3(ENERGY_NT), 0, s, b, nrow, ncol, alpha, e_out);
    return hipGetLastError();
}

}  // namespace pfc
