// pfc_fs.hip - field-split row and column passes for gfx950.
//
// The passes of pfc_step.hip give each thread the four fields of its
// elements (psi, con, and C2A/C2B x psi or the nonlinear terms), so one row or
// column is one chain of load -> four-field transform -> pointwise work ->
// transform -> store in one or two wavefronts.  A single instance of 512^2
// then launches 512 waves for the 1024 SIMDs, and each wave's chain (about
// 3500 VALU instructions in the FP64 row pass) is the kernel's time.
//
// Here each field has its own group of G threads (one wavefront from 256
// points up), so a row or column is four transforms side by side:
//
//   row pass    group f loads field f of S (psi, con) or AB (C2A psi, C2B psi)
//               -> c2r along m -> real values into LDS plane f; every thread
//               then evaluates the nonlinear terms (misfit.cpp:777-874) on
//               its share of the row from the four planes; groups 0 and 1 r2c
//               inn and inc and store them into IN.
//   column pass group f loads field f of S or IN -> FFT along l -> LDS plane
//               f; every thread updates its share of the column's modes
//               (misfit.cpp:1002-1010 via col_mode) and forms C2A/C2B x psi';
//               group f inverse-transforms plane f into S (psi', con') or
//               AB.
//
// The per-field arithmetic of the transforms, the nonlinear terms and the
// update is the template code the four-field passes run; the exchanges of a
// transform stay inside one wavefront (no workgroup barriers), and the only
// workgroup barriers are the two (row) or three (column) between the field
// phases.  Used where fs_row / fs_col (pfc_kernels.h) hold, at every batch
// size, so a batch member and the same fields alone take the same path.
#include "pfc_fft.hpp"
#include "pfc_kernels.h"
#include "pfc_step_dev.hpp"

namespace pfc {

// Elements per thread of one field's transform of length n: n/64 (one
// wavefront per field) from 256 points up, else 4.
constexpr int fs_e(int n) { return n >= 256 ? n / 64 : 4; }

#ifndef PFC_FS_ROW256_E
#define PFC_FS_ROW256_E 4  // elements per thread of 256-point field-split rows
#endif
template <typename R, int H>
struct FsRow {
    static constexpr int E = H == 256 ? PFC_FS_ROW256_E : fs_e(H), G = H / E, NT = 4 * G, PB = padded(H);
    static constexpr int EP = E / 4;  // row elements (pairs of real points) per thread, pointwise phase
    static_assert(G <= 64 && 64 % G == 0, "one field's transform inside one wavefront");
};

// TWL: the twiddle table is staged in LDS where that still leaves room for
// two workgroups per CU (FP64 1024: 74 KB of planes; with the table, one
// workgroup per CU: 33.7 -> 44.2 us per pass).
#ifndef PFC_FS_PAIR_MIN
#define PFC_FS_PAIR_MIN 256  // shortest column whose loads/stores go in lane pairs through the planes
#endif
#ifndef PFC_FS_COL512_E
#define PFC_FS_COL512_E 8  // elements per thread of 512-point field-split columns
#endif
template <typename R, int LL>
struct FsCol {
    static constexpr int E = LL == 512 ? PFC_FS_COL512_E : fs_e(LL), G = LL / E, NT = 4 * G, PB = padded(LL);
    static constexpr int MP = E / 4;  // modes per thread, update phase
    static constexpr bool WAVE = G <= 64 && 64 % G == 0;  // one field's transform inside one wavefront
    static_assert(WAVE || G % 64 == 0, "whole wavefronts per field");
    static constexpr bool TWL = (4 * PB + LL) * sizeof(Cx<R>) <= 80 * 1024;
    // PAIR: the column's loads and stores go in pieces of which lane pairs
    // cover one row's two fields, through the planes.  With one column per
    // workgroup every row is a separate cache line for each array; split by
    // field, each line is asked for twice (once per field group).  FP64
    // 512: 14.4 -> 12.9 us per pass, 1024: 37.7 -> 32.8; 128: +0.3 us.
    static constexpr bool PAIR = LL >= PFC_FS_PAIR_MIN;
};

// Row pass, one row of length M = 2H per workgroup (row blockIdx.x of
// instance blockIdx.y; rows are l*N + n).  flags: ROW_WRITE_IN, ROW_ENERGY
// (per-row sums of the pointwise energy terms into erow).
template <typename R, int H>
__global__ void __launch_bounds__((FsRow<R, H>::NT)) k_row_fs(DevBufs b, PhysArgs p, int flags) {
    using C = FsRow<R, H>;
    constexpr int E = C::E, G = C::G, NT = C::NT, PB = C::PB;
    __shared__ __attribute__((aligned(16))) Cx<R> pl[4][PB];
    __shared__ __attribute__((aligned(16))) Cx<R> twH[H], twM[H];  // exp(-2 pi i k/H), exp(-2 pi i k/M), k < H
    __shared__ double red[NT / 64 + 2];
    const int tid = threadIdx.x, f = tid / G, t = tid % G;
    const int inst = blockIdx.y;
    if (b.active && !b.active[inst]) return;  // idle instance (uniform per workgroup)
    const size_t rbase = ((size_t)inst * b.L * b.N + blockIdx.x) * H;
    constexpr uint32_t es = sizeof(Vec<R, 2>);
    const uint32_t fo = (uint32_t)(f & 1) * (uint32_t)sizeof(Cx<R>);  // field within the pair
    const LdsBuf<R, 1, 1, 0, true> lds{(char*)pl[f]};

    // c2r of field f along m: packed X[k], X[H-k] -> split -> H-point inverse
    // FFT, with both twiddle tables staged in LDS (every pass reads them
    // there).  (No compiler barriers order these loads: with one, the arrays
    // held across it stay in scratch memory.)
    {
        constexpr int NTW = H / NT;
        static_assert(H % NT == 0, "whole twiddle rows per thread");
        Cx<R> th[NTW], tm[NTW];
#pragma unroll
        for (int k2 = 0; k2 < NTW; k2++) {
            th[k2] = ((const Cx<R>*)b.twH)[tid + k2 * NT];
            tm[k2] = ((const Cx<R>*)b.twM)[tid + k2 * NT];
        }
        const Buf src = mkbuf((const Vec<R, 2>*)(f < 2 ? b.S : b.AB) + rbase, 0);
        Cx<R> x0[E], x1[E];
#pragma unroll
        for (int q = 0; q < E; q++) {
            const int k = t + G * q, kc = (H - k) % H;
            x0[q] = bload<Cx<R>>(src, (uint32_t)k * es + fo, 0u);
            x1[q] = bload<Cx<R>>(src, (uint32_t)kc * es + fo, 0u);
        }
#pragma unroll
        for (int k2 = 0; k2 < NTW; k2++) {
            twH[tid + k2 * NT] = th[k2];
            twM[tid + k2 * NT] = tm[k2];
        }
        __syncthreads();
        Vec<R, 1> z[E];
#pragma unroll
        for (int q = 0; q < E; q++) z[q].c[0] = c2r_split_packed(x0[q], x1[q], t + G * q, twM[t + G * q]);
        fft_reg<R, 1, H, E, +1>(z, lds, t, twH);
        xsync<true>();
#pragma unroll
        for (int q = 0; q < E; q++) pl[f][pidx(t + G * q)] = z[q].c[0];
    }
    __syncthreads();

    // Nonlinear terms (and energy terms) of row elements j = tid + NT*r, each
    // the real points 2j, 2j+1: inn into plane 2, inc into plane 3.
    const double* mu = b.means + 4 * inst;
    const R mph = (R)mu[0], mco = (R)mu[1], ma = (R)mu[2], mb = (R)mu[3];
    const R invM = R(1) / R(2 * H);
    const bool want_e = (flags & ROW_ENERGY) != 0;
    double esum = 0.0;
#pragma unroll
    for (int r = 0; r < C::EP; r++) {
        const int jj = pidx(tid + NT * r);
        const Cx<R> zp = pl[0][jj], zc = pl[1][jj], za = pl[2][jj], zb = pl[3][jj];
        const R ph[2] = {mph + zp.x * invM, mph + zp.y * invM};
        const R co[2] = {mco + zc.x * invM, mco + zc.y * invM};
        const R aa[2] = {ma + za.x * invM, ma + za.y * invM};
        const R bb[2] = {mb + zb.x * invM, mb + zb.y * invM};
        R inn[2], inc[2], in3[2], in2[2];
#pragma unroll
        for (int s2 = 0; s2 < 2; s2++) {
            nonlinear(p, ph[s2], co[s2], aa[s2], bb[s2], inn[s2], inc[s2], in3[s2], in2[s2]);
            if (want_e) esum += energy_point(p, ph[s2], co[s2], in3[s2], in2[s2]);
        }
        pl[2][jj] = Cx<R>{inn[0], inn[1]};
        pl[3][jj] = Cx<R>{inc[0], inc[1]};
    }
    __syncthreads();

    // r2c of inn (group 0) and inc (group 1); planes 0 and 1 are free now and
    // serve as their exchange buffers.
    if ((flags & ROW_WRITE_IN) && f < 2) {
        Vec<R, 1> w[E];
#pragma unroll
        for (int q = 0; q < E; q++) w[q].c[0] = pl[2 + f][pidx(t + G * q)];
        fft_reg<R, 1, H, E, -1>(w, lds, t, twH);
        xsync<true>();
#pragma unroll
        for (int q = 0; q < E; q++) pl[f][pidx(t + G * q)] = w[q].c[0];
        xsync<true>();
        const Buf dst = mkbuf((Vec<R, 2>*)b.IN + rbase, 0);
#pragma unroll
        for (int q = 0; q < E; q++) {
            const int k = t + G * q;
            const Cx<R> wk = w[q].c[0];
            Cx<R> X;
            if (k == 0) X = Cx<R>{wk.x + wk.y, wk.x - wk.y};  // X[0], X[H] (both real), packed
            else X = r2c_merge(wk, pl[f][pidx(H - k)], twM[k]);
            bstore(X, dst, (uint32_t)k * es + fo, 0u);
        }
    }
    if (want_e) {
        const double se = block_sum<NT>(esum, red);
        if (tid == 0) b.erow[(size_t)inst * gridDim.x + blockIdx.x] = se;
    }
}

// Column pass, plain step (WANT_E: with the gradient-energy partials of the
// updated state, and the previous state's energy pair when b.etrace is set),
// one packed m column (slot j, column kc) of length L per workgroup.
template <typename R, int LL, bool WANT_E>
__global__ void __launch_bounds__((FsCol<R, LL>::NT)) k_col_fs(DevBufs b, PhysArgs p) {
    using C = FsCol<R, LL>;
    constexpr int E = C::E, G = C::G, NT = C::NT, PB = C::PB;
    __shared__ __attribute__((aligned(16))) Cx<R> pl[4][PB];
    __shared__ __attribute__((aligned(16))) Cx<R> twS[C::TWL ? LL : 1];  // exp(-2 pi i k/L), staged
    __shared__ double red[NT / 64 + 2];
    const int tid = threadIdx.x, f = tid / G, t = tid % G;
    const int inst = blockIdx.y;
    // WANT_E: the last workgroup of each instance also reduces the previous
    // state's partial sums, after issuing its column's loads.
    const bool reduce_here = WANT_E && b.etrace && blockIdx.x == gridDim.x - 1;
    if (b.active && !b.active[inst]) {  // idle instance (uniform per workgroup)
        if constexpr (WANT_E) {
            if (reduce_here) energy_reduce_block<NT>(b, p, inst, gridDim.x, red);
        }
        return;
    }
    const int K = b.K, Nh = b.N / 2;
    const size_t CS = (size_t)b.N * K;  // elements between consecutive l
    // Adjacent columns on one XCD (workgroup x takes column (x % 8)*(n/8) + x/8),
    // so its L2 sees whole lines of each row.
    const int bx = gridDim.x % 8 == 0 ? (blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8 : blockIdx.x;
    const int j = bx / K, kc = bx - j * K;
    // Columns kc = 0 of slots 0 and N/2 carry two Hermitian-in-l sequences
    // (k_m = 0 and M/2); the kc = 0 column of other slots holds k_m = 0
    // (j < N/2) or k_m = M/2 with k_n = N - j (j > N/2).
    const bool packed = kc == 0 && (j == 0 || j == Nh);
    int kmi = kc, kni = j;
    if (kc == 0 && !packed) {
        kmi = j < Nh ? 0 : b.M / 2;
        kni = j < Nh ? j : b.N - j;
    }
    const size_t tbase = (size_t)inst * LL * CS + bx;
    constexpr uint32_t es = sizeof(Vec<R, 2>);
    const uint32_t fo = (uint32_t)(f & 1) * (uint32_t)sizeof(Cx<R>);
    const uint32_t rs = (uint32_t)(CS * es);  // bytes between consecutive l
    const LdsBuf<R, 1, 1, 0, C::WAVE> lds{(char*)pl[f]};

    // Loads: the twiddle table (staged in LDS for every pass), the
    // coefficients of this thread's modes (tables in the column layout, used
    // after the forward transform; for the packed column its k_m = 0 modes),
    // and field f of the column (psi, con from S; inn, inc from IN).  NT
    // divides L.
    constexpr int NTW = C::TWL ? LL / NT : 1;
    static_assert(LL % NT == 0, "whole twiddle rows per thread");
    Cx<R> tv[NTW];
    if constexpr (C::TWL) {
#pragma unroll
        for (int k2 = 0; k2 < NTW; k2++) tv[k2] = ((const Cx<R>*)b.twL)[tid + k2 * NT];
    }
    float4 cf[C::MP];
    float2 cc[C::MP];
    float klv[WANT_E ? C::MP : 1];  // kl of this thread's modes (energy terms)
    if constexpr (WANT_E) {
#pragma unroll
        for (int r = 0; r < C::MP; r++) klv[r] = b.kl[tid + NT * r];
    }
    {
        const Buf CA = mkbuf((const float4*)b.cA + bx, 0), CC = mkbuf((const float2*)b.cC + bx, 0);
#pragma unroll
        for (int r = 0; r < C::MP; r++) {
            const uint32_t e = (uint32_t)((tid + NT * r) * CS);
            cf[r] = bload<float4>(CA, e * 16u, 0u);
            cc[r] = bload<float2>(CC, e * 8u, 0u);
        }
    }
    Vec<R, 1> x[E];
    // PAIR: every thread loads pieces of S and IN in which lane pairs cover
    // one row's two fields (one request per row and array instead of two),
    // and the pieces reach their field groups through the planes.
    constexpr int NL = 4 * LL / NT;  // pieces per thread (2 arrays x L rows x 2 fields)
    if constexpr (C::PAIR) {
        const Buf s0 = mkbuf((const Vec<R, 2>*)b.S + tbase, 0), s1 = mkbuf((const Vec<R, 2>*)b.IN + tbase, 0);
        Cx<R> pc[NL];
#pragma unroll
        for (int u = 0; u < NL; u++) {
            const int gi = tid + NT * u, h = gi & 1, row = (gi >> 1) % LL;
            const uint32_t off = (uint32_t)row * rs + (uint32_t)h * (uint32_t)sizeof(Cx<R>);
            pc[u] = (gi >> 1) / LL ? bload<Cx<R>>(s1, off, 0u) : bload<Cx<R>>(s0, off, 0u);
        }
#pragma unroll
        for (int u = 0; u < NL; u++) {
            const int gi = tid + NT * u;
            pl[2 * ((gi >> 1) / LL) + (gi & 1)][pidx((gi >> 1) % LL)] = pc[u];
        }
    } else {
        const Buf src = mkbuf((const Vec<R, 2>*)(f < 2 ? b.S : b.IN) + tbase, 0);
#pragma unroll
        for (int q = 0; q < E; q++) x[q].c[0] = bload<Cx<R>>(src, (uint32_t)t * rs + fo, (uint32_t)(G * q) * rs);
    }
    if constexpr (C::TWL) {
#pragma unroll
        for (int k2 = 0; k2 < NTW; k2++) twS[tid + k2 * NT] = tv[k2];
    }
    if constexpr (WANT_E) {
        if (reduce_here) energy_reduce_block<NT>(b, p, inst, gridDim.x, red);
    }
    if (C::TWL || C::PAIR) __syncthreads();
    if constexpr (C::PAIR) {
#pragma unroll
        for (int q = 0; q < E; q++) x[q].c[0] = pl[f][pidx(t + G * q)];
    }
    const Cx<R>* twL = C::TWL ? twS : (const Cx<R>*)b.twL;

    // Forward FFT along l of field f.
    fft_reg<R, 1, LL, E, -1>(x, lds, t, twL);
    xsync<C::WAVE>();
#pragma unroll
    for (int q = 0; q < E; q++) pl[f][pidx(t + G * q)] = x[q].c[0];
    __syncthreads();

    // Modes i = tid + NT*r: semi-implicit update, DC removal, energy term,
    // C2A/C2B x psi'; the packed column splits each mode with its mirror -i.
    double g = 0.0;
    Cx<R> nv[C::MP][4];
    const Cx<R> z0{R(0), R(0)};
#pragma unroll
    for (int r = 0; r < C::MP; r++) {
        const int i = tid + NT * r, ii = pidx(i);
        Cx<R> ph = pl[0][ii], co = pl[1][ii], on = pl[2][ii], oc = pl[3][ii], a, bb;
        if (!packed) {
            // col_mode's update and C2 products from the tabulated coefficients
            // (the same float values: bit-identical).
            Vec<R, 2> sv, ov;
            sv.c[0] = ph; sv.c[1] = co;
            ov.c[0] = on; ov.c[1] = oc;
            if constexpr (sizeof(R) == 4) update_mode_coef(cf[r], sv, ov, z0);
            else update_mode_coef64(p, cf[r], sv, ov, z0);
            ph = sv.c[0];
            co = sv.c[1];
            if (i == 0 && kmi == 0 && kni == 0) ph = co = z0;
            if constexpr (WANT_E) g += mode_energy_kl(b, LL, i, klv[r], kmi, kni, co);
            a = Cx<R>{ph.x * (R)cc[r].x, ph.y * (R)cc[r].x};
            bb = Cx<R>{ph.x * (R)cc[r].y, ph.y * (R)cc[r].y};
        } else {
            // k_m = 0 from the column's own entries (cf, cc), k_m = M/2 from the table's tail
            const int im = pidx((LL - i) % LL);
            const size_t te = coef_packed_tail(b, j) + i;
            const float4 ch = ((const float4*)b.cA)[te];
            const float2 cch = ((const float2*)b.cC)[te];
            packed_modes_coef(b, p, WANT_E, LL, i, j, ph, co, on, oc, pl[0][im], pl[1][im], pl[2][im], pl[3][im],
                              cf[r], cc[r], ch, cch, a, bb, g);
        }
        nv[r][0] = ph;
        nv[r][1] = co;
        nv[r][2] = a;
        nv[r][3] = bb;
    }
    if (packed) __syncthreads();  // every mirror read before any write (uniform per workgroup)
#pragma unroll
    for (int r = 0; r < C::MP; r++) {
        const int ii = pidx(tid + NT * r);
#pragma unroll
        for (int k = 0; k < 4; k++) pl[k][ii] = nv[r][k];
    }
    if (b.tcount && blockIdx.x == 0 && tid == 0) b.tcount[inst] += 1;
    __syncthreads();

    // Inverse FFT of plane f: psi', con' -> S, C2A psi', C2B psi' -> AB.
#pragma unroll
    for (int q = 0; q < E; q++) x[q].c[0] = pl[f][pidx(t + G * q)];
    fft_reg<R, 1, LL, E, +1>(x, lds, t, twL);
    const R invL = R(1) / R(LL);
    if constexpr (C::PAIR) {
        xsync<C::WAVE>();
#pragma unroll
        for (int q = 0; q < E; q++) pl[f][pidx(t + G * q)] = Cx<R>{x[q].c[0].x * invL, x[q].c[0].y * invL};
        __syncthreads();
        const Buf d0 = mkbuf((Vec<R, 2>*)b.S + tbase, 0), d1 = mkbuf((Vec<R, 2>*)b.AB + tbase, 0);
#pragma unroll
        for (int u = 0; u < NL; u++) {
            const int gi = tid + NT * u, h = gi & 1, row = (gi >> 1) % LL, arr = (gi >> 1) / LL;
            const uint32_t off = (uint32_t)row * rs + (uint32_t)h * (uint32_t)sizeof(Cx<R>);
            const Cx<R> v = pl[2 * arr + h][pidx(row)];
            if (arr) bstore(v, d1, off, 0u);
            else bstore(v, d0, off, 0u);
        }
    } else {
        const Buf dst = mkbuf((Vec<R, 2>*)(f < 2 ? b.S : b.AB) + tbase, 0);
#pragma unroll
        for (int q = 0; q < E; q++)
            bstore(Cx<R>{x[q].c[0].x * invL, x[q].c[0].y * invL}, dst, (uint32_t)t * rs + fo, (uint32_t)(G * q) * rs);
    }
    if constexpr (WANT_E) {
        const double gs = block_sum<NT>(g, red);
        if (tid == 0) b.ecol[(size_t)inst * gridDim.x + blockIdx.x] = gs;
    }
}

// ---------------------------------------------------------------------------
// Host dispatch
// ---------------------------------------------------------------------------
template <typename R>
static hipError_t row_fs_t(const DevBufs& b, const PhysArgs& p, int flags, hipStream_t s, hipEvent_t t0,
                           hipEvent_t t1) {
    const dim3 grid(b.L * b.N, b.B);
    switch (b.M / 2) {
#define RF(H)                                                                                 \
    case H:                                                                                   \
        if constexpr (fs_row(sizeof(R) == 8 ? PREC_FP64 : PREC_FP32, H))                       \
            PFC_LAUNCH((k_row_fs<R, H>), grid, dim3(FsRow<R, H>::NT), s, t0, t1, b, p, flags);  \
        break;
        RF(16) RF(32) RF(64) RF(128) RF(256) RF(512)
#undef RF
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename R, bool WANT_E>
static hipError_t col_fs_t(const DevBufs& b, const PhysArgs& p, hipStream_t s, hipEvent_t t0, hipEvent_t t1) {
    const dim3 grid(b.N * b.K, b.B);
    switch (b.L) {
#define CF(LL)                                                                                 \
    case LL:                                                                                   \
        if constexpr (fs_col(sizeof(R) == 8 ? PREC_FP64 : PREC_FP32, LL))                        \
            PFC_LAUNCH((k_col_fs<R, LL, WANT_E>), grid, dim3(FsCol<R, LL>::NT), s, t0, t1, b, p); \
        break;
        CF(32) CF(64) CF(128) CF(256) CF(512) CF(1024)
#undef CF
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_row_fs(int prec, const DevBufs& b, const PhysArgs& p, int flags, hipStream_t s, hipEvent_t t0,
                         hipEvent_t t1) {
    if (!fs_row(prec, b.M / 2)) return hipErrorInvalidValue;
    return prec == PREC_FP64 ? row_fs_t<double>(b, p, flags, s, t0, t1) : row_fs_t<float>(b, p, flags, s, t0, t1);
}

hipError_t launch_col_fs(int prec, const DevBufs& b, const PhysArgs& p, bool want_e, hipStream_t s, hipEvent_t t0,
                         hipEvent_t t1) {
    if (!fs_col(prec, b.L)) return hipErrorInvalidValue;
    if (prec == PREC_FP64)
        return want_e ? col_fs_t<double, true>(b, p, s, t0, t1) : col_fs_t<double, false>(b, p, s, t0, t1);
    return want_e ? col_fs_t<float, true>(b, p, s, t0, t1) : col_fs_t<float, false>(b, p, s, t0, t1);
}

}  // namespace pfc
