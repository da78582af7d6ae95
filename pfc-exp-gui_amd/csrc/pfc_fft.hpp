// pfc_fft.hpp - register/LDS Stockham FFT building blocks for gfx950.
//
// Replaces the FFTW c2c plans of the reference (misfit.cpp:438-452,
// cvd.cpp:334-348).  Every transform in the step is a batch of 1-D transforms
// executed inside one workgroup: each thread owns E elements in VGPRs, each
// radix pass runs the radix-R butterfly in registers, and consecutive passes
// exchange data through LDS in Stockham (auto-sorted) order.  Lengths are
// compile-time so every loop unrolls and all per-thread data lives in VGPRs.
//
// An element is a Vec: NC complex values that share one index (the fields
// psi/con, the convolutions a/b, or all four), so one pass of index arithmetic
// and twiddle loads serves every field, and global memory moves 16 B (fp32)
// or 32 B (fp64) per pair of fields.
#pragma once
#include <hip/hip_runtime.h>

namespace pfc {

template <typename R>
struct Cx {
    R x, y;
};

template <typename R, int NC>
struct alignas(2 * NC * sizeof(R) >= 16 ? 16 : 2 * NC * sizeof(R)) Vec {
    Cx<R> c[NC];
};

template <typename R>
__device__ __forceinline__ Cx<R> cmul(Cx<R> a, Cx<R> w) {
    return Cx<R>{a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x};
}

template <typename R, int NC>
__device__ __forceinline__ Vec<R, NC> vadd(const Vec<R, NC>& a, const Vec<R, NC>& b) {
    Vec<R, NC> r;
#pragma unroll
    for (int i = 0; i < NC; i++) r.c[i] = Cx<R>{a.c[i].x + b.c[i].x, a.c[i].y + b.c[i].y};
    return r;
}
template <typename R, int NC>
__device__ __forceinline__ Vec<R, NC> vsub(const Vec<R, NC>& a, const Vec<R, NC>& b) {
    Vec<R, NC> r;
#pragma unroll
    for (int i = 0; i < NC; i++) r.c[i] = Cx<R>{a.c[i].x - b.c[i].x, a.c[i].y - b.c[i].y};
    return r;
}
template <typename R, int NC>
__device__ __forceinline__ Vec<R, NC> vmulw(const Vec<R, NC>& a, Cx<R> w) {
    Vec<R, NC> r;
#pragma unroll
    for (int i = 0; i < NC; i++) r.c[i] = cmul(a.c[i], w);
    return r;
}
// Multiply by s*i, s = +1 or -1.
template <int S, typename R, int NC>
__device__ __forceinline__ Vec<R, NC> vmuli(const Vec<R, NC>& a) {
    Vec<R, NC> r;
#pragma unroll
    for (int i = 0; i < NC; i++) r.c[i] = (S > 0) ? Cx<R>{-a.c[i].y, a.c[i].x} : Cx<R>{a.c[i].y, -a.c[i].x};
    return r;
}
template <typename R, int NC>
__device__ __forceinline__ Vec<R, NC> vscale(const Vec<R, NC>& a, R s) {
    Vec<R, NC> r;
#pragma unroll
    for (int i = 0; i < NC; i++) r.c[i] = Cx<R>{a.c[i].x * s, a.c[i].y * s};
    return r;
}
template <typename R, int NC>
__device__ __forceinline__ Vec<R, NC> vzero() {
    Vec<R, NC> r;
#pragma unroll
    for (int i = 0; i < NC; i++) r.c[i] = Cx<R>{R(0), R(0)};
    return r;
}

// ---------------------------------------------------------------------------
// Radix butterflies.  DIR = -1: forward (exp(-2 pi i ...)), +1: inverse.
// ---------------------------------------------------------------------------
template <int DIR, typename R, int NC>
__device__ __forceinline__ void bfly2(Vec<R, NC>* v) {
    Vec<R, NC> a = v[0], b = v[1];
    v[0] = vadd(a, b);
    v[1] = vsub(a, b);
}

template <int DIR, typename R, int NC>
__device__ __forceinline__ void bfly4(Vec<R, NC>* v) {
    Vec<R, NC> s0 = vadd(v[0], v[2]), d0 = vsub(v[0], v[2]);
    Vec<R, NC> s1 = vadd(v[1], v[3]), d1 = vsub(v[1], v[3]);
    Vec<R, NC> jd1 = vmuli<DIR>(d1);
    v[0] = vadd(s0, s1);
    v[2] = vsub(s0, s1);
    v[1] = vadd(d0, jd1);
    v[3] = vsub(d0, jd1);
}

template <int DIR, typename R, int NC>
__device__ __forceinline__ void bfly3(Vec<R, NC>* v) {
    const R c = R(-0.5);
    const R sn = R(DIR) * R(0.86602540378443864676372317075294);
    Vec<R, NC> s = vadd(v[1], v[2]), d = vsub(v[1], v[2]);
    Vec<R, NC> a = v[0];
    v[0] = vadd(a, s);
    Vec<R, NC> m = vadd(a, vscale(s, c));
    Vec<R, NC> jd;
#pragma unroll
    for (int i = 0; i < NC; i++) jd.c[i] = Cx<R>{-d.c[i].y * sn, d.c[i].x * sn};
    v[1] = vadd(m, jd);
    v[2] = vsub(m, jd);
}

template <int DIR, typename R, int NC>
__device__ __forceinline__ void bfly8(Vec<R, NC>* v) {
    Vec<R, NC> e[4] = {v[0], v[2], v[4], v[6]};
    Vec<R, NC> o[4] = {v[1], v[3], v[5], v[7]};
    bfly4<DIR>(e);
    bfly4<DIR>(o);
    const R h = R(0.70710678118654752440084436210485);
    // W8^k = exp(DIR*2*pi*i*k/8)
    o[1] = vmulw(o[1], Cx<R>{h, R(DIR) * h});
    o[2] = vmuli<DIR>(o[2]);
    o[3] = vmulw(o[3], Cx<R>{-h, R(DIR) * h});
#pragma unroll
    for (int k = 0; k < 4; k++) {
        v[k] = vadd(e[k], o[k]);
        v[k + 4] = vsub(e[k], o[k]);
    }
}

template <int RAD, int DIR, typename R, int NC>
__device__ __forceinline__ void bfly(Vec<R, NC>* v) {
    if constexpr (RAD == 2) bfly2<DIR>(v);
    else if constexpr (RAD == 3) bfly3<DIR>(v);
    else if constexpr (RAD == 4) bfly4<DIR>(v);
    else if constexpr (RAD == 8) bfly8<DIR>(v);
}

// Radix choice for the next pass: largest radix that divides both the
// remaining length and the per-thread element count.
constexpr int pick_radix(int rem, int E) {
    return (rem % 8 == 0 && E % 8 == 0)   ? 8
           : (rem % 4 == 0 && E % 4 == 0) ? 4
           : (rem % 2 == 0 && E % 2 == 0) ? 2
           : (rem % 3 == 0 && E % 3 == 0) ? 3
                                          : -1;
}

constexpr bool plan_ok(int n, int E, int ns = 1) {
    return ns == n ? true : (pick_radix(n / ns, E) > 0 && plan_ok(n, E, ns * pick_radix(n / ns, E)));
}

// Elements per thread for a transform of length n: `pow2` for powers of two
// (radix-8/4/2 passes), 12 for 3*2^k (radix-4/3/2 passes).
constexpr int elems_for(int n, int pow2) { return (n % 3 == 0) ? 12 : (n >= pow2 ? pow2 : n); }

// ---------------------------------------------------------------------------
// Register-resident Stockham transform.
//
// Thread t of a transform of length N owns elements t + G*q (q < E, G = N/E)
// in x[q] on entry and again on exit (natural order, unnormalised).  The first
// radix pass takes its inputs straight from x and the last pass writes its
// outputs straight back to x; only the passes in between exchange data
// through LDS.  The LDS buffer of one transform holds padded(N) elements with
// one pad element after every 8 (pidx), which makes the radix-8 Stockham
// scatter (stride-8 writes) conflict-free.  Every thread of the workgroup
// must call (the exchanges contain barriers); each exchange starts with a
// barrier, so the caller needs none before the call, and the call ends
// without LDS reads in flight only after the caller's next barrier.
// ---------------------------------------------------------------------------
constexpr int padded(int n) { return n + n / 8 + 1; }
__device__ __forceinline__ int pidx(int i) { return i + (i >> 3); }

template <typename R, int NC, int N, int E, int DIR, int NS, bool FIRST>
__device__ __forceinline__ void fft_reg_passes(Vec<R, NC> (&x)[E], Vec<R, NC>* lds, int t,
                                               const Cx<R>* __restrict__ tw) {
    if constexpr (NS < N) {
        constexpr int RAD = pick_radix(N / NS, E);
        static_assert(RAD > 0, "unsupported FFT length / elements-per-thread");
        constexpr int G = N / E, NB = E / RAD, STEP = N / RAD;
        constexpr bool LAST = (NS * RAD == N);
        Vec<R, NC> v[E];
        if constexpr (FIRST) {
#pragma unroll
            for (int b = 0; b < NB; b++)
#pragma unroll
                for (int r = 0; r < RAD; r++) v[b * RAD + r] = x[b + r * NB];
        } else {
#pragma unroll
            for (int b = 0; b < NB; b++)
#pragma unroll
                for (int r = 0; r < RAD; r++) v[b * RAD + r] = lds[pidx(t + G * b + r * STEP)];
        }
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const int j = t + G * b;
            const int k = j % NS;
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < RAD; r++) {
                    Cx<R> w = tw[r * k * (N / (NS * RAD))];
                    if constexpr (DIR > 0) w.y = -w.y;
                    v[b * RAD + r] = vmulw(v[b * RAD + r], w);
                }
            }
            bfly<RAD, DIR>(&v[b * RAD]);
        }
        if constexpr (LAST) {
#pragma unroll
            for (int b = 0; b < NB; b++)
#pragma unroll
                for (int r = 0; r < RAD; r++) x[b + r * NB] = v[b * RAD + r];
        } else {
            __syncthreads();
#pragma unroll
            for (int b = 0; b < NB; b++) {
                const int j = t + G * b;
                const int idx = (j / NS) * NS * RAD + j % NS;
#pragma unroll
                for (int r = 0; r < RAD; r++) lds[pidx(idx + r * NS)] = v[b * RAD + r];
            }
            __syncthreads();
            fft_reg_passes<R, NC, N, E, DIR, NS * RAD, false>(x, lds, t, tw);
        }
    }
}

template <typename R, int NC, int N, int E, int DIR>
__device__ __forceinline__ void fft_reg(Vec<R, NC> (&x)[E], Vec<R, NC>* lds, int t, const Cx<R>* __restrict__ tw) {
    static_assert(N % E == 0, "length must be a multiple of elements per thread");
    static_assert(plan_ok(N, E), "no radix plan");
    fft_reg_passes<R, NC, N, E, DIR, 1, true>(x, lds, t, tw);
}

}  // namespace pfc
