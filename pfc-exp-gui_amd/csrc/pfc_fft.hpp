// pfc_fft.hpp - LDS-resident Stockham FFT building blocks for gfx950.
//
// Replaces the FFTW c2c plans of the reference (misfit.cpp:438-452,
// cvd.cpp:334-348).  Every transform in the step is a batch of 1-D transforms
// executed inside one workgroup: the data sits in LDS, each thread owns E
// elements, and each radix pass reads its butterfly inputs from LDS, applies
// twiddles, runs the radix-R butterfly in registers and writes the outputs back
// in Stockham (auto-sorted) order.  Lengths are compile-time so every loop
// unrolls and all per-thread data lives in VGPRs.
//
// An element is a Vec: NC complex values that share one index (the two fields
// psi/con, or the two convolutions a/b), so one pass of index arithmetic and
// twiddle loads serves both fields, and global memory moves 16 B (fp32) or
// 32 B (fp64) per element.
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

// Elements per thread used for a transform of length n.
constexpr int elems_for(int n) { return (n % 3 == 0) ? 12 : (n >= 8 ? 8 : n); }

// One Stockham pass, in place in LDS (read all, barrier, write all, barrier).
// `lds` points at element 0 of this thread's transform; consecutive elements
// are STRIDE Vec apart.  tw[t] = exp(-2 pi i t / NLEN).  Barriers are taken
// unconditionally; `active` only guards the loads and stores, so threads that
// own no element of this transform never diverge around a barrier.
template <typename R, int NC, int NLEN, int E, int DIR, int STRIDE, int NS, int RAD>
__device__ __forceinline__ void fft_pass(Vec<R, NC>* lds, int t, bool active, const Cx<R>* __restrict__ tw) {
    constexpr int G = NLEN / E;
    constexpr int NB = E / RAD;
    constexpr int STEP = NLEN / RAD;
    Vec<R, NC> x[E];
    if (active) {
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const int j = t + G * b;
#pragma unroll
            for (int r = 0; r < RAD; r++) x[b * RAD + r] = lds[(j + r * STEP) * STRIDE];
        }
    }
    __syncthreads();
    if (active) {
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const int j = t + G * b;
            const int k = j % NS;
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < RAD; r++) {
                    Cx<R> w = tw[r * k * (NLEN / (NS * RAD))];
                    if constexpr (DIR > 0) w.y = -w.y;
                    x[b * RAD + r] = vmulw(x[b * RAD + r], w);
                }
            }
            bfly<RAD, DIR>(&x[b * RAD]);
            const int idx = (j / NS) * NS * RAD + k;
#pragma unroll
            for (int r = 0; r < RAD; r++) lds[(idx + r * NS) * STRIDE] = x[b * RAD + r];
        }
    }
    __syncthreads();
}

template <typename R, int NC, int NLEN, int E, int DIR, int STRIDE, int NS>
__device__ __forceinline__ void fft_passes(Vec<R, NC>* lds, int t, bool active, const Cx<R>* __restrict__ tw) {
    if constexpr (NS < NLEN) {
        constexpr int RAD = pick_radix(NLEN / NS, E);
        static_assert(RAD > 0, "unsupported FFT length / elements-per-thread");
        fft_pass<R, NC, NLEN, E, DIR, STRIDE, NS, RAD>(lds, t, active, tw);
        fft_passes<R, NC, NLEN, E, DIR, STRIDE, NS * RAD>(lds, t, active, tw);
    }
}

// Unnormalised length-NLEN transform of the data in LDS.  Every thread of the
// workgroup must call it (it contains barriers).  The caller must barrier
// before the call; the call ends with a barrier, so its results can be read
// right away.
template <typename R, int NC, int NLEN, int DIR, int STRIDE>
__device__ __forceinline__ void lds_fft(Vec<R, NC>* lds, int t, bool active, const Cx<R>* __restrict__ tw) {
    constexpr int E = elems_for(NLEN);
    static_assert(NLEN % E == 0, "length must be a multiple of elements per thread");
    static_assert(plan_ok(NLEN, E), "no radix plan");
    fft_passes<R, NC, NLEN, E, DIR, STRIDE, 1>(lds, t, active, tw);
}

}  // namespace pfc
