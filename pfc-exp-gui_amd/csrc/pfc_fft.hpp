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

// cos and sin of 2 pi m / 12 (exact multiples of 30 degrees).
__device__ __forceinline__ constexpr double cos12(int m) {
    constexpr double t[12] = {1.0, 0.86602540378443864676, 0.5, 0.0, -0.5, -0.86602540378443864676,
                              -1.0, -0.86602540378443864676, -0.5, 0.0, 0.5, 0.86602540378443864676};
    return t[((m % 12) + 12) % 12];
}
__device__ __forceinline__ constexpr double sin12(int m) { return cos12(m - 3); }

// Radix R1*R2 = 12 or 6 in registers (Cooley-Tukey inside the butterfly):
// R2 radix-R1 transforms over n1 of x[R2*n1 + n2], the twiddles
// exp(DIR 2 pi i n2 k1 / R), then R1 radix-R2 transforms over n2; the
// output index is k1 + R1*k2.
template <int R1, int R2, int DIR, typename R, int NC>
__device__ __forceinline__ void bfly_split(Vec<R, NC>* v) {
    constexpr int RR = R1 * R2;
    Vec<R, NC> y[RR];  // y[n2*R1 + k1]
#pragma unroll
    for (int n2 = 0; n2 < R2; n2++) {
        Vec<R, NC> a[R1];
#pragma unroll
        for (int n1 = 0; n1 < R1; n1++) a[n1] = v[R2 * n1 + n2];
        if constexpr (R1 == 2) bfly2<DIR>(a);
        else if constexpr (R1 == 3) bfly3<DIR>(a);
        else bfly4<DIR>(a);
#pragma unroll
        for (int k1 = 0; k1 < R1; k1++) {
            const int m = (n2 * k1 * (12 / RR)) % 12;  // angle in units of 2 pi / 12
            if (m == 0) {
                y[n2 * R1 + k1] = a[k1];
            } else if (m == 3 || m == 9) {
                y[n2 * R1 + k1] = (m == 3) == (DIR > 0) ? vmuli<+1>(a[k1]) : vmuli<-1>(a[k1]);
            } else {
                y[n2 * R1 + k1] = vmulw(a[k1], Cx<R>{R(cos12(m)), R(DIR) * R(sin12(m))});
            }
        }
    }
#pragma unroll
    for (int k1 = 0; k1 < R1; k1++) {
        Vec<R, NC> b[R2];
#pragma unroll
        for (int n2 = 0; n2 < R2; n2++) b[n2] = y[n2 * R1 + k1];
        if constexpr (R2 == 2) bfly2<DIR>(b);
        else if constexpr (R2 == 3) bfly3<DIR>(b);
        else bfly4<DIR>(b);
#pragma unroll
        for (int k2 = 0; k2 < R2; k2++) v[k1 + R1 * k2] = b[k2];
    }
}

template <int RAD, int DIR, typename R, int NC>
__device__ __forceinline__ void bfly(Vec<R, NC>* v) {
    if constexpr (RAD == 2) bfly2<DIR>(v);
    else if constexpr (RAD == 3) bfly3<DIR>(v);
    else if constexpr (RAD == 4) bfly4<DIR>(v);
    else if constexpr (RAD == 6) bfly_split<2, 3, DIR>(v);
    else if constexpr (RAD == 8) bfly8<DIR>(v);
    else if constexpr (RAD == 12) bfly_split<4, 3, DIR>(v);
}

// Radix choice for the next pass: largest radix that divides both the
// remaining length and the per-thread element count.
#ifndef PFC_MAX_RADIX
#define PFC_MAX_RADIX 12
#endif
constexpr int pick_radix(int rem, int E) {
    return (PFC_MAX_RADIX >= 12 && rem % 12 == 0 && E % 12 == 0) ? 12
           : (rem % 8 == 0 && E % 8 == 0)                         ? 8
           : (PFC_MAX_RADIX >= 6 && rem % 6 == 0 && E % 6 == 0)   ? 6
           : (rem % 4 == 0 && E % 4 == 0)                         ? 4
           : (rem % 2 == 0 && E % 2 == 0)                         ? 2
           : (rem % 3 == 0 && E % 3 == 0)                         ? 3
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

// Exchange buffer of one transform.  CH = 1: one Vec per slot (AoS).
// CH > 1: the Vec is stored as CH chunks of 16 bytes in CH arrays PS slots
// apart (structure of arrays), so each LDS access moves 16 bytes and lanes
// that touch consecutive slots touch consecutive 16-byte bank groups.
struct alignas(16) Chunk16 {
    unsigned int w[4];
};

// Exchange barrier.  WAVE: every transform sharing the buffer lives in one
// wavefront, whose LDS accesses execute in order, so only the compiler must
// be kept from moving accesses across (wavefront-scope fences); otherwise a
// workgroup barrier.
template <bool WAVE>
__device__ __forceinline__ void xsync() {
    if constexpr (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}
template <typename R, int NC, int CH, int PS, bool WAVE = false>
struct LdsBuf {
    static constexpr bool SPLIT = false;
    static constexpr bool WSYNC = WAVE;
    char* p;
    static_assert(CH == 1 || sizeof(Vec<R, NC>) == 16 * CH, "chunks are 16 bytes");
    __device__ __forceinline__ Vec<R, NC> ld(int i) const {
        if constexpr (CH == 1) {
            return ((const Vec<R, NC>*)p)[i];
        } else {
            Vec<R, NC> v;
#pragma unroll
            for (int c = 0; c < CH; c++) {
                const Chunk16 k = ((const Chunk16*)p)[c * PS + i];
                __builtin_memcpy((char*)&v + 16 * c, &k, 16);
            }
            return v;
        }
    }
    __device__ __forceinline__ void st(int i, const Vec<R, NC>& v) const {
        if constexpr (CH == 1) {
            ((Vec<R, NC>*)p)[i] = v;
        } else {
#pragma unroll
            for (int c = 0; c < CH; c++) {
                Chunk16 k;
                __builtin_memcpy(&k, (const char*)&v + 16 * c, 16);
                ((Chunk16*)p)[c * PS + i] = k;
            }
        }
    }
};

// Exchange buffer holding one 16-byte chunk of each slot: an exchange moves
// the Vec chunk by chunk (store chunk c, barrier, load chunk c, barrier), so
// the buffer is 16 bytes per slot whatever the Vec size, at the price of one
// barrier pair per chunk.
template <typename R, int NC, bool WAVE = false>
struct LdsSplit {
    static constexpr bool SPLIT = true;
    static constexpr bool WSYNC = WAVE;
    static constexpr int NCH = (int)(sizeof(Vec<R, NC>) / 16);
    static_assert(sizeof(Vec<R, NC>) % 16 == 0, "chunks are 16 bytes");
    char* p;
    __device__ __forceinline__ void stc(int c, int i, const Vec<R, NC>& v) const {
        Chunk16 k;
        __builtin_memcpy(&k, (const char*)&v + 16 * c, 16);
        ((Chunk16*)p)[i] = k;
    }
    __device__ __forceinline__ void ldc(int c, int i, Vec<R, NC>& v) const {
        const Chunk16 k = ((const Chunk16*)p)[i];
        __builtin_memcpy((char*)&v + 16 * c, &k, 16);
    }
};

// Slots per chunk array for an SoA buffer of n padded slots: the first
// PS >= padded(n) with (CH * PS) % 16 == MOD, so that buffers placed one
// after another start MOD 16-byte groups apart.
constexpr int soa_stride(int n, int CH, int MOD) {
    int ps = padded(n);
    for (int k = 0; k < 16 && (CH * ps) % 16 != MOD; k++) ps++;
    return ps;
}

// Twiddles of the pass after one with stride NS and radix RAD, per
// butterfly b < NB2 and r = 1 .. RAD2-1: loaded before the exchange barrier,
// so their latency overlaps the exchange instead of following it.
template <typename R, int N, int NS2, int RAD2, int NB2, int GP2>
__device__ __forceinline__ void tw_fetch(Cx<R>* w, int t, const Cx<R>* __restrict__ tw) {
#pragma unroll
    for (int b = 0; b < NB2; b++) {
        const int k = (t + GP2 * b) % NS2;
#pragma unroll
        for (int r = 1; r < RAD2; r++) w[b * (RAD2 - 1) + (r - 1)] = tw[r * k * (N / (NS2 * RAD2))];
    }
}

// SRC: where the pass's inputs are: 0 = x in natural order (first pass),
// 1 = the LDS buffer, 2 = `in`, already in this pass's butterfly order (after
// a chunked exchange).  wp: this pass's twiddles (passes after the first).
template <typename R, int NC, int N, int E, int DIR, int NS, int SRC, typename L>
__device__ __forceinline__ void fft_reg_passes(Vec<R, NC> (&x)[E], Vec<R, NC> (&in)[E], const L& lds, int t,
                                               const Cx<R>* __restrict__ tw, const Cx<R>* wp) {
    if constexpr (NS < N) {
        constexpr int RAD = pick_radix(N / NS, E);
        static_assert(RAD > 0, "unsupported FFT length / elements-per-thread");
        constexpr int G = N / E, NB = E / RAD, STEP = N / RAD;
        constexpr bool LAST = (NS * RAD == N);
        Vec<R, NC> v[E];
        if constexpr (SRC == 0) {
#pragma unroll
            for (int b = 0; b < NB; b++)
#pragma unroll
                for (int r = 0; r < RAD; r++) v[b * RAD + r] = x[b + r * NB];
        } else if constexpr (SRC == 1) {
#pragma unroll
            for (int b = 0; b < NB; b++)
#pragma unroll
                for (int r = 0; r < RAD; r++) v[b * RAD + r] = lds.ld(pidx(t + G * b + r * STEP));
        } else {
#pragma unroll
            for (int q = 0; q < E; q++) v[q] = in[q];
        }
#pragma unroll
        for (int b = 0; b < NB; b++) {
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < RAD; r++) {
                    Cx<R> w = wp[b * (RAD - 1) + (r - 1)];
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
            constexpr int NS2 = NS * RAD;
            constexpr int RAD2 = pick_radix(N / NS2, E);
            constexpr int NB2 = E / RAD2, STEP2 = N / RAD2;
            Cx<R> wn[NB2 * (RAD2 - 1)];
            tw_fetch<R, N, NS2, RAD2, NB2, G>(wn, t, tw);
            if constexpr (L::SPLIT) {
                Vec<R, NC> w[E];
#pragma unroll
                for (int c = 0; c < L::NCH; c++) {
                    xsync<L::WSYNC>();
#pragma unroll
                    for (int b = 0; b < NB; b++) {
                        const int j = t + G * b;
                        const int idx = (j / NS) * NS * RAD + j % NS;
#pragma unroll
                        for (int r = 0; r < RAD; r++) lds.stc(c, pidx(idx + r * NS), v[b * RAD + r]);
                    }
                    xsync<L::WSYNC>();
#pragma unroll
                    for (int b = 0; b < NB2; b++)
#pragma unroll
                        for (int r = 0; r < RAD2; r++) lds.ldc(c, pidx(t + G * b + r * STEP2), w[b * RAD2 + r]);
                }
                fft_reg_passes<R, NC, N, E, DIR, NS2, 2, L>(x, w, lds, t, tw, wn);
            } else {
                xsync<L::WSYNC>();
#pragma unroll
                for (int b = 0; b < NB; b++) {
                    const int j = t + G * b;
                    const int idx = (j / NS) * NS * RAD + j % NS;
#pragma unroll
                    for (int r = 0; r < RAD; r++) lds.st(pidx(idx + r * NS), v[b * RAD + r]);
                }
                xsync<L::WSYNC>();
                fft_reg_passes<R, NC, N, E, DIR, NS2, 1, L>(x, in, lds, t, tw, wn);
            }
        }
    }
}

template <typename R, int NC, int N, int E, int DIR, typename L>
__device__ __forceinline__ void fft_reg(Vec<R, NC> (&x)[E], const L& lds, int t, const Cx<R>* __restrict__ tw) {
    static_assert(N % E == 0, "length must be a multiple of elements per thread");
    static_assert(plan_ok(N, E), "no radix plan");
    fft_reg_passes<R, NC, N, E, DIR, 1, 0, L>(x, x, lds, t, tw, nullptr);
}

// ---------------------------------------------------------------------------
// Lengths N = 3*2^k with per-pass thread counts.
//
// Plan (forward order): radix 6, radix 8 while it divides, then 4 or 2; the
// reversed plan runs the same passes backwards (a Stockham transform may take
// its radices in any order).  A group of G = N/6 threads runs the transform;
// pass p uses GP = the largest divisor of N/R_p not above G threads, each
// doing N/(R_p*GP) butterflies, the other threads idle in that pass.  So a
// thread never holds more than 8 elements (E = 12 with a fixed thread count
// needs 12 and falls to one wave per SIMD).  Thread t enters the first pass
// with elements t + GP0*q (q < N/GP0) and leaves the last with t + GPL*q.
// ---------------------------------------------------------------------------
// Plans: 0 = radix 6, then 8s (the one above); 1 = radix 3, then 4s over
// N/3 threads (at most 4 elements per thread: half the registers of plan 0,
// one more exchange).  The row and column configurations choose per length
// and precision.
constexpr int gen_r0(int plan) { return plan ? 3 : 6; }  // first radix
constexpr int gen_rb(int plan) { return plan ? 4 : 8; }  // body radix
constexpr int gen_g(int n, int plan) { return n / gen_r0(plan); }  // threads of one transform
constexpr int gen_npass(int n, int plan) {
    int r = n / gen_r0(plan), c = 1;
    while (r % gen_rb(plan) == 0) { r /= gen_rb(plan); c++; }
    return r > 1 ? c + 1 : c;
}
constexpr int gen_radix_fwd(int n, int p, int plan) {
    if (p == 0) return gen_r0(plan);
    int r = n / gen_r0(plan);
    for (int i = 1; i < p; i++) r /= gen_rb(plan);
    return r % gen_rb(plan) == 0 ? gen_rb(plan) : r;
}
constexpr int gen_radix(int n, int p, bool rev, int plan) {
    return rev ? gen_radix_fwd(n, gen_npass(n, plan) - 1 - p, plan) : gen_radix_fwd(n, p, plan);
}
constexpr int gen_threads(int n, int rad, int g) {
    const int m = n / rad;
    int t = g < m ? g : m;
    while (m % t) t--;
    return t;
}
// Elements per thread entering (first) or leaving (last) a plan.
constexpr int gen_e_in(int n, bool rev, int plan) {
    return n / gen_threads(n, gen_radix(n, 0, rev, plan), gen_g(n, plan));
}
constexpr int gen_g_in(int n, bool rev, int plan) { return gen_threads(n, gen_radix(n, 0, rev, plan), gen_g(n, plan)); }
constexpr int gen_emax(int n, int plan) {
    int e = 0;
    for (int p = 0; p < gen_npass(n, plan); p++) {
        const int x = n / gen_threads(n, gen_radix(n, p, false, plan), gen_g(n, plan));
        e = x > e ? x : e;
    }
    return e;
}

template <typename R, int NC, int N, int DIR, bool REV, int PI, int NS, int EM, int PLAN, typename L>
__device__ __forceinline__ void fft_gen_passes(Vec<R, NC> (&x)[EM], const L& lds, int t,
                                               const Cx<R>* __restrict__ tw, const Cx<R>* wp) {
    if constexpr (NS < N) {
        constexpr int G = gen_g(N, PLAN);
        constexpr int RAD = gen_radix(N, PI, REV, PLAN);
        constexpr int GP = gen_threads(N, RAD, G);
        constexpr int NB = N / (RAD * GP), STEP = N / RAD;
        constexpr bool LAST = (NS * RAD == N);
        static_assert(RAD * NB <= EM, "register array too small for the plan");
        const bool act = GP == G || t < GP;
        Vec<R, NC> v[RAD * NB];
        if (act) {
            if constexpr (PI == 0) {
#pragma unroll
                for (int b = 0; b < NB; b++)
#pragma unroll
                    for (int r = 0; r < RAD; r++) v[b * RAD + r] = x[b + r * NB];
            } else {
#pragma unroll
                for (int b = 0; b < NB; b++)
#pragma unroll
                    for (int r = 0; r < RAD; r++) v[b * RAD + r] = lds.ld(pidx(t + GP * b + r * STEP));
            }
#pragma unroll
            for (int b = 0; b < NB; b++) {
                if constexpr (NS > 1) {
#pragma unroll
                    for (int r = 1; r < RAD; r++) {
                        Cx<R> w = wp[b * (RAD - 1) + (r - 1)];
                        if constexpr (DIR > 0) w.y = -w.y;
                        v[b * RAD + r] = vmulw(v[b * RAD + r], w);
                    }
                }
                bfly<RAD, DIR>(&v[b * RAD]);
            }
        }
        if constexpr (LAST) {
            if (act) {
#pragma unroll
                for (int b = 0; b < NB; b++)
#pragma unroll
                    for (int r = 0; r < RAD; r++) x[b + r * NB] = v[b * RAD + r];
            }
        } else {
            constexpr int NS2 = NS * RAD;
            constexpr int RAD2 = gen_radix(N, PI + 1, REV, PLAN);
            constexpr int GP2 = gen_threads(N, RAD2, G);
            constexpr int NB2 = N / (RAD2 * GP2);
            Cx<R> wn[NB2 * (RAD2 - 1)];
            if (GP2 == G || t < GP2) tw_fetch<R, N, NS2, RAD2, NB2, GP2>(wn, t, tw);
            xsync<L::WSYNC>();
            if (act) {
#pragma unroll
                for (int b = 0; b < NB; b++) {
                    const int j = t + GP * b;
                    const int idx = (j / NS) * NS * RAD + j % NS;
#pragma unroll
                    for (int r = 0; r < RAD; r++) lds.st(pidx(idx + r * NS), v[b * RAD + r]);
                }
            }
            xsync<L::WSYNC>();
            fft_gen_passes<R, NC, N, DIR, REV, PI + 1, NS2, EM, PLAN, L>(x, lds, t, tw, wn);
        }
    }
}

// A transform between the two element mappings of a length: side A (the
// loaded/stored side) and side B (the other).  Powers of two: fft_reg, the
// same mapping on both sides (E elements, G = N/E threads).  N = 3*2^k: the
// general plan, A -> B in forward order, B -> A reversed.
template <typename R, int NC, int N, int EM, int DIR, bool A2B, bool GEN, int PLAN, typename L>
__device__ __forceinline__ void fft_ab(Vec<R, NC> (&x)[EM], const L& lds, int t, const Cx<R>* __restrict__ tw) {
    if constexpr (GEN) {
        fft_gen_passes<R, NC, N, DIR, !A2B, 0, 1, EM, PLAN, L>(x, lds, t, tw, nullptr);
    } else {
        fft_reg<R, NC, N, EM, DIR>(x, lds, t, tw);
    }
}

// Mappings of a length N on its two sides: thread counts and elements.
// GEN: the general plan (N = 3*2^k); otherwise fft_reg with EPOW2 elements.
template <int N, int EPOW2, bool USEGEN, int PLAN>
struct Sides {
    static constexpr bool GEN = USEGEN && N % 3 == 0;
    static constexpr int EM = GEN ? gen_emax(N, PLAN) : EPOW2;  // register elements
    static constexpr int G = GEN ? gen_g(N, PLAN) : N / EPOW2;  // threads per transform
    static constexpr int GA = GEN ? gen_g_in(N, false, PLAN) : G, EA = GEN ? gen_e_in(N, false, PLAN) : EPOW2;
    static constexpr int GB = GEN ? gen_g_in(N, true, PLAN) : G, EB = GEN ? gen_e_in(N, true, PLAN) : EPOW2;
};

// AoS exchange buffer (one Vec per slot).
template <typename R, int NC, int N, int E, int DIR>
__device__ __forceinline__ void fft_reg(Vec<R, NC> (&x)[E], Vec<R, NC>* lds, int t, const Cx<R>* __restrict__ tw) {
    fft_reg<R, NC, N, E, DIR>(x, LdsBuf<R, NC, 1, 0>{(char*)lds}, t, tw);
}

}  // namespace pfc
