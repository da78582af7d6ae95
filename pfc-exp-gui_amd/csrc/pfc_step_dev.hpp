// pfc_step_dev.hpp - device helpers shared by the spectral step kernels
// (pfc_step.hip: row/column/depth passes; pfc_fs.hip: the field-split passes):
// buffer addressing, the pointwise physics (misfit.cpp:773-874, 983-1014,
// 1043-1092), the row real-to-complex helpers, the per-mode column work and
// the deterministic workgroup sums.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "pfc_fft.hpp"
#include "pfc_kernels.h"

namespace pfc {

// Launch; with a start/stop event pair the events take the kernel's own
// start and completion times (hipExtLaunchKernelGGL).
#define PFC_LAUNCH(kern, grid, block, s, t0, t1, ...)                                  \
    do {                                                                               \
        if (t0) hipExtLaunchKernelGGL(kern, grid, block, 0, s, t0, t1, 0, __VA_ARGS__); \
        else hipLaunchKernelGGL(kern, grid, block, 0, s, __VA_ARGS__);                 \
    } while (0)

// Accesses to workgroup-uniform arrays: a uniform base pointer plus a 32-bit
// per-thread byte offset and a 32-bit uniform byte offset (per unrolled
// element), so no 64-bit index arithmetic is done per element.
struct Buf {
    char* p;
};
__device__ __forceinline__ Buf mkbuf(const void* base, size_t) { return Buf{(char*)const_cast<void*>(base)}; }
template <typename T>
__device__ __forceinline__ T bload(const Buf& b, uint32_t voff, uint32_t soff) {
    return *(const T*)(b.p + soff + voff);
}
template <typename T>
__device__ __forceinline__ void bstore(const T& val, const Buf& b, uint32_t voff, uint32_t soff) {
    *(T*)(b.p + soff + voff) = val;
}

// ---------------------------------------------------------------------------
// Pointwise physics
// ---------------------------------------------------------------------------

// Nonlinear terms at one point (misfit.cpp:777, 792, 826, 840, 858-867).
// Returns inn, inc, in3 and the mixing term in2 = w (c ln(c/c0) + (1-c) ln((1-c)/(1-c0))).
__device__ __forceinline__ void nonlinear(const PhysArgs& p, double ph, double c, double a, double b,
                                          double& inn, double& inc, double& in3, double& in2) {
#pragma clang fp contract(off)
    const double c0 = (double)p.c0;
    const double lc = log(c / c0);
    const double lc1 = log((1.0 - c) / (1.0 - c0));
    const double ph2 = ph * ph;
    const double in1 = -(double)p.niu * ph2 / 2.0 + (double)p.ksi * (ph2 * ph) / 3.0;
    in2 = (double)p.w * (c * lc + (1.0 - c) * lc1);
    const double cc = c * c;
    const double cm = 1.0 - c;
    const double h1 = 1.0 - 3.0 * cc + 2.0 * (cc * c);
    const double h2 = 1.0 - 3.0 * (cm * cm) + 2.0 * (cm * cm * cm);
    in3 = -h1 * a - h2 * b;
    inn = in1 + in2 + in3;
    const double wp = (double)p.w * (ph + 1.0);
    const double t1 = wp * lc;
    const double t2 = wp * lc1;
    const double t3 = 3.0 * ph * (a - b) * cc;
    const double t4 = 3.0 * ph * (-1.0 * a + b) * c;
    if (p.narrow)
        inc = (double)((((float)t1 - (float)t2) - (float)t3) - (float)t4);
    else
        inc = t1 - t2 - t3 - t4;
}

// FP32: log(c/c0) = ln2 * log2(c) - ln(c0) with the hardware log2
// (v_log_f32, about 1 ulp) and ln(c0), ln(1 - c0) precomputed on the host.
__device__ __forceinline__ void nonlinear(const PhysArgs& p, float ph, float c, float a, float b, float& inn,
                                          float& inc, float& in3, float& in2) {
    constexpr float LN2 = 0.693147180559945309f;
    const float lc = __builtin_fmaf(LN2, __builtin_amdgcn_logf(c), -p.lnc0);
    const float lc1 = __builtin_fmaf(LN2, __builtin_amdgcn_logf(1.0f - c), -p.ln1c0);
    const float ph2 = ph * ph;
    const float in1 = -p.niu * ph2 * 0.5f + p.ksi * (ph2 * ph) * (1.0f / 3.0f);
    in2 = p.w * (c * lc + (1.0f - c) * lc1);
    const float cc = c * c;
    const float cm = 1.0f - c;
    const float h1 = 1.0f - 3.0f * cc + 2.0f * (cc * c);
    const float h2 = 1.0f - 3.0f * (cm * cm) + 2.0f * (cm * cm * cm);
    in3 = -h1 * a - h2 * b;
    inn = in1 + in2 + in3;
    const float wp = p.w * (ph + 1.0f);
    inc = wp * lc - wp * lc1 - 3.0f * ph * (a - b) * cc - 3.0f * ph * (b - a) * c;
}

// Per-point energy without the gradient term (ENERGYC2, misfit.cpp:1054-1058);
// the alpha*grac term is summed in k space by the column pass (Parseval).
// FP64 (and the reference's float narrowing of the three terms).
__device__ __forceinline__ double energy_point(const PhysArgs& p, double phr, double cr, double in3r, double) {
#pragma clang fp contract(off)
    const double ph = (double)phr, c = (double)cr, in3 = (double)in3r;
    const double c0 = (double)p.c0;
    const double ph2 = ph * ph;
    const double eid = ph2 / 2.0 - (double)p.niu * (ph2 * ph) / 6.0 + (double)p.ksi * (ph2 * ph2) / 12.0;
    const double emix = (ph + 1.0) * (double)p.w * (c * log(c / c0) + (1.0 - c) * log((1.0 - c) / (1.0 - c0)));
    const double eex = 0.5 * ph * in3;
    if (p.narrow) return (double)((float)eid + (float)emix + (float)eex);
    return eid + emix + eex;
}

// FP32 path: the three terms in float, the mixing term as (ph + 1) * in2
// with in2 = w (c ln(c/c0) + (1-c) ln((1-c)/(1-c0))) from the nonlinear terms
// (the same logarithms); the caller sums in FP64.
__device__ __forceinline__ double energy_point(const PhysArgs& p, float ph, float, float in3, float in2) {
    const float ph2 = ph * ph;
    const float eid = ph2 * 0.5f - p.niu * (ph2 * ph) * (1.0f / 6.0f) + p.ksi * (ph2 * ph2) * (1.0f / 12.0f);
    const float emix = (ph + 1.0f) * in2;
    const float eex = 0.5f * ph * in3;
    return (double)(eid + emix + eex);
}

// Noise term of MAINEQUATIONC2_NOI for one mode (misfit.cpp:1002-1003):
// dt*b0*noisf*sqrt(ks)*sqrt(|0.5 +- 0.5 sel|)*(0.5 - 0.5 tanh((sqrt(ks) - 2 pi)/0.1)),
// sel = 1 on the four self-conjugate bins (l in {0, L/2}, m in {0, M/2}).
__device__ __forceinline__ Cx<double> noise_term(const PhysArgs& p, float ks, bool sel, Cx<double> nz) {
#pragma clang fp contract(off)
    const float sks = sqrtf(ks);
    const double damp = 0.5 - 0.5 * tanh(((double)sks - 2.0 * 3.1415926) / 0.1);
    const double s = sel ? 1.0 : 0.0;
    const double a = (double)p.db * nz.x * (double)sks * sqrt(fabs(0.5 + 0.5 * s)) * damp;
    const double b = (double)p.db * nz.y * (double)sks * sqrt(fabs(0.5 - 0.5 * s)) * damp;
    return Cx<double>{a, b};
}

__device__ __forceinline__ Cx<float> noise_term(const PhysArgs& p, float ks, bool sel, Cx<float> nz) {
    const float sks = sqrtf(ks);
    const float damp = 0.5f - 0.5f * tanhf((sks - 6.2831852f) * 10.0f);
    const float ca = sel ? 1.0f : 0.70710678f, cb = sel ? 0.0f : 0.70710678f;
    return Cx<float>{p.db * nz.x * sks * ca * damp, p.db * nz.y * sks * cb * damp};
}

// FP64 coefficients of one mode as the reference forms them in float:
// {mk, mkc, den, 0} = {mn dt k^2, mc dt k^2, mc dt alpha k^4, 0}.
__device__ __forceinline__ float4 mode_coef64(const PhysArgs& p, float ks) {
#pragma clang fp contract(off)
    const float mk = p.mn_dt * ks;
    const float mkc = p.mc_dt * ks;
    const float den = (p.mc_dt_a * ks) * ks;
    return make_float4(mk, mkc, den, 0.0f);
}

// Semi-implicit update of one mode (misfit.cpp:1002-1010, cvd.cpp:872-880)
// from its coefficients c = mode_coef64(ks); nt is the noise term (zero
// when b0 = 0).
__device__ __forceinline__ void update_mode_coef64(const PhysArgs& p, float4 c, Vec<double, 2>& s,
                                                   const Vec<double, 2>& o, Cx<double> nt) {
#pragma clang fp contract(off)
    const float mk = c.x, mkc = c.y, den = c.z;
    double r1 = s.c[0].x - (double)mk * o.c[0].x + nt.x;
    double r2 = s.c[0].y - (double)mk * o.c[0].y + nt.y;
    double r3 = s.c[1].x - (double)mkc * o.c[1].x;
    double r4 = s.c[1].y - (double)mkc * o.c[1].y;
    if (p.narrow) {
        r1 = (double)(float)r1; r2 = (double)(float)r2;
        r3 = (double)(float)r3; r4 = (double)(float)r4;
    }
    const double d1 = 1.0 + (double)mk;
    const double d2 = 1.0 + (double)den;
    s.c[0].x = r1 / d1; s.c[0].y = r2 / d1;
    s.c[1].x = r3 / d2; s.c[1].y = r4 / d2;
}

__device__ __forceinline__ void update_mode(const PhysArgs& p, float ks, Vec<double, 2>& s,
                                            const Vec<double, 2>& o, Cx<double> nt) {
    update_mode_coef64(p, mode_coef64(p, ks), s, o, nt);
}

// FP32 coefficients of one mode: {mk, mkc, 1/(1 + mk), 1/(1 + den)}.  The
// FP32 arithmetic is spelled out (explicit FMAs, no implicit contraction) so
// that every path computing it gives the same bits.
__device__ __forceinline__ float4 mode_coef(const PhysArgs& p, float ks) {
#pragma clang fp contract(off)
    const float mk = p.mn_dt * ks;
    const float mkc = p.mc_dt * ks;
    const float den = (p.mc_dt_a * ks) * ks;
    return make_float4(mk, mkc, 1.0f / (1.0f + mk), 1.0f / (1.0f + den));
}

// The FP32 update from a mode's coefficients c = mode_coef(ks).
__device__ __forceinline__ void update_mode_coef(float4 c, Vec<float, 2>& s, const Vec<float, 2>& o,
                                                 Cx<float> nt) {
#pragma clang fp contract(off)
    const float mk = c.x, mkc = c.y, d1 = c.z, d2 = c.w;
    s.c[0].x = (__builtin_fmaf(-mk, o.c[0].x, s.c[0].x) + nt.x) * d1;
    s.c[0].y = (__builtin_fmaf(-mk, o.c[0].y, s.c[0].y) + nt.y) * d1;
    s.c[1].x = __builtin_fmaf(-mkc, o.c[1].x, s.c[1].x) * d2;
    s.c[1].y = __builtin_fmaf(-mkc, o.c[1].y, s.c[1].y) * d2;
}

__device__ __forceinline__ void update_mode(const PhysArgs& p, float ks, Vec<float, 2>& s,
                                            const Vec<float, 2>& o, Cx<float> nt) {
    update_mode_coef(mode_coef(p, ks), s, o, nt);
}

// |k|^2 in float, evaluated as the reference's pfc_ks (pfc_common.h:237-239).
__device__ __forceinline__ float ksq(float a, float b, float c) {
#pragma clang fp contract(off)
    return a * a + b * b + c * c;
}

// Deterministic workgroup sum (fixed shuffle tree, then waves in order).
// Result valid in thread 0.  All threads must call.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red) {
    constexpr int W = NT < 64 ? NT : 64;
#pragma unroll
    for (int off = W / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, W);
    const int lane = threadIdx.x % W, wid = threadIdx.x / W;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double tot = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < NT / W; i++) tot += red[i];
    return tot;
}

// ---------------------------------------------------------------------------
// Real-to-complex helpers for one row of length M = 2H.
// ---------------------------------------------------------------------------

// c2r split: Z[k] = (X[k] + X*[H-k]) + i (X[k] - X*[H-k]) e^{+2 pi i k/M};
// IFFT_H(Z) = M (x[2j] + i x[2j+1]).
template <typename R>
__device__ __forceinline__ Cx<R> c2r_split(Cx<R> xk, Cx<R> xc, Cx<R> w) {
    const R ax = xk.x + xc.x, ay = xk.y - xc.y;
    const R dx = xk.x - xc.x, dy = xk.y + xc.y;
    const R bx = dx * w.x + dy * w.y, by = dy * w.x - dx * w.y;
    return Cx<R>{ax - by, ay + bx};
}

// r2c merge: with W = FFT_H(x[2j] + i x[2j+1]),
// X[k] = (W[k] + W*[H-k])/2 + e^{-2 pi i k/M} (W[k] - W*[H-k])/(2i).
template <typename R>
__device__ __forceinline__ Cx<R> r2c_merge(Cx<R> wk, Cx<R> wc, Cx<R> tw) {
    const R half = R(0.5);
    const R ex = half * (wk.x + wc.x), ey = half * (wk.y - wc.y);
    const R dx = wk.x - wc.x, dy = wk.y + wc.y;
    const R ox = half * dy, oy = -half * dx;
    return Cx<R>{ex + (tw.x * ox - tw.y * oy), ey + (tw.x * oy + tw.y * ox)};
}

// c2r split input for packed index k: (X[k], X[H-k]) with X[0], X[H] unpacked.
template <typename R>
__device__ __forceinline__ Cx<R> c2r_split_packed(Cx<R> xk, Cx<R> xc, int k, Cx<R> w) {
    if (k == 0) return c2r_split(Cx<R>{xk.x, R(0)}, Cx<R>{xk.y, R(0)}, w);
    return c2r_split(xk, xc, w);
}

// ---------------------------------------------------------------------------
// Column per-mode work and energy sums
// ---------------------------------------------------------------------------

// Split of a packed column value (k_m = 0 and k_m = M/2 carried as
// Z = X0 + i XH, X0 and XH Hermitian along l): from Z(k) and Z(-k).
template <typename R>
__device__ __forceinline__ void hsplit(Cx<R> z, Cx<R> zm, Cx<R>& x0, Cx<R>& xh) {
    const R h = R(0.5);
    x0 = Cx<R>{h * (z.x + zm.x), h * (z.y - zm.y)};
    xh = Cx<R>{h * (z.y + zm.y), h * (zm.x - z.x)};
}
template <typename R>
__device__ __forceinline__ Cx<R> hjoin(Cx<R> x0, Cx<R> xh) {
    return Cx<R>{x0.x - xh.y, x0.y + xh.x};
}

// Gradient-energy term of one (k_l = i, k_m = kmi, k_n = kni) mode of the
// updated state: (kl'^2 + 2 km'^2) |con_hat|^2, own-direction Nyquist
// removed (Re of the reference's i k con_hat transform).  A mode whose
// conjugate is not stored (0 < km < M/2, or km in {0, M/2} with
// 0 < kn < N/2) counts twice.
// klv = b.kl[i] (loaded ahead by callers that prefetch it).
template <typename R>
__device__ __forceinline__ double mode_energy_kl(const DevBufs& b, int LL, int i, float klv, int kmi, int kni,
                                                 Cx<R> co) {
    const int Mh = b.M / 2, Nh = b.N / 2;
    const bool mself = kmi == 0 || kmi == Mh;
    const double kmd = (kmi == Mh) ? 0.0 : (double)b.km[kmi];
    const double mult = (mself && (kni == 0 || kni == Nh)) ? 1.0 : 2.0;
    const double kld = (i == LL / 2) ? 0.0 : (double)klv;
    return mult * (kld * kld + 2.0 * kmd * kmd) * ((double)co.x * co.x + (double)co.y * co.y);
}
template <typename R>
__device__ __forceinline__ double mode_energy(const DevBufs& b, int LL, int i, int kmi, int kni, Cx<R> co) {
    return mode_energy_kl(b, LL, i, b.kl[i], kmi, kni, co);
}

// Per-mode work of the column pass on one (k_l = i, k_m = kmi, k_n = kni)
// mode: update (mode STEP), DC removal, energy term, C2 products.
template <typename R>
__device__ __forceinline__ void col_mode(const DevBufs& b, const PhysArgs& p, int mode, bool want_e, bool use_nz, int LL, int i,
                                         int kmi, int kni, Cx<R>& ph, Cx<R>& co, Cx<R> on, Cx<R> oc, Cx<R> nz,
                                         Cx<R>& a, Cx<R>& bb, double& g) {
    const int Mh = b.M / 2, Nh = b.N / 2;
    const float kmv = b.km[kmi];
    const float ks = ksq(b.kl[i], kmv, b.kn[kni]);
    const bool mself = kmi == 0 || kmi == Mh;
    if (mode == COL_STEP) {
        Vec<R, 2> s, o;
        s.c[0] = ph; s.c[1] = co;
        o.c[0] = on; o.c[1] = oc;
        Cx<R> nt{R(0), R(0)};
        // self-conjugate selector of the reference: l and m only (misfit.cpp:991-1000)
        if (use_nz) nt = noise_term(p, ks, mself && (i == 0 || i == LL / 2), nz);
        update_mode(p, ks, s, o, nt);
        ph = s.c[0]; co = s.c[1];
    }
    if (i == 0 && kmi == 0 && kni == 0) {
        ph = Cx<R>{R(0), R(0)};
        co = Cx<R>{R(0), R(0)};
    }
    if (want_e) g += mode_energy(b, LL, i, kmi, kni, co);
    const int il = i <= LL / 2 ? i : LL - i;
    const int jn = kni <= Nh ? kni : b.N - kni;
    const size_t ci = 2 * (((size_t)il * (Mh + 1) + kmi) * (Nh + 1) + jn);
    const R ca = (R)b.c2[ci], cb = (R)b.c2[ci + 1];
    a = Cx<R>{ph.x * ca, ph.y * ca};
    bb = Cx<R>{ph.x * cb, ph.y * cb};
}

// Index of the first coefficient-table entry of the k_m = M/2 modes of the
// packed column of n slot j (0 or N/2): after the [L][N][K] entries, slot 0
// then slot N/2, L entries each (k_coef).  The packed column's own entries
// hold its k_m = 0 modes.
__device__ __forceinline__ size_t coef_packed_tail(const DevBufs& b, int j) {
    return (size_t)b.L * b.N * b.K + (j == 0 ? 0 : (size_t)b.L);
}

// Both modes of a packed-column element (k_m = 0 and M/2 at k_l = i, k_n =
// j) from the tabulated coefficients (c0, cc0: k_m = 0; ch, cch: k_m = M/2):
// the element and its mirror -i split into the two Hermitian sequences,
// each updated as col_mode does (update, DC removal, gradient energy, C2
// products) and joined back.  Bit-identical to two col_mode calls.
template <typename R>
__device__ __forceinline__ void packed_modes_coef(const DevBufs& b, const PhysArgs& p, bool want_e, int LL, int i,
                                                  int j, Cx<R>& ph, Cx<R>& co, Cx<R> on, Cx<R> oc, Cx<R> phm,
                                                  Cx<R> com, Cx<R> onm, Cx<R> ocm, float4 c0, float2 cc0, float4 ch,
                                                  float2 cch, Cx<R>& a, Cx<R>& bb, double& g) {
    Cx<R> ph0, phh, co0, coh, on0, onh, oc0, och;
    hsplit(ph, phm, ph0, phh);
    hsplit(co, com, co0, coh);
    hsplit(on, onm, on0, onh);
    hsplit(oc, ocm, oc0, och);
    const Cx<R> z0{R(0), R(0)};
    Vec<R, 2> s0, o0, sh, oh;
    s0.c[0] = ph0; s0.c[1] = co0; o0.c[0] = on0; o0.c[1] = oc0;
    sh.c[0] = phh; sh.c[1] = coh; oh.c[0] = onh; oh.c[1] = och;
    if constexpr (sizeof(R) == 4) {
        update_mode_coef(c0, s0, o0, z0);
        update_mode_coef(ch, sh, oh, z0);
    } else {
        update_mode_coef64(p, c0, s0, o0, z0);
        update_mode_coef64(p, ch, sh, oh, z0);
    }
    ph0 = s0.c[0]; co0 = s0.c[1];
    phh = sh.c[0]; coh = sh.c[1];
    if (i == 0 && j == 0) {
        ph0 = z0;
        co0 = z0;
    }
    if (want_e) g += mode_energy(b, LL, i, 0, j, co0);
    const Cx<R> a0{ph0.x * (R)cc0.x, ph0.y * (R)cc0.x}, b0{ph0.x * (R)cc0.y, ph0.y * (R)cc0.y};
    if (want_e) g += mode_energy(b, LL, i, b.M / 2, j, coh);
    const Cx<R> ah{phh.x * (R)cch.x, phh.y * (R)cch.x}, bh{phh.x * (R)cch.y, phh.y * (R)cch.y};
    ph = hjoin(ph0, phh);
    co = hjoin(co0, coh);
    a = hjoin(a0, ah);
    bb = hjoin(b0, bh);
}

// The energy pair of one instance from its partial sums (ENERGYC2,
// misfit.cpp:1043-1092): the pointwise terms plus alpha/P times the k-space
// gradient sum, both over the grid's points; the density is the FP64 k = 0
// mode, which no step changes.
__device__ __forceinline__ void energy_pair(const DevBufs& b, double alpha, int inst, double se, double sg,
                                            double* e_out) {
    const double P = (double)b.L * (double)b.M * (double)b.N;
    const double tot = se + alpha * sg / P;
    e_out[2 * inst] = tot / b.L / b.M / b.N;
    e_out[2 * inst + 1] = b.means[4 * inst];
}

// One workgroup of NT threads reduces instance inst's partial sums into
// b.etrace (energy-traced stepping; ncol = the column grid's width).
template <int NT>
__device__ void energy_reduce_block(const DevBufs& b, const PhysArgs& p, int inst, int ncol, double* red) {
    const double* er = b.erow + (size_t)inst * b.nrow;
    const double* ec = b.eprev + (size_t)inst * ncol;
    double se = 0.0, sg = 0.0;
    for (int i = threadIdx.x; i < b.nrow; i += NT) se += er[i];
    for (int i = threadIdx.x; i < ncol; i += NT) sg += ec[i];
    se = block_sum<NT>(se, red);
    sg = block_sum<NT>(sg, red);
    if (threadIdx.x == 0) energy_pair(b, (double)p.alpha, inst, se, sg, b.etrace);
}

}  // namespace pfc
