/*
 * pfc_oracle.c - CPU FP64 restatement of the pfc-exp-cli hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker; never part of the product).
 * Each function names the reference lines it follows.  The expression types
 * (float vs double) mirror the reference's C++ overload resolution: with
 * <cmath> and `using namespace std`, sqrt/cos/log/fabs of a float are the float
 * overloads, while pow(float,int) promotes to double.
 *
 * Compile with -ffp-contract=off (no fused multiply-add) so float/double
 * rounding matches an SSE2 x86-64 build of the reference.
 */
#include "pfc_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define PFO_PI 3.1415926 /* pfc_common.h:152 */
#define PFO_TWO_PI_EXACT 6.283185307179586476925286766559

/* ------------------------------------------------------------------------ */
/* Parameters                                                                 */
/* ------------------------------------------------------------------------ */

/* misfit.cpp:116-160, cvd.cpp:26-67: static initialisers (double expressions
 * rounded to float at initialisation). */
void pfo_defaults(pfo_model* p, int model, int L, int M, int N) {
    memset(p, 0, sizeof(*p));
    p->model = model;
    p->L = L; p->M = M; p->N = N;
    p->narrow = 1;
    p->seed = 20200604u;
    p->a_lattice = 1.0f;
    p->gaa1 = 0.8f;
    p->gaa2 = (float)(p->gaa1 * sqrt(2.));
    p->gab1 = 0.8f;
    p->gab2 = (float)(p->gaa2 * sqrt(2.));
    p->niu = 1.4f; p->ksi = 1.f;
    p->sigMa1 = 0.55f; p->sigMa2 = 0.55f; p->sigMb1 = 0.55f; p->sigMb2 = 0.55f;
    p->c0 = 0.5f; p->alpha = 0.5f; p->w = 0.02f;
    p->dt = 0.05f; p->dx = 0.125f; p->con0 = 0.2f;
    p->ka0 = (float)(2.0 * PFO_PI / p->a_lattice);
    p->ka1 = (float)(2.0 * PFO_PI);
    p->ka2 = (float)(sqrt(2.0) * p->ka1);
    if (model == PFO_MISFIT) {
        p->b_lattice = 0.88f;
        p->u0 = 0.05f; p->sig = 0.05f;
        p->kb0 = (float)(2.0 * PFO_PI / p->b_lattice);
        p->kb1 = p->kb0;
        p->kb2 = (float)(p->kb1 * sqrt(2.0));
        p->mn = 0.1f; p->mc = 10.f;
        p->grainx = 16; p->grainy = 16; p->grainz = 1;
        p->steps = 200; p->mod = 25;
    } else {
        p->b_lattice = (float)(sqrt(3.0) / 2.0);
        p->u0 = 0.01f; p->sig = 0.03f;
        p->kb0 = (float)(2.5 * PFO_PI / p->b_lattice);
        p->kb1 = (float)(5.0 * PFO_PI / sqrt(3.0));
        p->kb2 = (float)(p->kb1 * sqrt(2.0));
        p->mn = 1.f; p->mc = 1.f;
        p->grainx = 16; p->grainy = 16; p->grainz = 1;
        p->steps = 2000; p->mod = 200;
    }
    p->axx = (float)(0.0 * p->a_lattice);
    pfo_derive(p);
}

/* misfit.cpp:162-179 */
void pfo_derive(pfo_model* p) {
    p->dnx = (p->grainx > 0) ? (p->L / p->grainx) : 0;
    p->dny = (p->grainy > 0) ? (p->M / p->grainy) : 0;
    p->dnz = (p->grainz > 0) ? (p->N / p->grainz) : 0;
    if (p->dx > 0.0f && p->dt > 0.0f && p->axx >= 0.0f)
        p->b0 = (float)sqrt((double)(p->axx / p->dx / p->dx / p->dx / p->dt));
    else
        p->b0 = 0.0f;
}

/* ------------------------------------------------------------------------ */
/* RNG (pfc_common.h:165-181)                                                 */
/* ------------------------------------------------------------------------ */
uint64_t pfo_splitmix64(uint64_t* state) {
    uint64_t z = (*state += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

double pfo_rng_u01(uint64_t* state) {
    const double inv = 1.0 / 9007199254740992.0;
    return (double)(pfo_splitmix64(state) >> 11) * inv;
}

static double pfo_rng_u11(uint64_t* state) { return 2.0 * pfo_rng_u01(state) - 1.0; }

/* misfit.cpp:636-644 */
float pfo_grain_theta(const pfo_model* p, int iii, int jjj, int kkk) {
    const int dnx = p->dnx, dny = p->dny, dnz = p->dnz;
    unsigned int center_sum = (unsigned int)((dnx * iii + dnx / 2) + (dny * jjj + dny / 2) + (dnz * kkk + dnz / 2));
    uint64_t st = (uint64_t)(p->seed ^ (center_sum * 2654435761u));
    double v1 = 80.0 * (pfo_rng_u01(&st) - 0.5);
    (void)pfo_rng_u01(&st); /* v2, unused by the reference */
    return (float)(v1 * PFO_PI / 180);
}

/* ------------------------------------------------------------------------ */
/* k tables and C2 kernels                                                    */
/* ------------------------------------------------------------------------ */

/* misfit.cpp:455-462 */
void pfo_ktables(const pfo_model* p, float* kl, float* km, float* kn) {
    const int L = p->L, M = p->M, N = p->N;
    const float dx = p->dx;
    for (int l = 0; l < L / 2 + 1; l++) kl[l] = (float)(l * 2.0 * PFO_PI / (double)(L * dx));
    for (int l = L / 2 + 1; l < L; l++) kl[l] = (float)((l - L) * 2.0 * PFO_PI / (double)(L * dx));
    for (int m = 0; m < M / 2 + 1; m++) km[m] = (float)(m * 2.0 * PFO_PI / (double)(M * dx));
    for (int m = M / 2 + 1; m < M; m++) km[m] = (float)((m - M) * 2.0 * PFO_PI / (double)(M * dx));
    for (int n = 0; n < N / 2 + 1; n++) kn[n] = (float)(n * 2.0 * PFO_PI / (double)(N * dx));
    for (int n = N / 2 + 1; n < N; n++) kn[n] = (float)((n - N) * 2.0 * PFO_PI / (double)(N * dx));
}

/* pfc_common.h:237-239 */
static inline float pfo_ks(const float* kl, const float* km, const float* kn, int i, int j, int k) {
    return kl[i] * kl[i] + km[j] * km[j] + kn[k] * kn[k];
}

/* misfit.cpp:703-706 / 730-733: kc from the float quadratic. */
static float pfo_kc(float sig, float sg1, float sg2, float g1, float g2, float k1, float k2) {
    float aa1 = (float)pow((double)(sig / sg1), 2);
    float aa2 = (float)pow((double)(sig / sg2), 2);
    float bb1 = (float)(1. / (2. * g1 * g1));
    float bb2 = (float)(1. / (2. * g2 * g2));
    float t1 = bb1 * k1 - bb2 * k2;
    float t2 = (bb1 - bb2) * (aa1 - aa2 + bb1 * k1 * k1 - bb2 * k2 * k2);
    float kc = (float)(((double)t1 + sqrt(pow((double)t1, 2) - (double)t2)) / (double)(bb1 - bb2));
    if (g1 == g2) kc = (aa2 - aa1 + bb1 * (k2 * k2 - k1 * k1)) / (2 * bb1 * (k2 - k1));
    return kc;
}

/* Two-branch Gaussian peak kernel value at |k|=sk (misfit.cpp:714-715). */
static float pfo_c2val(float sk, float kc, float sig, float sg1, float sg2, float g1, float g2,
                       float k1, float k2) {
    if (sk <= kc)
        return (float)(exp(-pow((double)(sig / sg1), 2)) *
                       exp(-pow((double)(sk - k1), 2) / (double)(2 * g1 * g1)));
    return (float)(exp(-pow((double)(sig / sg2), 2)) *
                   exp(-pow((double)(sk - k2), 2) / (double)(2 * g2 * g2)));
}

/* misfit.cpp:700-750 (C2AK, C2BK). */
void pfo_c2(const pfo_model* p, float* c2a, float* c2b, float* kc1o, float* kc2o) {
    const int L = p->L, M = p->M, N = p->N;
    float* kl = (float*)malloc(sizeof(float) * L);
    float* km = (float*)malloc(sizeof(float) * M);
    float* kn = (float*)malloc(sizeof(float) * N);
    pfo_ktables(p, kl, km, kn);
    float kc1 = pfo_kc(p->sig, p->sigMa1, p->sigMa2, p->gaa1, p->gaa2, p->ka1, p->ka2);
    float kc2 = pfo_kc(p->sig, p->sigMb1, p->sigMb2, p->gab1, p->gab2, p->kb1, p->kb2);
    if (kc1o) *kc1o = kc1;
    if (kc2o) *kc2o = kc2;
    for (int l = 0; l < L; l++)
        for (int m = 0; m < M; m++)
            for (int n = 0; n < N; n++) {
                float sk = sqrtf(pfo_ks(kl, km, kn, l, m, n));
                size_t i = (size_t)n + (size_t)N * ((size_t)m + (size_t)M * l);
                if (c2a) c2a[i] = pfo_c2val(sk, kc1, p->sig, p->sigMa1, p->sigMa2, p->gaa1, p->gaa2, p->ka1, p->ka2);
                if (c2b) c2b[i] = pfo_c2val(sk, kc2, p->sig, p->sigMb1, p->sigMb2, p->gab1, p->gab2, p->kb1, p->kb2);
            }
    free(kl); free(km); free(kn);
}

/* ------------------------------------------------------------------------ */
/* Initial fields                                                             */
/* ------------------------------------------------------------------------ */

/* INITIAL_dis, misfit.cpp:610-682 (single rank: local_0_start=0, n0=L). */
static void pfo_init_misfit(const pfo_model* p, double* phi, double* con) {
    const int L = p->L, M = p->M, N = p->N;
    const float A1 = 0.1427f, A2 = 0.106f;
    const float u0 = p->u0, con0 = p->con0, dx = p->dx;
    const int seed_m_max = M < 160 ? M : 160;
    const size_t P = (size_t)L * M * N;
    for (size_t i = 0; i < P; i++) { phi[i] = u0; con[i] = con0; }
    const float kk = (1 - con0) * p->ka0 + con0 * p->kb0;
    const double lim_x = (double)(L * dx / (float)p->grainx) / 2. - 0.2;
    const double lim_y = (double)(M * dx / (float)p->grainy) / 2. - 0.2;
    for (int iii = 0; iii < p->grainx; iii++)
        for (int jjj = 0; jjj < p->grainy; jjj++)
            for (int kkk = 0; kkk < p->grainz; kkk++) {
                const float center_x = (float)(p->dnx * iii + p->dnx / 2);
                const float center_y = (float)(p->dny * jjj + p->dny / 2);
                const float center_z = (float)(p->dnz * kkk + p->dnz / 2);
                const float theta = pfo_grain_theta(p, iii, jjj, kkk);
                const float ct = cosf(theta);
                const double cs = cos(PFO_PI / 2 - (double)theta - (iii + jjj + kkk) * PFO_PI);
                for (int l = 0; l < L; l++)
                    for (int m = 0; m < seed_m_max; m++)
                        for (int n = 0; n < N; n++) {
                            const int lx = l;
                            float xx = (0 - center_x + l) * dx;
                            float yy = (m - center_y) * dx;
                            float zz = (n - center_z) * dx;
                            int X = (int)((double)(lx * ct) + m * cs);
                            int Y = (int)((double)(m * ct) - lx * cs);
                            if (fabsf(xx) < lim_x && fabsf(yy) < lim_y && fabsf(zz) < 4.3) {
                                float cx = cosf(kk * X * dx), cy = cosf(kk * Y * dx);
                                float v = 2 * A1 * (cx + cy) + 4 * A2 * cx * cy + u0;
                                phi[(size_t)n + (size_t)N * ((size_t)m + (size_t)M * l)] = v;
                            }
                        }
            }
}

/* INITIAL_C2, cvd.cpp:587-642 (w3 = 0). */
static void pfo_init_cvd(const pfo_model* p, double* phi, double* con) {
    const int L = p->L, M = p->M, N = p->N;
    const float A1 = 0.15f, A2 = 0.106f;
    const float w3 = (float)(0. * PFO_PI / 180);
    const float cw = cosf(w3), sw = sinf(w3);
    for (int l = 0; l < L; l++)
        for (int m = 0; m < M; m++)
            for (int n = 0; n < N; n++) {
                float x = (0 - L / 2 + l) * p->dx;
                float y = (m - M / 2) * p->dx;
                size_t i = (size_t)n + (size_t)N * ((size_t)m + (size_t)M * l);
                float k0 = (m < M / 2) ? p->ka0 : p->kb0;
                float a = cosf(k0 * x * cw + k0 * y * sw);
                float b = cosf(-x * k0 * sw + k0 * y * cw);
                float b2 = cosf(-x * k0 * sw + y * k0 * cw);
                float v = 2 * A1 * (a + b) + 4 * A2 * a * b2 + p->u0;
                phi[i] = v;
                con[i] = (m < M / 2) ? p->con0 : (float)(1 - p->con0);
            }
}

void pfo_init_fields(const pfo_model* p, double* phi, double* con) {
    if (p->model == PFO_MISFIT) pfo_init_misfit(p, phi, con);
    else pfo_init_cvd(p, phi, con);
}

/* ------------------------------------------------------------------------ */
/* FFT: mixed-radix Stockham autosort, double, unnormalised                   */
/* (stands in for FFTW's c2c fftw_plan_dft_3d, misfit.cpp:438-452)            */
/* ------------------------------------------------------------------------ */
typedef struct { double re, im; } cpx;

static void dft_small(cpx* v, int R, int sign, const cpx* wr) {
    /* wr[q] = exp(sign*2*pi*i*q/R) */
    cpx t[16];
    if (R == 2) {
        cpx a = v[0], b = v[1];
        v[0].re = a.re + b.re; v[0].im = a.im + b.im;
        v[1].re = a.re - b.re; v[1].im = a.im - b.im;
        return;
    }
    if (R == 4) {
        cpx a = v[0], b = v[1], c = v[2], d = v[3];
        cpx s0 = {a.re + c.re, a.im + c.im}, d0 = {a.re - c.re, a.im - c.im};
        cpx s1 = {b.re + d.re, b.im + d.im}, d1 = {b.re - d.re, b.im - d.im};
        /* multiply d1 by sign*i */
        cpx jd1 = {-sign * d1.im, sign * d1.re};
        v[0].re = s0.re + s1.re; v[0].im = s0.im + s1.im;
        v[2].re = s0.re - s1.re; v[2].im = s0.im - s1.im;
        v[1].re = d0.re + jd1.re; v[1].im = d0.im + jd1.im;
        v[3].re = d0.re - jd1.re; v[3].im = d0.im - jd1.im;
        return;
    }
    for (int k = 0; k < R; k++) {
        double sr = 0, si = 0;
        for (int q = 0; q < R; q++) {
            const cpx w = wr[(q * k) % R];
            sr += v[q].re * w.re - v[q].im * w.im;
            si += v[q].re * w.im + v[q].im * w.re;
        }
        t[k].re = sr; t[k].im = si;
    }
    for (int k = 0; k < R; k++) v[k] = t[k];
}

static int fft_factor(int n, int* f) {
    int nf = 0;
    while (n % 4 == 0) { f[nf++] = 4; n /= 4; }
    while (n % 2 == 0) { f[nf++] = 2; n /= 2; }
    for (int p = 3; p * p <= n; p += 2)
        while (n % p == 0) { f[nf++] = p; n /= p; }
    if (n > 1) f[nf++] = n;
    return nf;
}

/* Twiddles exp(sign 2 pi i t/n), t < n, kept per thread for the last few
 * (n, sign) pairs: the same values on every call, computed once. */
#define PFO_TW_SLOTS 8
typedef struct { int n, sign; cpx* tw; cpx* y; } pfo_tw_slot;
static _Thread_local pfo_tw_slot pfo_tw_cache[PFO_TW_SLOTS];
static _Thread_local int pfo_tw_next;

static pfo_tw_slot* pfo_twiddles(int n, int sign) {
    for (int i = 0; i < PFO_TW_SLOTS; i++)
        if (pfo_tw_cache[i].tw && pfo_tw_cache[i].n == n && pfo_tw_cache[i].sign == sign) return &pfo_tw_cache[i];
    pfo_tw_slot* s = &pfo_tw_cache[pfo_tw_next];
    pfo_tw_next = (pfo_tw_next + 1) % PFO_TW_SLOTS;
    free(s->tw);
    free(s->y);
    s->tw = (cpx*)malloc(sizeof(cpx) * n);
    s->y = (cpx*)malloc(sizeof(cpx) * n);
    if (!s->tw || !s->y) {
        fprintf(stderr, "pfo_fft1d: out of memory\n");
        abort();
    }
    s->n = n;
    s->sign = sign;
    for (int t = 0; t < n; t++) {
        double a = PFO_TWO_PI_EXACT * (double)t / (double)n;
        s->tw[t].re = cos(a); s->tw[t].im = sign * sin(a);
    }
    return s;
}

void pfo_fft1d(double* zd, int n, int sign) {
    if (n <= 1) return;
    cpx* x = (cpx*)zd;
    pfo_tw_slot* slot = pfo_twiddles(n, sign);
    cpx* y = slot->y;  /* work buffer of this (n, sign) */
    const cpx* tw = slot->tw;
    int f[64];
    int nf = fft_factor(n, f);
    cpx* in = x; cpx* out = y;
    int Ns = 1;
    for (int s = 0; s < nf; s++) {
        const int R = f[s];
        if (R > 16) { /* generic large prime: direct DFT (not used for supported sizes) */ }
        cpx wr[64];
        for (int q = 0; q < R && q < 64; q++) {
            double a = PFO_TWO_PI_EXACT * (double)q / (double)R;
            wr[q].re = cos(a); wr[q].im = sign * sin(a);
        }
        const int NB = n / R;
        const int tstep = n / (Ns * R);  /* k*r*tstep < n for k < Ns, r < R */
        for (int j = 0; j < NB; j++) {
            const int k = j % Ns;
            cpx v[64];
            for (int r = 0; r < R; r++) {
                cpx a = in[j + r * NB];
                const cpx w = tw[k * r * tstep];
                v[r].re = a.re * w.re - a.im * w.im;
                v[r].im = a.re * w.im + a.im * w.re;
            }
            dft_small(v, R, sign, wr);
            const int idx = (j / Ns) * Ns * R + k;
            for (int r = 0; r < R; r++) out[idx + r * Ns] = v[r];
        }
        cpx* t = in; in = out; out = t;
        Ns *= R;
    }
    if (in != x) memcpy(x, in, sizeof(cpx) * n);
}

void pfo_fft3d(double* zd, int L, int M, int N, int sign) {
    cpx* z = (cpx*)zd;
    int mx = L > M ? L : M; if (N > mx) mx = N;
    cpx* line = (cpx*)malloc(sizeof(cpx) * mx);
    /* along n (contiguous) */
    if (N > 1)
        for (size_t b = 0; b < (size_t)L * M; b++) pfo_fft1d((double*)(z + b * N), N, sign);
    /* along m (stride N) */
    if (M > 1)
        for (int l = 0; l < L; l++)
            for (int n = 0; n < N; n++) {
                cpx* base = z + (size_t)M * N * l + n;
                for (int m = 0; m < M; m++) line[m] = base[(size_t)m * N];
                pfo_fft1d((double*)line, M, sign);
                for (int m = 0; m < M; m++) base[(size_t)m * N] = line[m];
            }
    /* along l (stride M*N), gathered 8 columns at a time for cache reuse */
    if (L > 1) {
        enum { CB = 8 };
        cpx* blk = (cpx*)malloc(sizeof(cpx) * (size_t)L * CB);
        const size_t cols = (size_t)M * N;
        for (size_t b0 = 0; b0 < cols; b0 += CB) {
            const int nb = (int)(cols - b0 < CB ? cols - b0 : CB);
            for (int l = 0; l < L; l++)
                for (int c = 0; c < nb; c++) blk[(size_t)c * L + l] = z[(size_t)l * cols + b0 + c];
            for (int c = 0; c < nb; c++) pfo_fft1d((double*)(blk + (size_t)c * L), L, sign);
            for (int l = 0; l < L; l++)
                for (int c = 0; c < nb; c++) z[(size_t)l * cols + b0 + c] = blk[(size_t)c * L + l];
        }
        free(blk);
    }
    free(line);
}

/* ------------------------------------------------------------------------ */
/* Pointwise terms                                                            */
/* ------------------------------------------------------------------------ */

/* UPDATEINN3, misfit.cpp:777 */
double pfo_in1(const pfo_model* p, double phi) {
    return -p->niu * pow(phi, 2) / 2.0 + p->ksi * pow(phi, 3) / 3.0;
}

/* UPDATEINFMIX, misfit.cpp:792 */
double pfo_in2(const pfo_model* p, double c) {
    return p->w * (c * log(c / p->c0) + (1.0 - c) * log((1.0 - c) / (1.0 - p->c0)));
}

/* UPDATEINFCMIX second loop, misfit.cpp:826 */
double pfo_in3(double c, double a, double b) {
    return -(1.0 - 3.0 * pow(c, 2) + 2.0 * pow(c, 3)) * a -
           (1.0 - 3.0 * pow((1.0 - c), 2) + 2.0 * pow((1.0 - c), 3)) * b;
}

/* UPDATEINC2MIXC, misfit.cpp:858-867: float temporaries when narrowing. */
double pfo_inc(const pfo_model* p, double phi, double c, double a, double b) {
    double t1 = p->w * (phi + 1.0) * log(c / p->c0);
    double t2 = p->w * (phi + 1.0) * log((1.0 - c) / (1.0 - p->c0));
    double t3 = 3.0 * phi * (a - b) * pow(c, 2);
    double t4 = 3.0 * phi * (-1.0 * a + b) * c;
    if (p->narrow) {
        float c1 = (float)t1, c2 = (float)t2, c3 = (float)t3, c4 = (float)t4;
        return (double)(c1 - c2 - c3 - c4);
    }
    return t1 - t2 - t3 - t4;
}

/* MAINEQUATIONC2_NOI / MAINEQUATIONC2MIX for one mode, noise term 0.
 * misfit.cpp:1002-1005, cvd.cpp:872-875 */
void pfo_update_phi(const pfo_model* p, float ks, double* z, const double* outn) {
    const float mk = p->mn * p->dt * ks;
    double r1 = z[0] - mk * outn[0];
    double r2 = z[1] - mk * outn[1];
    if (p->narrow) { r1 = (float)r1; r2 = (float)r2; }
    z[0] = r1 / (1.0 + mk);
    z[1] = r2 / (1.0 + mk);
}

/* misfit.cpp:1007-1010, cvd.cpp:877-880 */
void pfo_update_con(const pfo_model* p, float ks, double* z, const double* outc) {
    const float mk = p->mc * p->dt * ks;
    const float den = p->mc * p->dt * p->alpha * ks * ks;
    double r3 = z[0] - mk * outc[0];
    double r4 = z[1] - mk * outc[1];
    if (p->narrow) { r3 = (float)r3; r4 = (float)r4; }
    z[0] = r3 / (1.0 + den);
    z[1] = r4 / (1.0 + den);
}

/* ------------------------------------------------------------------------ */
/* Simulation                                                                 */
/* ------------------------------------------------------------------------ */
struct pfo_sim {
    pfo_model p;
    size_t P;
    float *kl, *km, *kn, *c2a, *c2b;
    double *phi, *con;               /* complex interleaved, length 2P */
    double *phiHat, *conHat, *in1, *in2, *in3, *inn, *inc, *outn, *outc;
    double *c2adenp, *c2bdenp, *c2aden, *c2bden;
    double *gra1, *gra2, *gra3, *grac;
    double *nois;
};

static double* zalloc(size_t n) { return (double*)calloc(n, sizeof(double)); }

pfo_sim* pfo_create(const pfo_model* p) {
    pfo_sim* s = (pfo_sim*)calloc(1, sizeof(pfo_sim));
    s->p = *p;
    const size_t P = (size_t)p->L * p->M * p->N;
    s->P = P;
    s->kl = (float*)malloc(sizeof(float) * p->L);
    s->km = (float*)malloc(sizeof(float) * p->M);
    s->kn = (float*)malloc(sizeof(float) * p->N);
    pfo_ktables(p, s->kl, s->km, s->kn);
    s->c2a = (float*)malloc(sizeof(float) * P);
    s->c2b = (float*)malloc(sizeof(float) * P);
    pfo_c2(p, s->c2a, s->c2b, NULL, NULL);
    s->phi = zalloc(2 * P); s->con = zalloc(2 * P);
    s->phiHat = zalloc(2 * P); s->conHat = zalloc(2 * P);
    s->in1 = zalloc(2 * P); s->in2 = zalloc(2 * P); s->in3 = zalloc(2 * P);
    s->inn = zalloc(2 * P); s->inc = zalloc(2 * P);
    s->outn = zalloc(2 * P); s->outc = zalloc(2 * P);
    s->c2adenp = zalloc(2 * P); s->c2bdenp = zalloc(2 * P);
    s->c2aden = zalloc(2 * P); s->c2bden = zalloc(2 * P);
    s->gra1 = zalloc(2 * P); s->gra2 = zalloc(2 * P); s->gra3 = zalloc(2 * P);
    s->grac = zalloc(2 * P);
    s->nois = zalloc(2 * P);
    return s;
}

void pfo_destroy(pfo_sim* s) {
    if (!s) return;
    double* arr[] = {s->phi, s->con, s->phiHat, s->conHat, s->in1, s->in2, s->in3, s->inn, s->inc,
                     s->outn, s->outc, s->c2adenp, s->c2bdenp, s->c2aden, s->c2bden, s->gra1,
                     s->gra2, s->gra3, s->grac, s->nois};
    for (size_t i = 0; i < sizeof(arr) / sizeof(arr[0]); i++) free(arr[i]);
    free(s->kl); free(s->km); free(s->kn); free(s->c2a); free(s->c2b);
    free(s);
}

/* Sets real parts; Im(phi)=0, Im(con)=1e-10 as INITIAL_dis (misfit.cpp:624-627). */
void pfo_set_fields(pfo_sim* s, const double* phi, const double* con) {
    for (size_t i = 0; i < s->P; i++) {
        s->phi[2 * i] = phi[i]; s->phi[2 * i + 1] = 0.0;
        s->con[2 * i] = con[i]; s->con[2 * i + 1] = 1.0E-10;
    }
}

void pfo_get_fields(const pfo_sim* s, double* phi, double* con) {
    for (size_t i = 0; i < s->P; i++) {
        if (phi) phi[i] = s->phi[2 * i];
        if (con) con[i] = s->con[2 * i];
    }
}

static void fft(pfo_sim* s, const double* in, double* out, int sign) {
    if (out != in) memcpy(out, in, sizeof(double) * 2 * s->P);
    pfo_fft3d(out, s->p.L, s->p.M, s->p.N, sign);
}

/* UPDATEINFCMIX (misfit.cpp:801-833) */
static void updateinfcmix(pfo_sim* s) {
    const size_t P = s->P;
    for (size_t i = 0; i < P; i++) {
        s->c2adenp[2 * i] = s->c2a[i] * s->phiHat[2 * i];
        s->c2adenp[2 * i + 1] = s->c2a[i] * s->phiHat[2 * i + 1];
        s->c2bdenp[2 * i] = s->c2b[i] * s->phiHat[2 * i];
        s->c2bdenp[2 * i + 1] = s->c2b[i] * s->phiHat[2 * i + 1];
    }
    fft(s, s->c2adenp, s->c2aden, +1);
    fft(s, s->c2bdenp, s->c2bden, +1);
    const int LMN = s->p.L * s->p.M * s->p.N;
    for (size_t i = 0; i < P; i++) {
        s->c2aden[2 * i] /= LMN; s->c2aden[2 * i + 1] /= LMN;
        s->c2bden[2 * i] /= LMN; s->c2bden[2 * i + 1] /= LMN;
        s->in3[2 * i] = pfo_in3(s->con[2 * i], s->c2aden[2 * i], s->c2bden[2 * i]);
        s->in3[2 * i + 1] = 0.0;
    }
}

/* UPDATEINC2GRA (misfit.cpp:877-925), including the grac = g1^2 + 2 g2^2 quirk. */
static void updateinc2gra(pfo_sim* s) {
    const int L = s->p.L, M = s->p.M, N = s->p.N;
    for (int l = 0; l < L; l++)
        for (int m = 0; m < M; m++)
            for (int n = 0; n < N; n++) {
                size_t i = (size_t)n + (size_t)N * ((size_t)m + (size_t)M * l);
                const double cr = s->conHat[2 * i], ci = s->conHat[2 * i + 1];
                s->gra1[2 * i] = 1.0 * s->kl[l] * ci; s->gra1[2 * i + 1] = -1.0 * s->kl[l] * cr;
                s->gra2[2 * i] = 1.0 * s->km[m] * ci; s->gra2[2 * i + 1] = -1.0 * s->km[m] * cr;
                s->gra3[2 * i] = 1.0 * s->kn[n] * ci; s->gra3[2 * i + 1] = -1.0 * s->kn[n] * cr;
            }
    pfo_fft3d(s->gra1, L, M, N, +1);
    pfo_fft3d(s->gra2, L, M, N, +1);
    pfo_fft3d(s->gra3, L, M, N, +1);
    for (size_t i = 0; i < s->P; i++) {
        double g1 = s->gra1[2 * i] / L / M / N;
        double g2 = s->gra2[2 * i] / L / M / N;
        s->grac[2 * i] = pow(g1, 2) + pow(g2, 2) + pow(g2, 2);
        s->grac[2 * i + 1] = 0.0;
    }
}

/* Gauss (misfit.cpp:951-981): Marsaglia polar, sequential splitmix64 stream
 * seeded from (seed, timestep); n values at out[0], out[stride], ...  Returns
 * the number of uniform pairs drawn (accepted and rejected). */
static uint64_t gauss_fill(unsigned int seed, int timestep, size_t n, double* out, size_t stride) {
    uint64_t st = (uint64_t)seed;
    st ^= (uint64_t)(unsigned int)(0 + 1) * 2654435761ULL;
    st ^= (uint64_t)(unsigned int)timestep * 374761393ULL;
    uint64_t pairs = 0;
    for (size_t i = 0; i < n; i++) {
        float r = 2.0f, v1 = 0, v2 = 0;
        while (r >= 1.0f || r <= 0.0f) {
            v1 = (float)pfo_rng_u11(&st);
            v2 = (float)pfo_rng_u11(&st);
            r = v1 * v1 + v2 * v2;
            pairs++;
        }
        float f = (float)sqrt(-2. * (double)logf(r) / r);
        out[i * stride] = f * v1;
    }
    return pairs;
}

uint64_t pfo_gauss(unsigned int seed, int timestep, size_t n, double* out) {
    return gauss_fill(seed, timestep, n, out, 1);
}

static void gauss_noise(pfo_sim* s, int timestep) {
    gauss_fill(s->p.seed, timestep, s->P, s->nois, 2);
    for (size_t i = 0; i < s->P; i++) s->nois[2 * i + 1] = 0.0;
}

/* One step: run_misfit loop body misfit.cpp:479-498 / run_cvd cvd.cpp:400-419. */
static void step_core(pfo_sim* s, int t) {
    const pfo_model* p = &s->p;
    const size_t P = s->P;
    const int L = p->L, M = p->M, N = p->N;
    for (size_t i = 0; i < P; i++) {
        s->in1[2 * i] = pfo_in1(p, s->phi[2 * i]); s->in1[2 * i + 1] = 0.0;
        s->in2[2 * i] = pfo_in2(p, s->con[2 * i]); s->in2[2 * i + 1] = 0.0;
    }
    fft(s, s->phi, s->phiHat, -1);
    updateinfcmix(s);
    for (size_t i = 0; i < P; i++) {
        s->inn[2 * i] = s->in1[2 * i] + s->in2[2 * i] + s->in3[2 * i];
        s->inn[2 * i + 1] = 0.0;
        s->inc[2 * i] = pfo_inc(p, s->phi[2 * i], s->con[2 * i], s->c2aden[2 * i], s->c2bden[2 * i]);
        s->inc[2 * i + 1] = 0.0;
    }
    fft(s, s->inn, s->outn, -1);
    fft(s, s->inc, s->outc, -1);
    fft(s, s->con, s->conHat, -1);
    const int noisy = (p->model == PFO_MISFIT);
    if (noisy && p->b0 != 0.0f) {
        gauss_noise(s, t);
        pfo_fft3d(s->nois, L, M, N, -1);
    }
    for (int l = 0; l < L; l++)
        for (int m = 0; m < M; m++)
            for (int n = 0; n < N; n++) {
                size_t i = (size_t)n + (size_t)N * ((size_t)m + (size_t)M * l);
                const float ks = pfo_ks(s->kl, s->km, s->kn, l, m, n);
                if (noisy && p->b0 != 0.0f) {
                    /* MAINEQUATIONC2_NOI with its noise term (misfit.cpp:991-1005). */
                    float detrx = (l == 0) ? 1.0f : 0.0f, detrxn = (l == L / 2) ? 1.0f : 0.0f;
                    float detry = (m == 0) ? 1.0f : 0.0f, detryn = (m == M / 2) ? 1.0f : 0.0f;
                    const float mk = p->mn * p->dt * ks;
                    const float db = p->dt * p->b0;
                    const float sks = sqrtf(ks);
                    const double damp = 0.5 - 0.5 * tanh(((double)sks - 2.0 * PFO_PI) / 0.1);
                    const double sel = (double)((detrx + detrxn) * (detry + detryn));
                    double r1 = s->phiHat[2 * i] - mk * s->outn[2 * i] +
                                db * s->nois[2 * i] * sks * sqrt(fabs(0.5 + 0.5 * sel)) * damp;
                    double r2 = s->phiHat[2 * i + 1] - mk * s->outn[2 * i + 1] +
                                db * s->nois[2 * i + 1] * sks * sqrt(fabs(0.5 - 0.5 * sel)) * damp;
                    if (p->narrow) { r1 = (float)r1; r2 = (float)r2; }
                    s->phiHat[2 * i] = r1 / (1.0 + mk);
                    s->phiHat[2 * i + 1] = r2 / (1.0 + mk);
                } else {
                    pfo_update_phi(p, ks, s->phiHat + 2 * i, s->outn + 2 * i);
                }
                pfo_update_con(p, ks, s->conHat + 2 * i, s->outc + 2 * i);
            }
}

static void step_tail(pfo_sim* s, int faithful) {
    const size_t P = s->P;
    if (faithful) updateinc2gra(s);
    fft(s, s->phiHat, s->phi, +1);
    fft(s, s->conHat, s->con, +1);
    const int LMN = s->p.L * s->p.M * s->p.N;
    for (size_t i = 0; i < P; i++) {
        s->phi[2 * i] = s->phi[2 * i] / LMN; s->phi[2 * i + 1] = 0;
        s->con[2 * i] = s->con[2 * i] / LMN; s->con[2 * i + 1] = 0.0;
    }
    if (faithful) updateinfcmix(s); /* misfit.cpp:501 */
}

void pfo_step(pfo_sim* s, int t) {
    step_core(s, t);
    step_tail(s, 0);
}

void pfo_run(pfo_sim* s, int t0, int n, int faithful) {
    for (int t = t0; t < t0 + n; t++) {
        step_core(s, t);
        step_tail(s, faithful);
    }
}

/* ENERGYC2 (misfit.cpp:1043-1092): recomputes in3 and grac from the current
 * state (as the reference's per-step diagnostic calls leave them). */
void pfo_energy(pfo_sim* s, double* e_mean, double* rho_mean) {
    const pfo_model* p = &s->p;
    fft(s, s->phi, s->phiHat, -1);
    updateinfcmix(s);
    fft(s, s->con, s->conHat, -1);
    updateinc2gra(s);
    double E = 0., D = 0.;
    for (size_t i = 0; i < s->P; i++) {
        const double ph = s->phi[2 * i], c = s->con[2 * i];
        double eid = pow(ph, 2) / 2.0 - p->niu * pow(ph, 3) / 6.0 + p->ksi * pow(ph, 4) / 12.0;
        double emix = (ph + 1.0) * p->w * (c * log(c / p->c0) + (1.0 - c) * log((1.0 - c) / (1.0 - p->c0)));
        double eex = 0.5 * ph * s->in3[2 * i] + p->alpha * s->grac[2 * i];
        double e;
        if (p->narrow) e = (double)((float)eid + (float)emix + (float)eex);
        else e = eid + emix + eex;
        E += e;
        D += ph;
    }
    *e_mean = E / p->L / p->M / p->N;
    *rho_mean = D / p->L / p->M / p->N;
}

/* ------------------------------------------------------------------------ */
/* Phimax: Misfit (misfit.cpp:3261-3464)                                      */
/* ------------------------------------------------------------------------ */
static double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

int pfo_phimax_misfit(int L, int M, int N, float dx, float con0, const double* phi,
                      const double* con, pfo_atoms* out) {
    const int n0 = L;
    double* buf = (double*)calloc((size_t)M * (n0 + 2), sizeof(double));
    double* bufc = (double*)calloc((size_t)M * (n0 + 2), sizeof(double));
    /* dphi[m + M*l]: 2-D indexing of the flat array (ignores n). */
    for (int l = 1; l < n0 + 1; l++)
        for (int m = 0; m < M; m++) {
            buf[m + (size_t)M * l] = phi[m + (size_t)M * (l - 1)];
            bufc[m + (size_t)M * l] = con[m + (size_t)M * (l - 1)];
        }
    /* MPI_Sendrecv stubs with head==tail==0: edge replication. */
    memcpy(buf, buf + M, sizeof(double) * M);
    memcpy(buf + (size_t)(n0 + 1) * M, buf + (size_t)n0 * M, sizeof(double) * M);
    memcpy(bufc, bufc + M, sizeof(double) * M);
    memcpy(bufc + (size_t)(n0 + 1) * M, bufc + (size_t)n0 * M, sizeof(double) * M);

    double gmax = -INFINITY;
    int gm = 0, gl = 1;
    for (int l = 1; l < n0 + 1; ++l)
        for (int m = 0; m < M; ++m) {
            const double v = buf[m + (size_t)M * l];
            if (!isfinite(v)) continue;
            gmax = gmax > v ? gmax : v;
            if (v >= gmax) { gm = m; gl = l; }
        }

    const size_t max_atoms = (size_t)M * n0 * (N > 1 ? N : 1);
    int na = 0;
    for (int n = 0; n < N; n++)
        for (int m = 0; m < M; m++)
            for (int l = 1; l < n0 + 1; l++) {
                const double center = buf[m + (size_t)M * l];
                if (!isfinite(center) || center <= -0.07) continue;
                int lf = l - 1, lb = l + 1, mf = m - 1, mb = m + 1;
                if (mf < 0) mf = M - 1;
                if (mb > M - 1) mb = 0;
#define NBV(mm, ll) (isfinite(buf[(mm) + (size_t)M * (ll)]) ? buf[(mm) + (size_t)M * (ll)] : -INFINITY)
                const double nb[8] = {NBV(m, lf), NBV(m, lb), NBV(mf, l), NBV(mb, l),
                                      NBV(mb, lb), NBV(mb, lf), NBV(mf, lb), NBV(mf, lf)};
#undef NBV
                int ge_all = 1, gt_any = 0;
                for (int q = 0; q < 8; q++) {
                    if (center < nb[q]) ge_all = 0;
                    if (center > nb[q]) gt_any = 1;
                }
                if (!(ge_all && gt_any)) continue;
                const double xtemp = (double)((float)(0 + l + 1) * dx);
                const double ytemp = (double)((float)(m + 1) * dx);
                const double ztemp = (double)((float)(n + 1) * dx);
                const double z0 = center;
                const double z1 = buf[m + (size_t)M * lb];
                const double z3 = buf[m + (size_t)M * lf];
                const double z4 = buf[mb + (size_t)M * l];
                const double z2 = buf[mf + (size_t)M * l];
                double delta_x = 0.0, delta_y = 0.0, phi_local = z0;
                if (isfinite(z1) && isfinite(z3) && isfinite(z4) && isfinite(z2)) {
                    const double dx_d = (double)dx;
                    const double inv2dx = 0.5 / dx_d;
                    const double invdx2 = 1.0 / (dx_d * dx_d);
                    const double a1 = (z1 - z3) * inv2dx;
                    const double a2 = (z4 - z2) * inv2dx;
                    const double a3 = 0.5 * (z1 + z3 - 2.0 * z0) * invdx2;
                    const double a4 = 0.5 * (z2 + z4 - 2.0 * z0) * invdx2;
                    const double eps = 1e-12;
                    if (isfinite(a3) && fabs(a3) > eps) delta_x = -a1 / (2.0 * a3);
                    if (isfinite(a4) && fabs(a4) > eps) delta_y = -a2 / (2.0 * a4);
                    const double half_dx = 0.5 * dx_d;
                    delta_x = clampd(delta_x, -half_dx, half_dx);
                    delta_y = clampd(delta_y, -half_dx, half_dx);
                    double phi_interp = z0;
                    if (isfinite(a3) && fabs(a3) > eps && isfinite(a1)) phi_interp -= 0.25 * a1 * a1 / a3;
                    if (isfinite(a4) && fabs(a4) > eps && isfinite(a2)) phi_interp -= 0.25 * a2 * a2 / a4;
                    if (isfinite(phi_interp)) phi_local = phi_interp;
                }
                if ((size_t)na < max_atoms && na < out->cap) {
                    out->x[na] = xtemp + delta_x;
                    out->y[na] = ytemp + delta_y;
                    out->z[na] = ztemp;
                    out->phi[na] = phi_local;
                    const double c = bufc[m + (size_t)M * l];
                    out->con[na] = isfinite(c) ? (float)c : con0;
                    na++;
                }
            }
    if (na == 0 && out->cap > 0) {
        const double dx_d = (double)dx;
        out->x[0] = (0 + gl + 1) * dx_d;
        out->y[0] = (gm + 1) * dx_d;
        out->z[0] = dx_d;
        out->phi[0] = isfinite(gmax) ? gmax : 0.0;
        const double c0 = bufc[gm + (size_t)M * gl];
        out->con[0] = isfinite(c0) ? (float)c0 : con0;
        na = 1;
    }
    out->n = na;
    free(buf); free(bufc);
    return na;
}

/* ------------------------------------------------------------------------ */
/* Phimax: CVD (cvd.cpp:1076-1169)                                            */
/* ------------------------------------------------------------------------ */
int pfo_phimax_cvd(int L, int M, int N, float dx, float u0, const double* phi, float* pc_x,
                   float* pc_y, float* pc_z, float* pc_phi, int cap, float* phic_out,
                   float* phimax_out) {
    const int n0 = L;
    const size_t NM = (size_t)N * M;
    float phimax = -1.0f;
    for (size_t i = 0; i < (size_t)L * NM; i++)
        if (phi[i] > phimax) phimax = (float)phi[i];
    const float pii = phimax;
    const float phic = u0 + (pii - u0) * 0.618f;
    if (phic_out) *phic_out = phic;
    if (phimax_out) *phimax_out = phimax;
    float* buf = (float*)calloc(NM * (n0 + 2), sizeof(float));
    for (int l = 1; l < n0 + 1; l++)
        for (size_t j = 0; j < NM; j++) buf[j + NM * l] = (float)phi[j + NM * (l - 1)];
    memcpy(buf, buf + NM, sizeof(float) * NM);
    memcpy(buf + NM * (n0 + 1), buf + NM * n0, sizeof(float) * NM);
    const float frame_x = L * dx, frame_y = M * dx, frame_z = N * dx;
    int na = 0;
    for (int l = 1; l < n0 + 1; l++)
        for (int m = 0; m < M; m++)
            for (int n = 0; n < N; n++) {
                const float c = buf[n + (size_t)N * (m + (size_t)M * l)];
                if (!(c >= phic)) continue;
                if (!(c >= buf[n + (size_t)N * (m + (size_t)M * (l - 1))] &&
                      c >= buf[n + (size_t)N * (m + (size_t)M * (l + 1))]))
                    continue;
                const int xtemp = 0 + l;
                const long mf = (long)m - 1, mb = (long)m + 1;
                if (!(c >= buf[n + (long)N * (mf + (long)M * l)] && c >= buf[n + (long)N * (mb + (long)M * l)]))
                    continue;
                const int ytemp = m + 1;
                const long nf = (long)n - 1, nbi = (long)n + 1;
                const long base = (long)N * (m + (long)M * l);
                if (!(c >= buf[nf + base] && c >= buf[nbi + base])) continue;
                const int ztemp = n + 1;
                if (na < cap) {
                    pc_x[na] = (ytemp * dx) / frame_y;
                    pc_y[na] = (ztemp * dx) / frame_z;
                    pc_z[na] = (xtemp * dx) / frame_x;
                    pc_phi[na] = c;
                    na++;
                }
            }
    free(buf);
    return na;
}

/* ------------------------------------------------------------------------ */
/* Phimax: elastic (elastic.cpp:2809-2921), on float32 fields                 */
/* Out-of-range stencil reads (Q12: buf[-1], one past the end) are defined as */
/* "not greater".                                                             */
/* ------------------------------------------------------------------------ */
int pfo_phimax_elastic(int L, int M, int N, float dx, const float* phi, pfo_atoms* out) {
    const int n0 = L;
    const size_t NM = (size_t)N * M;
    const long total = (long)NM * (n0 + 2);
    float phimax = 0;
    for (size_t i = 0; i < (size_t)L * NM; i++)
        if (phi[i] > phimax) phimax = phi[i];
    const float pii = phimax;
    const float phic = (pii - pii * 0.40f);
    float* buf = (float*)calloc(NM * (n0 + 2), sizeof(float));
    for (int l = 1; l < n0 + 1; l++)
        for (size_t j = 0; j < NM; j++) buf[j + NM * l] = phi[j + NM * (l - 1)];
    memcpy(buf, buf + NM, sizeof(float) * NM);
    memcpy(buf + NM * (n0 + 1), buf + NM * n0, sizeof(float) * NM);
    int na = 0;
    for (int l = 1; l < n0 + 1; l++)
        for (int m = 0; m < M; m++)
            for (int n = 0; n < N; n++) {
                const float c = buf[n + (size_t)N * (m + (size_t)M * l)];
                if (!(c > phic)) continue;
                int lf = l - 1, lb = l + 1, mf = m - 1, mb = m + 1, nf = n - 1, nb = n + 1;
                if (mf < 0) mf = M - 1;
                if (mb > M - 1) mb = 0;
                if (nf < 0) nf = N - 1;
                if (nb > N - 1) nb = 0;
                int mark = 1;
                for (int lx = -1; lx < 2; lx++)
                    for (int ly = -1; ly < 2; ly++)
                        for (int lz = -1; lz < 2; lz++) {
                            int ix = l + lx, jy = m + ly, kz = n + lz;
                            if (jy > (M - 1)) jy = jy - M + 1;
                            if (kz > (N - 1)) kz = kz - N + 1;
                            if (jy < 0) jy = jy + M - 1;
                            if (kz < 0) kz = kz + N - 1;
                            const long idx = kz + (long)N * (jy + (long)M * ix);
                            if (idx < 0 || idx >= total) continue;
                            if (buf[idx] > c) mark = 0;
                        }
                if (mark != 1) continue;
                const double xtemp = (double)(0 + l), ytemp = (double)m, ztemp = (double)n;
                const double z0 = buf[n + (size_t)N * (m + (size_t)M * l)];
                const double z1 = buf[n + (size_t)N * (m + (size_t)M * lb)];
                const double z3 = buf[n + (size_t)N * (m + (size_t)M * lf)];
                const double z4 = buf[n + (size_t)N * (mb + (size_t)M * l)];
                const double z2 = buf[n + (size_t)N * (mf + (size_t)M * l)];
                const double z5 = buf[nb + (size_t)N * (m + (size_t)M * l)];
                const double z6 = buf[nf + (size_t)N * (m + (size_t)M * l)];
                const double a0 = z0;
                const double a1 = 0.5 * (z1 - z3) / dx + 10e-8;
                const double a2 = 0.5 * (z4 - z2) / dx + 10e-8;
                const double a3 = 0.5 * (z1 + z3 - 2. * z0) / (dx * dx) + 10e-8;
                const double a4 = 0.5 * (z2 + z4 - 2. * z0) / (dx * dx) + 10e-8;
                const double a5 = 0.5 * (z5 - z6) / dx + 10e-8;
                const double a6 = 0.5 * (z5 + z6 - 2. * z0) / (dx * dx) + 10e-8;
                const double delta_x = -a1 / (2. * a3);
                const double delta_y = -a2 / (2. * a4);
                const double delta_z = -a5 / (2. * a6);
                if (na < out->cap) {
                    out->x[na] = (xtemp + 1) * dx + delta_x;
                    out->y[na] = (ytemp + 1) * dx + delta_y;
                    out->z[na] = (ztemp + 1) * dx + delta_z;
                    out->phi[na] = a0 - 0.25 * a1 * a1 / a3 - 0.25 * a2 * a2 / a4 - 0.25 * a5 * a5 / a6;
                    na++;
                }
            }
    out->n = na;
    free(buf);
    return na;
}

/* ------------------------------------------------------------------------ */
/* elastic3D (elastic.cpp:3586-3729).  Terms for 1.3 <= r < benchel, which    */
/* the reference reads from uninitialised memory (Q13), are defined as 0.     */
/* ------------------------------------------------------------------------ */
static float min3f(float a, float b, float c) {
    float aa = fabsf(a), bb = fabsf(b), cc = fabsf(c);
    return ((aa < bb ? aa : bb) < cc) ? (aa < bb ? aa : bb) : cc;
}

void pfo_strain(int L, int M, int N, float dx, float benchel, int n, const double* x,
                const double* y, const double* z, double* strain, unsigned* nb) {
    const int CNmax = 16;
    const float frame_x = L * dx, frame_y = M * dx, frame_z = N * dx;
    for (int i = 0; i < n; i++) {
        unsigned Nb = 0;
        double sx = 0, sy = 0, sz = 0;
        float rjx = 0, rjy = 0, rjz = 0;
        for (int j = 0; j < n; j++) {
            if (i == j) continue;
            float r3[9];
            r3[0] = (float)(x[i] - x[j]);
            r3[1] = (float)(y[i] - y[j]);
            r3[2] = (float)(z[i] - z[j]);
            r3[3] = r3[0] + frame_x; r3[4] = r3[1] + frame_y; r3[5] = r3[2] + frame_z;
            r3[6] = r3[0] - frame_x; r3[7] = r3[1] - frame_y; r3[8] = r3[2] - frame_z;
            const float drx = min3f(fabsf(r3[0]), fabsf(r3[3]), fabsf(r3[6]));
            const float dry = min3f(fabsf(r3[1]), fabsf(r3[4]), fabsf(r3[7]));
            const float drz = min3f(fabsf(r3[2]), fabsf(r3[5]), fabsf(r3[8]));
            const float rr = sqrtf(drx * drx + dry * dry + drz * drz);
            if (!(rr < benchel && rr > 0.7)) continue;
            Nb = Nb + 1;
            if (!(Nb <= (unsigned)CNmax)) continue;
            const double rij = rr;
            for (int q = 0; q < 7; q += 3)
                if (fabsf(fabsf(r3[q]) - drx) < 1e-3) rjx = (float)(x[i] - r3[q]);
            for (int q = 1; q < 8; q += 3)
                if (fabsf(fabsf(r3[q]) - dry) < 1e-3) rjy = (float)(y[i] - r3[q]);
            for (int q = 2; q < 9; q += 3)
                if (fabsf(fabsf(r3[q]) - drz) < 1e-3) rjz = (float)(z[i] - r3[q]);
            const double cosn = (rjx - x[i]) / rij;
            const double sinn = (rjy - y[i]) / rij;
            const double than = (rjz - z[i]) / rij;
            double ex = 0.0, ey = 0.0, ez = 0.0;
            if (rij < 1.3) {
                ex = 1.0 * (rij - 1) * cosn;
                ey = 1.0 * (rij - 1) * sinn;
                ez = 1.0 * (rij - 1) * than;
            }
            sx = sx + ex; sy = sy + ey; sz = sz + ez;
        }
        strain[i] = sqrt(sx * sx + sy * sy + sz * sz);
        nb[i] = Nb;
    }
}

/* ------------------------------------------------------------------------ */
/* cryst_order (elastic.cpp:3013-3389) with the spherical harmonics of        */
/* pfc_common.h:274-346.  Float/double types as the reference: the harmonics  */
/* are float, angles and the averages double, pi the double 3.1415926.        */
/* Quirks kept: both "q4" and "q6" loops use l = 6 (:3155,3237) and the second */
/* loop keeps adding onto the first one's per-atom sum (qlmb_m is not reset), */
/* so q6 = sqrt(xishu * (S + S)) summed term by term; atoms with no or more   */
/* than CNmax neighbours get q = 0 but still count as neighbours of others.   */
/* ------------------------------------------------------------------------ */
static unsigned long pfo_facto(int h) { /* pfc_common.h:274-278 (int product) */
    int g = 1;
    for (int i = 1; i <= h; i++) g = g * i;
    return (unsigned long)g;
}

static float pfo_lg_poly(int l, float x) { /* pfc_common.h:280-291 */
    float plm2 = 1, plm1 = x, pl = 0;
    if (l == 0) return 1;
    if (l == 1) return x;
    for (int i = 2; i <= l; i++) {
        pl = ((2 * i - 1) * x * (float)plm1 - (i - 1) * plm2) / (float)i;
        plm2 = plm1;
        plm1 = pl;
    }
    return pl;
}

static float pfo_plgndr(int l, int m, float x) { /* pfc_common.h:293-320 */
    float fact, pll = 0, pmm, pmmp1, somx2;
    if (m < 0) m = -m;
    if (m == 0) return pfo_lg_poly(l, x);
    pmm = 1.0f;
    if (m > 0) {
        somx2 = (float)sqrt((1.0 - x) * (1.0 + x));
        fact = 1.0f;
        for (int i = 1; i <= m; i++) {
            pmm *= -fact * somx2;
            fact = (float)(fact + 2.0);
        }
    }
    if (l == m) return pmm;
    pmmp1 = x * (2 * m + 1) * pmm;
    if (l == m + 1) return pmmp1;
    for (int ll = m + 2; ll <= l; ll++) {
        pll = (x * (2 * ll - 1) * pmmp1 - (ll + m - 1) * pmm) / (ll - m);
        pmm = pmmp1;
        pmmp1 = pll;
    }
    return pll;
}

static float pfo_ylm_norm(int l, int m, double sgn) {
    return (float)(sgn * sqrt((2 * l + 1) / 4. / PFO_PI) *
                   (double)sqrtf((float)pfo_facto(l - m) / (float)pfo_facto(l + m)));
}

float pfo_ylm_r(int l, int m, float theta, float fai) { /* pfc_common.h:322-328 */
    if (m < 0) m = -m;
    const float fnorm = pfo_ylm_norm(l, m, pow(-1.0, m));
    const float plm = pfo_plgndr(l, m, cosf(theta));
    const float comp_m = cosf(m * fai);
    return fnorm * plm * comp_m;
}

float pfo_ylm_c(int l, int m, float theta, float fai) { /* pfc_common.h:330-344 */
    float fnorm, plm, comp_m;
    if (m < 0) {
        m = -m;
        fnorm = pfo_ylm_norm(l, m, pow(-1.0, m + m));
        plm = pfo_plgndr(l, m, cosf(theta));
        comp_m = -sinf(m * fai);
    } else {
        fnorm = pfo_ylm_norm(l, m, pow(-1.0, m));
        plm = pfo_plgndr(l, m, cosf(theta));
        comp_m = sinf(m * fai);
    }
    return fnorm * plm * comp_m;
}

void pfo_cryst_order(int L, int M, int N, float dx, int n, const double* x, const double* y, const double* z,
                     double* q4, double* q6, unsigned* nb_out) {
    const int CNmax = 16, LL = 6;
    const float bench100_h = 1.2f; /* elastic.cpp:105 */
    const float frame_x = L * dx, frame_y = M * dx, frame_z = N * dx;
    unsigned* Nb = (unsigned*)calloc((size_t)n, sizeof(unsigned));
    int* nbi = (int*)calloc((size_t)n * CNmax, sizeof(int));
    double* theta = (double*)calloc((size_t)n * CNmax, sizeof(double));
    double* fai = (double*)calloc((size_t)n * CNmax, sizeof(double));
    double* qr = (double*)calloc((size_t)n, sizeof(double));
    double* qc = (double*)calloc((size_t)n, sizeof(double));
    double* qm = (double*)calloc((size_t)n, sizeof(double));
    float rjx = 0, rjy = 0, rjz = 0;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            if (i == j) continue;
            float r3[9];
            r3[0] = (float)(x[i] - x[j]);
            r3[1] = (float)(y[i] - y[j]);
            r3[2] = (float)(z[i] - z[j]);
            r3[3] = r3[0] + frame_x; r3[4] = r3[1] + frame_y; r3[5] = r3[2] + frame_z;
            r3[6] = r3[0] - frame_x; r3[7] = r3[1] - frame_y; r3[8] = r3[2] - frame_z;
            const float drx = min3f(fabsf(r3[0]), fabsf(r3[3]), fabsf(r3[6]));
            const float dry = min3f(fabsf(r3[1]), fabsf(r3[4]), fabsf(r3[7]));
            const float drz = min3f(fabsf(r3[2]), fabsf(r3[5]), fabsf(r3[8]));
            const float rr = sqrtf(drx * drx + dry * dry + drz * drz);
            if (!(rr < bench100_h)) continue;
            Nb[i] = Nb[i] + 1;
            if (!(Nb[i] <= (unsigned)CNmax)) continue;
            for (int q = 0; q < 7; q += 3)
                if (fabsf(fabsf(r3[q]) - drx) < 1e-3) rjx = (float)(x[i] - r3[q]);
            for (int q = 1; q < 8; q += 3)
                if (fabsf(fabsf(r3[q]) - dry) < 1e-3) rjy = (float)(y[i] - r3[q]);
            for (int q = 2; q < 9; q += 3)
                if (fabsf(fabsf(r3[q]) - drz) < 1e-3) rjz = (float)(z[i] - r3[q]);
            const size_t s = Nb[i] - 1 + (size_t)CNmax * i;
            const double rij = rr;
            theta[s] = acos(((double)rjz - z[i]) / (rij + 1.e-4));
            fai[s] = atan(((double)rjy - y[i]) / ((double)rjx - x[i] + 1.e-4));
            if ((double)rjx - x[i] < 0.) fai[s] = fai[s] + PFO_PI;
            nbi[s] = j;
        }
    const double xishu = 4 * PFO_PI / (float)(2 * LL + 1);
    for (int pass = 0; pass < 2; pass++) {
        for (int m = -LL; m <= LL; m++) {
            for (int i = 0; i < n; i++) {
                qr[i] = 0.0; qc[i] = 0.0;
                if (Nb[i] > 0 && Nb[i] <= (unsigned)CNmax) {
                    for (unsigned j = 0; j < Nb[i]; j++) {
                        const size_t s = j + (size_t)CNmax * i;
                        qr[i] += pfo_ylm_r(LL, m, (float)theta[s], (float)fai[s]);
                        qc[i] += pfo_ylm_c(LL, m, (float)theta[s], (float)fai[s]);
                    }
                    qr[i] = qr[i] / (float)Nb[i];
                    qc[i] = qc[i] / (float)Nb[i];
                }
            }
            for (int i = 0; i < n; i++) {
                double br = 0.0, bc = 0.0;
                if (Nb[i] > 0 && Nb[i] <= (unsigned)CNmax) {
                    for (unsigned j = 0; j < Nb[i]; j++) {
                        const int jj = nbi[j + (size_t)CNmax * i];
                        br += qr[jj];
                        bc += qc[jj];
                    }
                    br = (qr[i] + br) / (float)(Nb[i] + 1);
                    bc = (qc[i] + bc) / (float)(Nb[i] + 1);
                }
                qm[i] += br * br + bc * bc;
            }
        }
        for (int i = 0; i < n; i++) {
            if (pass == 0) q4[i] = sqrt(xishu * qm[i]);
            else q6[i] = sqrt(xishu * qm[i]);
        }
    }
    if (nb_out)
        for (int i = 0; i < n; i++) nb_out[i] = Nb[i];
    free(Nb); free(nbi); free(theta); free(fai); free(qr); free(qc); free(qm);
}

/* q4q6_<kd>.txt (elastic.cpp:3328-3350): atoms with x >= 0. */
int pfo_write_q4q6(const char* path, const pfo_atoms* a, const double* q4, const double* q6) {
    FILE* f = fopen(path, "w");
    if (!f) return -1;
    int cnt = 0;
    for (int i = 0; i < a->n; i++)
        if (a->x[i] >= 0) cnt++;
    fprintf(f, "%d\nKommentar\n", cnt);
    for (int i = 0; i < a->n; i++)
        if (a->x[i] >= 0) fprintf(f, "qc %.4f %.4f %.4f %.4f %.4f\n", a->x[i], a->y[i], a->z[i], q4[i], q6[i]);
    fclose(f);
    return cnt;
}

/* ------------------------------------------------------------------------ */
/* CVD pair histogram and global Q6                                           */
/* ------------------------------------------------------------------------ */
#define PFO_CVD_A 1.0f   /* a_lattice (cvd.cpp:33), not a CLI parameter */
#define PFO_Q6_CNMAX 24  /* cvd.cpp:1348 */
#define PFO_Q6_NBIN 30   /* cvd.cpp:1496 */

/* delr = a_lattice/25. (cvd.cpp:1278), maxbin = int(L*dx/delr) (:1279). */
int pfo_cvd_gr_bins(int L, float dx) {
    const float delr = (float)(PFO_CVD_A / 25.);
    return (int)(L * dx / delr);
}

/* Pair distance of entries i, j as GR and Q6_GLO form it (cvd.cpp:1297-1309,
 * :1378-1390): coordinates are box fractions scaled per axis by the frame,
 * minimum image by |.| of the three shifts, 1e-6 added after the float sqrt.
 * r3 keeps the nine shifted components for the neighbour position. */
static float cvd_pair(const float* x, const float* y, const float* z, int i, int j, float fx, float fy,
                      float fz, float r3[9], float* drx, float* dry, float* drz) {
    r3[0] = (x[i] - x[j]) * fx;
    r3[1] = (y[i] - y[j]) * fy;
    r3[2] = (z[i] - z[j]) * fz;
    r3[3] = r3[0] + fx; r3[4] = r3[1] + fy; r3[5] = r3[2] + fz;
    r3[6] = r3[0] - fx; r3[7] = r3[1] - fy; r3[8] = r3[2] - fz;
    *drx = min3f(r3[0], r3[3], r3[6]);
    *dry = min3f(r3[1], r3[4], r3[7]);
    *drz = min3f(r3[2], r3[5], r3[8]);
    return (float)((double)sqrtf(*drx * *drx + *dry * *dry + *drz * *drz) + 1e-6);
}

/* GR (cvd.cpp:1277-1312): every pair i < j with rr > a/sqrt(2)/2 adds 2 to
 * bin int(rr/delr) when that bin is below maxbin. */
void pfo_cvd_gr(int L, int M, int N, float dx, int n, const float* x, const float* y, const float* z,
                unsigned* hist) {
    const float fx = L * dx, fy = M * dx, fz = N * dx;
    const float delr = (float)(PFO_CVD_A / 25.);
    const int maxbin = pfo_cvd_gr_bins(L, dx);
    for (int b = 0; b < maxbin; b++) hist[b] = 0;
    for (int i = 0; i < n - 1; i++)
        for (int j = i + 1; j < n; j++) {
            float r3[9], drx, dry, drz;
            const float rr = cvd_pair(x, y, z, i, j, fx, fy, fz, r3, &drx, &dry, &drz);
            const unsigned bin = (unsigned)(int)(rr / delr);
            if ((double)rr > PFO_CVD_A / sqrt(2.) * 0.5 && bin < (unsigned)maxbin) hist[bin] += 2;
        }
}

/* hist_<kd>.txt (cvd.cpp:1314-1322) and Grq6_<kd>.txt (:1324-1339, nstep = 1);
 * num_t counts the entries with x >= 0. */
int pfo_write_cvd_gr(int kd, int L, int M, int N, float dx, int n, const float* x, const unsigned* hist) {
    const float delr = (float)(PFO_CVD_A / 25.);
    const int maxbin = pfo_cvd_gr_bins(L, dx);
    int num_t = 0;
    for (int i = 0; i < n; i++)
        if (x[i] >= 0) num_t++;
    char name[64];
    snprintf(name, sizeof name, "hist_%d.txt", kd);
    FILE* f = fopen(name, "w");
    if (!f) return -1;
    for (int b = 0; b < maxbin; b++) fprintf(f, "%d %u %d\n", b, hist[b], maxbin);
    fclose(f);
    const float rho = (float)num_t / (float)(L * M * N * dx * dx * dx);
    const float preco = (float)(4. * PFO_PI * rho / 3.);
    snprintf(name, sizeof name, "Grq6_%d.txt", kd);
    f = fopen(name, "w");
    if (!f) return -1;
    for (int b = 0; b < maxbin; b++) {
        const float rlower = (float)b * delr;
        const float rupper = rlower + delr;
        const float nideal = preco * (rupper * rupper * rupper - rlower * rlower * rlower);
        const float gr = (float)hist[b] / (float)1 / (float)num_t / nideal;
        fprintf(f, "%.4f %.4f %u %.4f\n", b * delr, gr, hist[b], nideal);
    }
    fclose(f);
    return 0;
}

/* Q6_GLO (cvd.cpp:1347-1510).  Neighbours: a*sqrt(3)/2 * [0.9, 1.2] -/+ 1e-3,
 * at most 24 stored, counted beyond that.  The neighbour position takes,
 * per axis, every shift within 1e-5 of the minimum (the last wins); theta =
 * acos(dz/rr), fai = atan(dy/dx) (0 when dy == 0), no quadrant correction.
 * For each m: qlm = sum_j Y_6m (double sums of float terms), qlmb = qlm/Nb;
 * per entry qlmb_m += |qlmb|^2 (float); global Qlm = sum_i Nb_i qlmb_i /
 * sum_i Nb_i (float accumulators in entry order); Q = sqrt(4 pi/13 sum |Qlm|^2). */
void pfo_cvd_q6(int L, int M, int N, float dx, int n, const float* x, const float* y, const float* z,
                unsigned* nb, float* ql, float* qglobal, unsigned long long* histq) {
    const int CN = PFO_Q6_CNMAX, l = 6;
    const float fx = L * dx, fy = M * dx, fz = N * dx;
    const double rmax = PFO_CVD_A * sqrt(3.) / 2. * 1.2 + 1e-3, rmin = PFO_CVD_A * sqrt(3.) / 2. * 0.9 - 1e-3;
    float* rij = (float*)calloc((size_t)n * CN, sizeof(float));
    float* cx = (float*)calloc((size_t)n * CN, sizeof(float));
    float* cy = (float*)calloc((size_t)n * CN, sizeof(float));
    float* cz = (float*)calloc((size_t)n * CN, sizeof(float));
    float* theta = (float*)calloc((size_t)n * CN, sizeof(float));
    float* fai = (float*)calloc((size_t)n * CN, sizeof(float));
    double* qb = (double*)calloc((size_t)n * 2, sizeof(double));
    float* qbm = (float*)calloc((size_t)n, sizeof(float));
    float rjx = 0, rjy = 0, rjz = 0;
    for (int i = 0; i < n; i++) {
        nb[i] = 0;
        for (int j = 0; j < n; j++) {
            if (i == j) continue;
            float r3[9], drx, dry, drz;
            const float rr = cvd_pair(x, y, z, i, j, fx, fy, fz, r3, &drx, &dry, &drz);
            if (!((double)rr <= rmax && (double)rr >= rmin)) continue;
            nb[i]++;
            if (nb[i] > (unsigned)CN) continue;
            const size_t s = nb[i] - 1 + (size_t)CN * i;
            rij[s] = rr;
            for (int q = 0; q < 7; q += 3)
                if ((double)(fabsf(r3[q]) - drx) < 1e-5) rjx = x[i] * fx - r3[q];
            for (int q = 1; q < 8; q += 3)
                if ((double)(fabsf(r3[q]) - dry) < 1e-5) rjy = y[i] * fy - r3[q];
            for (int q = 2; q < 9; q += 3)
                if ((double)(fabsf(r3[q]) - drz) < 1e-5) rjz = z[i] * fz - r3[q];
            cx[s] = rjx; cy[s] = rjy; cz[s] = rjz;
        }
    }
    for (int i = 0; i < n; i++) {
        if (!(nb[i] > 0 && nb[i] <= (unsigned)CN)) continue;
        for (unsigned j = 0; j < nb[i]; j++) {
            const size_t s = j + (size_t)CN * i;
            theta[s] = acosf((cz[s] - z[i] * fz) / rij[s]);
            const float ddy = cy[s] - y[i] * fy;
            fai[s] = (ddy == 0.) ? 0.f : atanf(ddy / (cx[s] - x[i] * fx));
        }
    }
    float qm_glob = 0.f;
    for (int m = -l; m <= l; m++) {
        for (int i = 0; i < n; i++) {
            double re = 0.0, im = 0.0;
            qb[2 * i] = 0.0;
            qb[2 * i + 1] = 0.0;
            if (nb[i] > 0 && nb[i] <= (unsigned)CN) {
                for (unsigned j = 0; j < nb[i]; j++) {
                    const size_t s = j + (size_t)CN * i;
                    re += pfo_ylm_r(l, m, theta[s], fai[s]);
                    im += pfo_ylm_c(l, m, theta[s], fai[s]);
                }
                qb[2 * i] = re / (float)nb[i];
                qb[2 * i + 1] = im / (float)nb[i];
            }
        }
        float sum_nb = 0.f, qr = 0.f, qc = 0.f;
        for (int i = 0; i < n; i++) {
            qbm[i] = (float)((double)qbm[i] + (qb[2 * i] * qb[2 * i] + qb[2 * i + 1] * qb[2 * i + 1]));
            qr = (float)((double)qr + (double)nb[i] * qb[2 * i]);
            qc = (float)((double)qc + (double)nb[i] * qb[2 * i + 1]);
            sum_nb = sum_nb + (float)nb[i];
        }
        qr = qr / sum_nb;
        qc = qc / sum_nb;
        qm_glob = qm_glob + (qr * qr + qc * qc);
    }
    const float xishu = (float)(4 * PFO_PI / (float)(2 * l + 1));
    const float delq = 0.05f;
    *qglobal = sqrtf(xishu * qm_glob);
    for (int b = 0; b < PFO_Q6_NBIN; b++) histq[b] = 0;
    for (int i = 0; i < n; i++) {
        ql[i] = sqrtf(xishu * qbm[i]);
        const unsigned bin = (unsigned)(int)(ql[i] / delq);
        if (bin < PFO_Q6_NBIN) histq[bin]++;
    }
    free(rij); free(cx); free(cy); free(cz); free(theta); free(fai); free(qb); free(qbm);
}

/* Ql_<kd>.txt ("i Nb Ql_L Ql", 6 decimals) and HistQ_<kd>.txt ((i+1)*delq,
 * count) (cvd.cpp:1449-1510). */
int pfo_write_cvd_q6(int kd, int n, const unsigned* nb, const float* ql, float qglobal,
                     const unsigned long long* histq) {
    char name[64];
    snprintf(name, sizeof name, "Ql_%d.txt", kd);
    FILE* f = fopen(name, "w");
    if (!f) return -1;
    for (int i = 0; i < n; i++) fprintf(f, "%d %u %.6f %.6f\n", i, nb[i], ql[i], qglobal);
    fclose(f);
    snprintf(name, sizeof name, "HistQ_%d.txt", kd);
    f = fopen(name, "w");
    if (!f) return -1;
    for (int b = 0; b < PFO_Q6_NBIN; b++) fprintf(f, "%.4f %llu\n", (b + 1) * 0.05f, histq[b]);
    fclose(f);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* VTI I/O                                                                    */
/* ------------------------------------------------------------------------ */

/* sVTKwriteBianry (misfit.cpp:1616-1698), rank 0 of 1. */
int pfo_write_vti(const char* path, const char* name, int L, int M, int N, const double* field) {
    FILE* f = fopen(path, "wb");
    if (!f) return -1;
    fprintf(f, "<?xml version=\"1.0\"?>\n");
    fprintf(f, "<VTKFile type=\"ImageData\" version=\"0.1\"  byte_order=\"LittleEndian\">\n");
    fprintf(f, "<ImageData WholeExtent=\"%d %d %d %d %d %d\"   Origin=\"%d %d %d\" Spacing=\"%d %d %d\">\n",
            0, L - 1, 0, M - 1, 0, N - 1, 0, 0, 0, 1, 1, 1);
    fprintf(f, "<Piece Extent=\"%d %d %d %d %d %d\">\n", 0, L - 1, 0, M - 1, 0, N - 1);
    fprintf(f, "<PointData Scalars=\"%s\">\n", name);
    fprintf(f, "<DataArray type=\"Float32\" Name=\"%s\"   format=\"appended\" offset=\"%d\" />\n", name, 0);
    fprintf(f, "</PointData>\n");
    fprintf(f, "<CellData>\n");
    fprintf(f, "</CellData>\n");
    fprintf(f, "</Piece>\n");
    fprintf(f, "</ImageData>\n");
    fprintf(f, "<AppendedData encoding=\"raw\">\n");
    fprintf(f, "_");
    int bytes = (int)((size_t)N * M * L * sizeof(float)) + (int)sizeof(int);
    fwrite(&bytes, sizeof(int), 1, f);
    for (int n = 0; n < N; n++)
        for (int m = 0; m < M; m++)
            for (int l = 0; l < L; l++) {
                const float v = (float)field[(size_t)n + (size_t)N * ((size_t)m + (size_t)M * l)];
                fwrite(&v, sizeof(float), 1, f);
            }
    fprintf(f, "\n");
    fprintf(f, "</AppendedData>\n");
    fprintf(f, "</VTKFile>\n");
    fclose(f);
    return 0;
}

/* PVTKWRITE (misfit.cpp:1588-1614). */
int pfo_write_pvti(const char* path, const char* name, const char* prefix, int kd, int L, int M, int N) {
    FILE* f = fopen(path, "w");
    if (!f) return -1;
    fprintf(f, "<?xml version=\"1.0\"?>\n");
    fprintf(f, "<VTKFile type=\"PImageData\" version=\"0.1\" byte_order=\"LittleEndian\">\n");
    fprintf(f, "  <PImageData WholeExtent=\"0 %d 0 %d %d %d\" GhostLevel=\"#\" Origin=\"0  0  0\"  Spacing=\"1  1  1\">\n",
            L - 1, M - 1, 0, N - 1);
    fprintf(f, "    <PPointData>\n");
    fprintf(f, "      <PDataArray type=\"Float32\" Name=\"%s\"/>\n", name);
    fprintf(f, "    </PPointData>\n");
    fprintf(f, "    <Piece Extent=\"0 %d 0 %d %d %d\" Source=\"%s%d_0.vti\"/>\n", L - 1, M - 1, 0, N - 1, prefix, kd);
    fprintf(f, "  </PImageData>\n");
    fprintf(f, "</VTKFile>\n");
    fclose(f);
    return 0;
}

/* READPHIVTIN (elastic.cpp:2562-2608): skip 12 lines, one char, an int, then
 * payload (n outer, m, l inner) into index n+N*(m+M*l).  Returns values read. */
int pfo_read_vti(const char* path, int L, int M, int N, float* field) {
    FILE* f = fopen(path, "rb");
    if (!f) return -1;
    char line[100];
    for (int i = 1; i < 13; i++)
        if (!fgets(line, 100, f)) break;
    (void)fgetc(f);
    int bytes = 0;
    if (fread(&bytes, sizeof(int), 1, f) != 1) { /* ignored, as the reference */ }
    int got = 0;
    for (int n = 0; n < N; n++)
        for (int m = 0; m < M; m++)
            for (int l = 0; l < L; l++) {
                float v;
                if (fread(&v, sizeof(float), 1, f) == 1) {
                    field[(size_t)n + (size_t)N * ((size_t)m + (size_t)M * l)] = v;
                    got++;
                }
            }
    fclose(f);
    return got;
}

/* Phimax_<kd>.txt, Misfit (misfit.cpp:3495-3536). */
int pfo_write_phimax_misfit(const char* path, const pfo_atoms* a) {
    FILE* f = fopen(path, "w");
    if (!f) return -1;
    int cnt = 0;
    for (int i = 0; i < a->n; i++) {
        const double c = a->con[i];
        if (!isfinite(a->x[i]) || !isfinite(a->y[i]) || !isfinite(a->z[i]) || !isfinite(a->phi[i]) || !isfinite(c)) continue;
        if (a->x[i] < 0) continue;
        cnt++;
    }
    fprintf(f, "%d\nKommentar\n", 2 * cnt);
    for (int i = 0; i < a->n; i++) {
        const double c = a->con[i];
        if (!isfinite(a->x[i]) || !isfinite(a->y[i]) || !isfinite(a->z[i]) || !isfinite(a->phi[i]) || !isfinite(c)) continue;
        if (a->x[i] < 0) continue;
        fprintf(f, "%.4f %.4f %.4f %.4f %.4f\n", a->x[i], a->y[i], a->z[i], a->phi[i], c);
        fprintf(f, "%.4f %.4f %.4f %.4f %.4f\n", a->x[i], a->y[i], a->z[i] + 20, a->phi[i], c);
    }
    fclose(f);
    return 0;
}

/* Phimax_<kd>.txt, elastic (elastic.cpp:2953-2968). */
int pfo_write_phimax_elastic(const char* path, const pfo_atoms* a) {
    FILE* f = fopen(path, "w");
    if (!f) return -1;
    fprintf(f, "%d\nKommentar\n", a->n);
    for (int i = 0; i < a->n; i++)
        if (a->x[i] >= 0) fprintf(f, "%.4f %.4f %.4f %.4f\n", a->x[i], a->y[i], a->z[i], a->phi[i]);
    fclose(f);
    return 0;
}

/* strain_<kd>.txt (elastic.cpp:3738-3752). */
int pfo_write_strain(const char* path, const pfo_atoms* a, const double* strain, const unsigned* nb) {
    FILE* f = fopen(path, "w");
    if (!f) return -1;
    fprintf(f, "%d\nKommentar\n", a->n);
    for (int i = 0; i < a->n; i++)
        if (a->x[i] >= 0)
            fprintf(f, "st %.4f %.4f %.4f %.4f %u\n", a->x[i], a->y[i], a->z[i], strain[i], nb[i]);
    fclose(f);
    return 0;
}
