// pfc_model.cpp - host-side model constants, initial fields and spectral tables.
// Build with -ffp-contract=off: the float/double expression order below is the
// reference's and must not be fused.
#include "pfc_model.hpp"

#include <cmath>
#include <cstring>

namespace pfc {

static constexpr double kPi = 3.1415926;  // pfc_common.h:152

uint64_t splitmix64_next(uint64_t& state) {  // pfc_common.h:165-170
    uint64_t z = (state += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

double rng_u01(uint64_t& state) {  // pfc_common.h:172-176
    constexpr double inv = 1.0 / 9007199254740992.0;
    return static_cast<double>(splitmix64_next(state) >> 11) * inv;
}

void default_params(pfc_params& p, int model) {
    std::memset(&p, 0, sizeof(p));
    p.model = model;
    p.L = 256; p.M = 256; p.N = 1;
    p.precision = PFC_PREC_FP64;
    p.narrow = 1;
    p.dt = 0.05f; p.dx = 0.125f; p.con0 = 0.2f;
    p.seed = 20200604u;
    p.grainx = 16; p.grainy = 16; p.grainz = 1;
    p.axx = 0.0f;
    if (model == PFC_MODEL_MISFIT) { p.u0 = 0.05f; p.sig = 0.05f; }   // misfit.cpp:125
    else { p.u0 = 0.01f; p.sig = 0.03f; }                              // cvd.cpp:35
}

// File-static initialisers of misfit.cpp:116-141 and cvd.cpp:26-51.
Model make_model(const pfc_params& p) {
    Model m;
    std::memset(&m, 0, sizeof(m));
    m.model = p.model;
    m.L = p.L; m.M = p.M; m.N = p.N;
    m.u0 = p.u0; m.sig = p.sig; m.con0 = p.con0; m.dt = p.dt; m.dx = p.dx; m.axx = p.axx;
    m.grainx = p.grainx; m.grainy = p.grainy; m.grainz = p.grainz;
    m.seed = p.seed;
    m.a_lattice = 1.0f;
    m.gaa1 = 0.8f;
    m.gaa2 = static_cast<float>(m.gaa1 * std::sqrt(2.));
    m.gab1 = 0.8f;
    m.gab2 = static_cast<float>(m.gaa2 * std::sqrt(2.));
    m.niu = 1.4f; m.ksi = 1.f;
    m.sigMa1 = m.sigMa2 = m.sigMb1 = m.sigMb2 = 0.55f;
    m.c0 = 0.5f; m.alpha = 0.5f; m.w = 0.02f;
    m.ka0 = static_cast<float>(2.0 * kPi / m.a_lattice);
    m.ka1 = static_cast<float>(2.0 * kPi);
    m.ka2 = static_cast<float>(std::sqrt(2.0) * m.ka1);
    if (p.model == PFC_MODEL_MISFIT) {
        m.b_lattice = 0.88f;
        m.kb0 = static_cast<float>(2.0 * kPi / m.b_lattice);
        m.kb1 = m.kb0;
        m.kb2 = static_cast<float>(m.kb1 * std::sqrt(2.0));
        m.mn = 0.1f; m.mc = 10.f;
    } else {
        m.b_lattice = static_cast<float>(std::sqrt(3.0) / 2.0);
        m.kb0 = static_cast<float>(2.5 * kPi / m.b_lattice);
        m.kb1 = static_cast<float>(5.0 * kPi / std::sqrt(3.0));
        m.kb2 = static_cast<float>(m.kb1 * std::sqrt(2.0));
        m.mn = 1.f; m.mc = 1.f;
    }
    // RecomputeDerivedParams (misfit.cpp:162-179)
    m.dnx = m.grainx > 0 ? m.L / m.grainx : 0;
    m.dny = m.grainy > 0 ? m.M / m.grainy : 0;
    m.dnz = m.grainz > 0 ? m.N / m.grainz : 0;
    if (m.dx > 0.0f && m.dt > 0.0f && m.axx >= 0.0f)
        m.b0 = static_cast<float>(std::sqrt(static_cast<double>(m.axx / m.dx / m.dx / m.dx / m.dt)));
    else
        m.b0 = 0.0f;
    return m;
}

void k_tables(const Model& m, std::vector<float>& kl, std::vector<float>& km, std::vector<float>& kn) {
    auto fill = [&](std::vector<float>& k, int n) {
        k.assign(n, 0.0f);
        const double len = static_cast<double>(n * m.dx);  // int*float -> float, then promoted
        for (int i = 0; i < n / 2 + 1; i++) k[i] = static_cast<float>(i * 2.0 * kPi / len);
        for (int i = n / 2 + 1; i < n; i++) k[i] = static_cast<float>((i - n) * 2.0 * kPi / len);
    };
    fill(kl, m.L);
    fill(km, m.M);
    fill(kn, m.N);
}

// kc from the float quadratic of C2AK/C2BK (misfit.cpp:703-706).
static float crossover(float sig, float s1, float s2, float g1, float g2, float k1, float k2) {
    const float aa1 = static_cast<float>(std::pow(sig / s1, 2));
    const float aa2 = static_cast<float>(std::pow(sig / s2, 2));
    const float bb1 = static_cast<float>(1. / (2. * g1 * g1));
    const float bb2 = static_cast<float>(1. / (2. * g2 * g2));
    const float lin = bb1 * k1 - bb2 * k2;
    const float cst = (bb1 - bb2) * (aa1 - aa2 + bb1 * k1 * k1 - bb2 * k2 * k2);
    float kc = static_cast<float>((lin + std::sqrt(std::pow(lin, 2) - cst)) / (bb1 - bb2));
    if (g1 == g2) kc = (aa2 - aa1 + bb1 * (k2 * k2 - k1 * k1)) / (2 * bb1 * (k2 - k1));
    return kc;
}

static float peak(float sk, float kc, float sig, float s1, float s2, float g1, float g2, float k1, float k2) {
    if (sk <= kc) return static_cast<float>(std::exp(-std::pow(sig / s1, 2)) * std::exp(-std::pow(sk - k1, 2) / (2 * g1 * g1)));
    return static_cast<float>(std::exp(-std::pow(sig / s2, 2)) * std::exp(-std::pow(sk - k2, 2) / (2 * g2 * g2)));
}

void c2_tables_folded(const Model& m, std::vector<float>& out, float* c2a0, float* c2b0) {
    std::vector<float> kl, km, kn;
    k_tables(m, kl, km, kn);
    const float kc1 = crossover(m.sig, m.sigMa1, m.sigMa2, m.gaa1, m.gaa2, m.ka1, m.ka2);
    const float kc2 = crossover(m.sig, m.sigMb1, m.sigMb2, m.gab1, m.gab2, m.kb1, m.kb2);
    const int Lh = m.L / 2 + 1, K = m.M / 2 + 1, Nh = m.N / 2 + 1;
    out.assign(static_cast<size_t>(Lh) * K * Nh * 2, 0.0f);
    for (int il = 0; il < Lh; il++)
        for (int kc = 0; kc < K; kc++)
            for (int jn = 0; jn < Nh; jn++) {
                const float ks = kl[il] * kl[il] + km[kc] * km[kc] + kn[jn] * kn[jn];
                const float sk = std::sqrt(ks);
                const size_t i = ((static_cast<size_t>(il) * K + kc) * Nh + jn) * 2;
                out[i] = peak(sk, kc1, m.sig, m.sigMa1, m.sigMa2, m.gaa1, m.gaa2, m.ka1, m.ka2);
                out[i + 1] = peak(sk, kc2, m.sig, m.sigMb1, m.sigMb2, m.gab1, m.gab2, m.kb1, m.kb2);
            }
    if (c2a0) *c2a0 = out[0];
    if (c2b0) *c2b0 = out[1];
}

// INITIAL_dis (misfit.cpp:610-682), single rank.
static void init_misfit(const Model& m, double* phi, double* con) {
    const int L = m.L, M = m.M, N = m.N;
    const float A1 = 0.1427f, A2 = 0.106f;
    const int seed_m_max = M < 160 ? M : 160;
    const size_t P = static_cast<size_t>(L) * M * N;
    for (size_t i = 0; i < P; i++) { phi[i] = m.u0; con[i] = m.con0; }
    const float kk = (1 - m.con0) * m.ka0 + m.con0 * m.kb0;
    const double half_x = L * m.dx / m.grainx / 2. - 0.2;
    const double half_y = M * m.dx / m.grainy / 2. - 0.2;
    for (int iii = 0; iii < m.grainx; iii++)
        for (int jjj = 0; jjj < m.grainy; jjj++)
            for (int kkk = 0; kkk < m.grainz; kkk++) {
                const float center_x = m.dnx * iii + m.dnx / 2;
                const float center_y = m.dny * jjj + m.dny / 2;
                const float center_z = m.dnz * kkk + m.dnz / 2;
                const unsigned int center_sum =
                    static_cast<unsigned int>((m.dnx * iii + m.dnx / 2) + (m.dny * jjj + m.dny / 2) + (m.dnz * kkk + m.dnz / 2));
                uint64_t rng = static_cast<uint64_t>(m.seed ^ (center_sum * 2654435761u));
                const double v1 = 80.0 * (rng_u01(rng) - 0.5);
                const float theta = v1 * kPi / 180;
                const float ct = std::cos(theta);
                const double cs = std::cos(kPi / 2 - theta - (iii + jjj + kkk) * kPi);
                for (int l = 0; l < L; l++)
                    for (int mm = 0; mm < seed_m_max; mm++)
                        for (int n = 0; n < N; n++) {
                            const float xx = (0 - center_x + l) * m.dx;
                            const float yy = (mm - center_y) * m.dx;
                            const float zz = (n - center_z) * m.dx;
                            const int X = l * ct + mm * cs;
                            const int Y = mm * ct - l * cs;
                            if (std::fabs(xx) < half_x && std::fabs(yy) < half_y && std::fabs(zz) < 4.3) {
                                const float cx = std::cos(kk * X * m.dx), cy = std::cos(kk * Y * m.dx);
                                phi[n + static_cast<size_t>(N) * (mm + static_cast<size_t>(M) * l)] =
                                    2 * A1 * (cx + cy) + 4 * A2 * cx * cy + m.u0;
                            }
                        }
            }
}

// INITIAL_C2 (cvd.cpp:587-642); w3 = 0.
static void init_cvd(const Model& m, double* phi, double* con) {
    const int L = m.L, M = m.M, N = m.N;
    const float A1 = 0.15f, A2 = 0.106f;
    const float w3 = 0. * kPi / 180;
    for (int l = 0; l < L; l++)
        for (int mm = 0; mm < M; mm++)
            for (int n = 0; n < N; n++) {
                const float x = (0 - L / 2 + l) * m.dx;
                const float y = (mm - M / 2) * m.dx;
                const float k0 = mm < M / 2 ? m.ka0 : m.kb0;
                const float a = std::cos(k0 * x * std::cos(w3) + k0 * y * std::sin(w3));
                const float b = std::cos(-x * k0 * std::sin(w3) + k0 * y * std::cos(w3));
                const size_t i = n + static_cast<size_t>(N) * (mm + static_cast<size_t>(M) * l);
                phi[i] = 2 * A1 * (a + b) + 4 * A2 * a * b + m.u0;
                con[i] = mm < M / 2 ? m.con0 : 1 - m.con0;
            }
}

void init_fields(const Model& m, double* phi, double* con) {
    if (m.model == PFC_MODEL_MISFIT) init_misfit(m, phi, con);
    else init_cvd(m, phi, con);
}

}  // namespace pfc
