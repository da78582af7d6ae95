// pfc_api.cpp - C ABI of libpfc_hip.so: contexts, tables, step driver,
// energy, graphs and kernel timing.  See include/pfc_hip.h for the contract.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pfc_hip.h"
#include "pfc_kernels.h"
#include "pfc_model.hpp"
#include "pfc_post.h"

using namespace pfc;

// Failure text of the calling thread's last failed call (fixed buffer: no
// allocation on the error path).
static thread_local char g_err[512];

static int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define HIP_TRY(expr)                                                                            \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) return fail(PFC_E_HIP, "%s: %s", #expr, hipGetErrorString(e_));     \
    } while (0)

struct pfc_ctx {
    pfc_params p;
    Model m;
    int B = 1, prec = 0, L = 0, M = 0, N = 1, K = 0;
    size_t rsize = 8, vsize = 32;  // bytes per real / per Vec element
    hipStream_t stream = nullptr;
    void *S = nullptr, *AB = nullptr, *IN = nullptr;
    double* means_d = nullptr;
    std::vector<double> means_h;
    void *twH = nullptr, *twM = nullptr, *twL = nullptr, *twN = nullptr;
    float *kl_d = nullptr, *km_d = nullptr, *kn_d = nullptr, *c2_d = nullptr;
    void *coefA = nullptr, *coefC = nullptr;  // per-mode coefficients [L][N][K] + packed tail [2][L]
    int* active_d = nullptr;  // [B] step mask
    double *erow = nullptr, *ecol[2] = {nullptr, nullptr}, *eout = nullptr;
    int nrow = 0, ncol = 0;
    void *stage_a = nullptr, *stage_b = nullptr;
    void* host_stage = nullptr;  // pinned [2][L*M*N] reals for read-back, made at the first read
    long long* tcount = nullptr;  // [B] step index per instance (the noise RNG's t)
    // noise term (misfit --axx > 0)
    bool noise = false;
    unsigned* seeds_d = nullptr;
    void *nplanes = nullptr, *nz = nullptr, *nscratch = nullptr;
    int* nerr = nullptr;
    float c2a0 = 0.f, c2b0 = 0.f;
    PhysArgs phys{};
    DevBufs bufs{};
    // graphs
    int graph_chunk = 0;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    // energy trace of pfc_step_energy_enqueue: device entries, pinned host copy
    double* trace_d = nullptr;
    double* trace_h = nullptr;
    size_t trace_cap = 0;  // doubles in each
    int trace_n = 0;       // steps of the last enqueued trace
    // timing
    int time_kid = -1;
    std::vector<hipEvent_t> ev;
    size_t ev_used = 0;
};

extern "C" const char* pfc_last_error(void) { return g_err; }

// C entry points that allocate host memory return PFC_E_NOMEM instead of
// letting std::bad_alloc cross the C boundary.
#define PFC_HOST_CATCH                                                              \
    catch (const std::exception& e) {                                               \
        return fail(PFC_E_NOMEM, "host allocation failed: %s", e.what());            \
    }

namespace pfc {
int api_fail(int code, const char* msg) { return fail(code, "%s", msg); }

// Device bytes pfc_create allocates for `batch` instances (tables aside).
size_t ctx_bytes(const pfc_params& p, int batch) {
    const size_t rs = p.precision == PFC_PREC_FP64 ? 8 : 4;
    const size_t P = (size_t)p.L * p.M * p.N;
    const size_t nvec = (size_t)batch * p.L * p.N * (p.M / 2), nreal = (size_t)batch * P;
    size_t b = 3 * nvec * 4 * rs + 2 * nreal * rs;
    if (p.model == PFC_MODEL_MISFIT && p.axx > 0.0f)
        b += nreal * rs + nvec * 2 * rs + noise_scratch_bytes((long)P, batch);
    const size_t tables = ((size_t)(p.L / 2 + 1) * (p.M / 2 + 1) * (p.N / 2 + 1) * 2 + p.L + p.M + p.N) * 4 +
                          (size_t)(p.L + p.M + p.N + p.M / 2) * 2 * rs +
                          (col_coef(rs == 8, p.L) ? (size_t)p.L * p.N * (p.M / 2) * 24 : 0);
    return b + tables + (size_t)batch * 256 + (1u << 20);
}
}  // namespace pfc

extern "C" int pfc_copy_probe(size_t nbytes, int pattern, double* ms) {
    if (nbytes < 4096 || !ms || pattern < 0 || pattern > 2) return fail(PFC_E_ARG, "bad arguments");
    const size_t n = nbytes / 4096 * 256;  // whole rows of 256 words
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&a, n * 16);
    if (e == hipSuccess) e = hipMalloc(&b, n * 16);
    if (e == hipSuccess) e = hipMemset(a, 1, n * 16);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = launch_copy16(a, b, n, pattern, nullptr);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
    if (e == hipSuccess) e = launch_copy16(a, b, n, pattern, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (a) hipFree(a);
    if (b) hipFree(b);
    if (e != hipSuccess) return fail(PFC_E_HIP, "copy probe: %s", hipGetErrorString(e));
    *ms = t;
    return PFC_OK;
}

extern "C" int pfc_device_memory(size_t* free_bytes, size_t* total_bytes) {
    if (!free_bytes || !total_bytes) return fail(PFC_E_ARG, "bad arguments");
    HIP_TRY(hipMemGetInfo(free_bytes, total_bytes));
    return PFC_OK;
}

extern "C" int pfc_params_default(pfc_params* p, int model) {
    if (!p || (model != PFC_MODEL_MISFIT && model != PFC_MODEL_CVD)) return fail(PFC_E_ARG, "bad model");
    default_params(*p, model);
    return PFC_OK;
}

extern "C" int pfc_init_fields(const pfc_params* p, double* phi, double* con) try {
    if (!p || !phi || !con || p->L <= 0 || p->M <= 0 || p->N <= 0) return fail(PFC_E_ARG, "bad arguments");
    if (p->model == PFC_MODEL_MISFIT && (p->grainx <= 0 || p->grainy <= 0 || p->grainz <= 0))
        return fail(PFC_E_ARG, "grain counts must be > 0");
    Model m = make_model(*p);
    init_fields(m, phi, con);
    return PFC_OK;
} PFC_HOST_CATCH

template <typename R>
static void* upload_twiddles(int n, hipError_t* err) {
    std::vector<R> h(2 * (size_t)n);
    for (int t = 0; t < n; t++) {
        const long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)t / (long double)n;
        h[2 * t] = (R)cosl(a);
        h[2 * t + 1] = (R)sinl(a);
    }
    void* d = nullptr;
    *err = hipMalloc(&d, h.size() * sizeof(R));
    if (*err == hipSuccess) *err = hipMemcpy(d, h.data(), h.size() * sizeof(R), hipMemcpyHostToDevice);
    return d;
}

static void free_ctx(pfc_ctx* c) {
    if (!c) return;
    if (c->gexec) hipGraphExecDestroy(c->gexec);
    if (c->graph) hipGraphDestroy(c->graph);
    for (auto e : c->ev) hipEventDestroy(e);
    void* ptrs[] = {c->S, c->AB, c->IN, c->means_d, c->twH, c->twM, c->twL, c->twN, c->kl_d, c->km_d, c->kn_d,
                    c->c2_d, c->active_d, c->erow, c->ecol[0], c->ecol[1], c->eout, c->stage_a, c->stage_b,
                    c->tcount, c->seeds_d, c->nplanes, c->nz, c->nscratch, c->nerr, c->coefA, c->coefC};
    for (void* q : ptrs)
        if (q) hipFree(q);
    if (c->host_stage) hipHostFree(c->host_stage);
    if (c->trace_d) hipFree(c->trace_d);
    if (c->trace_h) hipHostFree(c->trace_h);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

// Column-tile autotuner.  The plain-step column pass is compiled for three
// tile widths (the default, twice and half as many columns per workgroup);
// all give bit-identical results, and which is fastest depends on the
// length, the precision and the batch (how many workgroups one launch has).
// Each candidate is timed on a scratch copy of the context's state; the
// choice is kept per (precision, grid, batch) for the process.
// PFC_AUTOTUNE=0 in the environment keeps the default.
namespace {
std::mutex g_tune_mu;
std::map<std::array<int, 5>, int> g_tune;
}  // namespace

static int tune_cols(pfc_ctx* c) {
    const char* env = std::getenv("PFC_AUTOTUNE");
    if (env && env[0] == '0') return 0;
    if (c->noise) return 0;  // steps with the noise term run the general kernel
    if (fs_col(c->prec, c->L)) return 0;  // the field-split column pass has one tile width
    if (col_coef(c->prec == PREC_FP64, c->L) && !c->bufs.cA) return 0;  // no plain-step kernel in use
    const std::array<int, 5> key{c->prec, c->L, c->M, c->N, c->B};
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find(key);
        if (it != g_tune.end()) return it->second;
    }
    int cand[3], nc = 0;
    for (int ts = 0; ts < 3; ts++) {
        const int t = col_tiles(c->prec, c->L, ts);
        bool dup = t <= 0 || c->K % t != 0;
        for (int i = 0; i < nc; i++) dup = dup || col_tiles(c->prec, c->L, cand[i]) == t;
        if (!dup) cand[nc++] = ts;
    }
    int best = 0;
    if (nc > 1) {
        const size_t bytes = (size_t)c->B * c->L * c->N * c->K * c->vsize;
        void *s2 = nullptr, *a2 = nullptr, *i2 = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        hipError_t e = hipMalloc(&s2, bytes);
        if (e == hipSuccess) e = hipMalloc(&a2, bytes);
        if (e == hipSuccess) e = hipMalloc(&i2, bytes);
        if (e == hipSuccess) e = hipEventCreate(&e0);
        if (e == hipSuccess) e = hipEventCreate(&e1);
        if (e == hipSuccess) e = hipMemcpyAsync(s2, c->S, bytes, hipMemcpyDeviceToDevice, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(i2, c->IN, bytes, hipMemcpyDeviceToDevice, c->stream);
        DevBufs b = c->bufs;
        b.S = s2; b.AB = a2; b.IN = i2;
        b.tcount = nullptr;
        // Rounds over the candidates in turn (so drift hits all of them), the
        // best round of each kept; after the first round the launch count is
        // raised until a candidate's run is about 0.3 ms.  A candidate
        // replaces the default only when it is at least 3% faster, so grids
        // small enough to be timed by their launch rate keep the default.
        // About 3 ms per new (precision, grid, batch).
        float tmin[3] = {1e30f, 1e30f, 1e30f};
        int reps = 5;
        for (int round = 0; round < 3 && e == hipSuccess; round++) {
            for (int k = 0; k < nc && e == hipSuccess; k++) {
                b.col_ts = cand[k];
                e = launch_col(c->prec, b, c->phys, COL_STEP, 0, c->stream);
                if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
                for (int r = 0; r < reps && e == hipSuccess; r++) e = launch_col(c->prec, b, c->phys, COL_STEP, 0, c->stream);
                if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
                if (e == hipSuccess) e = hipEventSynchronize(e1);
                float ms = 0.f;
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
                if (e == hipSuccess && round > 0 && ms / reps < tmin[k]) tmin[k] = ms / reps;
                if (e == hipSuccess && round == 0 && k == 0) {
                    const float per = ms / reps;  // the default's launch time
                    reps = per > 0.f ? (int)(0.3f / per) + 1 : 60;
                    reps = reps < 5 ? 5 : (reps > 60 ? 60 : reps);
                }
            }
        }
        int kb = 0;
        for (int k = 1; k < nc; k++)
            if (tmin[k] < tmin[kb]) kb = k;
        best = (kb != 0 && tmin[kb] < 0.97f * tmin[0]) ? cand[kb] : cand[0];
        if (e != hipSuccess) best = 0;
        hipStreamSynchronize(c->stream);
        if (e0) hipEventDestroy(e0);
        if (e1) hipEventDestroy(e1);
        for (void* q : {s2, a2, i2})
            if (q) hipFree(q);
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = best;
    return best;
}

extern "C" int pfc_create(const pfc_params* p, int batch, pfc_ctx** out) try {
    if (!p || !out || batch <= 0) return fail(PFC_E_ARG, "bad arguments");
    *out = nullptr;
    if (p->model != PFC_MODEL_MISFIT && p->model != PFC_MODEL_CVD) return fail(PFC_E_ARG, "bad model %d", p->model);
    if (p->precision != PFC_PREC_FP64 && p->precision != PFC_PREC_FP32) return fail(PFC_E_ARG, "bad precision");
    if (!size_supported(p->L) || !size_supported(p->M))
        return fail(PFC_E_UNSUPPORTED, "grid %dx%d not supported (sizes: 32,64,128,256,384,512,768,1024)", p->L, p->M);
    if (p->N != 1 && !depth_supported(p->N))
        return fail(PFC_E_UNSUPPORTED, "N = %d not supported (N: 1,16,32,64,128)", p->N);
    if (p->dt <= 0.0f || p->dx <= 0.0f) return fail(PFC_E_ARG, "dt and dx must be > 0");
    Model m = make_model(*p);
    const int prec = p->precision == PFC_PREC_FP64 ? PREC_FP64 : PREC_FP32;
    const int nrow = row_blocks(prec, p->M, p->L * p->N), ncol = col_blocks(prec, p->L, p->M, p->N);
    if (nrow <= 0 || ncol <= 0) return fail(PFC_E_UNSUPPORTED, "no launch shape for %dx%dx%d", p->L, p->M, p->N);

    // The context is freed on every failed return and on a host allocation
    // failure thrown anywhere below; released to the caller at the end.
    std::unique_ptr<pfc_ctx, void (*)(pfc_ctx*)> owner(new pfc_ctx(), free_ctx);
    pfc_ctx* c = owner.get();
    c->p = *p;
    c->m = m;
    c->B = batch;
    c->prec = prec;
    c->L = p->L; c->M = p->M; c->N = p->N; c->K = p->M / 2;  // packed: X[M/2] rides in Im X[0]
    c->rsize = prec == PREC_FP64 ? 8 : 4;
    c->vsize = 4 * c->rsize;
    c->nrow = nrow; c->ncol = ncol;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    const size_t nvec = (size_t)batch * c->L * c->N * c->K;
    const size_t nreal = (size_t)batch * c->L * c->M * c->N;
    if (e == hipSuccess) e = hipMalloc(&c->S, nvec * c->vsize);
    if (e == hipSuccess) e = hipMalloc(&c->AB, nvec * c->vsize);
    if (e == hipSuccess) e = hipMalloc(&c->IN, nvec * c->vsize);
    if (e == hipSuccess) e = hipMalloc(&c->stage_a, nreal * c->rsize);
    if (e == hipSuccess) e = hipMalloc(&c->stage_b, nreal * c->rsize);
    if (e == hipSuccess) e = hipMalloc(&c->means_d, sizeof(double) * 4 * batch);
    if (e == hipSuccess) e = hipMalloc(&c->erow, sizeof(double) * nrow * batch);
    if (e == hipSuccess) e = hipMalloc(&c->ecol[0], sizeof(double) * ncol * batch);
    if (e == hipSuccess) e = hipMalloc(&c->ecol[1], sizeof(double) * ncol * batch);
    if (e == hipSuccess) e = hipMalloc(&c->eout, sizeof(double) * 2 * batch);
    if (e == hipSuccess) e = hipMalloc(&c->tcount, sizeof(long long) * batch);
    if (e == hipSuccess) e = hipMemsetAsync(c->tcount, 0, sizeof(long long) * batch, c->stream);
    if (e == hipSuccess) e = hipMalloc(&c->active_d, sizeof(int) * batch);
    if (e == hipSuccess) {
        std::vector<int> on(batch, 1);
        e = hipMemcpy(c->active_d, on.data(), sizeof(int) * batch, hipMemcpyHostToDevice);
    }
    c->noise = (p->model == PFC_MODEL_MISFIT && m.b0 != 0.0f);
    if (c->noise) {
        std::vector<unsigned> sd(batch, p->seed);
        if (e == hipSuccess) e = hipMalloc(&c->seeds_d, sizeof(unsigned) * batch);
        if (e == hipSuccess) e = hipMemcpy(c->seeds_d, sd.data(), sizeof(unsigned) * batch, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMalloc(&c->nplanes, nreal * c->rsize);
        if (e == hipSuccess) e = hipMalloc(&c->nz, nvec * c->vsize / 2);
        if (e == hipSuccess) e = hipMalloc(&c->nscratch, noise_scratch_bytes((long)c->L * c->M * c->N, batch));
        if (e == hipSuccess) e = hipMalloc(&c->nerr, sizeof(int));
        if (e == hipSuccess) e = hipMemsetAsync(c->nerr, 0, sizeof(int), c->stream);
    }
    if (e == hipSuccess) e = hipMemsetAsync(c->S, 0, nvec * c->vsize, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(c->AB, 0, nvec * c->vsize, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(c->IN, 0, nvec * c->vsize, c->stream);
    if (e != hipSuccess) return fail(PFC_E_NOMEM, "device allocation failed: %s", hipGetErrorString(e));
    if (prec == PREC_FP64) {
        c->twH = upload_twiddles<double>(c->M / 2, &e);
        if (e == hipSuccess) c->twM = upload_twiddles<double>(c->M, &e);
        if (e == hipSuccess) c->twL = upload_twiddles<double>(c->L, &e);
        if (e == hipSuccess && c->N > 1) c->twN = upload_twiddles<double>(c->N, &e);
    } else {
        c->twH = upload_twiddles<float>(c->M / 2, &e);
        if (e == hipSuccess) c->twM = upload_twiddles<float>(c->M, &e);
        if (e == hipSuccess) c->twL = upload_twiddles<float>(c->L, &e);
        if (e == hipSuccess && c->N > 1) c->twN = upload_twiddles<float>(c->N, &e);
    }
    std::vector<float> kl, km, kn, c2;
    k_tables(m, kl, km, kn);
    c2_tables_folded(m, c2, &c->c2a0, &c->c2b0);
    if (e == hipSuccess) e = hipMalloc(&c->kl_d, sizeof(float) * kl.size());
    if (e == hipSuccess) e = hipMalloc(&c->km_d, sizeof(float) * km.size());
    if (e == hipSuccess) e = hipMalloc(&c->kn_d, sizeof(float) * kn.size());
    if (e == hipSuccess) e = hipMalloc(&c->c2_d, sizeof(float) * c2.size());
    if (e == hipSuccess) e = hipMemcpy(c->kl_d, kl.data(), sizeof(float) * kl.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->km_d, km.data(), sizeof(float) * km.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->kn_d, kn.data(), sizeof(float) * kn.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->c2_d, c2.data(), sizeof(float) * c2.size(), hipMemcpyHostToDevice);
    c->means_h.assign(4 * (size_t)batch, 0.0);
    if (e == hipSuccess) e = hipMemcpy(c->means_d, c->means_h.data(), sizeof(double) * 4 * batch, hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(PFC_E_HIP, "table upload failed: %s", hipGetErrorString(e));
    PhysArgs& ph = c->phys;
    ph.niu = m.niu; ph.ksi = m.ksi; ph.w = m.w; ph.c0 = m.c0; ph.alpha = m.alpha;
    ph.mn_dt = m.mn * m.dt;
    ph.mc_dt = m.mc * m.dt;
    ph.mc_dt_a = m.mc * m.dt * m.alpha;
    ph.narrow = p->narrow ? 1 : 0;
    ph.db = c->noise ? m.dt * m.b0 : 0.0f;
    ph.lnc0 = (float)std::log((double)m.c0);
    ph.ln1c0 = (float)std::log(1.0 - (double)m.c0);
    DevBufs& b = c->bufs;
    b.S = c->S; b.AB = c->AB; b.IN = c->IN;
    b.means = c->means_d;
    b.twH = c->twH; b.twM = c->twM; b.twL = c->twL; b.twN = c->twN;
    b.kl = c->kl_d; b.km = c->km_d; b.kn = c->kn_d; b.c2 = c->c2_d;
    b.erow = c->erow; b.ecol = c->ecol[0];
    b.NZ = c->noise ? c->nz : nullptr;
    b.tcount = c->tcount;
    b.active = c->active_d;
    b.L = c->L; b.M = c->M; b.N = c->N; b.K = c->K; b.B = batch;
    b.cA = b.cC = nullptr;
    // Plain steps read the per-mode coefficients from tables (coalesced,
    // 24 B per mode, shared by the batch) instead of gathering C2 and
    // forming k^2 per mode: faster at every batch size measured, one
    // instance included (FP32 512^2: 24.8 -> 19.4 us per step).
    if (col_coef(prec == PREC_FP64, c->L)) {
        // [L][N][K] entries, then the packed columns' k_m = M/2 modes (2 x L)
        const size_t nm = (size_t)c->L * c->N * c->K + 2 * (size_t)c->L;
        e = hipMalloc(&c->coefA, nm * 16);
        if (e == hipSuccess) e = hipMalloc(&c->coefC, nm * 8);
        if (e == hipSuccess) {
            b.cA = c->coefA;
            b.cC = c->coefC;
            e = launch_coef(prec, b, ph, c->stream);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return fail(PFC_E_NOMEM, "coefficient tables: %s", hipGetErrorString(e));
    }
    b.col_ts = 0;
    b.col_ts = tune_cols(c);
    *out = owner.release();
    return PFC_OK;
} PFC_HOST_CATCH

extern "C" int pfc_col_tiles(pfc_ctx* ctx, int* columns) {
    if (!ctx || !columns) return fail(PFC_E_ARG, "bad arguments");
    *columns = col_tiles(ctx->prec, ctx->L, ctx->bufs.col_ts);
    return PFC_OK;
}

extern "C" int pfc_set_col_tiles(pfc_ctx* ctx, int candidate) {
    if (!ctx || candidate < 0 || candidate > 2) return fail(PFC_E_ARG, "bad arguments");
    const int t = col_tiles(ctx->prec, ctx->L, candidate);
    if (t <= 0 || ctx->K % t) return fail(PFC_E_UNSUPPORTED, "tile candidate %d does not divide the columns", candidate);
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->bufs.col_ts = candidate;
    return pfc_set_graph(ctx, ctx->graph_chunk);  // recapture with the new width
}

extern "C" int pfc_destroy(pfc_ctx* ctx) {
    if (!ctx) return PFC_OK;
    hipStreamSynchronize(ctx->stream);
    free_ctx(ctx);
    return PFC_OK;
}

// Buffers of one instance only.
static DevBufs inst_bufs(const pfc_ctx* c, int inst) {
    DevBufs b = c->bufs;
    const size_t off = (size_t)inst * c->L * c->N * c->K * c->vsize;
    b.S = (char*)c->S + off;
    b.AB = (char*)c->AB + off;
    b.IN = (char*)c->IN + off;
    b.means = c->means_d + 4 * inst;
    b.erow = c->erow + (size_t)c->nrow * inst;
    b.ecol = c->ecol[0] + (size_t)c->ncol * inst;
    b.NZ = c->noise ? (char*)c->nz + off / 2 : nullptr;
    b.tcount = c->tcount + inst;
    b.active = nullptr;  // single-instance set-up and read-back run regardless of the step mask
    b.B = 1;
    return b;
}

extern "C" int pfc_set_fields(pfc_ctx* c, int inst, const double* phi, const double* con) try {
    if (!c || !phi || !con || inst < 0 || inst >= c->B) return fail(PFC_E_ARG, "bad arguments");
    const size_t P = (size_t)c->L * c->M * c->N;
    double sp = 0.0, sc = 0.0;
    for (size_t i = 0; i < P; i++) { sp += phi[i]; sc += con[i]; }
    // The reference rounds every spectral coefficient, the k=0 (mass) mode
    // included, to float each step (misfit.cpp:1002-1010); the rounded mass is
    // a fixed point of that map, so one rounding here reproduces it.
    if (c->p.narrow) { sp = (double)(float)sp; sc = (double)(float)sc; }
    const double mp = sp / (double)P, mc = sc / (double)P;
    double* mu = &c->means_h[4 * (size_t)inst];
    mu[0] = mp; mu[1] = mc;
    mu[2] = (double)c->c2a0 * mp;
    mu[3] = (double)c->c2b0 * mp;
    std::vector<char> ha(P * c->rsize), hb(P * c->rsize);
    for (size_t i = 0; i < P; i++) {
        const double dp = phi[i] - mp, dc = con[i] - mc;
        if (c->prec == PREC_FP64) { ((double*)ha.data())[i] = dp; ((double*)hb.data())[i] = dc; }
        else { ((float*)ha.data())[i] = (float)dp; ((float*)hb.data())[i] = (float)dc; }
    }
    const size_t poff = (size_t)inst * P * c->rsize;
    HIP_TRY(hipMemcpyAsync((char*)c->stage_a + poff, ha.data(), P * c->rsize, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync((char*)c->stage_b + poff, hb.data(), P * c->rsize, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->means_d + 4 * inst, mu, sizeof(double) * 4, hipMemcpyHostToDevice, c->stream));
    const long long zero = 0;
    HIP_TRY(hipMemcpyAsync(c->tcount + inst, &zero, sizeof(zero), hipMemcpyHostToDevice, c->stream));
    DevBufs b = inst_bufs(c, inst);
    HIP_TRY(launch_row_in(c->prec, b, (char*)c->stage_a + poff, (char*)c->stage_b + poff, c->stream));
    if (c->N > 1) HIP_TRY(launch_depth(c->prec, b, -1, b.S, b.S, nullptr, nullptr, false, c->stream));
    HIP_TRY(launch_col(c->prec, b, c->phys, COL_INIT, 0, c->stream));
    if (c->N > 1) HIP_TRY(launch_depth(c->prec, b, +1, b.S, b.S, b.AB, b.AB, false, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return PFC_OK;
} PFC_HOST_CATCH

static hipEvent_t next_event(pfc_ctx* c) {
    if (c->ev_used == c->ev.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->ev.push_back(e);
    }
    return c->ev[c->ev_used++];
}

// Start/stop events for one launch of the timed kernel class (none if
// another class is timed).
struct TimedLaunch {
    hipEvent_t t0 = nullptr, t1 = nullptr;
    TimedLaunch(pfc_ctx* c, int kid) {
        if (c->time_kid != kid) return;
        const size_t used = c->ev_used;
        t0 = next_event(c);
        t1 = t0 ? next_event(c) : nullptr;
        if (!t1) { t0 = nullptr; c->ev_used = used; }
    }
};

// One step on all active instances: row pass, then column pass; on 3-D
// grids the column pass sits between the two depth passes.
// etrace (energy-traced steps): where the column pass writes the energy pair
// of the state before this step, from this step's row partials and the other
// ecol slot (the previous column pass).
static hipError_t enqueue_step(pfc_ctx* c, int row_flags, int col_flags, int ecol_slot, double* etrace = nullptr) {
    DevBufs b = c->bufs;
    b.ecol = c->ecol[ecol_slot];
    if (etrace) {
        b.eprev = c->ecol[ecol_slot ^ 1];
        b.etrace = etrace;
        b.nrow = c->nrow;
    }
    hipError_t e = hipSuccess;
    if (c->noise) {
        e = launch_noise(c->prec, b, c->seeds_d, c->nplanes, c->nz, c->nscratch, c->nerr, c->stream);
        if (e != hipSuccess) return e;
    }
    TimedLaunch tr(c, KID_ROW);
    e = launch_row_step(c->prec, b, c->phys, row_flags, c->stream, tr.t0, tr.t1);
    if (e != hipSuccess) return e;
    if (c->N > 1) {
        TimedLaunch td(c, KID_DEPTH);
        e = launch_depth(c->prec, b, -1, b.S, b.S, b.IN, b.IN, false, c->stream, td.t0, td.t1);
        if (e == hipSuccess && c->noise) e = launch_depth(c->prec, b, -1, b.NZ, (void*)b.NZ, nullptr, nullptr, true, c->stream);
        if (e != hipSuccess) return e;
    }
    TimedLaunch tc(c, KID_COL);
    e = launch_col(c->prec, b, c->phys, COL_STEP, col_flags, c->stream, tc.t0, tc.t1);
    if (e != hipSuccess || c->N == 1) return e;
    TimedLaunch ti(c, KID_DEPTH);
    return launch_depth(c->prec, b, +1, b.S, b.S, b.AB, b.AB, false, c->stream, ti.t0, ti.t1);
}

extern "C" int pfc_set_graph(pfc_ctx* c, int chunk) try {
    if (!c || chunk < 0) return fail(PFC_E_ARG, "bad arguments");
    if (c->gexec) { hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
    if (c->graph) { hipGraphDestroy(c->graph); c->graph = nullptr; }
    c->graph_chunk = chunk;
    if (chunk == 0) return PFC_OK;
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int i = 0; i < chunk && e == hipSuccess; i++) e = enqueue_step(c, ROW_WRITE_IN, 0, 0);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(c->stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) hipGraphDestroy(g);
        c->graph_chunk = 0;
        return fail(PFC_E_HIP, "graph capture failed: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    }
    c->graph = g;
    HIP_TRY(hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0));
    HIP_TRY(hipGraphUpload(c->gexec, c->stream));  // not on the first launch's time
    HIP_TRY(hipStreamSynchronize(c->stream));
    return PFC_OK;
} PFC_HOST_CATCH

// Several contexts (the grid classes of an uneven sweep) as one graph: each
// context's `chunk` steps are a branch forked from and joined to the first
// context's stream, so one launch advances all of them and the device
// interleaves their kernels.
struct pfc_group {
    std::vector<pfc_ctx*> ctxs;
    int chunk = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<hipEvent_t> ev;  // fork, one join per branch, order events
};

static void free_group(pfc_group* g) {
    if (!g) return;
    if (g->gexec) hipGraphExecDestroy(g->gexec);
    if (g->graph) hipGraphDestroy(g->graph);
    for (auto e : g->ev)
        if (e) hipEventDestroy(e);
    delete g;
}

extern "C" int pfc_group_create(pfc_ctx** ctxs, int n, int chunk, pfc_group** out) try {
    if (!ctxs || n <= 0 || chunk <= 0 || !out) return fail(PFC_E_ARG, "bad arguments");
    *out = nullptr;
    for (int i = 0; i < n; i++) {
        if (!ctxs[i]) return fail(PFC_E_ARG, "null context %d", i);
        for (int j = 0; j < i; j++)
            if (ctxs[j] == ctxs[i]) return fail(PFC_E_ARG, "context %d given twice", i);
        if (ctxs[i]->time_kid >= 0) return fail(PFC_E_ARG, "kernel timing is on for context %d", i);
    }
    std::unique_ptr<pfc_group, void (*)(pfc_group*)> owner(new pfc_group(), free_group);
    pfc_group* g = owner.get();
    g->ctxs.assign(ctxs, ctxs + n);
    g->chunk = chunk;
    g->ev.assign(2 * (size_t)n + 1, nullptr);
    hipError_t e = hipSuccess;
    for (auto& ev : g->ev)
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    for (int i = 0; i < n && e == hipSuccess; i++) e = hipStreamSynchronize(ctxs[i]->stream);
    if (e != hipSuccess) return fail(PFC_E_HIP, "group set-up: %s", hipGetErrorString(e));
    // Steps are captured one at a time across the contexts, largest first
    // (points x instances) within each step: the device starts on the longest
    // chain and the shorter ones fill in beside it, instead of the longest
    // running alone at the end of every launch, and the host's submission of
    // a launch (about 0.5 ms for the benchmark sweep's 160 kernel nodes) feeds
    // every branch from its start instead of one branch after another
    // (20-step launches of that sweep 76.0-77.3 -> 78.0-78.2 G point-steps/s
    // median, back to back; 500-step regions unchanged).
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return (double)ctxs[a]->B * ctxs[a]->L * ctxs[a]->M * ctxs[a]->N >
               (double)ctxs[b]->B * ctxs[b]->L * ctxs[b]->M * ctxs[b]->N;
    });
    hipStream_t s0 = ctxs[0]->stream;
    e = hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        hipError_t ec = hipEventRecord(g->ev[0], s0);
        for (int i = 1; i < n && ec == hipSuccess; i++) ec = hipStreamWaitEvent(ctxs[i]->stream, g->ev[0], 0);
        for (int k = 0; k < chunk && ec == hipSuccess; k++)
            for (int i : order)
                if (ec == hipSuccess) ec = enqueue_step(ctxs[i], ROW_WRITE_IN, 0, 0);
        for (int i = 1; i < n && ec == hipSuccess; i++) {
            ec = hipEventRecord(g->ev[i], ctxs[i]->stream);
            if (ec == hipSuccess) ec = hipStreamWaitEvent(s0, g->ev[i], 0);
        }
        hipError_t e2 = hipStreamEndCapture(s0, &g->graph);
        e = ec != hipSuccess ? ec : e2;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&g->gexec, g->graph, nullptr, nullptr, 0);
    // Upload now, so that the first pfc_group_step does not carry it.
    if (e == hipSuccess) e = hipGraphUpload(g->gexec, ctxs[0]->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctxs[0]->stream);
    if (e != hipSuccess) return fail(PFC_E_HIP, "group graph capture: %s", hipGetErrorString(e));
    *out = owner.release();
    return PFC_OK;
} PFC_HOST_CATCH

// Advances every context of the group by nchunks * chunk steps.  Work queued
// on the member contexts' streams before the call runs first; work queued on
// them afterwards runs after the group's steps.
extern "C" int pfc_group_step(pfc_group* g, int nchunks) {
    if (!g || nchunks < 0) return fail(PFC_E_ARG, "bad arguments");
    if (nchunks == 0) return PFC_OK;
    const int n = (int)g->ctxs.size();
    hipStream_t s0 = g->ctxs[0]->stream;
    for (int i = 1; i < n; i++) {
        HIP_TRY(hipEventRecord(g->ev[n + i], g->ctxs[i]->stream));
        HIP_TRY(hipStreamWaitEvent(s0, g->ev[n + i], 0));
    }
    for (int k = 0; k < nchunks; k++) HIP_TRY(hipGraphLaunch(g->gexec, s0));
    HIP_TRY(hipEventRecord(g->ev[n], s0));
    for (int i = 1; i < n; i++) HIP_TRY(hipStreamWaitEvent(g->ctxs[i]->stream, g->ev[n], 0));
    return PFC_OK;
}

extern "C" int pfc_group_destroy(pfc_group* g) {
    if (!g) return PFC_OK;
    if (!g->ctxs.empty()) hipStreamSynchronize(g->ctxs[0]->stream);
    free_group(g);
    return PFC_OK;
}

extern "C" int pfc_step(pfc_ctx* c, int nsteps) {
    if (!c || nsteps < 0) return fail(PFC_E_ARG, "bad arguments");
    int done = 0;
    if (c->gexec && c->time_kid < 0) {
        while (nsteps - done >= c->graph_chunk) {
            HIP_TRY(hipGraphLaunch(c->gexec, c->stream));
            done += c->graph_chunk;
        }
    }
    for (; done < nsteps; done++) HIP_TRY(enqueue_step(c, ROW_WRITE_IN, 0, 0));
    return PFC_OK;
}

extern "C" int pfc_sync(pfc_ctx* c) {
    if (!c) return fail(PFC_E_ARG, "bad arguments");
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (c->noise) {
        int err = 0;
        HIP_TRY(hipMemcpy(&err, c->nerr, sizeof(int), hipMemcpyDeviceToHost));
        if (err) return fail(PFC_E_HIP, "noise stream: too few accepted Marsaglia pairs");
    }
    return PFC_OK;
}

extern "C" int pfc_set_active(pfc_ctx* c, int inst, int active) {
    if (!c || inst < 0 || inst >= c->B) return fail(PFC_E_ARG, "bad arguments");
    const int v = active ? 1 : 0;
    HIP_TRY(hipMemcpyAsync(c->active_d + inst, &v, sizeof(v), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return PFC_OK;
}

extern "C" int pfc_set_seed(pfc_ctx* c, int inst, unsigned int seed) {
    if (!c || inst < 0 || inst >= c->B) return fail(PFC_E_ARG, "bad arguments");
    if (!c->noise) return PFC_OK;
    HIP_TRY(hipMemcpyAsync(c->seeds_d + inst, &seed, sizeof(seed), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return PFC_OK;
}

template <typename R>
static void widen(const void* src, double* dst, size_t n) {
    const R* s = (const R*)src;
    for (size_t i = 0; i < n; i++) dst[i] = (double)s[i];
}

// S of instance `inst` -> real planes in stage_a/stage_b (device).
static int fields_to_stage(pfc_ctx* c, int inst) {
    const size_t P = (size_t)c->L * c->M * c->N;
    const size_t poff = (size_t)inst * P * c->rsize;
    DevBufs b = inst_bufs(c, inst);
    HIP_TRY(launch_row_out(c->prec, b, (char*)c->stage_a + poff, (char*)c->stage_b + poff, c->stream));
    return PFC_OK;
}

// Both real fields of instance `inst` into the pinned read-back buffer (one
// synchronisation); *a and *b point at them there.
static int fields_to_host(pfc_ctx* c, int inst, const void** a, const void** b) {
    int rc = fields_to_stage(c, inst);
    if (rc) return rc;
    const size_t bytes = (size_t)c->L * c->M * c->N * c->rsize;
    if (!c->host_stage) HIP_TRY(hipHostMalloc(&c->host_stage, 2 * bytes, hipHostMallocDefault));
    char* h = (char*)c->host_stage;
    const size_t poff = (size_t)inst * bytes;
    HIP_TRY(hipMemcpyAsync(h, (char*)c->stage_a + poff, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(h + bytes, (char*)c->stage_b + poff, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    *a = h;
    *b = h + bytes;
    return PFC_OK;
}

extern "C" int pfc_get_fields(pfc_ctx* c, int inst, double* phi, double* con) {
    if (!c || inst < 0 || inst >= c->B) return fail(PFC_E_ARG, "bad arguments");
    const void *ha = nullptr, *hb = nullptr;
    int rc = fields_to_host(c, inst, &ha, &hb);
    if (rc) return rc;
    const size_t P = (size_t)c->L * c->M * c->N;
    if (c->prec == PREC_FP64) {
        if (phi) std::memcpy(phi, ha, P * sizeof(double));
        if (con) std::memcpy(con, hb, P * sizeof(double));
    } else {
        if (phi) widen<float>(ha, phi, P);
        if (con) widen<float>(hb, con, P);
    }
    return PFC_OK;
}

extern "C" int pfc_get_fields_f32(pfc_ctx* c, int inst, float* phi, float* con) {
    if (!c || inst < 0 || inst >= c->B) return fail(PFC_E_ARG, "bad arguments");
    const void *ha = nullptr, *hb = nullptr;
    int rc = fields_to_host(c, inst, &ha, &hb);
    if (rc) return rc;
    const size_t P = (size_t)c->L * c->M * c->N;
    if (c->prec == PREC_FP64) {
        const double *a = (const double*)ha, *b = (const double*)hb;
        for (size_t i = 0; phi && i < P; i++) phi[i] = (float)a[i];
        for (size_t i = 0; con && i < P; i++) con[i] = (float)b[i];
    } else {
        if (phi) std::memcpy(phi, ha, P * sizeof(float));
        if (con) std::memcpy(con, hb, P * sizeof(float));
    }
    return PFC_OK;
}

extern "C" int pfc_energy(pfc_ctx* c, double* e_out) {
    if (!c || !e_out) return fail(PFC_E_ARG, "bad arguments");
    DevBufs b = c->bufs;
    b.ecol = c->ecol[0];
    b.active = nullptr;  // every instance's state has an energy
    HIP_TRY(launch_row_step(c->prec, b, c->phys, ROW_ENERGY, c->stream));
    if (c->N > 1) {
        // (m, n)-spectral copy of S in IN (free between steps); S is left untouched.
        HIP_TRY(launch_depth(c->prec, b, -1, b.S, b.IN, nullptr, nullptr, false, c->stream));
        b.S = b.IN;
    }
    HIP_TRY(launch_col(c->prec, b, c->phys, COL_ENERGY_ONLY, COL_ENERGY, c->stream));
    HIP_TRY(launch_energy_reduce(b, c->nrow, c->ncol, (double)c->m.alpha, c->eout, c->stream));
    HIP_TRY(hipMemcpyAsync(e_out, c->eout, sizeof(double) * 2 * c->B, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return PFC_OK;
}

extern "C" int pfc_step_energy_enqueue(pfc_ctx* c, int nsteps) {
    if (!c || nsteps < 0) return fail(PFC_E_ARG, "bad arguments");
    c->trace_n = nsteps;
    if (nsteps == 0) return PFC_OK;
    const size_t need = 2 * (size_t)c->B * (size_t)nsteps;
    if (need > c->trace_cap) {
        // Grown rarely (the buffers are kept): freeing waits for the device.
        HIP_TRY(hipStreamSynchronize(c->stream));
        if (c->trace_d) { hipFree(c->trace_d); c->trace_d = nullptr; }
        if (c->trace_h) { hipHostFree(c->trace_h); c->trace_h = nullptr; }
        c->trace_cap = 0;
        size_t cap = need < 4096 ? 4096 : need;
        HIP_TRY(hipMalloc(&c->trace_d, sizeof(double) * cap));
        HIP_TRY(hipHostMalloc(&c->trace_h, sizeof(double) * cap, hipHostMallocDefault));
        c->trace_cap = cap;
    }
    double* dtr = c->trace_d;
    hipError_t e = hipSuccess;
    for (int i = 0; i < nsteps && e == hipSuccess; i++) {
        // Row pass of step i also sums the pointwise energy of the state it
        // reads (after step i-1); the column pass of step i-1 left that
        // state's gradient sum in ecol[(i-1)&1], and the column pass of step
        // i reduces both into the trace entry of step i-1.
        const int slot = i & 1;
        e = enqueue_step(c, ROW_WRITE_IN | (i > 0 ? ROW_ENERGY : 0), COL_ENERGY, slot,
                         i > 0 ? dtr + 2 * c->B * (size_t)(i - 1) : nullptr);
    }
    if (e == hipSuccess) {
        DevBufs b = c->bufs;
        b.ecol = c->ecol[(nsteps - 1) & 1];
        e = launch_row_step(c->prec, b, c->phys, ROW_ENERGY, c->stream);
        if (e == hipSuccess)
            e = launch_energy_reduce(b, c->nrow, c->ncol, (double)c->m.alpha, dtr + 2 * c->B * (size_t)(nsteps - 1), c->stream);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(c->trace_h, dtr, sizeof(double) * need, hipMemcpyDeviceToHost, c->stream);
    if (e != hipSuccess) return fail(PFC_E_HIP, "step_energy: %s", hipGetErrorString(e));
    return PFC_OK;
}

extern "C" int pfc_step_energy_result(pfc_ctx* c, int nsteps, double* trace) {
    if (!c || !trace || nsteps < 0 || nsteps != c->trace_n) return fail(PFC_E_ARG, "bad arguments");
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (nsteps) std::memcpy(trace, c->trace_h, sizeof(double) * 2 * (size_t)c->B * (size_t)nsteps);
    return PFC_OK;
}

extern "C" int pfc_step_energy(pfc_ctx* c, int nsteps, double* trace) {
    if (!c || !trace || nsteps < 0) return fail(PFC_E_ARG, "bad arguments");
    int rc = pfc_step_energy_enqueue(c, nsteps);
    return rc ? rc : pfc_step_energy_result(c, nsteps, trace);
}

extern "C" int pfc_query(pfc_ctx* c, int* done) {
    if (!c || !done) return fail(PFC_E_ARG, "bad arguments");
    const hipError_t e = hipStreamQuery(c->stream);
    if (e == hipSuccess) { *done = 1; return PFC_OK; }
    if (e == hipErrorNotReady) { *done = 0; return PFC_OK; }
    return fail(PFC_E_HIP, "stream query: %s", hipGetErrorString(e));
}

extern "C" int pfc_time_kernel(pfc_ctx* c, int kernel, int enable) {
    if (!c || (kernel != KID_ROW && kernel != KID_COL && kernel != KID_DEPTH)) return fail(PFC_E_ARG, "bad arguments");
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->ev_used = 0;
    c->time_kid = enable ? kernel : -1;
    return PFC_OK;
}

extern "C" int pfc_kernel_time(pfc_ctx* c, double* mean_ms, int* count) {
    if (!c || !mean_ms || !count) return fail(PFC_E_ARG, "bad arguments");
    HIP_TRY(hipStreamSynchronize(c->stream));
    double tot = 0.0;
    int n = 0;
    for (size_t i = 0; i + 1 < c->ev_used; i += 2) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, c->ev[i], c->ev[i + 1]));
        tot += ms;
        n++;
    }
    *mean_ms = n ? tot / n : 0.0;
    *count = n;
    return PFC_OK;
}

extern "C" int pfc_kernel_bytes(pfc_ctx* c, int kernel, double* bytes) {
    if (!c || !bytes) return fail(PFC_E_ARG, "bad arguments");
    const double arr = (double)c->B * c->L * c->N * c->K * (double)c->vsize;
    if (kernel == KID_ROW) *bytes = 3.0 * arr;       // read S, AB; write IN
    else if (kernel == KID_COL) *bytes = 4.0 * arr;  // read S, IN; write S, AB
    else if (kernel == KID_DEPTH && c->N > 1) *bytes = 4.0 * arr;  // two field pairs in and out
    else return fail(PFC_E_ARG, "bad kernel id");
    return PFC_OK;
}

// ---------------------------------------------------------------------------
// Post-processing entry points (kernels in pfc_post.hip)
// ---------------------------------------------------------------------------
extern "C" int pfc_phimax_misfit(pfc_ctx* c, int inst, double* x, double* y, double* z, double* phi, float* con,
                                 int cap, int* n) {
    if (!c || inst < 0 || inst >= c->B || !n || cap < 0) return fail(PFC_E_ARG, "bad arguments");
    int rc = fields_to_stage(c, inst);
    if (rc) return rc;
    const size_t P = (size_t)c->L * c->M * c->N;
    const size_t poff = (size_t)inst * P * c->rsize;
    PostResult r;
    hipError_t e = post_phimax_misfit(c->prec, c->L, c->M, c->m.dx, c->m.con0, (char*)c->stage_a + poff,
                                      (char*)c->stage_b + poff, x, y, z, phi, con, cap, &r, c->stream);
    if (e != hipSuccess) return fail(PFC_E_HIP, "phimax_misfit: %s", hipGetErrorString(e));
    *n = r.n;
    if (r.n > cap) return fail(PFC_E_CAP, "capacity %d < %d atoms", cap, r.n);
    return PFC_OK;
}

extern "C" int pfc_phimax_cvd(pfc_ctx* c, int inst, float* pc_x, float* pc_y, float* pc_z, float* pc_phi, int cap,
                              int* n, float* phic, float* phimax) {
    if (!c || inst < 0 || inst >= c->B || !n || cap < 0) return fail(PFC_E_ARG, "bad arguments");
    int rc = fields_to_stage(c, inst);
    if (rc) return rc;
    const size_t P = (size_t)c->L * c->M * c->N;
    const size_t poff = (size_t)inst * P * c->rsize;
    PostResult r;
    hipError_t e = post_phimax_cvd(c->prec, c->L, c->M, c->N, c->m.dx, c->m.u0, (char*)c->stage_a + poff, pc_x, pc_y, pc_z,
                                   pc_phi, cap, &r, c->stream);
    if (e != hipSuccess) return fail(PFC_E_HIP, "phimax_cvd: %s", hipGetErrorString(e));
    *n = r.n;
    if (phic) *phic = r.phic;
    if (phimax) *phimax = r.phimax;
    if (r.n > cap) return fail(PFC_E_CAP, "capacity %d < %d atoms", cap, r.n);
    return PFC_OK;
}

extern "C" int pfc_phimax_elastic(int L, int M, int N, float dx, const float* phi, double* x, double* y, double* z,
                                  double* phimax, int cap, int* n) {
    if (L <= 0 || M <= 0 || N <= 0 || !phi || !n || cap < 0) return fail(PFC_E_ARG, "bad arguments");
    PostResult r;
    hipError_t e = post_phimax_elastic(L, M, N, dx, phi, x, y, z, phimax, cap, &r);
    if (e != hipSuccess) return fail(PFC_E_HIP, "phimax_elastic: %s", hipGetErrorString(e));
    *n = r.n;
    if (r.n > cap) return fail(PFC_E_CAP, "capacity %d < %d atoms", cap, r.n);
    return PFC_OK;
}

extern "C" int pfc_phimax_elastic_noalloc(int L, int M, int N, float dx, const float* phi, double* x, double* y,
                                          double* z, double* phimax, int cap, int* n, void* stream) {
    if (L <= 0 || M <= 0 || N <= 0 || !phi || !n || cap < 0 || (cap > 0 && (!x || !y || !z || !phimax)))
        return fail(PFC_E_ARG, "bad arguments");
    const long P = (long)L * M * N;
    if (P > NOALLOC_MAX_POINTS)
        return fail(PFC_E_UNSUPPORTED, "field of %ld points exceeds the allocation-free scratch (at most %ld points)", P,
                    NOALLOC_MAX_POINTS);
    hipError_t e = post_phimax_elastic_noalloc(L, M, N, dx, phi, x, y, z, phimax, cap, n, (hipStream_t)stream);
    if (e != hipSuccess) return fail(PFC_E_HIP, "phimax_elastic_noalloc: %s", hipGetErrorString(e));
    return PFC_OK;
}

extern "C" int pfc_cryst_order(int L, int M, int N, float dx, int n, const double* x, const double* y,
                               const double* z, double* q4, double* q6, unsigned int* nb) {
    if (L <= 0 || M <= 0 || N <= 0 || n < 0 || (n > 0 && (!x || !y || !z || !q4 || !q6)))
        return fail(PFC_E_ARG, "bad arguments");
    if (n == 0) return PFC_OK;
    hipError_t e = post_cryst(L, M, N, dx, n, x, y, z, q4, q6, nb);
    if (e != hipSuccess) return fail(PFC_E_HIP, "cryst_order: %s", hipGetErrorString(e));
    return PFC_OK;
}

extern "C" int pfc_cvd_gr(int L, int M, int N, float dx, int n, const float* pc_x, const float* pc_y,
                          const float* pc_z, unsigned int* hist, int* maxbin) {
    if (L <= 0 || M <= 0 || N <= 0 || !(dx > 0.0f) || n < 0 || !maxbin) return fail(PFC_E_ARG, "bad arguments");
    *maxbin = cvd_gr_bins(L, dx);
    if (!hist) return PFC_OK;
    if (n > 0 && (!pc_x || !pc_y || !pc_z)) return fail(PFC_E_ARG, "bad arguments");
    hipError_t e = post_cvd_gr(L, M, N, dx, n, pc_x, pc_y, pc_z, hist);
    if (e != hipSuccess) return fail(PFC_E_HIP, "cvd_gr: %s", hipGetErrorString(e));
    return PFC_OK;
}

extern "C" int pfc_cvd_q6(int L, int M, int N, float dx, int n, const float* pc_x, const float* pc_y,
                          const float* pc_z, unsigned int* nb, float* ql, float* qglobal,
                          unsigned long long* histq) {
    if (L <= 0 || M <= 0 || N <= 0 || n <= 0 || !pc_x || !pc_y || !pc_z || !nb || !ql || !qglobal || !histq)
        return fail(PFC_E_ARG, "bad arguments");
    hipError_t e = post_cvd_q6(L, M, N, dx, n, pc_x, pc_y, pc_z, nb, ql, qglobal, histq);
    if (e != hipSuccess) return fail(PFC_E_HIP, "cvd_q6: %s", hipGetErrorString(e));
    return PFC_OK;
}

extern "C" int pfc_strain(int L, int M, int N, float dx, float benchel, int n, const double* x, const double* y,
                          const double* z, double* strain, unsigned int* nb) {
    if (L <= 0 || M <= 0 || N <= 0 || n < 0 || (n > 0 && (!x || !y || !z || !strain || !nb)))
        return fail(PFC_E_ARG, "bad arguments");
    if (n == 0) return PFC_OK;
    hipError_t e = post_strain(L, M, N, dx, benchel, n, x, y, z, strain, nb);
    if (e != hipSuccess) return fail(PFC_E_HIP, "strain: %s", hipGetErrorString(e));
    return PFC_OK;
}
