// pfc_post.hip - Phimax (atom extraction) and strain kernels.
//
// The three Phimax variants of the reference differ in threshold, neighbour
// test and output order; each is a per-point predicate over the field plus an
// ordered stream compaction (count per block, exclusive scan of the counts,
// emit with a ballot prefix inside each block), so the atoms come out in the
// reference's loop order.  The halo semantics are the single-rank MPI stub's:
// the x neighbour of the first/last row is that row itself (pfc_common.h:
// 101-110).  Arithmetic on positions follows the reference's float/double
// expression types and is compiled without contraction.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <utility>
#include <vector>

#include "pfc_kernels.h"
#include "pfc_post.h"

namespace pfc {

constexpr int PB = 256;  // threads per compaction block

// Position of this thread's flag among the block's set flags (block order =
// thread order).  All threads must call.
__device__ __forceinline__ int block_rank(bool flag, int* wsum, int& total) {
    const unsigned long long bal = __ballot(flag);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int below = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[wid] = __popcll(bal);
    __syncthreads();
    int off = 0;
    total = 0;
    for (int w = 0; w < PB / 64; w++) {
        if (w < wid) off += wsum[w];
        total += wsum[w];
    }
    __syncthreads();
    return off + below;
}

// ---------------------------------------------------------------------------
// Predicates / emitters
// ---------------------------------------------------------------------------

// Misfit (misfit.cpp:3345-3443): scan order p = i + L*m (l fastest).
template <typename R>
struct MisfitOp {
    const R* phi;
    const R* con;
    int L, M;
    float dx, con0;
    double* x; double* y; double* z; double* ph; float* co;

    __device__ double v(int i, int m) const { return (double)phi[(size_t)i * M + m]; }
    __device__ double nbv(int i, int m) const {
        const double t = v(i, m);
        return isfinite(t) ? t : -INFINITY;
    }
    __device__ bool test(long p) const {
        const int i = (int)(p % L), m = (int)(p / L);
        const double center = v(i, m);
        if (!isfinite(center) || center <= -0.07) return false;
        const int lf = i > 0 ? i - 1 : 0, lb = i < L - 1 ? i + 1 : L - 1;
        const int mf = m > 0 ? m - 1 : M - 1, mb = m < M - 1 ? m + 1 : 0;
        const double nb[8] = {nbv(lf, m), nbv(lb, m), nbv(i, mf), nbv(i, mb),
                              nbv(lb, mb), nbv(lf, mb), nbv(lb, mf), nbv(lf, mf)};
        bool ge_all = true, gt_any = false;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            if (center < nb[q]) ge_all = false;
            if (center > nb[q]) gt_any = true;
        }
        return ge_all && gt_any;
    }
    __device__ void emit(long p, size_t o) const {
#pragma clang fp contract(off)
        const int i = (int)(p % L), m = (int)(p / L);
        const int lf = i > 0 ? i - 1 : 0, lb = i < L - 1 ? i + 1 : L - 1;
        const int mf = m > 0 ? m - 1 : M - 1, mb = m < M - 1 ? m + 1 : 0;
        const double xtemp = (double)((float)(i + 2) * dx);
        const double ytemp = (double)((float)(m + 1) * dx);
        const double ztemp = (double)((float)1 * dx);
        const double z0 = v(i, m), z1 = v(lb, m), z3 = v(lf, m), z4 = v(i, mb), z2 = v(i, mf);
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
            delta_x = delta_x < -half_dx ? -half_dx : (delta_x > half_dx ? half_dx : delta_x);
            delta_y = delta_y < -half_dx ? -half_dx : (delta_y > half_dx ? half_dx : delta_y);
            double phi_interp = z0;
            if (isfinite(a3) && fabs(a3) > eps && isfinite(a1)) phi_interp -= 0.25 * a1 * a1 / a3;
            if (isfinite(a4) && fabs(a4) > eps && isfinite(a2)) phi_interp -= 0.25 * a2 * a2 / a4;
            if (isfinite(phi_interp)) phi_local = phi_interp;
        }
        x[o] = xtemp + delta_x;
        y[o] = ytemp + delta_y;
        z[o] = ztemp;
        ph[o] = phi_local;
        const double c = (double)con[(size_t)i * M + m];
        co[o] = isfinite(c) ? (float)c : con0;
    }
};

// CVD (cvd.cpp:1125-1158): memory order p = i*M + m; float buffer; the m and
// n neighbours are not wrapped (for N = 1 the n test repeats the m test).
template <typename R>
struct CvdOp {
    const R* phi;
    int L, M, N;
    float dx, phic, fx, fy, fz;
    float* px; float* py; float* pz; float* pp;

    // Flat index F of the reference's halo buffer (rows 0..L+1 of N*M floats,
    // rows 0 and L+1 replicate the edge rows); F - 1 at n = 0 and F - N at
    // m = 0 step into the previous m or l, as the reference's index arithmetic.
    __device__ float bufv(long F) const {
        const long NM = (long)N * M;
        const long r = F / NM, rem = F - r * NM;
        const long fr = r == 0 ? 0 : (r == L + 1 ? L - 1 : r - 1);
        return (float)phi[fr * NM + rem];
    }
    __device__ bool test(long p) const {
        const long NM = (long)N * M;
        const long F = p + NM;  // field row l sits in halo row l + 1
        const float c = bufv(F);
        if (!(c >= phic)) return false;
        if (!(c >= bufv(F - NM) && c >= bufv(F + NM))) return false;
        if (!(c >= bufv(F - N) && c >= bufv(F + N))) return false;
        return c >= bufv(F - 1) && c >= bufv(F + 1);
    }
    __device__ void emit(long p, size_t o) const {
#pragma clang fp contract(off)
        const long NM = (long)N * M;
        const int i = (int)(p / NM);
        const long rem = p - (long)i * NM;
        const int m = (int)(rem / N), n = (int)(rem % N);
        px[o] = ((float)(m + 1) * dx) / fy;
        py[o] = ((float)(n + 1) * dx) / fz;
        pz[o] = ((float)(i + 1) * dx) / fx;
        pp[o] = bufv(p + NM);
    }
};

// Elastic (elastic.cpp:2853-2917): memory order p = n + N*(m + M*i); float
// field; 3x3x3 "none strictly greater" test with the reference's wrap
// arithmetic; stencil reads outside the halo buffer count as not greater.
struct ElasticOp {
    const float* phi;
    int L, M, N;
    float dx, phic;
    double* x; double* y; double* z; double* ph;

    // Halo buffer row `bix` (0..L+1) at in-row offset `rem`.
    __device__ float bufv(long bix, long rem) const {
        const long NM = (long)N * M;
        const long fr = bix == 0 ? 0 : (bix == L + 1 ? L - 1 : bix - 1);
        return phi[fr * NM + rem];
    }
    __device__ bool test(long p) const {
        const long NM = (long)N * M;
        const int i = (int)(p / NM);
        const long rem0 = p % NM;
        const int m = (int)(rem0 / N), n = (int)(rem0 % N);
        const float c = phi[p];
        if (!(c > phic)) return false;
        const long l = i + 1, total = NM * (L + 2);
        bool mark = true;
        for (int lx = -1; lx < 2; lx++)
            for (int ly = -1; ly < 2; ly++)
                for (int lz = -1; lz < 2; lz++) {
                    const long ix = l + lx;
                    long jy = m + ly, kz = n + lz;
                    if (jy > (M - 1)) jy = jy - M + 1;
                    if (kz > (N - 1)) kz = kz - N + 1;
                    if (jy < 0) jy = jy + M - 1;
                    if (kz < 0) kz = kz + N - 1;
                    const long idx = kz + (long)N * (jy + (long)M * ix);
                    if (idx < 0 || idx >= total) continue;
                    if (bufv(idx / NM, idx % NM) > c) mark = false;
                }
        return mark;
    }
    __device__ void emit(long p, size_t o) const {
#pragma clang fp contract(off)
        const long NM = (long)N * M;
        const int i = (int)(p / NM);
        const long rem0 = p % NM;
        const int m = (int)(rem0 / N), n = (int)(rem0 % N);
        const long l = i + 1;
        const int mf = m - 1 < 0 ? M - 1 : m - 1, mb = m + 1 > M - 1 ? 0 : m + 1;
        const int nf = n - 1 < 0 ? N - 1 : n - 1, nb = n + 1 > N - 1 ? 0 : n + 1;
        const double xtemp = (double)l, ytemp = (double)m, ztemp = (double)n;
        const double z0 = bufv(l, n + (long)N * m);
        const double z1 = bufv(l + 1, n + (long)N * m);
        const double z3 = bufv(l - 1, n + (long)N * m);
        const double z4 = bufv(l, n + (long)N * mb);
        const double z2 = bufv(l, n + (long)N * mf);
        const double z5 = bufv(l, nb + (long)N * m);
        const double z6 = bufv(l, nf + (long)N * m);
        const float dx2 = dx * dx;
        const double a0 = z0;
        const double a1 = 0.5 * (z1 - z3) / (double)dx + 10e-8;
        const double a2 = 0.5 * (z4 - z2) / (double)dx + 10e-8;
        const double a3 = 0.5 * (z1 + z3 - 2. * z0) / (double)dx2 + 10e-8;
        const double a4 = 0.5 * (z2 + z4 - 2. * z0) / (double)dx2 + 10e-8;
        const double a5 = 0.5 * (z5 - z6) / (double)dx + 10e-8;
        const double a6 = 0.5 * (z5 + z6 - 2. * z0) / (double)dx2 + 10e-8;
        const double delta_x = -a1 / (2. * a3);
        const double delta_y = -a2 / (2. * a4);
        const double delta_z = -a5 / (2. * a6);
        x[o] = (xtemp + 1) * (double)dx + delta_x;
        y[o] = (ytemp + 1) * (double)dx + delta_y;
        z[o] = (ztemp + 1) * (double)dx + delta_z;
        ph[o] = a0 - 0.25 * a1 * a1 / a3 - 0.25 * a2 * a2 / a4 - 0.25 * a5 * a5 / a6;
    }
};

template <typename Op>
__global__ void __launch_bounds__(PB) k_count(Op op, long P, int* counts) {
    __shared__ int wsum[PB / 64];
    const long p = (long)blockIdx.x * PB + threadIdx.x;
    const bool f = p < P && op.test(p);
    int total;
    (void)block_rank(f, wsum, total);
    if (threadIdx.x == 0) counts[blockIdx.x] = total;
}

// Exclusive scan of nb counts by one workgroup per blockIdx.y batch;
// offsets[nb] = total.  Batch y uses counts + y*nb and offsets + y*(nb+1).
__global__ void __launch_bounds__(1024) k_scan(const int* counts, int nb, int* offsets) {
    __shared__ int part[1024];
    counts += (size_t)blockIdx.y * nb;
    offsets += (size_t)blockIdx.y * (nb + 1);
    const int t = threadIdx.x;
    const int per = (nb + 1023) / 1024;
    const int b0 = t * per, b1 = min(nb, b0 + per);
    int s = 0;
    for (int i = b0; i < b1; i++) s += counts[i];
    part[t] = s;
    __syncthreads();
    if (t == 0) {
        int acc = 0;
        for (int i = 0; i < 1024; i++) { const int v = part[i]; part[i] = acc; acc += v; }
        offsets[nb] = acc;
    }
    __syncthreads();
    int acc = part[t];
    for (int i = b0; i < b1; i++) { offsets[i] = acc; acc += counts[i]; }
}

template <typename Op>
__global__ void __launch_bounds__(PB) k_emit(Op op, long P, const int* offsets, int cap) {
    __shared__ int wsum[PB / 64];
    const long p = (long)blockIdx.x * PB + threadIdx.x;
    const bool f = p < P && op.test(p);
    int total;
    const int r = block_rank(f, wsum, total);
    const long o = (long)offsets[blockIdx.x] + r;
    if (f && o < cap) op.emit(p, (size_t)o);
}

__global__ void k_max(const void* phi, int prec, long P, double* out) {
    __shared__ double red[256];
    double m = -INFINITY;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < P; i += (long)gridDim.x * blockDim.x) {
        const double v = prec == PREC_FP64 ? ((const double*)phi)[i] : (double)((const float*)phi)[i];
        m = v > m ? v : m;
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] > red[threadIdx.x + s] ? red[threadIdx.x] : red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// Device scratch of the post-processing calls, kept per host thread: a call
// takes blocks from its thread's pool (hipMalloc when none fits) and gives
// them back when it ends, after the stream it used has finished with them.
// hipFree waits for the whole device, which would tie a checkpoint's
// post-processing on one thread to the stepping another thread has queued.
// A block is reused for requests of half its size or more, on the device it
// was allocated on; the pool keeps at most 1 GiB (blocks beyond are freed).
struct PoolBlock {
    void* p;
    size_t n;
    int dev;
};
struct ThreadPool {
    std::vector<PoolBlock> blocks;
    size_t bytes = 0;
    ~ThreadPool() {
        for (const PoolBlock& b : blocks) hipFree(b.p);
    }
};
static thread_local ThreadPool t_pool;

struct Scratch {
    hipStream_t s;
    std::vector<PoolBlock> p;
    explicit Scratch(hipStream_t st = nullptr) : s(st) {}
    ~Scratch() {
        if (p.empty()) return;
        hipStreamSynchronize(s);
        for (const PoolBlock& b : p) {
            if (t_pool.bytes + b.n > (size_t(1) << 30)) {
                hipFree(b.p);
                continue;
            }
            t_pool.blocks.push_back(b);
            t_pool.bytes += b.n;
        }
    }
    template <typename T>
    T* get(size_t n, hipError_t& e) {
        if (e != hipSuccess) return nullptr;
        const size_t want = n * sizeof(T) + 16;
        int dev = 0;
        if ((e = hipGetDevice(&dev)) != hipSuccess) return nullptr;
        // Smallest pooled block of this device that fits and is at most twice the request.
        int best = -1;
        for (int i = 0; i < (int)t_pool.blocks.size(); i++) {
            const PoolBlock& c = t_pool.blocks[i];
            if (c.dev == dev && c.n >= want && c.n / 2 <= want && (best < 0 || c.n < t_pool.blocks[best].n)) best = i;
        }
        PoolBlock b{nullptr, want, dev};
        if (best >= 0) {
            b = t_pool.blocks[best];
            t_pool.blocks.erase(t_pool.blocks.begin() + best);
            t_pool.bytes -= b.n;
        } else if ((e = hipMalloc(&b.p, want)) != hipSuccess) {
            return nullptr;
        }
        p.push_back(b);
        return (T*)b.p;
    }
};

// The calling thread's stream for post-processing of host arrays, one per
// device (created on first use, non-blocking): it runs beside a context's
// stepping queued by another thread.
struct ThreadStreams {
    std::vector<std::pair<int, hipStream_t>> s;
    ~ThreadStreams() {
        for (auto& d : s) hipStreamDestroy(d.second);
    }
};
static thread_local ThreadStreams t_streams;
static hipError_t post_stream(hipStream_t* out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    for (auto& d : t_streams.s)
        if (d.first == dev) {
            *out = d.second;
            return hipSuccess;
        }
    hipStream_t st = nullptr;
    if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess) return e;
    t_streams.s.emplace_back(dev, st);
    *out = st;
    return hipSuccess;
}

template <typename T>
static hipError_t h2d(T* dst, const T* src, size_t n, hipStream_t s) {
    return hipMemcpyAsync(dst, src, sizeof(T) * n, hipMemcpyHostToDevice, s);
}

template <typename Op>
static hipError_t compact(Op op, long P, int cap, int* n_out, hipStream_t s) {
    hipError_t e = hipSuccess;
    Scratch sc(s);
    const int nb = (int)((P + PB - 1) / PB);
    int* counts = sc.get<int>(nb, e);
    int* offs = sc.get<int>(nb + 1, e);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_count<Op>, dim3(nb), dim3(PB), 0, s, op, P, counts);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, counts, nb, offs);
    hipLaunchKernelGGL(k_emit<Op>, dim3(nb), dim3(PB), 0, s, op, P, offs, cap);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    int tot = 0;
    if ((e = hipMemcpyAsync(&tot, offs + nb, sizeof(int), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    *n_out = tot;
    return hipSuccess;
}

template <typename T>
static hipError_t d2h(T* dst, const T* src, int n, hipStream_t s) {
    if (n <= 0 || !dst) return hipSuccess;
    hipError_t e = hipMemcpyAsync(dst, src, sizeof(T) * n, hipMemcpyDeviceToHost, s);
    return e;
}

template <typename R>
static hipError_t misfit_t(int L, int M, float dx, float con0, const void* phi, const void* con, double* x,
                           double* y, double* z, double* ph, float* co, int cap, PostResult* r, hipStream_t s) {
    const long P = (long)L * M;
    hipError_t e = hipSuccess;
    Scratch sc(s);
    MisfitOp<R> op;
    op.phi = (const R*)phi; op.con = (const R*)con; op.L = L; op.M = M; op.dx = dx; op.con0 = con0;
    op.x = sc.get<double>(P, e); op.y = sc.get<double>(P, e); op.z = sc.get<double>(P, e);
    op.ph = sc.get<double>(P, e); op.co = sc.get<float>(P, e);
    if (e != hipSuccess) return e;
    int n = 0;
    if ((e = compact(op, P, (int)P, &n, s)) != hipSuccess) return e;
    if (n == 0) {
        // Fallback of misfit.cpp:3449-3464: last position of the global max in
        // memory order, computed on the host copy of the fields.
        std::vector<R> hp(P), hc(P);
        if ((e = hipMemcpyAsync(hp.data(), phi, sizeof(R) * P, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(hc.data(), con, sizeof(R) * P, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
        if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
        double gmax = -INFINITY;
        int gm = 0, gl = 0;
        for (int i = 0; i < L; i++)
            for (int m = 0; m < M; m++) {
                const double v = (double)hp[(size_t)i * M + m];
                if (!std::isfinite(v)) continue;
                gmax = gmax > v ? gmax : v;
                if (v >= gmax) { gm = m; gl = i; }
            }
        r->n = 1;
        if (cap >= 1) {
            const double dxd = (double)dx;
            x[0] = (gl + 2) * dxd; y[0] = (gm + 1) * dxd; z[0] = dxd;
            ph[0] = std::isfinite(gmax) ? gmax : 0.0;
            const double c = (double)hc[(size_t)gl * M + gm];
            co[0] = std::isfinite(c) ? (float)c : con0;
        }
        return hipSuccess;
    }
    r->n = n;
    const int k = n < cap ? n : cap;
    if ((e = d2h(x, op.x, k, s)) != hipSuccess) return e;
    if ((e = d2h(y, op.y, k, s)) != hipSuccess) return e;
    if ((e = d2h(z, op.z, k, s)) != hipSuccess) return e;
    if ((e = d2h(ph, op.ph, k, s)) != hipSuccess) return e;
    if ((e = d2h(co, op.co, k, s)) != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

hipError_t post_phimax_misfit(int prec, int L, int M, float dx, float con0, const void* phi, const void* con,
                              double* x, double* y, double* z, double* ph, float* co, int cap, PostResult* r,
                              hipStream_t s) {
    return prec == PREC_FP64 ? misfit_t<double>(L, M, dx, con0, phi, con, x, y, z, ph, co, cap, r, s)
                             : misfit_t<float>(L, M, dx, con0, phi, con, x, y, z, ph, co, cap, r, s);
}

static hipError_t field_max(const void* phi, int prec, long P, double* out, hipStream_t s) {
    hipError_t e = hipSuccess;
    Scratch sc(s);
    const int nb = 256;
    double* part = sc.get<double>(nb, e);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_max, dim3(nb), dim3(256), 0, s, phi, prec, P, part);
    std::vector<double> h(nb);
    if ((e = hipMemcpyAsync(h.data(), part, sizeof(double) * nb, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    double m = -INFINITY;
    for (double v : h) m = v > m ? v : m;
    *out = m;
    return hipSuccess;
}

template <typename R>
static hipError_t cvd_t(int L, int M, int N, float dx, float u0, const void* phi, float* px, float* py, float* pz,
                        float* pp, int cap, PostResult* r, hipStream_t s) {
    const long P = (long)L * M * N;
    double gmax = 0.0;
    hipError_t e = field_max(phi, sizeof(R) == 8 ? PREC_FP64 : PREC_FP32, P, &gmax, s);
    if (e != hipSuccess) return e;
    // cvd.cpp:1077-1084: running float max starting at phimm = -1.
    float pmax = -1.0f;
    if (gmax > (double)pmax) pmax = (float)gmax;
    const float phic = u0 + (pmax - u0) * 0.618f;
    r->phic = phic;
    r->phimax = pmax;
    Scratch sc(s);
    CvdOp<R> op;
    op.phi = (const R*)phi; op.L = L; op.M = M; op.N = N; op.dx = dx; op.phic = phic;
    op.fx = L * dx; op.fy = M * dx; op.fz = N * dx;
    op.px = sc.get<float>(P, e); op.py = sc.get<float>(P, e); op.pz = sc.get<float>(P, e); op.pp = sc.get<float>(P, e);
    if (e != hipSuccess) return e;
    int n = 0;
    if ((e = compact(op, P, (int)P, &n, s)) != hipSuccess) return e;
    r->n = n;
    const int k = n < cap ? n : cap;
    if ((e = d2h(px, op.px, k, s)) != hipSuccess) return e;
    if ((e = d2h(py, op.py, k, s)) != hipSuccess) return e;
    if ((e = d2h(pz, op.pz, k, s)) != hipSuccess) return e;
    if ((e = d2h(pp, op.pp, k, s)) != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

hipError_t post_phimax_cvd(int prec, int L, int M, int N, float dx, float u0, const void* phi, float* px, float* py,
                           float* pz, float* pphi, int cap, PostResult* r, hipStream_t s) {
    return prec == PREC_FP64 ? cvd_t<double>(L, M, N, dx, u0, phi, px, py, pz, pphi, cap, r, s)
                             : cvd_t<float>(L, M, N, dx, u0, phi, px, py, pz, pphi, cap, r, s);
}

hipError_t post_phimax_elastic(int L, int M, int N, float dx, const float* phi, double* x, double* y, double* z,
                               double* ph, int cap, PostResult* r) {
    const long P = (long)L * M * N;
    hipStream_t s = nullptr;
    hipError_t e = post_stream(&s);
    if (e != hipSuccess) return e;
    Scratch sc(s);
    float* dphi = sc.get<float>(P, e);
    if (e != hipSuccess) return e;
    if ((e = h2d(dphi, phi, P, s)) != hipSuccess) return e;
    // elastic.cpp:2809-2816: float running max from 0.
    float pmax = 0.0f;
    for (long i = 0; i < P; i++)
        if (phi[i] > pmax) pmax = phi[i];
    ElasticOp op;
    op.phi = dphi; op.L = L; op.M = M; op.N = N; op.dx = dx;
    op.phic = (pmax - pmax * 0.40f);
    op.x = sc.get<double>(P, e); op.y = sc.get<double>(P, e); op.z = sc.get<double>(P, e); op.ph = sc.get<double>(P, e);
    if (e != hipSuccess) return e;
    int n = 0;
    if ((e = compact(op, P, (int)P, &n, s)) != hipSuccess) return e;
    r->n = n;
    const int k = n < cap ? n : cap;
    if ((e = d2h(x, op.x, k, s)) != hipSuccess) return e;
    if ((e = d2h(y, op.y, k, s)) != hipSuccess) return e;
    if ((e = d2h(z, op.z, k, s)) != hipSuccess) return e;
    if ((e = d2h(ph, op.ph, k, s)) != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

// ---------------------------------------------------------------------------
// Ordered neighbour search (elastic3D, cryst_order, Q6_GLO).  The reference
// walks every j in index order for each atom i, counts every j that passes a
// distance test and works on the first CAP of them, each step using values
// the previous accepted neighbour left (rjx/rjy/rjz).  Here the j range is
// cut into S slices of whole LDS tiles: k_nb_scan runs one thread per
// (i, slice) and records the slice's count and first CAP hits in j order;
// a finishing kernel per atom walks the slices in order, which yields the
// same first CAP neighbours and total count, and applies the reference's
// per-neighbour steps to them in order.  S * (n / 256) workgroups fill the
// chip where one thread per atom gave fewer workgroups than CUs.
// ---------------------------------------------------------------------------
constexpr int SB = 256;

__device__ __forceinline__ float min3f(float a, float b, float c) {
    const float aa = fabsf(a), bb = fabsf(b), cc = fabsf(c);
    return ((aa < bb ? aa : bb) < cc) ? (aa < bb ? aa : bb) : cc;
}

// Elastic pair (double positions): float differences, their +-frame shifts,
// the smallest |.| per axis and the float distance.
__device__ __forceinline__ float el_pair(double xi, double yi, double zi, double xj, double yj, double zj, float fx,
                                         float fy, float fz, float* r3, float& drx, float& dry, float& drz) {
#pragma clang fp contract(off)
    r3[0] = (float)(xi - xj);
    r3[1] = (float)(yi - yj);
    r3[2] = (float)(zi - zj);
    r3[3] = r3[0] + fx; r3[4] = r3[1] + fy; r3[5] = r3[2] + fz;
    r3[6] = r3[0] - fx; r3[7] = r3[1] - fy; r3[8] = r3[2] - fz;
    drx = min3f(fabsf(r3[0]), fabsf(r3[3]), fabsf(r3[6]));
    dry = min3f(fabsf(r3[1]), fabsf(r3[4]), fabsf(r3[7]));
    drz = min3f(fabsf(r3[2]), fabsf(r3[5]), fabsf(r3[8]));
    return sqrtf(drx * drx + dry * dry + drz * drz);
}

// The stored neighbour position of the elastic kernels: per axis, the last
// shift (in 0, +, - order) whose |.| is within 1e-3 of the minimum.
__device__ __forceinline__ void el_rj(double xi, double yi, double zi, const float* r3, float drx, float dry,
                                      float drz, float& rjx, float& rjy, float& rjz) {
#pragma clang fp contract(off)
#pragma unroll
    for (int q = 0; q < 7; q += 3)
        if ((double)fabsf(fabsf(r3[q]) - drx) < 1e-3) rjx = (float)(xi - (double)r3[q]);
#pragma unroll
    for (int q = 1; q < 8; q += 3)
        if ((double)fabsf(fabsf(r3[q]) - dry) < 1e-3) rjy = (float)(yi - (double)r3[q]);
#pragma unroll
    for (int q = 2; q < 9; q += 3)
        if ((double)fabsf(fabsf(r3[q]) - drz) < 1e-3) rjz = (float)(zi - (double)r3[q]);
}

// CVD pair (float fractions): fraction differences times the frame per
// axis, the +-frame shifts, the smallest |.| per axis and the float distance
// plus 1e-6 (cvd.cpp:1297-1309).
__device__ __forceinline__ float cvd_pair(float xi, float yi, float zi, float xj, float yj, float zj, float fx,
                                          float fy, float fz, float* r3, float& drx, float& dry, float& drz) {
#pragma clang fp contract(off)
    r3[0] = (xi - xj) * fx;
    r3[1] = (yi - yj) * fy;
    r3[2] = (zi - zj) * fz;
    r3[3] = r3[0] + fx; r3[4] = r3[1] + fy; r3[5] = r3[2] + fz;
    r3[6] = r3[0] - fx; r3[7] = r3[1] - fy; r3[8] = r3[2] - fz;
    drx = min3f(r3[0], r3[3], r3[6]);
    dry = min3f(r3[1], r3[4], r3[7]);
    drz = min3f(r3[2], r3[5], r3[8]);
    return (float)((double)sqrtf(drx * drx + dry * dry + drz * drz) + 1e-6);
}

// Neighbour tests.  Each gives the position type, the number of neighbours
// the reference stores and the acceptance test of one pair.
struct StrainTest {  // benchel > r > 0.7 (elastic.cpp:3671), 16 stored
    using T = double;
    static constexpr int CAP = 16;
    float fx, fy, fz, benchel;
    __device__ bool operator()(double xi, double yi, double zi, double xj, double yj, double zj) const {
        float r3[9], drx, dry, drz;
        const float rr = el_pair(xi, yi, zi, xj, yj, zj, fx, fy, fz, r3, drx, dry, drz);
        return rr < benchel && (double)rr > 0.7;
    }
};

struct CrystTest {  // r < bench100_h = 1.2 (elastic.cpp:105,3066), 16 stored
    using T = double;
    static constexpr int CAP = 16;
    float fx, fy, fz;
    __device__ bool operator()(double xi, double yi, double zi, double xj, double yj, double zj) const {
        float r3[9], drx, dry, drz;
        return el_pair(xi, yi, zi, xj, yj, zj, fx, fy, fz, r3, drx, dry, drz) < 1.2f;
    }
};

struct Q6Test {  // a*sqrt(3)/2 * [0.9, 1.2] -+ 1e-3 (cvd.cpp:1393), 24 stored
    using T = float;
    static constexpr int CAP = 24;
    float fx, fy, fz;
    double rlo, rhi;
    __device__ bool operator()(float xi, float yi, float zi, float xj, float yj, float zj) const {
        float r3[9], drx, dry, drz;
        const float rr = cvd_pair(xi, yi, zi, xj, yj, zj, fx, fy, fz, r3, drx, dry, drz);
        return (double)rr <= rhi && (double)rr >= rlo;
    }
};

// cnt[i*S + s] = hits of atom i in slice s; lst[(i*S + s)*CAP + k] its k-th
// hit (k < CAP) in j order.  Slice s covers j in [s*slice, (s+1)*slice).
template <class Test>
__global__ void __launch_bounds__(SB) k_nb_scan(Test test, int n, const typename Test::T* __restrict__ X,
                                                const typename Test::T* __restrict__ Y,
                                                const typename Test::T* __restrict__ Z, int slice, int* cnt,
                                                int* lst) {
    using T = typename Test::T;
    __shared__ T sx[SB], sy[SB], sz[SB];
    const int S = gridDim.y, s = blockIdx.y;
    const int i = blockIdx.x * SB + threadIdx.x;
    const bool act = i < n;
    const T xi = act ? X[i] : T(0), yi = act ? Y[i] : T(0), zi = act ? Z[i] : T(0);
    const int jb = s * slice, je = min(n, jb + slice);
    int* out = lst + ((size_t)(act ? i : 0) * S + s) * Test::CAP;
    int c = 0;
    for (int j0 = jb; j0 < je; j0 += SB) {
        __syncthreads();
        const int jl = j0 + threadIdx.x;
        if (jl < je) { sx[threadIdx.x] = X[jl]; sy[threadIdx.x] = Y[jl]; sz[threadIdx.x] = Z[jl]; }
        __syncthreads();
        if (!act) continue;
        const int m = min(SB, je - j0);
        for (int jj = 0; jj < m; jj++) {
            const int j = j0 + jj;
            if (j == i || !test(xi, yi, zi, sx[jj], sy[jj], sz[jj])) continue;
            if (c < Test::CAP) out[c] = j;
            c++;
        }
    }
    if (act) cnt[(size_t)i * S + s] = c;
}

// Calls f(k, j) for the first CAP neighbours of atom i in j order (k = 0,
// 1, ...); returns the total count.
template <int CAP, class F>
__device__ __forceinline__ unsigned nb_walk(int i, int S, const int* __restrict__ cnt, const int* __restrict__ lst,
                                            F&& f) {
    unsigned tot = 0;
    int k = 0;
    for (int s = 0; s < S; s++) {
        const int c = cnt[(size_t)i * S + s];
        tot += (unsigned)c;
        const int* src = lst + ((size_t)i * S + s) * CAP;
        for (int q = 0; q < c && q < CAP && k < CAP; q++) f(k++, src[q]);
    }
    return tot;
}

struct NbSlices {
    int S = 1, slice = SB;
};

// Slices of whole LDS tiles, about 2048 j each, at most 16.
static NbSlices nb_slices(int n) {
    NbSlices r;
    const int tiles = (n + SB - 1) / SB;
    int S = n / 2048;
    S = S < 1 ? 1 : (S > 16 ? 16 : S);
    const int per = (tiles + S - 1) / S;
    r.slice = (per < 1 ? 1 : per) * SB;
    r.S = (n + r.slice - 1) / r.slice;
    if (r.S < 1) r.S = 1;
    return r;
}

// elastic3D (elastic.cpp:3652-3729): Nb counts every neighbour, the first 16
// contribute; terms for 1.3 <= r < benchel (uninitialised reads in the
// reference) are 0.
__global__ void __launch_bounds__(SB) k_strain_fin(int n, const double* __restrict__ X, const double* __restrict__ Y,
                                                   const double* __restrict__ Z, float fx, float fy, float fz, int S,
                                                   const int* __restrict__ cnt, const int* __restrict__ lst,
                                                   double* strain, unsigned* nbo) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * SB + threadIdx.x;
    if (i >= n) return;
    const double xi = X[i], yi = Y[i], zi = Z[i];
    double ax = 0.0, ay = 0.0, az = 0.0;
    float rjx = 0.f, rjy = 0.f, rjz = 0.f;
    const unsigned Nb = nb_walk<StrainTest::CAP>(i, S, cnt, lst, [&](int, int j) {
        float r3[9], drx, dry, drz;
        const double rij = el_pair(xi, yi, zi, X[j], Y[j], Z[j], fx, fy, fz, r3, drx, dry, drz);
        el_rj(xi, yi, zi, r3, drx, dry, drz, rjx, rjy, rjz);
        if (rij < 1.3) {
            ax = ax + 1.0 * (rij - 1) * (((double)rjx - xi) / rij);
            ay = ay + 1.0 * (rij - 1) * (((double)rjy - yi) / rij);
            az = az + 1.0 * (rij - 1) * (((double)rjz - zi) / rij);
        } else {
            ax = ax + 0.0; ay = ay + 0.0; az = az + 0.0;
        }
    });
    strain[i] = sqrt(ax * ax + ay * ay + az * az);
    nbo[i] = Nb;
}

hipError_t post_strain(int L, int M, int N, float dx, float benchel, int n, const double* x, const double* y,
                       const double* z, double* strain, unsigned* nb) {
    if (n <= 0) return hipSuccess;
    hipStream_t s = nullptr;
    hipError_t e = post_stream(&s);
    if (e != hipSuccess) return e;
    const NbSlices ns = nb_slices(n);
    Scratch sc(s);
    double* dx_ = sc.get<double>(n, e);
    double* dy_ = sc.get<double>(n, e);
    double* dz_ = sc.get<double>(n, e);
    double* ds = sc.get<double>(n, e);
    unsigned* dn = sc.get<unsigned>(n, e);
    int* cnt = sc.get<int>((size_t)n * ns.S, e);
    int* lst = sc.get<int>((size_t)n * ns.S * StrainTest::CAP, e);
    if (e != hipSuccess) return e;
    if ((e = h2d(dx_, x, n, s)) != hipSuccess) return e;
    if ((e = h2d(dy_, y, n, s)) != hipSuccess) return e;
    if ((e = h2d(dz_, z, n, s)) != hipSuccess) return e;
    const float fx = L * dx, fy = M * dx, fz = N * dx;
    const StrainTest t{fx, fy, fz, benchel};
    const int gi = (n + SB - 1) / SB;
    hipLaunchKernelGGL(k_nb_scan<StrainTest>, dim3(gi, ns.S), dim3(SB), 0, s, t, n, dx_, dy_, dz_, ns.slice, cnt, lst);
    hipLaunchKernelGGL(k_strain_fin, dim3(gi), dim3(SB), 0, s, n, dx_, dy_, dz_, fx, fy, fz, ns.S, cnt, lst, ds, dn);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = d2h(strain, ds, n, s)) != hipSuccess) return e;
    if ((e = d2h(nb, dn, n, s)) != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

// ---------------------------------------------------------------------------
// cryst_order (elastic.cpp:3013-3389): Lechner-Dellago averaged q_l of every
// atom, with the reference's spherical harmonics (pfc_common.h:274-346, float)
// and its quirks: both loops use l = 6 and the second keeps adding onto the
// first one's per-atom sum, so q6 comes from the sum taken twice; atoms with
// no or more than 16 neighbours get q = 0 but still count as neighbours.
// Passes: neighbour lists in the reference's j order (the ordered scan
// above, then k_cryst_fin), per-atom q_6m, neighbour averages and the sums.
// ---------------------------------------------------------------------------
constexpr int QL = 6, QM = 2 * QL + 1, CNMAX = 16;

__device__ __forceinline__ float dv_lg_poly(int l, float x) {
#pragma clang fp contract(off)
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

__device__ __forceinline__ float dv_plgndr(int l, int m, float x) {
#pragma clang fp contract(off)
    float fact, pll = 0, pmm, pmmp1, somx2;
    if (m < 0) m = -m;
    if (m == 0) return dv_lg_poly(l, x);
    pmm = 1.0f;
    somx2 = (float)sqrt((1.0 - x) * (1.0 + x));
    fact = 1.0f;
    for (int i = 1; i <= m; i++) {
        pmm *= -fact * somx2;
        fact = (float)(fact + 2.0);
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

__device__ __forceinline__ float dv_ylm_norm(int l, int m, double sgn) {
#pragma clang fp contract(off)
    int f1 = 1, f2 = 1;
    for (int i = 1; i <= l - m; i++) f1 = f1 * i;
    for (int i = 1; i <= l + m; i++) f2 = f2 * i;
    return (float)(sgn * sqrt((2 * l + 1) / 4. / 3.1415926) * (double)sqrtf((float)f1 / (float)f2));
}

// Real and imaginary parts as the reference's Ylmtf_r / Ylmtf_c.
__device__ __forceinline__ void dv_ylm(int l, int m, float theta, float fai, float& yr, float& yc) {
#pragma clang fp contract(off)
    const int am = m < 0 ? -m : m;
    const float plm = dv_plgndr(l, am, cosf(theta));
    const float nr = dv_ylm_norm(l, am, (am & 1) ? -1.0 : 1.0);
    yr = nr * plm * cosf(am * fai);
    if (m < 0) yc = dv_ylm_norm(l, am, 1.0) * plm * (-sinf(am * fai));
    else yc = nr * plm * sinf(am * fai);
}

// First 16 neighbours (r < 1.2) in j order: index, theta and fai as the
// reference forms them from the stored neighbour position; Nb counts all.
__global__ void __launch_bounds__(SB) k_cryst_fin(int n, const double* __restrict__ X, const double* __restrict__ Y,
                                                  const double* __restrict__ Z, float fx, float fy, float fz, int S,
                                                  const int* __restrict__ cnt, const int* __restrict__ lst,
                                                  unsigned* nbo, int* nbi, double* theta, double* fai) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * SB + threadIdx.x;
    if (i >= n) return;
    const double xi = X[i], yi = Y[i], zi = Z[i];
    float rjx = 0.f, rjy = 0.f, rjz = 0.f;
    nbo[i] = nb_walk<CrystTest::CAP>(i, S, cnt, lst, [&](int k, int j) {
        float r3[9], drx, dry, drz;
        const double rij = el_pair(xi, yi, zi, X[j], Y[j], Z[j], fx, fy, fz, r3, drx, dry, drz);
        el_rj(xi, yi, zi, r3, drx, dry, drz, rjx, rjy, rjz);
        const size_t s = (size_t)i * CNMAX + k;
        theta[s] = acos(((double)rjz - zi) / (rij + 1.e-4));
        double f = atan(((double)rjy - yi) / ((double)rjx - xi + 1.e-4));
        if ((double)rjx - xi < 0.) f = f + 3.1415926;
        fai[s] = f;
        nbi[s] = j;
    });
}

// qlm[(i*QM + m+QL)*2 + {0,1}] = sum_j Y_6m(j) / Nb (double sums of float terms).
__global__ void __launch_bounds__(SB) k_cryst_qlm(int n, const unsigned* __restrict__ nbo,
                                                  const double* __restrict__ theta, const double* __restrict__ fai,
                                                  double* qlm) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * SB + threadIdx.x;
    if (i >= n) return;
    const unsigned Nb = nbo[i];
    const bool ok = Nb > 0 && Nb <= (unsigned)CNMAX;
    for (int m = -QL; m <= QL; m++) {
        double re = 0.0, im = 0.0;
        if (ok) {
            for (unsigned j = 0; j < Nb; j++) {
                const size_t s = (size_t)i * CNMAX + j;
                float yr, yc;
                dv_ylm(QL, m, (float)theta[s], (float)fai[s], yr, yc);
                re += yr;
                im += yc;
            }
            re = re / (float)Nb;
            im = im / (float)Nb;
        }
        qlm[((size_t)i * QM + (m + QL)) * 2] = re;
        qlm[((size_t)i * QM + (m + QL)) * 2 + 1] = im;
    }
}

__global__ void __launch_bounds__(SB) k_cryst_q(int n, const unsigned* __restrict__ nbo, const int* __restrict__ nbi,
                                                const double* __restrict__ qlm, double* q4, double* q6) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * SB + threadIdx.x;
    if (i >= n) return;
    const unsigned Nb = nbo[i];
    const bool ok = Nb > 0 && Nb <= (unsigned)CNMAX;
    const double xishu = 4 * 3.1415926 / (float)(2 * QL + 1);
    double acc = 0.0;
    for (int pass = 0; pass < 2; pass++) {
        for (int m = 0; m < QM; m++) {
            double br = 0.0, bc = 0.0;
            if (ok) {
                for (unsigned j = 0; j < Nb; j++) {
                    const int jj = nbi[(size_t)i * CNMAX + j];
                    br += qlm[((size_t)jj * QM + m) * 2];
                    bc += qlm[((size_t)jj * QM + m) * 2 + 1];
                }
                br = (qlm[((size_t)i * QM + m) * 2] + br) / (float)(Nb + 1);
                bc = (qlm[((size_t)i * QM + m) * 2 + 1] + bc) / (float)(Nb + 1);
            }
            acc += br * br + bc * bc;
        }
        if (pass == 0) q4[i] = sqrt(xishu * acc);
        else q6[i] = sqrt(xishu * acc);
    }
}

hipError_t post_cryst(int L, int M, int N, float dx, int n, const double* x, const double* y, const double* z,
                      double* q4, double* q6, unsigned* nb) {
    hipStream_t s = nullptr;
    hipError_t e = post_stream(&s);
    if (e != hipSuccess) return e;
    Scratch sc(s);
    double* dx_ = sc.get<double>(n, e);
    double* dy_ = sc.get<double>(n, e);
    double* dz_ = sc.get<double>(n, e);
    unsigned* dn = sc.get<unsigned>(n, e);
    int* di = sc.get<int>((size_t)n * CNMAX, e);
    double* dth = sc.get<double>((size_t)n * CNMAX, e);
    double* dfa = sc.get<double>((size_t)n * CNMAX, e);
    double* dq = sc.get<double>((size_t)n * QM * 2, e);
    double* d4 = sc.get<double>(n, e);
    double* d6 = sc.get<double>(n, e);
    const NbSlices ns = nb_slices(n);
    int* cnt = sc.get<int>((size_t)n * ns.S, e);
    int* lst = sc.get<int>((size_t)n * ns.S * CrystTest::CAP, e);
    if (e != hipSuccess) return e;
    if ((e = h2d(dx_, x, n, s)) != hipSuccess) return e;
    if ((e = h2d(dy_, y, n, s)) != hipSuccess) return e;
    if ((e = h2d(dz_, z, n, s)) != hipSuccess) return e;
    const float fx = L * dx, fy = M * dx, fz = N * dx;
    const dim3 g((n + SB - 1) / SB);
    const CrystTest t{fx, fy, fz};
    hipLaunchKernelGGL(k_nb_scan<CrystTest>, dim3(g.x, ns.S), dim3(SB), 0, s, t, n, dx_, dy_, dz_, ns.slice, cnt, lst);
    hipLaunchKernelGGL(k_cryst_fin, g, dim3(SB), 0, s, n, dx_, dy_, dz_, fx, fy, fz, ns.S, cnt, lst, dn, di, dth, dfa);
    hipLaunchKernelGGL(k_cryst_qlm, g, dim3(SB), 0, s, n, dn, dth, dfa, dq);
    hipLaunchKernelGGL(k_cryst_q, g, dim3(SB), 0, s, n, dn, di, dq, d4, d6);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = d2h(q4, d4, n, s)) != hipSuccess) return e;
    if ((e = d2h(q6, d6, n, s)) != hipSuccess) return e;
    if ((e = d2h(nb, dn, n, s)) != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

// ---------------------------------------------------------------------------
// CVD GR and Q6_GLO (cvd.cpp:1277-1510) on the gathered atom fractions pc_*.
// Pair distances: fractions times the frame per axis, minimum image by |.|
// of the three shifts, 1e-6 added after the float sqrt (a_lattice = 1).
// GR: every pair i < j adds 2 to bin int(rr/delr); blocks cover an i tile
// times a run of j tiles (the triangle's empty blocks return at once) and
// count into an LDS histogram when it fits.  Q6_GLO: neighbour lists in j
// order (window a*sqrt(3)/2*[0.9,1.2], 24 stored), per-entry q_6m and local
// Q6, then the global Q6 from float accumulators taken in entry order, as
// the reference adds them (one lane per m).
// ---------------------------------------------------------------------------
constexpr int GRJT = 8;          // j tiles per GR block
constexpr int GR_LDS_BINS = 12288;  // 48 KB of dynamic LDS
constexpr int Q6CN = 24, Q6NBIN = 30;


template <bool LDSH>
__global__ void __launch_bounds__(SB) k_cvd_gr(int n, const float* __restrict__ X, const float* __restrict__ Y,
                                               const float* __restrict__ Z, float fx, float fy, float fz,
                                               float delr, double rmin, int maxbin, unsigned* hist) {
#pragma clang fp contract(off)
    extern __shared__ unsigned sh[];
    __shared__ float sx[SB], sy[SB], sz[SB];
    const int i0 = blockIdx.x * SB;
    const int jt0 = blockIdx.y * GRJT * SB;  // first j of this block
    if (jt0 + GRJT * SB <= i0 + 1) return;   // every j <= every i: no pair
    if (LDSH)
        for (int b = threadIdx.x; b < maxbin; b += SB) sh[b] = 0;
    const int i = i0 + threadIdx.x;
    const bool act = i < n - 1;
    const float xi = act ? X[i] : 0.f, yi = act ? Y[i] : 0.f, zi = act ? Z[i] : 0.f;
    for (int t = 0; t < GRJT; t++) {
        const int j0 = jt0 + t * SB;
        if (j0 >= n) break;
        __syncthreads();
        const int jl = j0 + threadIdx.x;
        if (jl < n) { sx[threadIdx.x] = X[jl]; sy[threadIdx.x] = Y[jl]; sz[threadIdx.x] = Z[jl]; }
        __syncthreads();
        if (!act) continue;
        const int cnt = n - j0 < SB ? n - j0 : SB;
        const int jj0 = i + 1 - j0 > 0 ? i + 1 - j0 : 0;
        for (int jj = jj0; jj < cnt; jj++) {
            float r3[9], drx, dry, drz;
            const float rr = cvd_pair(xi, yi, zi, sx[jj], sy[jj], sz[jj], fx, fy, fz, r3, drx, dry, drz);
            const unsigned bin = (unsigned)(int)(rr / delr);
            if ((double)rr > rmin && bin < (unsigned)maxbin) atomicAdd(LDSH ? &sh[bin] : &hist[bin], 2u);
        }
    }
    if (LDSH) {
        __syncthreads();
        for (int b = threadIdx.x; b < maxbin; b += SB)
            if (sh[b]) atomicAdd(&hist[b], sh[b]);
    }
}

// First 24 neighbours of entry i in j order with their (theta, fai) as the
// reference forms them from the stored neighbour position; Nb counts all.
__global__ void __launch_bounds__(SB) k_q6_fin(int n, const float* __restrict__ X, const float* __restrict__ Y,
                                               const float* __restrict__ Z, float fx, float fy, float fz, int S,
                                               const int* __restrict__ cnt, const int* __restrict__ lst, unsigned* nbo,
                                               float* theta, float* fai) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * SB + threadIdx.x;
    if (i >= n) return;
    const float xi = X[i], yi = Y[i], zi = Z[i];
    float rjx = 0.f, rjy = 0.f, rjz = 0.f;
    nbo[i] = nb_walk<Q6Test::CAP>(i, S, cnt, lst, [&](int k, int j) {
        float r3[9], drx, dry, drz;
        const float rr = cvd_pair(xi, yi, zi, X[j], Y[j], Z[j], fx, fy, fz, r3, drx, dry, drz);
#pragma unroll
        for (int q = 0; q < 7; q += 3)
            if ((double)(fabsf(r3[q]) - drx) < 1e-5) rjx = xi * fx - r3[q];
#pragma unroll
        for (int q = 1; q < 8; q += 3)
            if ((double)(fabsf(r3[q]) - dry) < 1e-5) rjy = yi * fy - r3[q];
#pragma unroll
        for (int q = 2; q < 9; q += 3)
            if ((double)(fabsf(r3[q]) - drz) < 1e-5) rjz = zi * fz - r3[q];
        const size_t s = (size_t)i * Q6CN + k;
        theta[s] = acosf((rjz - zi * fz) / rr);
        const float ddy = rjy - yi * fy;
        fai[s] = (ddy == 0.) ? 0.f : atanf(ddy / (rjx - xi * fx));
    });
}

// Per entry: qb[m][i] = sum_j Y_6m / Nb (0 without stored neighbours), local
// Q6 = sqrt(xishu * sum_m |qb|^2) (float accumulator over m), its histogram.
__global__ void __launch_bounds__(SB) k_q6_local(int n, const unsigned* __restrict__ nbo,
                                                 const float* __restrict__ theta, const float* __restrict__ fai,
                                                 double2* qb, float* ql, unsigned long long* histq) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * SB + threadIdx.x;
    if (i >= n) return;
    const unsigned Nb = nbo[i];
    const bool ok = Nb > 0 && Nb <= (unsigned)Q6CN;
    float qbm = 0.f;
    for (int m = -QL; m <= QL; m++) {
        double re = 0.0, im = 0.0;
        if (ok) {
            for (unsigned j = 0; j < Nb; j++) {
                const size_t s = (size_t)i * Q6CN + j;
                float yr, yc;
                dv_ylm(QL, m, theta[s], fai[s], yr, yc);
                re += yr;
                im += yc;
            }
            re = re / (float)Nb;
            im = im / (float)Nb;
        }
        qb[(size_t)(m + QL) * n + i] = make_double2(re, im);
        qbm = (float)((double)qbm + (re * re + im * im));
    }
    const float xishu = (float)(4 * 3.1415926 / (float)(2 * QL + 1));
    const float q = sqrtf(xishu * qbm);
    ql[i] = q;
    const unsigned bin = (unsigned)(int)(q / 0.05f);
    if (bin < (unsigned)Q6NBIN) atomicAdd(&histq[bin], 1ull);
}

// Global Q6: lane m+6 runs the reference's float accumulators over the
// entries in order; lane 0 adds the 13 |Q_6m|^2 in m order.
__global__ void __launch_bounds__(64) k_q6_global(int n, const unsigned* __restrict__ nbo,
                                                  const double2* __restrict__ qb, float* qglobal) {
#pragma clang fp contract(off)
    __shared__ float part[QM];
    const int t = threadIdx.x;
    if (t < QM) {
        float qr = 0.f, qc = 0.f, sum_nb = 0.f;
        const double2* q = qb + (size_t)t * n;
        // Loads of a run of entries are issued before the in-order adds.
        constexpr int U = 16;
        int i = 0;
        for (; i + U <= n; i += U) {
            unsigned nb[U];
            double2 v[U];
#pragma unroll
            for (int k = 0; k < U; k++) {
                nb[k] = nbo[i + k];
                v[k] = q[i + k];
            }
#pragma unroll
            for (int k = 0; k < U; k++) {
                qr = (float)((double)qr + (double)nb[k] * v[k].x);
                qc = (float)((double)qc + (double)nb[k] * v[k].y);
                sum_nb = sum_nb + (float)nb[k];
            }
        }
        for (; i < n; i++) {
            const unsigned nb = nbo[i];
            const double2 v = q[i];
            qr = (float)((double)qr + (double)nb * v.x);
            qc = (float)((double)qc + (double)nb * v.y);
            sum_nb = sum_nb + (float)nb;
        }
        qr = qr / sum_nb;
        qc = qc / sum_nb;
        part[t] = qr * qr + qc * qc;
    }
    __syncthreads();
    if (t == 0) {
        float acc = 0.f;
        for (int m = 0; m < QM; m++) acc = acc + part[m];
        const float xishu = (float)(4 * 3.1415926 / (float)(2 * QL + 1));
        qglobal[0] = sqrtf(xishu * acc);
    }
}

int cvd_gr_bins(int L, float dx) {
    const float delr = (float)(1.0f / 25.);
    return (int)(L * dx / delr);
}

hipError_t post_cvd_gr(int L, int M, int N, float dx, int n, const float* x, const float* y, const float* z,
                       unsigned* hist) {
    const int maxbin = cvd_gr_bins(L, dx);
    if (maxbin <= 0) return hipSuccess;
    hipStream_t s = nullptr;
    hipError_t e = post_stream(&s);
    if (e != hipSuccess) return e;
    Scratch sc(s);
    float* dx_ = sc.get<float>(n, e);
    float* dy_ = sc.get<float>(n, e);
    float* dz_ = sc.get<float>(n, e);
    unsigned* dh = sc.get<unsigned>(maxbin, e);
    if (e != hipSuccess) return e;
    if ((e = h2d(dx_, x, n, s)) != hipSuccess) return e;
    if ((e = h2d(dy_, y, n, s)) != hipSuccess) return e;
    if ((e = h2d(dz_, z, n, s)) != hipSuccess) return e;
    if ((e = hipMemsetAsync(dh, 0, sizeof(unsigned) * maxbin, s)) != hipSuccess) return e;
    const float fx = L * dx, fy = M * dx, fz = N * dx;
    const float delr = (float)(1.0f / 25.);
    const double rmin = 1.0f / sqrt(2.) * 0.5;
    if (n > 1) {
        const dim3 g((n + SB - 1) / SB, (n + GRJT * SB - 1) / (GRJT * SB));
        if (maxbin <= GR_LDS_BINS)
            hipLaunchKernelGGL(k_cvd_gr<true>, g, dim3(SB), sizeof(unsigned) * maxbin, s, n, dx_, dy_, dz_, fx, fy,
                               fz, delr, rmin, maxbin, dh);
        else
            hipLaunchKernelGGL(k_cvd_gr<false>, g, dim3(SB), 0, s, n, dx_, dy_, dz_, fx, fy, fz, delr, rmin, maxbin,
                               dh);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if ((e = d2h(hist, dh, maxbin, s)) != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

hipError_t post_cvd_q6(int L, int M, int N, float dx, int n, const float* x, const float* y, const float* z,
                       unsigned* nb, float* ql, float* qglobal, unsigned long long* histq) {
    hipStream_t s = nullptr;
    hipError_t e = post_stream(&s);
    if (e != hipSuccess) return e;
    Scratch sc(s);
    float* dx_ = sc.get<float>(n, e);
    float* dy_ = sc.get<float>(n, e);
    float* dz_ = sc.get<float>(n, e);
    unsigned* dn = sc.get<unsigned>(n, e);
    float* dth = sc.get<float>((size_t)n * Q6CN, e);
    float* dfa = sc.get<float>((size_t)n * Q6CN, e);
    double2* dq = sc.get<double2>((size_t)n * QM, e);
    float* dql = sc.get<float>(n, e);
    float* dg = sc.get<float>(1, e);
    unsigned long long* dhq = sc.get<unsigned long long>(Q6NBIN, e);
    const NbSlices ns = nb_slices(n);
    int* cnt = sc.get<int>((size_t)n * ns.S, e);
    int* lst = sc.get<int>((size_t)n * ns.S * Q6Test::CAP, e);
    if (e != hipSuccess) return e;
    if ((e = h2d(dx_, x, n, s)) != hipSuccess) return e;
    if ((e = h2d(dy_, y, n, s)) != hipSuccess) return e;
    if ((e = h2d(dz_, z, n, s)) != hipSuccess) return e;
    if ((e = hipMemsetAsync(dhq, 0, sizeof(unsigned long long) * Q6NBIN, s)) != hipSuccess) return e;
    const float fx = L * dx, fy = M * dx, fz = N * dx;
    const double rhi = 1.0f * sqrt(3.) / 2. * 1.2 + 1e-3, rlo = 1.0f * sqrt(3.) / 2. * 0.9 - 1e-3;
    const Q6Test t{fx, fy, fz, rlo, rhi};
    const int gi = (n + SB - 1) / SB;
    hipLaunchKernelGGL(k_nb_scan<Q6Test>, dim3(gi, ns.S), dim3(SB), 0, s, t, n, dx_, dy_, dz_, ns.slice, cnt, lst);
    hipLaunchKernelGGL(k_q6_fin, dim3(gi), dim3(SB), 0, s, n, dx_, dy_, dz_, fx, fy, fz, ns.S, cnt, lst, dn, dth, dfa);
    hipLaunchKernelGGL(k_q6_local, dim3((n + SB - 1) / SB), dim3(SB), 0, s, n, dn, dth, dfa, dq, dql, dhq);
    hipLaunchKernelGGL(k_q6_global, dim3(1), dim3(64), 0, s, n, dn, dq, dg);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = d2h(nb, dn, n, s)) != hipSuccess) return e;
    if ((e = d2h(ql, dql, n, s)) != hipSuccess) return e;
    if ((e = d2h(qglobal, dg, 1, s)) != hipSuccess) return e;
    if ((e = d2h(histq, dhq, Q6NBIN, s)) != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

// ---------------------------------------------------------------------------
// Misfit noise field (Gauss, misfit.cpp:951-981).  The reference draws, point
// by point in memory order, Marsaglia polar pairs (v1, v2) = two successive
// u11 draws of one splitmix64 stream until 0 < v1^2 + v2^2 < 1.  splitmix64's
// k-th output is mix(s0 + k*gamma), so pair j is decided independently; point
// p takes the p-th accepted pair.  Count accepted pairs per block, scan, emit.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t sm64_at(uint64_t s0, uint64_t k) {
    uint64_t z = s0 + k * 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

__device__ __forceinline__ float u11f(uint64_t x) {
#pragma clang fp contract(off)
    const double u = (double)(x >> 11) * (1.0 / 9007199254740992.0);
    return (float)(2.0 * u - 1.0);
}

__device__ __forceinline__ uint64_t noise_state(unsigned seed, long long t) {
    uint64_t st = (uint64_t)seed;
    st ^= (uint64_t)(unsigned)(0 + 1) * 2654435761ULL;
    st ^= (uint64_t)(unsigned)t * 374761393ULL;
    return st;
}

struct NoisePair {
    float v1, r;
    bool ok;
};

__device__ __forceinline__ NoisePair noise_pair(uint64_t s0, long j) {
#pragma clang fp contract(off)
    NoisePair q;
    q.v1 = u11f(sm64_at(s0, (uint64_t)(2 * j + 1)));
    const float v2 = u11f(sm64_at(s0, (uint64_t)(2 * j + 2)));
    q.r = q.v1 * q.v1 + v2 * v2;
    q.ok = !(q.r >= 1.0f || q.r <= 0.0f);
    return q;
}

__global__ void __launch_bounds__(PB) k_noise_count(const unsigned* seeds, const long long* tcount, long NP,
                                                    int nb, int* counts) {
    __shared__ int wsum[PB / 64];
    const int inst = blockIdx.y;
    const long j = (long)blockIdx.x * PB + threadIdx.x;
    const uint64_t s0 = noise_state(seeds[inst], tcount[inst]);
    const bool f = j < NP && noise_pair(s0, j).ok;
    int total;
    (void)block_rank(f, wsum, total);
    if (threadIdx.x == 0) counts[(size_t)inst * nb + blockIdx.x] = total;
}

template <typename R>
__global__ void __launch_bounds__(PB) k_noise_emit(const unsigned* seeds, const long long* tcount, long NP, long P,
                                                   int nb, const int* offsets, R* planes, int* err) {
    __shared__ int wsum[PB / 64];
    const int inst = blockIdx.y;
    const long j = (long)blockIdx.x * PB + threadIdx.x;
    const uint64_t s0 = noise_state(seeds[inst], tcount[inst]);
    NoisePair q{0.f, 0.f, false};
    if (j < NP) q = noise_pair(s0, j);
    int total;
    const int rk = block_rank(q.ok, wsum, total);
    const int* off = offsets + (size_t)inst * (nb + 1);
    const long o = (long)off[blockIdx.x] + rk;
    if (q.ok && o < P) {
        const float f = (float)sqrt(-2. * (double)logf(q.r) / (double)q.r);
        planes[(size_t)inst * P + o] = (R)(f * q.v1);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && off[nb] < P) atomicOr(err, 1);
}

static long noise_pairs(long P) {
    // Acceptance probability pi/4; the margin is > 30 standard deviations.
    return (long)(P / 0.785398163397448 + 40.0 * sqrt((double)P) + 4096.0);
}

size_t noise_scratch_bytes(long P, int B) {
    const long NP = noise_pairs(P);
    const long nb = (NP + PB - 1) / PB;
    return sizeof(int) * (size_t)B * (2 * nb + 1) + 64;
}

hipError_t launch_noise(int prec, const DevBufs& b, const unsigned* seeds, void* planes, void* nz, void* scratch,
                        int* err_flag, hipStream_t s) {
    const long P = (long)b.L * b.M * b.N;
    const long NP = noise_pairs(P);
    const int nb = (int)((NP + PB - 1) / PB);
    int* counts = (int*)scratch;
    int* offs = counts + (size_t)b.B * nb;
    hipLaunchKernelGGL(k_noise_count, dim3(nb, b.B), dim3(PB), 0, s, seeds, (const long long*)b.tcount, NP, nb, counts);
    hipLaunchKernelGGL(k_scan, dim3(1, b.B), dim3(1024), 0, s, counts, nb, offs);
    if (prec == PREC_FP64)
        hipLaunchKernelGGL(k_noise_emit<double>, dim3(nb, b.B), dim3(PB), 0, s, seeds, (const long long*)b.tcount, NP,
                           P, nb, offs, (double*)planes, err_flag);
    else
        hipLaunchKernelGGL(k_noise_emit<float>, dim3(nb, b.B), dim3(PB), 0, s, seeds, (const long long*)b.tcount, NP,
                           P, nb, offs, (float*)planes, err_flag);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_row_r2c1(prec, b, planes, nz, s);
}

// Plain copy of n 16-byte words (reads n*16 bytes, writes n*16 bytes).
__global__ void __launch_bounds__(256) k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}

// The same words viewed as [n/256][256]: each workgroup copies 4 adjacent
// columns down all rows, so every row contributes one 64-byte piece per
// workgroup -- the access granularity of the column pass.  XCD: column tiles
// mapped as the column pass maps them (the workgroups of one XCD take
// adjacent tiles).
template <bool XCD>
__global__ void __launch_bounds__(256) k_copy16_cols(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                     size_t rows) {
    const unsigned bx = XCD ? (blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8 : blockIdx.x;
    const size_t col = (size_t)bx * 4 + threadIdx.x % 4;
    for (size_t r = (size_t)blockIdx.y * 64 + threadIdx.x / 4; r < rows; r += (size_t)gridDim.y * 64)
        dst[r * 256 + col] = src[r * 256 + col];
}

// ---------------------------------------------------------------------------
// Allocation-free elastic Phimax (pfc_phimax_elastic_noalloc) on a field the
// caller holds in device memory, with device outputs, on the caller's stream.
// Its scratch (the compaction's block counts and offsets, the threshold's
// partial maxima) is the file-scope array below; the library neither
// allocates nor frees device memory for it, so callers that cannot (graph
// capture, serving loops) can use it.  Every GPU holds its own copy for the
// life of the process and every call on that GPU uses it: calls that overlap
// on one GPU race.  Each kernel writes every scratch word it later reads, so
// nothing needs clearing between calls.
// ---------------------------------------------------------------------------
constexpr int NOALLOC_BLOCKS = (int)(NOALLOC_MAX_POINTS / PB);  // compaction blocks of the largest field
constexpr int NOALLOC_PARTS = 256;                                // partial maxima
constexpr size_t NOALLOC_BYTES = ((size_t)(2 * NOALLOC_BLOCKS + 1) * sizeof(int) + (NOALLOC_PARTS + 1) * sizeof(float) +
                                  255) / 256 * 256;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

// Device address of this GPU's `scratch`, looked up once per device with
// that device current (the array's name is a host-side stand-in).
static std::mutex g_scratch_mu;
static void* g_scratch_dev[64];
static hipError_t scratch_address(void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        if ((e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) != hipSuccess) return e;
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return hipSuccess;
}

// elastic.cpp:2809-2816: float running max from 0.  NaN never compares
// greater and max is exact, so partial maxima in any order give the same bits.
__global__ void __launch_bounds__(256) k_fmax_part(const float* __restrict__ phi, long P, float* part) {
    __shared__ float red[256];
    float m = 0.0f;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < P; i += (long)gridDim.x * 256) {
        const float v = phi[i];
        if (v > m) m = v;
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s && red[threadIdx.x + s] > red[threadIdx.x]) red[threadIdx.x] = red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// The threshold phic = pmax - pmax * 0.40 (elastic.cpp:2817) from the partials.
__global__ void __launch_bounds__(256) k_elastic_phic(const float* __restrict__ part, int n, float* phic) {
#pragma clang fp contract(off)
    __shared__ float red[256];
    float m = 0.0f;
    for (int i = threadIdx.x; i < n; i += 256)
        if (part[i] > m) m = part[i];
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s && red[threadIdx.x + s] > red[threadIdx.x]) red[threadIdx.x] = red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const float pmax = red[0];
        *phic = pmax - pmax * 0.40f;
    }
}

// ElasticOp with its threshold read from device memory.
struct ElasticDevOp {
    ElasticOp op;
    const float* phic;
    __device__ bool test(long p) const {
        ElasticOp o = op;
        o.phic = *phic;
        return o.test(p);
    }
    __device__ void emit(long p, size_t o) const { op.emit(p, o); }
};

__global__ void k_store_total(const int* __restrict__ offs, int nb, int* n_out) {
    if (threadIdx.x == 0) *n_out = offs[nb];
}

hipError_t post_phimax_elastic_noalloc(int L, int M, int N, float dx, const float* phi, double* x, double* y,
                                       double* z, double* ph, int cap, int* n_out, hipStream_t s) {
    const long P = (long)L * M * N;
    if (P <= 0 || P > NOALLOC_MAX_POINTS) return hipErrorInvalidValue;
    void* base = nullptr;
    hipError_t e = scratch_address(&base);
    if (e != hipSuccess) return e;
    int* counts = (int*)base;
    int* offs = counts + NOALLOC_BLOCKS;
    float* part = (float*)(offs + NOALLOC_BLOCKS + 1);
    float* phic = part + NOALLOC_PARTS;
    const int nb = (int)((P + PB - 1) / PB);
    const int np = nb < NOALLOC_PARTS ? nb : NOALLOC_PARTS;
    hipLaunchKernelGGL(k_fmax_part, dim3(np), dim3(256), 0, s, phi, P, part);
    hipLaunchKernelGGL(k_elastic_phic, dim3(1), dim3(256), 0, s, (const float*)part, np, phic);
    ElasticDevOp op;
    op.op.phi = phi; op.op.L = L; op.op.M = M; op.op.N = N; op.op.dx = dx; op.op.phic = 0.0f;
    op.op.x = x; op.op.y = y; op.op.z = z; op.op.ph = ph;
    op.phic = phic;
    hipLaunchKernelGGL(k_count<ElasticDevOp>, dim3(nb), dim3(PB), 0, s, op, P, counts);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, (const int*)counts, nb, offs);
    hipLaunchKernelGGL(k_emit<ElasticDevOp>, dim3(nb), dim3(PB), 0, s, op, P, (const int*)offs, cap);
    hipLaunchKernelGGL(k_store_total, dim3(1), dim3(64), 0, s, (const int*)offs, nb, n_out);
    return hipGetLastError();
}

hipError_t launch_copy16(const void* src, void* dst, size_t n, int pattern, hipStream_t s) {
    if (pattern == 0) {
        hipLaunchKernelGGL(k_copy16, dim3(4096), dim3(256), 0, s, (const uint4*)src, (uint4*)dst, n);
    } else {
        if (n % 256) return hipErrorInvalidValue;
        if (pattern == 1)
            hipLaunchKernelGGL(k_copy16_cols<false>, dim3(64, 256), dim3(256), 0, s, (const uint4*)src, (uint4*)dst,
                               n / 256);
        else
            hipLaunchKernelGGL(k_copy16_cols<true>, dim3(64, 256), dim3(256), 0, s, (const uint4*)src, (uint4*)dst,
                               n / 256);
    }
    return hipGetLastError();
}

}  // namespace pfc
