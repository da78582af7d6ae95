// pfc_cli.cpp - pfc-exp-cli: the process boundary the Qt GUI drives.
//
// Same argv, working-directory, progress-line, file and exit-code contract
// as the reference CLI (main_cli.cpp:24-88; misfit.cpp:181-531;
// cvd.cpp:127-454; elastic.cpp:188-521).  The numerics run on the GPU through
// libpfc_hip.so.  Additions (all optional): --L/--M/--N (runtime grid, the
// reference's compile-time PFC_GRID_*), --precision fp64|fp32, --narrow 0|1.
#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cmath>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <string>
#include <string_view>
#include <vector>

#include "../../include/pfc_hip.h"
#include "pfc_io.hpp"

namespace {

enum Kind { MISFIT = 0, CVD = 1, ELASTIC = 2 };

struct Opts {
    pfc_params p;
    int steps = 200, mod = 25;
    float benchel = 1.7f;
    std::string outdir, indir;
};

bool parse_int(const char* v, int& out) {
    if (!v) return false;
    char* end = nullptr;
    errno = 0;
    long x = std::strtol(v, &end, 10);
    if (errno || end == v || *end || x < INT_MIN || x > INT_MAX) return false;
    out = static_cast<int>(x);
    return true;
}
bool parse_uint(const char* v, unsigned& out) {
    if (!v) return false;
    char* end = nullptr;
    errno = 0;
    unsigned long long x = std::strtoull(v, &end, 10);
    if (errno || end == v || *end || x > 0xffffffffULL) return false;
    out = static_cast<unsigned>(x);
    return true;
}
bool parse_float(const char* v, float& out) {
    if (!v) return false;
    char* end = nullptr;
    errno = 0;
    float x = std::strtof(v, &end);
    if (errno || end == v || *end) return false;
    out = x;
    return true;
}

void usage(Kind k, const char* argv0) {
    const char* a0 = argv0 ? argv0 : "pfc-exp-cli";
    if (k == MISFIT) {
        std::fprintf(stderr,
                     "Usage: %s --model misfit [options]\nOptions:\n"
                     "  --u0 <float>        Density/phi mean (default: 0.05)\n"
                     "  --con0 <float>      Concentration mean (default: 0.2)\n"
                     "  --sig <float>       Kernel decay sigma (default: 0.05)\n"
                     "  --dt <float>        Time step (default: 0.05)\n"
                     "  --dx <float>        Space step (default: 0.125)\n"
                     "  --steps <int>       Total iterations (default: 200)\n"
                     "  --mod <int>         Output/analysis interval (default: 25)\n"
                     "  --seed <uint>       Random seed base (default: 20200604)\n"
                     "  --grainx <int>      Initial grains in x (default: 16)\n"
                     "  --grainy <int>      Initial grains in y (default: 16)\n"
                     "  --grainz <int>      Initial grains in z (default: 1)\n"
                     "  --axx <float>       Noise base intensity (default: 0)\n"
                     "  --outdir <path>     Output directory (default: current dir)\n"
                     "  --L/--M/--N <int>   Grid (default: 256 256 1)\n"
                     "  --precision <p>     fp64 (default) or fp32\n"
                     "  --narrow <0|1>      Reference float narrowing (default: 1)\n"
                     "  -h, --help          Show this help\n",
                     a0);
    } else if (k == CVD) {
        std::fprintf(stderr,
                     "Usage: %s --model cvd [options]\nOptions:\n"
                     "  --u0 <float>        Density/phi mean (default: 0.01)\n"
                     "  --con0 <float>      Concentration mean (default: 0.2)\n"
                     "  --sig <float>       Kernel decay sigma (default: 0.03)\n"
                     "  --dt <float>        Time step (default: 0.05)\n"
                     "  --dx <float>        Space step (default: 0.125)\n"
                     "  --steps <int>       Total iterations (default: 2000)\n"
                     "  --mod <int>         Output/analysis interval (default: 200)\n"
                     "  --seed <uint>       Random seed base (default: 20200604)\n"
                     "  --outdir <path>     Output directory (default: current dir)\n"
                     "  --L/--M/--N <int>   Grid (default: 256 256 1)\n"
                     "  --precision <p>     fp64 (default) or fp32\n"
                     "  -h, --help          Show this help\n",
                     a0);
    } else {
        std::fprintf(stderr,
                     "Usage: %s --model elastic [options]\nOptions:\n"
                     "  --dx <float>        Space step (default: 0.125)\n"
                     "  --steps <int>       Total iterations (default: 2002)\n"
                     "  --mod <int>         Output/analysis interval (default: 200)\n"
                     "  --benchel <float>   Neighbor cutoff for elastic analysis (default: 1.7)\n"
                     "  --input-dir <path>  Directory with pre-computed phi/con data\n"
                     "  --outdir <path>     Output directory (default: current dir)\n"
                     "  --L/--M/--N <int>   Grid (default: 256 256 1)\n"
                     "  -h, --help          Show this help\n",
                     a0);
    }
}

// Returns 0 ok, 1 help, 2 bad arguments.
int parse(Kind k, int argc, char** argv, Opts& o) {
    static const char* names[] = {"--u0", "--con0", "--sig", "--dt", "--dx", "--steps", "--mod", "--seed",
                                  "--grainx", "--grainy", "--grainz", "--axx", "--outdir", "--L", "--M", "--N",
                                  "--precision", "--narrow", "--benchel", "--input-dir", nullptr};
    for (int i = 1; i < argc; i++) {
        const char* raw = argv[i];
        if (!raw) continue;
        std::string_view a(raw);
        if (a == "-h" || a == "--help") return 1;
        int key = -1;
        const char* v = nullptr;
        bool missing = false;
        for (int q = 0; names[q]; q++) {
            const std::string_view nm(names[q]);
            if (a == nm) {
                key = q;
                if (i + 1 < argc) v = argv[++i];
                else missing = true;
                break;
            }
            if (a.size() > nm.size() && a.substr(0, nm.size()) == nm && a[nm.size()] == '=') {
                key = q;
                v = raw + nm.size() + 1;
                break;
            }
        }
        // Flags a model does not know behave as unknown flags: rejected by
        // misfit (misfit.cpp:269-272), ignored by cvd/elastic (cvd.cpp:192-193).
        const bool known = key >= 0 && !((k != ELASTIC && key >= 18) || (k == CVD && key >= 8 && key <= 11) ||
                                          (k == ELASTIC && (key == 11 || key == 17)));
        if (!known) {
            if (k == MISFIT) {
                std::fprintf(stderr, "Unknown option: %s\n", raw);
                return 2;
            }
            continue;
        }
        if (missing) {
            std::fprintf(stderr, "Missing value after %s\n", names[key]);
            return 2;
        }
        bool ok = true;
        switch (key) {
            case 0: ok = parse_float(v, o.p.u0); break;
            case 1: ok = parse_float(v, o.p.con0); break;
            case 2: ok = parse_float(v, o.p.sig); break;
            case 3: ok = parse_float(v, o.p.dt); break;
            case 4: ok = parse_float(v, o.p.dx); break;
            case 5: ok = parse_int(v, o.steps); break;
            case 6: ok = parse_int(v, o.mod); break;
            case 7: ok = parse_uint(v, o.p.seed); break;
            case 8: ok = parse_int(v, o.p.grainx); break;
            case 9: ok = parse_int(v, o.p.grainy); break;
            case 10: ok = parse_int(v, o.p.grainz); break;
            case 11: ok = parse_float(v, o.p.axx); break;
            case 12: ok = v && *v; if (ok) o.outdir = v; break;
            case 13: ok = parse_int(v, o.p.L); break;
            case 14: ok = parse_int(v, o.p.M); break;
            case 15: ok = parse_int(v, o.p.N); break;
            case 16:
                ok = v && (!std::strcmp(v, "fp64") || !std::strcmp(v, "fp32"));
                if (ok) o.p.precision = !std::strcmp(v, "fp64") ? PFC_PREC_FP64 : PFC_PREC_FP32;
                break;
            case 17: ok = parse_int(v, o.p.narrow); break;
            case 18: ok = parse_float(v, o.benchel); break;
            case 19: ok = v && *v; if (ok) o.indir = v; break;
        }
        if (!ok) return 2;
    }
    if (o.steps <= 0 || o.p.dt <= 0.0f || o.p.dx <= 0.0f || o.mod <= 0) {
        std::fprintf(stderr, "Invalid parameter values\n");
        return 2;
    }
    if (o.p.L <= 0 || o.p.M <= 0 || o.p.N <= 0) return 2;
    if (k == MISFIT) {
        if (o.p.grainx <= 0 || o.p.grainy <= 0 || o.p.grainz <= 0) {
            std::fprintf(stderr, "--grainx/--grainy/--grainz must be > 0\n");
            return 2;
        }
        if (o.p.grainx > o.p.L || o.p.grainy > o.p.M || o.p.grainz > o.p.N) {
            std::fprintf(stderr, "--grainx/--grainy/--grainz must not exceed grid dimensions (L=%d M=%d N=%d)\n",
                         o.p.L, o.p.M, o.p.N);
            return 2;
        }
    }
    return 0;
}

// PFC_SetupOutputDir (pfc_common.h:249-269): create, then chdir.
bool setup_outdir(const std::string& d) {
    if (d.empty()) return true;
    if (!pfcio::make_dirs(d)) {
        std::fprintf(stderr, "Failed to create --outdir '%s'\n", d.c_str());
        return false;
    }
    std::error_code ec;
    std::filesystem::current_path(d, ec);
    if (ec) {
        std::fprintf(stderr, "Failed to chdir to --outdir '%s': %s\n", d.c_str(), ec.message().c_str());
        return false;
    }
    return true;
}

int gpu_fail(const char* what) {
    std::fprintf(stderr, "%s failed: %s\n", what, pfc_last_error());
    return 1;
}

void write_run_config(Kind k, const Opts& o, float b0) {
    pfcio::RunConfig c;
    c.model = k == MISFIT ? "misfit" : (k == CVD ? "cvd" : "elastic");
    c.u0 = o.p.u0; c.con0 = o.p.con0; c.sig = o.p.sig; c.dt = o.p.dt; c.dx = o.p.dx;
    c.steps = o.steps; c.mod = o.mod; c.seed = o.p.seed;
    c.L = o.p.L; c.M = o.p.M; c.N = o.p.N;
    c.grainx = o.p.grainx; c.grainy = o.p.grainy; c.grainz = o.p.grainz;
    c.axx = o.p.axx; c.b0 = b0; c.benchel = o.benchel; c.indir = o.indir;
    pfcio::write_run_config("run_config.txt", c);
}

// VTI + PVTI pair of both fields at checkpoint kd (misfit.cpp:1588-1698).
void write_field_files(int kd, int L, int M, int N, const float* a, const float* b) {
    pfcio::write_vti("phi_" + std::to_string(kd) + "_0.vti", "phi", L, M, N, a);
    pfcio::write_pvti("phi_" + std::to_string(kd) + ".pvti", "phi", "phi_", kd, L, M, N);
    pfcio::write_vti("con_" + std::to_string(kd) + "_0.vti", "con", L, M, N, b);
    pfcio::write_pvti("con_" + std::to_string(kd) + ".pvti", "con", "con_", kd, L, M, N);
}

// Both fields of the context's instance 0 as float (what the VTI holds).
int fetch_fields(pfc_ctx* ctx, size_t P, std::vector<float>& a, std::vector<float>& b) {
    a.resize(P);
    b.resize(P);
    return pfc_get_fields_f32(ctx, 0, a.data(), b.data());
}

// CVD order results of a checkpoint: GR every time, Q6_GLO when kd % 20 == 0
// (cvd.cpp:1253-1254), over the atom fractions of pfc_phimax_cvd.  Computed
// on the GPU here; write() puts them into files.
struct CvdOrder {
    int kd = 0;
    std::vector<unsigned> hist;
    bool has_q6 = false;
    std::vector<unsigned> nb;
    std::vector<float> ql;
    float qg = 0.f;
    unsigned long long hq[30] = {};

    int compute(int k, const pfc_params& p, int na, const float* px, const float* py, const float* pz) {
        kd = k;
        int nbin = 0;
        if (pfc_cvd_gr(p.L, p.M, p.N, p.dx, na, px, py, pz, nullptr, &nbin)) return 1;
        hist.assign(static_cast<size_t>(nbin), 0u);
        if (pfc_cvd_gr(p.L, p.M, p.N, p.dx, na, px, py, pz, hist.data(), &nbin)) return 1;
        has_q6 = kd >= 0 && kd % 20 == 0 && na > 0;
        if (has_q6) {
            nb.resize(static_cast<size_t>(na));
            ql.resize(static_cast<size_t>(na));
            if (pfc_cvd_q6(p.L, p.M, p.N, p.dx, na, px, py, pz, nb.data(), ql.data(), &qg, hq)) return 1;
        }
        return 0;
    }
    void write(const pfc_params& p, int na, const float* px, const std::string& dir) const {
        pfcio::write_cvd_gr(kd, p.L, p.M, p.N, p.dx, na, px, hist.data(), static_cast<int>(hist.size()), dir);
        if (has_q6) pfcio::write_cvd_q6(kd, na, nb.data(), ql.data(), qg, hq, dir);
    }
};

float b0_of(const pfc_params& p) { return pfcio::noise_b0(p.axx, p.dx, p.dt); }

// run_misfit (misfit.cpp:311-531).
int run_sim(Kind k, int argc, char** argv) {
    Opts o;
    pfc_params_default(&o.p, k == MISFIT ? PFC_MODEL_MISFIT : PFC_MODEL_CVD);
    if (k == CVD) { o.steps = 2000; o.mod = 200; }
    const int rc = parse(k, argc, argv, o);
    if (rc == 1) { usage(k, argv[0]); return 0; }
    if (rc != 0) { usage(k, argv[0]); return 2; }
    if (!setup_outdir(o.outdir)) return 2;
    const float b0 = b0_of(o.p);
    std::fprintf(stderr, "%sRun config: u0=%s con0=%s sig=%s dt=%s dx=%s steps=%d mod=%d seed=%u grid=(%d,%d,%d) outdir=%s\n",
                 k == CVD ? "CVD " : "", pfcio::g6(o.p.u0).c_str(), pfcio::g6(o.p.con0).c_str(),
                 pfcio::g6(o.p.sig).c_str(), pfcio::g6(o.p.dt).c_str(), pfcio::g6(o.p.dx).c_str(), o.steps, o.mod,
                 o.p.seed, o.p.L, o.p.M, o.p.N, o.outdir.empty() ? "." : o.outdir.c_str());
    write_run_config(k, o, b0);
    FILE* ck = std::fopen("checkpoint_timestamps.txt", "w");
    if (ck) { std::fprintf(ck, "# Checkpoint timestamps - Step number and milliseconds since epoch\n"); std::fflush(ck); }

    const size_t P = static_cast<size_t>(o.p.L) * o.p.M * o.p.N;
    std::vector<double> phi(P), con(P);
    if (pfc_init_fields(&o.p, phi.data(), con.data())) return gpu_fail("init");
    pfc_ctx* ctx = nullptr;
    if (pfc_create(&o.p, 1, &ctx)) return gpu_fail("pfc_create");
    if (pfc_set_fields(ctx, 0, phi.data(), con.data())) return gpu_fail("pfc_set_fields");
    const int L = o.p.L, M = o.p.M, N = o.p.N;
    if (k == MISFIT) {
        std::vector<float> fa(P), fb(P);
        for (size_t i = 0; i < P; i++) { fa[i] = static_cast<float>(phi[i]); fb[i] = static_cast<float>(con[i]); }
        write_field_files(0, L, M, N, fa.data(), fb.data());
    }
    // Checkpoint files are formatted and written by `out` while the GPU
    // steps on; each task owns copies of what it writes (gr_fail, declared
    // first, outlives the writer's tasks).
    std::atomic<bool> gr_fail{false};
    pfcio::Writer out;
    std::vector<double> ax(P), ay(P), az(P), aphi(P);
    std::vector<float> acon(P), px(P), py(P), pz(P), pp(P);
    if (k == MISFIT) {
        const int every = std::max(1, o.steps / 200);
        // Steps between progress lines run as one captured graph (replays are
        // bit-identical to eager launches); the noise term keeps eager steps.
        if (every >= 8 && !(o.p.axx > 0.0f) && pfc_set_graph(ctx, every)) return gpu_fail("pfc_set_graph");
        int t = 0;
        while (t < o.steps) {
            // Advance to the next step that prints progress or checkpoints.
            int next = t + 1;
            while (!(next == 1 || next % every == 0 || next == o.steps || next % o.mod == 0)) next++;
            if (pfc_step(ctx, next - t)) return gpu_fail("pfc_step");
            t = next;
            const int os = t;
            if (os == 1 || os % every == 0 || os == o.steps) {
                if (pfc_sync(ctx)) return gpu_fail("pfc_sync");
                std::printf("PFC_PROGRESS step=%d total=%d\n", os, o.steps);
                std::fflush(stdout);
            }
            if (os % o.mod == 0 || os == o.steps) {
                const int kd = os;
                int n = 0;
                if (pfc_phimax_misfit(ctx, 0, ax.data(), ay.data(), az.data(), aphi.data(), acon.data(),
                                      static_cast<int>(P), &n))
                    return gpu_fail("pfc_phimax_misfit");
                // The 2-D extraction is repeated for each n (misfit.cpp:3345).
                std::vector<double> X, Y, Z, V;
                std::vector<float> C;
                for (int nn = 0; nn < N; nn++)
                    for (int i = 0; i < n; i++) {
                        X.push_back(ax[i]); Y.push_back(ay[i]);
                        Z.push_back(nn == 0 ? az[i] : static_cast<double>(static_cast<float>(nn + 1) * o.p.dx));
                        V.push_back(aphi[i]); C.push_back(acon[i]);
                    }
                std::vector<float> fa, fb;
                if (fetch_fields(ctx, P, fa, fb)) return gpu_fail("pfc_get_fields_f32");
                double e[2];
                if (pfc_energy(ctx, e)) return gpu_fail("pfc_energy");
                out.post([=, X = std::move(X), Y = std::move(Y), Z = std::move(Z), V = std::move(V),
                          C = std::move(C), fa = std::move(fa), fb = std::move(fb)] {
                    pfcio::write_phimax_misfit("Phimax_" + std::to_string(kd) + ".txt", static_cast<int>(X.size()),
                                               X.data(), Y.data(), Z.data(), V.data(), C.data());
                    write_field_files(kd, L, M, N, fa.data(), fb.data());
                    FILE* ef = std::fopen("energy.txt", "a");
                    if (ef) { std::fprintf(ef, "%d %s %s\n", kd, pfcio::g6(e[0]).c_str(), pfcio::g6(e[1]).c_str()); std::fclose(ef); }
                    if (ck) { std::fprintf(ck, "step %d timestamp_ms %lld\n", kd, pfcio::now_ms()); std::fflush(ck); }
                });
            }
        }
    } else {
        // run_cvd (cvd.cpp:389-454): ENERGYC2 after every step, checkpoint
        // outputs after step t when t % mod == 0 (kd = t).
        int t = 0;
        while (t < o.steps) {
            int last = t;  // chunk [t, last] ends at a checkpoint step or the end
            while (last < o.steps - 1 && last % o.mod != 0) last++;
            const int n = last - t + 1;
            std::vector<double> tr(2 * static_cast<size_t>(n));
            if (pfc_step_energy(ctx, n, tr.data())) return gpu_fail("pfc_step_energy");
            for (int s = t; s <= last; s++) std::fprintf(stderr, "PFC_PROGRESS step=%d total=%d\n", s + 1, o.steps);
            std::fflush(stderr);
            if (last % o.mod == 0) {
                int na = 0;
                float phic = 0.f, pmax = 0.f;
                if (pfc_phimax_cvd(ctx, 0, px.data(), py.data(), pz.data(), pp.data(), static_cast<int>(P), &na, &phic,
                                   &pmax))
                    return gpu_fail("pfc_phimax_cvd");
                std::vector<float> fa, fb;
                if (fetch_fields(ctx, P, fa, fb)) return gpu_fail("pfc_get_fields_f32");
                std::vector<float> qx(px.begin(), px.begin() + na), qy(py.begin(), py.begin() + na),
                    qz(pz.begin(), pz.begin() + na), qp(pp.begin(), pp.begin() + na);
                const pfc_params prm = o.p;
                // GR and Q6 work on the extracted atoms only: the writer
                // thread runs them (on its own stream) beside the next steps.
                out.post([=, &gr_fail, qx = std::move(qx), qy = std::move(qy), qz = std::move(qz), qp = std::move(qp),
                          fa = std::move(fa), fb = std::move(fb)] {
                    pfcio::write_cvd_atoms(last, L, M, N, prm.dx, phic, pmax, na, qx.data(), qy.data(), qz.data(),
                                           qp.data());
                    CvdOrder ord;
                    if (ord.compute(last, prm, na, qx.data(), qy.data(), qz.data())) {
                        gpu_fail("pfc_cvd_gr/q6");  // the message is this thread's
                        gr_fail = true;
                        return;
                    }
                    ord.write(prm, na, qx.data(), "");
                    write_field_files(last, L, M, N, fa.data(), fb.data());
                    if (ck) { std::fprintf(ck, "step %d timestamp_ms %lld\n", last, pfcio::now_ms()); std::fflush(ck); }
                });
                if (gr_fail) break;
            }
            out.post([t, n, tr = std::move(tr)] {
                FILE* ef = std::fopen("energy.txt", "a");
                if (!ef) return;
                for (int s = 0; s < n; s++)
                    std::fprintf(ef, "%d %s %s\n", t + s, pfcio::g6(tr[2 * s]).c_str(), pfcio::g6(tr[2 * s + 1]).c_str());
                std::fclose(ef);
            });
            t = last + 1;
        }
        out.drain();
        if (gr_fail) return 1;
    }
    out.drain();
    if (ck) std::fclose(ck);
    pfc_destroy(ctx);
    return 0;
}

// run_elastic (elastic.cpp:304-521).
int run_elastic(int argc, char** argv) {
    Opts o;
    pfc_params_default(&o.p, PFC_MODEL_MISFIT);
    o.p.dt = 0.1f;
    o.p.grainx = 2; o.p.grainy = 2; o.p.grainz = 1;
    o.steps = 2002; o.mod = 200;
    const int rc = parse(ELASTIC, argc, argv, o);
    if (rc == 1) { usage(ELASTIC, argv[0]); return 0; }
    if (rc != 0) { usage(ELASTIC, argv[0]); return 2; }
    if (!o.indir.empty()) {
        std::error_code ec;
        const auto a = std::filesystem::absolute(o.indir, ec);
        if (!ec) o.indir = a.string();
    }
    if (!setup_outdir(o.outdir)) return 2;
    write_run_config(ELASTIC, o, 0.0f);
    FILE* ck = std::fopen("checkpoint_timestamps.txt", "w");
    if (ck) { std::fprintf(ck, "# Checkpoint timestamps - Step number and milliseconds since epoch\n"); std::fclose(ck); }
    std::vector<int> steps = pfcio::checkpoint_steps(o.indir);
    if (!steps.empty()) {
        std::fprintf(stderr, "Elastic detected %zu checkpoint step(s)\n", steps.size());
    } else {
        for (int t = 0; t < o.steps; t++)
            if ((t + 1) % o.mod == 0 || t + 1 == o.steps) steps.push_back(t + 1);
    }
    const int L = o.p.L, M = o.p.M, N = o.p.N;
    const size_t P = static_cast<size_t>(L) * M * N;
    std::vector<float> phi(P);
    std::vector<double> x(P), y(P), z(P), pm(P), st(P), q4(P), q6(P);
    std::vector<unsigned> nb(P);
    pfcio::Writer out;
    for (size_t idx = 0; idx < steps.size(); idx++) {
        const int kd = steps[idx];
        std::fprintf(stderr, "PFC_PROGRESS step=%zu total=%zu\n", idx + 1, steps.size());
        std::fflush(stderr);
        const std::string base = o.indir.empty() ? std::string() : o.indir + "/";
        const std::string fp = base + "phi_" + std::to_string(kd) + "_0.vti";
        const std::string fc = base + "con_" + std::to_string(kd) + "_0.vti";
        if (pfcio::read_vti(fp, L, M, N, phi.data()) < 0) {
            std::fprintf(stderr, "can't open input file: %s\n", fp.c_str());
            return 2;
        }
        if (FILE* cf = std::fopen(fc.c_str(), "rb")) std::fclose(cf);
        else { std::fprintf(stderr, "can't open input file: %s\n", fc.c_str()); return 2; }
        int n = 0;
        if (pfc_phimax_elastic(L, M, N, o.p.dx, phi.data(), x.data(), y.data(), z.data(), pm.data(),
                               static_cast<int>(P), &n))
            return gpu_fail("pfc_phimax_elastic");
        if (pfc_strain(L, M, N, o.p.dx, o.benchel, n, x.data(), y.data(), z.data(), st.data(), nb.data()))
            return gpu_fail("pfc_strain");
        if (pfc_cryst_order(L, M, N, o.p.dx, n, x.data(), y.data(), z.data(), q4.data(), q6.data(), nullptr))
            return gpu_fail("pfc_cryst_order");
        // The three files of this checkpoint are written while the next one
        // is read and processed.
        auto head = [n](const std::vector<double>& v) { return std::vector<double>(v.begin(), v.begin() + n); };
        out.post([kd, n, X = head(x), Y = head(y), Z = head(z), PM = head(pm), ST = head(st), Q4 = head(q4),
                  Q6 = head(q6), NB = std::vector<unsigned>(nb.begin(), nb.begin() + n)] {
            pfcio::write_phimax_elastic("Phimax_" + std::to_string(kd) + ".txt", n, X.data(), Y.data(), Z.data(),
                                        PM.data());
            pfcio::write_strain("strain_" + std::to_string(kd) + ".txt", n, X.data(), Y.data(), Z.data(), ST.data(),
                                NB.data());
            pfcio::write_q4q6("q4q6_" + std::to_string(kd) + ".txt", kd, n, X.data(), Y.data(), Z.data(), Q4.data(),
                              Q6.data());
        });
    }
    out.drain();
    return 0;
}

void global_usage(const char* argv0) {
    std::fprintf(stderr,
                 "Usage: %s [--model <name>] [model-options...]\n\nModels:\n"
                 "  misfit    Phase-field crystal with misfit strain (default)\n"
                 "  cvd       Controlled vapor deposition / two-phase growth\n"
                 "  elastic   Post-processing: elastic strain & crystallographic order\n\n"
                 "Use --model <name> --help for model-specific options.\n",
                 argv0 ? argv0 : "pfc-exp-cli");
}

}  // namespace

// main_cli.cpp:24-88: strip --model/-m/--model=, dispatch.
int main(int argc, char** argv) {
    std::string model = "misfit";
    std::vector<char*> av{argv[0]};
    for (int i = 1; i < argc; i++) {
        if (!argv[i]) continue;
        std::string_view a(argv[i]);
        if (a == "--model" || a == "-m") {
            if (i + 1 < argc) { model = argv[++i]; continue; }
            std::fprintf(stderr, "Missing value after --model\n");
            global_usage(argv[0]);
            return 2;
        }
        if (a.substr(0, 8) == "--model=") { model = std::string(a.substr(8)); continue; }
        if ((a == "-h" || a == "--help") && model == "misfit" && i == 1) {
            bool has_model = false;
            for (int j = i + 1; j < argc; j++) {
                std::string_view b(argv[j] ? argv[j] : "");
                if (b == "--model" || b.substr(0, 8) == "--model=") { has_model = true; break; }
            }
            if (!has_model) { global_usage(argv[0]); return 0; }
        }
        av.push_back(argv[i]);
    }
    const int ac = static_cast<int>(av.size());
    if (model == "misfit") return run_sim(MISFIT, ac, av.data());
    if (model == "cvd") return run_sim(CVD, ac, av.data());
    if (model == "elastic") return run_elastic(ac, av.data());
    std::fprintf(stderr, "Unknown model: %s\n", model.c_str());
    global_usage(argv[0]);
    return 2;
}
