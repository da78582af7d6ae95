// pfc_io.cpp - file formats of pfc-exp-cli.  The byte layouts are the
// reference's, so the Qt GUI's Visualizer and the elastic reader accept them.
#include "pfc_io.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <exception>
#include <filesystem>
#include <set>
#include <thread>

namespace pfcio {

// printf-style append; the same bytes as fprintf with the same format.
static void appf(std::string& s, const char* fmt, ...) {
    char b[256];
    va_list a;
    va_start(a, fmt);
    const int k = std::vsnprintf(b, sizeof b, fmt, a);
    va_end(a);
    if (k < 0) return;
    if (k < static_cast<int>(sizeof b)) {
        s.append(b, static_cast<size_t>(k));
        return;
    }
    std::string big(static_cast<size_t>(k) + 1, '\0');
    va_start(a, fmt);
    std::vsnprintf(&big[0], big.size(), fmt, a);
    va_end(a);
    s.append(big.data(), static_cast<size_t>(k));
}

// Rows [0, n) of a table, formatted by row(i, out) (appending nothing skips
// the row) in contiguous chunks on up to 8 threads, then written in order:
// the file is byte-for-byte what a sequential loop of the same formats writes.
template <class Row>
static void put_rows(FILE* f, int n, const Row& row) {
    constexpr int kMinRows = 4096;
    const int hw = static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
    const int T = std::max(1, std::min({8, hw, n / kMinRows}));
    std::vector<std::string> out(static_cast<size_t>(T));
    auto chunk = [&](int t) {
        const int a = static_cast<int>(static_cast<long long>(n) * t / T);
        const int b = static_cast<int>(static_cast<long long>(n) * (t + 1) / T);
        out[t].reserve(static_cast<size_t>(b - a) * 48);
        for (int i = a; i < b; i++) row(i, out[t]);
    };
    std::vector<std::thread> th;
    for (int t = 1; t < T; t++) th.emplace_back(chunk, t);
    chunk(0);
    for (auto& x : th) x.join();
    for (const auto& o : out) std::fwrite(o.data(), 1, o.size(), f);
}

// A VTI payload (n, m, l with l fastest) from / to the n + N*(m + M*l)
// layout, in 32 x 32 (m, l) blocks so both sides stay in cache lines.
static void to_payload(int L, int M, int N, const float* field, float* pay) {
    constexpr int B = 32;
    for (int n = 0; n < N; n++)
        for (int m0 = 0; m0 < M; m0 += B)
            for (int l0 = 0; l0 < L; l0 += B) {
                const int m1 = std::min(M, m0 + B), l1 = std::min(L, l0 + B);
                for (int m = m0; m < m1; m++) {
                    float* dst = pay + (static_cast<size_t>(n) * M + m) * L;
                    for (int l = l0; l < l1; l++) dst[l] = field[n + static_cast<size_t>(N) * (m + static_cast<size_t>(M) * l)];
                }
            }
}

static void from_payload(int L, int M, int N, const float* pay, size_t count, float* field) {
    constexpr int B = 32;
    const size_t total = static_cast<size_t>(L) * M * N;
    for (int n = 0; n < N; n++)
        for (int l0 = 0; l0 < L; l0 += B)
            for (int m0 = 0; m0 < M; m0 += B) {
                const int m1 = std::min(M, m0 + B), l1 = std::min(L, l0 + B);
                for (int l = l0; l < l1; l++) {
                    float* dst = field + n + static_cast<size_t>(N) * static_cast<size_t>(M) * l;
                    const float* src = pay + static_cast<size_t>(n) * M * L + l;
                    if (count == total) {
                        for (int m = m0; m < m1; m++) dst[static_cast<size_t>(N) * m] = src[static_cast<size_t>(m) * L];
                    } else {
                        for (int m = m0; m < m1; m++) {
                            const size_t at = (static_cast<size_t>(n) * M + m) * L + l;
                            if (at < count) dst[static_cast<size_t>(N) * m] = src[static_cast<size_t>(m) * L];
                        }
                    }
                }
            }
}

std::string g6(double v) {
    char b[64];
    std::snprintf(b, sizeof b, "%g", v);
    return b;
}

bool write_vti(const std::string& path, const char* name, int L, int M, int N, const float* field) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    std::fprintf(f, "<?xml version=\"1.0\"?>\n");
    std::fprintf(f, "<VTKFile type=\"ImageData\" version=\"0.1\"  byte_order=\"LittleEndian\">\n");
    std::fprintf(f, "<ImageData WholeExtent=\"%d %d %d %d %d %d\"   Origin=\"%d %d %d\" Spacing=\"%d %d %d\">\n", 0,
                 L - 1, 0, M - 1, 0, N - 1, 0, 0, 0, 1, 1, 1);
    std::fprintf(f, "<Piece Extent=\"%d %d %d %d %d %d\">\n", 0, L - 1, 0, M - 1, 0, N - 1);
    std::fprintf(f, "<PointData Scalars=\"%s\">\n", name);
    std::fprintf(f, "<DataArray type=\"Float32\" Name=\"%s\"   format=\"appended\" offset=\"%d\" />\n", name, 0);
    std::fprintf(f, "</PointData>\n<CellData>\n</CellData>\n</Piece>\n</ImageData>\n");
    std::fprintf(f, "<AppendedData encoding=\"raw\">\n_");
    const int bytes = static_cast<int>(static_cast<size_t>(N) * M * L * sizeof(float)) + static_cast<int>(sizeof(int));
    std::fwrite(&bytes, sizeof(int), 1, f);
    // Payload order n, m, l (l fastest): a transpose of the n+N*(m+M*l) layout.
    std::vector<float> pay(static_cast<size_t>(N) * M * L);
    to_payload(L, M, N, field, pay.data());
    std::fwrite(pay.data(), sizeof(float), pay.size(), f);
    std::fprintf(f, "\n</AppendedData>\n</VTKFile>\n");
    return std::fclose(f) == 0;
}

bool write_pvti(const std::string& path, const char* name, const char* prefix, int kd, int L, int M, int N) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    std::fprintf(f, "<?xml version=\"1.0\"?>\n");
    std::fprintf(f, "<VTKFile type=\"PImageData\" version=\"0.1\" byte_order=\"LittleEndian\">\n");
    std::fprintf(f,
                 "  <PImageData WholeExtent=\"0 %d 0 %d %d %d\" GhostLevel=\"#\" Origin=\"0  0  0\"  Spacing=\"1  1  1\">\n",
                 L - 1, M - 1, 0, N - 1);
    std::fprintf(f, "    <PPointData>\n      <PDataArray type=\"Float32\" Name=\"%s\"/>\n    </PPointData>\n", name);
    std::fprintf(f, "    <Piece Extent=\"0 %d 0 %d %d %d\" Source=\"%s%d_0.vti\"/>\n", L - 1, M - 1, 0, N - 1, prefix, kd);
    std::fprintf(f, "  </PImageData>\n</VTKFile>\n");
    return std::fclose(f) == 0;
}

long read_vti(const std::string& path, int L, int M, int N, float* field) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return -1;
    char line[100];
    for (int i = 1; i < 13; i++)
        if (!std::fgets(line, 100, f)) break;
    (void)std::fgetc(f);
    int bytes = 0;
    if (std::fread(&bytes, sizeof(int), 1, f) != 1) bytes = 0;
    // Values in payload order; a short file fills the first `got` of them.
    std::vector<float> pay(static_cast<size_t>(N) * M * L);
    const size_t got = std::fread(pay.data(), sizeof(float), pay.size(), f);
    std::fclose(f);
    from_payload(L, M, N, pay.data(), got, field);
    return static_cast<long>(got);
}

static bool finite_atom(double x, double y, double z, double v, double c) {
    return std::isfinite(x) && std::isfinite(y) && std::isfinite(z) && std::isfinite(v) && std::isfinite(c);
}

bool write_phimax_misfit(const std::string& path, int n, const double* x, const double* y, const double* z,
                         const double* phi, const float* con) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    int cnt = 0;
    for (int i = 0; i < n; i++)
        if (finite_atom(x[i], y[i], z[i], phi[i], con[i]) && !(x[i] < 0)) cnt++;
    std::fprintf(f, "%d\nKommentar\n", 2 * cnt);
    put_rows(f, n, [&](int i, std::string& o) {
        const double c = con[i];
        if (!finite_atom(x[i], y[i], z[i], phi[i], c) || x[i] < 0) return;
        appf(o, "%.4f %.4f %.4f %.4f %.4f\n", x[i], y[i], z[i], phi[i], c);
        appf(o, "%.4f %.4f %.4f %.4f %.4f\n", x[i], y[i], z[i] + 20, phi[i], c);
    });
    return std::fclose(f) == 0;
}

bool write_phimax_elastic(const std::string& path, int n, const double* x, const double* y, const double* z,
                          const double* phi) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    std::fprintf(f, "%d\nKommentar\n", n);
    put_rows(f, n, [&](int i, std::string& o) {
        if (x[i] >= 0) appf(o, "%.4f %.4f %.4f %.4f\n", x[i], y[i], z[i], phi[i]);
    });
    return std::fclose(f) == 0;
}

bool write_strain(const std::string& path, int n, const double* x, const double* y, const double* z,
                  const double* strain, const unsigned* nb) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    std::fprintf(f, "%d\nKommentar\n", n);
    put_rows(f, n, [&](int i, std::string& o) {
        if (x[i] >= 0) appf(o, "st %.4f %.4f %.4f %.4f %u\n", x[i], y[i], z[i], strain[i], nb[i]);
    });
    return std::fclose(f) == 0;
}

bool write_q4q6(const std::string& path, int kd, int n, const double* x, const double* y, const double* z,
                const double* q4, const double* q6) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    int cnt = 0;
    for (int i = 0; i < n; i++) cnt += x[i] >= 0;
    std::fprintf(f, "%d\nKommentar\n", cnt);
    double x0 = 1e30, x1 = -1e30, y0 = 1e30, y1 = -1e30;
    for (int i = 0; i < n; i++) {
        if (!(x[i] >= 0)) continue;
        if (x[i] < x0) x0 = x[i];
        if (x[i] > x1) x1 = x[i];
        if (y[i] < y0) y0 = y[i];
        if (y[i] > y1) y1 = y[i];
    }
    put_rows(f, n, [&](int i, std::string& o) {
        if (x[i] >= 0) appf(o, "qc %.4f %.4f %.4f %.4f %.4f\n", x[i], y[i], z[i], q4[i], q6[i]);
    });
    std::fprintf(stderr, "q4q6_%d.txt: atoms=%d x=[%s,%s] y=[%s,%s]\n", kd, cnt, g6(x0).c_str(), g6(x1).c_str(),
                 g6(y0).c_str(), g6(y1).c_str());
    return std::fclose(f) == 0;
}

bool write_cvd_atoms(int kd, int L, int M, int N, float dx, float phic, float phimax, int n, const float* px,
                     const float* py, const float* pz, const float* pphi, const std::string& dir) {
    const std::string pre = dir.empty() ? std::string() : dir + "/";
    FILE* f = std::fopen((pre + "phimax.txt").c_str(), "a");
    if (!f) return false;
    std::fprintf(f, "%d , %s , %s , %d\n", kd, g6(phic).c_str(), g6(phimax).c_str(), 0);
    std::fprintf(f, "phimax is done myid=0\n buf is done \n pc_xyz is initializing \n pc_xyz is done \n");
    std::fprintf(f, " pc_xyz has been gathered \n");
    std::fclose(f);
    const float frame_x = L * dx, frame_y = M * dx, frame_z = N * dx;
    char name[64];
    std::snprintf(name, sizeof name, "phiball_%d.cfg", kd);
    f = std::fopen((pre + name).c_str(), "w");
    if (!f) return false;
    std::fprintf(f, "Number of particles = %d\n", n);
    std::fprintf(f, "A = %.4f Angstrom (basic length-scale)\n", 10.0);
    std::fprintf(f, "H0(1,1) = %.4f A \n", frame_y);
    std::fprintf(f, "H0(1,2) = 0 A \nH0(1,3) = 0 A \nH0(2,1) = 0 A \n");
    std::fprintf(f, "H0(2,2) = %.4f A \n", frame_z);
    std::fprintf(f, "H0(2,3) = 0 A \nH0(3,1) = 0 A \nH0(3,2) = 0 A \n");
    std::fprintf(f, "H0(3,3) = %.4f A \n", frame_x);
    std::fprintf(f, ".NO_VELOCITY.\nentry_count = 3\n63.546\nCu\n");
    put_rows(f, n, [&](int i, std::string& o) { appf(o, "%.4f %.4f %.4f\n", px[i], py[i], pz[i]); });
    std::fclose(f);
    std::snprintf(name, sizeof name, "Phimax_%d.txt", kd);
    f = std::fopen((pre + name).c_str(), "w");
    if (!f) return false;
    std::fprintf(f, "%d\nKommentar\n", n);
    put_rows(f, n, [&](int i, std::string& o) {
        appf(o, "%.4f %.4f %.4f %.4f\n", pz[i] * frame_x, px[i] * frame_y, py[i] * frame_z, pphi[i]);
    });
    return std::fclose(f) == 0;
}

// The reference runs on x86, where 0/0 is the negative default NaN and
// iostream prints it as "-nan"; NaNs are printed that way whatever sign the
// device produced.
static void put_fixed(FILE* f, double v, int prec) {
    if (std::isnan(v)) std::fputs("-nan", f);
    else std::fprintf(f, "%.*f", prec, v);
}

static void app_fixed(std::string& o, double v, int prec) {
    if (std::isnan(v)) o += "-nan";
    else appf(o, "%.*f", prec, v);
}

bool write_cvd_gr(int kd, int L, int M, int N, float dx, int n, const float* px, const unsigned* hist, int maxbin,
                  const std::string& dir) {
    const std::string pre = dir.empty() ? std::string() : dir + "/";
    const float delr = (float)(1.0f / 25.);
    int num_t = 0;
    for (int i = 0; i < n; i++) num_t += px[i] >= 0;
    char name[64];
    std::snprintf(name, sizeof name, "hist_%d.txt", kd);
    FILE* f = std::fopen((pre + name).c_str(), "w");
    if (!f) return false;
    for (int b = 0; b < maxbin; b++) std::fprintf(f, "%d %u %d\n", b, hist[b], maxbin);
    std::fclose(f);
    const float rho = (float)num_t / (float)(L * M * N * dx * dx * dx);
    const float preco = (float)(4. * 3.1415926 * rho / 3.);
    std::snprintf(name, sizeof name, "Grq6_%d.txt", kd);
    f = std::fopen((pre + name).c_str(), "w");
    if (!f) return false;
    for (int b = 0; b < maxbin; b++) {
        const float rlower = (float)b * delr;
        const float rupper = rlower + delr;
        const float nideal = preco * (rupper * rupper * rupper - rlower * rlower * rlower);
        const float gr = (float)hist[b] / (float)1 / (float)num_t / nideal;
        std::fprintf(f, "%.4f ", b * delr);
        put_fixed(f, gr, 4);
        std::fprintf(f, " %u ", hist[b]);
        put_fixed(f, nideal, 4);
        std::fputc('\n', f);
    }
    return std::fclose(f) == 0;
}

bool write_cvd_q6(int kd, int n, const unsigned* nb, const float* ql, float qglobal, const unsigned long long* histq,
                  const std::string& dir) {
    const std::string pre = dir.empty() ? std::string() : dir + "/";
    char name[64];
    std::snprintf(name, sizeof name, "Ql_%d.txt", kd);
    FILE* f = std::fopen((pre + name).c_str(), "w");
    if (!f) return false;
    put_rows(f, n, [&](int i, std::string& o) {
        appf(o, "%d %u ", i, nb[i]);
        app_fixed(o, ql[i], 6);
        o += ' ';
        app_fixed(o, qglobal, 6);
        o += '\n';
    });
    std::fclose(f);
    std::snprintf(name, sizeof name, "HistQ_%d.txt", kd);
    f = std::fopen((pre + name).c_str(), "w");
    if (!f) return false;
    for (int b = 0; b < 30; b++) std::fprintf(f, "%.4f %llu\n", (b + 1) * 0.05f, histq[b]);
    return std::fclose(f) == 0;
}

std::vector<int> checkpoint_steps(const std::string& dir) {
    std::vector<int> out;
    namespace fs = std::filesystem;
    const fs::path d = dir.empty() ? fs::current_path() : fs::path(dir);
    std::error_code ec;
    if (!fs::exists(d, ec) || ec) return out;
    std::set<int> ps, cs;
    for (const auto& e : fs::directory_iterator(d, ec)) {
        if (ec) break;
        if (!e.is_regular_file()) continue;
        const std::string nm = e.path().filename().string();
        int st = 0, rk = 0;
        if (std::sscanf(nm.c_str(), "phi_%d_%d.vti", &st, &rk) == 2) { ps.insert(st); continue; }
        if (std::sscanf(nm.c_str(), "con_%d_%d.vti", &st, &rk) == 2) cs.insert(st);
    }
    bool pos = false;
    for (int s : ps) {
        if (!cs.count(s)) continue;
        if (s > 0) pos = true;
        out.push_back(s);
    }
    std::sort(out.begin(), out.end());
    if (pos) out.erase(std::remove(out.begin(), out.end(), 0), out.end());
    return out;
}

bool make_dirs(const std::string& path) {
    std::error_code ec;
    std::filesystem::create_directories(path, ec);
    return !ec;
}

Writer::Writer(int depth) : depth_(depth < 1 ? 1 : depth), th_([this] { loop(); }) {}

Writer::~Writer() {
    {
        std::unique_lock<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    th_.join();
}

void Writer::post(std::function<void()> task) {
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [this] { return static_cast<int>(q_.size()) < depth_; });
    q_.push_back(std::move(task));
    cv_.notify_all();
}

void Writer::drain() {
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [this] { return q_.empty() && !busy_; });
}

void Writer::loop() {
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
        cv_.wait(lk, [this] { return stop_ || !q_.empty(); });
        if (q_.empty()) return;  // stop_ and nothing left
        std::function<void()> task = std::move(q_.front());
        q_.pop_front();
        busy_ = true;
        cv_.notify_all();
        lk.unlock();
        try {
            task();
        } catch (const std::exception& e) {  // a lost file, not a lost process
            std::fprintf(stderr, "checkpoint writer: %s\n", e.what());
        }
        lk.lock();
        busy_ = false;
        cv_.notify_all();
    }
}

bool write_run_config(const std::string& path, const RunConfig& c) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    const bool misfit = std::string(c.model) == "misfit", elastic = std::string(c.model) == "elastic";
    std::fprintf(f, "model %s\n", c.model);
    std::fprintf(f, "u0 %s\ncon0 %s\nsig %s\ndt %s\ndx %s\n", g6(c.u0).c_str(), g6(c.con0).c_str(), g6(c.sig).c_str(),
                 g6(c.dt).c_str(), g6(c.dx).c_str());
    std::fprintf(f, "steps %d\nmod %d\nseed %u\nL %d\nM %d\nN %d\n", c.steps, c.mod, c.seed, c.L, c.M, c.N);
    if (misfit)
        std::fprintf(f, "grainx %d\ngrainy %d\ngrainz %d\naxx %s\nb0 %s\n", c.grainx, c.grainy, c.grainz,
                     g6(c.axx).c_str(), g6(c.b0).c_str());
    if (elastic) {
        std::fprintf(f, "grainx %d\ngrainy %d\ngrainz %d\nbenchel %s\n", c.grainx, c.grainy, c.grainz,
                     g6(c.benchel).c_str());
        if (!c.indir.empty()) std::fprintf(f, "input_dir %s\n", c.indir.c_str());
    }
    return std::fclose(f) == 0;
}

float noise_b0(float axx, float dx, float dt) {
    if (dx > 0.0f && dt > 0.0f && axx >= 0.0f)
        return static_cast<float>(std::sqrt(static_cast<double>(axx / dx / dx / dx / dt)));
    return 0.0f;
}

bool start_timestamps(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    std::fprintf(f, "# Checkpoint timestamps - Step number and milliseconds since epoch\n");
    return std::fclose(f) == 0;
}

bool append_timestamp(const std::string& path, int step) {
    FILE* f = std::fopen(path.c_str(), "a");
    if (!f) return false;
    std::fprintf(f, "step %d timestamp_ms %lld\n", step, now_ms());
    return std::fclose(f) == 0;
}

bool append_energy(const std::string& path, int kd0, int n, const double* e) {
    FILE* f = std::fopen(path.c_str(), "a");
    if (!f) return false;
    for (int s = 0; s < n; s++) std::fprintf(f, "%d %s %s\n", kd0 + s, g6(e[2 * s]).c_str(), g6(e[2 * s + 1]).c_str());
    return std::fclose(f) == 0;
}

long long now_ms() {
    return std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::system_clock::now().time_since_epoch())
        .count();
}

}  // namespace pfcio
