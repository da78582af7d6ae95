// pfc_io.hpp - file formats of pfc-exp-cli (what the Qt GUI reads).
#pragma once
#include <condition_variable>
#include <cstdio>
#include <deque>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace pfcio {

// phi_<kd>_0.vti: 12 XML lines, '_', int32 (payload bytes + 4), float32
// payload with l fastest, then the closing tags (misfit.cpp:1616-1698).
bool write_vti(const std::string& path, const char* name, int L, int M, int N, const float* field);
// phi_<kd>.pvti (misfit.cpp:1588-1614).
bool write_pvti(const std::string& path, const char* name, const char* prefix, int kd, int L, int M, int N);
// READPHIVTIN (elastic.cpp:2562-2608): returns values read, -1 if unopenable.
long read_vti(const std::string& path, int L, int M, int N, float* field);

// Phimax_<kd>.txt of the Misfit model: header 2*count, each atom twice
// (z and z+20), columns x y z phi con (misfit.cpp:3495-3536).
bool write_phimax_misfit(const std::string& path, int n, const double* x, const double* y, const double* z,
                         const double* phi, const float* con);
// Elastic Phimax_<kd>.txt (elastic.cpp:2953-2968).
bool write_phimax_elastic(const std::string& path, int n, const double* x, const double* y, const double* z,
                          const double* phi);
// strain_<kd>.txt (elastic.cpp:3738-3752).
bool write_strain(const std::string& path, int n, const double* x, const double* y, const double* z,
                  const double* strain, const unsigned* nb);
// q4q6_<kd>.txt of cryst_order (elastic.cpp:3328-3350): atoms with x >= 0,
// "qc x y z q4 q6" with 4 decimals; plus its summary line on stderr.
bool write_q4q6(const std::string& path, int kd, int n, const double* x, const double* y, const double* z,
                const double* q4, const double* q6);
// CVD checkpoint outputs of atompositionCFGC2 (cvd.cpp:1086-1249):
// phimax.txt (append), phiball_<kd>.cfg and Phimax_<kd>.txt.
// Files go to `dir` (empty: the working directory).
bool write_cvd_atoms(int kd, int L, int M, int N, float dx, float phic, float phimax, int n, const float* px,
                     const float* py, const float* pz, const float* pphi, const std::string& dir = "");
// hist_<kd>.txt and Grq6_<kd>.txt of GR (cvd.cpp:1314-1339) from the pair
// histogram; num_t counts the px >= 0 entries, nstep = 1.
bool write_cvd_gr(int kd, int L, int M, int N, float dx, int n, const float* px, const unsigned* hist, int maxbin,
                  const std::string& dir = "");
// Ql_<kd>.txt and HistQ_<kd>.txt of Q6_GLO (cvd.cpp:1449-1510).
bool write_cvd_q6(int kd, int n, const unsigned* nb, const float* ql, float qglobal, const unsigned long long* histq,
                  const std::string& dir = "");

// Default iostream formatting of a double (%g, 6 significant digits).
std::string g6(double v);

// run_config.txt of one run ("key value" lines; misfit.cpp:369-389,
// cvd.cpp:260-286, elastic.cpp:380-391): model, u0 con0 sig dt dx, steps
// mod seed, the grid; Misfit adds grainx/y/z, axx and b0; elastic grainx/y/z,
// benchel and the input directory when set.
struct RunConfig {
    const char* model = "misfit";  // "misfit", "cvd" or "elastic"
    float u0 = 0, con0 = 0, sig = 0, dt = 0, dx = 0;
    int steps = 0, mod = 0;
    unsigned seed = 0;
    int L = 0, M = 0, N = 1;
    int grainx = 0, grainy = 0, grainz = 0;
    float axx = 0, b0 = 0, benchel = 0;
    std::string indir;
};
bool write_run_config(const std::string& path, const RunConfig& c);
// Noise amplitude b0 = sqrt(axx / dx^3 / dt) (misfit.cpp:174-178).
float noise_b0(float axx, float dx, float dt);
// checkpoint_timestamps.txt: the header, and one "step N timestamp_ms T" line.
bool start_timestamps(const std::string& path);
bool append_timestamp(const std::string& path, int step);
// energy.txt rows "kd E rho" appended (misfit.cpp:1084-1085), n rows of
// (energy, density) pairs labelled kd0, kd0 + 1, ...
bool append_energy(const std::string& path, int kd0, int n, const double* e);

// CollectAvailableCheckpointSteps (elastic.cpp:37-74).
std::vector<int> checkpoint_steps(const std::string& dir);

bool make_dirs(const std::string& path);
long long now_ms();

// Checkpoint files written behind the stepping: a background thread runs the
// posted tasks in order (so appends to one file keep their order); post()
// blocks while `depth` tasks wait, which bounds the host copies in flight.
// drain() returns when every posted task has run; the destructor drains.
class Writer {
public:
    explicit Writer(int depth = 2);
    ~Writer();
    Writer(const Writer&) = delete;
    Writer& operator=(const Writer&) = delete;
    void post(std::function<void()> task);
    void drain();

private:
    void loop();
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<std::function<void()>> q_;
    int depth_;
    bool busy_ = false, stop_ = false;
    std::thread th_;
};

}  // namespace pfcio
