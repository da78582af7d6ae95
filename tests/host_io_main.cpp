// Drives the checkpoint-file writers of pfc-exp-gui_amd/host/pfc_io.cpp on
// inputs from <dir>/in.bin (written by test_host_io.py), so that the test can
// compare every file with its own sequential formatting of the same values.
//
// in.bin: int32 n, L, M, N; float64 x, y, z, v, w [n]; uint32 nb [n];
//         float32 px, py, pz, pp, ql [n]; float32 field [L*M*N]
#include <cstdio>
#include <string>
#include <vector>

#include "../pfc-exp-gui_amd/host/pfc_io.hpp"

template <typename T>
static bool get(FILE* f, std::vector<T>& v, size_t n) {
    v.resize(n);
    return std::fread(v.data(), sizeof(T), n, f) == n;
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string d = argv[1];
    FILE* f = std::fopen((d + "/in.bin").c_str(), "rb");
    if (!f) return 2;
    int hdr[4];
    if (std::fread(hdr, sizeof(int), 4, f) != 4) return 2;
    const int n = hdr[0], L = hdr[1], M = hdr[2], N = hdr[3];
    const size_t P = static_cast<size_t>(L) * M * N;
    std::vector<double> x, y, z, v, w;
    std::vector<unsigned> nb;
    std::vector<float> px, py, pz, pp, ql, field;
    if (!get(f, x, n) || !get(f, y, n) || !get(f, z, n) || !get(f, v, n) || !get(f, w, n) || !get(f, nb, n) ||
        !get(f, px, n) || !get(f, py, n) || !get(f, pz, n) || !get(f, pp, n) || !get(f, ql, n) || !get(f, field, P))
        return 2;
    std::fclose(f);

    pfcio::write_phimax_elastic(d + "/pm.txt", n, x.data(), y.data(), z.data(), v.data());
    pfcio::write_strain(d + "/st.txt", n, x.data(), y.data(), z.data(), v.data(), nb.data());
    pfcio::write_q4q6(d + "/q.txt", 7, n, x.data(), y.data(), z.data(), v.data(), w.data());
    pfcio::write_cvd_atoms(7, L, M, N, 0.5f, 0.3f, 0.8f, n, px.data(), py.data(), pz.data(), pp.data(), d);
    unsigned long long hq[30];
    for (int b = 0; b < 30; b++) hq[b] = 1000ull * b + 7;
    pfcio::write_cvd_q6(7, n, nb.data(), ql.data(), 0.25f, hq, d);

    pfcio::write_vti(d + "/f.vti", "phi", L, M, N, field.data());
    std::vector<float> back(P, -7.0f);
    const long got = pfcio::read_vti(d + "/f.vti", L, M, N, back.data());
    f = std::fopen((d + "/back.bin").c_str(), "wb");
    if (!f) return 2;
    std::fwrite(back.data(), sizeof(float), P, f);
    std::fclose(f);

    // A short file (as the reference's reader allows): the values present fill
    // the field in payload order, the rest stay as they were.
    if (FILE* sf = std::fopen((d + "/short.vti").c_str(), "rb")) {
        std::fclose(sf);
        std::vector<float> part(P, -7.0f);
        const long sgot = pfcio::read_vti(d + "/short.vti", L, M, N, part.data());
        f = std::fopen((d + "/short_back.bin").c_str(), "wb");
        if (!f) return 2;
        std::fwrite(part.data(), sizeof(float), P, f);
        std::fclose(f);
        std::printf("short %ld\n", sgot);
    }

    // Tasks of the background writer run in the order they were posted.
    {
        pfcio::Writer out(2);
        for (int k = 0; k < 200; k++) {
            out.post([k, d] {
                FILE* a = std::fopen((d + "/order.txt").c_str(), "a");
                if (a) { std::fprintf(a, "%d\n", k); std::fclose(a); }
            });
            if (k % 64 == 0) out.drain();
        }
    }
    std::printf("read %ld\n", got);
    return 0;
}
