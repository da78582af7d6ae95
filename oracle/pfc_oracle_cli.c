/*
 * pfc_oracle_cli.c - reference-contract driver over the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY: produces the files the reference pfc-exp-cli would
 * write (run_config.txt, checkpoint_timestamps.txt, energy.txt, phi/con VTI and
 * PVTI, Phimax_*.txt, phimax.txt, phiball_*.cfg, strain_*.txt) so tests can
 * compare the product CLI against it.  Follows main_cli.cpp:24-88,
 * misfit.cpp:181-531, cvd.cpp:127-454, elastic.cpp:30-74,188-521.
 * Runtime grid flags --L/--M/--N replace the compile-time PFC_GRID_*;
 * --narrow 0 switches off the reference's float narrowing.
 */
#define _POSIX_C_SOURCE 200809L
#include "pfc_oracle.h"

#include <dirent.h>
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/time.h>
#include <unistd.h>

static int parse_int(const char* v, int* out) {
    if (!v) return 0;
    char* end = NULL;
    errno = 0;
    long x = strtol(v, &end, 10);
    if (errno || end == v || *end) return 0;
    *out = (int)x;
    return 1;
}
static int parse_uint(const char* v, unsigned* out) {
    if (!v) return 0;
    char* end = NULL;
    errno = 0;
    unsigned long long x = strtoull(v, &end, 10);
    if (errno || end == v || *end || x > 0xffffffffULL) return 0;
    *out = (unsigned)x;
    return 1;
}
static int parse_float(const char* v, float* out) {
    if (!v) return 0;
    char* end = NULL;
    errno = 0;
    float x = strtof(v, &end);
    if (errno || end == v || *end) return 0;
    *out = x;
    return 1;
}

static int mkdirs(const char* path) {
    char buf[4096];
    size_t n = strlen(path);
    if (n == 0 || n >= sizeof(buf)) return -1;
    memcpy(buf, path, n + 1);
    for (size_t i = 1; i <= n; i++) {
        if (buf[i] == '/' || buf[i] == 0) {
            char c = buf[i];
            buf[i] = 0;
            if (mkdir(buf, 0777) != 0 && errno != EEXIST) return -1;
            buf[i] = c;
        }
    }
    return 0;
}

static long long now_ms(void) {
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return (long long)tv.tv_sec * 1000 + tv.tv_usec / 1000;
}

typedef struct {
    pfo_model p;
    char outdir[4096];
    char indir[4096];
    float benchel;
} cli_t;

/* Shared flag parser: returns 0 ok, 1 help, 2 bad. `strict` rejects unknown
 * flags (misfit.cpp:269-272); cvd/elastic ignore them (cvd.cpp:192-193). */
static int parse_args(cli_t* c, int argc, char** argv, int strict, int model) {
    for (int i = 1; i < argc; i++) {
        const char* a = argv[i];
        if (!strcmp(a, "-h") || !strcmp(a, "--help")) return 1;
        const char* names[] = {"--u0", "--con0", "--sig", "--dt", "--dx", "--steps", "--mod", "--seed",
                               "--grainx", "--grainy", "--grainz", "--axx", "--outdir", "--L", "--M",
                               "--N", "--narrow", "--benchel", "--input-dir", NULL};
        int k = -1;
        const char* v = NULL;
        for (int q = 0; names[q]; q++) {
            size_t ln = strlen(names[q]);
            if (!strcmp(a, names[q])) {
                k = q;
                v = (i + 1 < argc) ? argv[++i] : NULL;
                break;
            }
            if (!strncmp(a, names[q], ln) && a[ln] == '=') {
                k = q;
                v = a + ln + 1;
                break;
            }
        }
        if (k < 0) {
            if (strict) { fprintf(stderr, "Unknown option: %s\n", a); return 2; }
            continue;
        }
        /* Flags a model does not know are ignored like unknown flags. */
        if (model != 2 && (k == 17 || k == 18)) { if (strict) return 2; continue; }
        if (model == PFO_CVD && (k >= 8 && k <= 11)) continue;
        int ok = 1;
        switch (k) {
            case 0: ok = parse_float(v, &c->p.u0); break;
            case 1: ok = parse_float(v, &c->p.con0); break;
            case 2: ok = parse_float(v, &c->p.sig); break;
            case 3: ok = parse_float(v, &c->p.dt); break;
            case 4: ok = parse_float(v, &c->p.dx); break;
            case 5: ok = parse_int(v, &c->p.steps); break;
            case 6: ok = parse_int(v, &c->p.mod); break;
            case 7: ok = parse_uint(v, &c->p.seed); break;
            case 8: ok = parse_int(v, &c->p.grainx); break;
            case 9: ok = parse_int(v, &c->p.grainy); break;
            case 10: ok = parse_int(v, &c->p.grainz); break;
            case 11: ok = parse_float(v, &c->p.axx); break;
            case 12: ok = v && *v; if (ok) snprintf(c->outdir, sizeof c->outdir, "%s", v); break;
            case 13: ok = parse_int(v, &c->p.L); break;
            case 14: ok = parse_int(v, &c->p.M); break;
            case 15: ok = parse_int(v, &c->p.N); break;
            case 16: ok = parse_int(v, &c->p.narrow); break;
            case 17: ok = parse_float(v, &c->benchel); break;
            case 18: ok = v && *v; if (ok) snprintf(c->indir, sizeof c->indir, "%s", v); break;
        }
        if (!ok) return 2;
    }
    return 0;
}

static int setup_outdir(const char* d) {
    if (!d[0]) return 0;
    if (mkdirs(d) != 0) { fprintf(stderr, "Failed to create --outdir '%s'\n", d); return -1; }
    if (chdir(d) != 0) { fprintf(stderr, "Failed to chdir to --outdir '%s'\n", d); return -1; }
    return 0;
}

static void write_fields(const pfo_model* p, int kd, const double* phi, const double* con) {
    char name[64];
    snprintf(name, sizeof name, "phi_%d_0.vti", kd);
    pfo_write_vti(name, "phi", p->L, p->M, p->N, phi);
    snprintf(name, sizeof name, "phi_%d.pvti", kd);
    pfo_write_pvti(name, "phi", "phi_", kd, p->L, p->M, p->N);
    snprintf(name, sizeof name, "con_%d_0.vti", kd);
    pfo_write_vti(name, "con", p->L, p->M, p->N, con);
    snprintf(name, sizeof name, "con_%d.pvti", kd);
    pfo_write_pvti(name, "con", "con_", kd, p->L, p->M, p->N);
}

static pfo_atoms atoms_alloc(int cap) {
    pfo_atoms a;
    a.n = 0; a.cap = cap;
    a.x = (double*)malloc(sizeof(double) * cap);
    a.y = (double*)malloc(sizeof(double) * cap);
    a.z = (double*)malloc(sizeof(double) * cap);
    a.phi = (double*)malloc(sizeof(double) * cap);
    a.con = (float*)malloc(sizeof(float) * cap);
    return a;
}
static void atoms_free(pfo_atoms* a) { free(a->x); free(a->y); free(a->z); free(a->phi); free(a->con); }

static void cvd_outputs(const pfo_model* p, int kd, const double* phi) {
    const int L = p->L, M = p->M, N = p->N;
    const int cap = L * M * N;
    float* px = (float*)malloc(sizeof(float) * cap);
    float* py = (float*)malloc(sizeof(float) * cap);
    float* pz = (float*)malloc(sizeof(float) * cap);
    float* pp = (float*)malloc(sizeof(float) * cap);
    float phic, phimax;
    int na = pfo_phimax_cvd(L, M, N, p->dx, p->u0, phi, px, py, pz, pp, cap, &phic, &phimax);
    FILE* f = fopen("phimax.txt", "a");
    fprintf(f, "%d , %g , %g , %d\n", kd, phic, phimax, 0);
    fprintf(f, "phimax is done myid=0\n");
    fprintf(f, " buf is done \n");
    fprintf(f, " pc_xyz is initializing \n");
    fprintf(f, " pc_xyz is done \n");
    fprintf(f, " pc_xyz has been gathered \n");
    fclose(f);
    const float frame_x = L * p->dx, frame_y = M * p->dx, frame_z = N * p->dx;
    char name[64];
    snprintf(name, sizeof name, "phiball_%d.cfg", kd);
    f = fopen(name, "w");
    fprintf(f, "Number of particles = %d\n", na);
    fprintf(f, "A = %.4f Angstrom (basic length-scale)\n", 10.0);
    fprintf(f, "H0(1,1) = %.4f A \n", frame_y);
    fprintf(f, "H0(1,2) = 0 A \nH0(1,3) = 0 A \nH0(2,1) = 0 A \n");
    fprintf(f, "H0(2,2) = %.4f A \n", frame_z);
    fprintf(f, "H0(2,3) = 0 A \nH0(3,1) = 0 A \nH0(3,2) = 0 A \n");
    fprintf(f, "H0(3,3) = %.4f A \n", frame_x);
    fprintf(f, ".NO_VELOCITY.\nentry_count = 3\n63.546\nCu\n");
    for (int i = 0; i < na; i++) fprintf(f, "%.4f %.4f %.4f\n", px[i], py[i], pz[i]);
    fclose(f);
    snprintf(name, sizeof name, "Phimax_%d.txt", kd);
    f = fopen(name, "w");
    fprintf(f, "%d\nKommentar\n", na);
    for (int i = 0; i < na; i++)
        fprintf(f, "%.4f %.4f %.4f %.4f\n", pz[i] * frame_x, px[i] * frame_y, py[i] * frame_z, pp[i]);
    fclose(f);
    /* GR every checkpoint, Q6_GLO when kd % 20 == 0 (cvd.cpp:1253-1254) */
    unsigned* hist = (unsigned*)malloc(sizeof(unsigned) * (size_t)(pfo_cvd_gr_bins(L, p->dx) + 1));
    pfo_cvd_gr(L, M, N, p->dx, na, px, py, pz, hist);
    pfo_write_cvd_gr(kd, L, M, N, p->dx, na, px, hist);
    free(hist);
    if (kd >= 0 && kd % 20 == 0) {
        unsigned* nb = (unsigned*)malloc(sizeof(unsigned) * (size_t)na);
        float* ql = (float*)malloc(sizeof(float) * (size_t)na);
        unsigned long long histq[30];
        float qg;
        pfo_cvd_q6(L, M, N, p->dx, na, px, py, pz, nb, ql, &qg, histq);
        pfo_write_cvd_q6(kd, na, nb, ql, qg, histq);
        free(nb); free(ql);
    }
    free(px); free(py); free(pz); free(pp);
}

static int run_sim(cli_t* c) {
    pfo_model* p = &c->p;
    const int misfit = (p->model == PFO_MISFIT);
    if (p->steps <= 0 || p->dt <= 0.0f || p->dx <= 0.0f || p->mod <= 0) return 2;
    if (misfit && (p->grainx <= 0 || p->grainy <= 0 || p->grainz <= 0)) return 2;
    if (misfit && (p->grainx > p->L || p->grainy > p->M || p->grainz > p->N)) return 2;
    pfo_derive(p);
    if (setup_outdir(c->outdir) != 0) return 2;
    FILE* f = fopen("run_config.txt", "w");
    fprintf(f, "model %s\n", misfit ? "misfit" : "cvd");
    fprintf(f, "u0 %g\ncon0 %g\nsig %g\ndt %g\ndx %g\n", p->u0, p->con0, p->sig, p->dt, p->dx);
    fprintf(f, "steps %d\nmod %d\nseed %u\nL %d\nM %d\nN %d\n", p->steps, p->mod, p->seed, p->L, p->M, p->N);
    if (misfit)
        fprintf(f, "grainx %d\ngrainy %d\ngrainz %d\naxx %g\nb0 %g\n", p->grainx, p->grainy, p->grainz, p->axx, p->b0);
    fclose(f);
    FILE* ck = fopen("checkpoint_timestamps.txt", "w");
    fprintf(ck, "# Checkpoint timestamps - Step number and milliseconds since epoch\n");
    fflush(ck);

    const size_t P = (size_t)p->L * p->M * p->N;
    double* phi = (double*)malloc(sizeof(double) * P);
    double* con = (double*)malloc(sizeof(double) * P);
    pfo_init_fields(p, phi, con);
    pfo_sim* s = pfo_create(p);
    pfo_set_fields(s, phi, con);
    if (misfit) write_fields(p, 0, phi, con);
    const int progress_every = p->steps / 200 > 1 ? p->steps / 200 : 1;
    pfo_atoms at = atoms_alloc((int)P);
    for (int t = 0; t < p->steps; t++) {
        if (!misfit) { fprintf(stderr, "PFC_PROGRESS step=%d total=%d\n", t + 1, p->steps); fflush(stderr); }
        pfo_step(s, t);
        pfo_get_fields(s, phi, con);
        if (misfit) {
            const int os = t + 1;
            if (os == 1 || os % progress_every == 0 || os == p->steps) {
                printf("PFC_PROGRESS step=%d total=%d\n", os, p->steps);
                fflush(stdout);
            }
            if (os % p->mod == 0 || os == p->steps) {
                const int kd = os;
                char name[64];
                pfo_phimax_misfit(p->L, p->M, p->N, p->dx, p->con0, phi, con, &at);
                snprintf(name, sizeof name, "Phimax_%d.txt", kd);
                pfo_write_phimax_misfit(name, &at);
                write_fields(p, kd, phi, con);
                double e, rho;
                pfo_energy(s, &e, &rho);
                FILE* ef = fopen("energy.txt", "a");
                fprintf(ef, "%d %g %g\n", kd, e, rho);
                fclose(ef);
                fprintf(ck, "step %d timestamp_ms %lld\n", os, now_ms());
                fflush(ck);
            }
        } else {
            const int kd = t;
            if (t % p->mod == 0) {
                cvd_outputs(p, kd, phi);
                write_fields(p, kd, phi, con);
                fprintf(ck, "step %d timestamp_ms %lld\n", t, now_ms());
                fflush(ck);
            }
            double e, rho;
            pfo_energy(s, &e, &rho);
            FILE* ef = fopen("energy.txt", "a");
            fprintf(ef, "%d %g %g\n", kd, e, rho);
            fclose(ef);
        }
    }
    fclose(ck);
    atoms_free(&at);
    pfo_destroy(s);
    free(phi); free(con);
    return 0;
}

static int cmp_int(const void* a, const void* b) { return (*(const int*)a > *(const int*)b) - (*(const int*)a < *(const int*)b); }

/* CollectAvailableCheckpointSteps (elastic.cpp:37-74). */
static int collect_steps(const char* dir, int* steps, int cap) {
    DIR* d = opendir(dir[0] ? dir : ".");
    if (!d) return 0;
    int phi_s[8192], con_s[8192], np = 0, nc = 0;
    struct dirent* e;
    while ((e = readdir(d))) {
        int st, rk;
        if (sscanf(e->d_name, "phi_%d_%d.vti", &st, &rk) == 2) { if (np < 8192) phi_s[np++] = st; continue; }
        if (sscanf(e->d_name, "con_%d_%d.vti", &st, &rk) == 2) { if (nc < 8192) con_s[nc++] = st; }
    }
    closedir(d);
    qsort(phi_s, np, sizeof(int), cmp_int);
    int n = 0, has_pos = 0;
    for (int i = 0; i < np; i++) {
        if (i > 0 && phi_s[i] == phi_s[i - 1]) continue;
        int found = 0;
        for (int j = 0; j < nc; j++) if (con_s[j] == phi_s[i]) { found = 1; break; }
        if (!found) continue;
        if (phi_s[i] > 0) has_pos = 1;
        if (n < cap) steps[n++] = phi_s[i];
    }
    if (has_pos) {
        int k = 0;
        for (int i = 0; i < n; i++) if (steps[i] != 0) steps[k++] = steps[i];
        n = k;
    }
    return n;
}

static int run_elastic(cli_t* c) {
    pfo_model* p = &c->p;
    if (p->steps <= 0 || p->dt <= 0.0f || p->dx <= 0.0f || p->mod <= 0) return 2;
    char indir[4096] = "";
    if (c->indir[0]) {
        if (c->indir[0] == '/') snprintf(indir, sizeof indir, "%s", c->indir);
        else {
            char cwd[2048];
            if (!getcwd(cwd, sizeof cwd)) return 2;
            const int w = snprintf(indir, sizeof indir, "%s/%s", cwd, c->indir);
            if (w < 0 || w >= (int)sizeof indir) return 2;
        }
    }
    if (setup_outdir(c->outdir) != 0) return 2;
    FILE* f = fopen("run_config.txt", "w");
    if (!f) return 2;
    fprintf(f, "model elastic\nu0 %g\ncon0 %g\nsig %g\ndt %g\ndx %g\n", p->u0, p->con0, p->sig, p->dt, p->dx);
    fprintf(f, "steps %d\nmod %d\nseed %u\nL %d\nM %d\nN %d\n", p->steps, p->mod, p->seed, p->L, p->M, p->N);
    fprintf(f, "grainx %d\ngrainy %d\ngrainz %d\nbenchel %g\n", p->grainx, p->grainy, p->grainz, c->benchel);
    if (indir[0]) fprintf(f, "input_dir %s\n", indir);
    fclose(f);
    FILE* ck = fopen("checkpoint_timestamps.txt", "w");
    fprintf(ck, "# Checkpoint timestamps - Step number and milliseconds since epoch\n");
    fclose(ck);
    static int steps[8192];
    int ns = collect_steps(indir, steps, 8192);
    int nlist = ns;
    if (ns == 0) {
        for (int t = 0; t < p->steps; t++) {
            const int os = t + 1;
            if (os % p->mod == 0 || os == p->steps) steps[nlist++] = os;
            if (nlist >= 8192) break;
        }
    }
    const size_t P = (size_t)p->L * p->M * p->N;
    float* phi = (float*)malloc(sizeof(float) * P);
    pfo_atoms at = atoms_alloc((int)P);
    double* strain = (double*)malloc(sizeof(double) * P);
    unsigned* nb = (unsigned*)malloc(sizeof(unsigned) * P);
    double* q4 = (double*)malloc(sizeof(double) * P);
    double* q6 = (double*)malloc(sizeof(double) * P);
    for (int i = 0; i < nlist; i++) {
        const int kd = steps[i];
        fprintf(stderr, "PFC_PROGRESS step=%d total=%d\n", i + 1, nlist);
        fflush(stderr);
        char path[8192];
        snprintf(path, sizeof path, "%s%sphi_%d_0.vti", indir, indir[0] ? "/" : "", kd);
        if (pfo_read_vti(path, p->L, p->M, p->N, phi) < 0) { fprintf(stderr, "can't open input file: %s\n", path); exit(2); }
        snprintf(path, sizeof path, "%s%scon_%d_0.vti", indir, indir[0] ? "/" : "", kd);
        FILE* cf = fopen(path, "rb");
        if (!cf) { fprintf(stderr, "can't open input file: %s\n", path); exit(2); }
        fclose(cf);
        pfo_phimax_elastic(p->L, p->M, p->N, p->dx, phi, &at);
        char name[64];
        snprintf(name, sizeof name, "Phimax_%d.txt", kd);
        pfo_write_phimax_elastic(name, &at);
        pfo_strain(p->L, p->M, p->N, p->dx, c->benchel, at.n, at.x, at.y, at.z, strain, nb);
        snprintf(name, sizeof name, "strain_%d.txt", kd);
        pfo_write_strain(name, &at, strain, nb);
        pfo_cryst_order(p->L, p->M, p->N, p->dx, at.n, at.x, at.y, at.z, q4, q6, NULL);
        snprintf(name, sizeof name, "q4q6_%d.txt", kd);
        pfo_write_q4q6(name, &at, q4, q6);
    }
    free(phi); free(strain); free(nb); free(q4); free(q6);
    atoms_free(&at);
    return 0;
}

int main(int argc, char** argv) {
    const char* model = "misfit";
    char** av = (char**)malloc(sizeof(char*) * (argc + 1));
    int ac = 0;
    av[ac++] = argv[0];
    for (int i = 1; i < argc; i++) {
        if (!strcmp(argv[i], "--model") || !strcmp(argv[i], "-m")) {
            if (i + 1 >= argc) { fprintf(stderr, "Missing value after --model\n"); free(av); return 2; }
            model = argv[++i];
            continue;
        }
        if (!strncmp(argv[i], "--model=", 8)) { model = argv[i] + 8; continue; }
        av[ac++] = argv[i];
    }
    int L = 256, M = 256, N = 1;
    for (int i = 1; i < ac; i++) {
        if (!strcmp(av[i], "--L") && i + 1 < ac) L = atoi(av[i + 1]);
        if (!strcmp(av[i], "--M") && i + 1 < ac) M = atoi(av[i + 1]);
        if (!strcmp(av[i], "--N") && i + 1 < ac) N = atoi(av[i + 1]);
    }
    cli_t c;
    memset(&c, 0, sizeof c);
    c.benchel = 1.7f;
    int kind;
    if (!strcmp(model, "misfit")) kind = PFO_MISFIT;
    else if (!strcmp(model, "cvd")) kind = PFO_CVD;
    else if (!strcmp(model, "elastic")) kind = 2;
    else { fprintf(stderr, "Unknown model: %s\n", model); free(av); return 2; }
    pfo_defaults(&c.p, kind == 2 ? PFO_MISFIT : kind, L, M, N);
    if (kind == 2) {
        /* elastic.cpp:83-105 defaults */
        c.p.dt = 0.1f; c.p.steps = 2002; c.p.mod = 200; c.p.grainx = 2; c.p.grainy = 2; c.p.grainz = 1;
    }
    int rc = parse_args(&c, ac, av, kind == PFO_MISFIT, kind);
    free(av);
    if (rc == 1) return 0;
    if (rc != 0) return 2;
    if (c.p.L <= 0 || c.p.M <= 0 || c.p.N <= 0) return 2;
    if (kind == 2) return run_elastic(&c);
    return run_sim(&c);
}
