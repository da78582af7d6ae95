// pfc_serve.cpp - continuous-batching request loop (pfc_server_* in pfc_hip.h).
//
// The reference runs one simulation per pfc-exp-cli process and the GUI
// queues jobs one at a time (MainWindow.cpp:2862-2976,3453-3521).  Here a
// job is a slot in a batch context: jobs of one class (model, precision,
// grid, and the parameters that enter the kernels) share contexts of up to
// max_slots instances, created when no slot of the class is free and sized
// against the HBM budget.  A scheduler thread admits queued jobs into free
// slots while the others keep stepping (idle slots are masked with
// pfc_set_active), steps every busy context concurrently (one HIP stream
// each) to the next job event, and hands checkpoints to a writer thread that
// writes the files the CLI would write (run_config.txt, energy.txt, phi/con
// VTI + PVTI, Phimax_<kd>.txt, checkpoint_timestamps.txt) while stepping
// continues.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pfc_hip.h"
#include "../host/pfc_io.hpp"

namespace pfc {
size_t ctx_bytes(const pfc_params& p, int batch);  // pfc_api.cpp
int api_fail(int code, const char* msg);           // pfc_api.cpp (sets pfc_last_error)
}

namespace {

using Clock = std::chrono::steady_clock;

bool same_class(const pfc_params& a, const pfc_params& b) {
    auto feq = [](float x, float y) { return std::memcmp(&x, &y, sizeof x) == 0; };
    return a.model == b.model && a.precision == b.precision && a.L == b.L && a.M == b.M && a.N == b.N &&
           a.narrow == b.narrow && feq(a.sig, b.sig) && feq(a.dt, b.dt) && feq(a.dx, b.dx) && feq(a.axx, b.axx);
}

struct Job {
    long long id = 0;
    pfc_job spec{};
    std::string outdir;
    std::vector<double> phi0, con0;  // initial fields, built at submit
    int state = PFC_JOB_QUEUED;
    int steps_done = 0;
    double t_submit = 0, t_start = -1, t_end = -1;
    double energy = NAN, density = NAN;
    std::vector<double> phi, con;  // final fields (keep_fields)
    bool cancel = false;
    std::string error;
    std::string write_error;  // set by the writer thread; the job fails at its final bundle
};

struct Batch {
    pfc_params key{};
    pfc_ctx* ctx = nullptr;
    int slots = 0;
    std::vector<Job*> slot_job;  // nullptr = free
    size_t bytes = 0;
    int running() const {
        int n = 0;
        for (Job* j : slot_job) n += j != nullptr;
        return n;
    }
};

// One checkpoint of one job, written by the writer thread.
struct Bundle {
    Job* job = nullptr;
    int kd = 0;
    bool final = false;
    bool first = false;  // also write run_config.txt and the timestamps header
    std::vector<float> phi, con;
    double energy = 0, density = 0;
    int na = 0;
    std::vector<double> x, y, z, ph;
    std::vector<float> co, px, py, pz, pp;
    float phic = 0, pmax = 0;
    // CVD GR histogram; Q6_GLO when kd % 20 == 0 (cvd.cpp:1253-1254)
    std::vector<unsigned> hist, q6nb;
    std::vector<float> q6l;
    float q6g = 0;
    unsigned long long q6h[30] = {};
    bool has_q6 = false;
};

int next_event(const Job& j) {
    const int steps = j.spec.steps;
    if (j.spec.mod <= 0 || j.outdir.empty()) return steps;
    const int nxt = (j.steps_done / j.spec.mod + 1) * j.spec.mod;
    return nxt < steps ? nxt : steps;
}

}  // namespace

struct pfc_server {
    pfc_server_opts o{};
    std::mutex mu;
    std::condition_variable cv_sched, cv_writer, cv_done;
    std::map<long long, std::unique_ptr<Job>> jobs;
    std::deque<Job*> pending;
    std::vector<std::unique_ptr<Batch>> batches;
    std::deque<Bundle> wq;
    std::thread sched, writer;
    bool stop = false, writer_stop = false;
    long long next_id = 1;
    Clock::time_point t0 = Clock::now();
    double budget = 0;
    size_t used = 0;
    int unfinished = 0;  // queued + running + awaiting writes
    double pt_steps = 0; // grid-point steps of finished jobs
    long long done = 0;
    std::string fatal;
    // Busy contexts step the whole graph chunks they share as one group graph
    // (scheduler thread only; rebuilt when the set of stepping contexts
    // changes, dropped before any member is destroyed).
    bool use_group = true;
    pfc_group* grp = nullptr;
    std::vector<pfc_ctx*> grp_key;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

void drop_group(pfc_server* s) {
    if (s->grp) pfc_group_destroy(s->grp);
    s->grp = nullptr;
    s->grp_key.clear();
}

void finish(pfc_server* s, Job* j, int state, double t) {
    // caller holds s->mu
    j->state = state;
    j->t_end = t;
    s->unfinished--;
    if (state == PFC_JOB_DONE) {
        s->done++;
        s->pt_steps += (double)j->spec.p.L * j->spec.p.M * j->spec.p.N * (double)j->spec.steps;
    }
    s->cv_done.notify_all();
}

void write_bundle(const Bundle& b) {
    const Job& j = *b.job;
    const pfc_params& p = j.spec.p;
    const std::string d = j.outdir + "/";
    const std::string kd = std::to_string(b.kd);
    if (b.first) {
        FILE* f = std::fopen((d + "run_config.txt").c_str(), "w");
        if (f) {
            using pfcio::g6;
            std::fprintf(f, "model %s\n", p.model == PFC_MODEL_MISFIT ? "misfit" : "cvd");
            std::fprintf(f, "u0 %s\ncon0 %s\nsig %s\ndt %s\ndx %s\n", g6(p.u0).c_str(), g6(p.con0).c_str(),
                         g6(p.sig).c_str(), g6(p.dt).c_str(), g6(p.dx).c_str());
            std::fprintf(f, "steps %d\nmod %d\nseed %u\nL %d\nM %d\nN %d\n", j.spec.steps, j.spec.mod, p.seed, p.L,
                         p.M, p.N);
            if (p.model == PFC_MODEL_MISFIT)
                std::fprintf(f, "grainx %d\ngrainy %d\ngrainz %d\naxx %s\n", p.grainx, p.grainy, p.grainz,
                             g6(p.axx).c_str());
            std::fprintf(f, "precision %s\n", p.precision == PFC_PREC_FP64 ? "fp64" : "fp32");
            std::fclose(f);
        }
        f = std::fopen((d + "checkpoint_timestamps.txt").c_str(), "w");
        if (f) {
            std::fprintf(f, "# Checkpoint timestamps - Step number and milliseconds since epoch\n");
            std::fclose(f);
        }
    }
    pfcio::write_vti(d + "phi_" + kd + "_0.vti", "phi", p.L, p.M, p.N, b.phi.data());
    pfcio::write_pvti(d + "phi_" + kd + ".pvti", "phi", "phi_", b.kd, p.L, p.M, p.N);
    pfcio::write_vti(d + "con_" + kd + "_0.vti", "con", p.L, p.M, p.N, b.con.data());
    pfcio::write_pvti(d + "con_" + kd + ".pvti", "con", "con_", b.kd, p.L, p.M, p.N);
    if (p.model == PFC_MODEL_MISFIT) {
        pfcio::write_phimax_misfit(d + "Phimax_" + kd + ".txt", (int)b.x.size(), b.x.data(), b.y.data(), b.z.data(),
                                   b.ph.data(), b.co.data());
    } else {
        pfcio::write_cvd_atoms(b.kd, p.L, p.M, p.N, p.dx, b.phic, b.pmax, b.na, b.px.data(), b.py.data(),
                               b.pz.data(), b.pp.data(), j.outdir);
        pfcio::write_cvd_gr(b.kd, p.L, p.M, p.N, p.dx, b.na, b.px.data(), b.hist.data(), (int)b.hist.size(),
                            j.outdir);
        if (b.has_q6) pfcio::write_cvd_q6(b.kd, b.na, b.q6nb.data(), b.q6l.data(), b.q6g, b.q6h, j.outdir);
    }
    FILE* f = std::fopen((d + "energy.txt").c_str(), "a");
    if (f) {
        std::fprintf(f, "%d %s %s\n", b.kd, pfcio::g6(b.energy).c_str(), pfcio::g6(b.density).c_str());
        std::fclose(f);
    }
    f = std::fopen((d + "checkpoint_timestamps.txt").c_str(), "a");
    if (f) {
        std::fprintf(f, "step %d timestamp_ms %lld\n", b.kd, pfcio::now_ms());
        std::fclose(f);
    }
}

// CVD GR every checkpoint, Q6_GLO when kd % 20 == 0 (cvd.cpp:1253-1254), on
// the captured atoms; runs on the writer thread (its own stream), beside
// the scheduler's stepping.
bool cvd_order(Bundle& b, std::string& err) {
    const pfc_params& p = b.job->spec.p;
    int nbin = 0;
    bool ok = pfc_cvd_gr(p.L, p.M, p.N, p.dx, b.na, b.px.data(), b.py.data(), b.pz.data(), nullptr, &nbin) == 0;
    if (ok) {
        b.hist.resize((size_t)nbin);
        ok = pfc_cvd_gr(p.L, p.M, p.N, p.dx, b.na, b.px.data(), b.py.data(), b.pz.data(), b.hist.data(), &nbin) == 0;
    }
    if (ok && b.kd >= 0 && b.kd % 20 == 0 && b.na > 0) {
        b.q6nb.resize((size_t)b.na);
        b.q6l.resize((size_t)b.na);
        ok = pfc_cvd_q6(p.L, p.M, p.N, p.dx, b.na, b.px.data(), b.py.data(), b.pz.data(), b.q6nb.data(),
                        b.q6l.data(), &b.q6g, b.q6h) == 0;
        b.has_q6 = ok;
    }
    if (!ok) err = pfc_last_error();
    return ok;
}

void writer_loop(pfc_server* s) {
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        s->cv_writer.wait(lk, [&] { return s->writer_stop || !s->wq.empty(); });
        if (s->wq.empty()) {
            if (s->writer_stop) return;
            continue;
        }
        Bundle b = std::move(s->wq.front());
        s->wq.pop_front();
        lk.unlock();
        std::string err;
        bool ok = false;
        try {  // a failed checkpoint fails its job, not the writer thread
            ok = b.job->spec.p.model != PFC_MODEL_CVD || cvd_order(b, err);
            if (ok) write_bundle(b);
        } catch (const std::exception& e) {
            ok = false;
            err = e.what();
        }
        lk.lock();
        if (!ok && b.job->write_error.empty()) b.job->write_error = "checkpoint " + std::to_string(b.kd) + ": " + err;
        if (b.final) {
            if (b.job->write_error.empty()) {
                finish(s, b.job, PFC_JOB_DONE, s->now());
            } else {
                b.job->error = b.job->write_error;
                finish(s, b.job, PFC_JOB_FAILED, s->now());
            }
        }
    }
}

// Fill a checkpoint bundle of slot `slot` (scheduler thread, no lock held).
bool capture(Batch& bt, int slot, Bundle& b, const double* e_all, std::string& err) {
    const pfc_params& p = bt.key;
    const size_t P = (size_t)p.L * p.M * p.N;
    b.phi.resize(P);
    b.con.resize(P);
    if (pfc_get_fields_f32(bt.ctx, slot, b.phi.data(), b.con.data())) { err = pfc_last_error(); return false; }
    b.energy = e_all[2 * slot];
    b.density = e_all[2 * slot + 1];
    if (p.model == PFC_MODEL_MISFIT) {
        std::vector<double> x(P), y(P), z(P), ph(P);
        std::vector<float> co(P);
        int n = 0;
        if (pfc_phimax_misfit(bt.ctx, slot, x.data(), y.data(), z.data(), ph.data(), co.data(), (int)P, &n)) {
            err = pfc_last_error();
            return false;
        }
        // The 2-D extraction is repeated for each n (misfit.cpp:3345).
        for (int nn = 0; nn < p.N; nn++)
            for (int i = 0; i < n; i++) {
                b.x.push_back(x[i]);
                b.y.push_back(y[i]);
                b.z.push_back(nn == 0 ? z[i] : (double)((float)(nn + 1) * p.dx));
                b.ph.push_back(ph[i]);
                b.co.push_back(co[i]);
            }
    } else {
        b.px.resize(P); b.py.resize(P); b.pz.resize(P); b.pp.resize(P);
        if (pfc_phimax_cvd(bt.ctx, slot, b.px.data(), b.py.data(), b.pz.data(), b.pp.data(), (int)P, &b.na, &b.phic,
                           &b.pmax)) {
            err = pfc_last_error();
            return false;
        }
        // GR and Q6 need only these atoms: the writer thread computes them.
        b.px.resize((size_t)b.na); b.py.resize((size_t)b.na); b.pz.resize((size_t)b.na); b.pp.resize((size_t)b.na);
    }
    return true;
}

// Admit queued jobs into free slots; creates contexts within the budget.
// Called with the lock held; releases it around GPU work.
void admit(pfc_server* s, std::unique_lock<std::mutex>& lk) {
    for (size_t qi = 0; qi < s->pending.size();) {
        Job* j = s->pending[qi];
        if (j->cancel) {
            s->pending.erase(s->pending.begin() + qi);
            finish(s, j, PFC_JOB_CANCELLED, s->now());
            continue;
        }
        Batch* bt = nullptr;
        int slot = -1;
        for (auto& b : s->batches) {
            if (!same_class(b->key, j->spec.p)) continue;
            for (int i = 0; i < b->slots && slot < 0; i++)
                if (!b->slot_job[i]) { bt = b.get(); slot = i; }
            if (bt) break;
        }
        if (!bt) {
            // New context for this class, as large as the budget allows.
            const size_t one = pfc::ctx_bytes(j->spec.p, 1);
            const size_t per = pfc::ctx_bytes(j->spec.p, 2) - one;
            if ((double)one > s->budget) {
                s->pending.erase(s->pending.begin() + qi);
                j->error = "job exceeds the HBM budget";
                finish(s, j, PFC_JOB_FAILED, s->now());
                continue;
            }
            // Free contexts with no running job (oldest first) until it fits.
            for (size_t bi = 0; bi < s->batches.size() && (double)(s->used + one) > s->budget;) {
                if (s->batches[bi]->running() == 0) {
                    drop_group(s);
                    pfc_destroy(s->batches[bi]->ctx);
                    s->used -= s->batches[bi]->bytes;
                    s->batches.erase(s->batches.begin() + bi);
                } else {
                    bi++;
                }
            }
            if ((double)(s->used + one) > s->budget) {
                qi++;  // waits until running jobs release memory
                continue;
            }
            int slots = s->o.max_slots;
            while (slots > 1 && (double)(s->used + one + per * (slots - 1)) > s->budget) slots--;
            auto nb = std::make_unique<Batch>();
            nb->key = j->spec.p;
            nb->slots = slots;
            nb->slot_job.assign(slots, nullptr);
            nb->bytes = pfc::ctx_bytes(j->spec.p, slots);
            lk.unlock();
            std::string err;
            int rc = pfc_create(&nb->key, slots, &nb->ctx);
            if (rc == 0 && s->o.graph_chunk > 0) rc = pfc_set_graph(nb->ctx, s->o.graph_chunk);
            for (int i = 0; rc == 0 && i < slots; i++) rc = pfc_set_active(nb->ctx, i, 0);
            if (rc) err = pfc_last_error();
            lk.lock();
            if (rc) {
                if (nb->ctx) pfc_destroy(nb->ctx);
                s->pending.erase(s->pending.begin() + qi);
                j->error = "context creation failed: " + err;
                finish(s, j, PFC_JOB_FAILED, s->now());
                continue;
            }
            s->used += nb->bytes;
            bt = nb.get();
            slot = 0;
            s->batches.push_back(std::move(nb));
        }
        // Load the job into its slot.
        s->pending.erase(s->pending.begin() + qi);
        bt->slot_job[slot] = j;
        j->state = PFC_JOB_RUNNING;
        j->t_start = s->now();
        lk.unlock();
        int rc = pfc_set_fields(bt->ctx, slot, j->phi0.data(), j->con0.data());
        if (rc == 0) rc = pfc_set_seed(bt->ctx, slot, j->spec.p.seed);
        if (rc == 0) rc = pfc_set_active(bt->ctx, slot, 1);
        std::string err = rc ? pfc_last_error() : "";
        lk.lock();
        std::vector<double>().swap(j->phi0);
        std::vector<double>().swap(j->con0);
        if (rc) {
            bt->slot_job[slot] = nullptr;
            j->error = "load failed: " + err;
            finish(s, j, PFC_JOB_FAILED, s->now());
        }
    }
}

void sched_loop(pfc_server* s) {
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        admit(s, lk);
        // Plan: every context with running jobs advances to its next job event.
        struct Plan { Batch* b; int n; };
        std::vector<Plan> plan;
        for (auto& b : s->batches) {
            int n = -1;
            for (Job* j : b->slot_job) {
                if (!j) continue;
                const int d = j->cancel ? 0 : next_event(*j) - j->steps_done;
                n = n < 0 ? d : std::min(n, d);
            }
            if (n >= 0) plan.push_back({b.get(), std::min(n, s->o.quantum)});
        }
        if (plan.empty()) {
            if (s->stop) return;
            s->cv_sched.wait(lk, [&] { return s->stop || !s->pending.empty(); });
            continue;
        }
        lk.unlock();
        std::string err;
        // Whole graph chunks every stepping context takes this round run as
        // one group graph; the rest of each context's steps on its own stream.
        const int chunk = s->o.graph_chunk;
        int common = -1, nstep = 0;
        std::vector<pfc_ctx*> key;
        for (auto& pl : plan)
            if (pl.n > 0) {
                common = common < 0 ? pl.n : std::min(common, pl.n);
                nstep++;
                key.push_back(pl.b->ctx);
            }
        int k = (s->use_group && chunk > 0 && nstep >= 2) ? common / chunk : 0;
        if (k > 0 && (!s->grp || key != s->grp_key)) {
            drop_group(s);
            if (pfc_group_create(key.data(), (int)key.size(), chunk, &s->grp) == 0) s->grp_key = key;
            else { s->grp = nullptr; k = 0; }
        }
        if (k > 0 && pfc_group_step(s->grp, k) && err.empty()) err = pfc_last_error();
        for (auto& pl : plan)
            if (pl.n - k * chunk > 0 && pfc_step(pl.b->ctx, pl.n - k * chunk) && err.empty()) err = pfc_last_error();
        for (auto& pl : plan)
            if (pl.n > 0 && pfc_sync(pl.b->ctx) && err.empty()) err = pfc_last_error();
        lk.lock();
        // Events: cancellations, checkpoints, completions.
        for (auto& pl : plan) {
            Batch& b = *pl.b;
            std::vector<int> ev;
            for (int i = 0; i < b.slots; i++) {
                Job* j = b.slot_job[i];
                if (!j) continue;
                const int target = next_event(*j);
                if (!j->cancel) j->steps_done += pl.n;
                if (j->cancel || !err.empty() || j->steps_done == target) ev.push_back(i);
            }
            if (ev.empty()) continue;
            lk.unlock();
            std::vector<double> e(2 * (size_t)b.slots, NAN);
            std::string eerr = err;
            if (eerr.empty() && pfc_energy(b.ctx, e.data())) eerr = pfc_last_error();
            std::vector<Bundle> out;
            std::vector<int> release, failed;
            for (int i : ev) {
                Job* j = b.slot_job[i];
                if (j->cancel || !eerr.empty()) {
                    release.push_back(i);
                    if (!j->cancel) failed.push_back(i);
                    continue;
                }
                const bool final = j->steps_done >= j->spec.steps;
                if (j->spec.keep_fields && final) {
                    const size_t P = (size_t)b.key.L * b.key.M * b.key.N;
                    j->phi.resize(P);
                    j->con.resize(P);
                    if (pfc_get_fields(b.ctx, i, j->phi.data(), j->con.data())) {
                        eerr = pfc_last_error();
                        release.push_back(i);
                        failed.push_back(i);
                        continue;
                    }
                }
                if (final) {
                    j->energy = e[2 * i];
                    j->density = e[2 * i + 1];
                }
                if (!j->outdir.empty()) {
                    Bundle bd;
                    bd.job = j;
                    bd.kd = j->steps_done;
                    bd.final = final;
                    bd.first = j->steps_done <= (j->spec.mod > 0 ? j->spec.mod : j->spec.steps);
                    std::string cerr;
                    bool got = false;
                    try {  // host buffers of the bundle: a failed allocation fails this job only
                        got = capture(b, i, bd, e.data(), cerr);
                    } catch (const std::exception& ex) {
                        cerr = std::string("checkpoint capture: ") + ex.what();
                    }
                    if (!got) {
                        eerr = cerr;
                        release.push_back(i);
                        failed.push_back(i);
                        continue;
                    }
                    out.push_back(std::move(bd));
                }
                if (final) release.push_back(i);
            }
            for (int i : release) pfc_set_active(b.ctx, i, 0);
            lk.lock();
            for (int i : release) {
                Job* j = b.slot_job[i];
                b.slot_job[i] = nullptr;
                const bool fail = std::find(failed.begin(), failed.end(), i) != failed.end();
                if (j->cancel) {
                    finish(s, j, PFC_JOB_CANCELLED, s->now());
                } else if (fail) {
                    j->error = eerr;
                    finish(s, j, PFC_JOB_FAILED, s->now());
                } else if (j->outdir.empty()) {
                    finish(s, j, PFC_JOB_DONE, s->now());
                }  // else: DONE once the writer has written the final bundle
            }
            for (auto& bd : out) s->wq.push_back(std::move(bd));
            if (!out.empty()) s->cv_writer.notify_one();
        }
    }
}

}  // namespace

// Host allocation failures (a job's initial fields, the server's state)
// return PFC_E_NOMEM instead of letting std::bad_alloc cross the C boundary.
#define PFC_SERVE_CATCH                                                                              \
    catch (const std::exception& e) {                                                                \
        return pfc::api_fail(PFC_E_NOMEM, (std::string("host allocation failed: ") + e.what()).c_str()); \
    }

extern "C" int pfc_server_create(const pfc_server_opts* o, pfc_server** out) try {
    if (!out) return pfc::api_fail(PFC_E_ARG, "bad arguments");
    *out = nullptr;
    pfc_server_opts d{};
    d.hbm_budget = 0;
    d.max_slots = 16;
    d.graph_chunk = 25;
    d.quantum = 500;
    if (o) d = *o;
    if (d.max_slots <= 0 || d.graph_chunk < 0 || d.quantum <= 0) return pfc::api_fail(PFC_E_ARG, "bad server options");
    pfc_server* s = new pfc_server();
    s->o = d;
    const char* ge = std::getenv("PFC_SERVE_GROUP");  // "0": each context on its own stream only
    s->use_group = !(ge && ge[0] == '0');
    if (d.hbm_budget > 0) {
        s->budget = d.hbm_budget;
    } else {
        size_t fr = 0, tot = 0;
        if (pfc_device_memory(&fr, &tot)) {
            delete s;
            return PFC_E_HIP;
        }
        s->budget = 0.8 * (double)fr;
    }
    s->sched = std::thread(sched_loop, s);
    s->writer = std::thread(writer_loop, s);
    *out = s;
    return PFC_OK;
} PFC_SERVE_CATCH

extern "C" int pfc_server_submit(pfc_server* s, const pfc_job* job, long long* id) try {
    if (!s || !job || !id || job->steps <= 0 || job->mod < 0) return pfc::api_fail(PFC_E_ARG, "bad arguments");
    const pfc_params& p = job->p;
    if (p.model != PFC_MODEL_MISFIT && p.model != PFC_MODEL_CVD) return pfc::api_fail(PFC_E_ARG, "bad model");
    auto j = std::make_unique<Job>();
    j->spec = *job;
    j->outdir = job->outdir ? job->outdir : "";
    j->spec.outdir = nullptr;
    if (!j->outdir.empty() && !pfcio::make_dirs(j->outdir))
        return pfc::api_fail(PFC_E_ARG, ("cannot create " + j->outdir).c_str());
    const size_t P = (size_t)p.L * p.M * p.N;
    if (p.L <= 0 || p.M <= 0 || p.N <= 0) return pfc::api_fail(PFC_E_ARG, "bad grid");
    j->phi0.resize(P);
    j->con0.resize(P);
    int rc = pfc_init_fields(&p, j->phi0.data(), j->con0.data());
    if (rc) return rc;
    std::lock_guard<std::mutex> g(s->mu);
    if (s->stop) return pfc::api_fail(PFC_E_ARG, "server is shutting down");
    j->id = s->next_id++;
    j->t_submit = s->now();
    *id = j->id;
    s->pending.push_back(j.get());
    s->unfinished++;
    s->jobs[j->id] = std::move(j);
    s->cv_sched.notify_one();
    return PFC_OK;
} PFC_SERVE_CATCH

extern "C" int pfc_server_poll(pfc_server* s, long long id, pfc_job_status* st) {
    if (!s || !st) return pfc::api_fail(PFC_E_ARG, "bad arguments");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return pfc::api_fail(PFC_E_ARG, "unknown job id");
    const Job& j = *it->second;
    st->state = j.state;
    st->steps_done = j.steps_done;
    st->t_submit = j.t_submit;
    st->t_start = j.t_start;
    st->t_end = j.t_end;
    st->energy = j.energy;
    st->density = j.density;
    return PFC_OK;
}

extern "C" int pfc_server_result(pfc_server* s, long long id, double* phi, double* con) {
    if (!s) return pfc::api_fail(PFC_E_ARG, "bad arguments");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return pfc::api_fail(PFC_E_ARG, "unknown job id");
    const Job& j = *it->second;
    if (j.state != PFC_JOB_DONE) return pfc::api_fail(PFC_E_ARG, "job is not done");
    if (!j.spec.keep_fields) return pfc::api_fail(PFC_E_ARG, "job was submitted without keep_fields");
    if (phi) std::copy(j.phi.begin(), j.phi.end(), phi);
    if (con) std::copy(j.con.begin(), j.con.end(), con);
    return PFC_OK;
}

extern "C" int pfc_server_error(pfc_server* s, long long id, char* buf, int cap) {
    if (!s || !buf || cap <= 0) return pfc::api_fail(PFC_E_ARG, "bad arguments");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return pfc::api_fail(PFC_E_ARG, "unknown job id");
    std::snprintf(buf, (size_t)cap, "%s", it->second->error.c_str());
    return PFC_OK;
}

extern "C" int pfc_server_cancel(pfc_server* s, long long id) {
    if (!s) return pfc::api_fail(PFC_E_ARG, "bad arguments");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return pfc::api_fail(PFC_E_ARG, "unknown job id");
    it->second->cancel = true;
    s->cv_sched.notify_one();
    return PFC_OK;
}

extern "C" int pfc_server_drain(pfc_server* s, double timeout_s) {
    if (!s) return pfc::api_fail(PFC_E_ARG, "bad arguments");
    std::unique_lock<std::mutex> lk(s->mu);
    auto pred = [&] { return s->unfinished == 0; };
    if (timeout_s < 0) {
        s->cv_done.wait(lk, pred);
        return PFC_OK;
    }
    if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred))
        return pfc::api_fail(PFC_E_ARG, "drain timed out");
    return PFC_OK;
}

extern "C" int pfc_server_stats(pfc_server* s, pfc_server_info* st) {
    if (!s || !st) return pfc::api_fail(PFC_E_ARG, "bad arguments");
    std::lock_guard<std::mutex> g(s->mu);
    st->now = s->now();
    st->jobs_done = s->done;
    st->jobs_unfinished = s->unfinished;
    st->point_steps = s->pt_steps;
    st->contexts = (int)s->batches.size();
    st->bytes_in_use = (double)s->used;
    st->budget = s->budget;
    return PFC_OK;
}

extern "C" int pfc_server_destroy(pfc_server* s) {
    if (!s) return PFC_OK;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
        for (auto& kv : s->jobs) kv.second->cancel = true;
        s->cv_sched.notify_all();
    }
    s->sched.join();
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->writer_stop = true;
        s->cv_writer.notify_all();
    }
    s->writer.join();
    drop_group(s);
    for (auto& b : s->batches) pfc_destroy(b->ctx);
    delete s;
    return PFC_OK;
}
