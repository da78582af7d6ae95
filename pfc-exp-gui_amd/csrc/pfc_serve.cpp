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
// writes the files the CLI would write, on the CLI's schedule (host/pfc_cli.cpp
// run_sim: Misfit kd = 0 fields, then a checkpoint every mod steps and at the
// end; CVD energy after every step and a checkpoint after step t when
// t % mod == 0, cvd.cpp:389-454) while stepping continues.
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
    double t_stepped = -1;    // last step done (before the final files)
    double create_s = 0;      // context creation this job waited for
    int energy_rows = 0;      // CVD: per-step energy rows handed to the writer
};

struct Batch {
    pfc_params key{};
    pfc_ctx* ctx = nullptr;
    int slots = 0;
    std::vector<Job*> slot_job;  // nullptr = free
    size_t bytes = 0;
    bool busy = false;    // steps in flight
    int n = 0;            // steps of the launch in flight
    bool traced = false;  // that launch traces the energy of every step
    int running() const {
        int n = 0;
        for (Job* j : slot_job) n += j != nullptr;
        return n;
    }
};

// Files of one job written by the writer thread: the start of the run
// (run_config.txt, the timestamps header, Misfit's kd = 0 fields), CVD
// energy rows, or a checkpoint.  `final` marks the job's last bundle.
enum { B_START = 0, B_ENERGY = 1, B_CKPT = 2 };
struct Bundle {
    Job* job = nullptr;
    int kind = B_CKPT;
    int kd = 0;
    bool final = false;
    std::vector<double> erows;  // B_ENERGY: (energy, density) per step from step index kd
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

// Steps done at the job's next checkpoint (or its end).  Misfit: every mod
// steps; CVD: after step index t with t % mod == 0, i.e. steps done = t + 1.
int next_event(const Job& j) {
    const int steps = j.spec.steps, mod = j.spec.mod, s = j.steps_done;
    if (mod <= 0 || j.outdir.empty()) return steps;
    int nxt;
    if (j.spec.p.model == PFC_MODEL_CVD) nxt = s < 1 ? 1 : ((s - 1) / mod + 1) * mod + 1;
    else nxt = (s / mod + 1) * mod;
    return nxt < steps ? nxt : steps;
}

// Whether steps_done = s is a checkpoint of the CLI's schedule.
bool is_checkpoint(const Job& j, int s) {
    const int mod = j.spec.mod;
    if (mod <= 0) return s == j.spec.steps;
    if (j.spec.p.model == PFC_MODEL_CVD) return (s - 1) % mod == 0;
    return s % mod == 0 || s == j.spec.steps;
}

}  // namespace

// P-sized host arrays of the checkpoint extraction, reused across
// checkpoints (scheduler thread only); a bundle gets copies of the atoms.
struct CaptureScratch {
    std::vector<double> x, y, z, ph;
    std::vector<float> co, px, py, pz, pp;
    void fit(size_t P) {
        if (x.size() >= P) return;
        x.resize(P); y.resize(P); z.resize(P); ph.resize(P);
        co.resize(P); px.resize(P); py.resize(P); pz.resize(P); pp.resize(P);
    }
};

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
    CaptureScratch cap;  // scheduler thread only

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

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

bool write_fields(const std::string& d, int kd, const pfc_params& p, const float* a, const float* b) {
    const std::string k = std::to_string(kd);
    bool ok = pfcio::write_vti(d + "phi_" + k + "_0.vti", "phi", p.L, p.M, p.N, a);
    ok = pfcio::write_pvti(d + "phi_" + k + ".pvti", "phi", "phi_", kd, p.L, p.M, p.N) && ok;
    ok = pfcio::write_vti(d + "con_" + k + "_0.vti", "con", p.L, p.M, p.N, b) && ok;
    ok = pfcio::write_pvti(d + "con_" + k + ".pvti", "con", "con_", kd, p.L, p.M, p.N) && ok;
    return ok;
}

// The files of one bundle, in the CLI's order (host/pfc_cli.cpp run_sim).
bool write_bundle(const Bundle& b) {
    const Job& j = *b.job;
    const pfc_params& p = j.spec.p;
    const std::string d = j.outdir + "/";
    const bool misfit = p.model == PFC_MODEL_MISFIT;
    if (b.kind == B_START) {
        pfcio::RunConfig c;
        c.model = misfit ? "misfit" : "cvd";
        c.u0 = p.u0; c.con0 = p.con0; c.sig = p.sig; c.dt = p.dt; c.dx = p.dx;
        c.steps = j.spec.steps; c.mod = j.spec.mod; c.seed = p.seed;
        c.L = p.L; c.M = p.M; c.N = p.N;
        c.grainx = p.grainx; c.grainy = p.grainy; c.grainz = p.grainz;
        c.axx = p.axx; c.b0 = pfcio::noise_b0(p.axx, p.dx, p.dt);
        bool ok = pfcio::write_run_config(d + "run_config.txt", c);
        ok = pfcio::start_timestamps(d + "checkpoint_timestamps.txt") && ok;
        if (misfit) ok = write_fields(d, 0, p, b.phi.data(), b.con.data()) && ok;  // kd = 0, before stepping
        return ok;
    }
    if (b.kind == B_ENERGY)
        return b.erows.empty() || pfcio::append_energy(d + "energy.txt", b.kd, (int)(b.erows.size() / 2), b.erows.data());
    bool ok = true;
    if (misfit) {
        ok = pfcio::write_phimax_misfit(d + "Phimax_" + std::to_string(b.kd) + ".txt", (int)b.x.size(), b.x.data(),
                                        b.y.data(), b.z.data(), b.ph.data(), b.co.data());
        ok = write_fields(d, b.kd, p, b.phi.data(), b.con.data()) && ok;
        const double e[2] = {b.energy, b.density};
        ok = pfcio::append_energy(d + "energy.txt", b.kd, 1, e) && ok;
    } else {
        ok = pfcio::write_cvd_atoms(b.kd, p.L, p.M, p.N, p.dx, b.phic, b.pmax, b.na, b.px.data(), b.py.data(),
                                    b.pz.data(), b.pp.data(), j.outdir);
        ok = pfcio::write_cvd_gr(b.kd, p.L, p.M, p.N, p.dx, b.na, b.px.data(), b.hist.data(), (int)b.hist.size(),
                                 j.outdir) && ok;
        if (b.has_q6) ok = pfcio::write_cvd_q6(b.kd, b.na, b.q6nb.data(), b.q6l.data(), b.q6g, b.q6h, j.outdir) && ok;
        ok = write_fields(d, b.kd, p, b.phi.data(), b.con.data()) && ok;
    }
    return pfcio::append_timestamp(d + "checkpoint_timestamps.txt", b.kd) && ok;
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
            const bool order = b.kind == B_CKPT && b.job->spec.p.model == PFC_MODEL_CVD;
            ok = !order || cvd_order(b, err);
            if (ok && !(ok = write_bundle(b))) err = "cannot write the files in " + b.job->outdir;
        } catch (const std::exception& e) {
            ok = false;
            err = e.what();
        }
        lk.lock();
        if (!ok && b.job->write_error.empty()) b.job->write_error = "files of step " + std::to_string(b.kd) + ": " + err;
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
bool capture(pfc_server* s, Batch& bt, int slot, Bundle& b, const double* e_all, std::string& err) {
    const pfc_params& p = bt.key;
    const size_t P = (size_t)p.L * p.M * p.N;
    b.phi.resize(P);
    b.con.resize(P);
    if (pfc_get_fields_f32(bt.ctx, slot, b.phi.data(), b.con.data())) { err = pfc_last_error(); return false; }
    b.energy = e_all[2 * slot];
    b.density = e_all[2 * slot + 1];
    CaptureScratch& c = s->cap;
    c.fit(P);
    if (p.model == PFC_MODEL_MISFIT) {
        int n = 0;
        if (pfc_phimax_misfit(bt.ctx, slot, c.x.data(), c.y.data(), c.z.data(), c.ph.data(), c.co.data(), (int)P, &n)) {
            err = pfc_last_error();
            return false;
        }
        // The 2-D extraction is repeated for each n (misfit.cpp:3345).
        const size_t k = (size_t)n * p.N;
        b.x.reserve(k); b.y.reserve(k); b.z.reserve(k); b.ph.reserve(k); b.co.reserve(k);
        for (int nn = 0; nn < p.N; nn++)
            for (int i = 0; i < n; i++) {
                b.x.push_back(c.x[i]);
                b.y.push_back(c.y[i]);
                b.z.push_back(nn == 0 ? c.z[i] : (double)((float)(nn + 1) * p.dx));
                b.ph.push_back(c.ph[i]);
                b.co.push_back(c.co[i]);
            }
    } else {
        if (pfc_phimax_cvd(bt.ctx, slot, c.px.data(), c.py.data(), c.pz.data(), c.pp.data(), (int)P, &b.na, &b.phic,
                           &b.pmax)) {
            err = pfc_last_error();
            return false;
        }
        // GR and Q6 need only these atoms: the writer thread computes them.
        b.px.assign(c.px.begin(), c.px.begin() + b.na);
        b.py.assign(c.py.begin(), c.py.begin() + b.na);
        b.pz.assign(c.pz.begin(), c.pz.begin() + b.na);
        b.pp.assign(c.pp.begin(), c.pp.begin() + b.na);
    }
    return true;
}

// The first bundle of a job with an output directory (lock held): run
// configuration, timestamps header, Misfit's kd = 0 fields.
void start_bundle(pfc_server* s, Job* j) {
    Bundle b;
    b.job = j;
    b.kind = B_START;
    if (j->spec.p.model == PFC_MODEL_MISFIT) {
        b.phi.assign(j->phi0.begin(), j->phi0.end());  // double -> float, as the CLI casts
        b.con.assign(j->con0.begin(), j->con0.end());
    }
    s->wq.push_back(std::move(b));
    s->cv_writer.notify_one();
}

// Admit queued jobs into free slots of idle contexts (a busy one takes its
// new jobs once its steps in flight are done); creates contexts within the
// budget.  Called with the lock held; releases it around GPU work.
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
        bool class_busy = false;
        for (auto& b : s->batches) {
            if (!same_class(b->key, j->spec.p)) continue;
            if (b->busy) { class_busy = true; continue; }
            for (int i = 0; i < b->slots && slot < 0; i++)
                if (!b->slot_job[i]) { bt = b.get(); slot = i; }
            if (bt) break;
        }
        double create_s = 0;
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
            // Free idle contexts with no running job (oldest first) until it fits.
            for (size_t bi = 0; bi < s->batches.size() && (double)(s->used + one) > s->budget;) {
                if (s->batches[bi]->running() == 0 && !s->batches[bi]->busy) {
                    pfc_destroy(s->batches[bi]->ctx);
                    s->used -= s->batches[bi]->bytes;
                    s->batches.erase(s->batches.begin() + bi);
                } else {
                    bi++;
                }
            }
            if ((double)(s->used + one) > s->budget || class_busy) {
                // waits until running jobs release memory, or for the busy
                // context of its class to take it when its steps are done
                qi++;
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
            const double t_c = s->now();
            std::string err;
            int rc = pfc_create(&nb->key, slots, &nb->ctx);
            if (rc == 0 && s->o.graph_chunk > 0) rc = pfc_set_graph(nb->ctx, s->o.graph_chunk);
            for (int i = 0; rc == 0 && i < slots; i++) rc = pfc_set_active(nb->ctx, i, 0);
            if (rc) err = pfc_last_error();
            create_s = s->now() - t_c;
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
        j->create_s = create_s;
        if (!j->outdir.empty()) start_bundle(s, j);
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
            if (j->outdir.empty()) {
                finish(s, j, PFC_JOB_FAILED, s->now());
            } else {  // after its start bundle: the writer ends it
                j->write_error = j->error;
                Bundle fb;
                fb.job = j;
                fb.kind = B_ENERGY;
                fb.final = true;
                s->wq.push_back(std::move(fb));
                s->cv_writer.notify_one();
            }
        }
    }
}

// Launch the next steps of an idle batch with running jobs: every job
// advances to its next event (quantum at most).  CVD batches whose jobs
// write files trace the energy of every step (energy.txt, cvd.cpp:442).
// Lock held; the launches only enqueue.
void launch_batch(pfc_server* s, Batch& b, std::string& err) {
    int n = -1;
    bool trace = false;
    for (Job* j : b.slot_job) {
        if (!j) continue;
        const int d = j->cancel ? 0 : next_event(*j) - j->steps_done;
        n = n < 0 ? d : std::min(n, d);
        trace = trace || (b.key.model == PFC_MODEL_CVD && !j->outdir.empty());
    }
    if (n < 0) return;
    n = std::min(n, s->o.quantum);
    b.n = n;
    b.traced = trace && n > 0;
    b.busy = true;
    int rc = 0;
    if (n > 0) rc = b.traced ? pfc_step_energy_enqueue(b.ctx, n) : pfc_step(b.ctx, n);
    if (rc) err = pfc_last_error();
}

// Events of a batch whose launched steps have finished (lock held on entry
// and exit, released around GPU work): energy rows, cancellations,
// checkpoints and completions.  `err` is a failure of the batch's stepping
// (fails all its jobs); other failures fail their own job only.
void batch_events(pfc_server* s, std::unique_lock<std::mutex>& lk, Batch& b, std::string err) {
    const int n = b.n;
    b.busy = false;
    std::vector<double> tr;
    lk.unlock();
    if (err.empty() && pfc_sync(b.ctx)) err = pfc_last_error();
    if (err.empty() && b.traced) {
        tr.resize(2 * (size_t)b.slots * n);
        if (pfc_step_energy_result(b.ctx, n, tr.data())) err = pfc_last_error();
    }
    lk.lock();
    const double t_done = s->now();
    std::vector<int> ev;
    std::vector<Bundle> out;
    for (int i = 0; i < b.slots; i++) {
        Job* j = b.slot_job[i];
        if (!j) continue;
        const int target = next_event(*j);
        const int before = j->steps_done;
        if (!j->cancel) j->steps_done += n;
        if (err.empty() && b.traced && !j->cancel && !j->outdir.empty() && n > 0) {
            Bundle eb;  // energy after each of these steps, labelled by step index
            eb.job = j;
            eb.kind = B_ENERGY;
            eb.kd = before;
            eb.erows.resize(2 * (size_t)n);
            for (int t = 0; t < n; t++) {
                eb.erows[2 * t] = tr[(2 * (size_t)t * b.slots) + 2 * i];
                eb.erows[2 * t + 1] = tr[(2 * (size_t)t * b.slots) + 2 * i + 1];
            }
            out.push_back(std::move(eb));
        }
        if (j->cancel || !err.empty() || j->steps_done == target) ev.push_back(i);
    }
    std::vector<int> release, stepped;  // stepped: slots whose jobs took their last step here
    std::map<int, std::string> failed;  // slot -> its job's failure
    // Cancellations as of this event, read under the lock: the unlocked work
    // below and the release after it act on the same answer (a cancel that
    // arrives meanwhile takes effect at the job's next event).
    std::vector<char> cancelled(b.slots, 0);
    for (int i : ev) cancelled[i] = b.slot_job[i]->cancel;
    std::vector<double> e(2 * (size_t)b.slots, NAN);
    if (!ev.empty()) {
        lk.unlock();
        std::string berr = err;  // failure of the whole batch
        if (berr.empty() && pfc_energy(b.ctx, e.data())) berr = pfc_last_error();
        for (int i : ev) {
            Job* j = b.slot_job[i];
            if (cancelled[i] || !berr.empty()) {
                release.push_back(i);
                if (!cancelled[i]) failed[i] = berr;
                continue;
            }
            const bool final = j->steps_done >= j->spec.steps;
            if (final) stepped.push_back(i);
            std::string jerr;
            try {  // host buffers: a failed allocation fails this job only
                if (j->spec.keep_fields && final) {
                    const size_t P = (size_t)b.key.L * b.key.M * b.key.N;
                    j->phi.resize(P);
                    j->con.resize(P);
                    if (pfc_get_fields(b.ctx, i, j->phi.data(), j->con.data())) jerr = pfc_last_error();
                }
                if (jerr.empty() && !j->outdir.empty() && is_checkpoint(*j, j->steps_done)) {
                    Bundle bd;
                    bd.job = j;
                    bd.kind = B_CKPT;
                    // CVD labels a checkpoint by the step index (kd = t after step t, cvd.cpp:425-426)
                    bd.kd = b.key.model == PFC_MODEL_CVD ? j->steps_done - 1 : j->steps_done;
                    bd.final = final;
                    if (capture(s, b, i, bd, e.data(), jerr)) out.push_back(std::move(bd));
                } else if (jerr.empty() && !j->outdir.empty() && final) {
                    Bundle fb;  // marks the end after the job's last files
                    fb.job = j;
                    fb.kind = B_ENERGY;
                    fb.final = true;
                    out.push_back(std::move(fb));
                }
            } catch (const std::exception& ex) {
                jerr = std::string("checkpoint: ") + ex.what();
            }
            if (!jerr.empty()) {
                failed[i] = jerr;
                release.push_back(i);
                continue;
            }
            if (final) release.push_back(i);
        }
        for (int i : release) pfc_set_active(b.ctx, i, 0);
        lk.lock();
    }
    for (int i : stepped) {  // under the lock: pfc_server_poll reads these
        Job* j = b.slot_job[i];
        j->t_stepped = t_done;
        j->energy = e[2 * i];
        j->density = e[2 * i + 1];
    }
    for (int i : release) {
        Job* j = b.slot_job[i];
        b.slot_job[i] = nullptr;
        auto f = failed.find(i);
        if (cancelled[i]) {
            finish(s, j, PFC_JOB_CANCELLED, s->now());
        } else if (f != failed.end()) {
            if (j->outdir.empty()) {
                j->error = f->second;
                finish(s, j, PFC_JOB_FAILED, s->now());
            } else {  // the writer ends it after the bundles already queued
                j->write_error = f->second;
                Bundle fb;
                fb.job = j;
                fb.kind = B_ENERGY;
                fb.final = true;
                out.push_back(std::move(fb));
            }
        } else if (j->outdir.empty()) {
            finish(s, j, PFC_JOB_DONE, s->now());
        }  // else: DONE once the writer has written the final bundle
    }
    for (auto& bd : out) s->wq.push_back(std::move(bd));
    if (!out.empty()) s->cv_writer.notify_one();
}

// The scheduler: contexts step independently.  Each idle context with
// running jobs is launched to its jobs' next event; the loop then handles
// whichever launched context finishes first (pfc_query) and launches it
// again, so a context with short intervals never waits for a long one.
void sched_loop(pfc_server* s) {
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        admit(s, lk);
        bool any_busy = false;
        for (auto& b : s->batches) {
            if (!b->busy && b->running() > 0) {
                std::string err;
                launch_batch(s, *b, err);
                if (!err.empty()) batch_events(s, lk, *b, err);
            }
            any_busy = any_busy || b->busy;
        }
        if (!any_busy) {
            if (s->stop && s->pending.empty()) return;
            bool running = false;
            for (auto& b : s->batches) running = running || b->running() > 0;
            if (!running) s->cv_sched.wait(lk, [&] { return s->stop || !s->pending.empty(); });
            if (s->stop && s->pending.empty()) {
                bool r = false;
                for (auto& b : s->batches) r = r || b->running() > 0;
                if (!r) return;
            }
            continue;
        }
        // Handle every finished context; if none is, wait briefly (new
        // submissions wake the loop early).
        bool handled = false;
        for (auto& b : s->batches) {
            if (!b->busy) continue;
            int done = 0;
            std::string err;
            if (pfc_query(b->ctx, &done)) { err = pfc_last_error(); done = 1; }
            if (!done) continue;
            batch_events(s, lk, *b, err);
            handled = true;
        }
        if (!handled) s->cv_sched.wait_for(lk, std::chrono::microseconds(200));
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
    std::unique_ptr<pfc_server> owner(new pfc_server());
    pfc_server* s = owner.get();
    s->o = d;
    if (d.hbm_budget > 0) {
        s->budget = d.hbm_budget;
    } else {
        size_t fr = 0, tot = 0;
        if (pfc_device_memory(&fr, &tot)) return PFC_E_HIP;
        s->budget = 0.8 * (double)fr;
    }
    s->sched = std::thread(sched_loop, s);
    try {
        s->writer = std::thread(writer_loop, s);
    } catch (...) {  // no writer thread: stop the scheduler before the server is freed
        {
            std::lock_guard<std::mutex> g(s->mu);
            s->stop = true;
            s->cv_sched.notify_all();
        }
        s->sched.join();
        throw;
    }
    *out = owner.release();
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
    st->t_stepped = j.t_stepped;
    st->create_s = j.create_s;
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
    for (auto& b : s->batches) pfc_destroy(b->ctx);
    delete s;
    return PFC_OK;
}
