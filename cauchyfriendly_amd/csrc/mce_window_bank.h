// Sliding-window bank behind the reference's window-manager C functions
// (pycauchy_initialize_{lti,ltv,nonlin}_window_manager, pycauchy_step, pycauchy_shutdown;
// scripts/swig/cauchy/pycauchy.hpp:80-221).
//
// The reference's SlidingWindowManager (include/cauchy_windows.hpp:729-1705) forks one child
// process per window and feeds it measurements through pipes.  Here every window is an
// estimator of this library in the calling process, on the GPU, with the same schedule:
//   * windows start one measurement apart; a window that has taken num_windows measurements
//     reports and resets (child_window_loop, cauchy_windows.hpp:397-706);
//   * at every measurement after the first, the window that has just reset (the
//     "initializee") restarts about the estimate of the best window, by Speyer's
//     initialisation (speyers_window_init, cauchy_util.hpp:23-113), and takes only the
//     step's last measurement row (SlidingWindowManager::step, cauchy_windows.hpp:1025-1165);
//   * the best window is the one with the most measurements whose fz is positive and whose
//     mean and covariance exist (select_best_stable_mean_cov, cauchy_windows.hpp:1536-1610).
// Each window keeps its own copy of the dynamics container, as each forked child holds one,
// and every window's estimator is built from the same random draws, as every child draws
// from its own copy of the parent's rand() state; the caller's rand() state is left as it
// was.  Callbacks are called for each window in window order, on the calling thread, with
// that window's container.  The histories are kept as the reference keeps them; log files
// are not written.
#pragma once

namespace {

// glibc rand() draws from random()'s state.  initstate / setstate swap the state buffer and
// record the outgoing buffer's position in its first word, so a copy of a buffer taken while
// it is not current is the whole generator state.
struct RandSnapshot {
    alignas(8) char scratch[128];
    alignas(8) char saved[128];
    char* global = nullptr;
    void capture() {
        global = initstate(1, scratch, sizeof(scratch));  // the caller's buffer, position recorded
        memcpy(saved, global, sizeof(saved));
        setstate(global);
    }
    void restore() {
        setstate(scratch);
        memcpy(global, saved, sizeof(saved));
        setstate(global);
    }
};

struct BankWindow {
    Handle* h = nullptr;
    int msmt_count = 0, win_msmt_count = 0, step_count = 0;
    bool first_loop = true;
    std::vector<double> x_bar;  // extended mode: x_bar before the window's last update
    int err = 0;                // numeric error bits of its last update
};

// How the windows of one measurement step: one after another; together as merged launches
// on the calling thread; or one host thread per window, each on its window's own streams
// (extended mode steps merged instead, so that its callbacks stay on the calling thread).
// A bank's few windows are large estimators whose updates are chains of dependent launches:
// on threads the chains of different windows overlap on the device, while merged launches
// on one stream still run each chain's steps one flush after another.
enum { kBankSequential = 0, kBankMerged = 1, kBankThreaded = 2 };

struct BankInit {  // WindowInitializationMessage
    std::vector<double> A0, p0, b0, x;
};

struct WindowBank {
    int W = 0, nsim = 0, n = 0, p = 0, pncc = 0, cmcc = 0;
    bool ext = false;
    void (*f_dyn)(CauchyDynamicsUpdateContainer*) = nullptr;
    void (*f_msmt)(CauchyDynamicsUpdateContainer*, double*) = nullptr;
    void (*f_ext)(CauchyDynamicsUpdateContainer*) = nullptr;
    Handle* mgr = nullptr;  // the manager's own container
    std::vector<BankWindow> win;
    std::vector<double> boost;
    bool has_boost = false;
    int initializer = 0, initializee = 0, msmt_count = 0;
    int mode = kBankThreaded;  // how the windows of a measurement step (mce_window_bank_set_mode)
    std::vector<char> active;
    std::vector<int> active_counts, active_errs, step_counts;
    // per window and per window step, as save_window_data keeps them
    std::vector<std::vector<double>> w_mean, w_var, w_fz, w_cfz, w_cmean, w_cvar, w_xbar;
    std::vector<std::vector<int>> w_err;
    // the best window at each measurement (save_best_window_data)
    std::vector<int> fw_idx, fw_err;
    std::vector<double> fw_mean, fw_var, fw_fz, fw_cfz, fw_cmean, fw_cvar;

    ~WindowBank() {
        for (BankWindow& w : win) pycauchy_single_step_shutdown(w.h);
        if (mgr) pycauchy_single_step_shutdown(mgr);
    }

    // A window (or the manager) with its own container, holding the manager's initial values.
    static Handle* make_handle(int num_steps, const Handle& src, const double* A0, const double* p0, const double* b0,
                               int eng_cmcc, bool with_engine) {
        Handle* h = new Handle;
        h->Phi = src.Phi;
        h->Gamma = src.Gamma;
        h->B = src.B;
        h->H = src.H;
        h->beta = src.beta;
        h->gamma = src.gamma;
        h->x = src.x;
        h->zbar = src.zbar;
        h->duc = src.duc;
        h->duc.x = h->x.data();
        h->duc.Phi = h->Phi.data();
        h->duc.Gamma = h->Gamma.data();
        h->duc.B = h->B.data();
        h->duc.H = h->H.data();
        h->duc.beta = h->beta.data();
        h->duc.gamma = h->gamma.data();
        h->duc.u = nullptr;
        if (with_engine)
            h->eng = new Engine(num_steps, src.duc.n, eng_cmcc, src.duc.pncc, src.duc.p, A0, p0, b0);
        return h;
    }

    void setup(int num_windows, int num_sim_steps, const Handle& tmpl, const double* A0, const double* p0,
               const double* b0, const double* wvb, int size_wvb) {
        if (num_windows < 1) throw std::runtime_error("the bank needs at least one window");
        if (num_sim_steps < 1) throw std::runtime_error("the bank needs at least one simulation step");
        if (tmpl.duc.step != 0 && tmpl.duc.step != 1) throw std::runtime_error("init_step must be 0 or 1");
        W = num_windows;
        nsim = num_sim_steps;
        n = tmpl.duc.n;
        p = tmpl.duc.p;
        pncc = tmpl.duc.pncc;
        cmcc = tmpl.duc.cmcc;
        has_boost = wvb && size_wvb > 0;
        if (has_boost) {
            if (size_wvb != n) throw std::runtime_error("win_var_boost must hold n entries");
            boost.assign(wvb, wvb + n);
        }
        mgr = make_handle(W, tmpl, A0, p0, b0, 0, false);
        // every window draws what a child forked now would draw; the caller's state stays
        RandSnapshot rs;
        rs.capture();
        win.resize(W);
        for (int j = 0; j < W; j++) {
            rs.restore();
            rs.capture();
            win[j].h = make_handle(W, tmpl, A0, p0, b0, ext ? 0 : cmcc, true);
            win[j].h->eng->win_num = j;
            win[j].step_count = tmpl.duc.step;
            win[j].x_bar.assign(n, 0.0);
        }
        rs.restore();
        active.assign(W, 0);
        active_counts.assign(W, 0);
        active_errs.assign(W, 0);
        step_counts.assign(W, 0);
        for (auto* v : {&w_mean, &w_var, &w_fz, &w_cfz, &w_cmean, &w_cvar, &w_xbar}) v->assign(W, {});
        w_err.assign(W, {});
        for (int j = 0; j < W; j++) {
            w_mean[j].assign((size_t)nsim * n, 0.0);
            w_var[j].assign((size_t)nsim * n * n, 0.0);
            w_fz[j].assign(nsim, 0.0);
            w_cfz[j].assign(nsim, 0.0);
            w_cmean[j].assign(nsim, 0.0);
            w_cvar[j].assign(nsim, 0.0);
            w_err[j].assign(nsim, 0);
            if (ext) w_xbar[j].assign((size_t)nsim * n, 0.0);
        }
        fw_idx.assign(nsim, 0);
        fw_err.assign(nsim, 0);
        fw_mean.assign((size_t)nsim * n, 0.0);
        fw_var.assign((size_t)nsim * n * n, 0.0);
        fw_fz.assign(nsim, 0.0);
        fw_cfz.assign(nsim, 0.0);
        fw_cmean.assign(nsim, 0.0);
        fw_cvar.assign(nsim, 0.0);
    }

    static void real_part(const std::vector<cplx>& v, double* out) {
        for (size_t i = 0; i < v.size(); i++) out[i] = v[i].re;
    }

    // Speyer restart of a window that has just reset: its start terms, and the container
    // state the child sets before its first update (child_window_loop, cauchy_windows.hpp:
    // 480-503 and 612-633).
    void apply_init(BankWindow& w, const BankInit& init) {
        Engine* e = w.h->eng;
        CauchyDynamicsUpdateContainer& duc = w.h->duc;
        e->A0 = init.A0;
        e->p0 = init.p0;
        e->b0 = init.b0;
        e->reset_terms();
        duc.step = w.step_count;
        if (ext) memcpy(duc.x, init.x.data(), n * sizeof(double));
        if (!ext && f_dyn) {
            memcpy(duc.x, init.x.data(), n * sizeof(double));
            f_dyn(&duc);
        }
    }

    // The window's updates of one measurement: every row, or only the last after a restart.
    // A task, so that a batch scheduler can step several windows' updates together; run on
    // its own it never suspends.
    Co<void> rows_co(BankWindow& w, std::vector<double> msm, bool restarted, bool keep_xbar) {
        Engine* e = w.h->eng;
        CauchyDynamicsUpdateContainer& duc = w.h->duc;
        for (int i = restarted ? p - 1 : 0; i < p; i++) {
            if (ext) {
                if (i == p - 1 && keep_xbar) memcpy(w.x_bar.data(), duc.x, n * sizeof(double));
                f_msmt(&duc, w.h->zbar.data());
                msm[i] -= w.h->zbar[i];
                f_ext(&duc);
                w.err = co_await e->update_co(msm[i], duc.Phi, duc.Gamma, duc.beta, duc.H + i * n, duc.gamma[i],
                                              nullptr, nullptr);
                co_await e->finalize_extended_co(duc.x);
            } else {
                w.err = co_await e->update_co(msm[i], duc.Phi, duc.Gamma, duc.beta, duc.H + i * n, duc.gamma[i],
                                              duc.B, duc.u);
            }
        }
        if (restarted) e->master_step = p;
    }

    void dynamics(BankWindow& w, bool set_x_from_mean) {
        if (!f_dyn) return;
        CauchyDynamicsUpdateContainer& duc = w.h->duc;
        if (ext) duc.is_xbar_set_for_ece = false;
        if (set_x_from_mean) real_part(w.h->eng->mean, duc.x);
        duc.step = w.step_count;
        f_dyn(&duc);
        if (ext && !duc.is_xbar_set_for_ece)
            throw std::runtime_error("extended mode: the dynamics callback must set is_xbar_set_for_ece");
    }

    // What window j does with one measurement, fixed before its updates run.
    struct Plan {
        bool run = false;         // it updates (and reports an estimate)
        bool restart = false;     // it restarts first and takes only the last row
        bool keep_xbar = false;   // extended mode: it keeps x_bar before its last row
        bool reset_after = false; // it has taken num_windows measurements and resets
    };

    // One measurement of window j (child_window_loop's loop body, cauchy_windows.hpp:448-702),
    // in three parts: begin (counters, the dynamics callback, a restart's initialisation),
    // the updates (rows_co) and end (the reset after a full window).  Only the updates touch
    // the device; every window's begin runs before any update of the measurement, and its
    // end after them.
    Plan window_begin(int j, const double* controls, const BankInit* init) {
        BankWindow& w = win[j];
        w.h->duc.u = cmcc > 0 ? const_cast<double*>(controls) : nullptr;
        Plan pl;
        if (w.first_loop) {
            w.msmt_count += 1;
            if (w.msmt_count > j) {
                w.win_msmt_count += 1;
                if (w.win_msmt_count != 1) dynamics(w, !ext);
                pl.run = true;
                pl.restart = (j != 0) && (w.win_msmt_count == 1);
                pl.keep_xbar = true;
            }
            if (w.msmt_count == W) w.first_loop = false;
        } else {
            w.msmt_count += 1;
            w.win_msmt_count += 1;
            if (w.win_msmt_count != 1) dynamics(w, true);
            pl.run = true;
            pl.restart = w.win_msmt_count == 1;
            pl.keep_xbar = w.win_msmt_count == W;
        }
        pl.reset_after = pl.run && w.win_msmt_count == W;
        if (pl.restart) {
            if (!init) throw std::runtime_error("window bank: window " + std::to_string(j) + " needs a restart");
            apply_init(w, *init);
        }
        return pl;
    }

    void window_end(int j, const Plan& pl) {
        BankWindow& w = win[j];
        if (pl.reset_after) {
            w.win_msmt_count = 0;
            w.h->eng->reset_terms();
        }
        w.step_count++;
    }

    // The updates of windows js (their plans in pl) with measurement msmts, as `mode` says.
    // Merged, the windows step together through the batch scheduler, which makes each kernel
    // of their updates one merged launch.  Each window's results are those of its own updates
    // run alone, in every mode.
    void run_windows(const std::vector<int>& js, const std::vector<Plan>& pl, const double* msmts) {
        std::vector<int> run;
        for (int j : js)
            if (pl[j].run) run.push_back(j);
        auto alone = [&](int j) {
            run_sync(rows_co(win[j], std::vector<double>(msmts, msmts + p), pl[j].restart, pl[j].keep_xbar));
        };
        if (run.size() < 2 || mode == kBankSequential) {
            for (int j : run) alone(j);
            return;
        }
        if (mode == kBankThreaded && !ext) {
            std::vector<std::string> errs(run.size());
            std::atomic<int> next{0};
            const std::function<void()> job = [&]() {
                int k;
                while ((k = next.fetch_add(1)) < (int)run.size()) {
                    try {
                        alone(run[k]);
                    } catch (const std::exception& ex) {
                        errs[k] = ex.what();
                    }
                }
            };
            worker_pool().run((int)run.size() - 1, job);
            for (size_t k = 0; k < run.size(); k++)
                if (!errs[k].empty()) throw std::runtime_error("window " + std::to_string(run[k]) + ": " + errs[k]);
            return;
        }
        std::vector<BatchRunner::Job> jobs(run.size());
        for (size_t k = 0; k < run.size(); k++) {
            BankWindow& w = win[run[k]];
            jobs[k].e = w.h->eng;
            jobs[k].task.emplace(rows_co(w, std::vector<double>(msmts, msmts + p), pl[run[k]].restart,
                                         pl[run[k]].keep_xbar));
        }
        batch_runner().run(jobs);
        for (size_t k = 0; k < run.size(); k++)
            if (!jobs[k].err.empty())
                throw std::runtime_error("window " + std::to_string(run[k]) + ": " + jobs[k].err);
    }

    // One measurement of window j on its own.
    bool window_msmt(int j, const double* msmts, const double* controls, const BankInit* init) {
        std::vector<Plan> pl(W);
        pl[j] = window_begin(j, controls, init);
        run_windows({j}, pl, msmts);
        window_end(j, pl[j]);
        return pl[j].run;
    }

    // save_window_data, cauchy_windows.hpp:1478-1505.
    void save_window(int j) {
        const Engine* e = win[j].h->eng;
        const int s = step_counts[j]++;
        if (s >= nsim) throw std::runtime_error("window bank: more measurements than num_sim_steps");
        w_fz[j][s] = e->fz.re;
        w_cfz[j][s] = e->fz.im;
        double cm = 0, cv = 0;
        for (int i = 0; i < n; i++) {
            w_mean[j][(size_t)s * n + i] = e->mean[i].re;
            cm = std::max(cm, fabs(e->mean[i].im));
        }
        for (int i = 0; i < n * n; i++) {
            w_var[j][(size_t)s * n * n + i] = e->cov[i].re;
            cv = std::max(cv, fabs(e->cov[i].im));
        }
        w_cmean[j][s] = cm;
        w_cvar[j][s] = cv;
        w_err[j][s] = active_errs[j];
        if (ext) memcpy(&w_xbar[j][(size_t)s * n], win[j].x_bar.data(), n * sizeof(double));
    }

    // select_best_stable_mean_cov + stratgey_choose_fullest_window_first,
    // cauchy_windows.hpp:1536-1610: active windows by measurement count, most first.
    int select_best() {
        std::vector<std::pair<int, int>> wc;  // (window, count)
        for (int i = 0; i < W; i++)
            if (active[i]) {
                wc.emplace_back(i, active_counts[i]);
                if (active_counts[i] == W) active_counts[i] = 0;
            }
        std::stable_sort(wc.begin(), wc.end(), [](const auto& a, const auto& b) { return a.second > b.second; });
        for (const auto& [j, c] : wc) {
            (void)c;
            const int err = active_errs[j];
            if (err & (1 << 9)) continue;  // ERROR_FZ_NEGATIVE
            if (err & (1 << 3)) continue;  // ERROR_COVARIANCE_AT_CURRENT_STEP_DNE
            if (err & (1 << 7)) continue;  // ERROR_MEAN_AT_CURRENT_STEP_DNE
            return j;
        }
        throw std::runtime_error("window bank: no window has a usable estimate (the reference stops here)");
    }

    void save_best(int best) {
        fw_idx[msmt_count] = best;
        const int s = step_counts[best] - 1;
        memcpy(&fw_mean[(size_t)msmt_count * n], &w_mean[best][(size_t)s * n], n * sizeof(double));
        memcpy(&fw_var[(size_t)msmt_count * n * n], &w_var[best][(size_t)s * n * n], n * n * sizeof(double));
        fw_cmean[msmt_count] = w_cmean[best][s];
        fw_cvar[msmt_count] = w_cvar[best][s];
        fw_fz[msmt_count] = w_fz[best][s];
        fw_cfz[msmt_count] = w_cfz[best][s];
        fw_err[msmt_count] = w_err[best][s];
    }

    // SlidingWindowManager::step, cauchy_windows.hpp:1025-1165.
    void step(const double* msmts, const double* controls) {
        if (msmt_count >= nsim) throw std::runtime_error("window bank: more measurements than num_sim_steps");
        std::vector<double> M(msmts, msmts + p);
        // controls (cmcc of them; zeros when none are given), as the children's u
        std::vector<double> C(std::max(cmcc, 1), 0.0);
        if (controls)
            for (int i = 0; i < cmcc; i++) C[i] = controls[i];
        if (msmt_count > 0) {
            if (msmt_count < W) {
                initializee += 1;
            } else {
                std::swap(initializer, initializee);
                initializer += 2;
                if (initializer >= W) initializer -= W;
            }
        }
        std::fill(active.begin(), active.end(), 0);
        // every window but the one that restarts, together
        std::vector<int> js;
        std::vector<Plan> pl(W);
        for (int j = 0; j < W; j++)
            if (!(j == initializee && msmt_count > 0)) js.push_back(j);
        for (int j : js) pl[j] = window_begin(j, C.data(), nullptr);
        run_windows(js, pl, msmts);
        for (int j : js) {
            window_end(j, pl[j]);
            if (j <= msmt_count) {
                active[j] = 1;
                active_counts[j] += 1;
                active_errs[j] = win[j].err;
                save_window(j);
            }
        }
        const int best = select_best();
        if (msmt_count > 0) {
            const int s = step_counts[best] - 1;
            std::vector<double> xr(&w_mean[best][(size_t)s * n], &w_mean[best][(size_t)s * n] + n);
            std::vector<double> Pr(&w_var[best][(size_t)s * n * n], &w_var[best][(size_t)s * n * n] + n * n);
            CauchyDynamicsUpdateContainer& duc = mgr->duc;
            BankInit init;
            init.A0.assign(n * n, 0.0);
            init.p0.assign(n, 0.0);
            init.b0.assign(n, 0.0);
            if (has_boost)
                for (int i = 0; i < n; i++) Pr[i * n + i] += boost[i];
            if (!ext) {
                if (f_dyn) {
                    double* keep = duc.x;
                    duc.x = xr.data();
                    f_dyn(&duc);
                    duc.x = keep;
                }
                host::speyers_window_init(n, xr.data(), Pr.data(), duc.H + (p - 1) * n, duc.gamma[p - 1], M[p - 1],
                                          init.A0.data(), init.p0.data(), init.b0.data(), kCovEigenvalueTol);
                init.x = xr;
            } else {
                std::vector<double> xb(&w_xbar[best][(size_t)s * n], &w_xbar[best][(size_t)s * n] + n);
                double* keep = duc.x;
                duc.x = xb.data();
                for (int i = 0; i < n; i++) xr[i] -= xb[i];
                f_ext(&duc);
                std::vector<double> zb(std::max(p, 1), 0.0);
                f_msmt(&duc, zb.data());
                M[p - 1] -= zb[p - 1];
                host::speyers_window_init(n, xr.data(), Pr.data(), duc.H + (p - 1) * n, duc.gamma[p - 1], M[p - 1],
                                          init.A0.data(), init.p0.data(), init.b0.data(), kCovEigenvalueTol);
                init.x = xb;
                duc.x = keep;
            }
            window_msmt(initializee, msmts, C.data(), &init);
            active_errs[initializee] = win[initializee].err;
            active_counts[initializee] += 1;
            save_window(initializee);
        }
        save_best(best);
        msmt_count += 1;
        mgr->duc.step++;
    }
};

WindowBank* g_bank = nullptr;

}  // namespace

extern "C" {

static void* bank_init(int num_windows, int num_sim_steps, Handle* tmpl, bool ext,
                       void (*f_dyn)(CauchyDynamicsUpdateContainer*),
                       void (*f_msmt)(CauchyDynamicsUpdateContainer*, double*),
                       void (*f_ext)(CauchyDynamicsUpdateContainer*), double* A0, double* p0, double* b0,
                       double* wvb, int size_wvb, const char* who) {
    try {
        if (g_bank) throw std::runtime_error("a window bank is already running; call pycauchy_shutdown first");
        auto* b = new WindowBank;
        b->ext = ext;
        b->f_dyn = f_dyn;
        b->f_msmt = f_msmt;
        b->f_ext = f_ext;
        try {
            b->setup(num_windows, num_sim_steps, *tmpl, A0, p0, b0, wvb, size_wvb);
        } catch (...) {
            delete b;
            throw;
        }
        g_bank = b;
        return b;
    } catch (const std::exception& ex) {
        set_err(who, ex);
        return nullptr;
    }
}

// The manager's container template: pycauchy_initialize_lti's, without an engine.
static Handle* bank_template(double* Phi, int size_Phi, double* Gamma, int size_Gamma, double* B, int size_B,
                             double* beta, int size_beta, double* H, int size_H, double* gamma, int size_gamma,
                             int n, double dt, int init_step) {
    const int pncc = size_beta, cmcc = n > 0 ? size_B / n : 0, p = size_gamma;
    Handle* h = new Handle;
    h->Phi.assign(std::max(n * n, 1), 0.0);
    h->Gamma.assign(std::max(n * pncc, 1), 0.0);
    h->B.assign(std::max(n * cmcc, 1), 0.0);
    h->H.assign(std::max(p * n, 1), 0.0);
    h->beta.assign(std::max(pncc, 1), 0.0);
    h->gamma.assign(std::max(p, 1), 0.0);
    if (Phi) std::copy(Phi, Phi + std::min(size_Phi, n * n), h->Phi.begin());
    if (Gamma) std::copy(Gamma, Gamma + std::min(size_Gamma, n * pncc), h->Gamma.begin());
    if (B) std::copy(B, B + std::min(size_B, n * cmcc), h->B.begin());
    if (H) std::copy(H, H + std::min(size_H, p * n), h->H.begin());
    if (beta) std::copy(beta, beta + size_beta, h->beta.begin());
    if (gamma) std::copy(gamma, gamma + size_gamma, h->gamma.begin());
    h->x.assign(n, 0.0);
    h->zbar.assign(std::max(p, 1), 0.0);
    CauchyDynamicsUpdateContainer& duc = h->duc;
    memset(&duc, 0, sizeof(duc));
    duc.dt = dt;
    duc.step = init_step;
    duc.n = n;
    duc.cmcc = cmcc;
    duc.pncc = pncc;
    duc.p = p;
    return h;
}

void pycauchy_initialize_lti_window_manager(int num_windows, int num_sim_steps, double* A0, int size_A0, double* p0,
                                            int size_p0, double* b0, int size_b0, double* Phi, int size_Phi,
                                            double* Gamma, int size_Gamma, double* B, int size_B, double* beta,
                                            int size_beta, double* H, int size_H, double* gamma, int size_gamma,
                                            bool debug_print, bool log_seq, bool log_full, char* log_dir, double dt,
                                            int init_step, double* win_var_boost, int size_wvb) {
    (void)size_A0, (void)size_p0, (void)debug_print, (void)log_seq, (void)log_full, (void)log_dir;
    Handle* t = bank_template(Phi, size_Phi, Gamma, size_Gamma, B, size_B, beta, size_beta, H, size_H, gamma,
                              size_gamma, size_b0, dt, init_step);
    bank_init(num_windows, num_sim_steps, t, false, nullptr, nullptr, nullptr, A0, p0, b0, win_var_boost, size_wvb,
              "pycauchy_initialize_lti_window_manager");
    delete t;
}

void pycauchy_initialize_ltv_window_manager(int num_windows, int num_sim_steps, double* A0, int size_A0, double* p0,
                                            int size_p0, double* b0, int size_b0, double* Phi, int size_Phi,
                                            double* Gamma, int size_Gamma, double* B, int size_B, double* beta,
                                            int size_beta, double* H, int size_H, double* gamma, int size_gamma,
                                            void (*f_dyn_update_callback)(CauchyDynamicsUpdateContainer*),
                                            bool debug_print, bool log_seq, bool log_full, char* log_dir, double dt,
                                            int init_step, double* win_var_boost, int size_wvb) {
    (void)size_A0, (void)size_p0, (void)debug_print, (void)log_seq, (void)log_full, (void)log_dir;
    Handle* t = bank_template(Phi, size_Phi, Gamma, size_Gamma, B, size_B, beta, size_beta, H, size_H, gamma,
                              size_gamma, size_b0, dt, init_step);
    bank_init(num_windows, num_sim_steps, t, false, f_dyn_update_callback, nullptr, nullptr, A0, p0, b0,
              win_var_boost, size_wvb, "pycauchy_initialize_ltv_window_manager");
    delete t;
}

void pycauchy_initialize_nonlin_window_manager(int num_windows, int num_sim_steps, double* x0, int size_x0,
                                               double* A0, int size_A0, double* p0, int size_p0, double* b0,
                                               int size_b0, double* beta, int size_beta, double* gamma,
                                               int size_gamma,
                                               void (*f_dyn_update_callback)(CauchyDynamicsUpdateContainer*),
                                               void (*f_nonlinear_msmt_model)(CauchyDynamicsUpdateContainer*,
                                                                              double*),
                                               void (*f_extended_msmt_update_callback)(CauchyDynamicsUpdateContainer*),
                                               int cmcc, bool debug_print, bool log_seq, bool log_full, char* log_dir,
                                               double dt, int init_step, double* win_var_boost, int size_wvb) {
    (void)size_A0, (void)size_p0, (void)debug_print, (void)log_seq, (void)log_full, (void)log_dir;
    const int n = size_b0;
    if (size_x0 != n || !f_dyn_update_callback || !f_nonlinear_msmt_model || !f_extended_msmt_update_callback) {
        set_err("pycauchy_initialize_nonlin_window_manager",
                std::runtime_error("x0 must hold n entries and all three callbacks are required"));
        return;
    }
    std::vector<double> B(std::max(n * cmcc, 1), 0.0);
    Handle* t = bank_template(nullptr, 0, nullptr, 0, B.data(), n * cmcc, beta, size_beta, nullptr, 0, gamma,
                              size_gamma, n, dt, init_step);
    std::copy(x0, x0 + n, t->x.begin());
    bank_init(num_windows, num_sim_steps, t, true, f_dyn_update_callback, f_nonlinear_msmt_model,
              f_extended_msmt_update_callback, A0, p0, b0, win_var_boost, size_wvb,
              "pycauchy_initialize_nonlin_window_manager");
    delete t;
}

void pycauchy_step(double* msmts, int size_msmts, double* controls, int size_controls, double* out_swm_fz,
                   double** out_xhat, int* size_out_xhat, double** out_Phat, int* size_out_Phat,
                   double* out_swm_cerr_fz, double* out_swm_cerr_xhat, double* out_swm_cerr_Phat,
                   int* out_swm_win_idx, int* out_swm_err_code) {
    (void)size_controls;
    WindowBank* b = g_bank;
    *out_xhat = nullptr;
    *out_Phat = nullptr;
    *size_out_xhat = *size_out_Phat = 0;
    *out_swm_win_idx = -1;
    *out_swm_err_code = -1;
    try {
        if (!b) throw std::runtime_error("no window bank: initialize one first");
        if (size_msmts != b->p) throw std::runtime_error("expected p measurements per step");
        b->step(msmts, controls);
    } catch (const std::exception& ex) {
        set_err("pycauchy_step", ex);
        return;
    }
    const int k = b->msmt_count - 1, n = b->n;
    *size_out_xhat = n;
    *size_out_Phat = n * n;
    *out_xhat = mallocd<double>(n);
    *out_Phat = mallocd<double>(n * n);
    memcpy(*out_xhat, &b->fw_mean[(size_t)k * n], n * sizeof(double));
    memcpy(*out_Phat, &b->fw_var[(size_t)k * n * n], n * n * sizeof(double));
    *out_swm_fz = b->fw_fz[k];
    *out_swm_cerr_fz = b->fw_cfz[k];
    *out_swm_cerr_xhat = b->fw_cmean[k];
    *out_swm_cerr_Phat = b->fw_cvar[k];
    *out_swm_win_idx = b->fw_idx[k];
    *out_swm_err_code = b->fw_err[k];
}

void pycauchy_shutdown() {
    delete g_bank;
    g_bank = nullptr;
}

int mce_window_bank_set_mode(int mode) {
    if (!g_bank) {
        set_err("mce_window_bank_set_mode", std::runtime_error("no window bank: initialize one first"));
        return -1;
    }
    if (mode < kBankSequential || mode > kBankThreaded) {
        set_err("mce_window_bank_set_mode", std::runtime_error("mode must be 0, 1 or 2"));
        return -1;
    }
    g_bank->mode = mode;
    return 0;
}

}  // extern "C"
