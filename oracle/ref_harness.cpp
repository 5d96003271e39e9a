// Test-infrastructure harness around the *reference* CauchyFriendly estimator.
//
// This file is NOT part of the product.  It is compiled (see oracle/Makefile)
// against the reference headers where they lie, into oracle/_ref/, and is used
// only by tests/golden/make_golden.py (to write golden fixtures) and by the
// cpu_baseline leg of bench.py (to time the reference CPU path on host cores).
//
// It drives `CauchyEstimator` (include/cauchy_estimator.hpp:47-1432) through the
// same call sequence as `CauchyEstimator::step` (cauchy_estimator.hpp:1211-1245)
// so that the moments *before* fast term reduction (the values that
// `step()` overwrites with fz = 1 in non-debug mode, cauchy_estimator.hpp:1172-1176)
// and the per-shape term counts after MU/MUC and after FTR can be recorded.
#include "cauchy_estimator.hpp"

#include <vector>

namespace {

struct RefHandle {
    CauchyEstimator* est;
    int d;
};

void put_c(double* out, C_COMPLEX_TYPE v) {
    out[0] = creal(v);
    out[1] = cimag(v);
}

}  // namespace

extern "C" {

// Constructs the reference estimator.  When use_seed != 0, glibc srand(seed)
// is called first so that root_point and the 8 b_pert vectors drawn in the
// constructor (cauchy_estimator.hpp:127-128, 150-161) are reproducible.
void* ref_create(int num_steps, int d, int cmcc, int pncc, int p,
                 const double* A0, const double* p0, const double* b0,
                 int use_seed, unsigned int seed) {
    if (use_seed) srand(seed);
    RefHandle* h = new RefHandle;
    h->d = d;
    h->est = new CauchyEstimator(const_cast<double*>(A0), const_cast<double*>(p0),
                                 const_cast<double*>(b0), num_steps, d, cmcc, pncc, p, false);
    return h;
}

void ref_destroy(void* hv) {
    RefHandle* h = (RefHandle*)hv;
    delete h->est;
    delete h;
}

int ref_shape_range(void* hv) { return ((RefHandle*)hv)->est->shape_range; }
int ref_master_step(void* hv) { return ((RefHandle*)hv)->est->master_step; }
int ref_num_terms(void* hv) { return ((RefHandle*)hv)->est->Nt; }

void ref_root_point(void* hv, double* out) {
    RefHandle* h = (RefHandle*)hv;
    memcpy(out, h->est->root_point, h->d * sizeof(double));
}

// b_pert of the DiffCellEnumHelper used by thread `tid` (cell_enumeration.hpp:467-470).
void ref_b_pert(void* hv, int tid, double* out) {
    RefHandle* h = (RefHandle*)hv;
    DiffCellEnumHelper* dce = h->est->dce_helper + tid;
    memcpy(out, dce->b_pert, dce->max_shape * sizeof(double));
}

// One scalar measurement update.
//   out_mom (len 6 + 2d + 2d^2): [fz_public(re,im), fz_pre_ftr(re,im), last_fz(re,im),
//                                 mean(re,im)*d, cov(re,im)*d*d]   (moments after MU, pre-FTR)
//   out_cnt (len 4 + 2*shape_range): [err_code, Nt_after_mu, Nt_after_ftr, threads_tp_to_muc,
//                                     terms_per_shape after MU/MUC..., terms_per_shape after FTR...]
// Returns numeric_moment_errors (or -1 when the estimator refused the step).
int ref_step(void* hv, double msmt, const double* Phi, const double* Gamma, const double* beta,
             const double* H, double gamma, const double* B, const double* u,
             double* out_mom, int* out_cnt) {
    RefHandle* h = (RefHandle*)hv;
    CauchyEstimator* e = h->est;
    const int d = h->d;
    const int sr = e->shape_range;
    e->set_function_pointers();
    if (e->numeric_moment_errors & (1 << ERROR_FZ_NEGATIVE)) return -1;
    if (e->master_step == e->num_estimation_steps) return -1;
    double* Phi_ = const_cast<double*>(Phi);
    double* Gam_ = const_cast<double*>(Gamma);
    double* beta_ = const_cast<double*>(beta);
    double* H_ = const_cast<double*>(H);
    double* B_ = const_cast<double*>(B);
    double* u_ = const_cast<double*>(u);

    e->skip_post_mu = SKIP_LAST_STEP && (e->master_step == (e->num_estimation_steps - 1));
    int threads_used = 1;
    if (e->master_step == 0) {
        // Body of CauchyEstimator::step_first (cauchy_estimator.hpp:1179-1208), with
        // the pre-FTR moments recorded between compute_moments and the g-table build.
        CauchyTerm* terms = e->terms_dp[d];
        e->Nt = terms[0].msmt_update(terms + 1, msmt, H_, gamma, true, false, &e->childterms_workspace) + 1;
        e->terms_per_shape[d] = e->Nt;
        e->compute_moments(true);
        put_c(out_mom + 2, e->fz);
        for (int i = 0; i < d; i++) put_c(out_mom + 6 + 2 * i, e->conditional_mean[i]);
        for (int i = 0; i < d * d; i++) put_c(out_mom + 6 + 2 * d + 2 * i, e->conditional_variance[i]);
        out_cnt[1] = e->Nt;
        for (int m = 0; m < sr; m++) out_cnt[4 + m] = e->terms_per_shape[m];
        for (int i = 0; i < e->Nt; i++) {
            e->reduce_store->set_term_ptrs(terms + i);
            terms[i].cells_gtable = e->dce_helper->cell_counts_cen[d] / (1 + HALF_STORAGE);
            e->gb_tables->set_term_btable_pointer(&(terms[i].enc_B), terms[i].cells_gtable, true);
            e->gb_tables->set_term_gtable_pointer(&(terms[i].gtable), terms[i].cells_gtable, true);
            make_gtable_first(terms + i, e->G_SCALE_FACTOR);
            terms[i].become_parent();
        }
        e->gb_tables->swap_gtables();
        e->fz = MAKE_CMPLX(1, 0);
        memcpy(e->last_conditional_mean, e->conditional_mean, d * sizeof(C_COMPLEX_TYPE));
        memcpy(e->last_conditional_variance, e->conditional_variance, d * d * sizeof(C_COMPLEX_TYPE));
        e->last_fz = e->fz;
    } else {
        if ((NUM_CPUS == 1) || (e->Nt < MIN_TERMS_PER_THREAD_TP_TO_MUC)) {
            e->step_tp_to_muc(msmt, Phi_, Gam_, beta_, H_, gamma, B_, u_);
        } else {
            e->threaded_step_tp_to_muc(msmt, Phi_, Gam_, beta_, H_, gamma, B_, u_);
            threads_used = e->num_threads_tp_to_muc;
        }
        put_c(out_mom + 2, e->fz);
        for (int i = 0; i < d; i++) put_c(out_mom + 6 + 2 * i, e->conditional_mean[i]);
        for (int i = 0; i < d * d; i++) put_c(out_mom + 6 + 2 * d + 2 * i, e->conditional_variance[i]);
        out_cnt[1] = e->Nt;
        for (int m = 0; m < sr; m++) out_cnt[4 + m] = e->terms_per_shape[m];
        e->fast_term_reduction_and_create_gtables();
    }
    e->master_step++;
    put_c(out_mom + 0, e->fz);
    put_c(out_mom + 4, e->last_fz);
    out_cnt[0] = e->numeric_moment_errors;
    out_cnt[2] = e->Nt;
    out_cnt[3] = threads_used;
    for (int m = 0; m < sr; m++) out_cnt[4 + sr + m] = e->terms_per_shape[m];
    return e->numeric_moment_errors;
}

// Public moments exactly as pycauchy_single_step_ltiv reports them
// (scripts/swig/cauchy/pycauchy.hpp:535-541): fz, xhat, Phat are real parts.
void ref_public_moments(void* hv, double* fz, double* xhat, double* Phat) {
    RefHandle* h = (RefHandle*)hv;
    CauchyEstimator* e = h->est;
    const int d = h->d;
    fz[0] = creal(e->fz);
    fz[1] = cimag(e->fz);
    for (int i = 0; i < d; i++) xhat[i] = creal(e->conditional_mean[i]);
    for (int i = 0; i < d * d; i++) Phat[i] = creal(e->conditional_variance[i]);
}

// Extended-mode finish of an update (CauchyEstimator::finalize_extended_moments,
// include/cauchy_estimator.hpp:1358-1394), x_bar in/out.
void ref_finalize_extended_moments(void* hv, double* x_bar) {
    RefHandle* h = (RefHandle*)hv;
    h->est->finalize_extended_moments(x_bar);
}

// Term introspection after a step that ran FTR: every term is a parent whose
// g-table (sorted by key) is gtable_p (cauchy_term.hpp:748-755).
//   info: [m, cells_gtable_p, phc]
int ref_term(void* hv, int shape, int idx, int* info, double* A, double* p, double* b,
             unsigned int* keys, double* gvals) {
    RefHandle* h = (RefHandle*)hv;
    CauchyEstimator* e = h->est;
    const int d = h->d;
    if (shape < 0 || shape >= e->shape_range) return -1;
    if (idx < 0 || idx >= e->terms_per_shape[shape]) return -1;
    CauchyTerm* t = e->terms_dp[shape] + idx;
    info[0] = t->m;
    info[1] = t->cells_gtable_p;
    info[2] = t->phc;
    if (A) memcpy(A, t->A, t->m * d * sizeof(double));
    if (p) memcpy(p, t->p, t->m * sizeof(double));
    if (b) memcpy(b, t->b, d * sizeof(double));
    if (keys && gvals) {
        for (int i = 0; i < t->cells_gtable_p; i++) {
            keys[i] = t->gtable_p[i].key;
            gvals[2 * i] = creal(t->gtable_p[i].value);
            gvals[2 * i + 1] = cimag(t->gtable_p[i].value);
        }
    }
    return 0;
}

int ref_set_tr_search_idxs_ordering(const int* ordering, int n) {
    if (n > 12) return 1;
    for (int i = 0; i < n; i++) TR_SEARCH_IDXS_ORDERING[i] = ordering[i];
    return 0;
}

}  // extern "C"
