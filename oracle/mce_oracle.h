/*
 * CPU restatement of the Multivariate Cauchy Estimator term-propagation step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: it may be loaded by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg, never by the product.
 *
 * The API mirrors oracle/ref_harness.cpp so that the same Python driver can run
 * either the restatement or the reference build on identical inputs.
 */
#ifndef MCE_ORACLE_H
#define MCE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

void* oracle_create(int num_steps, int d, int cmcc, int pncc, int p,
                    const double* A0, const double* p0, const double* b0,
                    int use_seed, unsigned int seed);
void oracle_destroy(void* h);
int oracle_shape_range(void* h);
int oracle_master_step(void* h);
int oracle_num_terms(void* h);
void oracle_root_point(void* h, double* out);
void oracle_b_pert(void* h, int tid, double* out);
/* Same output layout as ref_step (see oracle/ref_harness.cpp). */
int oracle_step(void* h, double msmt, const double* Phi, const double* Gamma, const double* beta,
                const double* H, double gamma, const double* B, const double* u,
                double* out_mom, int* out_cnt);
void oracle_public_moments(void* h, double* fz, double* xhat, double* Phat);
/* Extended-mode finish of the last update; x_bar (d) in/out. */
void oracle_finalize_extended_moments(void* h, double* x_bar);
int oracle_term(void* h, int shape, int idx, int* info, double* A, double* p, double* b,
                unsigned int* keys, double* gvals);
int oracle_set_tr_search_idxs_ordering(const int* ordering, int n);
/* Overrides the construction-time random draws (root point, 8 x max_shape b_pert). */
void oracle_set_random_draws(void* h, const double* root_point, const double* b_pert_8xmax);
/* Counters of rare flattening paths seen so far: [negligible roots, old-term members, cell-count
   mismatches, negligible roots with members, replacement roots found, child B-tables holding a key
   with its top bit set, total members, largest member list]. */
void oracle_path_counters(void* h, long long* out3);

#ifdef __cplusplus
}
#endif
#endif
