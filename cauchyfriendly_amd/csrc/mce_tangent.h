// Forward-mode (tangent-linear) derivatives of the measurement update.
//
// For P directions at once, every term carries the derivatives of its continuous fields
// (A, p, q, b, c, d) and every g-table cell the derivative of its value, with respect to
// the parameters the caller seeds per update (Phi, H, the process scales beta and the
// measurement scale gamma).  Every discrete decision (signs, coalignment, cells, term
// reduction, negligibility, root replacement) is the nominal forward's, taken from the
// same kernels or recomputed with the same arithmetic, so the derivative is that of the
// estimator with its term structure held fixed.  The derivative of a step's measurement
// likelihood fz comes out of the moment sum; nothing is stored between steps except the
// current terms' tangents (the forward is recomputed, step by step, with its tangents).
//
// Forward being differentiated: time_prop / normalize_hps / tp_coalign
// (cauchy_term.hpp:435-531), msmt_update and eval_g_yei (cauchy_term.hpp:84-401),
// mu_coalign (cauchy_term.hpp:533-745), make_gtable_first / make_gtable / add_gtables /
// make_gtables (flattening.hpp:14-568), cache_moments (cauchy_estimator.hpp:307-338).
#pragma once
#include "mce_kernels.h"

namespace mce {

constexpr int kMaxTan = 16;  // directions per estimator

// Per-update seeds: direction k moves Phi by Phi[k], the measurement row by H[k], the
// pre-coaligned process scales by beta[k] and the measurement scale by gamma[k].
struct StepTan {
    double Phi[kMaxTan][kMaxD * kMaxD];
    double H[kMaxTan][kMaxD];
    double beta[kMaxTan][kMaxM];
    double gamma[kMaxTan];
};

// Tangents of a term pool, [term][direction][field] (one thread owns a term's tangents).
struct TanPool {
    double *A, *p, *q, *b, *c, *dv;
    int P, MM, d;
    __device__ __host__ double* tA(int64_t i, int k) const { return A + ((i * P + k) * MM) * d; }
    __device__ __host__ double* tp(int64_t i, int k) const { return p + (i * P + k) * MM; }
    __device__ __host__ double* tq(int64_t i, int k) const { return q + (i * P + k) * MM; }
    __device__ __host__ double* tb(int64_t i, int k) const { return b + (i * P + k) * d; }
    __device__ __host__ double* tc(int64_t i, int k) const { return c + i * P + k; }
    __device__ __host__ double* tdv(int64_t i, int k) const { return dv + i * P + k; }
};

__device__ inline double sgn0(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : 0.0); }

// d(a / D) = (da - (a / D) dD) / D for complex a, D.
__device__ inline cplx dquot(cplx a, cplx da, cplx D, cplx dD) {
    const cplx w = cdiv(a, D);
    return cdiv(csub(da, cmul(w, dD)), D);
}

// Index of the parent g-table entry of a half-storage key (eval_gs.hpp:94-153): -1 when
// absent; *cj when the key was reversed (the value enters conjugated).
__device__ inline int lookup_idx(int enc, int top_bit, int rev_mask, const uint32_t* keys, int size, bool* cj) {
    *cj = (enc & top_bit) != 0;
    return bsearch_u32(keys, (uint32_t)(*cj ? (rev_mask ^ enc) : enc), size);
}

// ------------------------------------------------------------------ time propagation
// Tangents through time_prop / normalize_hps / tp_coalign, launched before k_tp on the
// same parents: the nominal rows are formed with k_tp's arithmetic, so the coalignment
// decisions are k_tp's.  Row r: u = A_r Phi^T, n1 = |u|_1, A'_r = u / n1, p'_r = p_r n1,
// then p'_k += beta_j for a coaligned Gamma row j, else the row (Gamma_j, beta_j) is
// appended; b' = Phi b (+ B u, which carries no tangent).
template <int D>
__global__ void k_tp_tan(TermPool P, TanPool T, int n, const StepConst* __restrict__ sc,
                         const StepTan* __restrict__ st) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !sc->with_tp) return;
    const int MM = P.MM, NP = T.P;
    const int m = P.m[i];
    const double* A = P.A + (int64_t)i * MM * D;
    const double* p = P.p + (int64_t)i * MM;
    const double* b = P.b + (int64_t)i * D;
    double An[kMaxM * D], n1s[kMaxM];
    for (int r = 0; r < m; r++) {
        double t[D];
        for (int k = 0; k < D; k++) t[k] = A[r * D + k];
        for (int j = 0; j < D; j++) {
            double s = 0.0;
            for (int k = 0; k < D; k++) s += t[k] * sc->Phi[k + j * D];
            An[r * D + j] = s;
        }
        double n1 = 0;
        for (int k = 0; k < D; k++) n1 += fabs(An[r * D + k]);
        n1s[r] = n1;
        for (int k = 0; k < D; k++) An[r * D + k] /= n1;
    }
    int crow[kMaxM];
    uint32_t F = (sc->ng >= 32) ? 0xffffffffu : ((1u << sc->ng) - 1u);
    for (int j = 0; j < sc->ng; j++) {
        crow[j] = -1;
        for (int k = 0; k < m; k++) {
            if (!((F >> j) & 1u)) continue;
            bool pos, neg;
            coalign_gates(&sc->G[j * D], &An[k * D], D, pos, neg);
            if (pos || neg) {
                F &= ~(1u << j);
                crow[j] = k;
            }
        }
    }
    for (int kk = 0; kk < NP; kk++) {
        const double* dPhi = st->Phi[kk];
        double* dA = T.tA(i, kk);
        double* dp = T.tp(i, kk);
        double* db = T.tb(i, kk);
        for (int r = 0; r < m; r++) {
            double du[D];
            double dn1 = 0;
            for (int j = 0; j < D; j++) {
                double s = 0.0;
                for (int k = 0; k < D; k++) s += dA[r * D + k] * sc->Phi[k + j * D] + A[r * D + k] * dPhi[k + j * D];
                du[j] = s;
                dn1 += sgn0(An[r * D + j]) * s;
            }
            for (int j = 0; j < D; j++) dA[r * D + j] = (du[j] - An[r * D + j] * dn1) / n1s[r];
            dp[r] = dp[r] * n1s[r] + p[r] * dn1;
        }
        double t[D];
        for (int k = 0; k < D; k++) t[k] = db[k];
        for (int r = 0; r < D; r++) {
            double s = 0.0;
            for (int j = 0; j < D; j++) s += dPhi[r * D + j] * b[j] + sc->Phi[r * D + j] * t[j];
            db[r] = s;
        }
        int mm = m;
        for (int j = 0; j < sc->ng; j++) {
            if (crow[j] >= 0) {
                dp[crow[j]] += st->beta[kk][j];
            } else if (mm < MM) {
                for (int k = 0; k < D; k++) dA[mm * D + k] = 0.0;
                dp[mm] = st->beta[kk][j];
                mm++;
            }
        }
    }
}

// ------------------------------------------------------------------ measurement update
// Row r of child t of a parent (parent row r below t, r + 1 from t on, parent row m
// standing for the measurement-noise row), nominal as k_mu forms it and its tangent.
template <int D>
__device__ inline void child_row_tan(const double* A, const double* p, const double* dA, const double* dp,
                                     const double* H, const double* dH, const double* mut, const double* dmut,
                                     double gamma, double dgamma, int m, int t, int r, double* row, double* drow,
                                     double* rho, double* drho, bool* horth) {
    const int src = (t < m && r >= t) ? r + 1 : r;
    *horth = false;
    if (src == m) {
        for (int k = 0; k < D; k++) {
            row[k] = 0.0 - mut[k];
            drow[k] = -dmut[k];
        }
        *rho = gamma;
        *drho = dgamma;
        return;
    }
    const double* Ar = A + src * D;
    const double* dAr = dA + src * D;
    const double h = dev_dot<D>(H, Ar);
    if (fabs(h) < kMuEps) {
        for (int k = 0; k < D; k++) {
            row[k] = Ar[k];
            drow[k] = dAr[k];
        }
        *horth = true;
        *rho = p[src];
        *drho = dp[src];
        return;
    }
    double dh = 0;
    for (int k = 0; k < D; k++) dh += dH[k] * Ar[k] + H[k] * dAr[k];
    const double rr = 1.0 / h;
    for (int k = 0; k < D; k++) {
        const double mu = Ar[k] * rr;
        row[k] = mu - mut[k];
        drow[k] = (dAr[k] - mu * dh) * rr - dmut[k];
    }
    *rho = p[src] * fabs(h);
    *drho = dp[src] * fabs(h) + p[src] * sgn0(h) * dh;
}

// mu_coalign's merges (cauchy_term.hpp:533-745) applied to one direction's tangents of
// the L1-normalised rows: the decisions come from the nominal normalised rows An and
// the H-orthogonality mask, exactly as dev_mu_coalign takes them.
template <int D>
__device__ inline void mu_coalign_tan(const double* An, uint32_t hf, int m, double* dA, double* dp, double* dq) {
    uint32_t hu = hf;
    uint32_t F = (m >= 32) ? 0xffffffffu : ((1u << m) - 1u);
    for (int j = 0; j < m - 1; j++) {
        if (!((F >> j) & 1u)) continue;
        for (int k = j + 1; k < m; k++) {
            if (!((F >> k) & 1u)) continue;
            bool pos, neg;
            coalign_gates(An + j * D, An + k * D, D, pos, neg);
            if (pos) {
                if (hf) {
                    const bool hj = (hu >> j) & 1u, hk = (hu >> k) & 1u;
                    if (!hj && !hk) dq[j] += dq[k];
                    else if (hj && !hk) {
                        dq[j] = dq[k];
                        hu &= ~(1u << j);
                    } else if (!hj && hk) {
                        hu &= ~(1u << k);
                    } else {
                        hu &= ~(1u << k);
                        dq[j] += dq[k];
                    }
                } else {
                    dq[j] += dq[k];
                }
                F &= ~(1u << k);
                dp[j] += dp[k];
            }
            if (neg) {
                if (hf) {
                    const bool hj = (hu >> j) & 1u, hk = (hu >> k) & 1u;
                    if (!hj && !hk) dq[j] -= dq[k];
                    else if (hj && !hk) {
                        dq[j] = -dq[k];
                        hu &= ~(1u << j);
                    } else if (!hj && hk) {
                        hu &= ~(1u << k);
                    }
                } else {
                    dq[j] -= dq[k];
                }
                F &= ~(1u << k);
                dp[j] += dp[k];
            }
        }
    }
    if (__popc(F) != m) {
        int w = 1;
        for (int j = 1; j < m; j++)
            if ((F >> j) & 1u) {
                if (w < j) {
                    for (int k = 0; k < D; k++) dA[w * D + k] = dA[j * D + k];
                    dp[w] = dp[j];
                    dq[w] = dq[j];
                }
                w++;
            }
    }
}

// Tangents of every measurement-update output slot (after k_mu, which wrote the nominal
// outputs; the parents are read post propagation), and of each slot's g value at the
// root point, whose sum is the tangent of fz (eval_g_yei, cauchy_term.hpp:312-401).
// dg: [slot][direction] complex.
template <int D>
__global__ void k_mu_tan(TermPool IN, TanPool TI, int np, const int* __restrict__ slot_base, TermPool OUT,
                         TanPool TO, int nslots, const StepConst* __restrict__ sc, const StepTan* __restrict__ st,
                         const uint32_t* __restrict__ gk_prev, const cplx* __restrict__ gv_prev,
                         const cplx* __restrict__ dgv_prev, cplx* __restrict__ dg) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const int NP = TO.P;
    if (OUT.flags[s] == 0) {  // a child of an H-orthogonal row: no term
        for (int kk = 0; kk < NP; kk++) dg[(int64_t)s * NP + kk] = cmake(0, 0);
        return;
    }
    int lo = 0, hi = np - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (slot_base[mid] <= s) lo = mid;
        else hi = mid - 1;
    }
    const int Pp = lo;
    const int MM = IN.MM;
    const int m = IN.m[Pp];
    const int ts = s - slot_base[Pp];
    const int t = (ts == 0) ? m : ts - 1;
    const double* A = IN.A + (int64_t)Pp * MM * D;
    const double* p = IN.p + (int64_t)Pp * MM;
    const double* b = IN.b + (int64_t)Pp * D;
    const double* H = sc->H;
    const double zeta = sc->msmt - dev_dot<D>(H, b);
    const bool first = sc->first != 0;
    const bool post = !sc->skip && !first;
    // nominal parent-table lookup of this slot's g (independent of the direction)
    const int phc = OUT.phc[s], lhp = OUT.lhp[s];
    const int64_t go = IN.gp_off[Pp];
    const int gn = IN.gp_n[Pp];
    double mut[D], dmut[D], row[kMaxM * D], drow[kMaxM * D], rho[kMaxM], drho[kMaxM], An[kMaxM * D], n1s[kMaxM];
    uint32_t hf = 0;
    for (int kk = 0; kk < NP; kk++) {
        const double* dH = st->H[kk];
        const double* dA = TI.tA(Pp, kk);
        const double* dp = TI.tp(Pp, kk);
        const double* db = TI.tb(Pp, kk);
        const double dgamma = st->gamma[kk];
        double dzeta = 0;
        for (int k = 0; k < D; k++) dzeta -= dH[k] * b[k] + H[k] * db[k];
        double rho_t, drho_t;
        if (t == m) {
            for (int k = 0; k < D; k++) mut[k] = dmut[k] = 0.0;
            rho_t = sc->gamma;
            drho_t = dgamma;
        } else {
            const double h = dev_dot<D>(H, A + t * D);
            double dh = 0;
            for (int k = 0; k < D; k++) dh += dH[k] * A[t * D + k] + H[k] * dA[t * D + k];
            const double r = 1.0 / h;
            for (int k = 0; k < D; k++) {
                mut[k] = A[t * D + k] * r;
                dmut[k] = (dA[t * D + k] - mut[k] * dh) * r;
            }
            rho_t = p[t] * fabs(h);
            drho_t = dp[t] * fabs(h) + p[t] * sgn0(h) * dh;
        }
        hf = 0;
        for (int r = 0; r < m; r++) {
            bool horth;
            child_row_tan<D>(A, p, dA, dp, H, dH, mut, dmut, sc->gamma, dgamma, m, t, r, row + r * D, drow + r * D,
                             rho + r, drho + r, &horth);
            if (horth) hf |= 1u << r;
        }
        double* tb_ = TO.tb(s, kk);
        double bo[D];
        for (int k = 0; k < D; k++) {
            bo[k] = b[k] + zeta * mut[k];
            tb_[k] = db[k] + dzeta * mut[k] + zeta * dmut[k];
        }
        // g at the root point: signs of the child rows, ygi over the rows that are not
        // H-orthogonal, the two parent-table numerators
        double ygi = 0, dygi = 0;
        uint32_t neg = 0;
        for (int l = 0; l < m; l++) {
            const double sl = dev_dot<D>(row + l * D, sc->root) > 0 ? 1 : -1;
            if (sl < 0) neg |= (1u << l);
            if (!((hf >> l) & 1u)) {
                ygi += rho[l] * sl;
                dygi += drho[l] * sl;
            }
        }
        cplx gp = cmake(1, 0), gm = cmake(1, 0), dgp = cmake(0, 0), dgm = cmake(0, 0);
        if (!first) {
            const int top = 1 << (phc - 1), rev = (phc >= 32) ? -1 : ((1 << phc) - 1);
            int lp = 0, lmv = 0, k = 0;
            for (int l = 0; l < m; l++) {
                if (k < phc) {
                    if (k == t) {
                        lmv |= (1 << k);
                        k++;
                        if (k == phc) continue;
                    }
                    if ((neg >> l) & 1u) {
                        lp |= (1 << k);
                        lmv |= (1 << k);
                    }
                    k++;
                }
            }
            bool cj;
            int ix = lookup_idx(lp ^ lhp, top, rev, gk_prev + go, gn, &cj);
            gp = ix < 0 ? cmake(0, 0) : (cj ? cconj(gv_prev[go + ix]) : gv_prev[go + ix]);
            dgp = ix < 0 ? cmake(0, 0) : (cj ? cconj(dgv_prev[(go + ix) * NP + kk]) : dgv_prev[(go + ix) * NP + kk]);
            ix = lookup_idx(lmv ^ lhp, top, rev, gk_prev + go, gn, &cj);
            gm = ix < 0 ? cmake(0, 0) : (cj ? cconj(gv_prev[go + ix]) : gv_prev[go + ix]);
            dgm = ix < 0 ? cmake(0, 0) : (cj ? cconj(dgv_prev[(go + ix) * NP + kk]) : dgv_prev[(go + ix) * NP + kk]);
        }
        const cplx D1 = cmake(ygi + rho_t, zeta), D2 = cmake(ygi - rho_t, zeta);
        const cplx dD1 = cmake(dygi + drho_t, dzeta), dD2 = cmake(dygi - drho_t, dzeta);
        const cplx v = csub(dquot(gp, dgp, D1, dD1), dquot(gm, dgm, D2, dD2));
        dg[(int64_t)s * NP + kk] = cscale(v, kInvTwoPi);
        *TO.tc(s, kk) = dzeta;
        *TO.tdv(s, kk) = drho_t;
        double* oA = TO.tA(s, kk);
        double* op = TO.tp(s, kk);
        double* oq = TO.tq(s, kk);
        if (!post) {
            for (int r = 0; r < m; r++) {
                for (int k = 0; k < D; k++) oA[r * D + k] = drow[r * D + k];
                op[r] = drho[r];
            }
            continue;
        }
        // normalize_hps(true): q = p, p *= |row|_1, row /= |row|_1
        for (int r = 0; r < m; r++) {
            double n1 = 0, dn1 = 0;
            for (int k = 0; k < D; k++) {
                n1 += fabs(row[r * D + k]);
                dn1 += sgn0(row[r * D + k]) * drow[r * D + k];
            }
            n1s[r] = n1;
            for (int k = 0; k < D; k++) {
                An[r * D + k] = row[r * D + k] / n1;
                drow[r * D + k] = (drow[r * D + k] - An[r * D + k] * dn1) / n1;
            }
            oq[r] = drho[r];
            drho[r] = drho[r] * n1 + rho[r] * dn1;
        }
        if (t < m) {
            for (int r = 0; r < m; r++) {
                for (int k = 0; k < D; k++) oA[r * D + k] = drow[r * D + k];
                op[r] = drho[r];
            }
            mu_coalign_tan<D>(An, hf, m, oA, op, oq);
        } else {
            for (int r = 0; r < m; r++) {
                for (int k = 0; k < D; k++) oA[r * D + k] = drow[r * D + k];
                op[r] = drho[r];
            }
        }
    }
}

// Column sums of the [slot][direction] tangents: out[k] = sum over slots of dg[s][k],
// one block per direction, a fixed strided partial per thread then a fixed tree.
__global__ void k_tan_colsum(const cplx* __restrict__ dg, int n, int NP, cplx* __restrict__ out) {
    __shared__ double sr[256], si[256];
    const int k = blockIdx.x;
    const int t = threadIdx.x;
    double a = 0, c = 0;
    for (int s = t; s < n; s += blockDim.x) {
        const cplx v = dg[(int64_t)s * NP + k];
        a += v.re;
        c += v.im;
    }
    sr[t] = a;
    si[t] = c;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if (t < h) {
            sr[t] += sr[t + h];
            si[t] += si[t + h];
        }
        __syncthreads();
    }
    if (t == 0) out[k] = cmake(sr[0], si[0]);
}

// ------------------------------------------------------------------ g-tables
// Tangents of the first update's g-tables and of the next parents (make_gtable_first,
// flattening.hpp:14-67): g_j = (1 / (ygi + d + i c) - 1 / (ygi - d + i c)) * gscale.
template <int D>
__global__ void k_gtable_first_tan(const TermPool OUT, TanPool TO, int n, const uint32_t* __restrict__ perm,
                                   double gscale, const double* __restrict__ dgscale, cplx* __restrict__ dgv,
                                   TanPool TN) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int64_t s = perm[r];
    const int NP = TO.P;
    const int cells = 1 << (D - 1);
    const int MM = OUT.MM;
    const double* p = OUT.p + s * MM;
    const uint32_t hf = OUT.hflag[s];
    const double c = OUT.c[s], dv = OUT.dv[s];
    for (int kk = 0; kk < NP; kk++) {
        const double* dp = TO.tp(s, kk);
        const double dc = *TO.tc(s, kk), ddv = *TO.tdv(s, kk);
        for (int j = 0; j < cells; j++) {
            double ygi = 0, dygi = 0;
            for (int k = 0; k < D; k++)
                if (!((hf >> k) & 1u)) {
                    const double sb = ((j >> k) & 1) == 0 ? 1 : -1;
                    ygi += p[k] * sb;
                    dygi += dp[k] * sb;
                }
            const cplx one = cmake(1.0, 0.0), zero = cmake(0, 0);
            const cplx D1 = cmake(ygi + dv, c), D2 = cmake(ygi - dv, c);
            const cplx v = csub(cdiv(one, D1), cdiv(one, D2));
            const cplx dvv = csub(dquot(one, zero, D1, cmake(dygi + ddv, dc)), dquot(one, zero, D2, cmake(dygi - ddv, dc)));
            dgv[((int64_t)r * cells + j) * NP + kk] = cadd(cscale(dvv, gscale), cscale(v, dgscale[kk]));
        }
        for (int k = 0; k < D * D; k++) TN.tA(r, kk)[k] = TO.tA(s, kk)[k];
        for (int k = 0; k < D; k++) TN.tp(r, kk)[k] = dp[k];
        for (int k = 0; k < D; k++) TN.tb(r, kk)[k] = TO.tb(s, kk)[k];
    }
}

// Parent-side key pair of a child cell (make_gtable's lambda_p / lambda_m,
// flattening.hpp:155-216): an old term keeps its key below phc; a new child gets
// parent bit z inserted (0 in lp, 1 in lm), through its coalignment maps if merged.
__device__ inline void parent_keys(uint32_t key, bool is_new, bool has_cmap, int z, int phc, const uint8_t* cm,
                                   const int8_t* cs, int* lp_, int* lm_) {
    const int rev_phc = (phc >= 32) ? -1 : ((1 << phc) - 1);
    int lp = 0, lmv = 0;
    if (!is_new) {
        lp = (int)key & rev_phc;
        lmv = lp;
    } else if (!has_cmap) {
        int k = 0, l = 0;
        while (k < phc) {
            if (k == z) {
                lmv |= (1 << k);
                k++;
                if (k == phc) break;
            }
            if ((key >> l) & 1u) {
                lp |= (1 << k);
                lmv |= (1 << k);
            }
            k++;
            l++;
        }
    } else {
        int k = 0, l = 0;
        while (k < phc) {
            if (k == z) {
                lmv |= (1 << k);
                k++;
                if (k == phc) break;
            }
            const int bv = (((key >> cm[l]) & 1u) ? -1 : 1) * (int)cs[l];
            if (bv < 0) {
                lp |= (1 << k);
                lmv |= (1 << k);
            }
            k++;
            l++;
        }
    }
    *lp_ = lp;
    *lm_ = lmv;
}

// Tangent of term s's g value at cell `key` (make_gtable's loop body,
// flattening.hpp:129-248), direction kk: g = (a / (ygi + d + i c) - b / (ygi - d + i c))
// * gscale, with ygi = sum of q_k s_k over the rows that are not H-orthogonal.
__device__ inline cplx cell_tan(const TermPool& O, const TanPool& T, int64_t s, int kk, uint32_t key,
                                const uint32_t* gk_prev, const cplx* gv_prev, const cplx* dgv_prev, double gscale,
                                double dgscale) {
    const int MM = O.MM, NP = T.P;
    const int m = O.m[s];
    const uint32_t hf = O.hflag[s];
    const double* q = O.q + s * MM;
    const double* dq = T.tq(s, kk);
    double ygi = 0, dygi = 0;
    for (int k = 0; k < m; k++)
        if (!((hf >> k) & 1u)) {
            const double sb = ((key >> k) & 1u) ? -1.0 : 1.0;
            ygi += q[k] * sb;
            dygi += dq[k] * sb;
        }
    const uint8_t fl = O.flags[s];
    const int phc = O.phc[s];
    int lp, lmv;
    parent_keys(key, fl & kIsNew, fl & kHasCmap, O.z[s], phc, O.cmap + s * MM, O.csmap + s * MM, &lp, &lmv);
    const int lhp = O.lhp[s];
    const int top = 1 << (phc - 1), rev = (phc >= 32) ? -1 : ((1 << phc) - 1);
    const int64_t go = O.gp_off[s];
    const int gn = O.gp_n[s];
    bool cj;
    cplx a = cmake(0, 0), da = cmake(0, 0), bm = cmake(0, 0), dbm = cmake(0, 0);
    int ix = lookup_idx(lp ^ lhp, top, rev, gk_prev + go, gn, &cj);
    if (ix >= 0) {
        a = cj ? cconj(gv_prev[go + ix]) : gv_prev[go + ix];
        da = cj ? cconj(dgv_prev[(go + ix) * NP + kk]) : dgv_prev[(go + ix) * NP + kk];
    }
    ix = lookup_idx(lmv ^ lhp, top, rev, gk_prev + go, gn, &cj);
    if (ix >= 0) {
        bm = cj ? cconj(gv_prev[go + ix]) : gv_prev[go + ix];
        dbm = cj ? cconj(dgv_prev[(go + ix) * NP + kk]) : dgv_prev[(go + ix) * NP + kk];
    }
    const double c = O.c[s], dv = O.dv[s];
    const double dc = *T.tc(s, kk), ddv = *T.tdv(s, kk);
    const cplx D1 = cmake(ygi + dv, c), D2 = cmake(ygi - dv, c);
    const cplx v = csub(cdiv(a, D1), cdiv(bm, D2));
    const cplx dvv = csub(dquot(a, da, D1, cmake(dygi + ddv, dc)), dquot(bm, dbm, D2, cmake(dygi - ddv, dc)));
    return cadd(cscale(dvv, gscale), cscale(v, dgscale));
}

// Tangents of each reduction group's g-table (make_gtables, flattening.hpp:318-568), one
// wavefront per group, after k_group_eval (and k_group_sort) settled it: the surviving
// root's values at B ^ X (X: the replacement chain's orientation), plus, in member order,
// every member k_group_eval / k_group_members added (mem_add), at bi ^ sigma reversed and
// conjugated when in the lower half, where the member's table holds that key
// (add_gtables, flattening.hpp:259-315).  Written at the cell's position in the final
// (key-sorted) table; dgv: [cell][direction].
__global__ __launch_bounds__(64) void k_group_tan(const TermPool OUT, TanPool TO, int ng, int d,
                                                 const int* __restrict__ grp_moff, const int* __restrict__ grp_mcnt,
                                                 const int* __restrict__ grp_m, const int* __restrict__ mem_slots,
                                                 const uint8_t* __restrict__ mem_add, const int* __restrict__ surv,
                                                 const int* __restrict__ gxor, const uint32_t* __restrict__ B2,
                                                 const int64_t* __restrict__ b2_off, const int* __restrict__ b2_n,
                                                 const int64_t* __restrict__ g_off, const uint32_t* __restrict__ gk_prev,
                                                 const cplx* __restrict__ gv_prev, const cplx* __restrict__ dgv_prev,
                                                 const uint32_t* __restrict__ gk_out, double gscale,
                                                 const double* __restrict__ dgscale, cplx* __restrict__ dgv_out) {
    const int g = blockIdx.x;
    if (g >= ng) return;
    const int r = surv[g];
    if (r < 0) return;
    const int lane = threadIdx.x;
    const int NP = TO.P;
    const int m = grp_m[g];
    const int MM = OUT.MM;
    const uint32_t* B = B2 + b2_off[g];
    const int n = b2_n[g];
    const uint32_t X = (uint32_t)gxor[g];
    const uint32_t top = 1u << (m - 1);
    const uint32_t rev = (m >= 32) ? 0xffffffffu : ((1u << m) - 1u);
    const uint32_t* ok = gk_out + g_off[g];
    const int* mem = mem_slots + grp_moff[g];
    const uint8_t* add = mem_add + grp_moff[g];
    const int nmem = grp_mcnt[g];
    const double* Ar = OUT.A + (int64_t)r * MM * d;
    cplx* out = dgv_out + g_off[g] * NP;
    for (int kk = 0; kk < NP; kk++) {
        // the root's values; each lane owns its cells through the member sums
        for (int i = lane; i < n; i += kWave) {
            const uint32_t bi = B[i] ^ X;
            const int pos = X == 0 ? i : bsearch_u32(ok, bi, n);
            if (pos >= 0)
                out[(int64_t)pos * NP + kk] = cell_tan(OUT, TO, r, kk, bi, gk_prev, gv_prev, dgv_prev, gscale, dgscale[kk]);
        }
        for (int k = 0; k < nmem; k++) {
            if (!add[k]) continue;
            const int sk = mem[k];
            const int sig = orientation_mask(Ar, OUT.A + (int64_t)sk * MM * d, m, d);
            const uint32_t sh = ((uint32_t)sig & top) ? ((uint32_t)sig ^ rev) : (uint32_t)sig;
            for (int i = lane; i < n; i += kWave) {
                const uint32_t bi = B[i] ^ X;
                uint32_t x = bi ^ (uint32_t)sig;
                bool cj = false;
                if (x & top) {
                    cj = true;
                    x ^= rev;
                }
                if (bsearch_u32(B, x ^ sh ^ X, n) < 0) continue;
                const int pos = X == 0 ? i : bsearch_u32(ok, bi, n);
                if (pos < 0) continue;
                const cplx v = cell_tan(OUT, TO, sk, kk, x, gk_prev, gv_prev, dgv_prev, gscale, dgscale[kk]);
                out[(int64_t)pos * NP + kk] = cadd(out[(int64_t)pos * NP + kk], cj ? cconj(v) : v);
            }
        }
    }
}

// Survivors' tangents follow k_compact into the next parents' slots.
template <int D>
__global__ void k_compact_tan(TanPool TO, int ng, const int* __restrict__ surv, const int* __restrict__ pos,
                              const int* __restrict__ grp_m, TanPool TN) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    const int r = surv[g];
    if (r < 0) return;
    const int o = pos[g];
    const int m = grp_m[g];
    for (int kk = 0; kk < TO.P; kk++) {
        for (int k = 0; k < m * D; k++) TN.tA(o, kk)[k] = TO.tA(r, kk)[k];
        for (int k = 0; k < m; k++) TN.tp(o, kk)[k] = TO.tp(r, kk)[k];
        for (int k = 0; k < D; k++) TN.tb(o, kk)[k] = TO.tb(r, kk)[k];
    }
}

}  // namespace mce
