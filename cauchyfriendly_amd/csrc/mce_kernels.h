// HIP kernels of the MCE term-propagation step (gfx950).
//
// Per scalar measurement update the host drives, in order:
//   k_tp               time propagation + Gamma coalignment, one thread per parent
//   k_btab_simple      B-tables that need no enumeration (copy / identity / empty)
//   k_btab_enum        time-propagation B-table by vertex enumeration, one block per parent
//   k_mu               measurement update, one thread per output slot (parent x child)
//   k_mom_terms/sum    fz / mean / covariance summands, then in-order sums per
//                      reference thread (second stream, overlapped with reduction)
//   k_ftr_*            fast term reduction of all shape buckets in single launches
//   k_group_btab       B-table of each reduction group's root, one wavefront per group
//   k_group_eval       g-tables of each group: root, replacement chain, member sums
//   k_group_sort       key order for groups whose root was replaced
//   k_compact          survivors become the next step's parents
#pragma once
#include <climits>

#include <hipcub/block/block_radix_sort.hpp>

#include "mce_device.h"
#include "mce_multi.h"
#include "mce_pool.h"

namespace mce {

constexpr int kWave = 64;
// Gather key of a slot that holds no term: above every term's ((m T + tid) 2 + child), which
// stays below 336 for shapes up to 20 hyperplanes and T <= 8, so the sort needs 10 bits.
constexpr uint32_t kGatherKeyNone = 0x3ffu;
constexpr int kGatherKeyBits = 10;
// Default: groups with more members than this add them with one wavefront per member.
constexpr int kBigGroupMembers = 16;

enum BtabMode : uint8_t { kModeNone = 0, kModeCopy = 1, kModeIdent = 2, kModeTrivial = 3, kModeEmpty = 4, kModeEnum = 5 };

// Per-update constants (one small device buffer).
struct StepConst {
    double Phi[kMaxD * kMaxD];
    double H[kMaxD];
    double G[kMaxM * kMaxD];   // pre-coaligned Gamma rows
    double beta[kMaxM];
    double Bu[kMaxD];          // B u, computed once per update
    double root[kMaxD];
    double msmt, gamma, gscale;
    int d, ng, cmcc, with_tp, skip, first, T, threaded;
    int order[kMaxD];          // fast-term-reduction axis order
    int in_off[kMaxM + 2];     // parents' shape bucket offsets (bucket m: [in_off[m], in_off[m+1]))
    int64_t rk_base[kMaxM + 2];  // first rank-arena word of each shape bucket's rank structures
};

// ------------------------------------------------------------------ small dense linear algebra
// Column-sum one-norm, cauchy_linalg.hpp:1301-1314.
template <int D>
__device__ inline double dev_one_norm(const double* A) {
    double best = -1;
#pragma unroll
    for (int c = 0; c < D; c++) {
        double s = 0;
#pragma unroll
        for (int r = 0; r < D; r++) s += fabs(A[r * D + c]);
        if (s > best) best = s;
    }
    return best;
}

// In-place partial-pivot LU, cauchy_linalg.hpp:1180-1232.  Row exchanges are
// written as predicated swaps over compile-time indices so the matrix stays in
// registers; the arithmetic and its order are the reference's.
template <int D>
__device__ inline bool dev_lu(double* A, int* piv) {
#pragma unroll
    for (int j = 0; j < D; ++j) {
        double pv = kPluEps;
        int pi = -1;
#pragma unroll
        for (int i = j; i < D; ++i)
            if (fabs(A[i * D + j]) > fabs(pv)) {
                pv = A[i * D + j];
                pi = i;
            }
        if (pi == -1) return false;
#pragma unroll
        for (int r = j + 1; r < D; ++r) {
            const bool sw = (r == pi);
#pragma unroll
            for (int k = 0; k < D; k++) {
                const double a = A[j * D + k], b = A[r * D + k];
                A[j * D + k] = sw ? b : a;
                A[r * D + k] = sw ? a : b;
            }
        }
        piv[j] = pi;
#pragma unroll
        for (int k = j + 1; k < D; ++k) {
            A[k * D + j] /= A[j * D + j];
            const double t = A[k * D + j];
#pragma unroll
            for (int q = j + 1; q < D; q++) A[k * D + q] -= t * A[j * D + q];
        }
    }
    return true;
}

template <int D>
__device__ inline void dev_fwd(const double* LU, double* b) {
#pragma unroll
    for (int i = 0; i < D; i++) {
        double s = b[i];
#pragma unroll
        for (int j = 0; j < i; j++) s -= LU[i * D + j] * b[j];
        b[i] = s;
    }
}
template <int D>
__device__ inline void dev_bwd(const double* LU, double* b) {
#pragma unroll
    for (int i = D - 1; i >= 0; i--) {
        double s = b[i];
#pragma unroll
        for (int j = D - 1; j > i; j--) s -= LU[i * D + j] * b[j];
        b[i] = s / LU[i * D + i];
    }
}
// Inverse of the LAPACK-style pivot sequence (cauchy_linalg.hpp:1258-1277):
// pt[r] = position row r of the original matrix ends up at.
template <int D>
__device__ inline void dev_pivot_inverse(const int* piv, int* pt) {
    int preg[D];
#pragma unroll
    for (int i = 0; i < D; i++) preg[i] = i;
#pragma unroll
    for (int i = 0; i < D; i++) {
        // swap preg[i] and preg[piv[i]] (piv[i] >= i)
        const int a = preg[i];
        int b = a;
#pragma unroll
        for (int r = i; r < D; r++) b = (r == piv[i]) ? preg[r] : b;
#pragma unroll
        for (int r = i; r < D; r++) preg[r] = (r == piv[i]) ? a : preg[r];
        preg[i] = b;
    }
#pragma unroll
    for (int r = 0; r < D; r++) {
        int v = 0;
#pragma unroll
        for (int i = 0; i < D; i++) v = (preg[i] == r) ? i : v;
        pt[r] = v;
    }
}

// One-norm condition number via explicit inverse, cauchy_linalg.hpp:1337-1392.
// The callers only compare it with kCondEps.  A cheap bound decides first: every
// entry of A^-1 is a cofactor over det A, and by Hadamard's inequality a cofactor is
// at most rn^(D-1), rn the largest row 1-norm, so cond1 <= ||A||_1 * D * rn^(D-1) /
// |det A|, with |det A| the product of the LU pivots.  When that bound is below
// kCondScreen (three orders of magnitude under kCondEps, far beyond the rounding of
// the pivots) the bound is returned; otherwise the reference's value is computed.
constexpr double kCondScreen = 1e9;
template <int D>
__device__ inline double dev_cond1(double* A, int* piv) {
    const double na = dev_one_norm<D>(A);
    double rn = 0;
#pragma unroll
    for (int r = 0; r < D; r++) {
        double s = 0;
#pragma unroll
        for (int c = 0; c < D; c++) s += fabs(A[r * D + c]);
        rn = fmax(rn, s);
    }
    if (!dev_lu<D>(A, piv)) return 1.7976931348623157e308;
    {
        double det = 1.0, bnd = na * (double)D;
#pragma unroll
        for (int i = 0; i < D; i++) det *= fabs(A[i * D + i]);
#pragma unroll
        for (int i = 0; i < D - 1; i++) bnd *= rn;
        if (bnd <= kCondScreen * det) return bnd / det;
    }
    int pt[D];
    dev_pivot_inverse<D>(piv, pt);
    // one-norm of the inverse = largest absolute row sum of W, whose row i solves
    // A w = e_i (W is the transpose of the inverse before reflect_array)
    double best = -1;
#pragma unroll
    for (int i = 0; i < D; i++) {
        double w[D];
#pragma unroll
        for (int c = 0; c < D; c++) w[c] = (c == pt[i]) ? 1.0 : 0.0;
        dev_fwd<D>(A, w);
        dev_bwd<D>(A, w);
        double s = 0;
#pragma unroll
        for (int r = 0; r < D; r++) s += fabs(w[r]);
        if (s > best) best = s;
    }
    return na * best;
}

template <int D>
__device__ inline void dev_lu_solve(const double* LU, const int* piv, const double* b, double* x) {
    int pt[D];
    dev_pivot_inverse<D>(piv, pt);
#pragma unroll
    for (int c = 0; c < D; c++) {
        double v = 0.0;
#pragma unroll
        for (int i = 0; i < D; i++) v = (pt[i] == c) ? b[i] : v;
        x[c] = v;
    }
    dev_fwd<D>(LU, x);
    dev_bwd<D>(LU, x);
}

template <int D>
__device__ inline double dev_dot(const double* x, const double* y) {
    double z = 0.0;
    for (int i = 0; i < D; i++) z += x[i] * y[i];
    return z;
}

// ------------------------------------------------------------------ wave helpers
__device__ inline int wave_lane() { return threadIdx.x & (kWave - 1); }

__device__ inline int wave_excl_scan(int v, int* total) {
    const int lane = wave_lane();
    int x = v;
    for (int o = 1; o < kWave; o <<= 1) {
        int y = __shfl_up(x, o, kWave);
        if (lane >= o) x += y;
    }
    *total = __shfl(x, kWave - 1, kWave);
    return x - v;
}

__device__ inline bool wave_all(bool pred) { return __all(pred); }

// Orders the LDS accesses of the lanes of one wavefront (for workgroups of one wavefront):
// a wavefront's LDS operations complete in issue order, so only the compiler has to be
// held back, and the wavefront does not wait for its global loads and stores as a
// workgroup barrier makes it.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Inclusive scan of one value per lane across a wavefront (shuffles, no barrier).
template <class T>
__device__ __forceinline__ T wave_incl_scan(T x) {
    const int lane = (int)(threadIdx.x & (kWave - 1));
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const T y = __shfl_up(x, o, kWave);
        if (lane >= o) x += y;
    }
    return x;
}

// Exclusive prefix of each thread's v over a block of whole wavefronts (at most 4) and the
// block total, through wsum[4] in LDS (two barriers).  Every thread of the block calls it.
__device__ __forceinline__ int block4_excl_scan(int v, int* total, int* wsum) {
    const int lane = (int)(threadIdx.x & (kWave - 1)), w = (int)(threadIdx.x / kWave);
    const int nw = (int)(blockDim.x / kWave);
    const int x = wave_incl_scan(v);
    __syncthreads();  // wsum of an earlier scan has been read
    if (lane == kWave - 1) wsum[w] = x;
    __syncthreads();
    int before = 0, tot = 0;
    for (int k = 0; k < nw; k++) {
        const int sk = wsum[k];
        before += k < w ? sk : 0;
        tot += sk;
    }
    *total = tot;
    return before + x - v;
}

// ------------------------------------------------------------------ time propagation
// cauchy_term.hpp:435-531 (time_prop, normalize_hps, tp_coalign), with the B-table
// mode decision of cauchy_estimator.hpp:678-708 (serial) / 1614-1637 (threaded).
template <int D>
__device__ __forceinline__ void k_tp(MG<1024, 1> mg, TermPool P, int n, const StepConst* __restrict__ sc, const int* __restrict__ cen_half,
                     int* __restrict__ slots, int* __restrict__ btcap, int* __restrict__ err) {
    const int i = mg.bx * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int MM = P.MM;
    int m = P.m[i];
    const int phc = m;
    P.phc[i] = phc;
    double* A = P.A + (int64_t)i * MM * D;
    double* p = P.p + (int64_t)i * MM;
    double* b = P.b + (int64_t)i * D;
    if (sc->with_tp) {
        double t[D];
        for (int r = 0; r < m; r++) {
            for (int k = 0; k < D; k++) t[k] = A[r * D + k];
            for (int j = 0; j < D; j++) {
                double s = 0.0;
                for (int k = 0; k < D; k++) s += t[k] * sc->Phi[k + j * D];
                A[r * D + j] = s;
            }
        }
        for (int k = 0; k < D; k++) t[k] = b[k];
        for (int r = 0; r < D; r++) {
            double s = 0.0;
            for (int j = 0; j < D; j++) s += sc->Phi[r * D + j] * t[j];
            b[r] = s;
        }
        if (sc->cmcc > 0)
            for (int r = 0; r < D; r++) b[r] += 1.0 * sc->Bu[r];
        for (int r = 0; r < m; r++) {
            double n1 = 0;
            for (int k = 0; k < D; k++) n1 += fabs(A[r * D + k]);
            p[r] *= n1;
            for (int k = 0; k < D; k++) A[r * D + k] /= n1;
        }
        uint32_t F = (sc->ng >= 32) ? 0xffffffffu : ((1u << sc->ng) - 1u);
        for (int j = 0; j < sc->ng; j++)
            for (int k = 0; k < m; k++) {
                if (!((F >> j) & 1u)) continue;
                bool pos, neg;
                coalign_gates(&sc->G[j * D], &A[k * D], D, pos, neg);
                if (pos || neg) {
                    F &= ~(1u << j);
                    p[k] += sc->beta[j];
                }
            }
        for (int j = 0; j < sc->ng; j++)
            if ((F >> j) & 1u) {
                if (m >= MM) {
                    atomicOr(err, 4);
                    break;
                }
                for (int k = 0; k < D; k++) A[m * D + k] = sc->G[j * D + k];
                p[m] = sc->beta[j];
                m++;
            }
        P.m[i] = m;
    }
    int tid = 0;
    if (sc->threaded) tid = (i - sc->in_off[phc]) % sc->T;
    P.tid[i] = tid;
    uint8_t mode;
    int cap;
    const int gn = P.gp_n[i];
    if (sc->skip) {
        mode = kModeNone;
        cap = 0;
    } else if (!sc->with_tp) {
        mode = kModeCopy;
        cap = gn;
    } else if (sc->threaded && m <= D) {
        mode = kModeIdent;
        cap = gn;
    } else if (m == phc) {
        mode = kModeCopy;
        cap = gn;
    } else if (m < D) {
        mode = kModeTrivial;
        cap = max(1 << (m - 1), gn);
    } else if (m == D) {
        mode = kModeEmpty;
        cap = 0;
    } else {
        mode = kModeEnum;
        cap = cen_half[m];
    }
    P.mode[i] = mode;
    btcap[i] = cap;
    slots[i] = m + 1;
}

// B-tables that need no enumeration; enumeration jobs are appended to a list.
__device__ __forceinline__ void k_btab_simple(MG<1024, 1> mg, TermPool P, int n, const int64_t* __restrict__ bt_base, const uint32_t* __restrict__ gkeys,
                              uint32_t* __restrict__ Barena, int* __restrict__ enum_list, int* __restrict__ enum_count) {
    const int i = mg.bx * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint8_t mode = P.mode[i];
    const int64_t off = bt_base[i];
    P.bt_off[i] = off;
    P.bt_sgn[i] = 0;
    const int gn = P.gp_n[i];
    uint32_t* B = Barena + off;
    switch (mode) {
        case kModeCopy: {
            const uint32_t* src = gkeys + P.gp_off[i];
            for (int k = 0; k < gn; k++) B[k] = src[k];
            P.bt_n[i] = gn;
        } break;
        case kModeIdent:
            for (int k = 0; k < gn; k++) B[k] = (uint32_t)k;
            P.bt_n[i] = gn;
            break;
        case kModeTrivial: {
            const int m = P.m[i];
            const int nn = max(1 << (m - 1), gn);
            for (int k = 0; k < nn; k++) B[k] = (uint32_t)k;
            P.bt_n[i] = gn;
        } break;
        case kModeEnum: {
            const int slot = atomicAdd(enum_count, 1);
            enum_list[slot] = i;
            // the enumerator's LDS is sized by the largest shape and parent hyperplane
            // count among the parents it visits
            atomicMax(enum_count + 4, P.m[i]);
            atomicMax(enum_count + 5, P.phc[i]);
            P.bt_n[i] = 0;
        } break;
        default:
            P.bt_n[i] = 0;
            break;
    }
}

// The sign vectors of the cells around one vertex v of the arrangement: every hyperplane
// not among the vertex's D (cmask) takes the side of v; the D of the vertex take both.
template <int D>
__device__ __forceinline__ void enum_vertex_cells(const double* __restrict__ A, const double* __restrict__ bp, int m,
                                                  const uint8_t* __restrict__ cb, const double* v, uint32_t* bits) {
    uint32_t cbit[D], cmask = 0;
#pragma unroll
    for (int k = 0; k < D; k++) {
        cbit[k] = 1u << cb[k];
        cmask |= cbit[k];
    }
    // every hyperplane's side is computed (m is uniform, so no lane waits on another);
    // those of the vertex are dropped
    uint32_t niv = 0;
    for (int ac = 0; ac < m; ac++)
        if ((dev_dot<D>(A + ac * D, v) - bp[ac]) < 0) niv |= (1u << ac);
    niv &= ~cmask;
    for (int s = 0; s < (1 << D); s++) {
        uint32_t sv = niv;
#pragma unroll
        for (int k = 0; k < D; k++)
            if ((s >> k) & 1) sv |= cbit[k];
        atomicOr(&bits[sv >> 5], 1u << (sv & 31));
    }
}

// Cells of the time-propagated arrangement by vertex enumeration,
// cell_enumeration.hpp:676-855.  One block per parent; the sign vectors seen are
// an LDS bitset of 2^m bits, so the output comes out sorted and de-duplicated.
// The parent's g-table keys are a second LDS bitset, `pofs` words after the first,
// or, when `psearch` is set (both sets would not fit in LDS: 2^20 + 2^18 bits), are
// looked up by binary search in the parent's sorted keys in global memory.
//
// A vertex is kept when the one-norm condition number of its D hyperplanes is at most
// kCondEps.  The block decides that for T = blockDim.x combinations at a time, cheapest
// test first, each test exact against the reference's value:
//   * the LU factorisation fails (a pivot under kPluEps): the reference's value is the
//     largest double, rejected;
//   * the upper bound of dev_cond1 is under kCondScreen: kept;
//   * ||A||_1 / |u_DD| > kCondEps, u_DD the last pivot: rejected.  The reference's inverse
//     has the entry fl(1 / u_DD) (the solve of the unit vector that the pivots move to the
//     last row: forward substitution leaves it a unit vector, back substitution divides
//     it by u_DD), its norm is a sum of absolute values no smaller than that entry, and
//     rounding is monotone, so the reference's product ||A||_1 ||A^-1||_1 is at least this
//     product;
//   * otherwise the reference's value is computed (dev_cond1).
// On the window workloads about 28% of the combinations fail to factor, 63% pass the
// bound, 8% are rejected by the last pivot and 2% need the full value.  The kept vertices
// (with their solution) and the undecided combinations go to two LDS queues, which are
// worked off T at a time, so that the vertex work and the full condition number run on
// full wavefronts rather than on the lanes that one combination in each wavefront keeps.
template <int D>
__device__ __forceinline__ void k_btab_enum(MG<256, 3> mg, TermPool P, const int* __restrict__ list,
                                                   const double* __restrict__ bpert, const uint8_t* __restrict__ combos,
                                                   const int64_t* __restrict__ combo_off, const int* __restrict__ combo_cnt,
                                                   const uint32_t* __restrict__ gkeys, uint32_t* __restrict__ Barena,
                                                   const int* __restrict__ btcap, int* __restrict__ err, int pofs,
                                                   int psearch) {
    extern __shared__ double enum_lds[];
    __shared__ int wsum[4];
    const int T = (int)blockDim.x, Q = 2 * T;
    // queues: kept vertices (combination, solution) and undecided combinations, 2T each
    double* qv = enum_lds;                    // [D][Q]
    int* qa = (int*)(qv + (size_t)D * Q);     // [Q]
    int* qs = qa + Q;                         // [Q]
    uint32_t* bits = (uint32_t*)(qs + Q);
    const int i = list[mg.bx];
    const int MM = P.MM;
    const int m = P.m[i];
    const int phc = P.phc[i];
    const int W = m >= 5 ? (1 << (m - 5)) : 1;
    // bitset of the parent's g-table keys (all below 2^(phc-1)), after the 2^m-bit
    // sets of candidate sign vectors of every parent of the launch
    uint32_t* pbits = bits + pofs;
    const uint32_t ptop = 1u << (phc - 1);
    const int PW = psearch ? 0 : (int)((ptop + 31u) >> 5);
    for (int w = threadIdx.x; w < W; w += blockDim.x) bits[w] = 0;
    for (int w = threadIdx.x; w < PW; w += blockDim.x) pbits[w] = 0;
    __syncthreads();
    const uint32_t* gk = gkeys + P.gp_off[i];
    const int gn = P.gp_n[i];
    // a parent key at or above 2^(phc-1) (a coaligned child's cell kept with bit phc-1
    // set) can never equal a looked-up sign vector, which is reversed below it
    if (!psearch)
        for (int q = threadIdx.x; q < gn; q += blockDim.x) {
            const uint32_t key = gk[q];
            if (key < ptop) atomicOr(&pbits[key >> 5], 1u << (key & 31));
        }
    const double* A = P.A + (int64_t)i * MM * D;
    const double* bp = bpert + (int64_t)P.tid[i] * MM;
    const uint8_t* cmb = combos + combo_off[m];
    const int nc = combo_cnt[m];
    auto load = [&](int ci, double* Ac, double* bc) {
        const uint8_t* cb = cmb + ci * D;
#pragma unroll
        for (int j = 0; j < D; j++) {
            const int r = cb[j];
#pragma unroll
            for (int k = 0; k < D; k++) Ac[j * D + k] = A[r * D + k];
            bc[j] = bp[r];
        }
    };
    // appends this thread's kept vertex and / or undecided combination (block-wide call)
    int na = 0, ns = 0;
    auto push = [&](bool keep, bool slow, int ci, const double* v) {
        int tot;
        const int pre = block4_excl_scan((keep ? 1 : 0) | (slow ? 1 << 16 : 0), &tot, wsum);
        if (keep) {
            const int q = na + (pre & 0xffff);
            qa[q] = ci;
#pragma unroll
            for (int k = 0; k < D; k++) qv[k * Q + q] = v[k];
        }
        if (slow) qs[ns + (pre >> 16)] = ci;
        na += tot & 0xffff;
        ns += tot >> 16;
        __syncthreads();
    };
    // the cells around the last `cnt` kept vertices (block-wide call)
    auto drain = [&](int cnt) {
        const int q = na - cnt + (int)threadIdx.x;
        if ((int)threadIdx.x < cnt) {
            double v[D];
#pragma unroll
            for (int k = 0; k < D; k++) v[k] = qv[k * Q + q];
            enum_vertex_cells<D>(A, bp, m, cmb + qa[q] * D, v, bits);
        }
        na -= cnt;
        __syncthreads();
    };
    // the reference's condition number of the last `cnt` undecided combinations
    auto settle = [&](int cnt) {
        const int q = ns - cnt + (int)threadIdx.x;
        bool keep = false;
        int ci = 0;
        double v[D];
        if ((int)threadIdx.x < cnt) {
            ci = qs[q];
            double Ac[D * D], bc[D];
            int piv[D];
            load(ci, Ac, bc);
            if (dev_cond1<D>(Ac, piv) <= kCondEps) {
                dev_lu_solve<D>(Ac, piv, bc, v);
                keep = true;
            }
        }
        ns -= cnt;
        push(keep, false, ci, v);
    };
    for (int base = 0; base < nc; base += T) {
        const int ci = base + (int)threadIdx.x;
        bool keep = false, slow = false;
        double v[D];
        if (ci < nc) {
            double Ac[D * D], bc[D];
            int piv[D];
            load(ci, Ac, bc);
            const double na1 = dev_one_norm<D>(Ac);
            double rn = 0;
#pragma unroll
            for (int r = 0; r < D; r++) {
                double sr = 0;
#pragma unroll
                for (int c = 0; c < D; c++) sr += fabs(Ac[r * D + c]);
                rn = fmax(rn, sr);
            }
            if (dev_lu<D>(Ac, piv)) {
                double det = 1.0, bnd = na1 * (double)D;
#pragma unroll
                for (int k = 0; k < D; k++) det *= fabs(Ac[k * D + k]);
#pragma unroll
                for (int k = 0; k < D - 1; k++) bnd *= rn;
                if (bnd <= kCondScreen * det) {
                    dev_lu_solve<D>(Ac, piv, bc, v);
                    keep = true;
                } else if (!(na1 * (1.0 / fabs(Ac[D * D - 1])) > kCondEps)) {
                    slow = true;
                }
            }
        }
        push(keep, slow, ci, v);
        if (na >= T) drain(T);
        if (ns >= T) {
            settle(T);
            if (na >= T) drain(T);
        }
    }
    if (ns > 0) settle(ns);
    while (na > 0) drain(min(na, T));
    __syncthreads();
    // keep sign vectors whose parent part is a cell of the parent arrangement
    const uint32_t top_phc = 1u << (phc - 1), rev_phc = (phc >= 32) ? 0xffffffffu : ((1u << phc) - 1u);
    const uint32_t valid_mask = (m >= 5) ? 0xffffffffu : ((1u << (1 << m)) - 1u);
    for (int w = threadIdx.x; w < W; w += blockDim.x) {
        uint32_t x = bits[w] & valid_mask, keep = 0;
        while (x) {
            const int bt = __ffs(x) - 1;
            x &= x - 1;
            uint32_t psv = ((uint32_t)(w * 32 + bt)) & rev_phc;
            if (psv & top_phc) psv ^= rev_phc;
            bool in_parent;
            if (!psearch) {
                in_parent = (pbits[psv >> 5] >> (psv & 31)) & 1u;
            } else {
                int lo = 0, hi = gn;  // first key >= psv
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (gk[mid] < psv) lo = mid + 1;
                    else hi = mid;
                }
                in_parent = lo < gn && gk[lo] == psv;
            }
            if (in_parent) keep |= (1u << bt);
        }
        bits[w] = keep;
    }
    __syncthreads();
    // antipodal pairs; the half-storage representative has bit m-1 clear
    const uint32_t rev_m = (m >= 32) ? 0xffffffffu : ((1u << m) - 1u);
    const int nlow = 1 << (m - 1);  // candidate keys [0, 2^(m-1))
    const int WL = (nlow + 31) >> 5;
    const int per = (WL + blockDim.x - 1) / blockDim.x;
    const int w0 = threadIdx.x * per, w1 = min(WL, w0 + per);
    int cnt = 0;
    for (int w = w0; w < w1; w++) {
        uint32_t x = bits[w];
        if (nlow < 32) x &= (1u << nlow) - 1u;
        while (x) {
            const int bt = __ffs(x) - 1;
            x &= x - 1;
            const uint32_t sv = (uint32_t)(w * 32 + bt);
            const uint32_t pr = sv ^ rev_m;
            if ((bits[pr >> 5] >> (pr & 31)) & 1u) cnt++;
        }
    }
    int total;
    int pos = block4_excl_scan(cnt, &total, wsum);
    uint32_t* out = Barena + P.bt_off[i];
    const int cap = btcap[i];
    for (int w = w0; w < w1; w++) {
        uint32_t x = bits[w];
        if (nlow < 32) x &= (1u << nlow) - 1u;
        while (x) {
            const int bt = __ffs(x) - 1;
            x &= x - 1;
            const uint32_t sv = (uint32_t)(w * 32 + bt);
            const uint32_t pr = sv ^ rev_m;
            if ((bits[pr >> 5] >> (pr & 31)) & 1u) {
                if (pos < cap) out[pos] = sv;
                pos++;
            }
        }
    }
    if (threadIdx.x == 0) {
        P.bt_n[i] = min(total, cap);
        if (total > cap) atomicOr(err, 8);
    }
}

// ------------------------------------------------------------------ g evaluation helpers
// g value and y_ei of one term at the estimator's root point, cauchy_term.hpp:312-401.
template <int D>
__device__ inline cplx dev_eval_g_yei(const double* A, const double* p, const double* b, double c, double dv, int m,
                                      uint32_t hf, int phc, int z, int lhp, const uint32_t* gk, const cplx* gv,
                                      int gn, const double* root, bool first, double* ty) {
    for (int k = 0; k < D; k++) ty[k] = 0.0;
    double ygi = 0.0;
    uint32_t neg = 0;
    for (int l = 0; l < m; l++) {
        const double s = dev_dot<D>(A + l * D, root) > 0 ? 1 : -1;
        if (s < 0) neg |= (1u << l);
        const double scl = p[l] * s;
        for (int k = 0; k < D; k++) ty[k] += scl * A[l * D + k];
        if (!((hf >> l) & 1u)) ygi += p[l] * s;
    }
    cplx gp, gm;
    if (first) {
        gp = cmake(1, 0);
        gm = cmake(1, 0);
    } else {
        const int top = 1 << (phc - 1), rev = (phc >= 32) ? -1 : ((1 << phc) - 1);
        int lp = 0, lmv = 0, k = 0;
        for (int l = 0; l < m; l++) {
            if (k < phc) {
                if (k == z) {
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
        gp = lookup_num(lp ^ lhp, top, rev, gk, gv, gn);
        gm = lookup_num(lmv ^ lhp, top, rev, gk, gv, gn);
    }
    cplx g = csub(cdiv(gp, cmake(ygi + dv, c)), cdiv(gm, cmake(ygi - dv, c)));
    return cscale(g, kInvTwoPi);
}

// Measurement-update coalignment in place, cauchy_term.hpp:533-745.  Returns the
// new shape; sets *cm when hyperplanes were merged (c_map kept).
template <int D>
__device__ inline int dev_mu_coalign(double* A, double* p, double* q, uint8_t* cmap, int8_t* csmap, uint32_t* hfp,
                                     int m, int* err) {
    for (int r = 0; r < m; r++) {
        q[r] = p[r];
        double n1 = 0;
        for (int k = 0; k < D; k++) n1 += fabs(A[r * D + k]);
        p[r] *= n1;
        for (int k = 0; k < D; k++) A[r * D + k] /= n1;
        cmap[r] = 255;
        csmap[r] = 1;
    }
    const uint32_t hf = *hfp;
    uint32_t hu = hf;
    uint32_t F = (m >= 32) ? 0xffffffffu : ((1u << m) - 1u);
    int u = 0;
    for (int j = 0; j < m - 1; j++) {
        if (!((F >> j) & 1u)) continue;
        cmap[j] = (uint8_t)u;
        for (int k = j + 1; k < m; k++) {
            if (!((F >> k) & 1u)) continue;
            bool pos, neg;
            coalign_gates(A + j * D, A + k * D, D, pos, neg);
            if (pos) {
                if (hf) {
                    const bool hj = (hu >> j) & 1u, hk = (hu >> k) & 1u;
                    if (!hj && !hk) q[j] += q[k];
                    else if (hj && !hk) {
                        q[j] = q[k];
                        hu &= ~(1u << j);
                    } else if (!hj && hk) {
                        hu &= ~(1u << k);
                    } else {
                        hu &= ~(1u << k);
                        q[j] += q[k];
                    }
                } else {
                    q[j] += q[k];
                }
                F &= ~(1u << k);
                p[j] += p[k];
                cmap[k] = (uint8_t)u;
                csmap[k] = 1;
            }
            if (neg) {
                if (hf) {
                    const bool hj = (hu >> j) & 1u, hk = (hu >> k) & 1u;
                    if (!hj && !hk) q[j] -= q[k];
                    else if (hj && !hk) {
                        q[j] = -q[k];
                        hu &= ~(1u << j);
                    } else if (!hj && hk) {
                        hu &= ~(1u << k);
                    } else {
                        atomicOr(err, 1);  // the reference exits here (cauchy_term.hpp:676-685)
                    }
                } else {
                    q[j] -= q[k];
                }
                F &= ~(1u << k);
                p[j] += p[k];
                cmap[k] = (uint8_t)u;
                csmap[k] = -1;
            }
        }
        u += 1;
    }
    if ((F >> (m - 1)) & 1u) cmap[m - 1] = (uint8_t)u;
    const int ns = __popc(F);
    if (ns != m) {
        int w = 1;
        for (int j = 1; j < m; j++)
            if ((F >> j) & 1u) {
                if (w < j) {
                    for (int k = 0; k < D; k++) A[w * D + k] = A[j * D + k];
                    p[w] = p[j];
                    q[w] = q[j];
                }
                w++;
            }
        if (hf) {
            uint32_t nh = 0;
            int cc = 0;
            for (int j = 0; j < m; j++)
                if ((F >> j) & 1u) nh |= (((hu >> j) & 1u) << cc++);
            *hfp = nh;
        }
    }
    return ns;
}

// ------------------------------------------------------------------ measurement update
// cauchy_term.hpp:84-310 (msmt_update) + cache_moments inputs + MU coalignment,
// one thread per output slot.  Slot t < m of parent P is its child t; slot m is
// the parent itself.  Writes the gather key (shape, reference thread, is_child)
// that reproduces the reference's term order (cauchy_estimator.hpp:898-937, 1683-1704).
template <int D>
__device__ __forceinline__ void k_mu(MG<1024, 1> mg, TermPool IN, int np, const int* __restrict__ slot_base, TermPool OUT, int nslots,
                     const StepConst* __restrict__ sc, const uint32_t* __restrict__ gk_prev,
                     const cplx* __restrict__ gv_prev, double* __restrict__ gy, uint32_t* __restrict__ keys_out,
                     uint32_t* __restrict__ chain_key, int* __restrict__ err) {
    const int s = mg.bx * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    int lo = 0, hi = np - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (slot_base[mid] <= s) lo = mid;
        else hi = mid - 1;
    }
    const int P = lo;
    const int MM = IN.MM;
    const int m = IN.m[P];
    // slot 0 of a parent holds the parent itself (child m), slots 1..m its children 0..m-1
    const int ts = s - slot_base[P];
    const int t = (ts == 0) ? m : ts - 1;
    const double* A = IN.A + (int64_t)P * MM * D;
    const double* p = IN.p + (int64_t)P * MM;
    const double* b = IN.b + (int64_t)P * D;
    const double* H = sc->H;
    uint32_t integ = 0, negs = 0;
    for (int l = 0; l < m; l++) {
        const double h = dev_dot<D>(H, A + l * D);
        if (!(fabs(h) < kMuEps)) {
            integ |= (1u << l);
            if (!(h > 0)) negs |= (1u << l);
        }
    }
    double* g_out = gy + (int64_t)s * (2 + 2 * D);
    if (t < m && !((integ >> t) & 1u)) {
        OUT.flags[s] = 0;
        keys_out[s] = kGatherKeyNone;
        chain_key[s] = 15u;
        for (int k = 0; k < 2 + 2 * D; k++) g_out[k] = 0.0;
        return;
    }
    chain_key[s] = (uint32_t)IN.tid[P];
    const double zeta = sc->msmt - dev_dot<D>(H, b);
    double mut[D];
    double rho_t;
    if (t == m) {
        for (int k = 0; k < D; k++) mut[k] = 0.0;
        rho_t = sc->gamma;
    } else {
        const double h = dev_dot<D>(H, A + t * D);
        const double r = 1.0 / h;
        for (int k = 0; k < D; k++) mut[k] = A[t * D + k] * r;
        rho_t = p[t] * fabs(h);
    }
    double* oA = OUT.A + (int64_t)s * MM * D;
    double* op = OUT.p + (int64_t)s * MM;
    double* oq = OUT.q + (int64_t)s * MM;
    double* ob = OUT.b + (int64_t)s * D;
    for (int k = 0; k < D; k++) ob[k] = b[k] + zeta * mut[k];
    int lo_row = 0;
    uint32_t hf = 0;
    for (int ll = 0; ll <= m; ll++) {
        if (ll == t) continue;
        double rho;
        if (ll == m) {
            rho = sc->gamma;
            for (int k = 0; k < D; k++) oA[lo_row * D + k] = 0.0 - mut[k];
        } else {
            const double h = dev_dot<D>(H, A + ll * D);
            if (fabs(h) < kMuEps) {
                rho = p[ll];
                for (int k = 0; k < D; k++) oA[lo_row * D + k] = A[ll * D + k];
                hf |= (1u << lo_row);
            } else {
                const double r = 1.0 / h;
                rho = p[ll] * fabs(h);
                for (int k = 0; k < D; k++) oA[lo_row * D + k] = (A[ll * D + k] * r) - mut[k];
            }
        }
        op[lo_row] = rho;
        lo_row++;
    }
    const int phc = IN.phc[P];
    OUT.m[s] = m;
    OUT.c[s] = zeta;
    OUT.dv[s] = rho_t;
    OUT.phc[s] = phc;
    OUT.pbc[s] = m;
    OUT.z[s] = t;
    OUT.gp_off[s] = IN.gp_off[P];
    OUT.rk_off[s] = IN.rk_off[P];
    OUT.gp_n[s] = IN.gp_n[P];
    const uint32_t enc = negs;
    uint32_t lhp = enc;
    if (phc < m) lhp &= (1u << phc) - 1u;
    uint32_t sg = enc;
    const uint32_t rev_m = (m >= 32) ? 0xffffffffu : ((1u << m) - 1u);
    if (sg & (1u << (m - 1))) sg ^= rev_m;
    OUT.lhp[s] = (int)lhp;
    OUT.bt_off[s] = IN.bt_off[P];
    OUT.bt_n[s] = IN.bt_n[P];
    OUT.bt_sgn[s] = sc->skip ? 0 : (int)sg;
    OUT.tid[s] = IN.tid[P];
    // moments: g and y_ei of this term (cache_moments, cauchy_estimator.hpp:307-338)
    {
        double ty[D];
        const int64_t go = IN.gp_off[P];
        const cplx g = dev_eval_g_yei<D>(oA, op, ob, zeta, rho_t, m, hf, phc, t, (int)lhp, gk_prev + go, gv_prev + go,
                                         IN.gp_n[P], sc->root, sc->first != 0, ty);
        g_out[0] = g.re;
        g_out[1] = g.im;
        for (int k = 0; k < D; k++) {
            g_out[2 + 2 * k] = -ty[k];
            g_out[3 + 2 * k] = ob[k];
        }
    }
    int mnew = m;
    uint8_t fl = kValid | ((t != m) ? kIsNew : 0);
    if (!sc->skip && !sc->first) {
        uint8_t* cm = OUT.cmap + (int64_t)s * MM;
        int8_t* cs = OUT.csmap + (int64_t)s * MM;
        if (t == m) {
            for (int r = 0; r < m; r++) {
                oq[r] = op[r];
                double n1 = 0;
                for (int k = 0; k < D; k++) n1 += fabs(oA[r * D + k]);
                op[r] *= n1;
                for (int k = 0; k < D; k++) oA[r * D + k] /= n1;
            }
        } else {
            mnew = dev_mu_coalign<D>(oA, op, oq, cm, cs, &hf, m, err);
            if (mnew < m) fl |= kHasCmap;
        }
    }
    OUT.hflag[s] = hf;
    OUT.flags[s] = fl;
    OUT.m[s] = mnew;
    keys_out[s] = (uint32_t)(((mnew * sc->T) + IN.tid[P]) * 2 + ((t != m) ? 1 : 0));
}

// Child row r of slot (P, t) as k_mu builds it: parent row r below t and row r + 1
// from t on, parent row m standing for the measurement-noise row (-mu_t, gamma).
// Returns the row's p value; sets *horth when the parent row is H-orthogonal.
template <int D>
__device__ inline double dev_child_row(const double* A, const double* p, const double* H, const double* mut,
                                       double gamma, int m, int t, int r, double* row, bool* horth) {
    const int src = (t < m && r >= t) ? r + 1 : r;
    *horth = false;
    if (src == m) {
#pragma unroll
        for (int kk = 0; kk < D; kk++) row[kk] = 0.0 - mut[kk];
        return gamma;
    }
    const double* Ar = A + src * D;
    const double h = dev_dot<D>(H, Ar);
    if (fabs(h) < kMuEps) {
#pragma unroll
        for (int kk = 0; kk < D; kk++) row[kk] = Ar[kk];
        *horth = true;
        return p[src];
    }
    const double rr = 1.0 / h;
#pragma unroll
    for (int kk = 0; kk < D; kk++) row[kk] = (Ar[kk] * rr) - mut[kk];
    return p[src] * fabs(h);
}

// Measurement update streamed row by row (KM >= m, compile-time bound).  The same
// operations in the same order as k_mu: each child row is built, enters the moment
// sums, is L1-normalised and written once; only the normalised first coordinates
// are kept, for the coalignment filter.  When the filter finds a possible merge (a
// rare case) the rows are rebuilt in memory and coaligned there, as k_mu does.
template <int D, int KM>
__device__ __forceinline__ void k_mu_r(MG<256, 3> mg, TermPool IN, int np, const int* __restrict__ slot_base, TermPool OUT,
                                             int nslots, const StepConst* __restrict__ sc,
                                             const uint32_t* __restrict__ gk_prev, const cplx* __restrict__ gv_prev,
                                             double* __restrict__ gy, uint32_t* __restrict__ keys_out,
                                             uint32_t* __restrict__ chain_key, int* __restrict__ err) {
    const int s = mg.bx * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    int lo = 0, hi = np - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (slot_base[mid] <= s) lo = mid;
        else hi = mid - 1;
    }
    const int P = lo;
    const int MM = IN.MM;
    const int m = IN.m[P];
    if (m > KM) {
        atomicOr(err, 2);
        return;
    }
    const int ts = s - slot_base[P];
    const int t = (ts == 0) ? m : ts - 1;
    const double* A = IN.A + (int64_t)P * MM * D;
    const double* p = IN.p + (int64_t)P * MM;
    const double* b = IN.b + (int64_t)P * D;
    const double* H = sc->H;
    uint32_t integ = 0, negs = 0;
    for (int l = 0; l < m; l++) {
        const double h = dev_dot<D>(H, A + l * D);
        if (!(fabs(h) < kMuEps)) {
            integ |= (1u << l);
            if (!(h > 0)) negs |= (1u << l);
        }
    }
    double* g_out = gy + (int64_t)s * (2 + 2 * D);
    if (t < m && !((integ >> t) & 1u)) {
        OUT.flags[s] = 0;
        keys_out[s] = kGatherKeyNone;
        chain_key[s] = 15u;
        for (int kk = 0; kk < 2 + 2 * D; kk++) g_out[kk] = 0.0;
        return;
    }
    chain_key[s] = (uint32_t)IN.tid[P];
    const double zeta = sc->msmt - dev_dot<D>(H, b);
    double mut[D];
    double rho_t;
    if (t == m) {
#pragma unroll
        for (int kk = 0; kk < D; kk++) mut[kk] = 0.0;
        rho_t = sc->gamma;
    } else {
        const double h = dev_dot<D>(H, A + t * D);
        const double r = 1.0 / h;
#pragma unroll
        for (int kk = 0; kk < D; kk++) mut[kk] = A[t * D + kk] * r;
        rho_t = p[t] * fabs(h);
    }
    double cb[D];
#pragma unroll
    for (int kk = 0; kk < D; kk++) cb[kk] = b[kk] + zeta * mut[kk];
    const int phc = IN.phc[P];
    OUT.c[s] = zeta;
    OUT.dv[s] = rho_t;
    OUT.phc[s] = phc;
    OUT.pbc[s] = m;
    OUT.z[s] = t;
    OUT.gp_off[s] = IN.gp_off[P];
    OUT.rk_off[s] = IN.rk_off[P];
    OUT.gp_n[s] = IN.gp_n[P];
    const uint32_t enc = negs;
    uint32_t lhp = enc;
    if (phc < m) lhp &= (1u << phc) - 1u;
    uint32_t sg = enc;
    const uint32_t rev_m = (m >= 32) ? 0xffffffffu : ((1u << m) - 1u);
    if (sg & (1u << (m - 1))) sg ^= rev_m;
    OUT.lhp[s] = (int)lhp;
    OUT.bt_off[s] = IN.bt_off[P];
    OUT.bt_n[s] = IN.bt_n[P];
    OUT.bt_sgn[s] = sc->skip ? 0 : (int)sg;
    OUT.tid[s] = IN.tid[P];
    double* ob = OUT.b + (int64_t)s * D;
#pragma unroll
    for (int kk = 0; kk < D; kk++) ob[kk] = cb[kk];
    double* oA = OUT.A + (int64_t)s * MM * D;
    double* op = OUT.p + (int64_t)s * MM;
    double* oq = OUT.q + (int64_t)s * MM;
    const bool normalise = !sc->skip && !sc->first;
    // rows: moment contributions (eval_g_yei, cauchy_term.hpp:312-401), then
    // q <- p and the L1 normalisation (normalize_hps, cauchy_term.hpp:458-474)
    double ty[D];
#pragma unroll
    for (int kk = 0; kk < D; kk++) ty[kk] = 0.0;
    double ygi = 0.0;
    uint32_t neg = 0, hf = 0;
    double x0[KM];
#pragma unroll
    for (int r = 0; r < KM; r++) {
        x0[r] = 0.0;
        if (r < m) {
            double row[D];
            bool ho;
            double pr = dev_child_row<D>(A, p, H, mut, sc->gamma, m, t, r, row, &ho);
            if (ho) hf |= (1u << r);
            double dr = 0.0;
#pragma unroll
            for (int kk = 0; kk < D; kk++) dr += row[kk] * sc->root[kk];
            const double sgn = dr > 0 ? 1 : -1;
            if (sgn < 0) neg |= (1u << r);
            const double scl = pr * sgn;
#pragma unroll
            for (int kk = 0; kk < D; kk++) ty[kk] += scl * row[kk];
            if (!ho) ygi += pr * sgn;
            if (normalise) {
                double n1 = 0;
#pragma unroll
                for (int kk = 0; kk < D; kk++) n1 += fabs(row[kk]);
                oq[r] = pr;
                pr *= n1;
#pragma unroll
                for (int kk = 0; kk < D; kk++) row[kk] /= n1;
                x0[r] = fabs(row[0]);
            }
#pragma unroll
            for (int kk = 0; kk < D; kk++) oA[r * D + kk] = row[kk];
            op[r] = pr;
        }
    }
    {
        cplx gpl, gmn;
        if (sc->first) {
            gpl = cmake(1, 0);
            gmn = cmake(1, 0);
        } else {
            const int top = 1 << (phc - 1), rev = (phc >= 32) ? -1 : ((1 << phc) - 1);
            int lp = 0, lmv = 0, kq = 0;
            for (int l = 0; l < m; l++) {
                if (kq < phc) {
                    if (kq == t) {
                        lmv |= (1 << kq);
                        kq++;
                        if (kq == phc) continue;
                    }
                    if ((neg >> l) & 1u) {
                        lp |= (1 << kq);
                        lmv |= (1 << kq);
                    }
                    kq++;
                }
            }
            const int64_t go = IN.gp_off[P];
            gpl = lookup_num(lp ^ (int)lhp, top, rev, gk_prev + go, gv_prev + go, IN.gp_n[P]);
            gmn = lookup_num(lmv ^ (int)lhp, top, rev, gk_prev + go, gv_prev + go, IN.gp_n[P]);
        }
        const cplx g = cscale(csub(cdiv(gpl, cmake(ygi + rho_t, zeta)), cdiv(gmn, cmake(ygi - rho_t, zeta))),
                              kInvTwoPi);
        g_out[0] = g.re;
        g_out[1] = g.im;
#pragma unroll
        for (int kk = 0; kk < D; kk++) {
            g_out[2 + 2 * kk] = -ty[kk];
            g_out[3 + 2 * kk] = cb[kk];
        }
    }
    int mnew = m;
    uint8_t fl = kValid | ((t != m) ? kIsNew : 0);
    if (normalise && t != m) {
        // Two normalised rows coalign (mu_coalign, cauchy_term.hpp:533-745) only if
        // |x_j - x_k| < eps or |x_j + x_k| < eps on every coordinate, so only if
        // ||x_j0| - |x_k0|| < eps on the first (in floating point the latter equals
        // one of the former exactly).  Without such a pair nothing merges.
        bool cand = false;
#pragma unroll
        for (int j = 0; j < KM - 1; j++)
#pragma unroll
            for (int kx = j + 1; kx < KM; kx++)
                if (kx < m && fabs(x0[j] - x0[kx]) < kCoalignEps) cand = true;
        uint8_t* cm = OUT.cmap + (int64_t)s * MM;
        int8_t* cs = OUT.csmap + (int64_t)s * MM;
        if (cand) {
            // possible merge: rebuild the rows as built before normalisation and run
            // the in-memory coalignment
            for (int r = 0; r < m; r++) {
                double row[D];
                bool ho;
                op[r] = dev_child_row<D>(A, p, H, mut, sc->gamma, m, t, r, row, &ho);
#pragma unroll
                for (int kk = 0; kk < D; kk++) oA[r * D + kk] = row[kk];
            }
            mnew = dev_mu_coalign<D>(oA, op, oq, cm, cs, &hf, m, err);
            if (mnew < m) fl |= kHasCmap;
        } else {
            // a new child's maps are the identity
            for (int r = 0; r < m; r++) {
                cm[r] = (uint8_t)r;
                cs[r] = 1;
            }
        }
    }
    OUT.hflag[s] = hf;
    OUT.flags[s] = fl;
    OUT.m[s] = mnew;
    keys_out[s] = (uint32_t)(((mnew * sc->T) + IN.tid[P]) * 2 + ((t != m) ? 1 : 0));
}

// ------------------------------------------------------------------ moments
// Moment sums in the reference's exact order.  Reference thread t owns parents
// i = t, t+T, ... of every shape bucket (cauchy_estimator.hpp:1575-1604) and adds,
// per parent, the parent's own term before its children (cache_moments,
// cauchy_estimator.hpp:307-338 and 1436-1468).  `chain` lists the output slots in
// exactly that order (slots are parent-first), so chain t is one contiguous range
// [chain_off[t], chain_off[t+1]) of chain positions.
//
// Two kernels: k_mom_terms writes every summand of every chain position (fully
// parallel), k_mom_sum adds them in order, one lane per double of a component.
// Summands are grouped in blocks of kSumCW doubles: X[(yb * npos + pos) * kSumCW + c]
// holds double yb * kSumCW + c (component / 2, real or imaginary part) of position
// pos.  Padding and slots without a term hold +0.0: an in-order sum that starts at
// +0.0 never becomes -0.0, so adding +0.0 leaves it unchanged.
constexpr int kSumCW = 8;     // doubles per summand block (lanes of the adding wave)
constexpr int kSumTP = 288;   // positions per LDS tile (tile = 12 doubles per loader thread)
constexpr int kSumDepth = 3;  // tiles in flight in loader registers

template <int D>
__host__ __device__ constexpr int mom_blocks() {
    return (2 * (1 + D + D * D) + kSumCW - 1) / kSumCW;
}

// Moments after term reduction (debug mode), cauchy_term.hpp:403-433 (eval_g_yei_after_ftr)
// over every parent term: the sign vector of its hyperplanes at the root point (bit l set
// where A_l . root <= 0), y_j = -(sum_l (p_l s_l) A_lj) + i b_j, and g its own g-table's value
// at that sign vector.  Writes (g, y) per term in the layout k_mom_terms reads.
template <int D>
__device__ __forceinline__ void k_after_ftr_gy(MG<1024, 1> mg, const TermPool P, int n,
                                               const StepConst* __restrict__ sc, const uint32_t* __restrict__ gk,
                                               const cplx* __restrict__ gv, double* __restrict__ gy) {
    const int t = mg.bx * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int m = P.m[t];
    const int MM = P.MM;
    const double* A = P.A + (int64_t)t * MM * D;
    const double* p = P.p + (int64_t)t * MM;
    const double* b = P.b + (int64_t)t * D;
    double tmp[D];
#pragma unroll
    for (int j = 0; j < D; j++) tmp[j] = 0.0;
    int enc = 0;
    for (int l = 0; l < m; l++) {
        double dp = 0.0;
#pragma unroll
        for (int k = 0; k < D; k++) dp += A[l * D + k] * sc->root[k];
        const double s = dp > 0 ? 1.0 : -1.0;
        const double w = p[l] * s;
#pragma unroll
        for (int j = 0; j < D; j++) tmp[j] += w * A[l * D + j];
        if (s < 0) enc |= 1 << l;
    }
    const int top = 1 << (m - 1);
    const int rev = (m >= 32) ? -1 : ((1 << m) - 1);
    const int64_t go = P.gp_off[t];
    const cplx g = lookup_num(enc, top, rev, gk + go, gv + go, P.gp_n[t]);
    double* rec = gy + (int64_t)t * (2 + 2 * D);
    rec[0] = g.re;
    rec[1] = g.im;
#pragma unroll
    for (int j = 0; j < D; j++) {
        rec[2 + 2 * j] = -tmp[j];
        rec[3 + 2 * j] = b[j];
    }
}

// Products with the reference's operations: g, g*y_j and the negated (g*y_j)*y_k
// (the reference subtracts the latter; a - b == a + (-b) exactly).
template <int D>
__device__ __forceinline__ void k_mom_terms(MG<1024, 1> mg, const double* __restrict__ gy, const uint8_t* __restrict__ flags,
                            const uint32_t* __restrict__ chain, const int* __restrict__ chain_off, int T, int nslots,
                            double* __restrict__ X) {
    constexpr int NS = 2 + 2 * D;
    constexpr int NC = 1 + D + D * D;
    constexpr int NCP = mom_blocks<D>() * kSumCW / 2;  // padded components per position
    const int64_t i = (int64_t)mg.bx * blockDim.x + threadIdx.x;
    const int pos = (int)(i / NCP);
    const int cc = (int)(i - (int64_t)pos * NCP);
    if (pos >= nslots) return;
    const int npos = chain_off[T];
    if (pos >= npos) return;
    const uint32_t s = chain[pos];
    double2 out = make_double2(0.0, 0.0);
    if (cc < NC && (flags[s] & kValid)) {
        const double* rec = gy + (int64_t)s * NS;
        const cplx g = cmake(rec[0], rec[1]);
        cplx x;
        if (cc == 0) {
            x = g;
        } else if (cc <= D) {
            x = cmul(g, cmake(rec[2 + 2 * (cc - 1)], rec[3 + 2 * (cc - 1)]));
        } else {
            const int j = (cc - 1 - D) / D, kk = (cc - 1 - D) % D;
            const cplx gyj = cmul(g, cmake(rec[2 + 2 * j], rec[3 + 2 * j]));
            const cplx q = cmul(gyj, cmake(rec[2 + 2 * kk], rec[3 + 2 * kk]));
            x = cmake(-q.re, -q.im);
        }
        out = make_double2(x.re, x.im);
    }
    const int c2 = 2 * cc;
    const int yb = c2 / kSumCW, c = c2 % kSumCW;
    *reinterpret_cast<double2*>(X + ((int64_t)yb * npos + pos) * kSumCW + c) = out;
}

// In-order sums.  Grid (T, mom_blocks), 256 threads: wave 0 adds, lane c < kSumCW
// owning double blockIdx.y * kSumCW + c; waves 1-3 stream the chain's summand block
// through a two-tile LDS ring, holding kSumDepth tiles in registers so that the
// loads of a tile are in flight while the tiles before it are added.  part[t][c2]
// receives the sum of chain t.
template <int D>
// part[T * NC2] also receives the measurement update's error bits (*mu_err), so one copy
// brings both back.
__device__ __forceinline__ void k_mom_sum(MG<256, 1> mg, const double* __restrict__ X, const int* __restrict__ chain_off,
                                                 int T, double* __restrict__ part, const int* __restrict__ mu_err) {
    constexpr int NC2 = 2 * (1 + D + D * D);
    constexpr int TE = kSumTP * kSumCW;          // doubles per tile
    constexpr int NL = 192;                      // loader threads
    constexpr int PER = TE / NL;                 // doubles per loader thread per tile
    static_assert(PER * NL == TE, "a tile is an exact multiple of the loader threads");
    __shared__ double ring[2][TE];
    const int t = mg.bx;
    const int yb = mg.by;
    const int tid = threadIdx.x;
    if (t == 0 && yb == 0 && tid == 0) part[T * NC2] = (double)mu_err[0];
    const int npos = chain_off[T];
    const int beg = chain_off[t], end = chain_off[t + 1];
    const int ntiles = (end - beg + kSumTP - 1) / kSumTP;
    const double* Xb = X + ((int64_t)yb * npos + beg) * kSumCW;  // this chain's block, tile 0 first
    const int64_t lim = (int64_t)(end - beg) * kSumCW;           // doubles of the chain
    const bool loader = __builtin_amdgcn_readfirstlane(tid >> 6) != 0;  // wave 0 adds
    const int l = tid - 64;
    double rg[kSumDepth][PER];

    // Loads are unconditional (address clamped into the chain) so that no load is
    // skipped and the wait for the oldest tile stays exact; values past the chain's
    // end are replaced by +0.0 when the tile is stored to LDS, not at load time,
    // so the loads stay in flight until then.
    auto gload = [&](double (&r)[PER], int tile) {
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int64_t o = (int64_t)tile * TE + l + NL * q;
            r[q] = Xb[o < lim ? o : lim - 1];
        }
    };
    auto lstore = [&](const double (&r)[PER], int tile, double* dst) {
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int64_t o = (int64_t)tile * TE + l + NL * q;
            dst[l + NL * q] = o < lim ? r[q] : 0.0;
        }
    };
    // LDS-only barrier: global loads stay in flight across it
    auto lds_barrier = [&]() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    };

    // The two roles run separate loops; each performs exactly nsteps + 1 barriers.
    // The step count is padded to a multiple of kSumDepth so that every unrolled
    // step is unconditional (the compiler then keeps the later tiles' loads in
    // flight); padded tiles hold +0.0 only.
    const int nsteps = (ntiles + kSumDepth - 1) / kSumDepth * kSumDepth;
    double acc = 0.0;
    const int lane = tid & 63;
    if (loader) {
        if (nsteps > 0) {
#pragma unroll
            for (int j = 0; j < kSumDepth; j++) gload(rg[j], j);
            lstore(rg[0], 0, ring[0]);
            gload(rg[0], kSumDepth);
        }
        lds_barrier();
        for (int k0 = 0; k0 < nsteps; k0 += kSumDepth) {
#pragma unroll
            for (int u = 0; u < kSumDepth; u++) {
                const int k = k0 + u;
                // u is a compile-time index once the loop is unrolled
                lstore(rg[(u + 1) % kSumDepth], k + 1, ring[(k + 1) & 1]);
                gload(rg[(u + 1) % kSumDepth], k + 1 + kSumDepth);
                lds_barrier();
            }
        }
    } else {
        lds_barrier();
        for (int k = 0; k < nsteps; k++) {
            if (lane < kSumCW) {
                const double* src = ring[k & 1] + lane;
                double cur[16], nxt[16];
#pragma unroll
                for (int q = 0; q < 16; q++) cur[q] = src[q * kSumCW];
                for (int r0 = 0; r0 < kSumTP; r0 += 16) {
                    if (r0 + 16 < kSumTP) {
#pragma unroll
                        for (int q = 0; q < 16; q++) nxt[q] = src[(r0 + 16 + q) * kSumCW];
                    }
#pragma unroll
                    for (int q = 0; q < 16; q++) acc += cur[q];
#pragma unroll
                    for (int q = 0; q < 16; q++) cur[q] = nxt[q];
                }
            }
            lds_barrier();
        }
    }
    const int c2 = yb * kSumCW + lane;
    if (!loader && lane < kSumCW && c2 < NC2) part[(int64_t)t * NC2 + c2] = acc;
}

// Histogram of small integer values privatised in LDS (nbins <= 64).  mode 0: valid
// slots by shape; mode 1: surviving groups by shape; mode 2: chain keys.
__device__ __forceinline__ void k_hist(MG<1024, 1> mg, int n, int nbins, int mode, const uint8_t* __restrict__ flags, const int* __restrict__ vals,
                       const int* __restrict__ surv, const uint32_t* __restrict__ ukeys, int* __restrict__ hist) {
    __shared__ int h[64];
    for (int b = threadIdx.x; b < nbins; b += blockDim.x) h[b] = 0;
    __syncthreads();
    for (int i = mg.bx * blockDim.x + threadIdx.x; i < n; i += mg.gx * blockDim.x) {
        int v = -1;
        if (mode == 0) {
            if (flags[i] & kValid) v = vals[i];
        } else if (mode == 1) {
            if (surv[i] >= 0) v = vals[i];
        } else {
            v = (int)ukeys[i];
        }
        if (v >= 0 && v < nbins) atomicAdd(&h[v], 1);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nbins; b += blockDim.x)
        if (h[b]) atomicAdd(&hist[b], h[b]);
}

// ------------------------------------------------------------------ gather
__device__ __forceinline__ void k_iota(MG<1024, 1> mg, uint32_t* __restrict__ v, int n) {
    const int i = mg.bx * blockDim.x + threadIdx.x;
    if (i < n) v[i] = (uint32_t)i;
}

// ------------------------------------------------------------------ fast term reduction
// All shape buckets are reduced together.  Global position p in [0, N) indexes the
// gathered order (bucket m occupies [boff[m], boff[m+1]), slot perm[p]); reduction
// decisions use bucket-local indices, as the reference reduces each bucket on its own
// (fast_term_reduction_and_create_gtables, cauchy_estimator.hpp:1001-1137).
__device__ inline int bucket_of(const int* __restrict__ boff, int nb, int p) {
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (boff[mid] <= p) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Chunk of local index l in a bucket of n terms: buckets above 10000 terms are reduced
// in 8 chunks of n/8 (the last takes the rest), term_reduction.hpp:651-681.
__device__ inline int ftr_chunk_of(int l, int n) {
    if (n <= 10000) return 0;
    const int c = l / (n / 8);
    return c > 7 ? 7 : c;
}

__device__ __forceinline__ void k_ftr_keys(MG<1024, 1> mg, const TermPool OUT, const uint32_t* __restrict__ perm, int n, int d, int axis,
                           uint64_t* __restrict__ keys, uint32_t* __restrict__ idx) {
    const int p = mg.bx * blockDim.x + threadIdx.x;
    if (p >= n) return;
    keys[p] = f64_radix_key(OUT.b[(int64_t)perm[p] * d + axis]);
    idx[p] = (uint32_t)p;
}

__device__ __forceinline__ void k_ftr_vals(MG<1024, 1> mg, const TermPool OUT, const uint32_t* __restrict__ perm, int n, int d, int axis,
                           const uint32_t* __restrict__ sidx, double* __restrict__ vals) {
    const int r = mg.bx * blockDim.x + threadIdx.x;
    if (r >= n) return;
    vals[r] = OUT.b[(int64_t)perm[sidx[r]] * d + axis];
}

// Bucket of each position of the value-sorted order (second, stable sort key).
__device__ __forceinline__ void k_ftr_bkeys(MG<1024, 1> mg, int n, const int* __restrict__ boff, int nb, const uint32_t* __restrict__ sidx,
                            uint32_t* __restrict__ bkey) {
    const int r = mg.bx * blockDim.x + threadIdx.x;
    if (r >= n) return;
    bkey[r] = (uint32_t)bucket_of(boff, nb, (int)sidx[r]);
}

// Candidate pairs (i, j > i) of the greedy reduction, term_reduction.hpp:159-276:
// b_j inside (b_i - ep, b_i + ep] on every axis (the window ftr_binary_search
// selects, term_reduction.hpp:89-157, which must hold two points), then the p check
// and the element-wise check over the first m flattened entries of A.  One thread
// per sorted position; the window is found by scanning the sorted neighbours.
// pass 0 counts, pass 1 writes the out-edges (global positions) of term i.
template <int D>
__device__ __forceinline__ void k_ftr_edges(MG<1024, 1> mg, const TermPool OUT, const uint32_t* __restrict__ perm, int n,
                            const int* __restrict__ boff, int nb, const StepConst* __restrict__ sc,
                            const double* __restrict__ vals, const uint32_t* __restrict__ sidx,
                            int* __restrict__ cnt, const int* __restrict__ eoff, int* __restrict__ edges, int pass) {
    const int sp = mg.bx * blockDim.x + threadIdx.x;
    if (sp >= n) return;
    const double ep = kReductionEps;
    const int bk = bucket_of(boff, nb, sp);
    const int beg = boff[bk], end = boff[bk + 1];
    const int m = bk;  // bucket index is the shape
    const int MM = OUT.MM;
    const int ig = (int)sidx[sp];
    const int il = ig - beg;
    const int64_t si = perm[ig];
    const double* bi = OUT.b + si * D;
    const double* pi = OUT.p + si * MM;
    const double* Ai = OUT.A + si * MM * D;
    const double v = vals[sp];
    const double lo = v - ep, hi = v + ep;
    // r0 = first sorted position in the bucket with value > lo; r1 = first with value > hi
    int r0 = sp;
    if (vals[r0] > lo) {
        while (r0 > beg && vals[r0 - 1] > lo) r0--;
    } else {
        while (r0 < end && vals[r0] <= lo) r0++;
    }
    int r1 = sp;
    if (vals[r1] <= hi) {
        while (r1 < end && vals[r1] <= hi) r1++;
    } else {
        while (r1 > beg && vals[r1 - 1] > hi) r1--;
    }
    int c = 0;
    int w = pass ? eoff[ig] : 0;
    if (r1 - r0 >= 2) {
        for (int r = r0; r < r1; r++) {
            const int jg = (int)sidx[r];
            if (jg - beg <= il) continue;
            const int64_t sj = perm[jg];
            const double* bj = OUT.b + sj * D;
            bool ok = true;
            for (int a = 1; a < D && ok; a++) {
                const int ax = sc->order[a];
                const double l2 = bi[ax] - ep, h2 = bi[ax] + ep;
                const double x = bj[ax];
                if (!(x > l2 && x <= h2)) ok = false;
            }
            if (!ok) continue;
            const double* pj = OUT.p + sj * MM;
            for (int k = 0; k < m; k++)
                if (fabs(pj[k] - pi[k]) > ep) {
                    ok = false;
                    break;
                }
            if (!ok) continue;
            const double* Aj = OUT.A + sj * MM * D;
            for (int k = 0; k < m; k++) {
                const double x = Ai[k], y = Aj[k];
                if (!((fabs(x - y) < ep) || (fabs(x + y) < ep))) {
                    ok = false;
                    break;
                }
            }
            if (!ok) continue;
            if (pass) edges[w++] = jg;
            c++;
        }
    }
    if (!pass) cnt[ig] = c;
}

// Pair-parallel candidate search.  k_ftr_gather copies what the predicate reads
// into the value-sorted order: b (d), p (first m) and the first m entries of the
// flattened A, row stride MM.
template <int D>
__device__ __forceinline__ void k_ftr_gather(MG<1024, 1> mg, const TermPool OUT, const uint32_t* __restrict__ perm, int n,
                             const int* __restrict__ boff, int nb, const uint32_t* __restrict__ sidx,
                             double* __restrict__ gb, double* __restrict__ gp, double* __restrict__ gA) {
    const int r = mg.bx * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int m = bucket_of(boff, nb, r);
    const int MM = OUT.MM;
    const int64_t s = perm[sidx[r]];
#pragma unroll
    for (int a = 0; a < D; a++) gb[(int64_t)r * D + a] = OUT.b[s * D + a];
    for (int q = 0; q < m; q++) {
        gp[(int64_t)r * MM + q] = OUT.p[s * MM + q];
        gA[(int64_t)r * MM + q] = OUT.A[s * MM * D + q];
    }
}

// Window of sorted position sp on the first search axis: positions r0 <= r < r1 of
// its bucket with value in (v - ep, v + ep] (a window must hold two points,
// ftr_binary_search, term_reduction.hpp:89-157).  wcnt = r1 - r0, or 0.
// Also clears each position's edge count for the edge passes that follow.
__device__ __forceinline__ void k_ftr_win(MG<1024, 1> mg, int n, const int* __restrict__ boff, int nb, const double* __restrict__ vals,
                          int* __restrict__ wlo, int64_t* __restrict__ wcnt, int* __restrict__ ecnt) {
    const int sp = mg.bx * blockDim.x + threadIdx.x;
    if (sp >= n) return;
    ecnt[sp] = 0;
    const int bk = bucket_of(boff, nb, sp);
    const int beg = boff[bk], end = boff[bk + 1];
    const double v = vals[sp];
    const double lo = v - kReductionEps, hi = v + kReductionEps;
    int a = beg, b = end;  // first position with value > lo
    while (a < b) {
        const int mid = (a + b) >> 1;
        if (vals[mid] > lo) b = mid;
        else a = mid + 1;
    }
    const int r0 = a;
    a = beg;
    b = end;  // first position with value > hi
    while (a < b) {
        const int mid = (a + b) >> 1;
        if (vals[mid] > hi) b = mid;
        else a = mid + 1;
    }
    const int r1 = a;
    wlo[sp] = r0;
    wcnt[sp] = (r1 - r0 >= 2) ? (int64_t)(r1 - r0) : 0;
}

// One thread per (position, window slot) pair: the rest of the greedy predicate,
// term_reduction.hpp:159-276 (j > i, every other axis inside (b_i - ep, b_i + ep],
// |p_j - p_i| <= ep on the first m, and the element-wise check on the first m
// entries of the flattened A).  pass 0 flags the pair and counts term i's out-edges;
// pass 1 writes the flagged pairs into term i's list (edge order within a list does
// not matter to the resolution that follows).
template <int D>
__device__ __forceinline__ void k_ftr_pairs(MG<1024, 1> mg, int64_t npairs, int n, const int* __restrict__ boff, int nb,
                            const StepConst* __restrict__ sc, const int64_t* __restrict__ woff,
                            const int* __restrict__ wlo, const uint32_t* __restrict__ sidx, int MM,
                            const double* __restrict__ gb, const double* __restrict__ gp,
                            const double* __restrict__ gA, uint8_t* __restrict__ flag, int* __restrict__ ecnt,
                            const int* __restrict__ eoff, int* __restrict__ cursor, int* __restrict__ edges,
                            int pass) {
    const int64_t q = (int64_t)mg.bx * blockDim.x + threadIdx.x;
    if (q >= npairs) return;
    // owning position: largest sp with woff[sp] <= q
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (woff[mid] <= q) lo = mid;
        else hi = mid - 1;
    }
    const int sp = lo;
    const int r = wlo[sp] + (int)(q - woff[sp]);
    const int ig = (int)sidx[sp];
    const int jg = (int)sidx[r];
    if (pass == 1) {
        if (flag[q]) {
            const int slot = atomicAdd(&cursor[ig], 1);
            edges[eoff[ig] + slot] = jg;
        }
        return;
    }
    bool ok = jg > ig;
    if (ok) {
        const double ep = kReductionEps;
        const double* bi = gb + (int64_t)sp * D;
        const double* bj = gb + (int64_t)r * D;
        for (int a = 1; a < D && ok; a++) {
            const int ax = sc->order[a];
            const double l2 = bi[ax] - ep, h2 = bi[ax] + ep;
            const double x = bj[ax];
            if (!(x > l2 && x <= h2)) ok = false;
        }
        const int m = bucket_of(boff, nb, sp);
        const double* pi = gp + (int64_t)sp * MM;
        const double* pj = gp + (int64_t)r * MM;
        for (int kk = 0; kk < m && ok; kk++)
            if (fabs(pj[kk] - pi[kk]) > ep) ok = false;
        const double* Ai = gA + (int64_t)sp * MM;
        const double* Aj = gA + (int64_t)r * MM;
        for (int kk = 0; kk < m && ok; kk++) {
            const double x = Ai[kk], y = Aj[kk];
            if (!((fabs(x - y) < ep) || (fabs(x + y) < ep))) ok = false;
        }
    }
    flag[q] = ok ? 1 : 0;
    if (ok) atomicAdd(&ecnt[ig], 1);
}

// In-degree over edges whose ends share a chunk (pass 0), or the in-edge lists (pass 1).
__device__ __forceinline__ void k_ftr_inedges(MG<1024, 1> mg, int n, const int* __restrict__ boff, int nb, const int* __restrict__ eoff,
                              const int* __restrict__ ecnt, const int* __restrict__ edges, int* __restrict__ indeg,
                              const int* __restrict__ inoff, int* __restrict__ cursor, int* __restrict__ inedges,
                              int pass) {
    const int ig = mg.bx * blockDim.x + threadIdx.x;
    if (ig >= n) return;
    const int bk = bucket_of(boff, nb, ig);
    const int beg = boff[bk], nbk = boff[bk + 1] - beg;
    const int ci = ftr_chunk_of(ig - beg, nbk);
    // Edges in batches of kInBatch: a batch's loads, then its atomics, are all in flight
    // before the first result is used.  The order of a term's in-edges is immaterial (the
    // rounds only ask whether any of them is a root, or all are claimed).
    constexpr int kInBatch = 4;
    const int e0 = eoff[ig], e1 = e0 + ecnt[ig];
    for (int e = e0; e < e1; e += kInBatch) {
        int jg[kInBatch];
#pragma unroll
        for (int u = 0; u < kInBatch; u++) jg[u] = e + u < e1 ? edges[e + u] : -1;
#pragma unroll
        for (int u = 0; u < kInBatch; u++)
            if (jg[u] >= 0 && ftr_chunk_of(jg[u] - beg, nbk) != ci) jg[u] = -1;
        if (pass == 0) {
#pragma unroll
            for (int u = 0; u < kInBatch; u++)
                if (jg[u] >= 0) atomicAdd(&indeg[jg[u]], 1);
        } else {
            int slot[kInBatch], base[kInBatch];
#pragma unroll
            for (int u = 0; u < kInBatch; u++)
                if (jg[u] >= 0) {
                    slot[u] = atomicAdd(&cursor[jg[u]], 1);
                    base[u] = inoff[jg[u]];
                }
#pragma unroll
            for (int u = 0; u < kInBatch; u++)
                if (jg[u] >= 0) inedges[base[u] + slot[u]] = ig;
        }
    }
}

enum { kUnknown = 0, kRoot = 1, kClaimed = 2 };

// One round of root resolution: a term is a root of its chunk's greedy pass when
// no same-chunk root before it claims it.
// Up to kFillRegions byte fills in one launch: region blockIdx.y gets `bytes` copies of
// `val`, by 32-bit words where the pointer allows, then the tail bytes.
constexpr int kFillRegions = 6;
struct FillArgs {
    uint8_t* ptr[kFillRegions];
    int64_t bytes[kFillRegions];
    uint32_t val[kFillRegions];
};
__device__ __forceinline__ void k_fill_multi(MG<1024, 1> mg, FillArgs f) {
    const int k = mg.by;
    uint8_t* p = f.ptr[k];
    const int64_t nb = f.bytes[k];
    const uint32_t v = f.val[k] & 0xffu;
    const int64_t stride = (int64_t)mg.gx * blockDim.x;
    const int64_t t = (int64_t)mg.bx * blockDim.x + threadIdx.x;
    int64_t words = 0;
    if (((uintptr_t)p & 3) == 0) {
        words = nb / 4;
        uint32_t* w = (uint32_t*)p;
        const uint32_t v4 = v * 0x01010101u;
        for (int64_t i = t; i < words; i += stride) w[i] = v4;
    }
    for (int64_t i = 4 * words + t; i < nb; i += stride) p[i] = (uint8_t)v;
}

// Round r counts its unresolved terms in slots[r % kRoundSlots] and clears the next
// round's slot (no other thread of round r touches it, and round r + 1 starts after round r
// ends), so the rounds need no fill between them.
constexpr int kRoundSlots = 64;
__device__ __forceinline__ void k_ftr_round(MG<1024, 1> mg, int n, const int* __restrict__ inoff, const int* __restrict__ indeg,
                            const int* __restrict__ inedges, const uint8_t* __restrict__ st_in,
                            uint8_t* __restrict__ st_out, int* __restrict__ slots, int r) {
    const int j = mg.bx * blockDim.x + threadIdx.x;
    if (j == 0) slots[(r + 1) % kRoundSlots] = 0;
    int* unresolved = slots + r % kRoundSlots;
    if (j >= n) return;
    uint8_t s = st_in[j];
    if (s == kUnknown) {
        // in-edges in batches of kRoundBatch: a batch's index loads, then its state loads,
        // are all in flight before the first is tested; padding counts as claimed
        constexpr int kRoundBatch = 4;
        bool any_root = false, all_claimed = true;
        const int e0 = inoff[j], e1 = e0 + indeg[j];
        for (int e = e0; e < e1 && !any_root; e += kRoundBatch) {
            int src[kRoundBatch];
#pragma unroll
            for (int u = 0; u < kRoundBatch; u++) src[u] = e + u < e1 ? inedges[e + u] : -1;
            uint8_t si[kRoundBatch];
#pragma unroll
            for (int u = 0; u < kRoundBatch; u++) si[u] = src[u] >= 0 ? st_in[src[u]] : (uint8_t)kClaimed;
#pragma unroll
            for (int u = 0; u < kRoundBatch; u++) {
                any_root |= si[u] == kRoot;
                all_claimed &= si[u] == kClaimed;
            }
        }
        s = any_root ? kClaimed : (all_claimed ? kRoot : kUnknown);
        if (s == kUnknown) atomicAdd(unresolved, 1);
    }
    st_out[j] = s;
}

// Flag arrays F_c (c < 8) over global positions, holding bucket-local indices:
// F_c[j] = smallest chunk-c root claiming j, else j itself.  Buckets reduced in one
// chunk use F_0 only.
__device__ __forceinline__ void k_ftr_finit(MG<1024, 1> mg, int* __restrict__ F, int n, const int* __restrict__ boff, int nb) {
    const int p = mg.bx * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int bk = bucket_of(boff, nb, p);
    const int l = p - boff[bk];
    const int nc = (boff[bk + 1] - boff[bk] > 10000) ? 8 : 1;
    for (int c = 0; c < nc; c++) F[(int64_t)c * n + p] = l;
}

__device__ __forceinline__ void k_ftr_claim(MG<1024, 1> mg, int n, const int* __restrict__ boff, int nb, const int* __restrict__ eoff,
                            const int* __restrict__ ecnt, const int* __restrict__ edges,
                            const uint8_t* __restrict__ st, int* __restrict__ F) {
    const int ig = mg.bx * blockDim.x + threadIdx.x;
    if (ig >= n || st[ig] != kRoot) return;
    const int bk = bucket_of(boff, nb, ig);
    const int beg = boff[bk], nbk = boff[bk + 1] - beg;
    const int il = ig - beg;
    int* Fc = F + (int64_t)ftr_chunk_of(il, nbk) * n;
    // edges in batches of four: the batch's loads are in flight together, then its minima
    const int e0 = eoff[ig], e1 = e0 + ecnt[ig];
    for (int e = e0; e < e1; e += 4) {
        int jg[4];
#pragma unroll
        for (int u = 0; u < 4; u++) jg[u] = e + u < e1 ? edges[e + u] : -1;
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (jg[u] >= 0) atomicMin(&Fc[jg[u]], il);
    }
}

// One level of the flag-array merge tree of every 8-chunk bucket,
// term_reduction.hpp:592-649.  Every value of a flag array is a fixed point of it,
// so a level is data parallel.
__device__ __forceinline__ void k_ftr_merge(MG<1024, 1> mg, int n, const int* __restrict__ boff, int nb, int* __restrict__ F, int level) {
    const int p = mg.bx * blockDim.x + threadIdx.x;
    const int k = mg.by;
    if (p >= n) return;
    const int bk = bucket_of(boff, nb, p);
    const int beg = boff[bk], nbk = boff[bk + 1] - beg;
    if (nbk <= 10000) return;
    const int pairs = 8 >> (level + 1);
    if (k >= pairs) return;
    const int rt = k * (1 << (level + 1));
    const int cm = rt + (1 << level);
    const int l = p - beg;
    if (l < cm * (nbk / 8)) return;
    int* F1 = F + (int64_t)rt * n + beg;
    const int* F2 = F + (int64_t)cm * n + beg;
    const int f2 = F2[l];
    if (f2 != -1 && f2 != l && F1[l] == l) F1[l] = (F1[f2] == f2) ? f2 : F1[f2];
}

// Root flags; members are keyed by their root's global position, roots by n.
__device__ __forceinline__ void k_ftr_rootflags(MG<1024, 1> mg, int n, const int* __restrict__ boff, int nb, const int* __restrict__ F,
                                uint8_t* __restrict__ isroot, uint32_t* __restrict__ memkey,
                                uint32_t* __restrict__ memval) {
    const int p = mg.bx * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int beg = boff[bucket_of(boff, nb, p)];
    const int f = F[p];
    const bool r = (f == p - beg);
    isroot[p] = r;
    memkey[p] = r ? (uint32_t)n : (uint32_t)(beg + f);
    memval[p] = (uint32_t)p;
}

// The root count is read on the device (grids are sized by the term count n).
__device__ __forceinline__ void k_ftr_memcount(MG<1024, 1> mg, int n, const int* __restrict__ nroots, const uint32_t* __restrict__ skey,
                               int* __restrict__ cnt) {
    const int k = mg.bx * blockDim.x + threadIdx.x;
    if (k >= n - *nroots) return;
    atomicAdd(&cnt[skey[k]], 1);
}

__device__ __forceinline__ void k_ftr_memslots(MG<1024, 1> mg, int n, const int* __restrict__ nroots, const uint32_t* __restrict__ sval,
                               const uint32_t* __restrict__ perm, int* __restrict__ mem_slots) {
    const int k = mg.bx * blockDim.x + threadIdx.x;
    if (k >= n - *nroots) return;
    mem_slots[k] = (int)perm[sval[k]];
}

// First reduction group of every bucket: gbase[m] = number of roots before boff[m].
__device__ __forceinline__ void k_ftr_gbase(MG<1024, 1> mg, const int* __restrict__ d_nroots, const uint32_t* __restrict__ roots,
                            const int* __restrict__ boff, int nb, int* __restrict__ gbase) {
    const int m = threadIdx.x;
    if (m > nb) return;
    const int nroots = *d_nroots;
    const uint32_t x = (uint32_t)boff[m];
    int lo = 0, hi = nroots;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (roots[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    gbase[m] = lo;
}

// Reduction groups in the reference's order: buckets ascending, roots ascending.
__device__ __forceinline__ void k_ftr_groups(MG<1024, 1> mg, const int* __restrict__ nroots, const uint32_t* __restrict__ roots,
                             const uint32_t* __restrict__ perm, const int* __restrict__ boff, int nb,
                             const int* __restrict__ memcnt, const int* __restrict__ memoff,
                             int* __restrict__ grp_root, int* __restrict__ grp_moff, int* __restrict__ grp_mcnt,
                             int* __restrict__ grp_m) {
    const int g = mg.bx * blockDim.x + threadIdx.x;
    if (g >= *nroots) return;
    const uint32_t r = roots[g];
    grp_root[g] = (int)perm[r];
    grp_moff[g] = memoff[r];
    grp_mcnt[g] = memcnt[r];
    grp_m[g] = bucket_of(boff, nb, (int)r);
}

// ------------------------------------------------------------------ group B-tables
// Upper bound of each group's B-table size.
__device__ __forceinline__ void k_group_ub(MG<1024, 1> mg, const TermPool OUT, int ng, const int* __restrict__ grp_root, const int* __restrict__ grp_m,
                           int d, int* __restrict__ ub) {
    const int g = mg.bx * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    const int s = grp_root[g];
    const int m = grp_m[g];
    int u = OUT.bt_n[s];
    if ((OUT.flags[s] & kIsNew) && m <= d) u = max(u, 1 << (m - 1));
    ub[g] = max(u, 1);
}

// B-table of a group root: new children pair the parent's post-update cells
// across their defining hyperplane and, when coaligned, project through c_map
// (make_new_child_btable, cell_enumeration.hpp:206-371); parents keep their own
// post-update cells.  One wavefront per group, an LDS bitset of 2^m bits per
// wavefront, sorted output.
__device__ __forceinline__ void k_group_btab(MG<1024, 1> mg, const TermPool OUT, int g0, int ngb, int m, int d, const int* __restrict__ grp_root,
                             const uint32_t* __restrict__ Barena, const int64_t* __restrict__ b2_off,
                             const int64_t* __restrict__ ub, uint32_t* __restrict__ B2, int* __restrict__ b2_n,
                             int words_per_wave, int* __restrict__ err) {
    extern __shared__ uint32_t sh[];
    const int wv = threadIdx.x / kWave;
    const int lane = wave_lane();
    const int g = g0 + mg.bx * (blockDim.x / kWave) + wv;
    if (g >= g0 + ngb) return;
    uint32_t* bits = sh + wv * words_per_wave;
    const int s = grp_root[g];
    uint32_t* out = B2 + b2_off[g];
    const bool is_new = OUT.flags[s] & kIsNew;
    if (is_new && m <= d) {
        const int nn = (int)min((int64_t)(1 << (m - 1)), ub[g]);
        for (int k = lane; k < nn; k += kWave) out[k] = (uint32_t)k;
        if (lane == 0) b2_n[g] = nn;
        return;
    }
    const int W = m >= 5 ? (1 << (m - 5)) : 1;
    for (int w = lane; w < W; w += kWave) bits[w] = 0;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const uint32_t* Bt = Barena + OUT.bt_off[s];
    const int nb = OUT.bt_n[s];
    const uint32_t sg = (uint32_t)OUT.bt_sgn[s];
    if (is_new) {
        const int pbc = OUT.pbc[s], z = OUT.z[s];
        const bool coal = m < pbc;
        int bitpos[kMaxM];
        int ncoal = 0;
        if (coal) {
            const uint8_t* cm = OUT.cmap + (int64_t)s * OUT.MM;
            uint32_t seen = 0;
            for (int j = 0; j < pbc; j++) {
                const int ci = cm[j];
                if (!((seen >> ci) & 1u)) {
                    seen |= (1u << ci);
                    bitpos[ncoal++] = j;
                }
            }
        }
        const uint32_t mask_z = 1u << z, top = 1u << (pbc - 1);
        const uint32_t rev = (pbc >= 32) ? 0xffffffffu : ((1u << pbc) - 1u);
        const uint32_t low = mask_z - 1u;
        for (int j = lane; j < nb; j += kWave) {
            const uint32_t bb = Bt[j] ^ sg;
            uint32_t qv = bb ^ mask_z;
            if (qv & top) qv ^= rev;
            if (bsearch_u32(Bt, qv ^ sg, nb) < 0) continue;
            const uint32_t zb = (bb & mask_z) >> z;
            const uint32_t c1 = ((bb >> (z + 1)) << z) | (bb & low) | (zb << (pbc - 1));
            const uint32_t c2 = c1 ^ top;
            uint32_t cs[2] = {(c1 & top) ? (c1 ^ rev) : c1, (c2 & top) ? (c2 ^ rev) : c2};
            for (int e = 0; e < 2; e++) {
                uint32_t key = cs[e];
                if (coal) {
                    uint32_t bc = 0;
                    for (int l = 0; l < ncoal; l++)
                        if ((key >> bitpos[l]) & 1u) bc |= (1u << l);
                    key = bc;
                }
                if (key >= (1u << m)) {
                    atomicOr(err, 16);
                    continue;
                }
                atomicOr(&bits[key >> 5], 1u << (key & 31));
            }
        }
    } else {
        for (int j = lane; j < nb; j += kWave) {
            const uint32_t key = Bt[j] ^ sg;
            if (key >= (1u << m)) {
                atomicOr(err, 16);
                continue;
            }
            atomicOr(&bits[key >> 5], 1u << (key & 31));
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const int per = (W + kWave - 1) / kWave;
    const int w0 = lane * per, w1 = min(W, w0 + per);
    int cnt = 0;
    for (int w = w0; w < w1; w++) cnt += __popc(bits[w]);
    int total;
    int pos = wave_excl_scan(cnt, &total);
    const int64_t cap = ub[g];
    for (int w = w0; w < w1; w++) {
        uint32_t x = bits[w];
        while (x) {
            const int bt = __ffs(x) - 1;
            x &= x - 1;
            if (pos < cap) out[pos] = (uint32_t)(w * 32 + bt);
            pos++;
        }
    }
    if (lane == 0) {
        b2_n[g] = (int)min((int64_t)total, cap);
        if (total > cap) atomicOr(err, 32);
    }
}

// ------------------------------------------------------------------ group g-tables
// Parent g-table lookup modes of the g-table kernel.
//   kLookupSearch     binary search over the parent keys in global memory
//   kLookupSearchLds  binary search over an LDS copy of the parent keys
//   kLookupRank       rank structure in LDS (bitset of the parent keys + per-word
//                     prefix counts), values read from global memory
//   kLookupRankVals   rank structure and an LDS copy of the parent values
// Every mode finds the same entry; they differ only in where the index comes from.
enum LookupMode { kLookupSearch = 0, kLookupSearchLds = 1, kLookupRank = 2, kLookupRankVals = 3 };

// Per-term data of one g-table evaluation, staged once per evaluated term: its
// q, coalignment maps and its parent g-table (keys, or rank structure and values)
// live in LDS, the rest in registers.
struct EvalTerm {
    const uint32_t* pkeys;  // parent g-table keys (LDS copy or global)
    const cplx* pvals;      // parent g-table values (LDS copy or global)
    const uint32_t* rbits;  // rank structure: key bitset
    const uint32_t* rpre;   // rank structure: keys in the words before each word
    uint32_t klim;          // 2^(phc-1): every stored key is below it (half storage)
    uint32_t fm;            // rows that enter ygi: below m and not H-orthogonal
    int pn;                 // parent g-table size
    int m, phc, z, lhp;
    uint32_t hf;
    bool is_new, has_cmap;
    double c, dv, gscale;
};

// Half-storage g numerator lookup in the parent g-table, eval_gs.hpp:94-153.
template <int MODE>
__device__ inline cplx lookup_g(const EvalTerm& T, int enc, int top_bit, int rev_mask) {
    if (MODE >= kLookupRank) {
        // branchless: both LDS reads are issued together, the value read uses
        // entry 0 when the key is absent and the result is then replaced by zero
        const bool cj = (enc & top_bit) != 0;
        const uint32_t key0 = (uint32_t)(cj ? (rev_mask ^ enc) : enc);
        const bool inr = key0 < T.klim;
        const uint32_t key = inr ? key0 : 0u;
        const uint32_t wv = T.rbits[key >> 5];
        const uint32_t pre = T.rpre[key >> 5];
        const uint32_t bit = 1u << (key & 31);
        const bool found = inr && (wv & bit) != 0;
        const int idx = found ? (int)(pre + __popc(wv & (bit - 1u))) : 0;
        // the LDS value copy always holds entry 0; a global table may be empty
        const cplx v = (MODE == kLookupRankVals || found) ? T.pvals[idx] : cmake(0, 0);
        const double re = found ? v.re : 0.0;
        const double im = found ? (cj ? -v.im : v.im) : 0.0;
        return cmake(re, im);
    }
    return lookup_num(enc, top_bit, rev_mask, T.pkeys, T.pvals, T.pn);
}

// g value of one term at one cell (make_gtable loop body, flattening.hpp:129-248).
// ygi is summed in row order as in the reference.  qr holds q for the rows that
// enter the sum and -0.0 for the others (KM >= m, compile-time bound): adding
// -0.0 leaves any double unchanged, and q * (-1) is q with its sign bit flipped.
template <int MODE, int KM>
__device__ inline cplx eval_cell(const EvalTerm& T, const double (&qr)[KM], const uint8_t* cm, const int8_t* cs,
                                 uint32_t key) {
    double ygi = 0;
    const uint32_t kb = key & T.fm;
#pragma unroll
    for (int k = 0; k < KM; k++) {
        const unsigned long long sb = (unsigned long long)((kb >> k) & 1u) << 63;
        ygi += __longlong_as_double((long long)((unsigned long long)__double_as_longlong(qr[k]) ^ sb));
    }
    const int phc = T.phc;
    const int rev_phc = (phc >= 32) ? -1 : ((1 << phc) - 1);
    const int top_phc = 1 << (phc - 1);
    int lp = 0, lmv = 0;
    if (!T.is_new) {
        lp = (int)key & rev_phc;
        lmv = lp;
    } else if (!T.has_cmap) {
        // the child's key with bit z of the parent inserted: parent bit k is child
        // bit k below z and child bit k-1 above it; bit z is 0 in lp and 1 in lmv
        if (T.z < phc) {
            const uint32_t low = key & ((1u << T.z) - 1u);
            const uint32_t high = (key >> T.z) << (T.z + 1);
            lp = (int)((low | high) & (uint32_t)rev_phc);
            lmv = lp | (1 << T.z);
        } else {
            lp = (int)key & rev_phc;
            lmv = lp;
        }
    } else {
        int k = 0, l = 0;
        while (k < phc) {
            if (k == T.z) {
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
    const cplx a = lookup_g<MODE>(T, lp ^ T.lhp, top_phc, rev_phc);
    const cplx bm = lookup_g<MODE>(T, lmv ^ T.lhp, top_phc, rev_phc);
    const cplx v = csub(cdiv(a, cmake(ygi + T.dv, T.c)), cdiv(bm, cmake(ygi - T.dv, T.c)));
    return cscale(v, T.gscale);
}

__device__ inline double dev_psum2(const TermPool& T, int64_t s, int m) {
    const double* p = T.p + s * T.MM;
    double ps = 0;
    for (int k = 0; k < m; k++) ps += p[k];
    return ps * ps;
}

// Stages term s for evaluation (a workgroup of one wavefront, blockDim == 64).  `lds` is the
// kernel's dynamic LDS: parent keys (kLookupSearchLds), or the rank structure of
// `wmax` words of bits and `wmax` prefix counts, then `wmax` words of the group's
// own key bitset (built by the kernel), then the parent values.
template <int MODE, int KM>
__device__ inline EvalTerm stage_term(const TermPool& O, int64_t s, int m, const uint32_t* gk_prev,
                                      const cplx* gv_prev, const uint2* rank, double gscale, uint32_t* lds, int wmax,
                                      double (&qr)[KM], uint8_t* cm_s, int8_t* cs_s, int* rare) {
    EvalTerm T;
    const int64_t go = O.gp_off[s];
    T.pn = O.gp_n[s];
    T.pvals = gv_prev + go;
    T.pkeys = gk_prev + go;
    T.rbits = nullptr;
    T.rpre = nullptr;
    T.m = m;
    T.phc = O.phc[s];
    T.z = O.z[s];
    T.lhp = O.lhp[s];
    T.hf = O.hflag[s];
    const uint8_t fl = O.flags[s];
    T.is_new = fl & kIsNew;
    T.has_cmap = fl & kHasCmap;
    T.c = O.c[s];
    T.dv = O.dv[s];
    T.gscale = gscale;
    T.klim = 1u << (T.phc - 1);
    wave_lds_sync();
    const int lane = threadIdx.x;
    if (MODE == kLookupSearchLds) {
        const uint32_t* src = gk_prev + go;
        for (int i = lane; i < T.pn; i += kWave) lds[i] = src[i];
        T.pkeys = lds;
    } else if (MODE >= kLookupRank) {
        // the parent's rank structure, built once per update by k_rank_build
        const int W = (int)((T.klim + 31u) >> 5);
        uint32_t* bits = lds;
        uint32_t* pre = lds + wmax;
        const uint2* R = rank + O.rk_off[s];
        for (int w = lane; w < W; w += kWave) {
            const uint2 v = R[w];
            bits[w] = v.x;
            pre[w] = v.y;
        }
        T.rbits = bits;
        T.rpre = pre;
        if (MODE == kLookupRankVals) {
            cplx* dst = reinterpret_cast<cplx*>(lds + 3 * wmax);
            const cplx* vs = gv_prev + go;
            for (int i = lane; i < T.pn; i += kWave) dst[i] = vs[i];
            if (T.pn == 0 && lane == 0) dst[0] = cmake(0, 0);
            T.pvals = dst;
        }
    }
    const int MM = O.MM;
    T.fm = (m >= 32 ? 0xffffffffu : ((1u << m) - 1u)) & ~T.hf;
#pragma unroll
    for (int kk = 0; kk < KM; kk++) qr[kk] = ((T.fm >> kk) & 1u) ? O.q[s * MM + kk] : -0.0;
    const int pbc = O.pbc[s];
    if (T.has_cmap && lane < pbc) {
        cm_s[lane] = O.cmap[s * MM + lane];
        cs_s[lane] = O.csmap[s * MM + lane];
    }
    wave_lds_sync();
    return T;
}

// Rank structure of each parent's g-table, once per update: a bitset of its keys below
// 2^(phc-1) and, per 32-bit word, the index of its first key at or after the word's first
// key value.  Every child of the parent evaluated by the g-table kernel reads it from here
// (stage_term) instead of rebuilding it from the keys.  One wavefront per parent; the
// structure is built in LDS (wmax words of bits, then wmax of prefix counts) and written
// out as (bits, prefix) pairs at sc->rk_base[phc] + (index in its shape bucket) * W.
__device__ __forceinline__ void k_rank_build(MG<64, 1> mg, const TermPool P, int n, const StepConst* __restrict__ sc,
                                             const uint32_t* __restrict__ gkeys, uint2* __restrict__ rank, int wmax) {
    extern __shared__ __align__(16) uint32_t lds_dyn[];
    const int t = mg.bx;
    if (t >= n) return;
    const int lane = threadIdx.x;
    const int phc = P.phc[t];
    const uint32_t klim = 1u << (phc - 1);
    const int W = (int)((klim + 31u) >> 5);
    const int64_t off = sc->rk_base[phc] + (int64_t)(t - sc->in_off[phc]) * W;
    if (lane == 0) P.rk_off[t] = off;
    const int pn = P.gp_n[t];
    const uint32_t* src = gkeys + P.gp_off[t];
    // keys are sorted ascending; W words cover those below klim.  A key at or above klim (a
    // coaligned child's cell whose bit m-1 was left set, kept as a parent key) sorts after
    // every key below it and is never looked up (lookups reverse their key below klim,
    // eval_gs.hpp:116-151), so it only ends the prefix counts.
    uint32_t* bits = lds_dyn;
    uint32_t* pre = lds_dyn + wmax;
    for (int w = lane; w < W; w += kWave) bits[w] = 0;
    __syncthreads();
    // eight keys per lane per round; all loads of a round are issued before the LDS updates
    constexpr int U = 8;
    for (int i0 = 0; i0 < pn; i0 += U * kWave) {
        uint32_t kc[U], kp[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int i = min(i0 + lane + u * kWave, pn - 1);
            kc[u] = src[i];
            kp[u] = src[i > 0 ? i - 1 : 0];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int i = i0 + lane + u * kWave;
            if (i >= pn) continue;
            const uint32_t key = kc[u];
            if (key < klim) atomicOr(&bits[key >> 5], 1u << (key & 31));
            // pre[x] = index of the first key in word >= x, filled for the words between the
            // previous key's word and this key's word
            const int w = min((int)(key >> 5), W - 1);
            const int wp = i ? min((int)(kp[u] >> 5), W - 1) : -1;
            for (int x = wp + 1; x <= w; x++) pre[x] = i;
        }
    }
    const int lastw = pn ? min((int)(src[pn - 1] >> 5), W - 1) : -1;
    for (int w = lane; w < W; w += kWave)
        if (w > lastw) pre[w] = pn;
    __syncthreads();
    for (int w = lane; w < W; w += kWave) rank[off + w] = make_uint2(bits[w], pre[w]);
}

// Flattening of one reduction group, make_gtables (flattening.hpp:351-564):
// the root's g-table; if the root is negligible (term approximation), the
// members in order re-orient the group B-table until one is not (root
// replacement); then every later non-negligible member's g-table is added into
// the root's in member order (add_gtables, flattening.hpp:259-315).
// One wavefront per group; lanes own cells.
template <int MODE, int KM>
__device__ __forceinline__ void k_group_eval(MG<64, 1> mg, const TermPool OUT, int ng, int d, int wmax,
                                                  const int* __restrict__ grp_root,
                                                  const int* __restrict__ grp_moff, const int* __restrict__ grp_mcnt,
                                                  const int* __restrict__ grp_m, const int* __restrict__ mem_slots,
                                                  const uint32_t* __restrict__ B2, const int64_t* __restrict__ b2_off,
                                                  const int* __restrict__ b2_n, const int64_t* __restrict__ g_off,
                                                  const uint32_t* __restrict__ gk_prev, const cplx* __restrict__ gv_prev,
                                                  uint32_t* __restrict__ gk_out, cplx* __restrict__ gv_out,
                                                  cplx* __restrict__ scratch, double gscale, int* __restrict__ surv,
                                                  int* __restrict__ gxor, int* __restrict__ sort_list,
                                                  int* __restrict__ sort_count, int* __restrict__ rare,
                                                  int big_members, int* __restrict__ defer_k0,
                                                  uint8_t* __restrict__ mem_add, const uint2* __restrict__ rank) {
    extern __shared__ __align__(16) uint32_t lds_dyn[];
    double qr[KM];
    __shared__ uint8_t cm_s[kMaxM];
    __shared__ int8_t cs_s[kMaxM];
    const int g = mg.bx;
    if (g >= ng) return;
    const int lane = threadIdx.x;
    const int m = grp_m[g];
    const int MM = OUT.MM;
    const uint32_t* B = B2 + b2_off[g];
    const int n = b2_n[g];
    uint32_t* ok = gk_out + g_off[g];
    cplx* ov = gv_out + g_off[g];
    cplx* sc = scratch + g_off[g];
    const int* mem = mem_slots + grp_moff[g];
    const int nmem = grp_mcnt[g];
    const uint32_t top = 1u << (m - 1);
    const uint32_t rev = (m >= 32) ? 0xffffffffu : ((1u << m) - 1u);
    int r = grp_root[g];
    uint32_t X = 0;
    if constexpr (MODE >= kLookupRank) {
        // bitset of the group's B-table keys below 2^(m-1), for add_gtables.  A coaligned
        // child's projected cells are not re-normalised (cell_enumeration.hpp:206-371), so
        // a key can have bit m-1 set; add_gtables only ever asks for keys below 2^(m-1)
        // (flattening.hpp:297-312), so such a key is never found and is left out.
        uint32_t* bset = lds_dyn + 2 * wmax;
        const int WB = (int)((top + 31u) >> 5);
        for (int w = lane; w < WB; w += kWave) bset[w] = 0;
        wave_lds_sync();
        for (int i = lane; i < n; i += kWave) {
            const uint32_t key = B[i];
            if (key < top) atomicOr(&bset[key >> 5], 1u << (key & 31));
        }
        // ordered before its reads by the LDS syncs of stage_term
    }
    {
        EvalTerm T = stage_term<MODE, KM>(OUT, r, m, gk_prev, gv_prev, rank, gscale, lds_dyn, wmax, qr, cm_s, cs_s, rare);
        const double ps = dev_psum2(OUT, r, m);
        // negligible iff every cell is small; the test stops at the first large cell
        bool negl = true;
        uint32_t bnext = n > 0 ? B[min(lane, n - 1)] : 0u;  // cell key of the next pass, loaded a pass ahead
        for (int i0 = 0; i0 < n; i0 += kWave) {
            const int i = i0 + lane;
            const uint32_t bcur = bnext;
            bnext = B[min(i + kWave, n - 1)];
            cplx v = cmake(0, 0);
            if (i < n) {
                v = eval_cell<MODE, KM>(T, qr, cm_s, cs_s, bcur);
                ov[i] = v;
            }
            if (negl) negl = wave_all(i >= n || !((ps * hypot(v.re, v.im)) > kApproxEps));
        }
        int k = 0;
        if (negl) {
            // replacement chain (flattening.hpp:412-489)
            int lfr = r;
            r = -1;
            while (k < nmem) {
                const int sk = mem[k++];
                if (!(OUT.flags[sk] & kIsNew) && lane == 0) atomicOr(rare, 1);
                int sig = orientation_mask(OUT.A + (int64_t)lfr * MM * d, OUT.A + (int64_t)sk * MM * d, m, d);
                if ((uint32_t)sig & top) sig ^= (int)rev;
                X ^= (uint32_t)sig;
                EvalTerm Tk = stage_term<MODE, KM>(OUT, sk, m, gk_prev, gv_prev, rank, gscale, lds_dyn, wmax, qr, cm_s, cs_s, rare);
                const double psk = dev_psum2(OUT, sk, m);
                bool sm = true;
                uint32_t bnext = n > 0 ? B[min(lane, n - 1)] : 0u;
                for (int i0 = 0; i0 < n; i0 += kWave) {
                    const int i = i0 + lane;
                    const uint32_t bcur = bnext;
                    bnext = B[min(i + kWave, n - 1)];
                    cplx v = cmake(0, 0);
                    if (i < n) {
                        v = eval_cell<MODE, KM>(Tk, qr, cm_s, cs_s, bcur ^ X);
                        ov[i] = v;
                    }
                    if (sm) sm = wave_all(i >= n || !((psk * hypot(v.re, v.im)) > kApproxEps));
                }
                if (!sm) {
                    r = sk;
                    break;
                }
                lfr = sk;
            }
            if (r < 0) {
                if (lane == 0) surv[g] = -1;
                return;
            }
        }
        // later members are added into the root's table; in a group with more than
        // big_members members that is left to k_group_members / k_group_sum
        const double* Ar = OUT.A + (int64_t)r * MM * d;
        if (MODE >= kLookupRank && nmem > big_members) {
            if (lane == 0) defer_k0[g] = k;
            k = nmem;
        }
        for (; k < nmem; k++) {
            const int sk = mem[k];
            if (!(OUT.flags[sk] & kIsNew) && lane == 0) atomicOr(rare, 2);
            const int sig = orientation_mask(Ar, OUT.A + (int64_t)sk * MM * d, m, d);
            const uint32_t sh = ((uint32_t)sig & top) ? ((uint32_t)sig ^ rev) : (uint32_t)sig;
            EvalTerm Tk = stage_term<MODE, KM>(OUT, sk, m, gk_prev, gv_prev, rank, gscale, lds_dyn, wmax, qr, cm_s, cs_s, rare);
            const double psk = dev_psum2(OUT, sk, m);
            if constexpr (MODE >= kLookupRank) {
                // negligibility over the member's own cells B ^ X ^ sh, stopping at the
                // first large cell
                bool sm = true;
                for (int i0 = 0; i0 < n && sm; i0 += kWave) {
                    const int i = i0 + lane;
                    cplx v = cmake(0, 0);
                    if (i < n) v = eval_cell<MODE, KM>(Tk, qr, cm_s, cs_s, B[i] ^ X ^ sh);
                    sm = wave_all(i >= n || !((psk * hypot(v.re, v.im)) > kApproxEps));
                }
                if (mem_add && lane == 0) mem_add[grp_moff[g] + k] = sm ? 0 : 1;
                if (sm) continue;
                // add_gtables: root cell bi takes the member's value at bi ^ sigma
                // (reversed and conjugated when that key is in the lower half), which
                // is evaluated directly; a key outside the member's table adds nothing
                const uint32_t* bset = lds_dyn + 2 * wmax;
                for (int i = lane; i < n; i += kWave) {
                    const uint32_t bi = B[i] ^ X;
                    uint32_t x = bi ^ (uint32_t)sig;
                    bool cj = false;
                    if (x & top) {
                        cj = true;
                        x ^= rev;
                    }
                    const uint32_t y = x ^ sh ^ X;  // the member's key x as a key of B
                    if (!((bset[y >> 5] >> (y & 31)) & 1u)) continue;
                    const cplx v = eval_cell<MODE, KM>(Tk, qr, cm_s, cs_s, x);
                    const cplx cur = ov[i];
                    ov[i] = cadd(cur, cj ? cconj(v) : v);
                }
            } else {
                bool sm = true;
                uint32_t bnext = n > 0 ? B[min(lane, n - 1)] : 0u;
                for (int i0 = 0; i0 < n; i0 += kWave) {
                    const int i = i0 + lane;
                    const uint32_t bcur = bnext;
                    bnext = B[min(i + kWave, n - 1)];
                    cplx v = cmake(0, 0);
                    if (i < n) {
                        v = eval_cell<MODE, KM>(Tk, qr, cm_s, cs_s, bcur ^ X ^ sh);
                        sc[i] = v;
                    }
                    if (sm) sm = wave_all(i >= n || !((psk * hypot(v.re, v.im)) > kApproxEps));
                }
                if (mem_add && lane == 0) mem_add[grp_moff[g] + k] = sm ? 0 : 1;
                if (sm) continue;
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
                __syncthreads();
                for (int i = lane; i < n; i += kWave) {
                    const uint32_t bi = B[i] ^ X;
                    uint32_t x = bi ^ (uint32_t)sig;
                    bool cj = false;
                    if (x & top) {
                        cj = true;
                        x ^= rev;
                    }
                    cplx v;
                    if (x == (bi ^ sh)) {
                        v = sc[i];
                    } else {
                        const int ii = bsearch_u32(B, x ^ sh ^ X, n);
                        if (ii < 0) continue;
                        v = sc[ii];
                    }
                    const cplx cur = ov[i];
                    ov[i] = cadd(cur, cj ? cconj(v) : v);
                }
                __syncthreads();
            }
        }
    }
    for (int i = lane; i < n; i += kWave) ok[i] = B[i] ^ X;
    if (lane == 0) {
        surv[g] = r;
        gxor[g] = (int)X;
        if (X != 0) {
            const int slot = atomicAdd(sort_count, 1);
            sort_list[slot] = g;
        }
    }
}

// Work lists of the groups whose members are added in parallel: per group the
// member count, the contribution entries (members x cells) and the 64-cell chunks.
__device__ __forceinline__ void k_big_prep(MG<1024, 1> mg, int ng, int big_members, const int* __restrict__ grp_mcnt, const int* __restrict__ b2_n,
                           int* __restrict__ icnt, int64_t* __restrict__ ccnt, int* __restrict__ hcnt) {
    const int g = mg.bx * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    const int nm = grp_mcnt[g];
    const bool big = nm > big_members;
    icnt[g] = big ? nm : 0;
    ccnt[g] = big ? (int64_t)nm * b2_n[g] : 0;
    hcnt[g] = big ? (b2_n[g] + kWave - 1) / kWave : 0;
}

// Largest g with off[g] <= item (off ascending, off[0] == 0).
__device__ inline int owner_of(const int* __restrict__ off, int n, int item) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// One wavefront per (large group, member k): the member's g-table contribution to
// every cell of the group's root, in the root's orientation (add_gtables,
// flattening.hpp:259-315), written to contrib[coff[g] + k * n + i].  A key outside
// the member's table contributes (-0.0, -0.0), which leaves any sum unchanged; a
// negligible member is flagged in cneg and contributes nothing.
template <int MODE, int KM>
__device__ __forceinline__ void k_group_members(MG<64, 1> mg, const TermPool OUT, int ng, int d, int wmax,
                                                     const int* __restrict__ ioff, int nitems,
                                                     const int64_t* __restrict__ coff,
                                                     const int* __restrict__ grp_moff, const int* __restrict__ grp_mcnt,
                                                     const int* __restrict__ grp_m, const int* __restrict__ mem_slots,
                                                     const uint32_t* __restrict__ B2, const int64_t* __restrict__ b2_off,
                                                     const int* __restrict__ b2_n, const uint32_t* __restrict__ gk_prev,
                                                     const cplx* __restrict__ gv_prev, double gscale,
                                                     const int* __restrict__ surv, const int* __restrict__ gxor,
                                                     const int* __restrict__ defer_k0, cplx* __restrict__ contrib,
                                                     uint8_t* __restrict__ cneg, int* __restrict__ rare,
                                                     uint8_t* __restrict__ mem_add, const uint2* __restrict__ rank) {
    static_assert(MODE >= kLookupRank, "parallel member sums use the rank lookup");
    extern __shared__ __align__(16) uint32_t lds_dyn[];
    double qr[KM];
    __shared__ uint8_t cm_s[kMaxM];
    __shared__ int8_t cs_s[kMaxM];
    const int item = mg.bx;
    if (item >= nitems) return;
    const int g = owner_of(ioff, ng, item);
    const int k = item - ioff[g];
    const int r = surv[g];
    if (r < 0 || k < defer_k0[g]) return;
    const int lane = threadIdx.x;
    const int m = grp_m[g];
    const int MM = OUT.MM;
    const uint32_t* B = B2 + b2_off[g];
    const int n = b2_n[g];
    const uint32_t X = (uint32_t)gxor[g];
    const uint32_t top = 1u << (m - 1);
    const uint32_t rev = (m >= 32) ? 0xffffffffu : ((1u << m) - 1u);
    uint32_t* bset = lds_dyn + 2 * wmax;
    {
        const int WB = (int)((top + 31u) >> 5);
        for (int w = lane; w < WB; w += kWave) bset[w] = 0;
        __syncthreads();
        for (int i = lane; i < n; i += kWave) {
            const uint32_t key = B[i];
            if (key < top) atomicOr(&bset[key >> 5], 1u << (key & 31));
        }
    }
    const int sk = mem_slots[grp_moff[g] + k];
    if (!(OUT.flags[sk] & kIsNew) && lane == 0) atomicOr(rare, 2);
    const int sig = orientation_mask(OUT.A + (int64_t)r * MM * d, OUT.A + (int64_t)sk * MM * d, m, d);
    const uint32_t sh = ((uint32_t)sig & top) ? ((uint32_t)sig ^ rev) : (uint32_t)sig;
    EvalTerm Tk = stage_term<MODE, KM>(OUT, sk, m, gk_prev, gv_prev, rank, gscale, lds_dyn, wmax, qr, cm_s, cs_s, rare);
    const double psk = dev_psum2(OUT, sk, m);
    bool sm = true;
    for (int i0 = 0; i0 < n && sm; i0 += kWave) {
        const int i = i0 + lane;
        cplx v = cmake(0, 0);
        if (i < n) v = eval_cell<MODE, KM>(Tk, qr, cm_s, cs_s, B[i] ^ X ^ sh);
        sm = wave_all(i >= n || !((psk * hypot(v.re, v.im)) > kApproxEps));
    }
    if (lane == 0) cneg[item] = sm ? 1 : 0;
    if (mem_add && lane == 0) mem_add[grp_moff[g] + k] = sm ? 0 : 1;
    if (sm) return;
    cplx* out = contrib + coff[g] + (int64_t)k * n;
    for (int i = lane; i < n; i += kWave) {
        const uint32_t bi = B[i] ^ X;
        uint32_t x = bi ^ (uint32_t)sig;
        bool cj = false;
        if (x & top) {
            cj = true;
            x ^= rev;
        }
        const uint32_t y = x ^ sh ^ X;
        cplx c = cmake(-0.0, -0.0);
        if ((bset[y >> 5] >> (y & 31)) & 1u) {
            const cplx v = eval_cell<MODE, KM>(Tk, qr, cm_s, cs_s, x);
            c = cj ? cconj(v) : v;
        }
        out[i] = c;
    }
}

// Adds the member contributions of each large group into its root's table, one
// lane per cell, members in order (the sum order of add_gtables).
__device__ __forceinline__ void k_group_sum(MG<64, 1> mg, int ng, const int* __restrict__ hoff, int nchunks,
                                                 const int* __restrict__ ioff, const int64_t* __restrict__ coff,
                                                 const int* __restrict__ grp_mcnt, const int* __restrict__ b2_n,
                                                 const int64_t* __restrict__ g_off, const int* __restrict__ surv,
                                                 const int* __restrict__ defer_k0, const cplx* __restrict__ contrib,
                                                 const uint8_t* __restrict__ cneg, cplx* __restrict__ gv_out) {
    const int item = mg.bx;
    if (item >= nchunks) return;
    const int g = owner_of(hoff, ng, item);
    if (surv[g] < 0) return;
    const int n = b2_n[g];
    const int i = (item - hoff[g]) * kWave + threadIdx.x;
    if (i >= n) return;
    const int nm = grp_mcnt[g];
    const cplx* cb = contrib + coff[g] + i;
    const uint8_t* ng_ = cneg + ioff[g];
    cplx acc = gv_out[g_off[g] + i];
    for (int k = defer_k0[g]; k < nm; k++) {
        if (ng_[k]) continue;
        const cplx c = cb[(int64_t)k * n];
        acc = cadd(acc, c);
    }
    gv_out[g_off[g] + i] = acc;
}

// Restores ascending key order of a group table whose keys were re-oriented by
// a root replacement.  One block per listed group; rank from an LDS bitset.
__device__ __forceinline__ void k_group_sort(MG<1024, 1> mg, int nlist, const int* __restrict__ list, const int* __restrict__ grp_m,
                             const int* __restrict__ b2_n, const int64_t* __restrict__ g_off,
                             uint32_t* __restrict__ gk, cplx* __restrict__ gv, uint32_t* __restrict__ tmpk,
                             cplx* __restrict__ tmpv) {
    extern __shared__ uint32_t bits[];
    __shared__ int wbase[1024];
    const int g = list[mg.bx];
    const int m = grp_m[g];
    const int n = b2_n[g];
    const int W = m >= 5 ? (1 << (m - 5)) : 1;
    uint32_t* k = gk + g_off[g];
    cplx* v = gv + g_off[g];
    uint32_t* tk = tmpk + g_off[g];
    cplx* tv = tmpv + g_off[g];
    for (int w = threadIdx.x; w < W; w += blockDim.x) bits[w] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        tk[i] = k[i];
        tv[i] = v[i];
        atomicOr(&bits[k[i] >> 5], 1u << (k[i] & 31));
    }
    __syncthreads();
    // prefix popcounts over words, in chunks of blockDim
    const int per = (W + blockDim.x - 1) / blockDim.x;
    const int w0 = threadIdx.x * per, w1 = min(W, w0 + per);
    int c = 0;
    for (int w = w0; w < w1; w++) c += __popc(bits[w]);
    wbase[threadIdx.x] = c;
    __syncthreads();
    for (int o = 1; o < (int)blockDim.x; o <<= 1) {
        int t = (threadIdx.x >= (unsigned)o) ? wbase[threadIdx.x - o] : 0;
        __syncthreads();
        wbase[threadIdx.x] += t;
        __syncthreads();
    }
    const int tb = wbase[threadIdx.x] - c;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const uint32_t key = tk[i];
        const int w = key >> 5;
        const int owner = w / per;
        int pos = (owner > 0 ? wbase[owner - 1] : 0);
        for (int ww = owner * per; ww < w; ww++) pos += __popc(bits[ww]);
        pos += __popc(bits[w] & ((1u << (key & 31)) - 1u));
        k[pos] = key;
        v[pos] = tv[i];
    }
    (void)tb;
}

// ------------------------------------------------------------------ compaction
template <int D>
__device__ __forceinline__ void k_compact(MG<1024, 1> mg, const TermPool OUT, int ng, const int* __restrict__ surv, const int* __restrict__ pos,
                          const int* __restrict__ grp_m, const int64_t* __restrict__ g_off,
                          const int* __restrict__ b2_n, TermPool NEXT) {
    const int g = mg.bx * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    const int r = surv[g];
    if (r < 0) return;
    const int o = pos[g];
    const int m = grp_m[g];
    const int MM = OUT.MM;
    const double* A = OUT.A + (int64_t)r * MM * D;
    const double* p = OUT.p + (int64_t)r * MM;
    const double* b = OUT.b + (int64_t)r * D;
    double* nA = NEXT.A + (int64_t)o * MM * D;
    double* np_ = NEXT.p + (int64_t)o * MM;
    double* nb = NEXT.b + (int64_t)o * D;
    for (int k = 0; k < m * D; k++) nA[k] = A[k];
    for (int k = 0; k < m; k++) np_[k] = p[k];
    for (int k = 0; k < D; k++) nb[k] = b[k];
    NEXT.m[o] = m;
    NEXT.gp_off[o] = g_off[g];
    NEXT.gp_n[o] = b2_n[g];
}

// g-tables of the first update's terms (make_gtable_first, flattening.hpp:14-67).
template <int D>
__device__ __forceinline__ void k_gtable_first(MG<1024, 1> mg, const TermPool OUT, int n, const uint32_t* __restrict__ perm, double gscale,
                               uint32_t* __restrict__ gk, cplx* __restrict__ gv, TermPool NEXT) {
    const int r = mg.bx * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int64_t s = perm[r];
    const int cells = 1 << (D - 1);
    const int MM = OUT.MM;
    const double* p = OUT.p + s * MM;
    const uint32_t hf = OUT.hflag[s];
    const double c = OUT.c[s], dv = OUT.dv[s];
    const int64_t off = (int64_t)r * cells;
    for (int j = 0; j < cells; j++) {
        double ygi = 0;
        for (int k = 0; k < D; k++)
            if (!((hf >> k) & 1u)) {
                const double sb = ((j >> k) & 1) == 0 ? 1 : -1;
                ygi += p[k] * sb;
            }
        cplx v = csub(cdiv(cmake(1.0, 0.0), cmake(ygi + dv, c)), cdiv(cmake(1.0, 0.0), cmake(ygi - dv, c)));
        gk[off + j] = (uint32_t)j;
        gv[off + j] = cscale(v, gscale);
    }
    const double* A = OUT.A + s * MM * D;
    const double* b = OUT.b + s * D;
    for (int k = 0; k < D * D; k++) NEXT.A[(int64_t)r * MM * D + k] = A[k];
    for (int k = 0; k < D; k++) NEXT.p[(int64_t)r * MM + k] = p[k];
    for (int k = 0; k < D; k++) NEXT.b[(int64_t)r * D + k] = b[k];
    NEXT.m[r] = D;
    NEXT.gp_off[r] = off;
    NEXT.gp_n[r] = cells;
}

// Algorithmic bytes of k_group_eval: per group, the B-table keys (4 B/cell) and the
// output table (key + complex value, 20 B/cell); per evaluated term (root and every
// member), its record (q, p, A rows, scalars, coalignment maps) and, per cell, the two
// parent g-table entries its numerators need (2 x 20 B).
__device__ __forceinline__ void k_geval_bytes(MG<1024, 1> mg, int ng, int d, const int* __restrict__ grp_m, const int* __restrict__ grp_mcnt,
                              const int* __restrict__ b2_n, const int* __restrict__ surv,
                              const int* __restrict__ defer_k0, int big_members, unsigned long long* __restrict__ acc) {
    const int g = mg.bx * blockDim.x + threadIdx.x;
    unsigned long long v = 0;
    if (g < ng) {
        const unsigned long long n = (unsigned long long)b2_n[g];
        const unsigned long long m = (unsigned long long)grp_m[g];
        const unsigned long long rec = m * (16ull + 8ull * d) + 2ull * m + 52ull;
        // terms evaluated by k_group_eval: the root and the members it adds itself (a
        // large group's later members are evaluated by k_group_members)
        const int nm = grp_mcnt[g];
        const int own = (nm > big_members && surv[g] >= 0) ? defer_k0[g] : nm;
        const unsigned long long evals = 1ull + (unsigned long long)own;
        v = 24ull * n + evals * (40ull * n + rec);
    }
    // one atomic per wavefront: the counter is shared by every group of the launch
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_down(v, o, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0 && v) atomicAdd(acc, v);
}

// acc += sum over i < n of a[i] * (b ? b[i] : 1), over the i with keep[i] >= 0 when keep is
// given: the cell counts of the update's byte accounting (profiling only).
__device__ __forceinline__ void k_prof_sum(MG<1024, 1> mg, int n, const int* __restrict__ a, const int* __restrict__ b,
                                           const int* __restrict__ keep, unsigned long long* __restrict__ acc) {
    const int i = mg.bx * blockDim.x + threadIdx.x;
    unsigned long long v = 0;
    if (i < n && (!keep || keep[i] >= 0))
        v = (unsigned long long)max(a[i], 0) * (unsigned long long)(b ? max(b[i], 0) : 1);
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_down(v, o, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0 && v) atomicAdd(acc, v);
}

// Start of each reference thread's chain in the sorted chain keys (T+1 entries).
__device__ __forceinline__ void k_chain_off(MG<1024, 1> mg, const uint32_t* __restrict__ skeys, int n, int T, int* __restrict__ off) {
    const int t = threadIdx.x;
    if (t > T) return;
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (skeys[mid] < (uint32_t)t) lo = mid + 1;
        else hi = mid;
    }
    off[t] = lo;
}

// ------------------------------------------------------------------ small helpers
// Total of an exclusive scan: off[n-1] + cnt[n-1].
template <class A, class B>
__device__ __forceinline__ void k_scan_total(MG<1024, 1> mg, const A* __restrict__ off, const B* __restrict__ cnt, int n, int64_t* __restrict__ out) {
    out[0] = n > 0 ? (int64_t)off[n - 1] + (int64_t)cnt[n - 1] : 0;
}

// The same total, also stored into a device-mapped host slot when one is given.
template <class A, class B>
__device__ __forceinline__ void k_scan_total_mapped(MG<1024, 1> mg, const A* __restrict__ off,
                                                    const B* __restrict__ cnt, int n, int64_t* __restrict__ out,
                                                    int64_t* host_out) {
    const int64_t v = n > 0 ? (int64_t)off[n - 1] + (int64_t)cnt[n - 1] : 0;
    out[0] = v;
    if (host_out) *(volatile int64_t*)host_out = v;
}

// Scan of up to kSmallScan items by one workgroup: the small arrays of small estimators,
// where a library scan costs more launches than work.  Each thread sums a contiguous run, a
// block scan of the run sums gives each run's offset, and the run is written out.  With
// `wide`, the input is also stored widened; with `out_total` / `host_total`, the sum of all
// items goes to a device scalar and a device-mapped host slot.  Integer sums, so the result
// is the same as any other summation order's.
constexpr int kSmallScan = 65536;
constexpr int kSmallScanThreads = 1024;

// Exclusive prefix of each thread's `v` over a block of kSmallScanThreads threads, and the
// block total: wavefront scans, then one wavefront scans the wavefront totals (two
// barriers).  Every thread of the block must call it.
template <class T>
__device__ __forceinline__ T block_excl_scan(T v, T* total) {
    constexpr int kWaves = kSmallScanThreads / kWave;
    __shared__ T wsum[kWaves];
    const int lane = (int)(threadIdx.x & (kWave - 1)), w = (int)(threadIdx.x / kWave);
    const T x = wave_incl_scan(v);
    if (lane == kWave - 1) wsum[w] = x;
    __syncthreads();
    if (w == 0) {
        T y = lane < kWaves ? wsum[lane] : (T)0;
        y = wave_incl_scan(y);
        if (lane < kWaves) wsum[lane] = y;
    }
    __syncthreads();
    *total = wsum[kWaves - 1];
    return (w > 0 ? wsum[w - 1] : (T)0) + x - v;
}

// Items per thread of one tile of the small scan.  Above one tile, tiles of
// kSmallScanThreads x kScanItems items are loaded coalesced (consecutive threads,
// consecutive items) into LDS, each thread scans its kScanItems consecutive items there,
// and the results go back out coalesced.
constexpr int kScanItems = 4;

template <class In, class Out, bool Inclusive>
__device__ __forceinline__ void k_small_scan(MG<kSmallScanThreads, 1> mg, const In* __restrict__ in,
                                             Out* __restrict__ out, int n, Out* __restrict__ wide,
                                             int64_t* __restrict__ out_total, int64_t* host_total) {
    (void)mg;
    constexpr int kTile = kSmallScanThreads * kScanItems;
    __shared__ Out tile[kTile];
    const int t = threadIdx.x;
    if (n <= kTile) {
        // one tile or less: each thread scans at most kScanItems consecutive items in place,
        // which costs fewer barriers than staging them.  The items are loaded once, all
        // loads issued before the first is used.
        const int per = (n + kSmallScanThreads - 1) / kSmallScanThreads;
        const int b = min(n, t * per), e = min(n, b + per);
        Out x[kScanItems];
#pragma unroll
        for (int j = 0; j < kScanItems; j++) x[j] = b + j < e ? (Out)in[b + j] : (Out)0;
        Out sum = 0;
#pragma unroll
        for (int j = 0; j < kScanItems; j++) sum += x[j];
        Out tot;
        Out run = block_excl_scan(sum, &tot);
#pragma unroll
        for (int j = 0; j < kScanItems; j++) {
            const int i = b + j;
            if (i < e) {
                if (wide) wide[i] = x[j];
                if (Inclusive) {
                    run += x[j];
                    out[i] = run;
                } else {
                    out[i] = run;
                    run += x[j];
                }
            }
        }
        if (t == kSmallScanThreads - 1) {
            if (out_total) out_total[0] = (int64_t)tot;
            if (host_total) *(volatile int64_t*)host_total = (int64_t)tot;
        }
        return;
    }
    // Tiles: the next tile's loads are issued before the current tile is scanned, so their
    // latency overlaps the scan's barriers and stores.
    Out nx[kScanItems];
#pragma unroll
    for (int j = 0; j < kScanItems; j++) {
        const int i = j * kSmallScanThreads + t;
        nx[j] = i < n ? (Out)in[i] : (Out)0;
    }
    Out carry = 0;
    for (int base = 0; base < n; base += kTile) {
#pragma unroll
        for (int j = 0; j < kScanItems; j++) {
            const int i = base + j * kSmallScanThreads + t;
            if (wide && i < n) wide[i] = nx[j];
            tile[j * kSmallScanThreads + t] = nx[j];
        }
        if (base + kTile < n) {
#pragma unroll
            for (int j = 0; j < kScanItems; j++) {
                const int i = base + kTile + j * kSmallScanThreads + t;
                nx[j] = i < n ? (Out)in[i] : (Out)0;
            }
        }
        __syncthreads();
        Out v[kScanItems], sum = 0;
#pragma unroll
        for (int j = 0; j < kScanItems; j++) {
            v[j] = tile[t * kScanItems + j];
            sum += v[j];
        }
        Out tot;
        Out run = block_excl_scan(sum, &tot) + carry;
#pragma unroll
        for (int j = 0; j < kScanItems; j++) {
            if (Inclusive) {
                run += v[j];
                tile[t * kScanItems + j] = run;
            } else {
                tile[t * kScanItems + j] = run;
                run += v[j];
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kScanItems; j++) {
            const int i = base + j * kSmallScanThreads + t;
            if (i < n) out[i] = tile[j * kSmallScanThreads + t];
        }
        carry += tot;
        __syncthreads();  // the tile is rewritten
    }
    if (t == kSmallScanThreads - 1) {
        if (out_total) out_total[0] = (int64_t)carry;
        if (host_total) *(volatile int64_t*)host_total = (int64_t)carry;
    }
}

__device__ __forceinline__ void k_widen(MG<1024, 1> mg, const int* __restrict__ src, int64_t* __restrict__ dst, int n) {
    const int i = mg.bx * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (int64_t)src[i];
}
__device__ __forceinline__ void k_surv_flags(MG<1024, 1> mg, int n, const int* __restrict__ surv, int* __restrict__ f) {
    const int i = mg.bx * blockDim.x + threadIdx.x;
    if (i < n) f[i] = surv[i] >= 0 ? 1 : 0;
}
template <int D>
__global__ void k_det_tp(TermPool P, int n, const double* __restrict__ Tr, const double* __restrict__ bias) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int m = P.m[i];
    double* A = P.A + (int64_t)i * P.MM * D;
    double* b = P.b + (int64_t)i * D;
    double t[D];
    for (int r = 0; r < m; r++) {
        for (int k = 0; k < D; k++) t[k] = A[r * D + k];
        for (int j = 0; j < D; j++) {
            double s = 0.0;
            for (int k = 0; k < D; k++) s += t[k] * Tr[k + j * D];
            A[r * D + j] = s;
        }
    }
    for (int k = 0; k < D; k++) t[k] = b[k];
    for (int r = 0; r < D; r++) {
        double s = 0.0;
        for (int j = 0; j < D; j++) s += Tr[r * D + j] * t[j];
        b[r] = s;
    }
    if (bias)
        for (int r = 0; r < D; r++) b[r] += bias[r];
}

// A d-vector passed by value as a kernel argument.
struct VecArg {
    double v[kMaxD];
};

// ------------------------------------------------------------ 1D marginal CPDF
// Per-term closed form of the 1D marginal density (cpdf_ndim.hpp:1233-1368): the
// arrangement is cut along state `marg`; w = sum_j p_j |A_j,marg|, the sign vectors of
// the two half-lines select the numerators, and the term contributes
// Re(g_lhs / (w + i(b_marg - x)) - g_rhs / (-w + i(b_marg - x))) at x.  The second form,
// (a x + bb) / (w^2 + (x - s)^2) with s = b_marg, a = Im(g_rhs) / pi and
// bb = (Re(g_rhs) w - Im(g_rhs) s) / pi, is the one every later grid point uses.
struct CpdfTerm1D {
    double a, bb, w2, s;
};

template <int D>
__global__ void k_cpdf1d_terms(TermPool P, int n, int marg, VecArg bar_nu, const uint32_t* __restrict__ gkeys,
                               const cplx* __restrict__ gvals, double x0, double* __restrict__ direct,
                               CpdfTerm1D* __restrict__ cached) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int m = P.m[i];
    const double* A = P.A + (int64_t)i * P.MM * D;
    const double* p = P.p + (int64_t)i * P.MM;
    const double bm = P.b[(int64_t)i * D + marg];
    double w = 0.0;
    int lhs = 0, rhs = 0;
    for (int j = 0; j < m; j++) {
        const double acj = A[j * D + marg];
        const double fa = fabs(acj);
        w += p[j] * fa;
        bool pos;
        if (fa > 1e-15) {
            pos = acj > 0;
        } else {
            // a row orthogonal to the state: its side is taken from the drawn direction
            double ip = 0.0;
            for (int k = 0; k < D; k++) ip += A[j * D + k] * bar_nu.v[k];
            pos = ip > 0;
        }
        if (pos) lhs |= 1 << j;
        else rhs |= 1 << j;
    }
    const int top = 1 << (m - 1), rev = (1 << m) - 1;
    const int64_t go = P.gp_off[i];
    const int gn = P.gp_n[i];
    const cplx gl = lookup_num(lhs, top, rev, gkeys + go, gvals + go, gn);
    const cplx gr = lookup_num(rhs, top, rev, gkeys + go, gvals + go, gn);
    const double bc = bm - x0;
    const cplx v = csub(cdiv(gl, cmake(w, bc)), cdiv(gr, cmake(-w, bc)));
    direct[i] = v.re;
    CpdfTerm1D c;
    c.s = bm;
    c.a = gr.im / kPi;
    c.bb = (gr.re * w - gr.im * bm) / kPi;
    c.w2 = w * w;
    cached[i] = c;
}

// Deterministic sum of n values by one block (fixed strided partials, then a fixed tree).
__global__ void k_sum_fixed(const double* __restrict__ v, int n, double* __restrict__ out) {
    __shared__ double sh[1024];
    const int t = threadIdx.x;
    double s = 0.0;
    for (int i = t; i < n; i += blockDim.x) s += v[i];
    sh[t] = s;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if (t < h) sh[t] += sh[t + h];
        __syncthreads();
    }
    if (t == 0) out[0] = sh[0];
}

constexpr int kCpdfChunk = 1024;   // terms staged in LDS per block
constexpr int kCpdfPoints = 256;   // grid points per block (one per thread)

// Partial sums over one chunk of terms for a block of grid points:
// part[chunk * npts + k] = sum over the chunk of (a x_k + bb) / (w2 + (x_k - s)^2).
__global__ void __launch_bounds__(kCpdfPoints) k_cpdf1d_grid(const CpdfTerm1D* __restrict__ cached, int n,
                                                             const double* __restrict__ xs, int npts,
                                                             double* __restrict__ part) {
    __shared__ CpdfTerm1D sh[kCpdfChunk];
    const int chunk = blockIdx.y;
    const int t0 = chunk * kCpdfChunk;
    const int nt = min(kCpdfChunk, n - t0);
    for (int j = threadIdx.x; j < nt; j += blockDim.x) sh[j] = cached[t0 + j];
    __syncthreads();
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npts) return;
    const double x = xs[k];
    double acc = 0.0;
    // terms in pairs, n0 / d0 + n1 / d1 = (n0 d1 + n1 d0) / (d0 d1): one division per two
    // terms where d0 d1 (each d = w^2 + (x - s)^2 > 0) stays in range, two otherwise
    int j = 0;
    for (; j + 1 < nt; j += 2) {
        const CpdfTerm1D c0 = sh[j], c1 = sh[j + 1];
        const double x0 = x - c0.s, x1 = x - c1.s;
        const double d0 = c0.w2 + x0 * x0, d1 = c1.w2 + x1 * x1;
        const double n0 = c0.a * x + c0.bb, n1 = c1.a * x + c1.bb;
        const double dd = d0 * d1;
        if (dd > 1e-280 && dd < 1e280) acc += (n0 * d1 + n1 * d0) / dd;
        else acc += n0 / d0 + n1 / d1;
    }
    if (j < nt) {
        const CpdfTerm1D c = sh[j];
        double xm = x - c.s;
        xm *= xm;
        acc += (c.a * x + c.bb) / (c.w2 + xm);
    }
    part[(int64_t)chunk * npts + k] = acc;
}

// y_k = (sum over chunks of part) / norm, the chunks added in order.
__global__ void k_cpdf1d_finish(const double* __restrict__ part, int nchunks, int npts, double norm,
                                double* __restrict__ y) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npts) return;
    double s = 0.0;
    for (int c = 0; c < nchunks; c++) s += part[(int64_t)c * npts + k];
    y[k] = s / norm;
}


// Grid abscissae of the 1D density, x_k = min(low + k res, high), as the host forms them
// (no multiply-add is fused), for callers that keep the grid in device memory.
__global__ void k_cpdf1d_xs(double* __restrict__ xs, int npts, double lo, double res, double hi) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npts) return;
    const double v = lo + (double)k * res;
    xs[k] = hi < v ? hi : v;
}

// The density at the first grid point from the direct sum, (sum / 2 pi) / norm, or
// zero everywhere when the estimator holds no terms.
__global__ void k_cpdf1d_first(const double* __restrict__ direct_sum, double norm, double* __restrict__ y,
                               int zero_n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (zero_n > 0) {
        if (k < zero_n) y[k] = 0.0;
        return;
    }
    if (k == 0) y[0] = direct_sum[0] * kInvTwoPi / norm;
}

// b <- b - dx for every term (cauchy_estimator.hpp:1371-1380, sub_vecs).
template <int D>
__device__ __forceinline__ void k_sub_b(MG<1024, 1> mg, TermPool P, int n, VecArg dx) {
    const int i = mg.bx * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double* b = P.b + (int64_t)i * D;
#pragma unroll
    for (int r = 0; r < D; r++) b[r] -= dx.v[r];
}

// ------------------------------------------------------------ 2D marginal CPDF
// Per-term cell data of the 2D marginal density (cpdf_ndim.hpp:1400-1700): the
// arrangement cut to states (i1, i2), zero rows removed and +/- equal rows merged
// (marg2d_remove_zeros_and_coalign), the cell walls' angles sorted over a full turn,
// and for each of the m' cells of a half turn its g numerator and the real parts of
// gamma1, gamma2.  Arrays have a stride of MM cells per term (MM + 1 angles).
// The point-independent parts of a cell's integral sin(t2 - t1) g / (D1 D2), with
// Dk = gamma1 cos tk + gamma2 sin tk: Re D1, Re D2 and the numerator.  Only Im gamma1 and
// Im gamma2 depend on the grid point.
struct Cell2D {
    double d1r, d2r, nre, nim;
};

struct Cpdf2D {
    int MM;
    int* mc;           // cells per term
    double* b2;        // b of the two states
    double* sn;        // sin of the wall angles
    double* cs;        // cos of the wall angles
    double* g1r;       // Re gamma1 per cell
    double* g2r;       // Re gamma2 per cell
    cplx* gc;          // g numerator per cell
    Cell2D* fc;        // point-independent parts per cell
};

__device__ inline Cell2D cell2d_parts(double s1, double c1, double s2, double c2, double g1re, double g2re, cplx g) {
    const double sd = s2 * c1 - s1 * c2;
    return Cell2D{g1re * c1 + g2re * s1, g1re * c2 + g2re * s2, g.re * sd, g.im * sd};
}

constexpr double kIntegralGammaEps = 1e-8;  // INTEGRAL_GAMMA_EPS, cauchy_constants.hpp:76
constexpr int kZeroHpMarker = 32;           // ZERO_HP_MARKER_VALUE, cpdf_ndim.hpp:371

// Integral of one cell at (x1, x2) (marg2d_eval_term_for_cpdf, cpdf_ndim.hpp:1603-1636);
// `singular` is set where the reference stops with a singularity error.
__device__ inline double cpdf2d_cell(double s1, double c1, double s2, double c2, double g1re, double g2re,
                                     double g1im, double g2im, cplx g, bool check_g1, bool& singular) {
    cplx ga1 = cmake(g1re, g1im), ga2 = cmake(g2re, g2im);
    if (check_g1 && fabs(g1re) < kIntegralGammaEps) {
        if (fabs(g2re) < kIntegralGammaEps && fabs(g2im) < kIntegralGammaEps) {
            singular = true;
            return 0.0;
        }
        const cplx lo = cdiv(csub(cscale(ga1, s1), cscale(ga2, c1)), cadd(cscale(ga1, c1), cscale(ga2, s1)));
        const cplx hi = cdiv(csub(cscale(ga1, s2), cscale(ga2, c2)), cadd(cscale(ga1, c2), cscale(ga2, s2)));
        const cplx sq = cadd(cmul(ga1, ga1), cmul(ga2, ga2));
        return cmul(csub(hi, lo), cdiv(g, sq)).re;
    }
    // The reference's form, sin t2 / (g1 (g1 cos t2 + g2 sin t2)) - (the same at t1), equals
    // sin(t2 - t1) / (D1 D2) with Dk = g1 cos tk + g2 sin tk exactly: one complex quotient
    // instead of two, and no cancellation between the two limits.
    const cplx d1 = cadd(cscale(ga1, c1), cscale(ga2, s1));
    const cplx d2 = cadd(cscale(ga1, c2), cscale(ga2, s2));
    const double sd = s2 * c1 - s1 * c2;
    return cdiv(cscale(g, sd), cmul(d1, d2)).re;
}

template <int D>
__global__ void k_cpdf2d_terms(TermPool P, int n, int i1, int i2, VecArg bar_nu, const uint32_t* __restrict__ gkeys,
                               const cplx* __restrict__ gvals, double x0, double y0, Cpdf2D C,
                               double* __restrict__ direct, int* __restrict__ err) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int m = P.m[i];
    const int MM = C.MM;
    const double* A = P.A + (int64_t)i * P.MM * D;
    const double* p = P.p + (int64_t)i * P.MM;
    double wa[2 * kMaxM], wp[kMaxM], wa2[2 * kMaxM], wp2[kMaxM], th[2 * kMaxM];
    int cmap[kMaxM], csm[kMaxM], fidx[kMaxM];
    bool F[kMaxM];
    // columns of the two states; zero rows marked, the others L1-normalised
    for (int r = 0; r < m; r++) {
        double a0 = A[r * D + i1], a1 = A[r * D + i2];
        const double f0 = fabs(a0), f1 = fabs(a1);
        wp[r] = p[r];
        if (f0 < kMuEps && f1 < kMuEps) {
            cmap[r] = csm[r] = fidx[r] = kZeroHpMarker;
            F[r] = false;
        } else {
            const double sa = f0 + f1;
            wa[2 * r] = a0 / sa;
            wa[2 * r + 1] = a1 / sa;
            wp[r] *= sa;
            F[r] = true;
        }
    }
    // +/- equal rows merged into the first of them
    int u = 0;
    for (int r = 0; r < m; r++) {
        if (!F[r]) continue;
        cmap[r] = u;
        csm[r] = 1;
        fidx[r] = r;
        wp2[u] = wp[r];
        for (int j = r + 1; j < m; j++) {
            if (!F[j]) continue;
            const bool pos = (fabs(wa[2 * r] - wa[2 * j]) < kCoalignEps) && (fabs(wa[2 * r + 1] - wa[2 * j + 1]) < kCoalignEps);
            if (pos) {
                cmap[j] = u;
                csm[j] = 1;
                F[j] = false;
                fidx[j] = r;
                wp2[u] += wp[j];
                continue;
            }
            const bool neg = (fabs(wa[2 * r] + wa[2 * j]) < kCoalignEps) && (fabs(wa[2 * r + 1] + wa[2 * j + 1]) < kCoalignEps);
            if (neg) {
                cmap[j] = u;
                csm[j] = -1;
                F[j] = false;
                fidx[j] = r;
                wp2[u] += wp[j];
            }
        }
        u++;
    }
    {
        int k = 0;
        for (int r = 0; r < m; r++)
            if (fidx[r] == r) {
                wa2[2 * k] = wa[2 * r];
                wa2[2 * k + 1] = wa[2 * r + 1];
                k++;
            }
    }
    // cell wall angles over a full turn, ascending (marg2d_get_cell_wall_angles)
    for (int r = 0; r < u; r++) {
        const double a0 = wa2[2 * r], a1 = wa2[2 * r + 1];
        double px, py;
        if (fabs(a0) < fabs(a1)) {
            px = 1.0;
            py = -a0 / a1;
        } else {
            px = -a1 / a0;
            py = 1.0;
        }
        double t1 = atan2(py, px);
        if (t1 < 0) t1 += kPi;
        th[2 * r] = t1;
        th[2 * r + 1] = t1 + kPi;
    }
    for (int a = 1; a < 2 * u; a++) {
        const double v = th[a];
        int b = a - 1;
        while (b >= 0 && th[b] > v) {
            th[b + 1] = th[b];
            b--;
        }
        th[b + 1] = v;
    }
    const double bx = P.b[(int64_t)i * D + i1], by = P.b[(int64_t)i * D + i2];
    const double g1im = bx - x0, g2im = by - y0;
    const bool check_g1 = fabs(g1im) < kIntegralGammaEps;
    const int top = 1 << (m - 1), rev = (1 << m) - 1;
    const int64_t go = P.gp_off[i];
    const int gn = P.gp_n[i];
    const int64_t co = (int64_t)i * MM;
    double total = 0.0;
    bool singular = false;
    double s1 = 0.0, c1 = 0.0;
    if (u > 0) {
        s1 = sin(th[0]);
        c1 = cos(th[0]);
    }
    for (int c = 0; c < u; c++) {
        // sign vector of the cell from its middle direction (marg2d_get_SVs)
        const double t = (th[c + 1] + th[c]) / 2.0;
        const double ptx = cos(t), pty = sin(t);
        uint32_t neg = 0;  // bit j: row j of the cut arrangement is negative in this cell
        for (int j = 0; j < u; j++)
            if (!(wa2[2 * j] * ptx + wa2[2 * j + 1] * pty > 0)) neg |= 1u << j;
        int enc = 0;
        for (int j = 0; j < m; j++) {
            if (cmap[j] == kZeroHpMarker) {
                double ip = 0.0;
                for (int k = 0; k < D; k++) ip += A[j * D + k] * bar_nu.v[k];
                if (ip < 0) enc |= 1 << j;
            } else {
                const int sv = ((neg >> cmap[j]) & 1u) ? -1 : 1;
                if (sv * csm[j] < 0) enc |= 1 << j;
            }
        }
        const cplx g = lookup_num(enc, top, rev, gkeys + go, gvals + go, gn);
        double g1re = 0.0, g2re = 0.0;
        for (int j = 0; j < u; j++) {
            const double sv = ((neg >> j) & 1u) ? -1.0 : 1.0;
            g1re -= (wa2[2 * j] * wp2[j]) * sv;
            g2re -= (wa2[2 * j + 1] * wp2[j]) * sv;
        }
        const double s2 = sin(th[c + 1]), c2 = cos(th[c + 1]);
        total += cpdf2d_cell(s1, c1, s2, c2, g1re, g2re, g1im, g2im, g, check_g1, singular);
        C.sn[co + i + c] = s1;
        C.cs[co + i + c] = c1;
        C.g1r[co + c] = g1re;
        C.g2r[co + c] = g2re;
        C.gc[co + c] = g;
        C.fc[co + c] = cell2d_parts(s1, c1, s2, c2, g1re, g2re, g);
        s1 = s2;
        c1 = c2;
    }
    C.sn[co + i + u] = s1;
    C.cs[co + i + u] = c1;
    C.mc[i] = u;
    C.b2[2 * i] = bx;
    C.b2[2 * i + 1] = by;
    direct[i] = total;
    if (singular) atomicOr(err, 1);
}

// Partial sums over one chunk of terms for a block of grid points (the cached cell
// integrals, marg2d_cached_eval_term_for_cpdf, cpdf_ndim.hpp:1654-1735).
// `idx`, when given, lists the terms to add (the kept ones of a relative system).
__global__ void __launch_bounds__(kCpdfPoints) k_cpdf2d_grid(Cpdf2D C, int n, const int* __restrict__ idx,
                                                             const double* __restrict__ px,
                                                             const double* __restrict__ py, int npts,
                                                             double* __restrict__ part, int* __restrict__ err) {
    const int chunk = blockIdx.y;
    const int t0 = chunk * kCpdfChunk;
    const int t1 = min(n, t0 + kCpdfChunk);
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npts) return;
    const double x = px[k], y = py[k];
    const int MM = C.MM;
    double acc = 0.0;
    bool singular = false;
    for (int tt = t0; tt < t1; tt++) {
        const int t = idx ? idx[tt] : tt;
        const int u = C.mc[t];
        const double g1im = C.b2[2 * t] - x, g2im = C.b2[2 * t + 1] - y;
        const bool check_g1 = fabs(g1im) < kIntegralGammaEps;
        const int64_t co = (int64_t)t * MM;
        const double* sn = C.sn + co + t;
        const double* cs = C.cs + co + t;
        double term = 0.0;
        for (int c = 0; c < u; c++) {
            const double s1 = sn[c], c1 = cs[c], s2 = sn[c + 1], c2 = cs[c + 1];
            if (check_g1 && fabs(C.g1r[co + c]) < kIntegralGammaEps) {
                term += cpdf2d_cell(s1, c1, s2, c2, C.g1r[co + c], C.g2r[co + c], g1im, g2im, C.gc[co + c], true,
                                    singular);
                continue;
            }
            // Re(n / (D1 D2)) from the cell's point-independent parts: one real division where
            // |D1 D2|^2 stays in range, the complex quotient otherwise
            const Cell2D f = C.fc[co + c];
            const double d1i = g1im * c1 + g2im * s1, d2i = g1im * c2 + g2im * s2;
            const cplx den = cmake(f.d1r * f.d2r - d1i * d2i, f.d1r * d2i + d1i * f.d2r);
            const double n2 = den.re * den.re + den.im * den.im;
            if (n2 > 1e-280 && n2 < 1e280) term += (f.nre * den.re + f.nim * den.im) / n2;
            else term += cdiv(cmake(f.nre, f.nim), den).re;
        }
        acc += term;
    }
    part[(int64_t)chunk * npts + k] = acc;
    if (singular) atomicOr(err, 1);
}


// ------------------------------------------------------------ relative system (two estimators)
// The relative system rel = secondary - primary, transformed to 2D by Trel (2 x d),
// as get_marg2d_relative_and_transformed_cpdf builds it (cauchy_prediction.hpp:497-650):
// one relative term per (primary term, secondary term) pair, A = [-A_p; A_s] Trel^T,
// p = [p_p; p_s], b = Trel (b_s - b_p).  Per pair: its moment contribution
// (eval_rel_sys_moments_for_term, 207-286), its cell data for the density
// (cache_rel_marg2d_term_for_cpdf, 306-443) and whether the term is approximated out
// (approx_out_rel_marg2d_term_for_cpdf, 445-471).
struct RsysArgs {
    double T[2 * kMaxD];   // Trel, row-major 2 x d
    double nu2[2];         // Trel bar_nu
    double nu[kMaxD];      // bar_nu (full state)
    double norm_fact;      // (1 / 2pi)^2 / (Re fz_p Re fz_s)
    double eps;            // term approximation threshold
};

__device__ inline double sgn_d(double x) { return (double)((x > 0) - (x < 0)); }

template <int D, int K>
__global__ void k_rsys_pairs(TermPool P, int np, const uint32_t* __restrict__ pkeys, const cplx* __restrict__ pvals,
                             TermPool S, int ns, const uint32_t* __restrict__ skeys, const cplx* __restrict__ svals,
                             RsysArgs R, Cpdf2D C, double* __restrict__ mom, uint8_t* __restrict__ keep,
                             int* __restrict__ err) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)np * ns) return;
    const int tp = (int)(q / ns), ts = (int)(q % ns);
    const int mp = P.m[tp], ms = S.m[ts];
    const int m = mp + ms;
    // K bounds the rows of every pair (the two pools' largest shapes); a pair above it
    // is reported, never evaluated
    if (m > K) {
        atomicOr(err, 2);
        keep[q] = 0;
        return;
    }
    const double* Ap = P.A + (int64_t)tp * P.MM * D;
    const double* As = S.A + (int64_t)ts * S.MM * D;
    // per-lane arrays, sized by K: the merged rows are compacted into a2 / p2 in place
    double a2[2 * K], p2[K], th[2 * K];
    int8_t cmap[K], csm[K], fidx[K];
    bool F[K];
    // A_2d = A_rel Trel^T (matmatmul with the second operand transposed), p_2d = p_rel
    for (int i = 0; i < m; i++) {
        const bool pr = i < mp;
        const double* ar = pr ? Ap + i * D : As + (i - mp) * D;
        const double sg = pr ? -1.0 : 1.0;
        for (int k = 0; k < 2; k++) {
            double acc = 0.0;
            for (int j = 0; j < D; j++) acc += (ar[j] * sg) * R.T[k * D + j];
            a2[2 * i + k] = acc;
        }
        p2[i] = pr ? P.p[(int64_t)tp * P.MM + i] : S.p[(int64_t)ts * S.MM + (i - mp)];
    }
    double b2[2];
    {
        double br[kMaxD];
        for (int j = 0; j < D; j++) br[j] = S.b[(int64_t)ts * D + j] - P.b[(int64_t)tp * D + j];
        for (int k = 0; k < 2; k++) {
            double acc = 0.0;
            for (int j = 0; j < D; j++) acc += R.T[k * D + j] * br[j];
            b2[k] = acc;
        }
    }
    // zero rows marked, the others L1-normalised in place; +/- equal rows merged
    for (int r = 0; r < m; r++) {
        const double f0 = fabs(a2[2 * r]), f1 = fabs(a2[2 * r + 1]);
        if (f0 < kMuEps && f1 < kMuEps) {
            cmap[r] = csm[r] = fidx[r] = kZeroHpMarker;
            F[r] = false;
        } else {
            const double sa = f0 + f1;
            a2[2 * r] /= sa;
            a2[2 * r + 1] /= sa;
            p2[r] *= sa;
            F[r] = true;
        }
    }
    const uint32_t* pk = pkeys + P.gp_off[tp];
    const cplx* pv = pvals + P.gp_off[tp];
    const int pn = P.gp_n[tp];
    const uint32_t* sk = skeys + S.gp_off[ts];
    const cplx* sv_ = svals + S.gp_off[ts];
    const int sn_ = S.gp_n[ts];
    // moment contribution: the cell of bar_nu (Trel bar_nu in 2D, bar_nu on zero rows);
    // taken before the merge, which leaves a2 and p2 as they are until the compaction
    {
        double ty0 = 0.0, ty1 = 0.0;
        int encp = 0, encs = 0;
        for (int i = 0; i < m; i++) {
            double ds;
            if (F[i]) {
                ds = a2[2 * i] * R.nu2[0] + a2[2 * i + 1] * R.nu2[1];
            } else {
                const bool pr = i < mp;
                const double* ar = pr ? Ap + i * D : As + (i - mp) * D;
                const double sg = pr ? -1.0 : 1.0;
                ds = 0.0;
                for (int j = 0; j < D; j++) ds += (ar[j] * sg) * R.nu[j];
            }
            const double sgi = sgn_d(ds);
            ty0 += a2[2 * i] * (p2[i] * sgi);
            ty1 += a2[2 * i + 1] * (p2[i] * sgi);
            if (sgi < 0) {
                if (i < mp) encp |= 1 << i;
                else encs |= 1 << (i - mp);
            }
        }
        const cplx gp = lookup_num(encp, 1 << (mp - 1), (1 << mp) - 1, pk, pv, pn);
        const cplx gs = lookup_num(encs, 1 << (ms - 1), (1 << ms) - 1, sk, sv_, sn_);
        const cplx g = cmul(gp, gs);
        const cplx y0 = cmake(-ty0, b2[0]), y1 = cmake(-ty1, b2[1]);
        // component c of pair q at mom[c * npairs + q] (column-major: contiguous column sums)
        const int64_t npq = (int64_t)np * ns;
        const cplx gy0 = cmul(g, y0), gy1 = cmul(g, y1);
        const cplx c00 = cmul(gy0, y0), c01 = cmul(gy0, y1), c10 = cmul(gy1, y0), c11 = cmul(gy1, y1);
        const double v[14] = {g.re,   g.im,   gy0.re, gy0.im, gy1.re, gy1.im, c00.re,
                              c00.im, c01.re, c01.im, c10.re, c10.im, c11.re, c11.im};
#pragma unroll
        for (int c = 0; c < 14; c++) mom[c * npq + q] = v[c];
    }
    int u = 0;
    for (int r = 0; r < m; r++) {
        if (!F[r]) continue;
        cmap[r] = u;
        csm[r] = 1;
        fidx[r] = r;
        double w = p2[r];
        for (int j = r + 1; j < m; j++) {
            if (!F[j]) continue;
            const bool pos = (fabs(a2[2 * r] - a2[2 * j]) < kCoalignEps) && (fabs(a2[2 * r + 1] - a2[2 * j + 1]) < kCoalignEps);
            if (pos) {
                cmap[j] = u;
                csm[j] = 1;
                F[j] = false;
                fidx[j] = r;
                w += p2[j];
                continue;
            }
            const bool neg = (fabs(a2[2 * r] + a2[2 * j]) < kCoalignEps) && (fabs(a2[2 * r + 1] + a2[2 * j + 1]) < kCoalignEps);
            if (neg) {
                cmap[j] = u;
                csm[j] = -1;
                F[j] = false;
                fidx[j] = r;
                w += p2[j];
            }
        }
        p2[u] = w;  // u <= r: no later read of the merge touches index u
        u++;
    }
    {
        int k = 0;
        for (int r = 0; r < m; r++)
            if (fidx[r] == r) {  // k <= r
                a2[2 * k] = a2[2 * r];
                a2[2 * k + 1] = a2[2 * r + 1];
                k++;
            }
    }
    // cell data over a half turn; rowof[i] is the compacted row whose line lies at th[i]
    int8_t rowof[2 * K];
    for (int r = 0; r < u; r++) {
        const double x0 = a2[2 * r], x1 = a2[2 * r + 1];
        double px, py;
        if (fabs(x0) < fabs(x1)) {
            px = 1.0;
            py = -x0 / x1;
        } else {
            px = -x1 / x0;
            py = 1.0;
        }
        double t1 = atan2(py, px);
        if (t1 < 0) t1 += kPi;
        th[2 * r] = t1;
        th[2 * r + 1] = t1 + kPi;
        rowof[2 * r] = rowof[2 * r + 1] = (int8_t)r;
    }
    for (int a = 1; a < 2 * u; a++) {
        const double v = th[a];
        const int8_t rv = rowof[a];
        int b = a - 1;
        while (b >= 0 && th[b] > v) {
            th[b + 1] = th[b];
            rowof[b + 1] = rowof[b];
            b--;
        }
        th[b + 1] = v;
        rowof[b + 1] = rv;
    }
    // The key bits of the original rows as a function of the compacted rows' sides: a zero
    // row's bit is fixed by bar_nu; row j on compacted row k = cmap[j] has its bit set when
    // k's side is negative and csm[j] = 1, or positive and csm[j] = -1.  With every side
    // positive the bits are enc_fix | enc_pos; turning row k negative flips the bits hit[k].
    uint32_t hit[K];
    for (int k = 0; k < u; k++) hit[k] = 0;
    int enc_fix = 0, enc_pos = 0;
    for (int j = 0; j < m; j++) {
        if (cmap[j] == kZeroHpMarker) {
            const bool pr = j < mp;
            const double* ar = pr ? Ap + j * D : As + (j - mp) * D;
            const double sg = pr ? -1.0 : 1.0;
            double ip = 0.0;
            for (int k = 0; k < D; k++) ip += (ar[k] * sg) * R.nu[k];
            if (ip < 0) enc_fix |= 1 << j;
        } else {
            hit[cmap[j]] |= 1u << j;
            if (csm[j] < 0) enc_pos |= 1 << j;
        }
    }
    // sides at the first cell's midpoint, then one crossing per cell: cell c lies between
    // th[c] and th[c + 1], so moving into it crosses the line at th[c] and flips its row.
    // Rows closer in angle than the merge tolerance were merged, so every midpoint is far
    // from every line compared with the rounding of its dot products, and the flips give
    // the sides a dot product at each midpoint gives.
    uint32_t negb = 0;
    int encv = 0;
    if (u > 0) {
        const double t = (th[1] + th[0]) / 2.0;
        const double ptx = cos(t), pty = sin(t);
        for (int j = 0; j < u; j++)
            if (!(a2[2 * j] * ptx + a2[2 * j + 1] * pty > 0)) {
                negb |= 1u << j;
                encv ^= (int)hit[j];
            }
    }
    // the g-integral terms a_j p_j (the products the sums below take, formed once)
    for (int j = 0; j < u; j++) {
        a2[2 * j] *= p2[j];
        a2[2 * j + 1] *= p2[j];
    }
    const int MM = C.MM;
    const int64_t co = q * MM;
    const int mpmask = (1 << mp) - 1;
    double total = 0.0;
    bool keep_term = false;
    double s_hi = 0.0, c_hi = 0.0;  // sin / cos of the previous cell's upper edge
    // a crossing usually changes the key of one estimator's term only: the other
    // term's g value is the previous cell's (same key, same table)
    int key_p = -1, key_s = -1;
    cplx gp = cmake(0, 0), gs = cmake(0, 0);
    for (int c = 0; c < u; c++) {
        if (c > 0) {
            const int k = rowof[c];
            negb ^= 1u << k;
            encv ^= (int)hit[k];
        }
        const int enc = enc_fix | (enc_pos ^ encv);
        if ((enc & mpmask) != key_p) {
            key_p = enc & mpmask;
            gp = lookup_num(key_p, 1 << (mp - 1), mpmask, pk, pv, pn);
        }
        if ((enc >> mp) != key_s) {
            key_s = enc >> mp;
            gs = lookup_num(key_s, 1 << (ms - 1), (1 << ms) - 1, sk, sv_, sn_);
        }
        const cplx g = cmul(gp, gs);
        double g1re = 0.0, g2re = 0.0;
        for (int j = 0; j < u; j++) {
            const double svd = ((negb >> j) & 1u) ? -1.0 : 1.0;
            g1re -= a2[2 * j] * svd;
            g2re -= a2[2 * j + 1] * svd;
        }
        // a cell's lower edge is the previous cell's upper edge: same argument, same value
        const double s1 = c > 0 ? s_hi : sin(th[c]), c1 = c > 0 ? c_hi : cos(th[c]);
        const double s2 = sin(th[c + 1]), c2 = cos(th[c + 1]);
        s_hi = s2;
        c_hi = c2;
        C.sn[co + q + c] = s1;
        C.cs[co + q + c] = c1;
        C.g1r[co + c] = g1re;
        C.g2r[co + c] = g2re;
        C.gc[co + c] = g;
        C.fc[co + c] = cell2d_parts(s1, c1, s2, c2, g1re, g2re, g);
        // the approximation's bound on the cell's contribution
        const double den1 = g1re * (g1re * c1 + g2re * s1);
        const double den2 = g1re * (g1re * c2 + g2re * s2);
        if (fabs(den1) < 1e-14 || fabs(den2) < 1e-14) keep_term = true;
        const double gr = hypot(g.re, g.im);
        total += fabs((gr * s2) / den2) + fabs((gr * s1) / den1);
    }
    if (u > 0) {  // th[u], the last cell's upper edge
        C.sn[co + q + u] = s_hi;
        C.cs[co + q + u] = c_hi;
    }
    C.mc[q] = u;
    C.b2[2 * q] = b2[0];
    C.b2[2 * q + 1] = b2[1];
    total *= (2 * R.norm_fact);
    keep[q] = (keep_term || !(fabs(total) < R.eps)) ? 1 : 0;
}

// Partial sums of the columns of a column-major n x cols array: block (b, c) adds the
// fixed slice b of column c; k_colsum_finish adds the slices in order.
constexpr int kColSlices = 256;
__global__ void k_colsum_part(const double* __restrict__ v, int64_t n, double* __restrict__ part) {
    __shared__ double sh[256];
    const int b = blockIdx.x, c = blockIdx.y, t = threadIdx.x;
    const int64_t per = (n + kColSlices - 1) / kColSlices;
    const int64_t lo = b * per, hi = min(n, lo + per);
    const double* col = v + c * n;
    double s = 0.0;
    for (int64_t i = lo + t; i < hi; i += blockDim.x) s += col[i];
    sh[t] = s;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if (t < h) sh[t] += sh[t + h];
        __syncthreads();
    }
    if (t == 0) part[c * kColSlices + b] = sh[0];
}

__global__ void k_colsum_finish(const double* __restrict__ part, int cols, double* __restrict__ out) {
    const int c = threadIdx.x;
    if (c >= cols) return;
    double s = 0.0;
    for (int b = 0; b < kColSlices; b++) s += part[c * kColSlices + b];
    out[c] = s;
}

// ------------------------------------------------------------------ merged-launch plumbing
// Byte copy of one staged payload or readback (8-byte words when both ends and the length
// allow).
__device__ __forceinline__ void k_copy_bytes(MG<256, 1> mg, const uint8_t* __restrict__ src, uint8_t* __restrict__ dst,
                                             int64_t bytes) {
    const int64_t t = (int64_t)mg.bx * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)mg.gx * blockDim.x;
    if ((((uintptr_t)src | (uintptr_t)dst | (uintptr_t)bytes) & 7) == 0) {
        const uint64_t* s = (const uint64_t*)src;
        uint64_t* d = (uint64_t*)dst;
        for (int64_t i = t; i < bytes / 8; i += stride) d[i] = s[i];
    } else {
        for (int64_t i = t; i < bytes; i += stride) dst[i] = src[i];
    }
}

// Stable sort of up to kSmallSort (key, value) pairs by key bits [begin_bit, end_bit) in one
// workgroup: the order a stable device radix sort over the same bits gives.  Slots past n
// hold all-ones keys, which sort after every real key of the same digits (stability keeps
// the real ones first), and are not written back.
constexpr int kSmallSortThreads = 256;
constexpr int kSmallSortItems = 16;
constexpr int kSmallSort = kSmallSortThreads * kSmallSortItems;
template <class K, class V>
__device__ __forceinline__ void k_small_sort(MG<kSmallSortThreads, 1> mg, const K* __restrict__ kin,
                                             K* __restrict__ kout, const V* __restrict__ vin, V* __restrict__ vout,
                                             int n, int begin_bit, int end_bit) {
    (void)mg;
    using Sort = hipcub::BlockRadixSort<K, kSmallSortThreads, kSmallSortItems, V>;
    __shared__ typename Sort::TempStorage tmp;
    K k[kSmallSortItems];
    V v[kSmallSortItems];
    const int base = threadIdx.x * kSmallSortItems;
#pragma unroll
    for (int j = 0; j < kSmallSortItems; j++) {
        const int i = base + j;
        k[j] = i < n ? kin[i] : (K)~(K)0;
        v[j] = i < n ? vin[i] : (V)0;
    }
    Sort(tmp).Sort(k, v, begin_bit, end_bit);
#pragma unroll
    for (int j = 0; j < kSmallSortItems; j++) {
        const int i = base + j;
        if (i < n) {
            kout[i] = k[j];
            vout[i] = v[j];
        }
    }
}

// Stable compaction of in[i] where flags[i] != 0 for up to kSmallScan items in one
// workgroup, with the selected count in *count (a device select's result).
__device__ __forceinline__ void k_small_select(MG<kSmallScanThreads, 1> mg, const uint32_t* __restrict__ in,
                                               const uint8_t* __restrict__ flags, uint32_t* __restrict__ out,
                                               int* __restrict__ count, int n) {
    (void)mg;
    const int t = threadIdx.x;
    const int per = (n + kSmallScanThreads - 1) / kSmallScanThreads;
    const int b = min(n, t * per), e = min(n, b + per);
    int c = 0;
    for (int i = b; i < e; i++) c += flags[i] != 0;
    int tot;
    int w = block_excl_scan(c, &tot);
    for (int i = b; i < e; i++)
        if (flags[i]) out[w++] = in[i];
    if (t == kSmallScanThreads - 1) count[0] = tot;
}

// Largest of n ints in one workgroup (INT_MIN when n = 0, as a device max
// reduction gives).
__device__ __forceinline__ void k_small_max(MG<1024, 1> mg, const int* __restrict__ in, int* __restrict__ out,
                                            int n) {
    (void)mg;
    __shared__ int part[1024];
    int m = INT_MIN;
    for (int i = threadIdx.x; i < n; i += blockDim.x) m = max(m, in[i]);
    part[threadIdx.x] = m;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) part[threadIdx.x] = max(part[threadIdx.x], part[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = part[0];
}

// Stable LSD radix sort of (key, value) pairs for inputs of kSmallSort .. kRadixMax items,
// one digit of up to 8 bits per pass, in chunks of kRadixChunk items per block, for many
// estimators at once: (1) each chunk's digit histogram, stored digit-major; (2) an exclusive
// scan over it (k_small_scan) gives every (digit, chunk) its first output slot; (3) each
// chunk sorts itself stably by the digit and writes every item to its digit's slot plus its
// rank among the chunk's items of that digit.  Equal digits keep chunk order and in-chunk
// order, so each pass is stable and the passes over bits [0, end) give the order of a stable
// sort over those bits.
constexpr int kRadixThreads = 256;
constexpr int kRadixItems = 16;
constexpr int kRadixChunk = kRadixThreads * kRadixItems;
constexpr int kRadixBins = 256;
constexpr int kRadixMax = kSmallScan / kRadixBins * kRadixChunk;  // the digit-major scan fits k_small_scan
template <class K>
__device__ __forceinline__ void k_radix_hist(MG<kRadixThreads, 1> mg, const K* __restrict__ keys, int n, int shift,
                                             int bits, int* __restrict__ hist, int nchunks) {
    __shared__ int h[kRadixBins];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int base = mg.bx * kRadixChunk;
    const K mask = (K)((1u << bits) - 1u);
    for (int j = threadIdx.x; j < kRadixChunk; j += kRadixThreads) {
        const int i = base + j;
        if (i < n) atomicAdd(&h[(int)((keys[i] >> shift) & mask)], 1);
    }
    __syncthreads();
    hist[threadIdx.x * nchunks + mg.bx] = h[threadIdx.x];
}
template <class K>
__device__ __forceinline__ void k_radix_scatter(MG<kRadixThreads, 1> mg, const K* __restrict__ kin,
                                                const uint32_t* __restrict__ vin, K* __restrict__ kout,
                                                uint32_t* __restrict__ vout, int n, int shift, int bits,
                                                const int* __restrict__ offs, int nchunks) {
    using Sort = hipcub::BlockRadixSort<K, kRadixThreads, kRadixItems, uint32_t>;
    __shared__ typename Sort::TempStorage tmp;
    __shared__ int first[kRadixBins];
    const int base = mg.bx * kRadixChunk;
    const int nreal = min(n - base, kRadixChunk);
    const K mask = (K)((1u << bits) - 1u);
    K k[kRadixItems];
    uint32_t v[kRadixItems];
    first[threadIdx.x] = 0;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kRadixItems; j++) {
        const int p = threadIdx.x * kRadixItems + j;
        const bool real = p < nreal;
        k[j] = real ? kin[base + p] : (K)~(K)0;
        v[j] = real ? vin[base + p] : 0u;
        if (real) atomicAdd(&first[(int)((k[j] >> shift) & mask)], 1);
    }
    __syncthreads();
    // exclusive scan of the chunk's digit counts: first[d] = items of smaller digits
    const int c = first[threadIdx.x];
    for (int o = 1; o < kRadixBins; o <<= 1) {
        const int add = (int)threadIdx.x >= o ? first[threadIdx.x - o] : 0;
        __syncthreads();
        first[threadIdx.x] += add;
        __syncthreads();
    }
    first[threadIdx.x] -= c;
    // padding keys are all ones: they sort after every real item of the last digit
    Sort(tmp).Sort(k, v, shift, shift + bits);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kRadixItems; j++) {
        const int p = threadIdx.x * kRadixItems + j;
        if (p < nreal) {
            const int dg = (int)((k[j] >> shift) & mask);
            const int dst = offs[dg * nchunks + mg.bx] + (p - first[dg]);
            kout[dst] = k[j];
            vout[dst] = v[j];
        }
    }
}

}  // namespace mce
