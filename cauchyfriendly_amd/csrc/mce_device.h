// Device-side numerics shared by the MCE kernels (gfx950, FP64 throughout).
//
// Every decision-bearing expression keeps the operation order of the reference
// CPU code it replaces (cited per function) and the library is compiled with
// -ffp-contract=off, so signs, coalignment tests and reduction decisions come
// out bit-identical to the reference.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mce {

constexpr int kMaxD = 8;    // largest state dimension the kernels are instantiated for
constexpr int kMaxM = 31;   // largest hyperplane count (the reference refuses > 31)

// include/cauchy_constants.hpp:25-73
constexpr double kCoalignEps = 1e-8;
constexpr double kMuEps = 1e-10;
constexpr double kApproxEps = 1e-15;
constexpr double kPluEps = 1e-15;
constexpr double kCondEps = 1e12;
constexpr double kCovEigenvalueTol = -1e-5;  // COV_EIGENVALUE_TOLERANCE, cauchy_constants.hpp:102
constexpr double kReductionEps = 1e-8;
constexpr double kPi = 3.14159265358979323846;
constexpr double kInvTwoPi = 1.0 / (2.0 * 3.14159265358979323846);

struct cplx {
    double re, im;
};

__device__ __host__ inline cplx cmake(double r, double i) {
    cplx v;
    v.re = r;
    v.im = i;
    return v;
}
__device__ __host__ inline cplx cconj(cplx a) { return cmake(a.re, -a.im); }
__device__ __host__ inline cplx cadd(cplx a, cplx b) { return cmake(a.re + b.re, a.im + b.im); }
__device__ __host__ inline cplx csub(cplx a, cplx b) { return cmake(a.re - b.re, a.im - b.im); }
__device__ __host__ inline cplx cmul(cplx a, cplx b) {
    return cmake((a.re * b.re) - (a.im * b.im), (a.re * b.im) + (a.im * b.re));
}
__device__ __host__ inline cplx cscale(cplx a, double s) { return cmake(a.re * s, a.im * s); }

// Complex quotient by Smith's method (R. L. Smith, "Algorithm 116: Complex
// division", CACM 5(8), 1962).  This is the scheme the reference's compiler
// runtime uses for `a / b` on GCC complex doubles, so quotients match it bit for bit.
// Both of Smith's branches are evaluated with selects (no divergence): every
// quotient, product and sum is the one the taken branch computes.
__device__ __host__ inline cplx cdiv(cplx a, cplx b) {
    const double c = b.re, d = b.im;
    const bool small_c = fabs(c) < fabs(d);
    const double p = small_c ? c : d, q = small_c ? d : c;
    const double r = p / q;
    const double den = (p * r) + q;
    const double pr = a.re * r, pi = a.im * r;
    const double xn = small_c ? (pr + a.im) : (pi + a.re);
    const double yn = small_c ? (pi - a.re) : (a.im - pr);
    return cmake(xn / den, yn / den);
}

// Binary search over an ascending key array, gtable.hpp:283-301.
__device__ inline int bsearch_u32(const uint32_t* keys, uint32_t target, int size) {
    int lo = 0, hi = size - 1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const uint32_t k = keys[mid];
        if (k == target) return mid;
        if (k > target) hi = mid - 1;
        else lo = mid + 1;
    }
    return -1;
}

// Half-storage g numerator lookup in a parent g-table, eval_gs.hpp:94-153.
__device__ inline cplx lookup_num(int enc, int top_bit, int rev_mask, const uint32_t* keys, const cplx* vals,
                                  int size) {
    if (enc & top_bit) {
        const int i = bsearch_u32(keys, (uint32_t)(rev_mask ^ enc), size);
        return i < 0 ? cmake(0, 0) : cconj(vals[i]);
    }
    const int i = bsearch_u32(keys, (uint32_t)enc, size);
    return i < 0 ? cmake(0, 0) : vals[i];
}

// Coalignment gate over d coordinates, cauchy_term.hpp:494-505.
__device__ inline void coalign_gates(const double* r, const double* c, int d, bool& pos, bool& neg) {
    pos = true;
    neg = true;
    for (int l = 0; l < d; l++) {
        if (pos) pos &= fabs(r[l] - c[l]) < kCoalignEps;
        if (neg) neg &= fabs(r[l] + c[l]) < kCoalignEps;
        if (!(pos || neg)) break;
    }
}

// Orientation mask between two +/- equal arrangements: per row, sign of the
// first element non-negligible in both, cell_enumeration.hpp:586-599.
__device__ inline int orientation_mask(const double* Ai, const double* Aj, int m, int d) {
    int sig = 0;
    for (int k = 0; k < m; k++) {
        int l = 0;
        const int kd = k * d;
        while (kd + l < m * d && ((fabs(Ai[kd + l]) < kReductionEps) || (fabs(Aj[kd + l]) < kReductionEps))) l++;
        if (kd + l >= m * d) continue;
        if ((Ai[kd + l] * Aj[kd + l]) < 0) sig |= (1 << k);
    }
    return sig;
}

// Order-preserving map of an FP64 value to an unsigned 64-bit radix key.
__device__ inline uint64_t f64_radix_key(double x) {
    uint64_t u = (uint64_t)__double_as_longlong(x);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

}  // namespace mce
