// Structure-of-arrays term storage in HBM.
//
// A TermPool replaces the reference's per-term CauchyTerm records and their
// page allocators (include/cauchy_term.hpp:46-68, include/cauchy_util.hpp:578-1461).
// Each field is its own array of `cap` entries; per-hyperplane fields have a
// fixed row stride MM (the estimator's largest shape), so a term of any shape m
// lives at the same index in every array.  Terms are kept bucketed by shape:
// survivors of a step are written back shape-ascending, and the measurement
// update's output slots are re-ordered by an index permutation (`perm`) instead
// of being moved.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mce {

enum TermFlags : uint8_t {
    kValid = 1,     // slot holds a term (H-orthogonal hyperplanes leave holes)
    kIsNew = 2,     // term was created by this measurement update (is_new_child)
    kHasCmap = 4,   // measurement-update coalignment merged hyperplanes (c_map != NULL)
};

struct TermPool {
    int64_t cap = 0;
    int MM = 0;      // row stride (largest shape)
    int d = 0;
    double* A = nullptr;      // [cap][MM*d]
    double* p = nullptr;      // [cap][MM]
    double* q = nullptr;      // [cap][MM]
    double* b = nullptr;      // [cap][d]
    double* c = nullptr;      // [cap]  c_val
    double* dv = nullptr;     // [cap]  d_val
    int* m = nullptr;         // [cap]  hyperplane count
    int* phc = nullptr;       // [cap]  hyperplane count of the parent g-table
    int* pbc = nullptr;       // [cap]  hyperplane count of the parent at the update
    int* z = nullptr;         // [cap]  index of the child's defining hyperplane
    int* lhp = nullptr;       // [cap]  enc_lhp
    uint32_t* hflag = nullptr;  // [cap] H-orthogonal hyperplane mask
    uint8_t* flags = nullptr;   // [cap] TermFlags
    uint8_t* cmap = nullptr;    // [cap][MM]
    int8_t* csmap = nullptr;    // [cap][MM]
    int64_t* gp_off = nullptr;  // [cap] parent g-table offset in the previous table arena
    int64_t* rk_off = nullptr;  // [cap] offset of the parent g-table's rank structure (rank arena words)
    int* gp_n = nullptr;        // [cap] parent g-table size (cells_gtable_p)
    int64_t* bt_off = nullptr;  // [cap] B-table (sorted) offset in the B arena
    int* bt_n = nullptr;        // [cap] B-table size (cells_gtable)
    int* bt_sgn = nullptr;      // [cap] xor turning the stored B-table into the post-update one
    int* tid = nullptr;         // [cap] reference thread that owns the term this step
    uint8_t* mode = nullptr;    // [cap] B-table construction mode after time propagation
};

}  // namespace mce
