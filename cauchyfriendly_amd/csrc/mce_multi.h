// One kernel body, two launch forms.
//
// The update's kernels are written as __device__ bodies that take their block coordinates
// as an MG<LB, MW> value instead of reading blockIdx / gridDim.  k_one launches a body the
// ordinary way, with its arguments in the kernel-argument segment.  k_many launches it for
// many estimators at once: each estimator's launch becomes an entry (its grid and its
// arguments) in a device array, the entries' blocks are numbered one after another, and
// every block finds its entry from its block index and runs the body with that entry's
// arguments and local block coordinates.  The work of each block is the one its own launch
// would have done, so results do not depend on the form.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mce {

// Block coordinates of a body: bx, by within its own grid of gx x (blocks / gx) blocks.
// LB / MW are the body's launch bounds (threads per block, minimum waves per SIMD).
template <int LB, int MW>
struct MG {
    unsigned bx, by, gx;
};

// Trivially copyable argument pack (byte-copied into the entry arrays).
template <class... A>
struct Pack;
template <>
struct Pack<> {};
template <class H, class... T>
struct Pack<H, T...> {
    H h;
    Pack<T...> t;
};

template <class... A>
struct PackOf;
template <>
struct PackOf<> {
    __host__ __device__ static Pack<> make() { return Pack<>{}; }
};
template <class H, class... T>
struct PackOf<H, T...> {
    template <class AH, class... AT>
    __host__ __device__ static Pack<H, T...> make(AH&& ah, AT&&... at) {
        return Pack<H, T...>{(H)ah, PackOf<T...>::make(static_cast<AT&&>(at)...)};
    }
};

template <class F, class... Got>
__device__ __forceinline__ void papply(const F& f, const Pack<>&, Got... g) {
    f(g...);
}
template <class F, class H, class... T, class... Got>
__device__ __forceinline__ void papply(const F& f, const Pack<H, T...>& p, Got... g) {
    papply(f, p.t, g..., p.h);
}

// One estimator's launch inside a merged one: its grid (gx x gy blocks) and arguments.
template <class... A>
struct MEntry {
    unsigned gx, gy;
    Pack<A...> a;
};

// Index of the entry that block `bid` belongs to: the last i with first[i] <= bid, where
// first[] (n values, strictly increasing, first[0] = 0) numbers each entry's first block.
// Each wavefront searches on its own: its lanes test up to 64 evenly spaced candidates per
// round, so n entries take ceil(log64 n) dependent loads.  Every lane of the wavefront
// must be active (it is: this runs before the body).
__device__ __forceinline__ int find_entry(const unsigned* __restrict__ first, int n, unsigned bid) {
    const int lanes = blockDim.x < 64 ? (int)blockDim.x : 64;
    const int lane = (int)(threadIdx.x & 63);
    int lo = 0, span = n;
    if (lanes == 1) {  // a one-thread block: plain bisection
        int hi = n - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (first[mid] <= bid)
                lo = mid;
            else
                hi = mid - 1;
        }
        return lo;
    }
    // with at least two lanes every round shrinks the span: step = ceil(span / lanes) < span
    while (span > 1) {
        const int step = (span + lanes - 1) / lanes;
        const int k = lane * step;
        const bool ok = k < span && first[lo + k] <= bid;
        const unsigned long long m = __ballot(ok);
        const int L = 63 - __clzll((long long)m);
        lo += L * step;
        span = min(step, span - L * step);
    }
    return lo;
}

template <int LB, int MW, auto F, class... A>
__global__ __launch_bounds__(LB, MW) void k_one(A... a) {
    F(MG<LB, MW>{blockIdx.x, blockIdx.y, gridDim.x}, a...);
}

template <int LB, int MW, auto F>
struct CallBody {
    MG<LB, MW> g;
    template <class... A>
    __device__ __forceinline__ void operator()(A... a) const {
        F(g, a...);
    }
};

template <int LB, int MW, auto F, class... A>
__global__ __launch_bounds__(LB, MW) void k_many(const unsigned* __restrict__ first,
                                                 const MEntry<A...>* __restrict__ ents, int n) {
    const unsigned bid = blockIdx.x;
    const int i = __builtin_amdgcn_readfirstlane(find_entry(first, n, bid));
    const unsigned l = bid - first[i];
    const unsigned gx = ents[i].gx;
    const CallBody<LB, MW, F> call{MG<LB, MW>{l % gx, l / gx, gx}};
    papply(call, ents[i].a);
}

// The two launch forms of body F and its entry type.
template <auto F>
struct Kern;
template <int LB, int MW, class... A, void (*F)(MG<LB, MW>, A...)>
struct Kern<F> {
    using Entry = MEntry<A...>;
    static_assert(__is_trivially_copyable(Entry), "merged-launch entries are copied as bytes");
    static constexpr int kMaxThreads = LB;
    template <class... Args>
    static Entry entry(unsigned gx, unsigned gy, Args&&... args) {
        return Entry{gx, gy, PackOf<A...>::make(static_cast<Args&&>(args)...)};
    }
    static const void* one() { return (const void*)&k_one<LB, MW, F, A...>; }
    static const void* many() { return (const void*)&k_many<LB, MW, F, A...>; }
    template <class... Args>
    static void launch(dim3 g, dim3 b, size_t lds, hipStream_t s, Args&&... args) {
        hipLaunchKernelGGL((k_one<LB, MW, F, A...>), g, b, lds, s, (A)args...);
    }
};

}  // namespace mce
