// Host side of the MI355X MCE engine: term pools in HBM, the per-update kernel
// sequence, and the C ABI of include/cauchy_mi355x.h.
//
// One Engine is one CauchyEstimator (include/cauchy_estimator.hpp:47-1432).  Its
// terms never leave the GPU; the host only sequences kernels, reads back a few
// counters (slot totals, per-shape counts, reduction rounds) and finalises the
// 1 + d + d^2 moment sums.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <coroutine>
#include <functional>
#include <initializer_list>
#include <mutex>
#include <optional>
#include <cstdio>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/cauchy_mi355x.h"
#include "mce_coro.h"
#include "mce_host_linalg.h"
#include "mce_kernels.h"
#include "mce_tangent.h"

using namespace mce;

namespace {

// per host thread, so that estimators stepped concurrently keep their own message
thread_local std::string g_last_error;

// Lanes per parent in the TP cell enumerator: 64, 128 or 256 (the kernel's launch bound);
// any other value keeps the default, one wavefront: most parents have far fewer vertex
// combinations than 256 (C(7, 5) = 21), and one wavefront per parent measured faster on
// MI355X (TP B-tables 4.50 against 4.62 ms per 5-state window).
int enum_threads_from_env() {
    const char* v = getenv("MCE_ENUM_THREADS");
    const int t = v ? atoi(v) : 0;
    return (t == 64 || t == 128 || t == 256) ? t : 64;
}

#define HIPCHK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess)                                                                       \
            throw std::runtime_error(std::string("HIP error '") + hipGetErrorString(e_) + "' at " + \
                                     #x);                                                           \
    } while (0)

int ceil_div(int64_t a, int b) { return (int)((a + b - 1) / b); }

// LDS of one workgroup on gfx950 (static and dynamic together).
constexpr size_t kLdsPerGroup = 160 * 1024;

// MCE_CHECK_LAUNCH=1: every launch is followed by an error check and a synchronisation of
// its stream, so a failing dispatch is named by its kernel.
bool check_launches() {
    static const bool on = [] {
        const char* v = getenv("MCE_CHECK_LAUNCH");
        return v && v[0] && v[0] != '0';
    }();
    return on;
}

// Validates a launch on the host before anything reaches the queue.  An empty grid has
// nothing to do and is skipped, so no dispatch packet is ever formed for it; a block or an
// LDS request outside the device's limits throws.
bool launch_ok(const dim3& g, const dim3& b, size_t lds, const char* name) {
    if (g.x == 0 || g.y == 0 || g.z == 0) return false;
    const uint64_t threads = (uint64_t)b.x * b.y * b.z;
    if (threads == 0 || threads > 1024 || g.y > 65535 || g.z > 65535 || g.x > 0x7fffffffu)
        throw std::runtime_error(std::string("launch of ") + name + ": grid " + std::to_string(g.x) + "x" +
                                 std::to_string(g.y) + "x" + std::to_string(g.z) + ", block " +
                                 std::to_string(threads) + " outside the device's limits");
    if (lds > kLdsPerGroup)
        throw std::runtime_error(std::string("launch of ") + name + ": " + std::to_string(lds) +
                                 " B of LDS requested, the device has " + std::to_string(kLdsPerGroup));
    return true;
}

void launch_done(hipStream_t s, const char* name) {
    if (!check_launches()) return;
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess)
        throw std::runtime_error(std::string("HIP error '") + hipGetErrorString(e) + "' in " + name);
}

#define MCE_LAUNCH(K_, G_, B_, L_, S_, ...)                      \
    do {                                                         \
        const dim3 g_ = (G_), b_ = (B_);                         \
        const size_t l_ = (L_);                                  \
        if (launch_ok(g_, b_, l_, #K_)) {                        \
            hipLaunchKernelGGL(K_, g_, b_, l_, S_, __VA_ARGS__); \
            launch_done(S_, #K_);                                \
        }                                                        \
    } while (0)

// Lets a kernel take more than the default 64 KiB of dynamic LDS.  The attribute belongs to
// the kernel symbol, which every engine shares; it is raised once per kernel, to everything
// its static LDS leaves, under a lock, and never again: host threads that launch the same
// kernel concurrently never see it change under them.
void allow_lds(const void* fn, size_t bytes) {
    if (bytes <= 64 * 1024) return;
    static std::mutex mu;
    static std::vector<std::pair<const void*, size_t>> raised;  // kernel, its static LDS
    std::lock_guard<std::mutex> lk(mu);
    size_t stat = 0;
    auto it = std::find_if(raised.begin(), raised.end(), [&](const auto& r) { return r.first == fn; });
    if (it != raised.end()) {
        stat = it->second;
    } else {
        hipFuncAttributes at;
        HIPCHK(hipFuncGetAttributes(&at, fn));
        stat = at.sharedSizeBytes;
    }
    if (stat + bytes > kLdsPerGroup)
        throw std::runtime_error("kernel needs " + std::to_string(bytes) + " B of dynamic LDS beside " +
                                 std::to_string(stat) + " B static; the device has " + std::to_string(kLdsPerGroup));
    if (it != raised.end()) return;
    HIPCHK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(kLdsPerGroup - stat)));
    raised.emplace_back(fn, stat);
}

// ---------------------------------------------------------------- merged launches
// What an engine stepped by a batch scheduler hands over instead of issuing it on a stream:
// kernel launches (the body's merged-launch symbol, its grid, and the entry holding its
// arguments), host-side stream work (library calls on large inputs, run in order at the
// scheduler's flush), uploads (bytes staged at once, copied in order), and the device
// reads that suspend its update until the scheduler has the data on the host.
struct BatchReq {
    enum Kind : uint8_t { kKernel, kHost, kUpload };
    Kind kind = kKernel;
    const void* fn = nullptr;  // k_many symbol
    const char* name = "";
    uint32_t gx = 0, gy = 0, threads = 0;
    size_t lds = 0;
    uint32_t esize = 0, ealign = 0;
    size_t off = 0;            // entry (kernel) or payload (upload) in BatchSlot::bytes
    void* dst = nullptr;       // upload destination
    size_t nbytes = 0;         // upload length
    std::function<void(hipStream_t)> host;
};
struct BatchRead {
    const void* src;
    size_t bytes;
    void* dst;
};
struct BatchSlot {
    std::vector<BatchReq> chain;
    std::vector<uint8_t> bytes;
    std::vector<BatchRead> reads;
    std::coroutine_handle<> waiting;
    // when the scheduler profiles: device counter of k_group_eval's algorithmic bytes
    unsigned long long* geval_bytes = nullptr;
    size_t append(const void* p, size_t n, size_t align) {
        size_t off = (bytes.size() + align - 1) / align * align;
        bytes.resize(off + n);
        memcpy(bytes.data() + off, p, n);
        return off;
    }
    template <class E>
    void push_kernel(const void* fn, const char* name, const dim3& g, const dim3& b, size_t lds, const E& e) {
        BatchReq r;
        r.kind = BatchReq::kKernel;
        r.fn = fn;
        r.name = name;
        r.gx = g.x;
        r.gy = g.y;
        r.threads = b.x * b.y * b.z;
        r.lds = lds;
        r.esize = sizeof(E);
        r.ealign = alignof(E);
        r.off = append(&e, sizeof(E), alignof(E));
        chain.push_back(std::move(r));
    }
    void push_upload(void* dst, const void* src, size_t n) {
        if (n == 0) return;
        BatchReq r;
        r.kind = BatchReq::kUpload;
        r.name = "upload";
        r.dst = dst;
        r.nbytes = n;
        r.off = append(src, n, 16);
        chain.push_back(std::move(r));
    }
    void push_host(std::function<void(hipStream_t)> f, const char* name) {
        BatchReq r;
        r.kind = BatchReq::kHost;
        r.name = name;
        r.host = std::move(f);
        chain.push_back(std::move(r));
    }
    void clear_chain() {
        chain.clear();
        bytes.clear();
    }
};

// Launches body F on grid g x block b: at once on stream s, or, when bs is set (a batch
// scheduler steps the engine), as an entry of the scheduler's next merged launch.
template <auto F, class... Args>
void mce_run(BatchSlot* bs, const dim3& g, const dim3& b, size_t lds, hipStream_t s, const char* name,
             Args&&... args) {
    using K = Kern<F>;
    if (!launch_ok(g, b, lds, name)) return;
    if (g.z != 1) throw std::runtime_error(std::string("launch of ") + name + ": 3D grids are not supported");
    if (bs) {
        bs->push_kernel(K::many(), name, g, b, lds, K::entry(g.x, g.y, std::forward<Args>(args)...));
        return;
    }
    allow_lds(K::one(), lds);
    K::launch(g, b, lds, s, std::forward<Args>(args)...);
    launch_done(s, name);
}

// Launch sites outside an engine have no scheduler.
inline BatchSlot* bslot() { return nullptr; }

#define MCE_RUN(K_, G_, B_, L_, S_, ...) mce_run<&K_>(bslot(), (G_), (B_), (L_), (S_), #K_, __VA_ARGS__)

// Device blocks of up to kMaxBlock bytes come from a process-wide cache with one free list
// per power-of-two size.  An estimator that outgrows its buffers, or a new estimator of a
// batch, takes blocks that others gave back instead of calling hipMalloc, which holds a
// device-wide lock and slows down as live allocations multiply (thousands of small
// estimators stepped from many host threads otherwise spend most of their time there).
// Blocks of up to kSlabMax bytes are carved from slabs of several hundred blocks and stay
// with the process; larger ones are allocated one by one and freed when the cache holds more
// than kMaxCached.  A block is given back only when no queued work can touch it: after the
// streams of the engine that used it were synchronised.  Larger blocks bypass the cache.
class BlockCache {
   public:
    static constexpr size_t kMinBlock = 256;
    static constexpr size_t kSlabMax = size_t(4) << 20;
    static constexpr size_t kMaxBlock = size_t(256) << 20;
    static constexpr size_t kMaxCached = size_t(16) << 30;  // unslabbed blocks beyond this are freed

    static size_t block_size(size_t bytes) {
        if (bytes > kMaxBlock) return bytes;
        size_t s = kMinBlock;
        while (s < bytes) s <<= 1;
        return s;
    }
    // Bytes of the slab that blocks of `bytes` (a block_size() value <= kSlabMax) are cut from.
    static size_t slab_size(size_t bytes) {
        return std::min<size_t>(size_t(256) << 20, std::max<size_t>(size_t(4) << 20, bytes * 512));
    }
    // `bytes` is a block_size() value.
    void* get(size_t bytes) {
        if (bytes <= kMaxBlock) {
            std::lock_guard<std::mutex> lk(mu_);
            auto& l = free_[cls(bytes)];
            if (l.empty() && bytes <= kSlabMax) {
                const size_t sb = slab_size(bytes);
                char* slab = (char*)raw_alloc(sb, true);
                slab_bytes_ += sb;
                for (size_t off = sb; off >= bytes; off -= bytes) l.push_back(slab + off - bytes);
            }
            if (!l.empty()) {
                void* p = l.back();
                l.pop_back();
                if (bytes > kSlabMax) cached_ -= bytes;
                return p;
            }
        }
        return raw_alloc(bytes, false);
    }
    // The caller guarantees that no queued work reads or writes the block.
    void put(void* p, size_t bytes) {
        if (!p) return;
        if (bytes <= kMaxBlock) {
            std::lock_guard<std::mutex> lk(mu_);
            if (bytes <= kSlabMax) {
                free_[cls(bytes)].push_back(p);
                return;
            }
            if (cached_ + bytes <= kMaxCached) {
                free_[cls(bytes)].push_back(p);
                cached_ += bytes;
                return;
            }
        }
        (void)hipFree(p);
    }

   private:
    // `locked`: the caller holds mu_.
    void* raw_alloc(size_t bytes, bool locked) {
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, bytes);
        if (e == hipErrorOutOfMemory) {
            if (locked) {
                trim_unslabbed();
            } else {
                std::lock_guard<std::mutex> lk(mu_);
                trim_unslabbed();
            }
            e = hipMalloc(&p, bytes);
        }
        if (e != hipSuccess)
            throw std::runtime_error(std::string("HIP error '") + hipGetErrorString(e) + "' at hipMalloc(" +
                                     std::to_string(bytes) + " bytes)");
        return p;
    }
    // Frees the cached blocks that were allocated one by one (hipFree synchronises the
    // device); the lists it empties belong to classes above kSlabMax.  Needs mu_.
    void trim_unslabbed() {
        for (int c = cls(kSlabMax) + 1; c < 64; c++) {
            for (void* p : free_[c]) (void)hipFree(p);
            free_[c].clear();
        }
        cached_ = 0;
    }
    static int cls(size_t bytes) { return __builtin_ctzll(bytes); }
    std::mutex mu_;
    std::vector<void*> free_[64];
    size_t cached_ = 0, slab_bytes_ = 0;
};

BlockCache& block_cache() {
    static BlockCache* c = new BlockCache();  // never destroyed: engines may outlive static teardown
    return *c;
}

// Allocations outgrown while an engine works on this host thread.  The engine gives them
// back once its streams are idle: blocks of the cache's sizes at the end of the call, larger
// ones (hipFree synchronises the whole device, which would stall every other estimator
// stepped concurrently) when they add up to more than a bound or the engine goes.  Outside an
// engine they are given back at once, after a device synchronisation.
struct Graveyard {
    std::vector<std::pair<void*, size_t>> blocks;
    size_t bytes = 0, big_bytes = 0;
    void add(void* p, size_t b) {
        blocks.emplace_back(p, b);
        bytes += b;
        if (b > BlockCache::kMaxBlock) big_bytes += b;
    }
    // Gives back the cache-sized blocks, and the larger ones too if `all`.
    void release(bool all) {
        std::vector<std::pair<void*, size_t>> keep;
        for (auto& [p, b] : blocks) {
            if (all || b <= BlockCache::kMaxBlock)
                block_cache().put(p, b);
            else
                keep.emplace_back(p, b);
        }
        blocks.swap(keep);
        bytes = big_bytes = 0;
        for (auto& [p, b] : blocks) bytes += b, big_bytes += b;
    }
};
thread_local Graveyard* t_graveyard = nullptr;

struct GraveyardScope {
    Graveyard* saved;
    explicit GraveyardScope(Graveyard* g) : saved(t_graveyard) { t_graveyard = g; }
    ~GraveyardScope() { t_graveyard = saved; }
};

// Grow-only device buffer; growing discards the contents.
// Grow-only device buffer; growing discards the contents.  Its block comes from the block
// cache; the owner's streams are idle whenever it is destroyed (every engine call ends with
// its streams synchronised), so the block goes straight back.
template <class T>
struct DBuf {
    T* ptr = nullptr;
    size_t cap = 0, bytes = 0;
    void reserve(size_t n) {
        if (n <= cap) return;
        // doubling, except for buffers beyond 1 GiB (the g-tables of a window's largest
        // updates reach tens of GiB): those get an eighth over what is asked
        size_t nn = n * sizeof(T) > (size_t(1) << 30) ? n + n / 8 : std::max(n, 2 * cap);
        nn = std::max<size_t>(nn, 64);
        if (ptr) {
            if (t_graveyard) {
                // kernels still in flight may read the old allocation: it stays valid
                t_graveyard->add(ptr, bytes);
            } else {
                HIPCHK(hipDeviceSynchronize());
                block_cache().put(ptr, bytes);
            }
            ptr = nullptr;
            cap = bytes = 0;
        }
        const size_t b = BlockCache::block_size(nn * sizeof(T));
        ptr = (T*)block_cache().get(b);
        bytes = b;
        cap = b / sizeof(T);
    }
    void release() {
        block_cache().put(ptr, bytes);
        ptr = nullptr;
        cap = bytes = 0;
    }
    ~DBuf() { release(); }
};

struct PoolStore {
    int MM = 0, d = 0;
    int64_t cap = 0;
    DBuf<double> A, p, q, b, c, dv;
    DBuf<int> m, phc, pbc, z, lhp, gp_n, bt_n, bt_sgn, tid;
    DBuf<uint32_t> hflag;
    DBuf<uint8_t> flags, cmap, mode;
    DBuf<int8_t> csmap;
    DBuf<int64_t> gp_off, bt_off, rk_off;
    void reserve(int64_t n) {
        if (n <= cap) return;
        int64_t nn = std::max<int64_t>(n, cap + cap / 2);
        A.reserve((size_t)nn * MM * d);
        p.reserve((size_t)nn * MM);
        q.reserve((size_t)nn * MM);
        b.reserve((size_t)nn * d);
        c.reserve(nn);
        dv.reserve(nn);
        m.reserve(nn);
        phc.reserve(nn);
        pbc.reserve(nn);
        z.reserve(nn);
        lhp.reserve(nn);
        gp_n.reserve(nn);
        bt_n.reserve(nn);
        bt_sgn.reserve(nn);
        tid.reserve(nn);
        hflag.reserve(nn);
        flags.reserve(nn);
        cmap.reserve((size_t)nn * MM);
        mode.reserve(nn);
        csmap.reserve((size_t)nn * MM);
        gp_off.reserve(nn);
        bt_off.reserve(nn);
        rk_off.reserve(nn);
        cap = nn;
    }
    TermPool view() const {
        TermPool t;
        t.cap = cap;
        t.MM = MM;
        t.d = d;
        t.A = A.ptr;
        t.p = p.ptr;
        t.q = q.ptr;
        t.b = b.ptr;
        t.c = c.ptr;
        t.dv = dv.ptr;
        t.m = m.ptr;
        t.phc = phc.ptr;
        t.pbc = pbc.ptr;
        t.z = z.ptr;
        t.lhp = lhp.ptr;
        t.hflag = hflag.ptr;
        t.flags = flags.ptr;
        t.cmap = cmap.ptr;
        t.csmap = csmap.ptr;
        t.gp_off = gp_off.ptr;
        t.rk_off = rk_off.ptr;
        t.gp_n = gp_n.ptr;
        t.bt_off = bt_off.ptr;
        t.bt_n = bt_n.ptr;
        t.bt_sgn = bt_sgn.ptr;
        t.tid = tid.ptr;
        t.mode = mode.ptr;
        return t;
    }
};

// Tangents of a term pool in tangent mode (P directions), grown with it; growing discards
// the contents, as for the pool.
struct TanStore {
    int P = 0, MM = 0, d = 0;
    int64_t cap = 0;
    DBuf<double> A, p, q, b, c, dv;
    void reserve(int64_t n) {
        if (n <= cap) return;
        const int64_t nn = std::max<int64_t>(n, cap + cap / 2);
        A.reserve((size_t)nn * P * MM * d);
        p.reserve((size_t)nn * P * MM);
        q.reserve((size_t)nn * P * MM);
        b.reserve((size_t)nn * P * d);
        c.reserve((size_t)nn * P);
        dv.reserve((size_t)nn * P);
        cap = nn;
    }
    TanPool view() const {
        TanPool t;
        t.A = A.ptr;
        t.p = p.ptr;
        t.q = q.ptr;
        t.b = b.ptr;
        t.c = c.ptr;
        t.dv = dv.ptr;
        t.P = P;
        t.MM = MM;
        t.d = d;
        return t;
    }
};

// ---------------------------------------------------------------- host numerics
unsigned long long binom(int n, int k) {
    unsigned long long r = 1;
    if (k > n - k) k = n - k;
    for (int i = 0; i < k; ++i) {
        r *= (n - i);
        r /= (i + 1);
    }
    return r;
}
int cells_central(int m, int d) {
    if (m < d) return 1 << m;
    unsigned long long s = 0;
    for (int i = 0; i < d; i++) s += binom(m - 1, i);
    return (int)(2 * s);
}

double rand_u() { return ((double)rand() + 1.0) / ((double)RAND_MAX + 1.0); }  // random_variables.hpp:22-26

int g_search_order[12] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};

// ---------------------------------------------------------------- profiling
enum Phase { kPhTp, kPhEnum, kPhMu, kPhMom, kPhGather, kPhFtr, kPhGBtab, kPhGEval, kPhCompact, kNumPhases };

struct Profiler {
    bool on = false;
    double ms[kNumPhases] = {0};
    double total = 0;
    long long geval_launches = 0;
    double geval_bytes = 0;
    double geval_ms = 0;
    double alg_bytes = 0;   // algorithmic bytes of the profiled updates (whole update)
    double alg_updates = 0; // updates profiled
    double groups = 0;      // reduction groups evaluated
    double members = 0;     // members in those groups (terms evaluated = groups + members)
    double max_group = 0;   // largest group (root + members)
    hipEvent_t k0 = nullptr, k1 = nullptr;
    std::vector<std::pair<int, hipEvent_t>> marks;
    hipEvent_t start_ev = nullptr;
    void begin(hipStream_t s) {
        if (!on) return;
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        HIPCHK(hipEventRecord(e, s));
        start_ev = e;
        marks.clear();
    }
    void mark(int ph, hipStream_t s) {
        if (!on) return;
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        HIPCHK(hipEventRecord(e, s));
        marks.push_back(std::make_pair(ph, e));
    }
    void end(hipStream_t s) {
        if (!on || !start_ev) return;
        HIPCHK(hipStreamSynchronize(s));
        hipEvent_t prev = start_ev;
        for (auto& pe : marks) {
            float t = 0;
            HIPCHK(hipEventElapsedTime(&t, prev, pe.second));
            ms[pe.first] += t;
            total += t;
            if (prev != start_ev) (void)hipEventDestroy(prev);
            prev = pe.second;
        }
        if (prev != start_ev) (void)hipEventDestroy(prev);
        (void)hipEventDestroy(start_ev);
        start_ev = nullptr;
        marks.clear();
    }
};

// ---------------------------------------------------------------- per-thread streams
// An engine runs on the streams and pinned staging buffer of the host thread that calls
// it (calls on one engine never overlap, and every call leaves its streams idle), so a
// process holds two streams per calling thread instead of two per estimator: thousands of
// live estimators do not multiply the streams that every device-wide synchronisation
// (hipFree, hipDeviceSynchronize) has to visit.  They live as long as the process.
constexpr size_t kStageBytes = 4096;
struct ThreadStreams {
    hipStream_t st = nullptr, st2 = nullptr;
    uint8_t* stage = nullptr;
    // a pinned, device-mapped slot that kernels write scan totals into directly
    int64_t* total = nullptr;
    int64_t* total_dev = nullptr;
    void init() {
        if (st) return;
        HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        HIPCHK(hipStreamCreateWithFlags(&st2, hipStreamNonBlocking));
        HIPCHK(hipHostMalloc((void**)&stage, kStageBytes, hipHostMallocDefault));
        HIPCHK(hipHostMalloc((void**)&total, 64, hipHostMallocMapped));
        HIPCHK(hipHostGetDevicePointer((void**)&total_dev, total, 0));
    }
};
thread_local ThreadStreams t_streams;

// ---------------------------------------------------------------- engine
struct Engine {
    int d, cmcc, pncc, p, n_mu, sr, MM;
    int master_step = 0, Nt = 1, errs = 0, win_num = 0;
    bool skip = false;
    std::vector<double> A0, p0, b0, root, bpert;
    cplx fz, last_fz;
    cplx fz_raw = {0, 0};  // the update's fz before the numeric checks may replace it by last_fz
    std::vector<cplx> mean, cov, last_mean, last_cov;
    double gscale = 0;
    std::vector<int> in_count;  // per-shape term counts of the parent pool
    // largest hyperplane count among the parent pool's terms (0 when it is empty)
    int largest_shape() const {
        for (int m = (int)in_count.size() - 1; m > 0; m--)
            if (in_count[m] > 0) return m;
        return 0;
    }
    // diagnostics of the last update
    bool have_diag = false;
    std::vector<double> diag_mom;
    std::vector<int> diag_cnt;
    int threads_used = 1;

    hipStream_t st = nullptr;
    hipStream_t st2 = nullptr;  // moment chains, overlapped with the reduction phase
    hipEvent_t ev_mu = nullptr, ev_c0 = nullptr, ev_c1 = nullptr;
    DBuf<uint32_t> chain_key, chain_skey, chain_perm;
    DBuf<int> chain_off;
    int cur_T = 1;
    bool moments_pending = false;
    int max_gp_cells = 1;  // largest parent g-table of the current parents
    int geval_mode = -1;   // parent lookup mode of the last g-table launch
    long long terms_generated = 0;  // terms after MU/MUC (after MU on a window's last update), summed over updates
    // MCE_GEVAL_MODE=0..3 selects the g-table lookup mode (tuning; unsupported choices are ignored)
    int geval_mode_override = getenv("MCE_GEVAL_MODE") ? atoi(getenv("MCE_GEVAL_MODE")) : -1;
    // MCE_BIG_GROUP_MEMBERS: groups with more members add them one wavefront per member
    int big_group_members = getenv("MCE_BIG_GROUP_MEMBERS") ? std::max(0, atoi(getenv("MCE_BIG_GROUP_MEMBERS")))
                                                            : kBigGroupMembers;
    // MCE_ENUM_THREADS=64|128|256: lanes per parent in the TP cell enumerator
    int enum_threads = enum_threads_from_env();
    PoolStore in, out;
    DBuf<uint32_t> gk[2];
    DBuf<cplx> gv[2];
    // ---- tangent mode (P directions): the pools' tangents, the g-tables' tangents
    // ([cell][direction], same offsets as gv), the update's seeds and results
    int ntan = 0;
    TanStore tin, tout;
    DBuf<cplx> dgv[2];
    DBuf<cplx> dg_slot, dfz_dev;
    DBuf<double> dgscale_dev;
    DBuf<StepTan> dstan;
    DBuf<uint8_t> mem_add;
    StepTan htan;                // seeds of the update being stepped
    bool tan_seeded = false;
    std::vector<cplx> dfz_tan;   // d fz of the last update, per direction
    int gcur = 0;
    DBuf<uint32_t> barena, b2arena, tmpk;
    DBuf<cplx> scratch, tmpv;
    DBuf<StepConst> dconst;
    DBuf<double> dbpert, gy, mom_part, mom_out, mom_x;
    DBuf<uint8_t> dcombos;
    DBuf<int64_t> dcombo_off;
    DBuf<int> dcombo_cnt, dcen_half;
    DBuf<int> slots, slot_base, enum_list, counters, hist;
    DBuf<int64_t> btcap64, bt_base;
    DBuf<int> btcap;
    DBuf<uint32_t> keys, skeys, iota, perm;
    // reduction
    DBuf<uint64_t> fkeys, fskeys;
    DBuf<uint32_t> fidx, fsidx;
    DBuf<double> fvals;
    DBuf<int> ecnt, eoff, edges, indeg, inoff, cursor, cursor2, round_slots, Fc;
    DBuf<uint8_t> st0, st1, isroot;
    DBuf<uint32_t> memkey, memval, smkey, smval, roots, jiota;
    DBuf<int> memcnt, memoff, inedges, dboff, gbase_dev, wlo;
    DBuf<int64_t> wcnt, woff;
    DBuf<double> fgb, fgp, fgA;
    DBuf<uint8_t> pflag;
    DBuf<uint32_t> bkey, bskey;
    std::vector<int> h_boff;
    // groups (all buckets of an update)
    DBuf<int> grp_root, grp_moff, grp_mcnt, grp_m, mem_slots, b2n, surv, gxor, sort_list, spos, sflag;
    DBuf<int64_t> ub64, b2off, goff, b2n64;
    DBuf<uint8_t> cub_tmp;
    DBuf<unsigned long long> bytes_acc;
    DBuf<int> pstat;
    uint8_t* h_stage = nullptr;  // pinned staging for device-to-host scalar reads (the thread's)
    int64_t* h_total = nullptr;      // the thread's mapped total slot, host and device views
    int64_t* h_total_dev = nullptr;

    // Set while a batch scheduler steps this engine: launches, uploads and library calls
    // become requests to it, and device reads suspend the update until it has the data.
    BatchSlot* bs = nullptr;
    BatchSlot* bslot() const { return bs; }

    // Takes the calling thread's streams and staging buffer; every entry point starts here.
    // Under a batch scheduler everything runs in order on its one stream.
    void bind() {
        t_streams.init();
        st = t_streams.st;
        st2 = bs ? t_streams.st : t_streams.st2;
        h_stage = t_streams.stage;
        h_total = t_streams.total;
        h_total_dev = t_streams.total_dev;
    }
    DBuf<int64_t> d_scal;        // device scalars (scan totals)
    DBuf<int> defer_k0, big_icnt, big_ioff, big_hcnt, big_hoff, hist_mx;
    DBuf<int64_t> big_ccnt, big_coff;
    DBuf<cplx> contrib;
    DBuf<uint8_t> cneg;
    Profiler prof;
    Graveyard graveyard;
    static constexpr size_t kGraveyardBytes = size_t(4) << 30;  // released early beyond this

    // Gives back the outgrown allocations once this engine's streams are idle: every one if
    // `all`, else the cache-sized ones, and the larger ones beyond kGraveyardBytes.
    void bury(bool all = true) {
        if (graveyard.blocks.empty()) return;
        bind();
        HIPCHK(hipStreamSynchronize(st));
        HIPCHK(hipStreamSynchronize(st2));
        graveyard.release(all || graveyard.big_bytes > kGraveyardBytes);
    }

    Engine(int num_steps, int d_, int cmcc_, int pncc_, int p_, const double* A0_, const double* p0_,
           const double* b0_)
        : d(d_), cmcc(cmcc_), pncc(pncc_), p(p_) {
        if (d < 1 || d > kMaxD) throw std::runtime_error("state dimension outside 1..8 is not supported");
        n_mu = p * num_steps;
        const int maxhp = d > 1 ? (num_steps - 1) * pncc + d : d + pncc;
        sr = maxhp + 1;
        MM = maxhp;
        if (maxhp > 31) throw std::runtime_error("largest hyperplane count exceeds 31");
        if (maxhp > 20) throw std::runtime_error("largest hyperplane count above 20 is not supported by the LDS cell tables");
        A0.assign(A0_, A0_ + d * d);
        p0.assign(p0_, p0_ + d);
        b0.assign(b0_, b0_ + d);
        // construction-time draws in the reference's order (cauchy_estimator.hpp:127-161)
        root.resize(d);
        for (int i = 0; i < d; i++) root[i] = 1.0 + rand_u();
        bpert.assign(8 * MM, 0.0);
        for (int h = 0; h < 8; h++)
            for (int i = 0; i < MM; i++) bpert[h * MM + i] = 2 * rand_u() - 1;
        // the first HIP call of a process may itself draw from rand(): streams only after
        // the draws above
        bind();
        mean.assign(d, cmake(0, 0));
        cov.assign(d * d, cmake(0, 0));
        last_mean = mean;
        last_cov = cov;
        fz = cmake(0, 0);
        last_fz = cmake(0, 0);
        d_scal.reserve(16);
        HIPCHK(hipEventCreateWithFlags(&ev_mu, hipEventDisableTiming));
        HIPCHK(hipEventCreate(&ev_c0));
        HIPCHK(hipEventCreate(&ev_c1));
        in.MM = out.MM = MM;
        in.d = out.d = d;
        // combinations in lexicographic order per shape (cell_enumeration.hpp:71-92)
        std::vector<uint8_t> cmb;
        std::vector<int64_t> coff(sr + 1, 0);
        std::vector<int> ccnt(sr, 0), cen(sr + 1, 0);
        for (int m = 0; m < sr; m++) {
            coff[m] = (int64_t)cmb.size();
            cen[m] = cells_central(m, d) / 2;
            if (m <= d) continue;
            std::vector<int> c(d);
            for (int i = 0; i < d; i++) c[i] = i;
            while (true) {
                for (int i = 0; i < d; i++) cmb.push_back((uint8_t)c[i]);
                ccnt[m]++;
                int i = d - 1;
                while (i >= 0 && c[i] == m - d + i) i--;
                if (i < 0) break;
                c[i]++;
                for (int j = i + 1; j < d; j++) c[j] = c[j - 1] + 1;
            }
        }
        coff[sr] = (int64_t)cmb.size();
        dcombos.reserve(std::max<size_t>(cmb.size(), 1));
        if (!cmb.empty()) copy_sync(dcombos.ptr, cmb.data(), cmb.size(), hipMemcpyHostToDevice);
        dcombo_off.reserve(sr + 1);
        copy_sync(dcombo_off.ptr, coff.data(), (sr + 1) * sizeof(int64_t), hipMemcpyHostToDevice);
        dcombo_cnt.reserve(sr);
        copy_sync(dcombo_cnt.ptr, ccnt.data(), sr * sizeof(int), hipMemcpyHostToDevice);
        dcen_half.reserve(sr + 1);
        copy_sync(dcen_half.ptr, cen.data(), (sr + 1) * sizeof(int), hipMemcpyHostToDevice);
        dbpert.reserve(8 * MM);
        upload_draws();
        dconst.reserve(1);
        counters.reserve(32);
        hist.reserve(sr + 1);
        hist_mx.reserve(sr + 2);
        reset_terms();
    }

    ~Engine() {
        bind();
        (void)hipStreamSynchronize(st);
        (void)hipStreamSynchronize(st2);
        graveyard.release(true);
        if (ev_mu) (void)hipEventDestroy(ev_mu);
        if (ev_c0) (void)hipEventDestroy(ev_c0);
        if (ev_c1) (void)hipEventDestroy(ev_c1);
    }

    void upload_draws() {
        bind();
        copy_sync(dbpert.ptr, bpert.data(), 8 * MM * sizeof(double), hipMemcpyHostToDevice);
    }

    // First term (setup_first_term, cauchy_term.hpp:772-785).
    void reset_terms() {
        bind();
        bury();
        GraveyardScope gs(&graveyard);
        in.reserve(1);
        std::vector<double> A(MM * d, 0.0), pp(MM, 0.0);
        memcpy(A.data(), A0.data(), d * d * sizeof(double));
        memcpy(pp.data(), p0.data(), d * sizeof(double));
        copy_sync(in.A.ptr, A.data(), MM * d * sizeof(double), hipMemcpyHostToDevice);
        copy_sync(in.p.ptr, pp.data(), MM * sizeof(double), hipMemcpyHostToDevice);
        copy_sync(in.b.ptr, b0.data(), d * sizeof(double), hipMemcpyHostToDevice);
        int mm = d, zero = 0;
        int64_t z64 = 0;
        copy_sync(in.m.ptr, &mm, sizeof(int), hipMemcpyHostToDevice);
        copy_sync(in.gp_n.ptr, &zero, sizeof(int), hipMemcpyHostToDevice);
        copy_sync(in.gp_off.ptr, &z64, sizeof(int64_t), hipMemcpyHostToDevice);
        in_count.assign(sr, 0);
        in_count[d] = 1;
        Nt = 1;
        master_step = 0;
        errs = 0;
        if (ntan) zero_prior_tangents();
    }

    // The first term (A0, p0, b0) carries no tangent.
    void zero_prior_tangents() {
        tin.reserve(1);
        std::vector<double> z((size_t)ntan * MM * d, 0.0);
        copy_sync(tin.A.ptr, z.data(), (size_t)ntan * MM * d * sizeof(double), hipMemcpyHostToDevice);
        copy_sync(tin.p.ptr, z.data(), (size_t)ntan * MM * sizeof(double), hipMemcpyHostToDevice);
        copy_sync(tin.b.ptr, z.data(), (size_t)ntan * d * sizeof(double), hipMemcpyHostToDevice);
    }

    // Tangent mode with P directions, from the first update of a window on (the first
    // update after construction or a reset).
    void enable_tangents(int P) {
        if (P < 1 || P > kMaxTan) throw std::runtime_error("tangent directions outside 1.." + std::to_string(kMaxTan));
        if (master_step != 0) throw std::runtime_error("tangents are enabled before a window's first update");
        bind();
        GraveyardScope gs(&graveyard);
        ntan = P;
        for (TanStore* t : {&tin, &tout}) {
            t->P = P;
            t->MM = MM;
            t->d = d;
            t->cap = 0;
        }
        zero_prior_tangents();
        memset(&htan, 0, sizeof(htan));
        dfz_tan.assign(P, cmake(0, 0));
    }
    std::vector<double> tan_dbeta;  // raw process-scale seeds of the update, ntan x pncc

    // A device-wide library call (hipcub) with its temporary storage: run now, or, under a
    // batch scheduler, handed to it with every argument captured by value and run in order
    // at its next flush.  The size query touches no stream.
    template <class F>
    void cub(const char* name, F f) {
        size_t bytes = 0;
        HIPCHK(f((void*)nullptr, bytes, st));
        cub_tmp.reserve(bytes + 16);
        void* tmp = cub_tmp.ptr;
        if (bs) {
            bs->push_host([f, tmp, bytes](hipStream_t s) mutable { HIPCHK(f(tmp, bytes, s)); }, name);
            return;
        }
        HIPCHK(f(tmp, bytes, st));
    }
    // Stable sort of n (key, value) pairs by key bits [b0, b1): one workgroup for small inputs
    // (one launch, merged with other estimators' under a batch scheduler), the multi-block
    // radix sort up to kRadixMax items under a batch scheduler, else the library radix sort
    // (the same order: all are stable sorts over the same bits).
    template <class K, class V>
    void sort_pairs(const K* kin, K* kout, const V* vin, V* vout, int n, int b0, int b1) {
        if (n <= 0) return;
        if (n <= kSmallSort) {
            MCE_RUN((k_small_sort<K, V>), dim3(1), dim3(kSmallSortThreads), 0, st, kin, kout, vin, vout, n, b0, b1);
            return;
        }
        if constexpr (std::is_same_v<V, uint32_t>) {
            if (bs && n <= kRadixMax && b0 == 0) {
                radix_sort(kin, kout, vin, vout, n, b1);
                return;
            }
        }
        cub("radix sort", [=](void* t, size_t& bytes, hipStream_t s) {
            return hipcub::DeviceRadixSort::SortPairs(t, bytes, kin, kout, vin, vout, n, b0, b1, s);
        });
    }
    // The multi-block radix sort (k_radix_hist, k_small_scan, k_radix_scatter per 8-bit
    // digit) over bits [0, nbits), ping-ponging through rs_keys / rs_vals so that the last
    // pass writes kout / vout.  Its launches merge with other estimators' under a scheduler.
    DBuf<uint8_t> rs_keys;
    DBuf<uint32_t> rs_vals;
    DBuf<int> rs_hist, rs_offs;
    template <class K>
    void radix_sort(const K* kin, K* kout, const uint32_t* vin, uint32_t* vout, int n, int nbits) {
        const int nc = ceil_div(n, kRadixChunk);
        const int passes = (nbits + 7) / 8;
        rs_keys.reserve((size_t)n * sizeof(K));
        rs_vals.reserve(n);
        rs_hist.reserve((size_t)kRadixBins * nc);
        rs_offs.reserve((size_t)kRadixBins * nc);
        K* tk = (K*)rs_keys.ptr;
        uint32_t* tv = rs_vals.ptr;
        const K* sk = kin;
        const uint32_t* sv = vin;
        for (int p = 0; p < passes; p++) {
            const bool to_out = ((passes - 1 - p) % 2) == 0;
            K* dk = to_out ? kout : tk;
            uint32_t* dv = to_out ? vout : tv;
            const int shift = 8 * p, bits = std::min(8, nbits - 8 * p);
            MCE_RUN((k_radix_hist<K>), dim3(nc), dim3(kRadixThreads), 0, st, sk, n, shift, bits, rs_hist.ptr, nc);
            MCE_RUN((k_small_scan<int, int, false>), dim3(1), dim3(kSmallScanThreads), 0, st, (const int*)rs_hist.ptr,
                    rs_offs.ptr, kRadixBins * nc, (int*)nullptr, (int64_t*)nullptr, (int64_t*)nullptr);
            MCE_RUN((k_radix_scatter<K>), dim3(nc), dim3(kRadixThreads), 0, st, sk, sv, dk, dv, n, shift, bits,
                    (const int*)rs_offs.ptr, nc);
            sk = dk;
            sv = dv;
        }
    }
    // out <- in[i] where flags[i] != 0, in order; the count into *count.
    void select_flagged(const uint32_t* in_, const uint8_t* flags, uint32_t* out_, int* count, int n) {
        if (n <= kSmallScan) {
            MCE_RUN(k_small_select, dim3(1), dim3(kSmallScanThreads), 0, st, in_, flags, out_, count, n);
            return;
        }
        cub("select", [=](void* t, size_t& bytes, hipStream_t s) {
            return hipcub::DeviceSelect::Flagged(t, bytes, in_, flags, out_, count, n, s);
        });
    }
    // *out <- max of n ints.
    void reduce_max(const int* in_, int* out_, int n) {
        if (n <= kSmallScan) {
            MCE_RUN(k_small_max, dim3(1), dim3(1024), 0, st, in_, out_, n);
            return;
        }
        cub("reduce max", [=](void* t, size_t& bytes, hipStream_t s) {
            return hipcub::DeviceReduce::Max(t, bytes, in_, out_, n, s);
        });
    }
    // Host-to-device copy ordered on the engine's stream (staged by the scheduler in batch
    // mode, so the source may change once this returns).
    void upload(void* dst, const void* src, size_t bytes) {
        if (bytes == 0) return;
        if (bs) {
            bs->push_upload(dst, src, bytes);
            return;
        }
        HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
    }

    struct FillSpec {
        void* ptr;
        size_t bytes;
        int val;
    };
    // Up to kFillRegions fills of this engine's stream in one launch.
    void fill_multi(std::initializer_list<FillSpec> regions) {
        if (regions.size() > (size_t)kFillRegions) throw std::runtime_error("fill_multi: too many regions");
        FillArgs f{};
        int n = 0;
        int64_t most = 0;
        for (const FillSpec& r : regions) {
            if (r.bytes == 0) continue;
            f.ptr[n] = (uint8_t*)r.ptr;
            f.bytes[n] = (int64_t)r.bytes;
            f.val[n] = (uint32_t)(r.val & 0xff);
            most = std::max(most, (int64_t)r.bytes);
            n++;
        }
        if (n == 0) return;
        const int blocks = (int)std::min<int64_t>(1024, std::max<int64_t>(1, (most / 4 + 255) / 256));
        MCE_RUN(k_fill_multi, dim3(blocks, n), dim3(256), 0, st, f);
        HIPCHK(hipGetLastError());
    }

    // Device scalars are read through pinned host memory, one copy and one stream
    // synchronisation per read (several values at once where possible).
    template <class T>
    T read1(const T* dptr) {
        HIPCHK(hipMemcpyAsync(h_stage, dptr, sizeof(T), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        T v;
        memcpy(&v, h_stage, sizeof(T));
        return v;
    }
    template <class T>
    void readn(const T* dptr, T* out_, int n) {
        HIPCHK(hipMemcpyAsync(h_stage, dptr, n * sizeof(T), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        memcpy(out_, h_stage, n * sizeof(T));
    }
    // Awaitable device read of `bytes` bytes into host memory `dst` (nothing: just wait until
    // the engine's queued work is done).  On its own the engine copies and synchronises its
    // stream where the read is awaited; under a batch scheduler the update suspends and the
    // scheduler reads every waiting engine's ranges after one flush and one synchronisation.
    struct ReadAw {
        Engine* e;
        const void* src;
        size_t bytes;
        void* dst;
        bool await_ready() {
            if (e->bs) return false;
            if (bytes == 0) {
                HIPCHK(hipStreamSynchronize(e->st));
            } else if (bytes <= kStageBytes) {
                HIPCHK(hipMemcpyAsync(e->h_stage, src, bytes, hipMemcpyDeviceToHost, e->st));
                HIPCHK(hipStreamSynchronize(e->st));
                memcpy(dst, e->h_stage, bytes);
            } else {
                HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, e->st));
                HIPCHK(hipStreamSynchronize(e->st));
            }
            return true;
        }
        void await_suspend(std::coroutine_handle<> h) {
            if (bytes) e->bs->reads.push_back(BatchRead{src, bytes, dst});
            e->bs->waiting = h;
        }
        void await_resume() {}
    };
    template <class T>
    ReadAw rd(const T* src, T* dst, int n) {
        return ReadAw{this, src, n * sizeof(T), dst};
    }
    ReadAw sync_point() { return ReadAw{this, nullptr, 0, nullptr}; }
    template <class T>
    Co<T> rd1(const T* src) {
        T v{};
        co_await rd(src, &v, 1);
        co_return v;
    }
    // Blocking copy ordered on the engine's stream.
    void copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
        if (bytes == 0) return;
        HIPCHK(hipMemcpyAsync(dst, src, bytes, kind, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    // off[n-1] + cnt[n-1] (the total of an exclusive scan) into device scalar `slot`.
    template <class A, class B>
    void enqueue_total(const A* off, const B* cnt, int n, int slot) {
        MCE_RUN((k_scan_total<A, B>), dim3(1), dim3(1), 0, st, off, cnt, n, d_scal.ptr + slot);
    }
    // Where a total computed on the device goes: d_scal[0], and on its own the engine also
    // has the kernel store it into the thread's device-mapped host slot (so no copy is
    // launched to read it).  The slot holds -1 until the kernel's value arrives.
    int64_t* mapped_slot() {
        if (bs) return nullptr;
        *(volatile int64_t*)h_total = -1;
        return h_total_dev;
    }
    // The total a kernel just stored through mapped_slot(): from the mapped slot once the
    // stream is done (d_scal[0] should the value not have arrived), or from d_scal[0] by a
    // scheduler read.
    Co<int64_t> take_total() {
        if (bs) co_return co_await rd1(d_scal.ptr);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
        const int64_t v = *(volatile int64_t*)h_total;
        co_return v >= 0 ? v : read1(d_scal.ptr);
    }
    // off[n-1] + cnt[n-1] read by the host.
    template <class A, class B>
    Co<int64_t> read_total(const A* off, const B* cnt, int n) {
        MCE_RUN((k_scan_total_mapped<A, B>), dim3(1), dim3(1), 0, st, off, cnt, n, d_scal.ptr, mapped_slot());
        co_return co_await take_total();
    }

    // Transpose, L1-normalise and pre-coalign Gamma/beta, cauchy_util.hpp:128-210.
    // In tangent mode `dbeta` (ntan x pncc) is carried through the same scaling and merges
    // into tan->beta.
    int precoalign(const double* Gamma, const double* beta, StepConst& sc, const double* dbeta = nullptr,
                   StepTan* tan = nullptr) {
        std::vector<double> G(pncc * d), bt(pncc);
        const int NP = (dbeta && tan) ? ntan : 0;
        std::vector<double> dbt((size_t)NP * pncc);
        for (int i = 0; i < pncc; i++) {
            for (int j = 0; j < d; j++) G[i * d + j] = Gamma[j * pncc + i];
            bt[i] = beta[i];
            for (int k = 0; k < NP; k++) dbt[k * pncc + i] = dbeta[k * pncc + i];
        }
        for (int i = 0; i < pncc; i++) {
            double nf = 0;
            for (int j = 0; j < d; j++) nf += fabs(G[i * d + j]);
            for (int j = 0; j < d; j++) G[i * d + j] /= nf;
            bt[i] *= nf;
            for (int k = 0; k < NP; k++) dbt[k * pncc + i] *= nf;
        }
        std::vector<char> F(pncc, 1);
        for (int i = 0; i < pncc - 1; i++) {
            if (!F[i]) continue;
            for (int j = i + 1; j < pncc; j++) {
                if (!F[j]) continue;
                bool pos = true, neg = true;
                for (int l = 0; l < d; l++) {
                    if (pos) pos &= fabs(G[i * d + l] - G[j * d + l]) < kCoalignEps;
                    if (neg) neg &= fabs(G[i * d + l] + G[j * d + l]) < kCoalignEps;
                    if (!(pos || neg)) break;
                }
                if (pos || neg) {
                    F[j] = 0;
                    bt[i] += bt[j];
                    for (int k = 0; k < NP; k++) dbt[k * pncc + i] += dbt[k * pncc + j];
                }
            }
        }
        int cnt = 0;
        for (int i = 0; i < pncc; i++) cnt += F[i];
        if (cnt != pncc) {
            int u = 1;
            for (int j = 1; j < pncc; j++)
                if (F[j]) {
                    if (u < j) {
                        memcpy(&G[u * d], &G[j * d], d * sizeof(double));
                        bt[u] = bt[j];
                        for (int k = 0; k < NP; k++) dbt[k * pncc + u] = dbt[k * pncc + j];
                    }
                    u++;
                }
        }
        for (int i = 0; i < cnt; i++) {
            for (int j = 0; j < d; j++) sc.G[i * d + j] = G[i * d + j];
            sc.beta[i] = bt[i];
            for (int k = 0; k < NP; k++) tan->beta[k][i] = dbt[k * pncc + i];
        }
        return cnt;
    }

    // finalize_cached_moments + moments_numerical_check, cauchy_estimator.hpp:340-461.
    void finalize_and_check() {
        gscale = kInvTwoPi / fz.re;
        const cplx ifz = cmake(0, fz.re);
        for (int i = 0; i < d; i++) mean[i] = cdiv(mean[i], ifz);
        for (int i = 0; i < d; i++)
            for (int j = 0; j < d; j++) cov[i * d + j] = csub(cdiv(cov[i * d + j], fz), cmul(mean[i], mean[j]));
        const bool first_of_step = (master_step % p) == 0;
        const bool not_last = (master_step % p) != (p - 1);
        const bool last = (master_step % p) == (p - 1);
        if (first_of_step) {
            errs |= (1 << 7);
            errs |= (1 << 3);
            errs &= ~(1 << 5);
            errs &= ~(1 << 6);
            errs &= ~(1 << 2);
            errs &= ~(1 << 1);
        }
        int ne = 0;
        if (fz.re <= 0) ne |= (1 << 9);
        if (fabs(fz.im / (1e-15 + fz.re)) > 1e-3) ne |= (1 << 8);
        bool mean_ok = true;
        for (int i = 0; i < d; i++) {
            double r = fabs(mean[i].re), im = fabs(mean[i].im);
            double ratio = im / (1e-15 + r);
            if ((ratio > 1e-1) || (im > 0.001)) {
                ne |= (1 << 4);
                if (not_last) ne |= (1 << 6);
                if (last) ne |= (1 << 5);
                mean_ok = false;
            }
        }
        int cf = 0;
        {
            std::vector<double> cr(d * d), ci(d * d), ev(d);
            for (int i = 0; i < d * d; i++) {
                cr[i] = cov[i].re;
                ci[i] = cov[i].im;
            }
            {
                std::vector<double> evec((size_t)d * d);
                host::sym_eig(cr.data(), ev.data(), evec.data(), d);
            }
            for (int i = 0; i < d; i++)
                if (ev[i] < kCovEigenvalueTol) cf |= 1;
            for (int i = 0; i < d - 1; i++) {
                double sii = sqrt(cr[i * d + i]);
                for (int j = i + 1; j < d; j++) {
                    double sjj = sqrt(cr[j * d + j]);
                    double corr = cr[i * d + j] / (sii * sjj);
                    if (fabs(corr) > 1) cf |= 2;
                }
            }
            for (int i = 0; i < d; i++)
                for (int j = i; j < d; j++)
                    if (fabs(ci[i * d + j]) / (fabs(cr[i * d + j]) + 1e-15) > 10) cf |= 4;
            for (int i = 0; i < d * d; i++)
                if (fabs(ci[i]) > 2000) cf |= 8;
        }
        bool cov_ok = true;
        if (cf) {
            ne |= 1;
            if (not_last) ne |= (1 << 2);
            if (last) ne |= (1 << 1);
            cov_ok = false;
        }
        errs |= ne;
        if (mean_ok) errs &= ~(1 << 7);
        if (cov_ok) errs &= ~(1 << 3);
        if (mean_ok) last_mean = mean;
        else mean = last_mean;
        if (cov_ok) last_cov = cov;
        else cov = last_cov;
        if (!mean_ok && !cov_ok) fz = last_fz;
        else last_fz = fz;
    }

    // ------------------------------------------------------------ dispatch helpers
#define MCE_DISPATCH_D(D_, ...)                         \
    switch (D_) {                                       \
        case 1: { constexpr int DD = 1; __VA_ARGS__; } break; \
        case 2: { constexpr int DD = 2; __VA_ARGS__; } break; \
        case 3: { constexpr int DD = 3; __VA_ARGS__; } break; \
        case 4: { constexpr int DD = 4; __VA_ARGS__; } break; \
        case 5: { constexpr int DD = 5; __VA_ARGS__; } break; \
        case 6: { constexpr int DD = 6; __VA_ARGS__; } break; \
        case 7: { constexpr int DD = 7; __VA_ARGS__; } break; \
        case 8: { constexpr int DD = 8; __VA_ARGS__; } break; \
        default: throw std::runtime_error("unsupported state dimension"); \
    }

    // Scans.  Up to kSmallScan items take one workgroup and one launch (k_small_scan, which
    // also widens and forms the total); larger inputs the library's device scan.
    // Inclusive scan of n ints; returns the total (the last inclusive value).
    Co<int> scan_i32_incl_total(const int* in_, int* out_, int n) {
        if (n <= kSmallScan) {
            int64_t* host = mapped_slot();
            MCE_RUN((k_small_scan<int, int, true>), dim3(1), dim3(kSmallScanThreads), 0, st, in_, out_, n,
                    (int*)nullptr, d_scal.ptr, host);
            co_return (int)(co_await take_total());
        }
        cub("inclusive scan", [=](void* t, size_t& b, hipStream_t s) {
            return hipcub::DeviceScan::InclusiveSum(t, b, in_, out_, n, s);
        });
        co_return co_await rd1(out_ + n - 1);
    }
    // Exclusive scan of n ints widened to 64 bits (the widened input kept in `wide`); returns
    // the total.
    Co<int64_t> widen_scan_total(const int* in_, int64_t* wide, int64_t* out_, int n) {
        if (n <= kSmallScan) {
            int64_t* host = mapped_slot();
            MCE_RUN((k_small_scan<int, int64_t, false>), dim3(1), dim3(kSmallScanThreads), 0, st, in_, out_, n, wide,
                    d_scal.ptr, host);
            co_return co_await take_total();
        }
        widen(in_, wide, n);
        scan_i64_excl(wide, out_, n);
        co_return co_await read_total(out_, wide, n);
    }
    template <class T>
    void scan_excl(const T* in_, T* out_, int n) {
        if (n <= kSmallScan) {
            MCE_RUN((k_small_scan<T, T, false>), dim3(1), dim3(kSmallScanThreads), 0, st, in_, out_, n, (T*)nullptr,
                    (int64_t*)nullptr, (int64_t*)nullptr);
            return;
        }
        cub("exclusive scan", [=](void* t, size_t& b, hipStream_t s) {
            return hipcub::DeviceScan::ExclusiveSum(t, b, in_, out_, n, s);
        });
    }
    // Exclusive scan with its total.
    template <class T>
    Co<int64_t> scan_excl_total(const T* in_, T* out_, int n) {
        if (n <= kSmallScan) {
            int64_t* host = mapped_slot();
            MCE_RUN((k_small_scan<T, T, false>), dim3(1), dim3(kSmallScanThreads), 0, st, in_, out_, n, (T*)nullptr,
                    d_scal.ptr, host);
            co_return co_await take_total();
        }
        scan_excl(in_, out_, n);
        co_return co_await read_total(out_, in_, n);
    }
    void scan_i64_excl(const int64_t* in_, int64_t* out_, int n) { scan_excl(in_, out_, n); }
    void scan_i32_excl(const int* in_, int* out_, int n) { scan_excl(in_, out_, n); }

    // Joins the moment chains, adds the per-thread chunks in thread order
    // (cauchy_estimator.hpp:886-895) and finalises the moments.
    // Tangent mode, once the update's fz (and so its g scale 1 / (2 pi Re fz)) is known:
    // d fz per direction, and the tangent of the g scale, -gscale d(Re fz) / Re fz.
    std::vector<double> dgscale_host;
    void tan_after_moments() {
        dfz_tan.assign(ntan, cmake(0, 0));
        HIPCHK(hipMemcpyAsync(dfz_tan.data(), dfz_dev.ptr, ntan * sizeof(cplx), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        dgscale_host.assign(ntan, 0.0);
        // gscale = 1 / (2 pi Re fz) of the update's own fz, before the numeric checks
        for (int k = 0; k < ntan; k++) dgscale_host[k] = -gscale * dfz_tan[k].re / fz_raw.re;
        dgscale_dev.reserve(ntan);
        HIPCHK(hipMemcpyAsync(dgscale_dev.ptr, dgscale_host.data(), ntan * sizeof(double), hipMemcpyHostToDevice,
                              st));
    }

    Co<void> collect_moments(bool first) {
        if (!moments_pending) co_return;
        const int NC = 1 + d + d * d;
        const int T = cur_T;
        // the partial sums, then the measurement update's error bits (k_mom_sum appends them)
        std::vector<double> part((size_t)T * NC * 2 + 1);
        if (bs) {
            co_await rd(mom_part.ptr, part.data(), (int)part.size());
        } else {
            HIPCHK(hipMemcpyAsync(part.data(), mom_part.ptr, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                                  st2));
            HIPCHK(hipStreamSynchronize(st2));
        }
        const int mu_err = (int)part[(size_t)T * NC * 2];
        moments_pending = false;
        if (prof.on) {
            float t = 0;
            HIPCHK(hipEventElapsedTime(&t, ev_c0, ev_c1));
            prof.ms[kPhMom] += t;
        }
        if (mu_err & 1) throw std::runtime_error("two H-orthogonal hyperplanes coalign negatively (reference exits here)");
        if (mu_err) throw std::runtime_error("measurement update: a term's shape exceeds the kernel's row bound");
        std::vector<double> mo(NC * 2, 0.0);
        for (int t = 0; t < T; t++)
            for (int c = 0; c < NC; c++) {
                mo[2 * c] += part[((size_t)t * NC + c) * 2];
                mo[2 * c + 1] += part[((size_t)t * NC + c) * 2 + 1];
            }
        fz = cmake(mo[0], mo[1]);
        fz_raw = fz;
        for (int j = 0; j < d; j++) mean[j] = cmake(mo[2 + 2 * j], mo[3 + 2 * j]);
        for (int j = 0; j < d * d; j++) cov[j] = cmake(mo[2 + 2 * d + 2 * j], mo[3 + 2 * d + 2 * j]);
        if (first) {
            // compute_moments(true), cauchy_estimator.hpp:524-602 (no numeric checks at step 1)
            gscale = kInvTwoPi / fz.re;
            const cplx ifz = cmake(0, fz.re);
            for (int i = 0; i < d; i++) mean[i] = cdiv(mean[i], ifz);
            for (int i = 0; i < d; i++)
                for (int j = 0; j < d; j++) cov[i * d + j] = csub(cdiv(cov[i * d + j], fz), cmul(mean[i], mean[j]));
        } else {
            finalize_and_check();
        }
        if (ntan) tan_after_moments();
        diag_mom[2] = fz.re;
        diag_mom[3] = fz.im;
        for (int i = 0; i < d; i++) {
            diag_mom[6 + 2 * i] = mean[i].re;
            diag_mom[7 + 2 * i] = mean[i].im;
        }
        for (int i = 0; i < d * d; i++) {
            diag_mom[6 + 2 * d + 2 * i] = cov[i].re;
            diag_mom[7 + 2 * d + 2 * i] = cov[i].im;
        }
    }

    // Parent lookup of this update's g-table kernel: the rank structure over the 2^(MM-1)
    // half-storage keys when it fits in LDS (with the parent values too when those fit), else
    // binary search; with its dynamic LDS and the rank structure's word capacity.  It depends
    // only on the largest shape, the parents' largest g-table and the override, all fixed
    // for the update.
    void pick_lookup(int& mode, size_t& lds, int& wmax) const {
        wmax = (std::max(4, (((1 << (MM - 1)) + 31) >> 5)) + 3) & ~3;  // multiple of 4 words
        const size_t rank_b = (size_t)wmax * 12;  // parent bitset, prefix counts, group bitset
        const size_t vals_b = (size_t)std::max(max_gp_cells, 1) * sizeof(cplx);
        const size_t keys_b = (size_t)std::max(max_gp_cells, 1) * sizeof(uint32_t);
        const size_t budget = 48 * 1024;
        // the rank structure with values read from global memory is the default: the
        // parent values are L2-resident and the smaller LDS footprint keeps more wavefronts
        // resident than staging the values (measured on MI355X)
        if (MM <= 24 && rank_b <= budget) {
            mode = kLookupRank;
            lds = rank_b;
        } else if (keys_b <= budget) {
            mode = kLookupSearchLds;
            lds = keys_b;
        } else {
            mode = kLookupSearch;
            lds = 0;
        }
        if (geval_mode_override >= 0 && geval_mode_override <= kLookupRankVals &&
            (geval_mode_override < kLookupRank || mode == kLookupRank) &&
            (geval_mode_override != kLookupRankVals || rank_b + vals_b <= budget)) {
            mode = geval_mode_override;
            lds = mode == kLookupRankVals ? rank_b + vals_b
                  : mode == kLookupRank   ? rank_b
                  : mode == kLookupSearchLds ? keys_b
                                             : 0;
        }
    }
    // Words of one parent's rank structure: its keys lie below 2^(phc-1).
    static int rank_words(int phc) { return phc >= 1 ? (int)((((uint32_t)1 << (phc - 1)) + 31u) >> 5) : 0; }
    DBuf<uint2> rank_arena;  // rank structures of the current parents' g-tables (k_rank_build)
    int upd_mode = -1;       // this update's parent lookup (pick_lookup at the update's start)

    // Debug mode (the reference's print_basic_info / debug_print): after term reduction the
    // public fz, mean and covariance are the moments of the reduced terms, each evaluated in
    // its own g-table at the root point's sign vector (compute_moments(false),
    // cauchy_estimator.hpp:524-602 with eval_g_yei_after_ftr, cauchy_term.hpp:403-433), summed
    // in term order, and printed.  Without it fz is 1 after reduction.
    bool debug_after_ftr = false;
    DBuf<double> dbg_gy, dbg_x, dbg_part;
    DBuf<uint8_t> dbg_flags;
    DBuf<uint32_t> dbg_chain;
    DBuf<int> dbg_off;
    Co<void> moments_after_ftr() {
        const int n = Nt;
        const int NC = 1 + d + d * d;
        TermPool P = in.view();
        dbg_gy.reserve((size_t)std::max(n, 1) * (2 + 2 * d));
        dbg_flags.reserve(std::max(n, 1));
        dbg_chain.reserve(std::max(n, 1));
        dbg_off.reserve(2);
        dbg_part.reserve((size_t)NC * 2 + 1);
        std::vector<double> part((size_t)NC * 2 + 1, 0.0);
        if (n > 0) {
            MCE_DISPATCH_D(d, MCE_RUN(k_after_ftr_gy<DD>, dim3(ceil_div(n, 256)), dim3(256), 0, st, P, n, dconst.ptr,
                                      gk[gcur].ptr, gv[gcur].ptr, dbg_gy.ptr));
            fill_multi({{dbg_flags.ptr, (size_t)n, kValid}});
            MCE_RUN(k_iota, dim3(ceil_div(n, 256)), dim3(256), 0, st, dbg_chain.ptr, n);
            const int off2[2] = {0, n};
            upload(dbg_off.ptr, off2, sizeof(off2));
            int ncp = 0, nyb = 0;
            MCE_DISPATCH_D(d, (nyb = mom_blocks<DD>(), ncp = nyb * kSumCW / 2));
            dbg_x.reserve((size_t)nyb * n * kSumCW);
            // one chain over the terms in order: the reference's sequential sums
            MCE_DISPATCH_D(d, MCE_RUN(k_mom_terms<DD>, dim3(ceil_div((int64_t)n * ncp, 256)), dim3(256), 0, st,
                                      dbg_gy.ptr, dbg_flags.ptr, dbg_chain.ptr, dbg_off.ptr, 1, n, dbg_x.ptr));
            MCE_DISPATCH_D(d, MCE_RUN(k_mom_sum<DD>, dim3(1, nyb), dim3(256), 0, st, dbg_x.ptr, dbg_off.ptr, 1,
                                      dbg_part.ptr, counters.ptr + 17));
            co_await rd(dbg_part.ptr, part.data(), (int)part.size());
        }
        fz = cmake(part[0], part[1]);
        for (int j = 0; j < d; j++) mean[j] = cmake(part[2 + 2 * j], part[3 + 2 * j]);
        for (int j = 0; j < d * d; j++) cov[j] = cmake(part[2 + 2 * d + 2 * j], part[3 + 2 * d + 2 * j]);
        if (!(fz.re > 0)) throw std::runtime_error("debug moments after term reduction: Re fz <= 0");
        const cplx ifz = cmake(0, fz.re);
        for (int i = 0; i < d; i++) mean[i] = cdiv(mean[i], ifz);
        for (int i = 0; i < d; i++)
            for (int j = 0; j < d; j++) cov[i * d + j] = csub(cdiv(cov[i * d + j], fz), cmul(mean[i], mean[j]));
        printf("Moment Information (after FTR) at step %d, MU %d/%d\n", (master_step + 1) / p, (master_step % p) + 1, p);
        printf("fz: %.16lf + %.16lfj\n", fz.re, fz.im);
        printf("Conditional Mean:\n");
        for (int i = 0; i < d; i++) printf("%.16lf + %.16lfj%s", mean[i].re, mean[i].im, i + 1 < d ? ", " : "\n");
        printf("Conditional Variance:\n");
        for (int i = 0; i < d; i++)
            for (int j = 0; j < d; j++)
                printf("%.16lf + %.16lfj%s", cov[i * d + j].re, cov[i * d + j].im, j + 1 < d ? ", " : "\n");
    }

    // ------------------------------------------------------------ algorithmic bytes
    // Bytes an update must move at least, counted from its parity-pinned term and cell
    // counts (profiling only): every parent's record (A, p, q, b, c, d) and g-table read
    // once; every child written by MU/MUC and read by the reduction; every surviving
    // g-table written (20 B per cell) with its B-table (4 B); every merged member's table
    // read and added (24 B per cell).
    double term_record_bytes(int m) const { return 8.0 * (m * (d + 2) + d + 2); }
    double prof_device_sum(int n, const int* a, const int* b, const int* keep) {
        if (n <= 0) return 0;
        bytes_acc.reserve(1);
        fill_multi({{bytes_acc.ptr, sizeof(unsigned long long), 0}});
        MCE_RUN(k_prof_sum, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, a, b, keep, bytes_acc.ptr);
        return (double)read1(bytes_acc.ptr);
    }

    // ------------------------------------------------------------ one scalar measurement update
    int update(double msmt, const double* Phi, const double* Gamma, const double* beta, const double* Hrow,
               double gamma, const double* Bm, const double* u) {
        return run_sync(update_co(msmt, Phi, Gamma, beta, Hrow, gamma, Bm, u));
    }
    // The update as a task.  On its own it never suspends, keeps the allocations it outgrows
    // until its streams are idle and gives them back at its end; under a batch scheduler it
    // suspends at each device read, and the scheduler (which sets the graveyard while the
    // task runs) gives them back once its stream is idle.
    Co<int> update_co(double msmt, const double* Phi, const double* Gamma, const double* beta, const double* Hrow,
                      double gamma, const double* Bm, const double* u) {
        bind();
        if (bs) co_return co_await update_terms(msmt, Phi, Gamma, beta, Hrow, gamma, Bm, u);
        int r;
        {
            GraveyardScope gs(&graveyard);
            r = co_await update_terms(msmt, Phi, Gamma, beta, Hrow, gamma, Bm, u);
        }
        bury(false);
        co_return r;
    }

    Co<int> update_terms(double msmt, const double* Phi, const double* Gamma, const double* beta, const double* Hrow,
                     double gamma, const double* Bm, const double* u) {
        if (errs & (1 << 9)) {
            fprintf(stderr, "[Window %d:] ERROR_FZ_NEGATIVE triggered. Cannot continue stepping until this estimator has been reset!\n", win_num);
            co_return errs;
        }
        if (bs && ntan) throw std::runtime_error("tangent mode steps one estimator at a time");
        if (master_step == n_mu)
            throw std::runtime_error("master_step == num_estimation_steps: the estimator must be reset before stepping again");
        skip = (master_step == n_mu - 1);
        const bool first = (master_step == 0);
        const bool with_tp = !first && (master_step % p) == 0;
        const bool threaded = !first && Nt >= 200;
        int T = 1;
        if (threaded) {
            int chunks = (Nt + 199) / 200;
            T = chunks > 8 ? 8 : chunks;
        }
        threads_used = threaded ? T : 1;
        StepConst sc;
        memset(&sc, 0, sizeof(sc));
        for (int i = 0; i < d * d; i++) sc.Phi[i] = Phi ? Phi[i] : 0.0;
        for (int i = 0; i < d; i++) sc.H[i] = Hrow[i];
        if (ntan && !tan_seeded) throw std::runtime_error("tangent mode: step with mce_estimator_step_tangent");
        sc.ng = with_tp ? precoalign(Gamma, beta, sc, ntan ? tan_dbeta.data() : nullptr, ntan ? &htan : nullptr) : 0;
        if (with_tp && cmcc > 0 && Bm && u)
            for (int i = 0; i < d; i++) {
                double s = 0.0;
                for (int j = 0; j < cmcc; j++) s += Bm[i * cmcc + j] * u[j];
                sc.Bu[i] = s;
            }
        for (int i = 0; i < d; i++) sc.root[i] = root[i];
        sc.msmt = msmt;
        sc.gamma = gamma;
        sc.d = d;
        sc.cmcc = (with_tp && Bm && u) ? cmcc : 0;
        sc.with_tp = with_tp;
        sc.skip = skip;
        sc.first = first;
        sc.T = T;
        sc.threaded = threaded;
        for (int i = 0; i < d; i++) {
            sc.order[i] = g_search_order[i];
            // the reduction reads b[order[i]] of every term
            if (sc.order[i] < 0 || sc.order[i] >= d)
                throw std::runtime_error("term-reduction search ordering names axis " + std::to_string(sc.order[i]) +
                                         " of a " + std::to_string(d) + "-state estimator");
        }
        // parents' shape bucket offsets travel with the constants (one upload), and with a
        // rank-structure lookup this update, each bucket's first word in the rank arena
        sc.in_off[0] = 0;
        for (int m = 0; m < sr; m++) sc.in_off[m + 1] = sc.in_off[m] + in_count[m];
        {
            size_t lds_;
            int wmax_;
            pick_lookup(upd_mode, lds_, wmax_);
        }
        const bool build_rank = !first && !skip && upd_mode >= kLookupRank;
        sc.rk_base[0] = 0;
        for (int m = 0; m < sr; m++) sc.rk_base[m + 1] = sc.rk_base[m] + (int64_t)in_count[m] * rank_words(m);
        upload(dconst.ptr, &sc, sizeof(sc));
        if (ntan) {
            dstan.reserve(1);
            upload(dstan.ptr, &htan, sizeof(StepTan));
        }
        const int np = Nt;
        prof.begin(st);
        if (prof.on && !bs) {
            // the parents' records and g-tables
            double b = 0;
            for (int m = 0; m < sr; m++) b += in_count[m] * term_record_bytes(m);
            b += 24.0 * prof_device_sum(np, in.view().gp_n, nullptr, nullptr);
            prof.alg_bytes += b;
            prof.alg_updates += 1;
        }

        // ---- time propagation and B-tables of the parents
        slots.reserve(np);
        slot_base.reserve(np + 1);
        btcap64.reserve(np);
        bt_base.reserve(np);
        btcap.reserve(np);
        TermPool IN = in.view();
        // counters: [5] root count, [8] enumeration guard, [16] parents to enumerate, [17] MU
        // errors, [18] groups to sort, [19] rare-path flags; the shape histograms of the gather
        // (hist) and of the compaction (hist_mx, with the largest g-table after it): all
        // cleared here, once per update
        fill_multi({{counters.ptr, 32 * sizeof(int), 0},
                    {slot_base.ptr, sizeof(int), 0},
                    {hist.ptr, (sr + 1) * sizeof(int), 0},
                    {hist_mx.ptr, (sr + 2) * sizeof(int), 0}});
        // tangents first: k_tp_tan reads the parents before k_tp propagates them in place
        if (ntan)
            MCE_DISPATCH_D(d, MCE_LAUNCH(k_tp_tan<DD>, dim3(ceil_div(np, 256)), dim3(256), 0, st, IN, tin.view(), np,
                                         dconst.ptr, dstan.ptr));
        MCE_DISPATCH_D(d, MCE_RUN(k_tp<DD>, dim3(ceil_div(np, 256)), dim3(256), 0, st, IN, np, dconst.ptr,
                                             dcen_half.ptr, slots.ptr, btcap.ptr, counters.ptr + 8));
        HIPCHK(hipGetLastError());
        // under a batch scheduler the slot and B-table scans (both of k_tp's counts) go out
        // together and their totals come back in one read
        const bool need_bt = !skip && !first;
        const bool paired = bs && np <= kSmallScan && need_bt;
        int nslots;
        int64_t bt_total = 0;
        if (paired) {
            MCE_RUN((k_small_scan<int, int, true>), dim3(1), dim3(kSmallScanThreads), 0, st, slots.ptr,
                    slot_base.ptr + 1, np, (int*)nullptr, d_scal.ptr + 1, (int64_t*)nullptr);
            MCE_RUN((k_small_scan<int, int64_t, false>), dim3(1), dim3(kSmallScanThreads), 0, st, btcap.ptr,
                    bt_base.ptr, np, btcap64.ptr, d_scal.ptr + 2, (int64_t*)nullptr);
            int64_t tt[2];
            co_await rd(d_scal.ptr + 1, tt, 2);
            nslots = (int)tt[0];
            bt_total = tt[1];
        } else {
            nslots = co_await scan_i32_incl_total(slots.ptr, slot_base.ptr + 1, np);
        }
        if (build_rank) {
            // each parent's rank structure once, before its children copy its offset
            int wmax_;
            size_t lds_;
            int mode_;
            pick_lookup(mode_, lds_, wmax_);
            rank_arena.reserve((size_t)std::max<int64_t>(sc.rk_base[sr], 1));
            MCE_RUN(k_rank_build, dim3(np), dim3(64), (size_t)wmax_ * 8, st, IN, np, dconst.ptr, gk[gcur].ptr,
                    rank_arena.ptr, wmax_);
        }
        prof.mark(kPhTp, st);
        if (need_bt) {
            // B arena offsets (64-bit)
            if (!paired) bt_total = co_await widen_scan_total(btcap.ptr, btcap64.ptr, bt_base.ptr, np);
            barena.reserve((size_t)bt_total + 1);
            enum_list.reserve(np);
            MCE_RUN(k_btab_simple, dim3(ceil_div(np, 256)), dim3(256), 0, st, IN, np, bt_base.ptr,
                               gk[gcur].ptr, barena.ptr, enum_list.ptr, counters.ptr + 16);
            HIPCHK(hipGetLastError());
            int ec[6];  // parents to enumerate, -, -, -, their largest shape and parent hyperplane count
            co_await rd(counters.ptr + 16, ec, 6);
            const int nenum = ec[0];
            if (nenum > 0) {
                // the 2^m-bit candidate set of the largest shape, then the parent key set of
                // 2^(phc-1) bits; when both do not fit beside the kernel's static LDS (m = 20
                // with phc = 19), parent keys are binary-searched in global memory instead
                const int mt = ec[4], ph = ec[5];
                if (mt < 1 || mt > MM || ph < 1 || ph >= mt)
                    throw std::runtime_error("cell enumerator: parent shapes outside the estimator's range");
                // queues of 2T kept vertices (combination, solution) and 2T undecided
                // combinations come first; with the largest sets, fewer lanes per parent
                const int words = mt >= 5 ? (1 << (mt - 5)) : 1;
                const int pwords = ph - 1 >= 5 ? (1 << (ph - 6)) : 1;
                const size_t enum_static = 4 * sizeof(int);
                int threads = enum_threads;
                auto queues = [&](int t) { return (size_t)2 * t * (8 * (size_t)d + 8); };
                int psearch = 0;
                size_t lds = queues(threads) + (size_t)(words + pwords) * 4;
                if (lds + enum_static > kLdsPerGroup) {
                    psearch = 1;
                    lds = queues(threads) + (size_t)words * 4;
                }
                while (lds + enum_static > kLdsPerGroup && threads > 64) {
                    threads /= 2;
                    lds = queues(threads) + (size_t)words * 4;
                }
                if (lds + enum_static > kLdsPerGroup)
                    throw std::runtime_error("hyperplane count too large for the cell enumerator's LDS sign-vector set");
                MCE_DISPATCH_D(d, {
                    MCE_RUN(k_btab_enum<DD>, dim3(nenum), dim3(threads), lds, st, IN, enum_list.ptr, dbpert.ptr,
                               dcombos.ptr, dcombo_off.ptr, dcombo_cnt.ptr, gk[gcur].ptr, barena.ptr, btcap.ptr,
                               counters.ptr + 8, words, psearch);
                });
                HIPCHK(hipGetLastError());
            }
        }
        prof.mark(kPhEnum, st);

        // ---- measurement update
        out.reserve(nslots);
        gy.reserve((size_t)nslots * (2 + 2 * d));
        keys.reserve(nslots);
        chain_key.reserve(nslots);
        chain_skey.reserve(nslots);
        chain_perm.reserve(nslots);
        iota.reserve(nslots);
        chain_off.reserve(16);
        mom_part.reserve((size_t)T * (1 + d + d * d) * 2 + 1);
        TermPool OUT = out.view();
        // child rows in registers when the largest shape fits a compile-time bound
        const int mu_km = MM <= 8 ? 8 : MM <= 14 ? 14 : 0;
        if (mu_km == 8) {
            MCE_DISPATCH_D(d, MCE_RUN((k_mu_r<DD, 8>), dim3(ceil_div(nslots, 256)), dim3(256), 0, st, IN,
                                                 np, slot_base.ptr, OUT, nslots, dconst.ptr, gk[gcur].ptr,
                                                 gv[gcur].ptr, gy.ptr, keys.ptr, chain_key.ptr, counters.ptr + 17));
        } else if (mu_km == 14 && d <= 5) {
            MCE_DISPATCH_D(d, if constexpr (DD <= 5) MCE_RUN(
                                  (k_mu_r<DD, 14>), dim3(ceil_div(nslots, 256)), dim3(256), 0, st, IN, np,
                                  slot_base.ptr, OUT, nslots, dconst.ptr, gk[gcur].ptr, gv[gcur].ptr, gy.ptr, keys.ptr,
                                  chain_key.ptr, counters.ptr + 17));
        } else {
            MCE_DISPATCH_D(d, MCE_RUN(k_mu<DD>, dim3(ceil_div(nslots, 256)), dim3(256), 0, st, IN, np,
                                                 slot_base.ptr, OUT, nslots, dconst.ptr, gk[gcur].ptr, gv[gcur].ptr,
                                                 gy.ptr, keys.ptr, chain_key.ptr, counters.ptr + 17));
        }
        HIPCHK(hipGetLastError());
        if (ntan) {
            tout.reserve(nslots);
            dg_slot.reserve((size_t)nslots * ntan);
            dfz_dev.reserve(ntan);
            MCE_DISPATCH_D(d, MCE_LAUNCH(k_mu_tan<DD>, dim3(ceil_div(nslots, 128)), dim3(128), 0, st, IN, tin.view(),
                                         np, slot_base.ptr, OUT, tout.view(), nslots, dconst.ptr, dstan.ptr,
                                         gk[gcur].ptr, gv[gcur].ptr, dgv[gcur].ptr, dg_slot.ptr));
            MCE_LAUNCH(k_tan_colsum, dim3(ntan), dim3(256), 0, st, dg_slot.ptr, nslots, ntan, dfz_dev.ptr);
        }
        prof.mark(kPhMu, st);

        // ---- moments: slots in reference-thread chain order, then the chains on st2
        MCE_RUN(k_iota, dim3(ceil_div(nslots, 256)), dim3(256), 0, st, iota.ptr, nslots);
        sort_pairs(chain_key.ptr, chain_skey.ptr, iota.ptr, chain_perm.ptr, nslots, 0, 4);
        MCE_RUN(k_chain_off, dim3(1), dim3(16), 0, st, chain_skey.ptr, nslots, T, chain_off.ptr);
        if (!bs) {
            HIPCHK(hipEventRecord(ev_mu, st));
            HIPCHK(hipStreamWaitEvent(st2, ev_mu, 0));
        }
        if (prof.on) HIPCHK(hipEventRecord(ev_c0, st2));
        {
            int ncp = 0, nyb = 0;
            MCE_DISPATCH_D(d, (nyb = mom_blocks<DD>(), ncp = nyb * kSumCW / 2));
            mom_x.reserve((size_t)nyb * nslots * kSumCW);
            MCE_DISPATCH_D(d, MCE_RUN(k_mom_terms<DD>, dim3(ceil_div((int64_t)nslots * ncp, 256)), dim3(256),
                                                 0, st2, gy.ptr, out.flags.ptr, chain_perm.ptr, chain_off.ptr, T,
                                                 nslots, mom_x.ptr));
            MCE_DISPATCH_D(d, MCE_RUN(k_mom_sum<DD>, dim3(T, nyb), dim3(256), 0, st2, mom_x.ptr,
                                                 chain_off.ptr, T, mom_part.ptr, counters.ptr + 17));
        }
        HIPCHK(hipGetLastError());
        if (prof.on) HIPCHK(hipEventRecord(ev_c1, st2));
        cur_T = T;
        moments_pending = true;
        diag_mom.assign(6 + 2 * d + 2 * d * d, 0.0);
        diag_cnt.assign(4 + 2 * sr, 0);
        if (first || skip) co_await collect_moments(first);

        if (skip) {
            MCE_RUN(k_hist, dim3(std::min(ceil_div(nslots, 256), 1024)), dim3(256), 0, st, nslots, sr + 1, 0,
                               out.flags.ptr, out.m.ptr, (const int*)nullptr, (const uint32_t*)nullptr, hist.ptr);
            std::vector<int> h(sr + 1);
            co_await rd(hist.ptr, h.data(), sr + 1);
            if (prof.on && !bs)
                for (int m = 0; m < sr; m++) prof.alg_bytes += h[m] * (2 * term_record_bytes(m) + 2.0 * m);
            Nt = 0;
            for (int m = 0; m < sr; m++) {
                diag_cnt[4 + m] = h[m];
                diag_cnt[4 + sr + m] = h[m];
                Nt += h[m];
            }
            diag_cnt[1] = Nt;
            terms_generated += Nt;
        } else {
            // ---- gather into shape buckets in the reference's term order
            iota.reserve(nslots);
            perm.reserve(nslots);
            skeys.reserve(nslots);
            // iota still holds 0..nslots-1: the chain sort only read it
            sort_pairs(keys.ptr, skeys.ptr, iota.ptr, perm.ptr, nslots, 0, kGatherKeyBits);
            MCE_RUN(k_hist, dim3(std::min(ceil_div(nslots, 256), 1024)), dim3(256), 0, st, nslots, sr + 1, 0,
                               out.flags.ptr, out.m.ptr, (const int*)nullptr, (const uint32_t*)nullptr, hist.ptr);
            std::vector<int> cnt(sr + 1);
            co_await rd(hist.ptr, cnt.data(), sr + 1);
            if (prof.on && !bs)
                for (int m = 0; m < sr; m++) prof.alg_bytes += cnt[m] * (2 * term_record_bytes(m) + 2.0 * m);
            prof.mark(kPhGather, st);
            int nt_mu = 0;
            for (int m = 0; m < sr; m++) {
                diag_cnt[4 + m] = cnt[m];
                nt_mu += cnt[m];
            }
            diag_cnt[1] = nt_mu;
            terms_generated += nt_mu;
            if (first) {
                co_await first_gtables(nt_mu);
            } else {
                co_await reduce_and_flatten(cnt);
            }
            for (int m = 0; m < sr; m++) diag_cnt[4 + sr + m] = in_count[m];
        }
        {
            const int guard = co_await rd1(counters.ptr + 8);
            if (guard) {
                char msg[160];
                snprintf(msg, sizeof(msg), "device bound guard tripped (flags 0x%x) at update %d", guard, master_step);
                throw std::runtime_error(msg);
            }
        }
        if (!skip) {
            if (debug_after_ftr)
                co_await moments_after_ftr();
            else
                fz = cmake(1, 0);
        }
        if (first) {
            last_mean = mean;
            last_cov = cov;
            last_fz = fz;
        }
        master_step++;
        prof.end(st);
        diag_mom[0] = fz.re;
        diag_mom[1] = fz.im;
        diag_mom[4] = last_fz.re;
        diag_mom[5] = last_fz.im;
        diag_cnt[0] = errs;
        diag_cnt[2] = Nt;
        diag_cnt[3] = threads_used;
        have_diag = true;
        co_return errs;
    }

    void widen(const int* src, int64_t* dst, int n);

    // g-tables of the first update's d+1 terms, which become the parents.
    Co<void> first_gtables(int n) {
        const int cells = 1 << (d - 1);
        if (prof.on && !bs) prof.alg_bytes += 24.0 * n * cells;
        const int nx = gcur ^ 1;
        gk[nx].reserve((size_t)n * cells);
        gv[nx].reserve((size_t)n * cells);
        in.reserve(n);
        TermPool OUT = out.view();
        TermPool NEXT = in.view();
        MCE_DISPATCH_D(d, MCE_RUN(k_gtable_first<DD>, dim3(1), dim3(64), 0, st, OUT, n, perm.ptr, gscale,
                                             gk[nx].ptr, gv[nx].ptr, NEXT));
        if (ntan) {
            dgv[nx].reserve((size_t)n * cells * ntan);
            tin.reserve(n);
            MCE_DISPATCH_D(d, MCE_LAUNCH(k_gtable_first_tan<DD>, dim3(ceil_div(n, 64)), dim3(64), 0, st, OUT,
                                         tout.view(), n, perm.ptr, gscale, dgscale_dev.ptr, dgv[nx].ptr, tin.view()));
        }
        HIPCHK(hipGetLastError());
        co_await sync_point();
        gcur = nx;
        in_count.assign(sr, 0);
        in_count[d] = n;
        Nt = n;
        max_gp_cells = cells;
    }

    // Fast term reduction and g-tables for every shape bucket
    // (fast_term_reduction_and_create_gtables, cauchy_estimator.hpp:981-1177).
    Co<void> reduce_and_flatten(const std::vector<int>& cnt) {
        TermPool OUT = out.view();
        std::vector<int> off(sr + 1, 0);
        for (int m = 0; m < sr; m++) off[m + 1] = off[m] + cnt[m];
        const int total = off[sr];
        grp_root.reserve(total);
        grp_moff.reserve(total);
        grp_mcnt.reserve(total);
        grp_m.reserve(total);
        mem_slots.reserve(total);
        std::vector<int> gbase_of;
        const int ng = co_await reduce_all(OUT, off, gbase_of);
        prof.mark(kPhFtr, st);

        // ---- group B-tables
        ub64.reserve(ng);
        b2off.reserve(ng);
        b2n.reserve(ng);
        b2n64.reserve(ng);
        goff.reserve(ng);
        MCE_RUN(k_group_ub, dim3(ceil_div(ng, 256)), dim3(256), 0, st, OUT, ng, grp_root.ptr, grp_m.ptr, d,
                           (int*)b2n.ptr);
        const int64_t ub_total = co_await widen_scan_total(b2n.ptr, ub64.ptr, b2off.ptr, ng);
        b2arena.reserve((size_t)ub_total + 1);
        for (int m = 1; m < sr; m++) {
            const int ngb = gbase_of[m + 1] - gbase_of[m];
            if (ngb == 0) continue;
            const int words = m >= 5 ? (1 << (m - 5)) : 1;
            int wpb = 4;
            while (wpb > 1 && (size_t)wpb * words * 4 > 65536) wpb--;
            const size_t lds = (size_t)wpb * words * 4;
            MCE_RUN(k_group_btab, dim3(ceil_div(ngb, wpb)), dim3(64 * wpb), lds, st, OUT, gbase_of[m], ngb,
                               m, d, grp_root.ptr, barena.ptr, b2off.ptr, ub64.ptr, b2arena.ptr, b2n.ptr, words,
                               counters.ptr + 8);
            HIPCHK(hipGetLastError());
        }
        const int64_t cells_total = co_await widen_scan_total(b2n.ptr, b2n64.ptr, goff.ptr, ng);
        if (prof.on && !bs) prof.alg_bytes += 24.0 * (double)cells_total;
        prof.mark(kPhGBtab, st);

        // ---- group g-tables need the update's g scale 1/(2 pi Re fz)
        co_await collect_moments(false);
        const int nx = gcur ^ 1;
        gk[nx].reserve((size_t)cells_total + 1);
        gv[nx].reserve((size_t)cells_total + 1);
        scratch.reserve((size_t)cells_total + 1);
        if (ntan) {
            // which members each group adds, written by k_group_eval / k_group_members
            mem_add.reserve(std::max(total, 1));
            fill_multi({{mem_add.ptr, (size_t)std::max(total, 1), 0}});
        }
        surv.reserve(ng);
        gxor.reserve(ng);
        sort_list.reserve(ng);
        if (prof.on && !bs) {
            if (!prof.k0) {
                HIPCHK(hipEventCreate(&prof.k0));
                HIPCHK(hipEventCreate(&prof.k1));
            }
            HIPCHK(hipEventRecord(prof.k0, st));
        }
        int ge_mode = 0, ge_wmax = 0;
        size_t ge_lds = 0;
        {
            int mode, wmax;
            size_t lds;
            pick_lookup(mode, lds, wmax);
            if (mode != upd_mode) throw std::logic_error("g-table lookup mode changed within an update");
            geval_mode = mode;
            ge_mode = mode;
            ge_lds = lds;
            ge_wmax = wmax;
            const int big_members = mode >= kLookupRank ? big_group_members : INT_MAX;
            defer_k0.reserve(std::max(ng, 1));
#define MCE_GEVAL(M_, K_)                                                                                        \
    do {                                                                                                          \
        MCE_RUN((k_group_eval<M_, K_>), dim3(ng), dim3(64), lds, st, OUT, ng, d, wmax, grp_root.ptr,    \
                           grp_moff.ptr, grp_mcnt.ptr, grp_m.ptr, mem_slots.ptr, b2arena.ptr, b2off.ptr,           \
                           b2n.ptr, goff.ptr, gk[gcur].ptr, gv[gcur].ptr, gk[nx].ptr, gv[nx].ptr, scratch.ptr,     \
                           gscale, surv.ptr, gxor.ptr, sort_list.ptr, counters.ptr + 18, counters.ptr + 19,          \
                           big_members, defer_k0.ptr, ntan ? mem_add.ptr : nullptr, (const uint2*)rank_arena.ptr);  \
    } while (0)
            // rows held in registers: the smallest compile-time bound >= the largest shape
            if (mode == kLookupRankVals && MM <= 8) MCE_GEVAL(kLookupRankVals, 8);
            else if (mode == kLookupRankVals && MM <= 11) MCE_GEVAL(kLookupRankVals, 11);
            else if (mode == kLookupRankVals && MM <= 14) MCE_GEVAL(kLookupRankVals, 14);
            else if (mode == kLookupRankVals && MM <= 16) MCE_GEVAL(kLookupRankVals, 16);
            else if (mode == kLookupRankVals && MM <= 24) MCE_GEVAL(kLookupRankVals, 24);
            else if (mode == kLookupRankVals) MCE_GEVAL(kLookupRankVals, 32);
            else if (mode == kLookupRank && MM <= 8) MCE_GEVAL(kLookupRank, 8);
            else if (mode == kLookupRank && MM <= 14) MCE_GEVAL(kLookupRank, 14);
            else if (mode == kLookupRank && MM <= 16) MCE_GEVAL(kLookupRank, 16);
            else if (mode == kLookupRank) MCE_GEVAL(kLookupRank, 32);
            else if (mode == kLookupSearchLds) MCE_GEVAL(kLookupSearchLds, 32);
            else MCE_GEVAL(kLookupSearch, 32);
#undef MCE_GEVAL
        }
        HIPCHK(hipGetLastError());
        if (prof.on && !bs) {
            HIPCHK(hipEventRecord(prof.k1, st));
            bytes_acc.reserve(1);
            fill_multi({{bytes_acc.ptr, sizeof(unsigned long long), 0}});
            MCE_RUN(k_geval_bytes, dim3(ceil_div(ng, 256)), dim3(256), 0, st, ng, d, grp_m.ptr,
                               grp_mcnt.ptr, b2n.ptr, surv.ptr, defer_k0.ptr,
                               ge_mode >= kLookupRank ? big_group_members : INT_MAX, bytes_acc.ptr);
            const unsigned long long bytes = read1(bytes_acc.ptr);
            float t = 0;
            HIPCHK(hipEventElapsedTime(&t, prof.k0, prof.k1));
            prof.geval_ms += t;
            prof.geval_bytes += (double)bytes;
            prof.geval_launches++;
            pstat.reserve(2);
            const int* mc = grp_mcnt.ptr;
            int* ps = pstat.ptr;
            reduce_max(mc, ps, ng);
            cub("reduce sum", [=](void* t_, size_t& b_, hipStream_t s_) {
                return hipcub::DeviceReduce::Sum(t_, b_, mc, ps + 1, ng, s_);
            });
            int ps_[2];
            HIPCHK(hipMemcpyAsync(ps_, pstat.ptr, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            prof.groups += ng;
            prof.members += ps_[1];
            prof.max_group = std::max(prof.max_group, (double)(ps_[0] + 1));
        }
        if (bs && bs->geval_bytes)
            MCE_RUN(k_geval_bytes, dim3(ceil_div(ng, 256)), dim3(256), 0, st, ng, d, grp_m.ptr, grp_mcnt.ptr, b2n.ptr,
                    surv.ptr, defer_k0.ptr, ge_mode >= kLookupRank ? big_group_members : INT_MAX, bs->geval_bytes);
        if (ge_mode >= kLookupRank) co_await big_groups(ng, ge_mode, ge_lds, ge_wmax, gscale, nx);
        if (prof.on && !bs) prof.alg_bytes += 24.0 * prof_device_sum(ng, grp_mcnt.ptr, b2n.ptr, surv.ptr);
        int nr_[2];
        co_await rd(counters.ptr + 18, nr_, 2);
        const int nsort = nr_[0];
        const int rare = nr_[1];
        // bits 1 and 2 note parent (old) terms among a group's members, which stage_term
        // evaluates through their own tables; bit 4 is a broken invariant
        if (rare & 4) throw std::runtime_error("parent g-table key outside its half-storage range");
        if (nsort > 0) {
            tmpk.reserve((size_t)cells_total + 1);
            tmpv.reserve((size_t)cells_total + 1);
            const int words = MM >= 5 ? (1 << (MM - 5)) : 1;
            const size_t lds = (size_t)words * 4;
            MCE_RUN(k_group_sort, dim3(nsort), dim3(256), lds, st, nsort, sort_list.ptr, grp_m.ptr, b2n.ptr,
                               goff.ptr, gk[nx].ptr, gv[nx].ptr, tmpk.ptr, tmpv.ptr);
            HIPCHK(hipGetLastError());
        }
        if (ntan) {
            dgv[nx].reserve(((size_t)cells_total + 1) * ntan);
            MCE_LAUNCH(k_group_tan, dim3(ng), dim3(64), 0, st, OUT, tout.view(), ng, d, grp_moff.ptr, grp_mcnt.ptr,
                       grp_m.ptr, mem_slots.ptr, mem_add.ptr, surv.ptr, gxor.ptr, b2arena.ptr, b2off.ptr, b2n.ptr,
                       goff.ptr, gk[gcur].ptr, gv[gcur].ptr, dgv[gcur].ptr, gk[nx].ptr, gscale, dgscale_dev.ptr,
                       dgv[nx].ptr);
        }
        prof.mark(kPhGEval, st);

        // ---- survivors become the next parents (shape ascending, root index ascending)
        sflag.reserve(ng);
        spos.reserve(ng);
        MCE_RUN(k_surv_flags, dim3(ceil_div(ng, 256)), dim3(256), 0, st, ng, surv.ptr, sflag.ptr);
        scan_i32_excl(sflag.ptr, spos.ptr, ng);
        in.reserve(std::max(ng, 1));  // survivors <= groups; their count comes with the histogram
        TermPool NEXT = in.view();
        MCE_DISPATCH_D(d, MCE_RUN(k_compact<DD>, dim3(ceil_div(ng, 256)), dim3(256), 0, st, OUT, ng,
                                             surv.ptr, spos.ptr, grp_m.ptr, goff.ptr, b2n.ptr, NEXT));
        if (ntan) {
            tin.reserve(std::max(ng, 1));
            MCE_DISPATCH_D(d, MCE_LAUNCH(k_compact_tan<DD>, dim3(ceil_div(ng, 256)), dim3(256), 0, st, tout.view(), ng,
                                         surv.ptr, spos.ptr, grp_m.ptr, tin.view()));
        }
        HIPCHK(hipGetLastError());
        MCE_RUN(k_hist, dim3(std::min(ceil_div(ng, 256), 1024)), dim3(256), 0, st, ng, sr + 1, 1,
                           (const uint8_t*)nullptr, grp_m.ptr, surv.ptr, (const uint32_t*)nullptr, hist_mx.ptr);
        std::vector<int> h(sr + 2);
        {
            // histogram and the largest g-table (the next update's parent tables) in one read
            reduce_max(b2n.ptr, hist_mx.ptr + sr + 1, ng);
            co_await rd(hist_mx.ptr, h.data(), sr + 2);
            max_gp_cells = h[sr + 1];
        }
        prof.mark(kPhCompact, st);
        in_count.assign(sr, 0);
        Nt = 0;
        for (int m = 0; m < sr; m++) {
            in_count[m] = h[m];
            Nt += h[m];
        }
        gcur = nx;
    }

    // Groups with more than big_group_members members: their members' contributions are
    // evaluated one wavefront per member (k_group_members) and added in member order
    // (k_group_sum), after k_group_eval has settled each group's root.
    Co<void> big_groups(int ng, int mode, size_t lds, int wmax, double gscale, int nx) {
        if (ng <= 0) co_return;
        big_icnt.reserve(ng);
        big_ioff.reserve(ng);
        big_hcnt.reserve(ng);
        big_hoff.reserve(ng);
        big_ccnt.reserve(ng);
        big_coff.reserve(ng);
        MCE_RUN(k_big_prep, dim3(ceil_div(ng, 256)), dim3(256), 0, st, ng, big_group_members, grp_mcnt.ptr,
                           b2n.ptr, big_icnt.ptr, big_ccnt.ptr, big_hcnt.ptr);
        scan_i32_excl(big_icnt.ptr, big_ioff.ptr, ng);
        scan_i32_excl(big_hcnt.ptr, big_hoff.ptr, ng);
        scan_i64_excl(big_ccnt.ptr, big_coff.ptr, ng);
        enqueue_total(big_ioff.ptr, big_icnt.ptr, ng, 0);
        enqueue_total(big_hoff.ptr, big_hcnt.ptr, ng, 1);
        enqueue_total(big_coff.ptr, big_ccnt.ptr, ng, 2);
        int64_t tt[3];
        co_await rd(d_scal.ptr, tt, 3);
        const int ni = (int)tt[0];
        if (ni == 0) co_return;
        const int nh = (int)tt[1];
        const int64_t nc = tt[2];
        contrib.reserve((size_t)std::max<int64_t>(nc, 1));
        cneg.reserve(ni);
        TermPool OUT = out.view();
#define MCE_GMEM(M_, K_)                                                                                         \
    do {                                                                                                          \
        MCE_RUN((k_group_members<M_, K_>), dim3(ni), dim3(64), lds, st, OUT, ng, d, wmax, big_ioff.ptr, \
                           ni, big_coff.ptr, grp_moff.ptr, grp_mcnt.ptr, grp_m.ptr, mem_slots.ptr, b2arena.ptr,    \
                           b2off.ptr, b2n.ptr, gk[gcur].ptr, gv[gcur].ptr, gscale, surv.ptr, gxor.ptr,             \
                           defer_k0.ptr, contrib.ptr, cneg.ptr, counters.ptr + 19, ntan ? mem_add.ptr : nullptr,    \
                           (const uint2*)rank_arena.ptr);                                                          \
    } while (0)
        if (mode == kLookupRankVals && MM <= 8) MCE_GMEM(kLookupRankVals, 8);
        else if (mode == kLookupRankVals && MM <= 11) MCE_GMEM(kLookupRankVals, 11);
        else if (mode == kLookupRankVals && MM <= 14) MCE_GMEM(kLookupRankVals, 14);
        else if (mode == kLookupRankVals && MM <= 16) MCE_GMEM(kLookupRankVals, 16);
        else if (mode == kLookupRankVals && MM <= 24) MCE_GMEM(kLookupRankVals, 24);
        else if (mode == kLookupRankVals) MCE_GMEM(kLookupRankVals, 32);
        else if (mode == kLookupRank && MM <= 8) MCE_GMEM(kLookupRank, 8);
        else if (mode == kLookupRank && MM <= 14) MCE_GMEM(kLookupRank, 14);
        else if (mode == kLookupRank && MM <= 16) MCE_GMEM(kLookupRank, 16);
        else MCE_GMEM(kLookupRank, 32);
#undef MCE_GMEM
        HIPCHK(hipGetLastError());
        MCE_RUN(k_group_sum, dim3(nh), dim3(64), 0, st, ng, big_hoff.ptr, nh, big_ioff.ptr, big_coff.ptr,
                           grp_mcnt.ptr, b2n.ptr, goff.ptr, surv.ptr, defer_k0.ptr, contrib.ptr, cneg.ptr, gv[nx].ptr);
        HIPCHK(hipGetLastError());
    }

    // Fast term reduction of every shape bucket at once.  off[m] is the first gathered
    // position of shape m; fills the reduction groups (shape ascending, roots ascending,
    // members ascending per root) and gbase_of[m], the first group of shape m.
    Co<int> reduce_all(const TermPool& OUT, const std::vector<int>& off, std::vector<int>& gbase_of) {
        const int N = off[sr];
        gbase_of.assign(sr + 1, 0);
        if (N == 0) co_return 0;
        const int a0 = g_search_order[0];
        if (a0 < 0 || a0 >= d)
            throw std::runtime_error("term-reduction search ordering names axis " + std::to_string(a0) + " of a " +
                                     std::to_string(d) + "-state estimator");
        h_boff = off;
        dboff.reserve(sr + 1);
        upload(dboff.ptr, h_boff.data(), (sr + 1) * sizeof(int));
        const int* boff = dboff.ptr;
        fkeys.reserve(N);
        fskeys.reserve(N);
        fidx.reserve(N);
        fsidx.reserve(N);
        bkey.reserve(N);
        bskey.reserve(N);
        fvals.reserve(N);
        ecnt.reserve(N);
        eoff.reserve(N);
        const dim3 g1(ceil_div(N, 256)), b1(256);
        // value order on the first search axis, then (stable) bucket order
        MCE_RUN(k_ftr_keys, g1, b1, 0, st, OUT, perm.ptr, N, d, a0, fkeys.ptr, fidx.ptr);
        sort_pairs(fkeys.ptr, fskeys.ptr, fidx.ptr, fsidx.ptr, N, 0, 64);
        MCE_RUN(k_ftr_bkeys, g1, b1, 0, st, N, boff, sr, fsidx.ptr, bkey.ptr);
        int bbits = 1;
        while ((1 << bbits) < sr) bbits++;
        sort_pairs(bkey.ptr, bskey.ptr, fsidx.ptr, fidx.ptr, N, 0, bbits);
        const uint32_t* sidx = fidx.ptr;
        MCE_RUN(k_ftr_vals, g1, b1, 0, st, OUT, perm.ptr, N, d, a0, sidx, fvals.ptr);
        // candidate windows on the first search axis; the predicate runs one thread per
        // (term, window slot) pair unless the pairs would not fit, then one per term
        wlo.reserve(N);
        wcnt.reserve(N);
        woff.reserve(N);
        MCE_RUN(k_ftr_win, g1, b1, 0, st, N, boff, sr, fvals.ptr, wlo.ptr, wcnt.ptr, ecnt.ptr);
        const int64_t npairs = co_await scan_excl_total(wcnt.ptr, woff.ptr, N);
        const bool by_pairs = npairs < ((int64_t)1 << 31);
        if (by_pairs && npairs > 0) {
            fgb.reserve((size_t)N * d);
            fgp.reserve((size_t)N * MM);
            fgA.reserve((size_t)N * MM);
            pflag.reserve((size_t)npairs);
            MCE_DISPATCH_D(d, MCE_RUN(k_ftr_gather<DD>, g1, b1, 0, st, OUT, perm.ptr, N, boff, sr, sidx,
                                                 fgb.ptr, fgp.ptr, fgA.ptr));
            MCE_DISPATCH_D(d, MCE_RUN(k_ftr_pairs<DD>, dim3(ceil_div(npairs, 256)), b1, 0, st, npairs, N,
                                                 boff, sr, dconst.ptr, woff.ptr, wlo.ptr, sidx, MM, fgb.ptr, fgp.ptr,
                                                 fgA.ptr, pflag.ptr, ecnt.ptr, eoff.ptr, cursor.ptr, edges.ptr, 0));
        } else if (!by_pairs) {
            MCE_DISPATCH_D(d, MCE_RUN(k_ftr_edges<DD>, g1, b1, 0, st, OUT, perm.ptr, N, boff, sr,
                                                 dconst.ptr, fvals.ptr, sidx, ecnt.ptr, eoff.ptr, edges.ptr, 0));
        }
        HIPCHK(hipGetLastError());
        const int E = (int)(co_await scan_excl_total(ecnt.ptr, eoff.ptr, N));
        Fc.reserve((size_t)8 * N);
        MCE_RUN(k_ftr_finit, g1, b1, 0, st, Fc.ptr, N, boff, sr);
        if (E > 0) {
            edges.reserve(E);
            inedges.reserve(E);
            cursor.reserve(N);
            cursor2.reserve(N);
            indeg.reserve(N);
            inoff.reserve(N);
            st0.reserve(N);
            st1.reserve(N);
            round_slots.reserve(kRoundSlots);
            // every buffer the edge passes and the rounds start from, in one launch
            fill_multi({{cursor.ptr, N * sizeof(int), 0},
                        {cursor2.ptr, N * sizeof(int), 0},
                        {indeg.ptr, N * sizeof(int), 0},
                        {st0.ptr, (size_t)N, kUnknown},
                        {round_slots.ptr, kRoundSlots * sizeof(int), 0}});
            if (by_pairs) {
                MCE_DISPATCH_D(d, MCE_RUN(k_ftr_pairs<DD>, dim3(ceil_div(npairs, 256)), b1, 0, st, npairs,
                                                     N, boff, sr, dconst.ptr, woff.ptr, wlo.ptr, sidx, MM, fgb.ptr,
                                                     fgp.ptr, fgA.ptr, pflag.ptr, ecnt.ptr, eoff.ptr, cursor.ptr,
                                                     edges.ptr, 1));
            } else {
                MCE_DISPATCH_D(d, MCE_RUN(k_ftr_edges<DD>, g1, b1, 0, st, OUT, perm.ptr, N, boff, sr,
                                                     dconst.ptr, fvals.ptr, sidx, ecnt.ptr, eoff.ptr, edges.ptr, 1));
            }
            MCE_RUN(k_ftr_inedges, g1, b1, 0, st, N, boff, sr, eoff.ptr, ecnt.ptr, edges.ptr, indeg.ptr,
                               (const int*)nullptr, (int*)nullptr, (int*)nullptr, 0);
            scan_i32_excl(indeg.ptr, inoff.ptr, N);
            MCE_RUN(k_ftr_inedges, g1, b1, 0, st, N, boff, sr, eoff.ptr, ecnt.ptr, edges.ptr, indeg.ptr,
                               inoff.ptr, cursor2.ptr, inedges.ptr, 1);
            uint8_t* sa = st0.ptr;
            uint8_t* sb = st1.ptr;
            // Rounds are idempotent once every term is resolved, so several run between
            // reads of the unresolved count (taken from the last round of a batch).
            int batch = 2, round = 0;
            for (int done = 0; done <= N + 1;) {
                for (int r = 0; r < batch; r++, round++) {
                    MCE_RUN(k_ftr_round, g1, b1, 0, st, N, inoff.ptr, indeg.ptr, inedges.ptr, sa, sb,
                                       round_slots.ptr, round);
                    std::swap(sa, sb);
                }
                done += batch;
                if ((co_await rd1(round_slots.ptr + (round - 1) % kRoundSlots)) == 0) break;
                batch = std::min(batch * 2, 16);
            }
            MCE_RUN(k_ftr_claim, g1, b1, 0, st, N, boff, sr, eoff.ptr, ecnt.ptr, edges.ptr, sa, Fc.ptr);
            if (N > 10000) {
                for (int lvl = 0; lvl < 3; lvl++)
                    MCE_RUN(k_ftr_merge, dim3(ceil_div(N, 256), 8 >> (lvl + 1)), b1, 0, st, N, boff, sr,
                                       Fc.ptr, lvl);
            }
            HIPCHK(hipGetLastError());
        }
        // reduction groups
        isroot.reserve(N);
        memkey.reserve(N);
        memval.reserve(N);
        smkey.reserve(N);
        smval.reserve(N);
        roots.reserve(N);
        jiota.reserve(N);
        memcnt.reserve(N + 1);
        memoff.reserve(N + 1);
        MCE_RUN(k_ftr_rootflags, g1, b1, 0, st, N, boff, sr, Fc.ptr, isroot.ptr, memkey.ptr, memval.ptr);
        MCE_RUN(k_iota, g1, b1, 0, st, jiota.ptr, N);
        select_flagged(jiota.ptr, isroot.ptr, roots.ptr, counters.ptr + 5, N);
        int kbits = 1;
        while (((int64_t)1 << kbits) <= N) kbits++;
        sort_pairs(memkey.ptr, smkey.ptr, memval.ptr, smval.ptr, N, 0, kbits);
        gbase_dev.reserve(sr + 1);
        const int* d_nroots = counters.ptr + 5;
        MCE_RUN(k_ftr_gbase, dim3(1), dim3(64), 0, st, d_nroots, roots.ptr, boff, sr, gbase_dev.ptr);
        fill_multi({{memcnt.ptr, (N + 1) * sizeof(int), 0}});
        MCE_RUN(k_ftr_memcount, g1, b1, 0, st, N, d_nroots, smkey.ptr, memcnt.ptr);
        MCE_RUN(k_ftr_memslots, g1, b1, 0, st, N, d_nroots, smval.ptr, perm.ptr, mem_slots.ptr);
        scan_i32_excl(memcnt.ptr, memoff.ptr, N);
        MCE_RUN(k_ftr_groups, g1, b1, 0, st, d_nroots, roots.ptr, perm.ptr, boff, sr, memcnt.ptr,
                           memoff.ptr, grp_root.ptr, grp_moff.ptr, grp_mcnt.ptr, grp_m.ptr);
        HIPCHK(hipGetLastError());
        // group offsets per shape; the last one is the root count
        co_await rd(gbase_dev.ptr, gbase_of.data(), sr + 1);
        co_return gbase_of[sr];
    }

    // Deterministic transform of every parent term (cauchy_estimator.hpp:1312-1355).
    void deterministic_transform(const double* Trans, const double* bias);

    // 1D marginal density of state `marg` on the grid lo, lo + res, ... (clamped to hi), as
    // CauchyCPDFGridDispatcher1D::evaluate_point_grid does (cpdf_ndim.hpp:2061-2130): the
    // first point by the direct per-term formula, every later point by the cached form.
    // Fills xy with (x, f(x)) pairs; returns 1 (no points) after a window's last update,
    // whose g-tables are not built.
    int cpdf_1d(int marg, double lo, double hi, double res, std::vector<double>& xy);
    // The same density written to device arrays xs / ys (capacity points each) on `s`,
    // with every intermediate in `work` (work_bytes): no device allocation.
    int cpdf_1d_into(int marg, double lo, double hi, double res, double* xs_dev, double* ys_dev, int capacity,
                     hipStream_t s, char* work, size_t work_bytes, int* npts_out);
    // Checks shared by both forms; returns the point count, or 0 after the window's last
    // update (whose g-tables were not built).
    int cpdf_1d_points(int marg, double lo, double hi, double res);
    // 2D marginal density of states (i1, i2) on the grid of CauchyCPDFGridDispatcher2D
    // (cpdf_ndim.hpp:1816-1900): x runs fastest; fills xyz with (x, y, f) triples.
    int cpdf_2d(int i1, int i2, double lox, double hix, double resx, double loy, double hiy, double resy,
                std::vector<double>& xyz, int& nx, int& ny);
    void draw_bar_nu();
    std::vector<double> bar_nu;  // drawn at the first density evaluation (cpdf_ndim.hpp:384-387)
    DBuf<double> cpdf_direct, cpdf_x, cpdf_y, cpdf_part, cpdf_sum;
    DBuf<CpdfTerm1D> cpdf_cache;
    // relative-system scratch (held by the primary estimator of a pair)
    DBuf<double> rs_mom, rs_sums, rs_part;
    DBuf<uint8_t> rs_keep;
    DBuf<int> rs_idx, rs_cnt;
    DBuf<int> c2_mc, c2_err;
    DBuf<double> c2_b2, c2_sn, c2_cs, c2_g1r, c2_g2r, c2_px, c2_py;
    DBuf<cplx> c2_g;
    DBuf<Cell2D> c2_fc;

    // Extended-mode finish of a measurement update (cauchy_estimator.hpp:1358-1394): the
    // terms' b-vectors move by -Re(mean) (skipped on a window's last update), the mean
    // becomes x_bar + mean, and x_bar takes its real part.
    void finalize_extended_moments(double* x_bar) { run_sync(finalize_extended_co(x_bar)); }
    Co<void> finalize_extended_co(double* x_bar);

    // Host copy of parent term `idx` of bucket `shape`.
    int term(int shape, int idx, int* info, double* A, double* p_, double* b, unsigned int* keys_, double* g) {
        bind();
        if (shape < 0 || shape >= sr || idx < 0 || idx >= in_count[shape]) return -1;
        int o = 0;
        for (int m = 0; m < shape; m++) o += in_count[m];
        o += idx;
        int mm;
        int gn;
        int64_t go;
        copy_sync(&mm, in.m.ptr + o, sizeof(int), hipMemcpyDeviceToHost);
        copy_sync(&gn, in.gp_n.ptr + o, sizeof(int), hipMemcpyDeviceToHost);
        copy_sync(&go, in.gp_off.ptr + o, sizeof(int64_t), hipMemcpyDeviceToHost);
        info[0] = mm;
        info[1] = gn;
        info[2] = mm;
        if (A) copy_sync(A, in.A.ptr + (int64_t)o * MM * d, mm * d * sizeof(double), hipMemcpyDeviceToHost);
        if (p_) copy_sync(p_, in.p.ptr + (int64_t)o * MM, mm * sizeof(double), hipMemcpyDeviceToHost);
        if (b) copy_sync(b, in.b.ptr + (int64_t)o * d, d * sizeof(double), hipMemcpyDeviceToHost);
        if (keys_ && g && gn > 0) {
            copy_sync(keys_, gk[gcur].ptr + go, gn * sizeof(uint32_t), hipMemcpyDeviceToHost);
            copy_sync(g, gv[gcur].ptr + go, gn * sizeof(cplx), hipMemcpyDeviceToHost);
        }
        return 0;
    }
};

}  // namespace

void Engine::widen(const int* src, int64_t* dst, int n) {
    if (n <= 0) return;
    MCE_RUN(k_widen, dim3(ceil_div(n, 256)), dim3(256), 0, st, src, dst, n);
}

void Engine::deterministic_transform(const double* Trans, const double* bias) {
    if (ntan) throw std::runtime_error("the deterministic transform carries no tangents");
    bind();
    DBuf<double> dT;
    dT.reserve(d * d + d);
    copy_sync(dT.ptr, Trans, d * d * sizeof(double), hipMemcpyHostToDevice);
    const bool shift = !skip && bias != nullptr;
    if (shift) copy_sync(dT.ptr + d * d, bias, d * sizeof(double), hipMemcpyHostToDevice);
    TermPool IN = in.view();
    MCE_DISPATCH_D(d, MCE_LAUNCH(k_det_tp<DD>, dim3(ceil_div(Nt, 256)), dim3(256), 0, st, IN, Nt, dT.ptr,
                                         shift ? dT.ptr + d * d : nullptr));
    HIPCHK(hipStreamSynchronize(st));
}

int Engine::cpdf_1d_points(int marg, double lo, double hi, double res) {
    if (master_step == 0) throw std::runtime_error("the density needs at least one measurement update");
    if (marg < 0 || marg >= d) throw std::runtime_error("marginal index outside the state");
    if (!(hi > lo) || !(res > 0)) throw std::runtime_error("the grid needs high > low and resolution > 0");
    if (master_step == n_mu) {
        fprintf(stderr, "[cauchy_mi355x] the density cannot be evaluated after the window's last update: its g-tables "
                        "were not built\n");
        return 0;
    }
    if ((hi - lo) / res > 1e8) throw std::runtime_error("the grid would exceed 1e8 points");
    return (int)((hi - lo + res - 1e-15) / res) + 1;
}

int Engine::cpdf_1d_into(int marg, double lo, double hi, double res, double* xs_dev, double* ys_dev, int capacity,
                         hipStream_t s, char* work, size_t work_bytes, int* npts_out) {
    *npts_out = 0;
    const int npts = cpdf_1d_points(marg, lo, hi, res);
    if (npts == 0) return 1;
    if (!xs_dev || !ys_dev || capacity < npts)
        throw std::runtime_error("the grid has " + std::to_string(npts) + " points; the output arrays hold " +
                                 std::to_string(capacity));
    // working memory: direct values and cached terms (40 B per term), the chunk partial
    // sums (8 B per chunk of terms and grid point) and the direct sum
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    const int nchunks = ceil_div(Nt, kCpdfChunk);
    const size_t b_direct = up((size_t)Nt * sizeof(double)), b_cache = up((size_t)Nt * sizeof(CpdfTerm1D));
    const size_t b_part = up((size_t)nchunks * (npts - 1) * sizeof(double));
    const size_t need = b_direct + b_cache + b_part + 256;
    if (need > work_bytes)
        throw std::runtime_error("the density needs " + std::to_string(need) + " B of working memory (" +
                                 std::to_string(Nt) + " terms, " + std::to_string(npts) +
                                 " points); the no-allocation form has " + std::to_string(work_bytes));
    draw_bar_nu();
    double* direct = (double*)work;
    CpdfTerm1D* cache = (CpdfTerm1D*)(work + b_direct);
    double* part = (double*)(work + b_direct + b_cache);
    double* dsum = (double*)(work + b_direct + b_cache + b_part);
    MCE_LAUNCH(k_cpdf1d_xs, dim3(ceil_div(npts, 256)), dim3(256), 0, s, xs_dev, npts, lo, res, hi);
    const double norm = fz.re;
    if (Nt > 0) {
        const double x0 = std::min(lo + 0 * res, hi);  // the first abscissa, formed as the host grid forms it
        VecArg bn;
        for (int i = 0; i < d; i++) bn.v[i] = bar_nu[i];
        TermPool IN = in.view();
        MCE_DISPATCH_D(d, MCE_LAUNCH(k_cpdf1d_terms<DD>, dim3(ceil_div(Nt, 256)), dim3(256), 0, s, IN, Nt, marg, bn,
                                     gk[gcur].ptr, gv[gcur].ptr, x0, direct, cache));
        MCE_LAUNCH(k_sum_fixed, dim3(1), dim3(1024), 0, s, direct, Nt, dsum);
        if (npts > 1) {
            MCE_LAUNCH(k_cpdf1d_grid, dim3(ceil_div(npts - 1, kCpdfPoints), nchunks), dim3(kCpdfPoints), 0, s, cache,
                       Nt, xs_dev + 1, npts - 1, part);
            MCE_LAUNCH(k_cpdf1d_finish, dim3(ceil_div(npts - 1, 256)), dim3(256), 0, s, part, nchunks, npts - 1, norm,
                       ys_dev + 1);
        }
        MCE_LAUNCH(k_cpdf1d_first, dim3(1), dim3(64), 0, s, dsum, norm, ys_dev, 0);
    } else {
        MCE_LAUNCH(k_cpdf1d_first, dim3(ceil_div(npts, 256)), dim3(256), 0, s, dsum, norm, ys_dev, npts);
    }
    HIPCHK(hipGetLastError());
    *npts_out = npts;
    return 0;
}

int Engine::cpdf_1d(int marg, double lo, double hi, double res, std::vector<double>& xy) {
    const int npts = cpdf_1d_points(marg, lo, hi, res);
    if (npts == 0) return 1;
    draw_bar_nu();
    bind();  // after the draws: a thread's first stream creation must not come between them
    std::vector<double> xs(npts), ys(npts, 0.0);
    for (int i = 0; i < npts; i++) xs[i] = std::min(lo + i * res, hi);
    const double norm = fz.re;
    if (Nt > 0) {
        GraveyardScope gs(&graveyard);
        const int nchunks = ceil_div(Nt, kCpdfChunk);
        cpdf_direct.reserve(Nt);
        cpdf_cache.reserve(Nt);
        cpdf_x.reserve(npts);
        cpdf_y.reserve(npts);
        cpdf_part.reserve((size_t)nchunks * npts);
        cpdf_sum.reserve(1);
        HIPCHK(hipMemcpyAsync(cpdf_x.ptr, xs.data(), npts * sizeof(double), hipMemcpyHostToDevice, st));
        VecArg bn;
        for (int i = 0; i < d; i++) bn.v[i] = bar_nu[i];
        TermPool IN = in.view();
        MCE_DISPATCH_D(d, MCE_LAUNCH(k_cpdf1d_terms<DD>, dim3(ceil_div(Nt, 256)), dim3(256), 0, st, IN, Nt,
                                             marg, bn, gk[gcur].ptr, gv[gcur].ptr, xs[0], cpdf_direct.ptr,
                                             cpdf_cache.ptr));
        MCE_LAUNCH(k_sum_fixed, dim3(1), dim3(1024), 0, st, cpdf_direct.ptr, Nt, cpdf_sum.ptr);
        if (npts > 1) {
            MCE_LAUNCH(k_cpdf1d_grid, dim3(ceil_div(npts - 1, kCpdfPoints), nchunks), dim3(kCpdfPoints), 0, st,
                               cpdf_cache.ptr, Nt, cpdf_x.ptr + 1, npts - 1, cpdf_part.ptr);
            MCE_LAUNCH(k_cpdf1d_finish, dim3(ceil_div(npts - 1, 256)), dim3(256), 0, st, cpdf_part.ptr,
                               nchunks, npts - 1, norm, cpdf_y.ptr + 1);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(ys.data() + 1, cpdf_y.ptr + 1, (npts - 1) * sizeof(double), hipMemcpyDeviceToHost,
                                  st));
        }
        double direct_sum = 0.0;
        HIPCHK(hipMemcpyAsync(&direct_sum, cpdf_sum.ptr, sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        ys[0] = direct_sum * kInvTwoPi / norm;
    }
    xy.resize(2 * (size_t)npts);
    for (int i = 0; i < npts; i++) {
        xy[2 * i] = xs[i];
        xy[2 * i + 1] = ys[i];
    }
    return 0;
}

void Engine::draw_bar_nu() {
    if (!bar_nu.empty()) return;
    bar_nu.resize(d);
    for (int i = 0; i < d; i++) bar_nu[i] = 2 * (((double)rand() + 1.0) / ((double)RAND_MAX + 2.0));
}

int Engine::cpdf_2d(int i1, int i2, double lox, double hix, double resx, double loy, double hiy, double resy,
                    std::vector<double>& xyz, int& nx, int& ny) {
    nx = ny = 0;
    if (master_step == 0) throw std::runtime_error("the density needs at least one measurement update");
    if (!(0 <= i1 && i1 < i2 && i2 < d)) throw std::runtime_error("marginal indices need 0 <= i1 < i2 < n");
    if (!(hix > lox) || !(resx > 0) || !(hiy > loy) || !(resy > 0))
        throw std::runtime_error("the grid needs high > low and resolution > 0 on both axes");
    if (master_step == n_mu) {
        fprintf(stderr, "[cauchy_mi355x] the density cannot be evaluated after the window's last update: its g-tables "
                        "were not built\n");
        return 1;
    }
    if (((hix - lox) / resx + 1) * ((hiy - loy) / resy + 1) > 1e8) throw std::runtime_error("the grid would exceed 1e8 points");
    draw_bar_nu();
    bind();
    nx = (int)((hix - lox + resx - 1e-15) / resx) + 1;
    ny = (int)((hiy - loy + resy - 1e-15) / resy) + 1;
    const int npts = nx * ny;
    std::vector<double> xs(npts), ys(npts), zs(npts, 0.0);
    for (int i = 0; i < ny; i++) {
        const double y = std::min(loy + i * resy, hiy);
        for (int j = 0; j < nx; j++) {
            xs[i * nx + j] = std::min(lox + j * resx, hix);
            ys[i * nx + j] = y;
        }
    }
    const double norm = fz.re;
    const double R = kInvTwoPi;
    if (Nt > 0) {
        GraveyardScope gs(&graveyard);
        const int nchunks = ceil_div(Nt, kCpdfChunk);
        const size_t cells = (size_t)Nt * MM;
        c2_mc.reserve(Nt);
        c2_b2.reserve(2 * (size_t)Nt);
        c2_sn.reserve(cells + Nt);
        c2_cs.reserve(cells + Nt);
        c2_g1r.reserve(cells);
        c2_g2r.reserve(cells);
        c2_g.reserve(cells);
        c2_fc.reserve(cells);
        c2_err.reserve(1);
        c2_px.reserve(npts);
        c2_py.reserve(npts);
        cpdf_direct.reserve(Nt);
        cpdf_y.reserve(npts);
        cpdf_part.reserve((size_t)nchunks * npts);
        cpdf_sum.reserve(1);
        HIPCHK(hipMemcpyAsync(c2_px.ptr, xs.data(), npts * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(c2_py.ptr, ys.data(), npts * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemsetAsync(c2_err.ptr, 0, sizeof(int), st));
        Cpdf2D C;
        C.MM = MM;
        C.mc = c2_mc.ptr;
        C.b2 = c2_b2.ptr;
        C.sn = c2_sn.ptr;
        C.cs = c2_cs.ptr;
        C.g1r = c2_g1r.ptr;
        C.g2r = c2_g2r.ptr;
        C.gc = c2_g.ptr;
        C.fc = c2_fc.ptr;
        VecArg bn;
        for (int i = 0; i < d; i++) bn.v[i] = bar_nu[i];
        TermPool IN = in.view();
        MCE_DISPATCH_D(d, MCE_LAUNCH(k_cpdf2d_terms<DD>, dim3(ceil_div(Nt, 128)), dim3(128), 0, st, IN, Nt,
                                             i1, i2, bn, gk[gcur].ptr, gv[gcur].ptr, xs[0], ys[0], C,
                                             cpdf_direct.ptr, c2_err.ptr));
        MCE_LAUNCH(k_sum_fixed, dim3(1), dim3(1024), 0, st, cpdf_direct.ptr, Nt, cpdf_sum.ptr);
        if (npts > 1) {
            MCE_LAUNCH(k_cpdf2d_grid, dim3(ceil_div(npts - 1, kCpdfPoints), nchunks), dim3(kCpdfPoints), 0, st,
                               C, Nt, (const int*)nullptr, c2_px.ptr + 1, c2_py.ptr + 1, npts - 1, cpdf_part.ptr,
                               c2_err.ptr);
            MCE_LAUNCH(k_cpdf1d_finish, dim3(ceil_div(npts - 1, 256)), dim3(256), 0, st, cpdf_part.ptr,
                               nchunks, npts - 1, 1.0, cpdf_y.ptr + 1);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(zs.data() + 1, cpdf_y.ptr + 1, (npts - 1) * sizeof(double), hipMemcpyDeviceToHost,
                                  st));
        }
        double direct_sum = 0.0;
        int singular = 0;
        HIPCHK(hipMemcpyAsync(&direct_sum, cpdf_sum.ptr, sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(&singular, c2_err.ptr, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (singular)
            throw std::runtime_error("a grid point meets a term's singularity (gamma1 = gamma2 = 0); move the grid");
        zs[0] = direct_sum;
        // fx = 2 * sum * (1 / 2pi)^2 / Re fz, cpdf_ndim.hpp:1451
        for (int k = 0; k < npts; k++) zs[k] = 2 * zs[k] * R * R / norm;
    }
    xyz.resize(3 * (size_t)npts);
    for (int k = 0; k < npts; k++) {
        xyz[3 * k] = xs[k];
        xyz[3 * k + 1] = ys[k];
        xyz[3 * k + 2] = zs[k];
    }
    return 0;
}

// The relative system sec - prim, transformed by Trel (2 x d): its moments and its density
// on a grid (pycauchy_single_step_eval_2d_rsys_cpdf, pycauchy.hpp:673-750, with
// get_marg2d_relative_and_transformed_cpdf and grid_eval_marg2d_relative_and_transformed_cpdf,
// cauchy_prediction.hpp:653-972, 1102-1200).  The grid limits are relative to the relative
// mean.  Points are ordered with y fastest (point i * ny + j is (x_i, y_j)), as the reference's
// grid is.  Returns the number of relative terms kept by the approximation.
static int64_t rsys_2d(Engine& P, Engine& S, const double* Trel, double eps, double xlow, double xhigh, double dx,
                       double ylow, double yhigh, double dy, cplx& nf, cplx mean[2], cplx cov[4],
                       std::vector<double>& xyz, int& nx, int& ny) {
    const int d = P.d;
    if (S.d != d) throw std::runtime_error("the two estimators have different state dimensions");
    if (P.master_step == 0 || S.master_step == 0) throw std::runtime_error("both estimators need a measurement update");
    if (P.master_step == P.n_mu || S.master_step == S.n_mu)
        throw std::runtime_error("an estimator is past its window's last update, whose g-tables are not built");
    if (!(xhigh > xlow) || !(yhigh > ylow) || !(dx > 0) || !(dy > 0))
        throw std::runtime_error("the grid needs high > low and resolution > 0 on both axes");
    // the direction used on rows that vanish under Trel (cauchy_prediction.hpp:703-713)
    RsysArgs R;
    memset(&R, 0, sizeof(R));
    for (int i = 0; i < 2 * d; i++) R.T[i] = Trel[i];
    for (int i = 0; i < d; i++) R.nu[i] = 2 * (((double)rand() + 1.0) / ((double)RAND_MAX + 2.0)) - 1;
    for (int k = 0; k < 2; k++) {
        double acc = 0.0;
        for (int j = 0; j < d; j++) acc += Trel[k * d + j] * R.nu[j];
        R.nu2[k] = acc;
    }
    P.bind();  // after the draws
    S.bind();
    R.norm_fact = kInvTwoPi * kInvTwoPi / P.fz.re / S.fz.re;
    R.eps = eps;
    const int np = P.Nt, ns = S.Nt;
    const int64_t npairs = (int64_t)np * ns;
    const int CM = P.MM + S.MM;
    if (CM > kMaxM) throw std::runtime_error("the relative arrangement would exceed 31 hyperplanes");
    const double bytes = (double)npairs * (CM * 72.0 + 8.0 * (CM + 1) * 2 + 14 * 8 + 32);
    if (bytes > 64.0 * (1ull << 30))
        throw std::runtime_error("the relative system's pairs need more than 64 GiB; use estimators with fewer terms");
    GraveyardScope gs(&P.graveyard);
    hipStream_t st = P.st;
    HIPCHK(hipStreamSynchronize(S.st));
    P.c2_mc.reserve(npairs);
    P.c2_b2.reserve(2 * npairs);
    P.c2_sn.reserve(npairs * (CM + 1));
    P.c2_cs.reserve(npairs * (CM + 1));
    P.c2_g1r.reserve(npairs * CM);
    P.c2_g2r.reserve(npairs * CM);
    P.c2_g.reserve(npairs * CM);
    P.c2_fc.reserve(npairs * CM);
    P.rs_mom.reserve(npairs * 14);
    P.rs_keep.reserve(npairs);
    P.rs_idx.reserve(npairs);
    P.rs_cnt.reserve(1);
    P.rs_sums.reserve(14);
    P.rs_part.reserve(14 * kColSlices);
    P.c2_err.reserve(1);
    Cpdf2D C;
    C.MM = CM;
    C.mc = P.c2_mc.ptr;
    C.b2 = P.c2_b2.ptr;
    C.sn = P.c2_sn.ptr;
    C.cs = P.c2_cs.ptr;
    C.g1r = P.c2_g1r.ptr;
    C.g2r = P.c2_g2r.ptr;
    C.gc = P.c2_g.ptr;
    C.fc = P.c2_fc.ptr;
    TermPool PI = P.in.view(), SI = S.in.view();
    if (npairs > 0) {
        const unsigned blocks = (unsigned)((npairs + 127) / 128);
        // per-lane arrays sized by the pairs' largest row count: the smaller the bound,
        // the smaller each lane's private memory
        const int KM = P.largest_shape() + S.largest_shape();
        HIPCHK(hipMemsetAsync(P.c2_err.ptr, 0, sizeof(int), st));
#define MCE_RSYS(K_)                                                                                             \
    MCE_DISPATCH_D(d, MCE_LAUNCH((k_rsys_pairs<DD, K_>), dim3(blocks), dim3(128), 0, st, PI, np,         \
                                         P.gk[P.gcur].ptr, P.gv[P.gcur].ptr, SI, ns, S.gk[S.gcur].ptr,           \
                                         S.gv[S.gcur].ptr, R, C, P.rs_mom.ptr, P.rs_keep.ptr, P.c2_err.ptr))
        if (KM <= 16) {
            MCE_RSYS(16);
        } else if (KM <= 24) {
            MCE_RSYS(24);
        } else {
            MCE_RSYS(kMaxM);
        }
#undef MCE_RSYS
        MCE_LAUNCH(k_colsum_part, dim3(kColSlices, 14), dim3(256), 0, st, P.rs_mom.ptr, npairs,
                           P.rs_part.ptr);
        MCE_LAUNCH(k_colsum_finish, dim3(1), dim3(64), 0, st, P.rs_part.ptr, 14, P.rs_sums.ptr);
        {
            const uint8_t* keep = P.rs_keep.ptr;
            int* idx = P.rs_idx.ptr;
            int* cnt = P.rs_cnt.ptr;
            const int np2 = (int)npairs;
            P.cub("select", [=](void* t, size_t& b, hipStream_t s_) {
                return hipcub::DeviceSelect::Flagged(t, b, hipcub::CountingInputIterator<int>(0), keep, idx, cnt, np2,
                                                     s_);
            });
        }
        HIPCHK(hipGetLastError());
    }
    double sums[14] = {0};
    int nkeep = 0, pair_err = 0;
    if (npairs > 0) {
        HIPCHK(hipMemcpyAsync(sums, P.rs_sums.ptr, sizeof(sums), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(&nkeep, P.rs_cnt.ptr, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(&pair_err, P.c2_err.ptr, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    if (pair_err) throw std::runtime_error("a relative term has more rows than the pair kernel was launched for");
    // finalize_rel_sys_moments (cauchy_prediction.hpp:288-304)
    nf = cmake(sums[0], sums[1]);
    const cplx ifz = cmake(0.0, nf.re);
    for (int i = 0; i < 2; i++) mean[i] = cdiv(cmake(sums[2 + 2 * i], sums[3 + 2 * i]), ifz);
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
            const cplx c = cmake(-sums[6 + 2 * (2 * i + j)], -sums[7 + 2 * (2 * i + j)]);
            cov[2 * i + j] = csub(cdiv(c, nf), cmul(mean[i], mean[j]));
        }
    // the grid about the relative mean (grid_eval_marg2d_relative_and_transformed_cpdf)
    const double x0 = xlow + mean[0].re, x1 = xhigh + mean[0].re;
    const double y0 = ylow + mean[1].re, y1 = yhigh + mean[1].re;
    if (((x1 - x0) / dx + 2) * ((y1 - y0) / dy + 2) > 1e8) throw std::runtime_error("the grid would exceed 1e8 points");
    nx = (int)ceil((x1 - x0) / dx + 1);
    ny = (int)ceil((y1 - y0) / dy + 1);
    std::vector<double> xs(nx), ys(ny);
    for (int i = 0; i < nx; i++) xs[i] = x0 + dx * i;
    xs[nx - 1] = x1;
    for (int j = 0; j < ny; j++) ys[j] = y0 + dy * j;
    ys[ny - 1] = y1;
    const int npts = nx * ny;
    std::vector<double> px(npts), py(npts), z(npts, 0.0);
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            px[i * ny + j] = xs[i];
            py[i * ny + j] = ys[j];
        }
    if (nkeep > 0) {
        const int nchunks = ceil_div(nkeep, kCpdfChunk);
        P.c2_px.reserve(npts);
        P.c2_py.reserve(npts);
        P.cpdf_part.reserve((size_t)nchunks * npts);
        P.cpdf_y.reserve(npts);
        HIPCHK(hipMemcpyAsync(P.c2_px.ptr, px.data(), npts * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(P.c2_py.ptr, py.data(), npts * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemsetAsync(P.c2_err.ptr, 0, sizeof(int), st));
        MCE_LAUNCH(k_cpdf2d_grid, dim3(ceil_div(npts, kCpdfPoints), nchunks), dim3(kCpdfPoints), 0, st, C,
                           nkeep, P.rs_idx.ptr, P.c2_px.ptr, P.c2_py.ptr, npts, P.cpdf_part.ptr, P.c2_err.ptr);
        MCE_LAUNCH(k_cpdf1d_finish, dim3(ceil_div(npts, 256)), dim3(256), 0, st, P.cpdf_part.ptr, nchunks, npts,
                           1.0, P.cpdf_y.ptr);
        HIPCHK(hipGetLastError());
        int singular = 0;
        HIPCHK(hipMemcpyAsync(z.data(), P.cpdf_y.ptr, npts * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(&singular, P.c2_err.ptr, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (singular)
            throw std::runtime_error("a grid point meets a relative term's singularity (gamma1 = gamma2 = 0)");
        // 2 * sum * (1 / 2pi)^2 / (Re fz_p Re fz_s), cauchy_prediction.hpp:1072
        for (int k = 0; k < npts; k++) z[k] = 2 * z[k] * kInvTwoPi * kInvTwoPi / (P.fz.re * S.fz.re);
    }
    xyz.resize(3 * (size_t)npts);
    for (int k = 0; k < npts; k++) {
        xyz[3 * k] = px[k];
        xyz[3 * k + 1] = py[k];
        xyz[3 * k + 2] = z[k];
    }
    return nkeep;
}

Co<void> Engine::finalize_extended_co(double* x_bar) {
    if (ntan) throw std::runtime_error("the extended-mode finish carries no tangents");
    bind();
    if (!skip) {
        VecArg dx;
        for (int i = 0; i < d; i++) dx.v[i] = mean[i].re;
        TermPool IN = in.view();
        MCE_DISPATCH_D(d, MCE_RUN(k_sub_b<DD>, dim3(ceil_div(Nt, 256)), dim3(256), 0, st, IN, Nt, dx));
        HIPCHK(hipGetLastError());
    }
    for (int i = 0; i < d; i++) mean[i].re += x_bar[i];
    for (int i = 0; i < d; i++) x_bar[i] = mean[i].re;
    if (!skip) co_await sync_point();
}

// ======================================================================== C ABI
namespace {

struct Handle {
    CauchyDynamicsUpdateContainer duc;
    Engine* eng = nullptr;
    void (*cb)(CauchyDynamicsUpdateContainer*) = nullptr;
    // extended (nonlinear) mode: h(x) into zbar, then the measurement Jacobian
    void (*msmt_model)(CauchyDynamicsUpdateContainer*, double*) = nullptr;
    void (*ext_update)(CauchyDynamicsUpdateContainer*) = nullptr;
    std::vector<double> Phi, Gamma, B, H, beta, gamma, x, zbar;
    // per scalar update of the last step: fz before term reduction, and the term counts after
    // MU/MUC and after reduction
    std::vector<double> step_fz;
    std::vector<int> step_cnt;
    void record(int i, const Engine* e) {
        if ((int)step_fz.size() <= i) {
            step_fz.resize(i + 1, 0.0);
            step_cnt.resize(2 * (i + 1), 0);
        }
        step_fz[i] = e->fz_raw.re;
        step_cnt[2 * i] = e->diag_cnt[1];
        step_cnt[2 * i + 1] = e->diag_cnt[2];
    }
};

void set_err(const char* where, const std::exception& e) {
    g_last_error = std::string(where) + ": " + e.what();
    fprintf(stderr, "[cauchy_mi355x] %s\n", g_last_error.c_str());
}

// ---------------------------------------------------------------- no-allocation working memory
// The working memory of mce_cpdf1d_noalloc, for callers that cannot allocate device memory:
// one array per GPU, in the library's code object, held until the process ends.  Every
// call on a GPU shares its array, so calls that overlap on one GPU race.
constexpr size_t kNoAllocBytes = size_t(128) << 20;
__device__ __align__(256) char g_noalloc_scratch[kNoAllocBytes];

// Device address of this GPU's array, looked up once per device on its first use (the
// array's name is a host-side stand-in, never a device address).
char* noalloc_scratch() {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    static std::mutex mu;
    static char* ptr[64] = {};
    if (dev < 0 || dev >= 64) throw std::runtime_error("device index beyond 63");
    std::lock_guard<std::mutex> lk(mu);
    if (!ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(g_noalloc_scratch));
        if (e != hipSuccess || !p)
            throw std::runtime_error(std::string("hipGetSymbolAddress of the no-allocation array: ") +
                                     hipGetErrorString(e));
        ptr[dev] = (char*)p;
    }
    return ptr[dev];
}

// Helper threads for the batch calls.  They live as long as the process (each keeps its
// streams), so a batch call starts no thread; one batch call runs at a time.
class WorkerPool {
   public:
    // Runs work() on the calling thread and on `helpers` pool threads; returns when all end.
    void run(int helpers, const std::function<void()>& work) {
        std::lock_guard<std::mutex> call(call_mu_);
        {
            std::unique_lock<std::mutex> lk(mu_);
            while ((int)threads_.size() < helpers) {
                const int id = (int)threads_.size();
                threads_.emplace_back([this, id] { loop(id); });
                threads_.back().detach();
            }
            job_ = &work;
            active_ = helpers;
            pending_ = helpers;
            gen_++;
        }
        cv_.notify_all();
        work();
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [this] { return pending_ == 0; });
        job_ = nullptr;
    }

   private:
    void loop(int id) {
        unsigned long long seen = 0;
        for (;;) {
            const std::function<void()>* job;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
                if (id >= active_) continue;
                job = job_;
            }
            (*job)();
            std::lock_guard<std::mutex> lk(mu_);
            if (--pending_ == 0) done_.notify_all();
        }
    }
    std::mutex call_mu_, mu_;
    std::condition_variable cv_, done_;
    std::vector<std::thread> threads_;
    const std::function<void()>* job_ = nullptr;
    int active_ = 0, pending_ = 0;
    unsigned long long gen_ = 0;
};

WorkerPool& worker_pool() {
    static WorkerPool* pool = new WorkerPool();  // never destroyed: its threads outlive main
    return *pool;
}

// Runs step(i) for i in [0, count) on up to num_threads host threads (0: one per hardware
// thread), estimators with the most terms first.  Each estimator is stepped by one thread
// at a time, on its own streams.  Returns the number of estimators whose step reported an
// error; the calling thread's last error is the first of them.
template <class F>
int run_concurrently(void** handles, int count, int num_threads, int (*terms_of)(void*), F step) {
    if (count <= 0) return 0;
    std::vector<int> order(count);
    for (int i = 0; i < count; i++) order[i] = i;
    std::vector<int> nt(count);
    for (int i = 0; i < count; i++) nt[i] = terms_of(handles[i]);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return nt[a] > nt[b]; });
    int nthreads = num_threads > 0 ? num_threads : (int)std::max(1u, std::thread::hardware_concurrency());
    nthreads = std::min(nthreads, count);
    std::atomic<int> next{0};
    std::vector<std::string> errors(count);
    auto work = [&]() {
        int k;
        while ((k = next.fetch_add(1)) < count) {
            const int i = order[k];
            g_last_error.clear();
            step(i);
            errors[i] = g_last_error;
        }
    };
    if (nthreads > 1) {
        const std::function<void()> job = work;
        worker_pool().run(nthreads - 1, job);
    } else {
        work();
    }
    int failed = 0;
    std::string first;
    for (int i = 0; i < count; i++)
        if (!errors[i].empty()) {
            if (!failed) first = "estimator " + std::to_string(i) + ": " + errors[i];
            failed++;
        }
    g_last_error = first;
    return failed;
}

template <class T>
T* mallocd(size_t n) {
    T* p = (T*)malloc(std::max<size_t>(n, 1) * sizeof(T));
    if (!p) {
        fprintf(stderr, "Pointer allocated by malloc has returned with NULL, indicating FAILURE!\n");
        exit(1);
    }
    return p;
}

// Records a failed update: into *err for a batched estimator (its first failure), else as
// the calling thread's last error.
void note_failure(std::exception_ptr fail, const char* where, std::string* err) {
    try {
        std::rethrow_exception(fail);
    } catch (const std::exception& ex) {
        if (!err) {
            set_err(where, ex);
        } else if (err->empty()) {
            *err = std::string(where) + ": " + ex.what();
            fprintf(stderr, "[cauchy_mi355x] %s\n", err->c_str());
        }
    }
}

Co<void> step_ltiv_task(void* handle, double* msmts, int size_msmts, double* controls, int size_controls,
                               double** out_Phi, int* size_out_Phi, double** out_Gamma, int* size_out_Gamma,
                               double** out_B, int* size_out_B, double** out_H, int* size_out_H, double** out_beta,
                               int* size_out_beta, double** out_gamma, int* size_out_gamma, double** out_fz,
                               int* size_out_fz, double** out_xhat, int* size_out_xhat, double** out_Phat,
                               int* size_out_Phat, double** out_cerr_fz, int* size_out_cerr_fz,
                               double** out_cerr_xhat, int* size_out_cerr_xhat, double** out_cerr_Phat,
                               int* size_out_cerr_Phat, int** out_err_code, int* size_out_err_code,
                        std::string* err) {
    (void)size_controls;
    Handle* h = (Handle*)handle;
    Engine* e = h->eng;
    CauchyDynamicsUpdateContainer& duc = h->duc;
    duc.u = controls;
    const int p = duc.p, n = duc.n, cmcc = duc.cmcc, pncc = duc.pncc;
    const int nm = size_msmts;
    const bool is_ltv = h->cb != nullptr;
    // dynamics outputs (pycauchy.hpp:379-419)
    if (!is_ltv) {
        *size_out_Phi = *size_out_Gamma = *size_out_B = *size_out_H = *size_out_beta = *size_out_gamma = 0;
        *out_Phi = mallocd<double>(0);
        *out_Gamma = mallocd<double>(0);
        *out_B = mallocd<double>(0);
        *out_H = mallocd<double>(0);
        *out_beta = mallocd<double>(0);
        *out_gamma = mallocd<double>(0);
    } else {
        *size_out_Phi = n * n;
        *out_Phi = mallocd<double>(n * n);
        *size_out_Gamma = n * pncc;
        *out_Gamma = mallocd<double>(n * pncc);
        *size_out_B = n * cmcc;
        *out_B = mallocd<double>(n * cmcc);
        *size_out_H = nm * n;
        *out_H = mallocd<double>(nm * n);
        *size_out_beta = pncc;
        *out_beta = mallocd<double>(pncc);
        *size_out_gamma = nm;
        *out_gamma = mallocd<double>(nm);
    }
    *size_out_fz = nm;
    *out_fz = mallocd<double>(nm);
    *size_out_xhat = nm * n;
    *out_xhat = mallocd<double>(nm * n);
    *size_out_Phat = nm * n * n;
    *out_Phat = mallocd<double>(nm * n * n);
    *size_out_cerr_fz = nm;
    *out_cerr_fz = mallocd<double>(nm);
    *size_out_cerr_xhat = nm;
    *out_cerr_xhat = mallocd<double>(nm);
    *size_out_cerr_Phat = nm;
    *out_cerr_Phat = mallocd<double>(nm);
    *size_out_err_code = nm;
    *out_err_code = mallocd<int>(nm);
    const int start = p - nm;
    if (is_ltv) {
        for (int i = 0; i < n; i++) duc.x[i] = e->mean[i].re;
        h->cb(&duc);
        memcpy(*out_Phi, duc.Phi, n * n * sizeof(double));
        memcpy(*out_Gamma, duc.Gamma, n * pncc * sizeof(double));
        memcpy(*out_B, duc.B, n * cmcc * sizeof(double));
        memcpy(*out_H, duc.H + start * n, nm * n * sizeof(double));
        memcpy(*out_beta, duc.beta, pncc * sizeof(double));
        memcpy(*out_gamma, duc.gamma + start, nm * sizeof(double));
    }
    h->step_fz.clear();
    h->step_cnt.clear();
    for (int i = 0; i < nm; i++) {
        const int idx = start + i;
        std::exception_ptr fail;
        try {
            co_await e->update_co(msmts[i], duc.Phi, duc.Gamma, duc.beta, duc.H + idx * n, duc.gamma[idx], duc.B,
                                  duc.u);
        } catch (...) {
            fail = std::current_exception();
        }
        if (fail) note_failure(fail, "pycauchy_single_step_ltiv", err);
        h->record(i, e);
        (*out_fz)[i] = e->fz.re;
        for (int j = 0; j < n; j++) (*out_xhat)[i * n + j] = e->mean[j].re;
        for (int j = 0; j < n * n; j++) (*out_Phat)[i * n * n + j] = e->cov[j].re;
        (*out_cerr_fz)[i] = e->fz.im;
        double mx = fabs(e->mean[0].im);
        for (int j = 1; j < n; j++) mx = std::max(mx, fabs(e->mean[j].im));
        (*out_cerr_xhat)[i] = mx;
        double mp = fabs(e->cov[0].im);
        for (int j = 1; j < n * n; j++) mp = std::max(mp, fabs(e->cov[j].im));
        (*out_cerr_Phat)[i] = mp;
        (*out_err_code)[i] = e->errs;
    }
}


// ---------------------------------------------------------------- batch scheduler
// Steps many engines on one host thread and one stream, launching each kernel of their
// updates once for all of them.  Every engine's step runs as a task; a task runs until it
// needs device data (a count, a total, the moment sums), handing over its launches, uploads
// and library calls on the way.  Once every task waits or has ended, flush() issues the
// handed-over work: requests of one engine stay in their order, and the heads of all
// engines' request chains that name the same kernel become one merged launch (k_many), so
// an update's ~70 launches are made once per group of engines instead of once per engine.
// The reads of all waiting engines are then gathered into one buffer, copied to the host
// with one copy and delivered after one synchronisation, and the tasks resume.
// Profile of the merged g-table launches (mce_batch_profile_*): HIP events around every
// merged k_group_eval launch on the scheduler's stream, and its algorithmic bytes counted on
// the device by k_geval_bytes for each estimator's part, summed over all schedulers.
struct BatchProfile {
    std::mutex mu;
    std::atomic<bool> on{false};
    double launches = 0, bytes = 0, ms = 0, flushes = 0, merged = 0, library = 0;
};
BatchProfile& batch_profile() {
    static BatchProfile* p = new BatchProfile();
    return *p;
}

class BatchRunner {
   public:
    // MCE_BATCH_STATS=1: each run prints its flushes, launches and where its host time went.
    struct Stats {
        long flushes = 0, launches = 0, host_steps = 0, resumes = 0, reads = 0;
        double t_resume = 0, t_pack = 0, t_wait = 0;
    };
    Stats stats;
    static bool stats_on() {
        static const bool on = [] {
            const char* v = getenv("MCE_BATCH_STATS");
            return v && v[0] && v[0] != '0';
        }();
        return on;
    }
    struct Job {
        Engine* e = nullptr;
        std::optional<Co<void>> task;
        BatchSlot slot;
        std::string err;
        bool prof_on = false;
    };

    void run(std::vector<Job>& jobs) {
        t_streams.init();
        hipStream_t s = t_streams.st;
        profiling_ = batch_profile().on.load();
        if (profiling_) {
            prof_acc_.reserve(1);
            HIPCHK(hipMemsetAsync(prof_acc_.ptr, 0, sizeof(unsigned long long), s));
            p_launches_ = p_ms_ = p_flushes_ = p_merged_ = p_library_ = 0;
        }
        for (Job& jb : jobs) {
            jb.e->bs = &jb.slot;
            jb.prof_on = jb.e->prof.on;
            jb.e->prof.on = false;
            jb.slot.geval_bytes = profiling_ ? prof_acc_.ptr : nullptr;
        }
        try {
            drive(jobs, s);
        } catch (...) {
            (void)hipStreamSynchronize(s);
            for (Job& jb : jobs) {
                jb.e->bs = nullptr;
                jb.e->prof.on = jb.prof_on;
                jb.slot.clear_chain();
                jb.slot.reads.clear();
            }
            throw;
        }
        for (Job& jb : jobs) {
            jb.e->bs = nullptr;
            jb.e->prof.on = jb.prof_on;
            jb.e->graveyard.release(jb.e->graveyard.big_bytes > Engine::kGraveyardBytes);
        }
        if (profiling_) {
            unsigned long long b = 0;
            HIPCHK(hipMemcpy(&b, prof_acc_.ptr, sizeof(b), hipMemcpyDeviceToHost));
            BatchProfile& bp = batch_profile();
            std::lock_guard<std::mutex> lk(bp.mu);
            bp.launches += p_launches_;
            bp.ms += p_ms_;
            bp.bytes += (double)b;
            bp.flushes += p_flushes_;
            bp.merged += p_merged_;
            bp.library += p_library_;
        }
        if (stats_on()) {
            fprintf(stderr,
                    "[batch] %zu engines: %ld flushes, %ld merged launches, %ld library calls, %ld resumes, %ld reads; "
                    "host ms: resume %.2f pack %.2f wait %.2f\n",
                    jobs.size(), stats.flushes, stats.launches, stats.host_steps, stats.resumes, stats.reads,
                    stats.t_resume * 1e3, stats.t_pack * 1e3, stats.t_wait * 1e3);
            stats = Stats{};
        }
    }

   private:
    void drive(std::vector<Job>& jobs, hipStream_t s) {
        for (Job& jb : jobs) resume(jb, {});
        for (;;) {
            bool busy = false;
            for (Job& jb : jobs)
                if (jb.slot.waiting || !jb.slot.chain.empty()) {
                    busy = true;
                    break;
                }
            if (!busy) break;
            flush(jobs, s);
            for (Job& jb : jobs)
                if (jb.slot.waiting) {
                    const std::coroutine_handle<> h = jb.slot.waiting;
                    jb.slot.waiting = {};
                    resume(jb, h);
                }
        }
        HIPCHK(hipStreamSynchronize(s));
    }

    static double now() {
        return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    }
    void resume(Job& jb, std::coroutine_handle<> h) {
        GraveyardScope gs(&jb.e->graveyard);
        const double t0 = stats_on() ? now() : 0;
        if (h)
            h.resume();
        else
            jb.task->start();
        if (stats_on()) {
            stats.t_resume += now() - t0;
            stats.resumes++;
        }
        if (jb.task->done()) {
            try {
                jb.task->result();
            } catch (const std::exception& ex) {
                if (jb.err.empty()) jb.err = ex.what();
            }
        }
    }

    struct Item {
        int j;
        int r;
    };
    struct Step {
        BatchReq::Kind kind;
        const void* fn;
        uint32_t threads;
        std::vector<Item> items;
    };

    static size_t up(size_t v, size_t a) { return (v + a - 1) / a * a; }

    // Grows a pinned host buffer (the stream is idle whenever this runs).
    static void grow_pinned(uint8_t*& p, size_t& cap, size_t need) {
        if (need <= cap) return;
        if (p) HIPCHK(hipHostFree(p));
        cap = std::max(need, 2 * cap);
        HIPCHK(hipHostMalloc((void**)&p, cap, hipHostMallocDefault));
    }
    // Grows a device buffer; the old block goes straight back (the stream is idle).
    static void grow_dev(DBuf<uint8_t>& b, size_t need) {
        if (need <= b.cap) return;
        const size_t n = std::max(need, 2 * b.cap);
        b.release();
        b.reserve(n);
    }

    void plan_steps(std::vector<Job>& jobs) {
        steps_.clear();
        const int J = (int)jobs.size();
        head_.assign(J, 0);
        size_t left = 0;
        for (Job& jb : jobs) left += jb.slot.chain.size();
        while (left) {
            // library calls at a chain's head run first, one by one, in job order
            bool host = false;
            for (int j = 0; j < J; j++) {
                auto& ch = jobs[j].slot.chain;
                while (head_[j] < ch.size() && ch[head_[j]].kind == BatchReq::kHost) {
                    steps_.push_back(Step{BatchReq::kHost, nullptr, 0, {Item{j, (int)head_[j]}}});
                    head_[j]++;
                    left--;
                    host = true;
                }
            }
            if (host) continue;
            // then the kernel (or the uploads) at the most chain heads, for all those chains
            keys_.clear();
            for (int j = 0; j < J; j++) {
                auto& ch = jobs[j].slot.chain;
                if (head_[j] >= ch.size()) continue;
                const BatchReq& r = ch[head_[j]];
                const void* fn = r.kind == BatchReq::kUpload ? nullptr : r.fn;
                bool found = false;
                for (auto& k : keys_)
                    if (k.fn == fn && k.threads == r.threads && k.kind == r.kind) {
                        k.count++;
                        found = true;
                        break;
                    }
                if (!found) keys_.push_back(Key{r.kind, fn, r.threads, 1});
            }
            const Key* best = &keys_[0];
            for (const Key& k : keys_)
                if (k.count > best->count) best = &k;
            Step st{best->kind, best->fn, best->threads, {}};
            for (int j = 0; j < J; j++) {
                auto& ch = jobs[j].slot.chain;
                if (head_[j] >= ch.size()) continue;
                const BatchReq& r = ch[head_[j]];
                const void* fn = r.kind == BatchReq::kUpload ? nullptr : r.fn;
                if (r.kind == best->kind && fn == best->fn && r.threads == best->threads) {
                    st.items.push_back(Item{j, (int)head_[j]});
                    head_[j]++;
                    left--;
                }
            }
            steps_.push_back(std::move(st));
        }
    }

    void flush(std::vector<Job>& jobs, hipStream_t s) {
        const double t0 = stats_on() ? now() : 0;
        plan_steps(jobs);
        using Copy = Kern<&k_copy_bytes>;
        using CopyEntry = Copy::Entry;
        auto copy_blocks = [](size_t bytes) {
            return (unsigned)std::min<size_t>(64, std::max<size_t>(1, (bytes / 8 + 255) / 256));
        };
        // layout: upload payloads, then per step its first-block array and entries, then the
        // read gather's
        size_t off = 0;
        pay_off_.clear();
        for (const Step& st : steps_)
            if (st.kind == BatchReq::kUpload)
                for (const Item& it : st.items) {
                    off = up(off, 16);
                    pay_off_.push_back(off);
                    off += jobs[it.j].slot.chain[it.r].nbytes;
                }
        step_off_.clear();
        for (const Step& st : steps_) {
            if (st.kind == BatchReq::kHost) {
                step_off_.push_back({0, 0});
                continue;
            }
            const size_t n = st.items.size();
            size_t esize, ealign;
            if (st.kind == BatchReq::kUpload) {
                esize = sizeof(CopyEntry);
                ealign = alignof(CopyEntry);
            } else {
                const BatchReq& r0 = jobs[st.items[0].j].slot.chain[st.items[0].r];
                esize = r0.esize;
                ealign = r0.ealign;
            }
            const size_t f = up(off, 16);
            const size_t e = up(f + 4 * n, std::max<size_t>(16, ealign));
            step_off_.push_back({f, e});
            off = e + n * esize;
        }
        size_t nreads = 0, rbytes = 0;
        for (Job& jb : jobs)
            for (const BatchRead& rd : jb.slot.reads) {
                nreads++;
                rbytes = up(rbytes, 16) + rd.bytes;
            }
        const size_t rf = up(off, 16);
        const size_t re = up(rf + 4 * nreads, std::max<size_t>(16, alignof(CopyEntry)));
        if (nreads) off = re + nreads * sizeof(CopyEntry);
        const size_t total = off;
        grow_pinned(stage_, stage_cap_, total);
        grow_dev(arena_, total);
        if (nreads) {
            grow_dev(rdev_, rbytes);
            grow_pinned(rhost_, rhost_cap_, rbytes);
        }
        uint8_t* dev = arena_.ptr;
        // fill the stage
        size_t pi = 0;
        for (size_t si = 0; si < steps_.size(); si++) {
            const Step& st = steps_[si];
            if (st.kind == BatchReq::kHost) continue;
            unsigned* first = (unsigned*)(stage_ + step_off_[si].first);
            uint8_t* ents = stage_ + step_off_[si].second;
            unsigned blocks = 0;
            for (size_t k = 0; k < st.items.size(); k++) {
                const Item& it = st.items[k];
                const BatchReq& r = jobs[it.j].slot.chain[it.r];
                first[k] = blocks;
                if (st.kind == BatchReq::kUpload) {
                    const size_t po = pay_off_[pi++];
                    memcpy(stage_ + po, jobs[it.j].slot.bytes.data() + r.off, r.nbytes);
                    const unsigned gx = copy_blocks(r.nbytes);
                    const CopyEntry ce = Copy::entry(gx, 1u, (const uint8_t*)(dev + po), (uint8_t*)r.dst,
                                                     (int64_t)r.nbytes);
                    memcpy(ents + k * sizeof(CopyEntry), &ce, sizeof(CopyEntry));
                    blocks += gx;
                } else {
                    memcpy(ents + k * r.esize, jobs[it.j].slot.bytes.data() + r.off, r.esize);
                    blocks += r.gx * r.gy;
                }
            }
            step_blocks_.resize(steps_.size());
            step_blocks_[si] = blocks;
        }
        unsigned rblocks = 0;
        if (nreads) {
            unsigned* first = (unsigned*)(stage_ + rf);
            uint8_t* ents = stage_ + re;
            size_t ro = 0, k = 0;
            for (Job& jb : jobs)
                for (const BatchRead& rd : jb.slot.reads) {
                    ro = up(ro, 16);
                    first[k] = rblocks;
                    const unsigned gx = copy_blocks(rd.bytes);
                    const CopyEntry ce =
                        Copy::entry(gx, 1u, (const uint8_t*)rd.src, (uint8_t*)(rdev_.ptr + ro), (int64_t)rd.bytes);
                    memcpy(ents + k * sizeof(CopyEntry), &ce, sizeof(CopyEntry));
                    rblocks += gx;
                    ro += rd.bytes;
                    k++;
                }
        }
        // issue: the staged bytes, the steps in order, the read gather and its copy
        if (total) HIPCHK(hipMemcpyAsync(dev, stage_, total, hipMemcpyHostToDevice, s));
        for (size_t si = 0; si < steps_.size(); si++) {
            const Step& st = steps_[si];
            if (st.kind == BatchReq::kHost) {
                const Item& it = st.items[0];
                jobs[it.j].slot.chain[it.r].host(s);
                continue;
            }
            size_t lds = 0;
            for (const Item& it : st.items) lds = std::max(lds, jobs[it.j].slot.chain[it.r].lds);
            const void* fn = st.kind == BatchReq::kUpload ? Copy::many() : st.fn;
            const uint32_t threads = st.kind == BatchReq::kUpload ? 256u : st.threads;
            const bool timed = profiling_ && st.kind == BatchReq::kKernel &&
                               strstr(jobs[st.items[0].j].slot.chain[st.items[0].r].name, "k_group_eval") != nullptr;
            if (timed) {
                if (timed_.size() <= timed_n_) timed_.resize(timed_n_ + 1, {nullptr, nullptr});
                auto& evp = timed_[timed_n_++];
                if (!evp.first) {
                    HIPCHK(hipEventCreate(&evp.first));
                    HIPCHK(hipEventCreate(&evp.second));
                }
                HIPCHK(hipEventRecord(evp.first, s));
                launch_many(fn, step_blocks_[si], threads, lds, dev + step_off_[si].first,
                            dev + step_off_[si].second, (int)st.items.size(), s);
                HIPCHK(hipEventRecord(evp.second, s));
            } else {
                launch_many(fn, step_blocks_[si], threads, lds, dev + step_off_[si].first,
                            dev + step_off_[si].second, (int)st.items.size(), s);
            }
        }
        if (nreads) {
            launch_many(Copy::many(), rblocks, 256, 0, dev + rf, dev + re, (int)nreads, s);
            HIPCHK(hipMemcpyAsync(rhost_, rdev_.ptr, rbytes, hipMemcpyDeviceToHost, s));
        }
        const double t1 = stats_on() ? now() : 0;
        HIPCHK(hipStreamSynchronize(s));
        if (profiling_) {
            for (size_t k = 0; k < timed_n_; k++) {
                float t = 0;
                HIPCHK(hipEventElapsedTime(&t, timed_[k].first, timed_[k].second));
                p_ms_ += t;
                p_launches_ += 1;
            }
            timed_n_ = 0;
            p_flushes_ += 1;
            for (const Step& st : steps_) (st.kind == BatchReq::kHost ? p_library_ : p_merged_) += 1;
        }
        if (stats_on()) {
            const double t2 = now();
            stats.t_pack += t1 - t0;
            stats.t_wait += t2 - t1;
            stats.flushes++;
            stats.reads += (long)nreads;
            for (const Step& st : steps_) (st.kind == BatchReq::kHost ? stats.host_steps : stats.launches)++;
        }
        if (nreads) {
            size_t ro = 0;
            for (Job& jb : jobs)
                for (const BatchRead& rd : jb.slot.reads) {
                    ro = up(ro, 16);
                    memcpy(rd.dst, rhost_ + ro, rd.bytes);
                    ro += rd.bytes;
                }
        }
        for (Job& jb : jobs) {
            jb.slot.clear_chain();
            jb.slot.reads.clear();
        }
    }

    void launch_many(const void* fn, unsigned blocks, uint32_t threads, size_t lds, uint8_t* first, uint8_t* ents,
                     int n, hipStream_t s) {
        if (blocks == 0 || n == 0) return;
        if (!launch_ok(dim3(blocks), dim3(threads), lds, "merged launch")) return;
        allow_lds(fn, lds);
        void* args[] = {(void*)&first, (void*)&ents, (void*)&n};
        HIPCHK(hipLaunchKernel(fn, dim3(blocks), dim3(threads), args, lds, s));
        launch_done(s, "merged launch");
    }

    struct Key {
        BatchReq::Kind kind;
        const void* fn;
        uint32_t threads;
        int count;
    };
    bool profiling_ = false;
    DBuf<unsigned long long> prof_acc_;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> timed_;
    size_t timed_n_ = 0;
    double p_launches_ = 0, p_ms_ = 0, p_flushes_ = 0, p_merged_ = 0, p_library_ = 0;
    std::vector<Step> steps_;
    std::vector<size_t> head_;
    std::vector<Key> keys_;
    std::vector<size_t> pay_off_;
    std::vector<std::pair<size_t, size_t>> step_off_;
    std::vector<unsigned> step_blocks_;
    DBuf<uint8_t> arena_, rdev_;
    uint8_t* stage_ = nullptr;
    uint8_t* rhost_ = nullptr;
    size_t stage_cap_ = 0, rhost_cap_ = 0;
};

// The calling thread's scheduler; it keeps its buffers for the next batch call and lives as
// long as the process (pool threads outlive main, so it is never destroyed).
BatchRunner& batch_runner() {
    static thread_local BatchRunner* r = new BatchRunner();
    return *r;
}


}  // namespace

extern "C" {

void* pycauchy_initialize_lti(int num_steps, double* A0, int size_A0, double* p0, int size_p0, double* b0,
                              int size_b0, double* Phi, int size_Phi, double* Gamma, int size_Gamma, double* B,
                              int size_B, double* beta, int size_beta, double* H, int size_H, double* gamma,
                              int size_gamma, double dt, int init_step, bool debug_print) {
    (void)size_A0;
    (void)size_p0;
    try {
        const int n = size_b0;
        const int pncc = size_beta;
        const int cmcc = n > 0 ? size_B / n : 0;
        const int p = size_gamma;
        Handle* h = new Handle;
        h->Phi.assign(Phi, Phi + size_Phi);
        h->Gamma.assign(Gamma, Gamma + size_Gamma);
        h->B.assign(B, B + size_B);
        h->H.assign(H, H + size_H);
        h->beta.assign(beta, beta + size_beta);
        h->gamma.assign(gamma, gamma + size_gamma);
        h->x.assign(n, 0.0);
        h->Phi.resize(std::max(n * n, 1));
        h->Gamma.resize(std::max(n * pncc, 1));
        h->B.resize(std::max(n * cmcc, 1));
        h->H.resize(std::max(p * n, 1));
        h->beta.resize(std::max(pncc, 1));
        h->gamma.resize(std::max(p, 1));
        CauchyDynamicsUpdateContainer& duc = h->duc;
        memset(&duc, 0, sizeof(duc));
        duc.x = h->x.data();
        duc.u = nullptr;
        duc.dt = dt;
        duc.step = init_step;
        duc.n = n;
        duc.cmcc = cmcc;
        duc.pncc = pncc;
        duc.p = p;
        duc.Phi = h->Phi.data();
        duc.Gamma = h->Gamma.data();
        duc.B = h->B.data();
        duc.H = h->H.data();
        duc.beta = h->beta.data();
        duc.gamma = h->gamma.data();
        h->eng = new Engine(num_steps, n, cmcc, pncc, p, A0, p0, b0);
        h->eng->debug_after_ftr = debug_print;
        return h;
    } catch (const std::exception& e) {
        set_err("pycauchy_initialize_lti", e);
        return nullptr;
    }
}

void* pycauchy_initialize_ltv(int num_steps, double* A0, int size_A0, double* p0, int size_p0, double* b0,
                              int size_b0, double* Phi, int size_Phi, double* Gamma, int size_Gamma, double* B,
                              int size_B, double* beta, int size_beta, double* H, int size_H, double* gamma,
                              int size_gamma, void (*f_dyn_update_callback)(CauchyDynamicsUpdateContainer*),
                              double dt, int init_step, bool debug_print) {
    Handle* h = (Handle*)pycauchy_initialize_lti(num_steps, A0, size_A0, p0, size_p0, b0, size_b0, Phi, size_Phi,
                                                 Gamma, size_Gamma, B, size_B, beta, size_beta, H, size_H, gamma,
                                                 size_gamma, dt, init_step, debug_print);
    if (h) h->cb = f_dyn_update_callback;
    return h;
}

void pycauchy_single_step_ltiv(void* handle, double* msmts, int size_msmts, double* controls, int size_controls,
                               double** out_Phi, int* size_out_Phi, double** out_Gamma, int* size_out_Gamma,
                               double** out_B, int* size_out_B, double** out_H, int* size_out_H, double** out_beta,
                               int* size_out_beta, double** out_gamma, int* size_out_gamma, double** out_fz,
                               int* size_out_fz, double** out_xhat, int* size_out_xhat, double** out_Phat,
                               int* size_out_Phat, double** out_cerr_fz, int* size_out_cerr_fz,
                               double** out_cerr_xhat, int* size_out_cerr_xhat, double** out_cerr_Phat,
                               int* size_out_cerr_Phat, int** out_err_code, int* size_out_err_code) {
    run_sync(step_ltiv_task(handle, msmts, size_msmts, controls, size_controls, out_Phi, size_out_Phi, out_Gamma,
                            size_out_Gamma, out_B, size_out_B, out_H, size_out_H, out_beta, size_out_beta, out_gamma,
                            size_out_gamma, out_fz, size_out_fz, out_xhat, size_out_xhat, out_Phat, size_out_Phat,
                            out_cerr_fz, size_out_cerr_fz, out_cerr_xhat, size_out_cerr_xhat, out_cerr_Phat,
                            size_out_cerr_Phat, out_err_code, size_out_err_code, nullptr));
}

void* pycauchy_initialize_nonlin(int num_steps, double* x0, int size_x0, double* A0, int size_A0, double* p0,
                                 int size_p0, double* b0, int size_b0, double* beta, int size_beta, double* gamma,
                                 int size_gamma, void (*f_dyn_update_callback)(CauchyDynamicsUpdateContainer*),
                                 void (*f_nonlinear_msmt_model)(CauchyDynamicsUpdateContainer*, double*),
                                 void (*f_extended_msmt_update_callback)(CauchyDynamicsUpdateContainer*), int cmcc,
                                 double dt, int init_step, bool debug_print) {
    (void)size_A0;
    (void)size_p0;
    try {
        const int n = size_b0;
        const int pncc = size_beta;
        const int p = size_gamma;
        if (size_x0 != n) throw std::runtime_error("x0 must hold n = size_b0 entries");
        if (!f_dyn_update_callback || !f_nonlinear_msmt_model || !f_extended_msmt_update_callback)
            throw std::runtime_error("all three callbacks are required");
        Handle* h = new Handle;
        h->x.assign(x0, x0 + n);
        h->Phi.assign(std::max(n * n, 1), 0.0);
        h->Gamma.assign(std::max(n * pncc, 1), 0.0);
        h->B.assign(std::max(n * cmcc, 1), 0.0);
        h->H.assign(std::max(p * n, 1), 0.0);
        h->beta.assign(std::max(pncc, 1), 0.0);
        std::copy(beta, beta + size_beta, h->beta.begin());
        h->gamma.assign(std::max(p, 1), 0.0);
        std::copy(gamma, gamma + size_gamma, h->gamma.begin());
        h->zbar.assign(std::max(p, 1), 0.0);
        CauchyDynamicsUpdateContainer& duc = h->duc;
        memset(&duc, 0, sizeof(duc));
        duc.x = h->x.data();
        duc.dt = dt;
        duc.step = init_step;
        duc.n = n;
        duc.cmcc = cmcc;
        duc.pncc = pncc;
        duc.p = p;
        duc.Phi = h->Phi.data();
        duc.Gamma = h->Gamma.data();
        duc.B = h->B.data();
        duc.H = h->H.data();
        duc.beta = h->beta.data();
        duc.gamma = h->gamma.data();
        h->cb = f_dyn_update_callback;
        h->msmt_model = f_nonlinear_msmt_model;
        h->ext_update = f_extended_msmt_update_callback;
        // the estimator propagates the state deviation, which has no control term
        h->eng = new Engine(num_steps, n, 0, pncc, p, A0, p0, b0);
        h->eng->debug_after_ftr = debug_print;
        return h;
    } catch (const std::exception& e) {
        set_err("pycauchy_initialize_nonlin", e);
        return nullptr;
    }
}

void pycauchy_single_step_nonlin(void* handle, double* msmts, int size_msmts, double* controls, int size_controls,
                                 bool with_propagate, double** out_Phi, int* size_out_Phi, double** out_Gamma,
                                 int* size_out_Gamma, double** out_B, int* size_out_B, double** out_H,
                                 int* size_out_H, double** out_beta, int* size_out_beta, double** out_gamma,
                                 int* size_out_gamma, double** out_fz, int* size_out_fz, double** out_xhat,
                                 int* size_out_xhat, double** out_Phat, int* size_out_Phat, double** out_xbar,
                                 int* size_out_xbar, double** out_zbar, int* size_out_zbar, double** out_cerr_fz,
                                 int* size_out_cerr_fz, double** out_cerr_xhat, int* size_out_cerr_xhat,
                                 double** out_cerr_Phat, int* size_out_cerr_Phat, int** out_err_code,
                                 int* size_out_err_code) {
    (void)size_controls;
    Handle* h = (Handle*)handle;
    Engine* e = h->eng;
    CauchyDynamicsUpdateContainer& duc = h->duc;
    duc.u = controls;
    const int p = duc.p, n = duc.n, cmcc = duc.cmcc, pncc = duc.pncc;
    const int nm = size_msmts;
    // outputs (pycauchy.hpp:580-604); the dynamics are this step's, or the current
    // ones when the step does not propagate
    *size_out_Phi = n * n;
    *out_Phi = mallocd<double>(n * n);
    *size_out_Gamma = n * pncc;
    *out_Gamma = mallocd<double>(n * pncc);
    *size_out_B = n * cmcc;
    *out_B = mallocd<double>(n * cmcc);
    *size_out_H = nm * n;
    *out_H = mallocd<double>(nm * n);
    *size_out_beta = pncc;
    *out_beta = mallocd<double>(pncc);
    *size_out_gamma = nm;
    *out_gamma = mallocd<double>(nm);
    *size_out_fz = nm;
    *out_fz = mallocd<double>(nm);
    *size_out_xhat = nm * n;
    *out_xhat = mallocd<double>(nm * n);
    *size_out_Phat = nm * n * n;
    *out_Phat = mallocd<double>(nm * n * n);
    *size_out_xbar = nm * n;
    *out_xbar = mallocd<double>(nm * n);
    *size_out_zbar = nm;
    *out_zbar = mallocd<double>(nm);
    *size_out_cerr_fz = nm;
    *out_cerr_fz = mallocd<double>(nm);
    *size_out_cerr_xhat = nm;
    *out_cerr_xhat = mallocd<double>(nm);
    *size_out_cerr_Phat = nm;
    *out_cerr_Phat = mallocd<double>(nm);
    *size_out_err_code = nm;
    *out_err_code = mallocd<int>(nm);
    if (with_propagate) {
        // x_bar(k+1) = f(x_bar(k)) and the Jacobians, from the caller (pycauchy.hpp:606-619)
        duc.is_xbar_set_for_ece = false;
        h->cb(&duc);
        if (!duc.is_xbar_set_for_ece) {
            g_last_error = "pycauchy_single_step_nonlin: the dynamics callback did not set is_xbar_set_for_ece";
            fprintf(stderr, "[cauchy_mi355x] %s\n", g_last_error.c_str());
        }
    }
    memcpy(*out_Phi, duc.Phi, n * n * sizeof(double));
    memcpy(*out_B, duc.B, n * cmcc * sizeof(double));
    memcpy(*out_Gamma, duc.Gamma, n * pncc * sizeof(double));
    memcpy(*out_beta, duc.beta, pncc * sizeof(double));
    const int start = p - nm;
    double* zbar = h->zbar.data();
    h->step_fz.clear();
    h->step_cnt.clear();
    for (int i = 0; i < nm; i++) {
        const int idx = start + i;
        // zbar = h(x), then H = dh/dx at x (pycauchy.hpp:628-629); x is x_bar before the
        // first update and the updated estimate after it
        h->msmt_model(&duc, zbar);
        h->ext_update(&duc);
        memcpy(*out_xbar + i * n, duc.x, n * sizeof(double));
        (*out_zbar)[i] = zbar[idx];
        const double dz = msmts[i] - zbar[idx];
        memcpy(*out_H + i * n, duc.H + idx * n, n * sizeof(double));
        (*out_gamma)[i] = duc.gamma[idx];
        try {
            e->update(dz, duc.Phi, duc.Gamma, duc.beta, duc.H + idx * n, duc.gamma[idx], nullptr, nullptr);
            e->finalize_extended_moments(duc.x);
        } catch (const std::exception& ex) {
            set_err("pycauchy_single_step_nonlin", ex);
        }
        h->record(i, e);
        (*out_fz)[i] = e->fz.re;
        for (int j = 0; j < n; j++) (*out_xhat)[i * n + j] = e->mean[j].re;
        for (int j = 0; j < n * n; j++) (*out_Phat)[i * n * n + j] = e->cov[j].re;
        (*out_cerr_fz)[i] = e->fz.im;
        double mx = fabs(e->mean[0].im);
        for (int j = 1; j < n; j++) mx = std::max(mx, fabs(e->mean[j].im));
        (*out_cerr_xhat)[i] = mx;
        double mp = fabs(e->cov[0].im);
        for (int j = 1; j < n * n; j++) mp = std::max(mp, fabs(e->cov[j].im));
        (*out_cerr_Phat)[i] = mp;
        (*out_err_code)[i] = e->errs;
    }
    duc.step += 1;
}

int pycauchy_batch_single_step_ltiv(void** handles, int count, double** msmts, const int* size_msmts,
                                    double** controls, const int* size_controls, void** outs, int* out_sizes,
                                    int num_threads) {
    return run_concurrently(handles, count, num_threads, pycauchy_single_step_get_number_of_terms, [&](int i) {
        void** o = outs + 13 * (int64_t)i;
        int* n = out_sizes + 13 * (int64_t)i;
        double* u = controls ? controls[i] : nullptr;
        const int nu = size_controls ? size_controls[i] : 0;
        pycauchy_single_step_ltiv(handles[i], msmts[i], size_msmts[i], u, nu, (double**)&o[0], &n[0],
                                  (double**)&o[1], &n[1], (double**)&o[2], &n[2], (double**)&o[3], &n[3],
                                  (double**)&o[4], &n[4], (double**)&o[5], &n[5], (double**)&o[6], &n[6],
                                  (double**)&o[7], &n[7], (double**)&o[8], &n[8], (double**)&o[9], &n[9],
                                  (double**)&o[10], &n[10], (double**)&o[11], &n[11], (int**)&o[12], &n[12]);
    });
}

int pycauchy_batch_single_step_ltiv_merged(void** handles, int count, double** msmts, const int* size_msmts,
                                           double** controls, const int* size_controls, void** outs, int* out_sizes,
                                           int num_groups) {
    if (count <= 0) return 0;
    const int G = std::max(1, std::min(num_groups, count));
    // estimators with the most terms first, dealt round-robin so each group gets a share
    std::vector<int> order(count);
    std::vector<int> nt(count);
    for (int i = 0; i < count; i++) {
        order[i] = i;
        nt[i] = pycauchy_single_step_get_number_of_terms(handles[i]);
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return nt[a] > nt[b]; });
    std::vector<std::vector<int>> groups(G);
    for (int k = 0; k < count; k++) groups[k % G].push_back(order[k]);
    std::vector<std::string> errors(count);
    std::atomic<int> next{0};
    auto work = [&]() {
        int g;
        while ((g = next.fetch_add(1)) < G) {
            const std::vector<int>& mem = groups[g];
            std::vector<BatchRunner::Job> jobs(mem.size());
            for (size_t k = 0; k < mem.size(); k++) {
                const int i = mem[k];
                jobs[k].e = ((Handle*)handles[i])->eng;
                void** o = outs + 13 * (int64_t)i;
                int* n = out_sizes + 13 * (int64_t)i;
                double* u = controls ? controls[i] : nullptr;
                const int nu = size_controls ? size_controls[i] : 0;
                jobs[k].task.emplace(step_ltiv_task(
                    handles[i], msmts[i], size_msmts[i], u, nu, (double**)&o[0], &n[0], (double**)&o[1], &n[1],
                    (double**)&o[2], &n[2], (double**)&o[3], &n[3], (double**)&o[4], &n[4], (double**)&o[5], &n[5],
                    (double**)&o[6], &n[6], (double**)&o[7], &n[7], (double**)&o[8], &n[8], (double**)&o[9], &n[9],
                    (double**)&o[10], &n[10], (double**)&o[11], &n[11], (int**)&o[12], &n[12], &jobs[k].err));
            }
            try {
                batch_runner().run(jobs);
            } catch (const std::exception& ex) {
                for (auto& jb : jobs)
                    if (jb.err.empty()) jb.err = std::string("batch scheduler: ") + ex.what();
            }
            for (size_t k = 0; k < mem.size(); k++) errors[mem[k]] = jobs[k].err;
        }
    };
    if (G > 1) {
        const std::function<void()> job = work;
        worker_pool().run(G - 1, job);
    } else {
        work();
    }
    int failed = 0;
    std::string first;
    for (int i = 0; i < count; i++)
        if (!errors[i].empty()) {
            if (!failed) first = "estimator " + std::to_string(i) + ": " + errors[i];
            failed++;
        }
    g_last_error = first;
    return failed;
}

void mce_batch_profile_enable(int on) {
    BatchProfile& bp = batch_profile();
    std::lock_guard<std::mutex> lk(bp.mu);
    bp.on = on != 0;
    bp.launches = bp.bytes = bp.ms = bp.flushes = bp.merged = bp.library = 0;
}

void mce_batch_profile_read(double* out) {
    BatchProfile& bp = batch_profile();
    std::lock_guard<std::mutex> lk(bp.mu);
    out[0] = bp.launches;
    out[1] = bp.bytes;
    out[2] = bp.ms;
    out[3] = bp.flushes;
    out[4] = bp.merged;
    out[5] = bp.library;
}

int pycauchy_batch_single_step_nonlin(void** handles, int count, double** msmts, const int* size_msmts,
                                     double** controls, const int* size_controls, const int* with_propagate,
                                     void** outs, int* out_sizes, int num_threads) {
    return run_concurrently(handles, count, num_threads, pycauchy_single_step_get_number_of_terms, [&](int i) {
        void** o = outs + 15 * (int64_t)i;
        int* n = out_sizes + 15 * (int64_t)i;
        double* u = controls ? controls[i] : nullptr;
        const int nu = size_controls ? size_controls[i] : 0;
        pycauchy_single_step_nonlin(handles[i], msmts[i], size_msmts[i], u, nu, with_propagate[i] != 0,
                                    (double**)&o[0], &n[0], (double**)&o[1], &n[1], (double**)&o[2], &n[2],
                                    (double**)&o[3], &n[3], (double**)&o[4], &n[4], (double**)&o[5], &n[5],
                                    (double**)&o[6], &n[6], (double**)&o[7], &n[7], (double**)&o[8], &n[8],
                                    (double**)&o[9], &n[9], (double**)&o[10], &n[10], (double**)&o[11], &n[11],
                                    (double**)&o[12], &n[12], (double**)&o[13], &n[13], (int**)&o[14], &n[14]);
    });
}

void pycauchy_get_marginal_1D_pointwise_cpdf(void* handle, int marg_idx1, double gridx_low, double gridx_high,
                                             double gridx_resolution, char* log_dir, double** out_cpdf_data,
                                             int* size_out_cpdf_data, int* out_num_gridx) {
    (void)log_dir;
    std::vector<double> xy;
    int rc = 1;
    try {
        rc = ((Handle*)handle)->eng->cpdf_1d(marg_idx1, gridx_low, gridx_high, gridx_resolution, xy);
    } catch (const std::exception& ex) {
        set_err("pycauchy_get_marginal_1D_pointwise_cpdf", ex);
    }
    if (rc != 0) xy.clear();
    *out_num_gridx = (int)(xy.size() / 2);
    *size_out_cpdf_data = (int)xy.size();
    *out_cpdf_data = mallocd<double>(xy.size());
    if (!xy.empty()) memcpy(*out_cpdf_data, xy.data(), xy.size() * sizeof(double));
}

int mce_cpdf1d_noalloc(void* handle, int marg_idx, double gridx_low, double gridx_high, double gridx_resolution,
                       double* out_x, double* out_f, int capacity, void* stream, int* out_num_gridx) {
    if (out_num_gridx) *out_num_gridx = 0;
    if (!handle || !out_num_gridx) return -1;
    try {
        char* work = noalloc_scratch();
        return ((Handle*)handle)->eng->cpdf_1d_into(marg_idx, gridx_low, gridx_high, gridx_resolution, out_x, out_f,
                                                    capacity, (hipStream_t)stream, work, kNoAllocBytes,
                                                    out_num_gridx);
    } catch (const std::exception& ex) {
        set_err("mce_cpdf1d_noalloc", ex);
        return -1;
    }
}

void pycauchy_get_1D_pointwise_cpdf(void* handle, double gridx_low, double gridx_high, double gridx_resolution,
                                    char* log_dir, double** out_cpdf_data, int* size_out_cpdf_data,
                                    int* out_num_gridx) {
    pycauchy_get_marginal_1D_pointwise_cpdf(handle, 0, gridx_low, gridx_high, gridx_resolution, log_dir, out_cpdf_data,
                                            size_out_cpdf_data, out_num_gridx);
}

void pycauchy_get_marginal_2D_pointwise_cpdf(void* handle, int marg_idx1, int marg_idx2, double gridx_low,
                                             double gridx_high, double gridx_resolution, double gridy_low,
                                             double gridy_high, double gridy_resolution, char* log_dir,
                                             bool reset_cache, double** out_cpdf_data, int* size_out_cpdf_data,
                                             int* out_num_gridx, int* out_num_gridy) {
    (void)log_dir;
    (void)reset_cache;
    std::vector<double> xyz;
    int rc = 1, nx = 0, ny = 0;
    try {
        rc = ((Handle*)handle)->eng->cpdf_2d(marg_idx1, marg_idx2, gridx_low, gridx_high, gridx_resolution, gridy_low,
                                            gridy_high, gridy_resolution, xyz, nx, ny);
    } catch (const std::exception& ex) {
        set_err("pycauchy_get_marginal_2D_pointwise_cpdf", ex);
    }
    if (rc != 0) {
        xyz.clear();
        nx = ny = 0;
    }
    *out_num_gridx = nx;
    *out_num_gridy = ny;
    *size_out_cpdf_data = (int)xyz.size();
    *out_cpdf_data = mallocd<double>(xyz.size());
    if (!xyz.empty()) memcpy(*out_cpdf_data, xyz.data(), xyz.size() * sizeof(double));
}

void pycauchy_get_2D_pointwise_cpdf(void* handle, double gridx_low, double gridx_high, double gridx_resolution,
                                    double gridy_low, double gridy_high, double gridy_resolution, char* log_dir,
                                    bool reset_cache, double** out_cpdf_data, int* size_out_cpdf_data,
                                    int* out_num_gridx, int* out_num_gridy) {
    pycauchy_get_marginal_2D_pointwise_cpdf(handle, 0, 1, gridx_low, gridx_high, gridx_resolution, gridy_low,
                                            gridy_high, gridy_resolution, log_dir, reset_cache, out_cpdf_data,
                                            size_out_cpdf_data, out_num_gridx, out_num_gridy);
}

void pycauchy_single_step_eval_2d_rsys_cpdf(double* Trel, int size_Trel, void* _s_pcdh, void* _p_pcdh,
                                            double RSYS_APPROX_EPS, double xlow, double xhigh, double delta_x,
                                            double ylow, double yhigh, double delta_y, double** out_rsys_fz,
                                            int* size_out_rsys_fz, double** out_rsys_xhat, int* size_out_rsys_xhat,
                                            double** out_rsys_Phat, int* size_out_rsys_Phat,
                                            double** out_rsys_cerr_fz, int* size_out_rsys_cerr_fz,
                                            double** out_rsys_cerr_xhat, int* size_out_rsys_cerr_xhat,
                                            double** out_rsys_cerr_Phat, int* size_out_rsys_cerr_Phat,
                                            double** out_cpdf_data, int* size_out_cpdf_data, int* out_num_gridx,
                                            int* out_num_gridy) {
    *size_out_rsys_fz = *size_out_rsys_cerr_fz = 1;
    *size_out_rsys_xhat = *size_out_rsys_cerr_xhat = 2;
    *size_out_rsys_Phat = *size_out_rsys_cerr_Phat = 4;
    *out_rsys_fz = mallocd<double>(1);
    *out_rsys_cerr_fz = mallocd<double>(1);
    *out_rsys_xhat = mallocd<double>(2);
    *out_rsys_cerr_xhat = mallocd<double>(2);
    *out_rsys_Phat = mallocd<double>(4);
    *out_rsys_cerr_Phat = mallocd<double>(4);
    cplx nf = cmake(0, 0), mean[2] = {cmake(0, 0), cmake(0, 0)}, cov[4];
    for (auto& c : cov) c = cmake(0, 0);
    std::vector<double> xyz;
    int nx = 0, ny = 0;
    try {
        Engine* s = ((Handle*)_s_pcdh)->eng;
        Engine* p = ((Handle*)_p_pcdh)->eng;
        if (size_Trel != 2 * p->d) throw std::runtime_error("Trel must be 2 x n");
        rsys_2d(*p, *s, Trel, RSYS_APPROX_EPS, xlow, xhigh, delta_x, ylow, yhigh, delta_y, nf, mean, cov, xyz, nx, ny);
    } catch (const std::exception& ex) {
        set_err("pycauchy_single_step_eval_2d_rsys_cpdf", ex);
        xyz.clear();
        nx = ny = 0;
    }
    (*out_rsys_fz)[0] = nf.re;
    (*out_rsys_cerr_fz)[0] = nf.im;
    for (int i = 0; i < 2; i++) {
        (*out_rsys_xhat)[i] = mean[i].re;
        (*out_rsys_cerr_xhat)[i] = mean[i].im;
    }
    for (int i = 0; i < 4; i++) {
        (*out_rsys_Phat)[i] = cov[i].re;
        (*out_rsys_cerr_Phat)[i] = cov[i].im;
    }
    *size_out_cpdf_data = (int)xyz.size();
    *out_cpdf_data = mallocd<double>(xyz.size());
    if (!xyz.empty()) memcpy(*out_cpdf_data, xyz.data(), xyz.size() * sizeof(double));
    *out_num_gridx = nx;
    *out_num_gridy = ny;
}

void pycauchy_single_step_reset(void* handle, double* A0, int size_A0, double* p0, int size_p0, double* b0,
                                int size_b0, double* xbar, int size_xbar) {
    Handle* h = (Handle*)handle;
    Engine* e = h->eng;
    const int n = e->d;
    if (size_A0 > 0) memcpy(e->A0.data(), A0, n * n * sizeof(double));
    if (size_p0 > 0) memcpy(e->p0.data(), p0, n * sizeof(double));
    if (size_b0 > 0) memcpy(e->b0.data(), b0, n * sizeof(double));
    if (size_xbar > 0) memcpy(h->duc.x, xbar, n * sizeof(double));
    try {
        e->reset_terms();
    } catch (const std::exception& ex) {
        set_err("pycauchy_single_step_reset", ex);
    }
}

void pycauchy_single_step_set_master_step(void* handle, int step) { ((Handle*)handle)->eng->master_step = step; }

void pycauchy_single_step_set_window_number(void* handle, int win_num) { ((Handle*)handle)->eng->win_num = win_num; }

int pycauchy_single_step_get_number_of_terms(void* handle) { return ((Handle*)handle)->eng->Nt; }

void pycauchy_single_step_deterministic_transform(void* handle, double* Trans, int size_Trans, double* bias,
                                                  int size_bias) {
    (void)size_Trans;
    try {
        ((Handle*)handle)->eng->deterministic_transform(Trans, size_bias > 0 ? bias : nullptr);
    } catch (const std::exception& ex) {
        set_err("pycauchy_single_step_deterministic_transform", ex);
    }
}

void* pycauchy_single_step_get_duc(void* handle) { return &((Handle*)handle)->duc; }

void pycauchy_single_step_shutdown(void* handle) {
    Handle* h = (Handle*)handle;
    if (!h) return;
    delete h->eng;
    delete h;
}

int pycauchy_set_tr_search_idxs_ordering(int* ordering, int size_ordering) {
    if (size_ordering > 12) {
        fprintf(stderr, "[ERROR pycauchy_set_tr_search_idxs_ordering:] Cannot run estimation problems over 12 dimensions currently!\n");
        return 1;
    }
    // every entry becomes an axis index into a term's b vector: refuse what cannot be one
    if (size_ordering < 0 || (size_ordering > 0 && !ordering)) {
        fprintf(stderr, "[ERROR pycauchy_set_tr_search_idxs_ordering:] invalid ordering array\n");
        return 1;
    }
    for (int i = 0; i < size_ordering; i++)
        if (ordering[i] < 0 || ordering[i] >= 12) {
            fprintf(stderr, "[ERROR pycauchy_set_tr_search_idxs_ordering:] axis %d at position %d is outside 0..11\n",
                    ordering[i], i);
            return 1;
        }
    for (int i = 0; i < size_ordering; i++) g_search_order[i] = ordering[i];
    return 0;
}

void pycauchy_speyers_window_init(double* xhat, int size_xhat, double* Phat, int size_Phat, double* H, int size_H,
                                  double gamma, double z, double** out_A0, int* size_out_A0, double** out_p0,
                                  int* size_out_p0, double** out_b0, int* size_out_b0) {
    const int n = size_xhat;
    if (size_Phat != n * n || size_H != n) {
        fprintf(stderr, "[cauchy_mi355x] pycauchy_speyers_window_init: size mismatch\n");
        exit(1);
    }
    *out_A0 = mallocd<double>((size_t)n * n);
    *size_out_A0 = n * n;
    *out_p0 = mallocd<double>(n);
    *size_out_p0 = n;
    *out_b0 = mallocd<double>(n);
    *size_out_b0 = n;
    std::vector<double> var(Phat, Phat + (size_t)n * n);
    host::speyers_window_init(n, xhat, var.data(), H, gamma, z, *out_A0, *out_p0, *out_b0, kCovEigenvalueTol);
}

void pycauchy_get_reinitialization_statistics(void* handle, double z, double* xhat, int size_xhat, double* Phat,
                                               int size_Phat, double* H, int size_H, double gamma, double** out_A0,
                                               int* size_out_A0, double** out_p0, int* size_out_p0, double** out_b0,
                                               int* size_out_b0) {
    Handle* h = (Handle*)handle;
    const int n = h->duc.n;
    if (size_xhat != n || size_Phat != n * n || size_H != n) {
        fprintf(stderr, "[cauchy_mi355x] pycauchy_get_reinitialization_statistics: size mismatch\n");
        exit(1);
    }
    pycauchy_speyers_window_init(xhat, size_xhat, Phat, size_Phat, H, size_H, gamma, z, out_A0, size_out_A0, out_p0,
                                 size_out_p0, out_b0, size_out_b0);
}

void* mce_estimator_create(double* A0, double* p0, double* b0, int steps, int d, int cmcc, int pncc, int p,
                           bool print_basic_info) {
    try {
        Handle* h = new Handle;
        h->x.assign(d, 0.0);
        memset(&h->duc, 0, sizeof(h->duc));
        h->duc.x = h->x.data();
        h->duc.n = d;
        h->duc.cmcc = cmcc;
        h->duc.pncc = pncc;
        h->duc.p = p;
        h->eng = new Engine(steps, d, cmcc, pncc, p, A0, p0, b0);
        h->eng->debug_after_ftr = print_basic_info;
        return h;
    } catch (const std::exception& ex) {
        set_err("mce_estimator_create", ex);
        return nullptr;
    }
}

int mce_estimator_step(void* handle, double msmt, double* Phi, double* Gamma, double* beta, double* H,
                       double gamma, double* B, double* u) {
    Engine* e = ((Handle*)handle)->eng;
    if ((e->errs & (1 << 9)) || e->master_step == e->n_mu) return -1;
    try {
        return e->update(msmt, Phi, Gamma, beta, H, gamma, B, u);
    } catch (const std::exception& ex) {
        set_err("mce_estimator_step", ex);
        return -2;
    }
}

int mce_estimator_enable_tangents(void* handle, int num_directions) {
    if (!handle) return -1;
    try {
        ((Handle*)handle)->eng->enable_tangents(num_directions);
        return 0;
    } catch (const std::exception& ex) {
        set_err("mce_estimator_enable_tangents", ex);
        return -1;
    }
}

int mce_estimator_step_tangent(void* handle, double msmt, double* Phi, double* Gamma, double* beta, double* H,
                               double gamma, double* B, double* u, const double* dPhi, const double* dbeta,
                               const double* dH, const double* dgamma, double* dfz) {
    Engine* e = ((Handle*)handle)->eng;
    if (!e->ntan) {
        set_err("mce_estimator_step_tangent", std::runtime_error("tangent mode is not enabled"));
        return -2;
    }
    if ((e->errs & (1 << 9)) || e->master_step == e->n_mu) return -1;
    const int P = e->ntan, d = e->d;
    memset(&e->htan, 0, sizeof(e->htan));
    for (int k = 0; k < P; k++) {
        for (int i = 0; i < d * d; i++) e->htan.Phi[k][i] = dPhi ? dPhi[(size_t)k * d * d + i] : 0.0;
        for (int i = 0; i < d; i++) e->htan.H[k][i] = dH ? dH[(size_t)k * d + i] : 0.0;
        e->htan.gamma[k] = dgamma ? dgamma[k] : 0.0;
    }
    e->tan_dbeta.assign((size_t)P * e->pncc, 0.0);
    if (dbeta)
        for (size_t i = 0; i < e->tan_dbeta.size(); i++) e->tan_dbeta[i] = dbeta[i];
    e->tan_seeded = true;
    int rc;
    try {
        rc = e->update(msmt, Phi, Gamma, beta, H, gamma, B, u);
    } catch (const std::exception& ex) {
        e->tan_seeded = false;
        set_err("mce_estimator_step_tangent", ex);
        return -2;
    }
    e->tan_seeded = false;
    for (int k = 0; k < P; k++) {
        dfz[2 * k] = e->dfz_tan[k].re;
        dfz[2 * k + 1] = e->dfz_tan[k].im;
    }
    return rc;
}

int mce_estimator_moments(void* handle, double* fz, double* mean, double* cov) {
    Engine* e = ((Handle*)handle)->eng;
    fz[0] = e->fz.re;
    fz[1] = e->fz.im;
    for (int i = 0; i < e->d; i++) {
        mean[2 * i] = e->mean[i].re;
        mean[2 * i + 1] = e->mean[i].im;
    }
    for (int i = 0; i < e->d * e->d; i++) {
        cov[2 * i] = e->cov[i].re;
        cov[2 * i + 1] = e->cov[i].im;
    }
    return e->Nt;
}

void mce_estimator_reset(void* handle) {
    try {
        ((Handle*)handle)->eng->reset_terms();
    } catch (const std::exception& ex) {
        set_err("mce_estimator_reset", ex);
    }
}

void mce_estimator_destroy(void* handle) { pycauchy_single_step_shutdown(handle); }

int mce_estimator_finalize_extended_moments(void* handle, double* x_bar) {
    try {
        ((Handle*)handle)->eng->finalize_extended_moments(x_bar);
        return 0;
    } catch (const std::exception& ex) {
        set_err("mce_estimator_finalize_extended_moments", ex);
        return -1;
    }
}

int mce_last_update(void* handle, double* out_mom, int* out_cnt) {
    Engine* e = ((Handle*)handle)->eng;
    if (!e->have_diag) return -1;
    memcpy(out_mom, e->diag_mom.data(), e->diag_mom.size() * sizeof(double));
    memcpy(out_cnt, e->diag_cnt.data(), e->diag_cnt.size() * sizeof(int));
    return e->errs;
}

int mce_shape_range(void* handle) { return ((Handle*)handle)->eng->sr; }

long long mce_terms_generated(void* handle) { return ((Handle*)handle)->eng->terms_generated; }

int mce_term(void* handle, int shape, int idx, int* info, double* A, double* p, double* b, unsigned int* keys,
             double* gvals) {
    try {
        return ((Handle*)handle)->eng->term(shape, idx, info, A, p, b, keys, gvals);
    } catch (const std::exception& ex) {
        set_err("mce_term", ex);
        return -1;
    }
}

void mce_set_random_draws(void* handle, const double* root_point, const double* b_pert_8xmax) {
    Engine* e = ((Handle*)handle)->eng;
    memcpy(e->root.data(), root_point, e->d * sizeof(double));
    memcpy(e->bpert.data(), b_pert_8xmax, 8 * e->MM * sizeof(double));
    e->upload_draws();
}

void mce_get_random_draws(void* handle, double* root_point, double* b_pert_8xmax) {
    Engine* e = ((Handle*)handle)->eng;
    memcpy(root_point, e->root.data(), e->d * sizeof(double));
    memcpy(b_pert_8xmax, e->bpert.data(), 8 * e->MM * sizeof(double));
}

void mce_profile_enable(void* handle, int on) { ((Handle*)handle)->eng->prof.on = on != 0; }

void mce_profile_read(void* handle, double* out) {
    Profiler& p = ((Handle*)handle)->eng->prof;
    for (int i = 0; i < kNumPhases; i++) out[i] = p.ms[i];
    out[kNumPhases] = p.total;
    out[kNumPhases + 1] = (double)p.geval_launches;
    out[kNumPhases + 2] = p.geval_bytes;
    out[kNumPhases + 3] = p.geval_ms;
    out[kNumPhases + 4] = p.groups;
    out[kNumPhases + 5] = p.members;
    out[kNumPhases + 6] = p.max_group;
    out[kNumPhases + 7] = p.alg_bytes;
    out[kNumPhases + 8] = p.alg_updates;
}

void mce_last_fz_raw(void* handle, double* out2) {
    const Engine* e = ((Handle*)handle)->eng;
    out2[0] = e->fz_raw.re;
    out2[1] = e->fz_raw.im;
}

int mce_estimator_set_shapes(void* handle, int enum_threads, int big_group_members, int geval_mode) {
    try {
        Engine* e = ((Handle*)handle)->eng;
        if (enum_threads != 0) {
            if (enum_threads != 64 && enum_threads != 128 && enum_threads != 256)
                throw std::runtime_error("enumerator lanes per parent must be 64, 128 or 256");
            e->enum_threads = enum_threads;
        }
        if (big_group_members > 0) e->big_group_members = big_group_members;
        if (geval_mode != 0) {
            if (geval_mode < -1 || geval_mode > kLookupRankVals + 1)
                throw std::runtime_error("g-table lookup mode must be -1 (automatic) or 1..4 (mode + 1)");
            e->geval_mode_override = geval_mode < 0 ? -1 : geval_mode - 1;
        }
        return 0;
    } catch (const std::exception& ex) {
        set_err("mce_estimator_set_shapes", ex);
        return -1;
    }
}

void mce_profile_reset(void* handle) {
    Profiler& p = ((Handle*)handle)->eng->prof;
    for (int i = 0; i < kNumPhases; i++) p.ms[i] = 0;
    p.total = 0;
    p.geval_launches = 0;
    p.groups = 0;
    p.members = 0;
    p.max_group = 0;
    p.geval_bytes = 0;
    p.geval_ms = 0;
    p.alg_bytes = 0;
    p.alg_updates = 0;
}

const char* mce_last_error(void) { return g_last_error.c_str(); }

int mce_step_record(void* handle, double* out_fz, int* out_counts, int max_updates) {
    const Handle* h = (const Handle*)handle;
    const int n = (int)h->step_fz.size();
    for (int i = 0; i < std::min(n, max_updates); i++) {
        out_fz[i] = h->step_fz[i];
        out_counts[2 * i] = h->step_cnt[2 * i];
        out_counts[2 * i + 1] = h->step_cnt[2 * i + 1];
    }
    return n;
}

long long mce_collect_outputs(void** outs, const int* sizes, long long count, int nout, double* dbuf, int* ibuf) {
    long long nd = 0, ni = 0;
    for (long long i = 0; i < count; i++)
        for (int j = 0; j < nout; j++) {
            const long long k = i * nout + j;
            const int n = std::max(sizes[k], 0);
            if (j == nout - 1) {
                if (n) memcpy(ibuf + ni, outs[k], n * sizeof(int));
                ni += n;
            } else {
                if (n) memcpy(dbuf + nd, outs[k], n * sizeof(double));
                nd += n;
            }
            free(outs[k]);
            outs[k] = nullptr;
        }
    return nd;
}

}  // extern "C"

#include "mce_window_bank.h"
