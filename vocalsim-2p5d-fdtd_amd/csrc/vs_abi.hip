// vs_abi.hip -- host side of the library: the float32 geometry prelude,
// device buffer management and the extern "C" entry points declared in
// include/vocalsim_abi.h.

#include "vs_common.h"
#include "../../include/vocalsim_abi.h"

#include <algorithm>
#include <atomic>
#include <thread>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <vector>
#include <mutex>
#include <chrono>
#include <array>

namespace vs {
namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

#define VS_HIP(call)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail(VS_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));        \
    } while (0)

// A copy between host and device memory on a handle's stream, waited for:
// ordered after the handle's earlier work and complete on return.  A plain
// hipMemcpy runs on the null stream, which the handles' non-blocking streams
// do not wait for, and from pageable memory it may return before the data
// has landed; a kernel on the handle's stream could then read the buffer
// before it is written.
hipError_t copy_on(hipStream_t st, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e;
}

// Runs the body of an entry point.  A C++ exception (host memory exhausted,
// or a size no standard container can hold) becomes a status code instead of
// crossing the C ABI.
template <typename F>
int guarded(const char* name, F&& body) noexcept {
    try {
        return body();
    } catch (const std::bad_alloc&) {
        return fail(VS_E_HOST_MEMORY, std::string(name) + ": host memory allocation failed");
    } catch (const std::exception& e) {
        return fail(VS_E_STATE, std::string(name) + ": " + e.what());
    } catch (...) {
        return fail(VS_E_STATE, std::string(name) + ": unexpected exception");
    }
}

// Excitation table (GlottalExcitations.cu:9-52) in Real.
template <typename Real>
void source_table(const vs_tube_spec& s, const PhysTables& ph, int num_steps, std::vector<Real>& out) {
    out.assign(num_steps, Real(0));
    const double dt = ph.dt;
    if (s.source_type == 0) {
        const Real f = 440.0f, amp = 25;
        for (int i = 0; i < num_steps; ++i) {
            const Real t = i * dt;
            out[i] = amp * std::sin(2 * M_PI * f * t);
        }
    } else if (s.source_type == 1) {
        Real f0 = 10.0f * 1e3f;
        Real peak = 0.646f / f0;
        Real width = 0.29f * peak;
        Real amp = 0.025f;
        if (s.gauss_amp > 0) amp = (Real)s.gauss_amp;
        if (s.gauss_t0 > 0) peak = (Real)s.gauss_t0;
        if (s.gauss_width > 0) width = (Real)s.gauss_width;
        for (int i = 0; i < num_steps; ++i) {
            const Real t = i * dt;
            const Real e = (t - peak) / width;
            out[i] = amp * std::exp(-1 * (e * e));
        }
    }
}

}  // namespace
}  // namespace vs

using namespace vs;

// ---------------------------------------------------------------------------
// Cache of large device memory across handles.  A handle takes its field
// memory in chunks, and the on-chip reverse pass of a batch its record slabs
// (64 tubes x 512 steps: about 14 GB); mapping and unmapping them costs
// milliseconds (about 0.1 ms per allocation), paid on every pass when the
// caller creates one handle per forward pass, as the autograd Function does.
// Pieces of at least kSlabCacheMin bytes that a handle releases are kept per
// device, up to kSlabCacheCap bytes over all devices (the oldest are freed
// beyond it), and handed to the next request of between half their size and
// their size.  Every user of a piece zeroes or writes it before reading it,
// as with fresh hipMalloc memory.  The cache is never destroyed, so no
// device call runs at process exit; vs_release_cached_memory frees what it
// holds, and a failed allocation frees the device's cached pieces before its
// one retry.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kSlabCacheMin = size_t(4) << 20;
constexpr size_t kSlabCacheCap = size_t(64) << 30;
// a handle's field chunks: its estimated set-up field memory, within these bounds
constexpr size_t kArenaChunkMin = kSlabCacheMin, kArenaChunkMax = size_t(64) << 20;
struct CachedSlab {
    int dev;
    size_t bytes;
    void* p;
};
struct SlabCache {
    std::mutex mu;
    std::vector<CachedSlab> slabs;   // oldest first
    size_t held = 0;
};
SlabCache& slab_cache() {
    static SlabCache* c = new SlabCache;   // intentionally never destroyed
    return *c;
}
void free_on(int dev, void* p) {
    int cur = -1;
    const bool other = hipGetDevice(&cur) == hipSuccess && cur != dev;
    if (other) (void)hipSetDevice(dev);
    (void)hipFree(p);
    if (other) (void)hipSetDevice(cur);
}
size_t slab_cached_bytes(int dev) {
    SlabCache& c = slab_cache();
    std::lock_guard<std::mutex> g(c.mu);
    size_t b = 0;
    for (const CachedSlab& sl : c.slabs)
        if (sl.dev == dev) b += sl.bytes;
    return b;
}
// The smallest cached slab of device dev holding bytes, at most twice them.
void* slab_take(int dev, size_t bytes, size_t* got) {
    if (bytes < kSlabCacheMin) return nullptr;
    SlabCache& c = slab_cache();
    std::lock_guard<std::mutex> g(c.mu);
    size_t best = c.slabs.size();
    for (size_t i = 0; i < c.slabs.size(); ++i) {
        const CachedSlab& sl = c.slabs[i];
        if (sl.dev == dev && sl.bytes >= bytes && sl.bytes / 2 <= bytes &&
            (best == c.slabs.size() || sl.bytes < c.slabs[best].bytes))
            best = i;
    }
    if (best == c.slabs.size()) return nullptr;
    void* p = c.slabs[best].p;
    *got = c.slabs[best].bytes;
    c.held -= *got;
    c.slabs.erase(c.slabs.begin() + (long)best);
    return p;
}
void slab_give(int dev, void* p, size_t bytes) {
    if (!p) return;
    if (bytes < kSlabCacheMin) {
        free_on(dev, p);
        return;
    }
    SlabCache& c = slab_cache();
    std::lock_guard<std::mutex> g(c.mu);
    c.slabs.push_back({dev, bytes, p});
    c.held += bytes;
    while (c.held > kSlabCacheCap) {
        free_on(c.slabs.front().dev, c.slabs.front().p);
        c.held -= c.slabs.front().bytes;
        c.slabs.erase(c.slabs.begin());
    }
}
// Frees the cached slabs of device dev (of every device when dev < 0).
void slab_release(int dev) {
    SlabCache& c = slab_cache();
    std::lock_guard<std::mutex> g(c.mu);
    std::vector<CachedSlab> keep;
    for (const CachedSlab& sl : c.slabs) {
        if (dev < 0 || sl.dev == dev) {
            free_on(sl.dev, sl.p);
            c.held -= sl.bytes;
        } else {
            keep.push_back(sl);
        }
    }
    c.slabs.swap(keep);
}
// Pinned host blocks for the gradient chain's copies back, kept across
// calls (pinning tens of MB costs milliseconds per call): up to
// kPinnedCacheCap bytes, the oldest freed beyond it.  Like the device cache,
// never destroyed; vs_release_cached_memory frees them.
constexpr size_t kPinnedCacheCap = size_t(4) << 30;
struct PinnedCache {
    std::mutex mu;
    std::vector<std::pair<void*, size_t>> blocks;   // oldest first
    size_t held = 0;
};
PinnedCache& pinned_cache() {
    static PinnedCache* c = new PinnedCache;   // intentionally never destroyed
    return *c;
}
// A kept block of bytes to four times bytes (the smallest), else a new one.
hipError_t pinned_take(size_t bytes, void** p, size_t* got) {
    {
        PinnedCache& c = pinned_cache();
        std::lock_guard<std::mutex> g(c.mu);
        size_t best = c.blocks.size();
        for (size_t i = 0; i < c.blocks.size(); ++i)
            if (c.blocks[i].second >= bytes && c.blocks[i].second / 4 <= bytes &&
                (best == c.blocks.size() || c.blocks[i].second < c.blocks[best].second))
                best = i;
        if (best < c.blocks.size()) {
            *p = c.blocks[best].first;
            *got = c.blocks[best].second;
            c.held -= *got;
            c.blocks.erase(c.blocks.begin() + (long)best);
            return hipSuccess;
        }
    }
    *got = bytes;
    return hipHostMalloc(p, bytes, hipHostMallocDefault);
}
void pinned_give(void* p, size_t bytes) {
    if (!p) return;
    PinnedCache& c = pinned_cache();
    std::lock_guard<std::mutex> g(c.mu);
    c.blocks.emplace_back(p, bytes);
    c.held += bytes;
    while (c.held > kPinnedCacheCap) {
        (void)hipHostFree(c.blocks.front().first);
        c.held -= c.blocks.front().second;
        c.blocks.erase(c.blocks.begin());
    }
}
void pinned_release() {
    PinnedCache& c = pinned_cache();
    std::lock_guard<std::mutex> g(c.mu);
    for (const auto& b : c.blocks) (void)hipHostFree(b.first);
    c.blocks.clear();
    c.held = 0;
}
// Returns a pinned block to the cache when the scope ends.
struct PinnedBlock {
    void* p = nullptr;
    size_t bytes = 0;
    ~PinnedBlock() { pinned_give(p, bytes); }
};

// Device memory for a slab: a cached one, else hipMalloc, retried once after
// the device's cached slabs are freed.
hipError_t slab_alloc(int dev, size_t bytes, void** p, size_t* got) {
    if ((*p = slab_take(dev, bytes, got))) return hipSuccess;
    *got = bytes;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess && slab_cached_bytes(dev) > 0) {
        (void)hipGetLastError();
        slab_release(dev);
        e = hipMalloc(p, bytes);
    }
    return e;
}
}  // namespace

struct vs_handle;
namespace {
void release_adj_fast_slabs(vs_handle* h);
}  // namespace

struct vs_handle {
    int n = 0;
    int precision = 64;
    int device = 0;            // the GPU current at vs_create
    std::vector<TubeDesc> desc;
    std::vector<vs_dims> dims;
    int max_steps = 0;
    int steps_done = 0;
    int graph_t0 = 0;          // host source of the graph replays' first-step copy
    long long cells_per_step = 0;
    // options
    int steps_per_launch = 256;
    int threads = 256;
    int use_graph = 0;
    int kernel = VS_K_EXACT;
    int timing = 0;
    // device memory: fields are 256-byte aligned pieces of chunks of
    // chunk_bytes or more (from the slab cache, returned to it by the
    // destructor), so a handle maps a few chunks instead of one allocation
    // per field
    std::vector<std::pair<void*, size_t>> chunks;   // device pointer, bytes
    size_t chunk_used = 0;                          // bytes handed out of chunks.back()
    size_t chunk_bytes = kArenaChunkMax;            // size of a new chunk (set by the build from the batch)
    TubeDesc* d_desc = nullptr;
    int* d_order = nullptr;
    void* bufs = nullptr;   // Bufs<float> or Bufs<double>
    Bufs<float> bf{};
    Bufs<double> bd{};
    long long total_cells = 0, total_cols = 0, total_walls = 0, total_secs = 0, total_src = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool ev_recorded = false;  // the last run recorded ev0 .. ev1 (a run that launches nothing does not)
    double last_ms = 0, last_mean_launch_ms = 0;
    int last_launches = 0;
    std::vector<TwoMassState> tm0;   // initial two-mass states (for vs_reset)
    // FP32 flux-form path
    FastBufs fb{};
    bool fast_ready = false;   // coefficients built
    bool flux_valid = false;   // Uxf/Uyf hold the current state
    bool fast_classes = false; // class bytes and tables built
    int fast_pairs = 0;        // pair classes of the wide-access form (0 until the class bytes are built)
    bool fast_wide_ok = true;  // the batch's pair classes and source columns suit the wide-access form
    std::vector<int> src_cols; // per tube: interior column of its source faces, -1 if none
    int max_interior = 0;      // largest (H-2)*(W-2) of the batch
    int max_padded = 0;        // largest (H-2)*(W-1): interior with one pad column per row
    int max_wp = 0;            // widest padded row W-1
    int max_padded_w[3] = {};  // the same with rows padded to a multiple of K = kFastWideKs[i] (wide on-chip kernel)
    int max_wp_w[3] = {};
    int max_padded4 = 0;       // ... and to a multiple of 4 (its compact rows)
    int max_wp4 = 0;
    int fast_wide = 1;         // VS_OPT_FAST_WIDE
    int lds_limit = 0;
    // materialised export buffers (coefficients, types, z_inv) per tube
    std::vector<void*> coeff_dev, types_dev, zinv_dev;
    std::vector<std::pair<long long, void*>> f32_dev;   // (tube << 32 | field) -> float32 copy
    // captured step blocks
    int graph_blocks = 4;
    int* d_t0 = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    long long graph_key = -1;
    // per-section inputs kept for the gradient chain (tube-major, sec_off)
    std::vector<double> sec_area, sec_mu, sec_scale;
    // reverse-time pass
    AdjBufs<float> af{};
    AdjBufs<double> ad{};
    bool adj_ready = false;    // per-cell buffers allocated
    int adj_S = 0;             // checkpoint spacing of the slabs held
    bool adj_forward = false;  // checkpoints hold a completed forward pass
    std::vector<void*> adj_slabs;
    int adj_kernel = 0;        // VS_OPT_ADJOINT_KERNEL
    int exact_groups = 0;      // VS_OPT_EXACT_GROUPS (0: automatic, coop_plan)
    int exact_steps = 0;       // VS_OPT_EXACT_STEPS (0: automatic, coop_plan)
    int barrier_spin = 0;      // VS_OPT_BARRIER_SPIN (0: kCoopSpinDefault)
    unsigned* d_bar = nullptr; // grid barrier state of the cooperative exact step
    void* d_state2 = nullptr;  // its second state buffers: P, Vx, Vy (Real, total_cells each)
    int coop_ok = -1;          // device supports cooperative launches (-1: not asked yet)
    int adj_block = 0;         // VS_OPT_ADJOINT_BLOCK
    bool adj_onchip = false;   // the held forward pass ran on chip
    // on-chip FP32 reverse pass: lanes shape, slabs and scratch
    int adj_K = 0, adj_G = 0, adj_B = 0, adj_fast_S = 0, adj_sum_groups = 0;
    int adj_fwd_threads = 0;   // work-group size of the checkpointing forward (timed once)
    int adj_fwd_wide = 0;      // the checkpointing forward runs the wide-access form
    FastAdj fa{};
    float* adj_rec[2] = {nullptr, nullptr};   // double-buffered record slabs
    float* adj_adj[2] = {nullptr, nullptr};
    float* adj_start[2] = {nullptr, nullptr};
    hipStream_t adj_rec_stream = nullptr, adj_rev_stream = nullptr, adj_sum_stream = nullptr;
    hipEvent_t adj_ev[7] = {};               // init, rec x2, rev x2, sum x2
    float* adj_ck = nullptr;    // checkpoints: P, Ux, Uy (total_cells each) every S steps
    float* adj_sub = nullptr;   // block starts inside one segment when the record block B < S
    double* adj_strace = nullptr;
    double* adj_pexc = nullptr;
    TwoMassState* adj_tm = nullptr;
    std::vector<std::pair<void*, size_t>> adj_fast_slabs;   // device pointer, bytes
    double adj_reverse_ms = 0, adj_chain_ms = 0;   // last vs_adjoint_backward: device part, host chain

    ~vs_handle() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        // the chunks and slabs go back to the cache once no stream can use them
        for (hipStream_t st : {stream, adj_rec_stream, adj_rev_stream, adj_sum_stream})
            if (st) (void)hipStreamSynchronize(st);
        for (const auto& c : chunks) slab_give(device, c.first, c.second);
        for (void* p : adj_slabs) (void)hipFree(p);
        release_adj_fast_slabs(this);
        for (hipEvent_t e : adj_ev)
            if (e) (void)hipEventDestroy(e);
        if (adj_rec_stream) (void)hipStreamDestroy(adj_rec_stream);
        if (adj_rev_stream) (void)hipStreamDestroy(adj_rev_stream);
        if (adj_sum_stream) (void)hipStreamDestroy(adj_sum_stream);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
    // count zeroed elements of T (on the handle's stream)
    template <typename T>
    int alloc(T** p, size_t count) {
        const size_t bytes = (std::max<size_t>(count, 1) * sizeof(T) + 255) & ~size_t(255);
        if (chunks.empty() || chunks.back().second - chunk_used < bytes) {
            void* c = nullptr;
            size_t got = 0;
            const hipError_t e = slab_alloc(device, std::max(bytes, chunk_bytes), &c, &got);
            if (e != hipSuccess) return fail(VS_E_DEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
            chunks.emplace_back(c, got);
            chunk_used = 0;
        }
        void* q = static_cast<char*>(chunks.back().first) + chunk_used;
        chunk_used += bytes;
        *p = static_cast<T*>(q);
        const hipError_t e = hipMemsetAsync(q, 0, bytes, stream);
        if (e != hipSuccess) return fail(VS_E_DEVICE, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
        return VS_OK;
    }
};

namespace {
// Hands the on-chip pass's slabs of h to the cache once the work of every
// stream that may still use them has finished.
void release_adj_fast_slabs(vs_handle* h) {
    if (h->adj_fast_slabs.empty()) return;
    for (hipStream_t st : {h->stream, h->adj_rec_stream, h->adj_rev_stream, h->adj_sum_stream})
        if (st) (void)hipStreamSynchronize(st);
    for (const auto& sl : h->adj_fast_slabs) slab_give(h->device, sl.first, sl.second);
    h->adj_fast_slabs.clear();
}
}  // namespace

namespace {

long long graph_key(const vs_handle* h) {
    return ((long long)h->fast_wide << 48) ^ ((long long)h->kernel << 40) ^ ((long long)h->threads << 24) ^
           ((long long)h->steps_per_launch << 8) ^ (long long)h->graph_blocks;
}

template <typename Real>
Bufs<Real>& bufs_of(vs_handle* h);
template <>
Bufs<float>& bufs_of<float>(vs_handle* h) { return h->bf; }
template <>
Bufs<double>& bufs_of<double>(vs_handle* h) { return h->bd; }

// Per tube, the distinct (cur, right, top, segment-wall) classes of its
// interior cells (enumerate_exact_classes) and a class byte per cell; the
// class coefficients are computed on the device in the step's own arithmetic.
template <typename Real>
int build_exact_classes(vs_handle* h, Bufs<Real>& b) {
    std::vector<uint8_t> cls((size_t)h->total_cells), xc((size_t)h->total_cells, 0);
    VS_HIP(copy_on(h->stream, cls.data(), b.cls, cls.size(), hipMemcpyDeviceToHost));
    std::vector<int> keys((size_t)h->n * kExactClasses, 0), nk(h->n, 0);
    for (int k = 0; k < h->n; ++k) {
        const TubeDesc& d = h->desc[k];
        std::string err;
        const int rc = enumerate_exact_classes(cls.data() + d.cell_off, d.H, d.W, xc.data() + d.cell_off,
                                               keys.data() + (size_t)k * kExactClasses, nk[k], err);
        if (rc != VS_OK) return fail(rc, "tube " + std::to_string(k) + ": " + err);
    }
    VS_HIP(copy_on(h->stream, b.xcls, xc.data(), xc.size(), hipMemcpyHostToDevice));
    VS_HIP(copy_on(h->stream, b.xkey, keys.data(), keys.size() * sizeof(int), hipMemcpyHostToDevice));
    VS_HIP(copy_on(h->stream, b.xn, nk.data(), nk.size() * sizeof(int), hipMemcpyHostToDevice));
    hipError_t e = launch_exact_classes<Real>(h->d_desc, h->n, b, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail(VS_E_DEVICE, std::string("class coefficients: ") + hipGetErrorString(e));
    return VS_OK;
}

template <typename Real>
int build(vs_handle* h, const vs_tube_spec* specs) {
    const bool fp64 = sizeof(Real) == 8;
    std::vector<Prelude> pre(h->n);
    {
        // the host preludes are independent per tube and run on up to 16
        // host threads; single-plane tubes search their chords on the
        // device, so a batch with one runs them in tube order on this thread
        std::vector<int> rcs(h->n, VS_OK);
        std::vector<std::string> errs(h->n);
        const hipStream_t st = h->stream;
        const MaxChordFn chord = [st](const std::vector<float>& y, const std::vector<float>& z, int sections, int npt,
                                      std::vector<float>& dist, std::vector<int>& far_idx, std::string& e) {
            return max_chord_on_device(st, y, z, sections, npt, dist, far_idx, e);
        };
        bool any_plane = false;
        for (int k = 0; k < h->n; ++k) any_plane = any_plane || specs[k].cross_sectional_shape == 3;
        std::atomic<int> next{0};
        auto worker = [&] {
            for (int k; (k = next.fetch_add(1)) < h->n;) {
                try {
                    rcs[k] = prelude(specs[k], fp64, pre[k], chord, errs[k]);
                } catch (const std::exception& ex) {
                    rcs[k] = VS_E_STATE;
                    errs[k] = ex.what();
                } catch (...) {
                    rcs[k] = VS_E_STATE;
                    errs[k] = "unexpected exception";
                }
            }
        };
        const int nthr = any_plane ? 1 : std::max(1, std::min({h->n, 16, (int)std::thread::hardware_concurrency()}));
        if (nthr == 1) {
            worker();
        } else {
            std::vector<std::thread> pool;
            for (int t = 0; t < nthr; ++t) pool.emplace_back(worker);
            for (std::thread& t : pool) t.join();
        }
        for (int k = 0; k < h->n; ++k)
            if (rcs[k] != VS_OK) return fail(rcs[k], "tube " + std::to_string(k) + ": " + errs[k]);
    }
    h->desc.resize(h->n);
    h->dims.resize(h->n);
    long long cell = 0, col = 0, wall = 0, sec = 0, src = 0;
    for (int k = 0; k < h->n; ++k) {
        const Prelude& p = pre[k];
        const vs_tube_spec& s = specs[k];
        TubeDesc& d = h->desc[k];
        std::memset(&d, 0, sizeof(d));
        const int nsec = s.open_space ? 0 : s.num_sections;
        d.H = p.H; d.W = p.W; d.L = p.L; d.nsec = nsec;
        d.open_space = s.open_space ? 1 : 0;
        d.ts = p.ts; d.te = p.te; d.midY = p.midY;
        d.listener = p.listenerY * p.W + p.listenerX;
        d.excitation = p.excY * p.W + p.excX;
        d.num_steps = p.num_steps;
        d.source_type = s.source_type; d.dirichlet = !s.open_space && s.mouth_radiation_cond == 1;
        d.variant = s.engine_variant; d.srate_mul = s.srate_mul;
        d.shape = s.cross_sectional_shape; d.solver_type = s.fdtd_solver_type;
        d.ds_f32 = p.ds32; d.pml = s.pml_flag ? 1 : 0;
        d.cell_off = cell; d.col_off = col; d.wall_off = wall; d.sec_off = sec; d.src_off = src;
        d.ph = p.ph;
        cell += ((long long)p.H * p.W + 63) / 64 * 64;   // 256-byte aligned tube blocks
        col += p.W; wall += p.L; sec += nsec; src += p.num_steps;
        h->max_steps = std::max(h->max_steps, p.num_steps);
        h->cells_per_step += (long long)(p.H - 2) * (p.W - 2);
        h->max_interior = std::max(h->max_interior, (p.H - 2) * (p.W - 2));
        h->max_padded = std::max(h->max_padded, (p.H - 2) * (p.W - 1));
        h->max_wp = std::max(h->max_wp, p.W - 1);
        for (int i = 0; i < 3; ++i) {
            const int wp = fast_wide_wp(p.W - 2, kFastWideKs[i]);
            h->max_padded_w[i] = std::max(h->max_padded_w[i], (p.H - 2) * wp);
            h->max_wp_w[i] = std::max(h->max_wp_w[i], wp);
        }
        h->max_padded4 = std::max(h->max_padded4, (p.H - 2) * fast_wide_wp(p.W - 2, 4));
        h->max_wp4 = std::max(h->max_wp4, fast_wide_wp(p.W - 2, 4));
        vs_dims& dm = h->dims[k];
        dm.frameH = p.H; dm.frameW = p.W; dm.domainH = p.domainH; dm.domainW = p.domainW;
        dm.listenerX = p.listenerX; dm.listenerY = p.listenerY;
        dm.excitationX = p.excX; dm.excitationY = p.excY; dm.num_steps = p.num_steps;
        dm.tube_start_x = p.ts; dm.tube_end_x = p.te; dm.tube_len_cells = p.L;
        dm.midY = p.midY; dm.num_sections = nsec; dm.mic_cells = p.mic;
    }
    h->total_cells = cell; h->total_cols = col; h->total_walls = wall; h->total_secs = sec; h->total_src = src;
    {
        // chunk size: about the set-up fields below (P, V, depths, classes,
        // columns, walls, sections, per-step source, probe and trace), so a
        // one-tube handle takes a few MB and a large batch 64 MiB chunks
        const size_t R = sizeof(Real);
        const size_t est = (size_t)cell * (6 * R + 10) + (size_t)col * R + (size_t)wall * 12 +
                           (size_t)sec * (8 + 3 * R) + (size_t)src * (R + 16);
        h->chunk_bytes = std::min(kArenaChunkMax, std::max(kArenaChunkMin, est));
    }
    Bufs<Real>& b = bufs_of<Real>(h);
    int rc;
#define ALLOC(ptr, n) if ((rc = h->alloc(&(ptr), (n))) != VS_OK) return rc
    ALLOC(b.P, cell); ALLOC(b.Vx, cell); ALLOC(b.Vy, cell);
    ALLOC(b.dP, cell); ALLOC(b.dX, cell); ALLOC(b.dY, cell);
    ALLOC(b.cls, cell); ALLOC(b.bst, cell);
    ALLOC(b.zcol, col);
    ALLOC(b.upper, wall); ALLOC(b.lower, wall); ALLOC(b.seg, wall);
    ALLOC(b.cum, sec); ALLOC(b.diam, sec); ALLOC(b.minor, sec); ALLOC(b.dscale, sec); ALLOC(b.zsec, sec);
    ALLOC(b.source, src); ALLOC(b.probe, src); ALLOC(b.strace, src);
    ALLOC(b.tm, h->n); ALLOC(b.pexc, h->n); ALLOC(b.status, 1);
    bool any_plane = false;
    for (int k = 0; k < h->n; ++k) any_plane = any_plane || specs[k].cross_sectional_shape == 3;
    const long long npoly = any_plane ? sec * kPlanePoints : 1;
    ALLOC(b.poly_y, npoly); ALLOC(b.poly_z, npoly); ALLOC(b.pbox, any_plane ? sec * 3 : 1);
    ALLOC(b.xcls, cell); ALLOC(b.xkey, (long long)h->n * kExactClasses); ALLOC(b.xn, h->n);
    ALLOC(b.xtab, (long long)h->n * kExactClasses * kExactSlots);
    ALLOC(h->d_desc, h->n); ALLOC(h->d_order, h->n);
    // host staging of the per-section and per-step inputs
    std::vector<float> cum; std::vector<int> diam; std::vector<Real> minor, dsc, zs, srct;
    std::vector<TwoMassState> tms(h->n);
    srct.reserve((size_t)src);
    // tubes whose table source has the same parameters share one table: the
    // first such tube's values are copied (a table is num_steps evaluations
    // of exp or sin on the host)
    struct SrcKey {
        int type, steps;
        double amp, t0, width, dt;
        size_t off;
    };
    std::vector<SrcKey> tables;
    for (int k = 0; k < h->n; ++k) {
        const Prelude& p = pre[k];
        cum.insert(cum.end(), p.cum.begin(), p.cum.end());
        diam.insert(diam.end(), p.diam.begin(), p.diam.end());
        const float mu_def = specs[k].mu3d_default > 0 ? specs[k].mu3d_default
                                                       : (specs[k].fdtd_solver_type == 1 ? 0.025f : 0.005f);
        for (int i = 0; i < h->desc[k].nsec; ++i) {
            minor.push_back((Real)p.minor[i]);
            dsc.push_back((Real)p.dscale[i]);
            zs.push_back((Real)p.zsec[i]);
            h->sec_area.push_back((double)specs[k].area_cm2[i]);
            h->sec_mu.push_back((double)(specs[k].mu3d ? specs[k].mu3d[i] : mu_def));
            h->sec_scale.push_back(p.dscale[i]);
        }
        const vs_tube_spec& s = specs[k];
        const SrcKey key{s.source_type, p.num_steps, s.source_type == 1 ? s.gauss_amp : 0.0,
                         s.source_type == 1 ? s.gauss_t0 : 0.0, s.source_type == 1 ? s.gauss_width : 0.0, p.ph.dt,
                         srct.size()};
        const SrcKey* same = nullptr;
        for (const SrcKey& t : tables)
            if (t.type == key.type && t.steps == key.steps && t.amp == key.amp && t.t0 == key.t0 &&
                t.width == key.width && t.dt == key.dt) {
                same = &t;
                break;
            }
        if (same) {
            // srct has room for every tube's table (reserved above), so
            // appending from its own front does not reallocate
            const size_t off = same->off;
            for (int i = 0; i < p.num_steps; ++i) srct.push_back(srct[off + i]);
        } else {
            std::vector<Real> tab;
            source_table<Real>(s, p.ph, p.num_steps, tab);
            srct.insert(srct.end(), tab.begin(), tab.end());
            tables.push_back(key);
        }
        const double q = specs[k].tm_q > 0 ? specs[k].tm_q : 1.0;
        const double ps = specs[k].tm_ps > 0 ? specs[k].tm_ps : (double)686.46f;
        twomass_init(tms[k], q, ps, specs[k].srate_mul);
    }
    // largest tubes first
    std::vector<int> order(h->n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
        return (long long)h->desc[a].H * h->desc[a].W > (long long)h->desc[c].H * h->desc[c].W;
    });
    h->tm0 = tms;
    hipStream_t st = h->stream;
    VS_HIP(hipMemcpyAsync(b.cum, cum.data(), cum.size() * sizeof(float), hipMemcpyHostToDevice, st));
    VS_HIP(hipMemcpyAsync(b.diam, diam.data(), diam.size() * sizeof(int), hipMemcpyHostToDevice, st));
    VS_HIP(hipMemcpyAsync(b.minor, minor.data(), minor.size() * sizeof(Real), hipMemcpyHostToDevice, st));
    VS_HIP(hipMemcpyAsync(b.dscale, dsc.data(), dsc.size() * sizeof(Real), hipMemcpyHostToDevice, st));
    VS_HIP(hipMemcpyAsync(b.zsec, zs.data(), zs.size() * sizeof(Real), hipMemcpyHostToDevice, st));
    VS_HIP(hipMemcpyAsync(b.source, srct.data(), srct.size() * sizeof(Real), hipMemcpyHostToDevice, st));
    VS_HIP(hipMemcpyAsync(b.tm, tms.data(), tms.size() * sizeof(TwoMassState), hipMemcpyHostToDevice, st));
    if (any_plane) {
        std::vector<float> py((size_t)sec * kPlanePoints, 0.0f), pz((size_t)sec * kPlanePoints, 0.0f),
            box((size_t)sec * 3, 0.0f);
        for (int k = 0; k < h->n; ++k) {
            const long long so = h->desc[k].sec_off;
            for (size_t i = 0; i < pre[k].plane.size(); ++i) {
                const PlaneSection& P = pre[k].plane[i];
                std::copy(P.y.begin(), P.y.end(), py.begin() + (so + i) * kPlanePoints);
                std::copy(P.z.begin(), P.z.end(), pz.begin() + (so + i) * kPlanePoints);
                box[3 * (so + i)] = P.minY;
                box[3 * (so + i) + 1] = P.minZ;
                box[3 * (so + i) + 2] = P.maxZ;
            }
        }
        VS_HIP(copy_on(st, b.poly_y, py.data(), py.size() * sizeof(float), hipMemcpyHostToDevice));
        VS_HIP(copy_on(st, b.poly_z, pz.data(), pz.size() * sizeof(float), hipMemcpyHostToDevice));
        VS_HIP(copy_on(st, b.pbox, box.data(), box.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    VS_HIP(hipMemcpyAsync(h->d_desc, h->desc.data(), h->n * sizeof(TubeDesc), hipMemcpyHostToDevice, st));
    VS_HIP(hipMemcpyAsync(h->d_order, order.data(), h->n * sizeof(int), hipMemcpyHostToDevice, st));
    // setup kernels; depth scratch is one maxHW slice per tube
    long long maxHW = 0;
    for (auto& d : h->desc) maxHW = std::max(maxHW, (long long)d.H * d.W);
    Real* tmp = nullptr;
    void* tmp_raw = nullptr;
    size_t tmp_got = 0;
    VS_HIP(slab_alloc(h->device, sizeof(Real) * (size_t)maxHW * h->n * 4, &tmp_raw, &tmp_got));
    tmp = static_cast<Real*>(tmp_raw);
    const size_t slab = (size_t)maxHW * h->n;
    hipError_t e = launch_setup<Real>(h->d_desc, h->desc.data(), h->n, b, tmp, tmp + slab, tmp + 2 * slab,
                                      tmp + 3 * slab, st);
    hipError_t e2 = hipStreamSynchronize(st);
    slab_give(h->device, tmp_raw, tmp_got);
    if (e != hipSuccess) return fail(VS_E_DEVICE, std::string("setup launch: ") + hipGetErrorString(e));
    if (e2 != hipSuccess) return fail(VS_E_DEVICE, std::string("setup: ") + hipGetErrorString(e2));
    unsigned status = 0;
    VS_HIP(copy_on(st, &status, b.status, sizeof(unsigned), hipMemcpyDeviceToHost));
    if (status & kStatusSqrtDomain) return fail(VS_E_GEOMETRY, "depth sqrt domain error");
    return build_exact_classes<Real>(h, b);
}

int check_handle(const vs_handle* h) {
    if (!h) return fail(VS_E_INVALID, "null handle");
    return VS_OK;
}

// Class bytes and class table of the FP32 flux-form kernel from the face
// codes its setup kernel wrote (enumerate_fast_classes), each tube's damping
// codes read through its own table: 1, 0, the wall damping 1 / (1 + dt) and
// the five PML rings 1 / (1 + sigma_k).  The step kernels read s and r from
// the sign bits of their ax, ay entries.
int build_fast_classes(vs_handle* h) {
    FastBufs& f = h->fb;
    const size_t n = (size_t)h->total_cells;
    // tube k's block of cells (its frame and the alignment gap after it)
    std::vector<size_t> bounds(h->n + 1, n);
    std::vector<float> luts(8 * (size_t)h->n);
    for (int k = 0; k < h->n; ++k) {
        const TubeDesc& d = h->desc[k];
        bounds[k] = (size_t)d.cell_off;
        float* lut = &luts[8 * (size_t)k];
        lut[0] = 1.0f;
        lut[1] = 0.0f;
        lut[2] = (float)(1.0 / (1.0 + d.ph.dt));
        for (int r = 1; r <= 5; ++r) lut[2 + r] = (float)(1.0 / (1.0 + d.ph.sigma[kPml0 + r]));
    }
    std::vector<uint16_t> code(n);
    VS_HIP(hipStreamSynchronize(h->stream));
    VS_HIP(copy_on(h->stream, code.data(), f.code, n * sizeof(uint16_t), hipMemcpyDeviceToHost));
    std::vector<uint8_t> c8(n);
    std::vector<float> tab(5 * kFastClasses, 0.0f);
    int ncls = 0;
    std::string err;
    int rc = enumerate_fast_classes(code.data(), bounds.data(), luts.data(), h->n, c8.data(), tab.data(), ncls, err);
    if (rc != VS_OK) return fail(rc, err);
    VS_HIP(copy_on(h->stream, f.cls8, c8.data(), n, hipMemcpyHostToDevice));
    VS_HIP(copy_on(h->stream, f.ctab, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice));
    // pair classes of the wide-access form (used when there are at most kFastPairs)
    std::vector<int> Hs(h->n), Ws(h->n);
    std::vector<long long> offs(h->n);
    for (int k = 0; k < h->n; ++k) {
        Hs[k] = h->desc[k].H;
        Ws[k] = h->desc[k].W;
        offs[k] = h->desc[k].cell_off;
    }
    std::vector<uint8_t> p8(n, 0);
    std::vector<float> ptab(6 * kFastPairs, 0.0f);
    std::vector<int> src_col(h->n, -1);
    int npairs = 0;
    rc = enumerate_fast_pairs(c8.data(), tab.data(), Hs.data(), Ws.data(), offs.data(), h->n, p8.data(), ptab.data(),
                              npairs, src_col.data(), err);
    if (rc != VS_OK) return fail(rc, err);
    h->fast_pairs = npairs;
    h->fast_wide_ok = npairs <= kFastPairs &&
                      std::none_of(src_col.begin(), src_col.end(), [](int c) { return c == -2; });
    h->src_cols = src_col;
    if (h->fast_wide_ok) {
        VS_HIP(copy_on(h->stream, f.pcls8, p8.data(), n, hipMemcpyHostToDevice));
        VS_HIP(copy_on(h->stream, f.ptab, ptab.data(), ptab.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    return VS_OK;
}

// Launch shape of the wide-access on-chip kernel (wide_shape).
struct WideShape {
    int K = 0;    // cells per lane
    int GW = 0;   // zero guard floats (the widest padded row)
    int q = 0;    // rows padded to a multiple of q (K, or 4 for compact rows)
    int ms = -2;  // lane-run group of the batch's source column (-1: none, -2: every group)
};

// The lane-run group (4 cells) holding the source column of every tube of
// the batch in the wide-access form at K cells per lane (rows padded to a
// multiple of K): -1 when no tube has a source face, -2 when the tubes'
// groups differ (the kernel then handles the source in every group).
int wide_source_group(const vs_handle* h, int K) {
    int ms = -1;
    for (int c : h->src_cols) {
        if (c < 0) continue;
        const int g = (c % K) / 4;
        if (ms == -1) ms = g;
        else if (ms != g) return -2;
    }
    return ms;
}

// Shape of the wide-access on-chip kernel for the batch at `threads`
// threads: the smallest K of kFastWideKs whose rows fit, padded to a
// multiple of K (q = K: a column sits at one slot of its lanes' runs in
// every row, so the source is handled in one group, ms) or, when those rows
// do not fit, to a multiple of 4 (q = 4, ms = -2: every group handles it);
// the guard GW is in floats.  False when the batch does not fit
// (VS_OPT_FAST_WIDE off, K > kFastWideMaxK, LDS beyond the device limit, or
// pair classes / source columns the kernel cannot take).
bool wide_shape(const vs_handle* h, int threads, WideShape& w) {
    if (!h->fast_wide || h->precision != VS_FP32 || threads < 128) return false;
    // the pair tables hold kFastPairs pair classes, and each tube's source
    // faces lie in one column (known once the class bytes are built)
    if (h->fast_classes && !h->fast_wide_ok) return false;
    const int T = threads - 64;
    for (int i = 0; i < 3; ++i) {
        const int K = kFastWideKs[i];
        WideShape c;
        if (h->max_padded_w[i] <= K * T) c = WideShape{K, h->max_wp_w[i], K, wide_source_group(h, K)};
        else if (h->max_padded4 <= K * T) c = WideShape{K, h->max_wp4, 4, -2};
        else continue;
        if ((long long)(fast_wide_lds_bytes(c.K, threads, c.GW) + fast_wide_lds_static_bytes()) > h->lds_limit)
            continue;
        w = c;
        return true;
    }
    return false;
}

// Whether the batch fits the wide-access on-chip kernel at `threads` threads.
bool fast_wide_fits(const vs_handle* h, int threads, int* K_out, int* GW_out) {
    WideShape w;
    if (!wide_shape(h, threads, w)) return false;
    if (K_out) *K_out = w.K;
    if (GW_out) *GW_out = w.GW;
    return true;
}

// Whether the batch fits the on-chip kernels at `threads` threads per tube
// (cells per lane K <= kFastMaxK, LDS within the device limit).
bool fast_fits(const vs_handle* h, int threads, int* K_out, int* G_out) {
    if (h->precision != VS_FP32 || threads < 128) return false;
    const int K = fast_cells_per_lane(h->max_padded, threads);
    if (K > kFastMaxK) return false;
    const int G = fast_guard_lanes(K, h->max_wp);
    if ((long long)(fast_lds_bytes(K, threads, G) + fast_lds_static_bytes()) > h->lds_limit) return false;
    if (K_out) *K_out = K;
    if (G_out) *G_out = G;
    return true;
}

// Lazily builds the per-cell coefficients of the FP32 flux-form kernel.
int prepare_fast(vs_handle* h, int threads) {
    if (h->precision != VS_FP32) return fail(VS_E_INVALID, "the flux-form kernel is FP32 only");
    if (threads < 128) return fail(VS_E_INVALID, "the on-chip kernel needs at least 128 threads");
    const int K = fast_cells_per_lane(h->max_padded, threads);
    const int G = fast_guard_lanes(K, h->max_wp);
    const bool odd_fits = K <= kFastMaxK && (long long)(fast_lds_bytes(K, threads, G) + fast_lds_static_bytes()) <= h->lds_limit;
    // before the class bytes exist, fast_wide_fits judges the size only
    if (!odd_fits && !fast_wide_fits(h, threads, nullptr, nullptr)) {
        if (K > kFastMaxK) return fail(VS_E_INVALID, "tube too large for the on-chip kernel at this work-group size");
        return fail(VS_E_INVALID, "on-chip kernel needs " + std::to_string(fast_lds_bytes(K, threads, G)) +
                    " bytes of LDS, device allows " + std::to_string(h->lds_limit));
    }
    if (!h->fast_ready) {
        FastBufs& f = h->fb;
        const long long n = h->total_cells;
        int rc;
        if ((rc = h->alloc(&f.kP, n)) || (rc = h->alloc(&f.gX, n)) || (rc = h->alloc(&f.gY, n)) ||
            (rc = h->alloc(&f.Uxf, n)) || (rc = h->alloc(&f.Uyf, n)) || (rc = h->alloc(&f.code, n)) ||
            (rc = h->alloc(&f.cls8, n)) || (rc = h->alloc(&f.ctab, 5 * kFastClasses)) ||
            (rc = h->alloc(&f.pcls8, n)) || (rc = h->alloc(&f.ptab, 6 * kFastPairs)))
            return rc;
        h->fast_ready = true;
    }
    if (!h->flux_valid) {
        hipError_t e = launch_fast_setup(h->d_desc, h->desc.data(), h->n, h->bf, h->fb, h->stream);
        if (e != hipSuccess) return fail(VS_E_DEVICE, std::string("flux setup: ") + hipGetErrorString(e));
        if (!h->fast_classes) {
            const int rc = build_fast_classes(h);
            if (rc != VS_OK) return rc;
            h->fast_classes = true;
        }
        h->flux_valid = true;
    }
    // the wide-access form also needs the batch's pair classes to fit its
    // tables and each tube's source faces in one column
    WideShape ws;
    const bool wide = wide_shape(h, threads, ws);
    if (!wide && !odd_fits)
        return fail(VS_E_INVALID, "batch has " + std::to_string(h->fast_pairs) + " pair classes (the wide-access form holds " +
                    std::to_string(kFastPairs) + ", with each tube's source faces in one column) and is too large "
                    "for the odd-run form at this work-group size");
    hipError_t e = wide ? configure_fast_wide(ws.K, ws.ms) : configure_fast(K);
    if (e != hipSuccess) return fail(VS_E_DEVICE, std::string("configure: ") + hipGetErrorString(e));
    return VS_OK;
}

template <typename Real>
hipError_t launch_any(vs_handle* h, Bufs<Real>& b, int t0, int ns, const int* d_t0 = nullptr);
// The exact step of a single tube over several work-groups of one
// cooperative launch, several steps per grid barrier; graph replays keep the
// one-work-group kernel (the results are bit-identical either way).
struct CoopPlan {
    int groups = 1, steps = 1;
    bool one = false;   // one cell per lane (coop_one_cell_per_lane)
    size_t tile = 0;
};
// Work-groups and steps per barrier for this handle (groups 1: none).  By
// default: the most steps per barrier up to kCoopSteps for which at most
// kCoopAutoGroups work-groups hold their first pressure regions (own
// columns and 2s - 1 more) at one cell per lane (coop_one_cell_per_lane),
// with the fewest such work-groups; tubes of at most one cell per lane stay
// on one work-group.  VS_OPT_EXACT_GROUPS / VS_OPT_EXACT_STEPS override
// either; with the groups given and the steps not, the steps are the most
// (up to kCoopSteps) that keep one cell per lane, else 1.  Steps are then
// lowered until the LDS tile fits and every source work-group sees the
// excitation column (coop_source_ok).
constexpr int kCoopSteps = 8;
constexpr int kCoopAutoGroups = 48;
constexpr int kCoopThreads = 1024;   // work-group size of the cooperative step (VS_OPT_THREADS is for the others)
template <typename Real>
CoopPlan coop_plan(const vs_handle* h) {
    CoopPlan p;
    const TubeDesc& T = h->desc[0];
    const int H = T.H, W = T.W, Hi = H - 2, Wi = W - 2, ex = T.excitation % W;
    if (Hi < 1 || Wi < 2 || h->exact_groups == 1) return p;
    int s = h->exact_steps > 0 ? h->exact_steps : kCoopSteps;
    int G = h->exact_groups;
    if (G == 0) {
        if ((long long)Hi * Wi <= kCoopThreads) return p;
        for (int ss = s; ss >= 1 && G == 0; --ss) {
            for (int g = 2; g <= kCoopAutoGroups; ++g) {
                if (coop_one_cell_per_lane(H, W, g, ss, kCoopThreads)) { G = g; s = ss; break; }
            }
        }
        if (G == 0) { G = kCoopAutoGroups; s = 1; }
    }
    G = std::min(G, Wi);
    G = (Wi + coop_chunk(W, G) - 1) / coop_chunk(W, G);   // no empty work-groups
    if (G < 2) return p;
    if (h->exact_groups > 0 && h->exact_steps == 0) {
        // groups given, steps not: the most steps that keep one cell per lane
        s = 1;
        for (int ss = kCoopSteps; ss > 1; --ss)
            if (coop_one_cell_per_lane(H, W, G, ss, kCoopThreads)) { s = ss; break; }
    }
    while (s > 1 && (!coop_source_ok(W, G, s, ex) || coop_tile_bytes<Real>(H, W, G, s) > kCoopTileMax)) --s;
    if (coop_tile_bytes<Real>(H, W, G, s) > kCoopTileMax) return p;
    p.groups = G;
    p.steps = s;
    p.one = coop_one_cell_per_lane(H, W, G, s, kCoopThreads);
    p.tile = coop_tile_bytes<Real>(H, W, G, s);
    return p;
}
template <typename Real>
bool use_coop(vs_handle* h, const int* d_t0) {
    if (h->n != 1 || d_t0 || coop_plan<Real>(h).groups < 2) return false;
    if (h->coop_ok < 0) {
        int dev = 0, v = 0;
        h->coop_ok = (hipGetDevice(&dev) == hipSuccess &&
                      hipDeviceGetAttribute(&v, hipDeviceAttributeCooperativeLaunch, dev) == hipSuccess && v) ? 1 : 0;
    }
    return h->coop_ok == 1;
}
template <typename Real>
hipError_t launch_exact(vs_handle* h, Bufs<Real>& b, int t0, int ns, const int* d_t0) {
    if (use_coop<Real>(h, d_t0)) {
        if (!h->d_bar) {
            Real* st2 = nullptr;
            if (h->alloc(&h->d_bar, 2) != VS_OK || h->alloc(&st2, 3 * (size_t)h->total_cells) != VS_OK)
                return hipErrorOutOfMemory;
            h->d_state2 = st2;
        }
        Real* st2 = static_cast<Real*>(h->d_state2);
        const long long C = h->total_cells;
        const CoopPlan p = coop_plan<Real>(h);
        hipError_t e = launch_step_coop<Real>(h->d_desc, h->d_order, b, t0, ns, kCoopThreads, p.groups, p.steps, p.one,
                                              p.tile, h->d_bar, st2, st2 + C, st2 + 2 * C,
                                              h->barrier_spin > 0 ? h->barrier_spin : kCoopSpinDefault, h->stream);
        if (e != hipErrorCooperativeLaunchTooLarge) return e;
        // the device cannot hold the groups at once: one work-group from now on
        (void)hipGetLastError();
        h->coop_ok = 0;
    }
    return launch_step<Real>(h->d_desc, h->d_order, h->n, b, t0, ns, h->threads, h->kernel, d_t0, h->stream);
}
template <>
hipError_t launch_any<double>(vs_handle* h, Bufs<double>& b, int t0, int ns, const int* d_t0) {
    return launch_exact<double>(h, b, t0, ns, d_t0);
}
template <>
hipError_t launch_any<float>(vs_handle* h, Bufs<float>& b, int t0, int ns, const int* d_t0) {
    if (h->kernel == VS_K_FAST) {
        WideShape w;
        if (wide_shape(h, h->threads, w))
            return launch_fast_wide(w.K, w.ms, h->d_desc, h->d_order, h->n, b, h->fb, t0, ns, h->threads, w.GW, w.q,
                                    d_t0, h->stream);
        const int K = fast_cells_per_lane(h->max_padded, h->threads);
        const int G = fast_guard_lanes(K, h->max_wp);
        return launch_fast_step(K, h->d_desc, h->d_order, h->n, b, h->fb, t0, ns, h->threads, G, d_t0, h->stream);
    }
    h->flux_valid = false;
    return launch_exact<float>(h, b, t0, ns, d_t0);
}

// Captures G step blocks (step kernel + counter bump) into one graph.
template <typename Real>
int build_graph(vs_handle* h, Bufs<Real>& b) {
    if (h->graph_exec) {
        (void)hipGraphExecDestroy(h->graph_exec);
        h->graph_exec = nullptr;
    }
    if (!h->d_t0) {
        int rc = h->alloc(&h->d_t0, 1);
        if (rc != VS_OK) return rc;
    }
    hipGraph_t g = nullptr;
    VS_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int k = 0; k < h->graph_blocks && e == hipSuccess; ++k) {
        e = launch_any<Real>(h, b, 0, h->steps_per_launch, h->d_t0);
        if (e == hipSuccess) e = launch_bump(h->d_t0, h->steps_per_launch, h->stream);
    }
    hipError_t e2 = hipStreamEndCapture(h->stream, &g);
    if (e != hipSuccess) return fail(VS_E_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
    if (e2 != hipSuccess) return fail(VS_E_DEVICE, std::string("graph capture end: ") + hipGetErrorString(e2));
    e = hipGraphInstantiate(&h->graph_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(VS_E_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(e));
    h->graph_key = graph_key(h);
    return VS_OK;
}

template <typename Real>
int run_steps(vs_handle* h, int n_steps) {
    Bufs<Real>& b = bufs_of<Real>(h);
    if (h->kernel == VS_K_FAST) {
        const int rc = prepare_fast(h, h->threads);
        if (rc != VS_OK) return rc;
    }
    bool graph_ran = false;
    h->ev_recorded = false;
    h->last_launches = 0;
    if (h->use_graph && !h->timing) {
        // whole graph replays (graph_blocks x steps_per_launch steps each), then
        // the remainder with ordinary launches
        const int per = h->graph_blocks * h->steps_per_launch;
        const int start = h->steps_done;
        const int end = std::min(h->max_steps, start + n_steps);
        const int reps = (end - start) / per;
        if (reps > 0) {
            if (!h->graph_exec || h->graph_key != graph_key(h)) {
                const int rc = build_graph<Real>(h, b);
                if (rc != VS_OK) return rc;
            }
            // the first step of the replays, staged in a host field that nothing
            // touches until the copy has run (the synchronize below)
            h->graph_t0 = start;
            VS_HIP(hipMemcpyAsync(h->d_t0, &h->graph_t0, sizeof(int), hipMemcpyHostToDevice, h->stream));
            VS_HIP(hipEventRecord(h->ev0, h->stream));
            for (int k = 0; k < reps; ++k) VS_HIP(hipGraphLaunch(h->graph_exec, h->stream));
            VS_HIP(hipEventRecord(h->ev1, h->stream));
            h->ev_recorded = true;
            if (h->kernel != VS_K_FAST) h->flux_valid = false;
            h->last_launches = reps * h->graph_blocks;
            graph_ran = true;
            VS_HIP(hipStreamSynchronize(h->stream));
            h->steps_done = start + reps * per;
        }
        if (end - h->steps_done <= 0) return VS_OK;
        n_steps = end - h->steps_done;
    }
    const int start = h->steps_done;
    const int end = std::min(h->max_steps, start + n_steps);
    if (end <= start) return VS_OK;   // nothing to run (no graph replay ran either)
    int launches = 0;
    double launch_ms = 0;
    std::vector<hipEvent_t> evs;
    if (!graph_ran) VS_HIP(hipEventRecord(h->ev0, h->stream));
    for (int t0 = start; t0 < end; t0 += h->steps_per_launch) {
        const int ns = std::min(h->steps_per_launch, end - t0);
        hipEvent_t a = nullptr, c = nullptr;
        if (h->timing) {
            VS_HIP(hipEventCreate(&a)); VS_HIP(hipEventCreate(&c));
            VS_HIP(hipEventRecord(a, h->stream));
        }
        hipError_t e = launch_any<Real>(h, b, t0, ns, nullptr);
        if (e != hipSuccess) return fail(VS_E_DEVICE, std::string("step launch: ") + hipGetErrorString(e));
        if (h->timing) { VS_HIP(hipEventRecord(c, h->stream)); evs.push_back(a); evs.push_back(c); }
        ++launches;
    }
    VS_HIP(hipEventRecord(h->ev1, h->stream));
    h->ev_recorded = true;
    h->steps_done = end;
    h->last_launches = launches + (graph_ran ? h->last_launches : 0);
    if (h->timing) {
        VS_HIP(hipEventSynchronize(h->ev1));
        for (size_t i = 0; i < evs.size(); i += 2) {
            float ms = 0;
            VS_HIP(hipEventElapsedTime(&ms, evs[i], evs[i + 1]));
            launch_ms += ms;
            (void)hipEventDestroy(evs[i]); (void)hipEventDestroy(evs[i + 1]);
        }
        h->last_mean_launch_ms = launches ? launch_ms / launches : 0;
    }
    return VS_OK;
}

size_t elem_size(const vs_handle* h) { return h->precision == 64 ? 8 : 4; }

// Field location: device pointer, element count and element size.
int locate(vs_handle* h, int tube, int field, void** ptr, size_t* count, size_t* esz) {
    const TubeDesc& d = h->desc[tube];
    const size_t HW = (size_t)d.H * d.W;
    const bool f64 = h->precision == 64;
    auto R = [&](void* base_f, void* base_d, long long off, size_t n) {
        *esz = f64 ? 8 : 4;
        *ptr = f64 ? (void*)((double*)base_d + off) : (void*)((float*)base_f + off);
        *count = n;
        return VS_OK;
    };
    switch (field) {
    case VS_F_CELL_TYPE: return fail(VS_E_INVALID, "cell types are derived on copy");
    case VS_F_BOUNDARY_SEG:
        *ptr = (f64 ? h->bd.bst : h->bf.bst) + d.cell_off; *count = HW; *esz = 4; return VS_OK;
    case VS_F_CELL_CLASS:
        *ptr = (f64 ? h->bd.cls : h->bf.cls) + d.cell_off; *count = HW; *esz = 1; return VS_OK;
    case VS_F_DEPTH_P: return R(h->bf.dP, h->bd.dP, d.cell_off, HW);
    case VS_F_DEPTH_X: return R(h->bf.dX, h->bd.dX, d.cell_off, HW);
    case VS_F_DEPTH_Y: return R(h->bf.dY, h->bd.dY, d.cell_off, HW);
    case VS_F_PRESSURE: return R(h->bf.P, h->bd.P, d.cell_off, HW);
    case VS_F_VX: return R(h->bf.Vx, h->bd.Vx, d.cell_off, HW);
    case VS_F_VY: return R(h->bf.Vy, h->bd.Vy, d.cell_off, HW);
    case VS_F_SOURCE: return R(h->bf.source, h->bd.source, d.src_off, d.num_steps);
    case VS_F_UPPER_WALL:
        *ptr = (f64 ? h->bd.upper : h->bf.upper) + d.wall_off; *count = d.L; *esz = 4; return VS_OK;
    case VS_F_LOWER_WALL:
        *ptr = (f64 ? h->bd.lower : h->bf.lower) + d.wall_off; *count = d.L; *esz = 4; return VS_OK;
    case VS_F_SEGMENT:
        *ptr = (f64 ? h->bd.seg : h->bf.seg) + d.wall_off; *count = d.L; *esz = 4; return VS_OK;
    default: break;
    }
    return fail(VS_E_INVALID, "unknown field id " + std::to_string(field));
}

}  // namespace

extern "C" {

const char* vs_last_error(void) { return g_last_error.c_str(); }
int vs_abi_version(void) { return VS_ABI_VERSION; }

int vs_create(const vs_tube_spec* specs, int n_tubes, int precision, vs_handle** out) {
    return guarded("vs_create", [&]() -> int {
        g_last_error.clear();
        if (!specs || !out || n_tubes <= 0) return fail(VS_E_INVALID, "vs_create: need specs, n_tubes > 0 and an output handle");
        if (precision != VS_FP32 && precision != VS_FP64) return fail(VS_E_INVALID, "precision must be 32 or 64");
        *out = nullptr;
        std::unique_ptr<vs_handle> h(new vs_handle());
        h->n = n_tubes;
        h->precision = precision;
        VS_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        VS_HIP(hipEventCreate(&h->ev0));
        VS_HIP(hipEventCreate(&h->ev1));
        {
            int dev = 0;
            VS_HIP(hipGetDevice(&dev));
            h->device = dev;
            hipDeviceProp_t prop;
            VS_HIP(hipGetDeviceProperties(&prop, dev));
            h->lds_limit = (int)std::max(prop.sharedMemPerBlockOptin, prop.sharedMemPerBlock);
        }
        const int rc = precision == VS_FP64 ? build<double>(h.get(), specs) : build<float>(h.get(), specs);
        if (rc != VS_OK) return rc;
        *out = h.release();
        return VS_OK;
    });
}

int vs_destroy(vs_handle* h) {
    return guarded("vs_destroy", [&]() -> int {
        if (!h) return VS_OK;
        (void)hipStreamSynchronize(h->stream);
        delete h;
        return VS_OK;
    });
}

int vs_release_cached_memory(void) {
    return guarded("vs_release_cached_memory", [&]() -> int {
        slab_release(-1);
        pinned_release();
        return VS_OK;
    });
}

int vs_num_tubes(const vs_handle* h, int* n) {
    return guarded("vs_num_tubes", [&]() -> int {
        if (check_handle(h) || !n) return fail(VS_E_INVALID, "vs_num_tubes: bad argument");
        *n = h->n;
        return VS_OK;
    });
}

int vs_get_dims(const vs_handle* h, int tube, vs_dims* out) {
    return guarded("vs_get_dims", [&]() -> int {
        if (check_handle(h) || !out) return fail(VS_E_INVALID, "vs_get_dims: bad argument");
        if (tube < 0 || tube >= h->n) return fail(VS_E_RANGE, "tube index out of range");
        *out = h->dims[tube];
        return VS_OK;
    });
}

int vs_field_ptr(vs_handle* h, int tube, int field, void** dptr, size_t* bytes) {
    return guarded("vs_field_ptr", [&]() -> int {
        if (check_handle(h) || !dptr) return fail(VS_E_INVALID, "vs_field_ptr: bad argument");
        if (tube < 0 || tube >= h->n) return fail(VS_E_RANGE, "tube index out of range");
        if (field & VS_F_AS_F32) {
            field &= ~VS_F_AS_F32;
            const bool real_field = field == VS_F_DEPTH_P || field == VS_F_DEPTH_X || field == VS_F_DEPTH_Y ||
                                    field == VS_F_Z_INV || field == VS_F_SOURCE || field == VS_F_PRESSURE ||
                                    field == VS_F_VX || field == VS_F_VY ||
                                    (field >= VS_F_COEFF0 && field < VS_F_COEFF0 + 17);
            if (!real_field) return fail(VS_E_INVALID, "VS_F_AS_F32 applies to floating-point fields only");
            if (h->precision == 64) {
                const long long key = ((long long)tube << 32) | field;
                for (auto& kv : h->f32_dev)
                    if (kv.first == key) {
                        *dptr = kv.second;
                        return VS_OK;
                    }
                void* p64 = nullptr;
                size_t nb = 0;
                int rc = vs_field_ptr(h, tube, field, &p64, &nb);
                if (rc != VS_OK) return rc;
                const size_t n = nb / sizeof(double);
                std::vector<double> host(n);
                VS_HIP(hipStreamSynchronize(h->stream));
                VS_HIP(copy_on(h->stream, host.data(), p64, nb, hipMemcpyDeviceToHost));
                std::vector<float> f(host.begin(), host.end());
                float* p = nullptr;
                if ((rc = h->alloc(&p, n)) != VS_OK) return rc;
                VS_HIP(copy_on(h->stream, p, f.data(), n * sizeof(float), hipMemcpyHostToDevice));
                h->f32_dev.push_back({key, p});
                *dptr = p;
                if (bytes) *bytes = n * sizeof(float);
                return VS_OK;
            }
        }
        const TubeDesc& d = h->desc[tube];
        const size_t es = elem_size(h), HW = (size_t)d.H * d.W, ni = (size_t)(d.H - 2) * (d.W - 2);
        if (h->coeff_dev.empty()) {
            h->coeff_dev.assign(h->n, nullptr);
            h->types_dev.assign(h->n, nullptr);
            h->zinv_dev.assign(h->n, nullptr);
        }
        if (field >= VS_F_COEFF0 && field < VS_F_COEFF0 + 17) {
            if (!h->coeff_dev[tube]) {
                void* p = nullptr;
                int rc = h->alloc((char**)&p, ni * 17 * es);
                if (rc != VS_OK) return rc;
                hipError_t e = h->precision == 64
                    ? launch_coeffs<double>(h->d_desc, &d, tube, h->bd, (double*)p, h->stream)
                    : launch_coeffs<float>(h->d_desc, &d, tube, h->bf, (float*)p, h->stream);
                if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
                if (e != hipSuccess) return fail(VS_E_DEVICE, std::string("coefficients: ") + hipGetErrorString(e));
                h->coeff_dev[tube] = p;
            }
            *dptr = (char*)h->coeff_dev[tube] + (size_t)(field - VS_F_COEFF0) * ni * es;
            if (bytes) *bytes = ni * es;
            return VS_OK;
        }
        if (field == VS_F_CELL_TYPE || field == VS_F_Z_INV) {
            std::vector<void*>& slot = field == VS_F_CELL_TYPE ? h->types_dev : h->zinv_dev;
            const size_t nb = HW * (field == VS_F_CELL_TYPE ? sizeof(int) : es);
            if (!slot[tube]) {
                std::vector<char> host(nb);
                int rc = vs_get_field(h, tube, field, host.data(), nb);
                if (rc != VS_OK) return rc;
                void* p = nullptr;
                if ((rc = h->alloc((char**)&p, nb)) != VS_OK) return rc;
                VS_HIP(copy_on(h->stream, p, host.data(), nb, hipMemcpyHostToDevice));
                slot[tube] = p;
            }
            *dptr = slot[tube];
            if (bytes) *bytes = nb;
            return VS_OK;
        }
        size_t count = 0, esz = 0;
        int rc = locate(h, tube, field, dptr, &count, &esz);
        if (rc != VS_OK) return rc;
        if (bytes) *bytes = count * esz;
        return VS_OK;
    });
}

int vs_get_field(vs_handle* h, int tube, int field, void* dst, size_t bytes) {
    return guarded("vs_get_field", [&]() -> int {
        if (check_handle(h) || !dst) return fail(VS_E_INVALID, "vs_get_field: bad argument");
        if (tube < 0 || tube >= h->n) return fail(VS_E_RANGE, "tube index out of range");
        const TubeDesc& d = h->desc[tube];
        const size_t HW = (size_t)d.H * d.W;
        if (field == VS_F_CELL_TYPE) {
            if (bytes < HW * sizeof(int)) return fail(VS_E_RANGE, "destination too small");
            std::vector<uint8_t> cls(HW);
            VS_HIP(hipStreamSynchronize(h->stream));
            const uint8_t* src = (h->precision == 64 ? h->bd.cls : h->bf.cls) + d.cell_off;
            VS_HIP(copy_on(h->stream, cls.data(), src, HW, hipMemcpyDeviceToHost));
            int* o = static_cast<int*>(dst);
            for (size_t i = 0; i < HW; ++i) o[i] = cls[i] & kTypeMask;
            return VS_OK;
        }
        if (field == VS_F_Z_INV) {
            const size_t es = elem_size(h);
            if (bytes < HW * es) return fail(VS_E_RANGE, "destination too small");
            std::vector<uint8_t> cls(HW);
            VS_HIP(hipStreamSynchronize(h->stream));
            const uint8_t* src = (h->precision == 64 ? h->bd.cls : h->bf.cls) + d.cell_off;
            VS_HIP(copy_on(h->stream, cls.data(), src, HW, hipMemcpyDeviceToHost));
            std::vector<double> zc(d.W);
            if (h->precision == 64) {
                VS_HIP(copy_on(h->stream, zc.data(), h->bd.zcol + d.col_off, d.W * 8, hipMemcpyDeviceToHost));
            } else {
                std::vector<float> zf(d.W);
                VS_HIP(copy_on(h->stream, zf.data(), h->bf.zcol + d.col_off, d.W * 4, hipMemcpyDeviceToHost));
                for (int c = 0; c < d.W; ++c) zc[c] = zf[c];
            }
            for (size_t i = 0; i < HW; ++i) {
                const double z = (cls[i] & kSegWallBit) ? zc[i % d.W] : 0.0;
                if (h->precision == 64) static_cast<double*>(dst)[i] = z;
                else static_cast<float*>(dst)[i] = (float)z;
            }
            return VS_OK;
        }
        if (field >= VS_F_COEFF0 && field < VS_F_COEFF0 + 17) {
            const size_t ni = (size_t)(d.H - 2) * (d.W - 2);
            const size_t es = elem_size(h);
            if (bytes < ni * es) return fail(VS_E_RANGE, "destination too small");
            void* tmp = nullptr;
            VS_HIP(hipMalloc(&tmp, ni * 17 * es));
            hipError_t e = h->precision == 64
                ? launch_coeffs<double>(h->d_desc, &d, tube, h->bd, (double*)tmp, h->stream)
                : launch_coeffs<float>(h->d_desc, &d, tube, h->bf, (float*)tmp, h->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
            if (e == hipSuccess)
                e = copy_on(h->stream, dst, (char*)tmp + (size_t)(field - VS_F_COEFF0) * ni * es, ni * es, hipMemcpyDeviceToHost);
            (void)hipFree(tmp);
            if (e != hipSuccess) return fail(VS_E_DEVICE, std::string("coefficients: ") + hipGetErrorString(e));
            return VS_OK;
        }
        void* p = nullptr;
        size_t count = 0, esz = 0;
        int rc = locate(h, tube, field, &p, &count, &esz);
        if (rc != VS_OK) return rc;
        if (bytes < count * esz) return fail(VS_E_RANGE, "destination too small");
        VS_HIP(hipStreamSynchronize(h->stream));
        VS_HIP(copy_on(h->stream, dst, p, count * esz, hipMemcpyDeviceToHost));
        return VS_OK;
    });
}

int vs_run(vs_handle* h, int n_steps) {
    return guarded("vs_run", [&]() -> int {
        if (check_handle(h)) return VS_E_INVALID;
        if (n_steps < 0) return fail(VS_E_INVALID, "n_steps must be >= 0");
        return h->precision == 64 ? run_steps<double>(h, n_steps) : run_steps<float>(h, n_steps);
    });
}

int vs_sync(vs_handle* h) {
    return guarded("vs_sync", [&]() -> int {
        if (check_handle(h)) return VS_E_INVALID;
        VS_HIP(hipStreamSynchronize(h->stream));
        if (h->d_bar) {
            unsigned* st = h->precision == 64 ? h->bd.status : h->bf.status;
            unsigned status = 0;
            VS_HIP(copy_on(h->stream, &status, st, sizeof(unsigned), hipMemcpyDeviceToHost));
            if (status & kStatusBarrierTimeout) {
                const unsigned clear = status & ~kStatusBarrierTimeout;
                VS_HIP(copy_on(h->stream, st, &clear, sizeof(unsigned), hipMemcpyHostToDevice));
                return fail(VS_E_DEVICE, "cooperative exact step: a grid barrier timed out (results invalid)");
            }
        }
        float ms = 0;
        if (h->ev_recorded) VS_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        h->last_ms = ms;
        return VS_OK;
    });
}

int vs_steps_done(const vs_handle* h, int* steps) {
    return guarded("vs_steps_done", [&]() -> int {
        if (check_handle(h) || !steps) return fail(VS_E_INVALID, "vs_steps_done: bad argument");
        *steps = h->steps_done;
        return VS_OK;
    });
}

int vs_reset(vs_handle* h) {
    return guarded("vs_reset", [&]() -> int {
        if (check_handle(h)) return VS_E_INVALID;
        const size_t es = elem_size(h);
        void* P = h->precision == 64 ? (void*)h->bd.P : (void*)h->bf.P;
        void* Vx = h->precision == 64 ? (void*)h->bd.Vx : (void*)h->bf.Vx;
        void* Vy = h->precision == 64 ? (void*)h->bd.Vy : (void*)h->bf.Vy;
        VS_HIP(hipMemsetAsync(P, 0, h->total_cells * es, h->stream));
        VS_HIP(hipMemsetAsync(Vx, 0, h->total_cells * es, h->stream));
        VS_HIP(hipMemsetAsync(Vy, 0, h->total_cells * es, h->stream));
        double* pexc = h->precision == 64 ? h->bd.pexc : h->bf.pexc;
        VS_HIP(hipMemsetAsync(pexc, 0, h->n * sizeof(double), h->stream));
        h->flux_valid = false;
        TwoMassState* tm = h->precision == 64 ? h->bd.tm : h->bf.tm;
        VS_HIP(hipMemcpyAsync(tm, h->tm0.data(), h->n * sizeof(TwoMassState), hipMemcpyHostToDevice, h->stream));
        VS_HIP(hipStreamSynchronize(h->stream));
        h->steps_done = 0;
        return VS_OK;
    });
}

int vs_get_probe(vs_handle* h, int tube, double* dst, size_t n) {
    return guarded("vs_get_probe", [&]() -> int {
        if (check_handle(h) || !dst) return fail(VS_E_INVALID, "vs_get_probe: bad argument");
        if (tube < 0 || tube >= h->n) return fail(VS_E_RANGE, "tube index out of range");
        const TubeDesc& d = h->desc[tube];
        const size_t m = std::min<size_t>(n, (size_t)d.num_steps);
        double* p = (h->precision == 64 ? h->bd.probe : h->bf.probe) + d.src_off;
        VS_HIP(hipStreamSynchronize(h->stream));
        VS_HIP(copy_on(h->stream, dst, p, m * sizeof(double), hipMemcpyDeviceToHost));
        return VS_OK;
    });
}

int vs_get_source_trace(vs_handle* h, int tube, double* dst, size_t n) {
    return guarded("vs_get_source_trace", [&]() -> int {
        if (check_handle(h) || !dst) return fail(VS_E_INVALID, "vs_get_source_trace: bad argument");
        if (tube < 0 || tube >= h->n) return fail(VS_E_RANGE, "tube index out of range");
        const TubeDesc& d = h->desc[tube];
        const size_t m = std::min<size_t>(n, (size_t)d.num_steps);
        double* p = (h->precision == 64 ? h->bd.strace : h->bf.strace) + d.src_off;
        VS_HIP(hipStreamSynchronize(h->stream));
        VS_HIP(copy_on(h->stream, dst, p, m * sizeof(double), hipMemcpyDeviceToHost));
        return VS_OK;
    });
}

int vs_set_option(vs_handle* h, int option, int value) {
    return guarded("vs_set_option", [&]() -> int {
        if (check_handle(h)) return VS_E_INVALID;
        switch (option) {
        case VS_OPT_STEPS_PER_LAUNCH:
            if (value < 1) return fail(VS_E_INVALID, "steps per launch must be >= 1");
            h->steps_per_launch = value; return VS_OK;
        case VS_OPT_THREADS:
            if (value < 64 || value > 1024 || value % 64) return fail(VS_E_INVALID, "threads must be a multiple of 64 in [64, 1024]");
            h->threads = value; return VS_OK;
        case VS_OPT_USE_GRAPH: h->use_graph = value ? 1 : 0; return VS_OK;
        case VS_OPT_KERNEL:
            if (value != VS_K_EXACT && value != VS_K_FAST) return fail(VS_E_INVALID, "unknown kernel family");
            if (value == VS_K_FAST && h->precision != VS_FP32) return fail(VS_E_INVALID, "the flux-form kernel is FP32 only");
            if (value != h->kernel && value == VS_K_FAST) h->flux_valid = false;
            h->kernel = value; return VS_OK;
        case VS_OPT_TIMING: h->timing = value ? 1 : 0; return VS_OK;
        case VS_OPT_ADJOINT_KERNEL:
            if (value != 0 && value != 1) return fail(VS_E_INVALID, "adjoint kernel must be 0 (on chip) or 1 (per step)");
            if (value != h->adj_kernel) h->adj_forward = false;
            h->adj_kernel = value; return VS_OK;
        case VS_OPT_EXACT_GROUPS:
            if (value < 0 || value > kCoopMaxGroups) return fail(VS_E_INVALID, "exact groups must be in [0, 64]");
            h->exact_groups = value; return VS_OK;
        case VS_OPT_EXACT_STEPS:
            if (value < 0 || value > kCoopMaxSteps) return fail(VS_E_INVALID, "exact steps must be in [0, 32]");
            h->exact_steps = value; return VS_OK;
        case VS_OPT_BARRIER_SPIN:
            if (value < 0) return fail(VS_E_INVALID, "barrier spin must be >= 0");
            h->barrier_spin = value; return VS_OK;
        case VS_OPT_FAST_LAYOUT_K: return fail(VS_E_INVALID, "VS_OPT_FAST_LAYOUT_K is read-only");
        case VS_OPT_FAST_PAIR_CLASSES: return fail(VS_E_INVALID, "VS_OPT_FAST_PAIR_CLASSES is read-only");
        case VS_OPT_FAST_ROW_QUANTUM: return fail(VS_E_INVALID, "VS_OPT_FAST_ROW_QUANTUM is read-only");
        case VS_OPT_FAST_WIDE:
            if (value != 0 && value != 1) return fail(VS_E_INVALID, "fast wide must be 0 or 1");
            h->fast_wide = value; return VS_OK;
        case VS_OPT_ADJOINT_BLOCK:
            if (value < 0) return fail(VS_E_INVALID, "adjoint block must be >= 0");
            if (value != h->adj_block) {   // slabs are re-sized by the next forward pass
                h->adj_fast_S = 0;
                if (h->adj_onchip) h->adj_forward = false;
            }
            h->adj_block = value; return VS_OK;
        case VS_OPT_ENGINE_VARIANT: {
            if (value != 0 && value != 1) return fail(VS_E_INVALID, "engine variant must be 0 or 1");
            for (TubeDesc& d : h->desc) d.variant = value;
            VS_HIP(hipMemcpyAsync(h->d_desc, h->desc.data(), h->n * sizeof(TubeDesc), hipMemcpyHostToDevice, h->stream));
            VS_HIP(hipStreamSynchronize(h->stream));
            return VS_OK;
        }
        default: return fail(VS_E_INVALID, "unknown option");
        }
    });
}

int vs_get_option(const vs_handle* h, int option, int* value) {
    return guarded("vs_get_option", [&]() -> int {
        if (check_handle(h) || !value) return fail(VS_E_INVALID, "vs_get_option: bad argument");
        switch (option) {
        case VS_OPT_STEPS_PER_LAUNCH: *value = h->steps_per_launch; return VS_OK;
        case VS_OPT_THREADS: *value = h->threads; return VS_OK;
        case VS_OPT_USE_GRAPH: *value = h->use_graph; return VS_OK;
        case VS_OPT_KERNEL: *value = h->kernel; return VS_OK;
        case VS_OPT_TIMING: *value = h->timing; return VS_OK;
        case VS_OPT_ENGINE_VARIANT: *value = h->desc.empty() ? 0 : h->desc[0].variant; return VS_OK;
        case VS_OPT_ADJOINT_KERNEL: *value = h->adj_kernel; return VS_OK;
        case VS_OPT_ADJOINT_BLOCK: *value = h->adj_block; return VS_OK;
        case VS_OPT_EXACT_GROUPS: *value = h->exact_groups; return VS_OK;
        case VS_OPT_EXACT_STEPS: *value = h->exact_steps; return VS_OK;
        case VS_OPT_BARRIER_SPIN: *value = h->barrier_spin; return VS_OK;
        case VS_OPT_FAST_WIDE: *value = h->fast_wide; return VS_OK;
        case VS_OPT_FAST_PAIR_CLASSES: *value = h->fast_classes ? h->fast_pairs : 0; return VS_OK;
        case VS_OPT_FAST_ROW_QUANTUM: {
            WideShape w;
            *value = wide_shape(h, h->threads, w) ? w.q : 0;
            return VS_OK;
        }
        case VS_OPT_FAST_LAYOUT_K: {
            int Kw = 0, K = 0;
            if (fast_wide_fits(h, h->threads, &Kw, nullptr)) *value = Kw;
            else if (fast_fits(h, h->threads, &K, nullptr)) *value = K;
            else *value = 0;
            return VS_OK;
        }
        default: return fail(VS_E_INVALID, "unknown option");
        }
    });
}

int vs_stream(vs_handle* h, void** stream) {
    return guarded("vs_stream", [&]() -> int {
        if (check_handle(h) || !stream) return fail(VS_E_INVALID, "vs_stream: bad argument");
        *stream = (void*)h->stream;
        return VS_OK;
    });
}

int vs_last_run_timing(const vs_handle* h, double* total_ms, int* launches, double* mean_launch_ms) {
    return guarded("vs_last_run_timing", [&]() -> int {
        if (check_handle(h)) return VS_E_INVALID;
        if (total_ms) *total_ms = h->last_ms;
        if (launches) *launches = h->last_launches;
        if (mean_launch_ms) *mean_launch_ms = h->last_mean_launch_ms;
        return VS_OK;
    });
}

namespace {
int dtft_device(const double* d_x, int n, double fs, int hann, const double* freqs, int nf, double* re,
                double* im, hipStream_t st) {
    double *d_f = nullptr, *d_re = nullptr, *d_im = nullptr;
    VS_HIP(hipMalloc(&d_f, sizeof(double) * nf));
    VS_HIP(hipMalloc(&d_re, sizeof(double) * nf));
    VS_HIP(hipMalloc(&d_im, sizeof(double) * nf));
    hipError_t e = hipMemcpyAsync(d_f, freqs, sizeof(double) * nf, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_dtft(d_x, 0, 1, n, fs, hann, d_f, nf, d_re, d_im, st);
    if (e == hipSuccess) e = hipMemcpyAsync(re, d_re, sizeof(double) * nf, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(im, d_im, sizeof(double) * nf, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d_f); (void)hipFree(d_re); (void)hipFree(d_im);
    if (e != hipSuccess) return fail(VS_E_DEVICE, std::string("dtft: ") + hipGetErrorString(e));
    return VS_OK;
}
}  // namespace

int vs_dtft(const double* x, int n, double fs, int hann, const double* freqs, int nf, double* re, double* im) {
    return guarded("vs_dtft", [&]() -> int {
        if (!x || !freqs || !re || !im || n <= 0 || nf <= 0 || !(fs > 0))
            return fail(VS_E_INVALID, "vs_dtft: bad argument");
        double* d_x = nullptr;
        VS_HIP(hipMalloc(&d_x, sizeof(double) * n));
        hipError_t e = hipMemcpy(d_x, x, sizeof(double) * n, hipMemcpyHostToDevice);
        int rc = e == hipSuccess ? dtft_device(d_x, n, fs, hann, freqs, nf, re, im, nullptr)
                                 : fail(VS_E_DEVICE, std::string("vs_dtft copy: ") + hipGetErrorString(e));
        (void)hipFree(d_x);
        return rc;
    });
}

int vs_dtft_noalloc(const double* x, int n, double fs, int hann, const double* freqs, int nf, double* re,
                    double* im) {
    return guarded("vs_dtft_noalloc", [&]() -> int {
        if (!x || !freqs || !re || !im || n <= 0 || nf <= 0 || !(fs > 0))
            return fail(VS_E_INVALID, "vs_dtft_noalloc: bad argument");
        if (n > kNoAllocMaxSamples || nf > kNoAllocMaxFreqs)
            return fail(VS_E_RANGE, "vs_dtft_noalloc: at most " + std::to_string(kNoAllocMaxSamples) + " samples and " +
                                        std::to_string(kNoAllocMaxFreqs) + " frequencies");
        const hipError_t e = dtft_noalloc(x, n, fs, hann, freqs, nf, re, im);
        if (e != hipSuccess) return fail(VS_E_DEVICE, std::string("vs_dtft_noalloc: ") + hipGetErrorString(e));
        return VS_OK;
    });
}

int vs_probe_dtft(vs_handle* h, int tube, int kind, int hann, const double* freqs, int nf, double* re, double* im) {
    return guarded("vs_probe_dtft", [&]() -> int {
        if (check_handle(h) || !freqs || !re || !im || nf <= 0) return fail(VS_E_INVALID, "vs_probe_dtft: bad argument");
        if (tube < 0 || tube >= h->n) return fail(VS_E_RANGE, "tube index out of range");
        if (kind != 0 && kind != 1) return fail(VS_E_INVALID, "kind must be 0 (probe) or 1 (source)");
        const TubeDesc& d = h->desc[tube];
        const int n = std::min(h->steps_done, d.num_steps);
        if (n <= 0) return fail(VS_E_STATE, "no steps advanced yet");
        const double* base = kind == 0 ? (h->precision == 64 ? h->bd.probe : h->bf.probe)
                                        : (h->precision == 64 ? h->bd.strace : h->bf.strace);
        const double fs = 1.0 / d.ph.dt;
        VS_HIP(hipStreamSynchronize(h->stream));
        return dtft_device(base + d.src_off, n, fs, hann, freqs, nf, re, im, h->stream);
    });
}

namespace {
struct TuneKey {
    int n, max_interior, precision, kernel, wide;
    long long cells;
    bool operator==(const TuneKey& o) const {
        return n == o.n && max_interior == o.max_interior && precision == o.precision && kernel == o.kernel &&
               wide == o.wide && cells == o.cells;
    }
};
std::vector<std::pair<TuneKey, std::pair<int, int>>> g_tune_cache;
std::mutex g_tune_mutex;
}  // namespace

int vs_autotune(vs_handle* h, int timed_steps, char* report, size_t report_len) {
    return guarded("vs_autotune", [&]() -> int {
        if (check_handle(h)) return VS_E_INVALID;
        if (h->steps_done != 0) return fail(VS_E_STATE, "vs_autotune needs a handle with no steps advanced");
        if (timed_steps <= 0) timed_steps = 64;
        const TuneKey key{h->n, h->max_interior, h->precision, h->kernel, h->fast_wide, h->cells_per_step};
        std::string rep = "{\"candidates\": [";
        std::unique_lock<std::mutex> cache_lock(g_tune_mutex);
        for (auto& kv : g_tune_cache)
            if (kv.first == key) {
                h->threads = kv.second.first;
                h->steps_per_launch = kv.second.second;
                if (report && report_len) {
                    std::string r = "{\"cached\": true, \"threads\": " + std::to_string(h->threads) +
                                    ", \"steps_per_launch\": " + std::to_string(h->steps_per_launch) + "}";
                    std::snprintf(report, report_len, "%s", r.c_str());
                }
                return VS_OK;
            }
        std::vector<int> threads;
        if (h->kernel == VS_K_FAST) {
            for (int t = 1024; t >= 128; t -= 64) {
                const int K = fast_cells_per_lane(h->max_padded, t);
                if (fast_wide_fits(h, t, nullptr, nullptr) ||
                    (K <= kFastMaxK && (long long)(fast_lds_bytes(K, t, fast_guard_lanes(K, h->max_wp)) +
                                           fast_lds_static_bytes()) <= h->lds_limit))
                    threads.push_back(t);
            }
        } else {
            threads = {1024, 768, 512, 256};
        }
        cache_lock.unlock();
        if (threads.empty()) return fail(VS_E_INVALID, "no work-group shape fits this batch");
        const int spls[3] = {128, 256, 512};
        const int warm = 16;
        const int old_timing = h->timing, old_graph = h->use_graph;
        h->timing = 0;
        h->use_graph = 0;
        double best = 1e300;
        int best_t = h->threads, best_s = h->steps_per_launch;
        bool first = true;
        for (int t : threads) {
            for (int spl : spls) {
                if (spl > 2 * timed_steps && spl != 128) continue;
                h->threads = t;
                h->steps_per_launch = spl;
                int rc = vs_reset(h);
                if (rc == VS_OK) rc = vs_run(h, warm);
                if (rc == VS_OK) rc = vs_sync(h);
                if (rc == VS_OK) rc = vs_run(h, timed_steps);
                if (rc == VS_OK) rc = vs_sync(h);
                if (rc != VS_OK) {
                    h->timing = old_timing; h->use_graph = old_graph;
                    return rc;
                }
                const int ran = h->steps_done - warm;
                if (ran <= 0) continue;
                const double us = h->last_ms * 1e3 / ran;
                rep += std::string(first ? "" : ", ") + "{\"threads\": " + std::to_string(t) +
                       ", \"steps_per_launch\": " + std::to_string(spl) + ", \"us_per_step\": " + std::to_string(us) + "}";
                first = false;
                if (us < best) { best = us; best_t = t; best_s = spl; }
            }
        }
        h->threads = best_t;
        h->steps_per_launch = best_s;
        h->timing = old_timing;
        h->use_graph = old_graph;
        const int rc = vs_reset(h);
        {
            std::lock_guard<std::mutex> lk(g_tune_mutex);
            g_tune_cache.push_back({key, {best_t, best_s}});
        }
        rep += "], \"threads\": " + std::to_string(best_t) + ", \"steps_per_launch\": " + std::to_string(best_s) +
               ", \"us_per_step\": " + std::to_string(best) + "}";
        if (report && report_len) std::snprintf(report, report_len, "%s", rep.c_str());
        return rc;
    });
}

int vs_copy_to_host(void* dst, const void* src, size_t bytes) {
    return guarded("vs_copy_to_host", [&]() -> int {
        if (!dst || !src) return fail(VS_E_INVALID, "vs_copy_to_host: null pointer");
        if (bytes == 0) return VS_OK;
        hipPointerAttribute_t attr{};
        const hipError_t e = hipPointerGetAttributes(&attr, src);
        if (e == hipSuccess && attr.type == hipMemoryTypeDevice) {
            VS_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
        } else {
            (void)hipGetLastError();
            std::memcpy(dst, src, bytes);
        }
        return VS_OK;
    });
}

int vs_cells_per_step(const vs_handle* h, long long* cells) {
    return guarded("vs_cells_per_step", [&]() -> int {
        if (check_handle(h) || !cells) return fail(VS_E_INVALID, "vs_cells_per_step: bad argument");
        *cells = h->cells_per_step;
        return VS_OK;
    });
}

int vs_measure_lds_bandwidth(int width_bytes, double* bytes_per_s) {
    return guarded("vs_measure_lds_bandwidth", [&]() -> int {
        if (!bytes_per_s || (width_bytes != 4 && width_bytes != 8 && width_bytes != 16))
            return fail(VS_E_INVALID, "vs_measure_lds_bandwidth: width must be 4, 8 or 16 bytes");
        int dev = 0, cus = 0;
        VS_HIP(hipGetDevice(&dev));
        VS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        const int groups = 8 * std::max(cus, 1), iters = 512, words = width_bytes / 4;
        float* d_out = nullptr;
        VS_HIP(hipMalloc(&d_out, sizeof(float) * (size_t)groups * 1024));
        hipStream_t st = nullptr;
        hipError_t e = hipStreamCreate(&st);
        float best = 0.f;
        for (int rep = 0; rep < 4 && e == hipSuccess; ++rep) {   // first launch warms up
            float ms = 0.f;
            e = measure_lds_read(words, groups, iters, d_out, st, &ms);
            if (e == hipSuccess && rep > 0 && (best == 0.f || ms < best)) best = ms;
        }
        if (st) (void)hipStreamDestroy(st);
        (void)hipFree(d_out);
        if (e != hipSuccess) return fail(VS_E_DEVICE, std::string("lds measurement: ") + hipGetErrorString(e));
        *bytes_per_s = (double)groups * 1024.0 * iters * 8.0 * width_bytes / (best * 1e-3);
        return VS_OK;
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Reverse-time pass: gradients of J = sum_n g[n] probe[n] with respect to
// each section's area, depth scale and wall admittance.
// ---------------------------------------------------------------------------
namespace {

template <typename Real>
AdjBufs<Real>& adj_of(vs_handle* h);
template <>
AdjBufs<float>& adj_of<float>(vs_handle* h) { return h->af; }
template <>
AdjBufs<double>& adj_of<double>(vs_handle* h) { return h->ad; }

// d z_inv / d mu of the admittance map z_inv(mu) (FDTDSolver.cu:613-617):
// z = (1 - s) / (rho c (1 + s)), s = sqrt(1 - a), a = 1 / (1/2 + (m + 1/m) / 4),
// m = mu pi / 1.84.
double dz_dmu(double mu, double rho_c) {
    const double kap = (2.0 * 0.5 * M_PI) / 1.84;
    const double m2 = mu * kap;
    const double a = 1.0 / (0.5 + 0.25 * (m2 + 1.0 / m2));
    const double sq = std::sqrt(1.0 - a);
    const double da_dm = -a * a * 0.25 * (1.0 - 1.0 / (m2 * m2));
    const double ds_da = -0.5 / sq;
    const double dz_ds = -2.0 / (rho_c * (1.0 + sq) * (1.0 + sq));
    return dz_ds * ds_da * da_dm * kap;
}

// Allocates the per-cell buffers of the reverse-time pass (once).
template <typename Real>
int adj_cells(vs_handle* h) {
    AdjBufs<Real>& a = adj_of<Real>(h);
    Bufs<Real>& b = bufs_of<Real>(h);
    int rc;
    if (!h->adj_ready) {
        const long long n = h->total_cells;
        if ((rc = h->alloc(&a.aP, n)) || (rc = h->alloc(&a.kP, n)) || (rc = h->alloc(&a.ax, n)) ||
            (rc = h->alloc(&a.ay, n)) || (rc = h->alloc(&a.cx, n)) || (rc = h->alloc(&a.cy, n)) ||
            (rc = h->alloc(&a.fx, n)) || (rc = h->alloc(&a.fy, n)) || (rc = h->alloc(&a.P, n)) ||
            (rc = h->alloc(&a.Ux, n)) || (rc = h->alloc(&a.Uy, n)) || (rc = h->alloc(&a.LP, n)) ||
            (rc = h->alloc(&a.LX, n)) || (rc = h->alloc(&a.LY, n)) || (rc = h->alloc(&a.M, n)) ||
            (rc = h->alloc(&a.GkP, n)) || (rc = h->alloc(&a.Gx, n)) || (rc = h->alloc(&a.Gy, n)) ||
            (rc = h->alloc(&a.probe, h->total_src)) || (rc = h->alloc(&a.gprobe, h->total_src)) ||
            (rc = h->alloc(&a.ck_off, h->n)) || (rc = h->alloc(&a.seg_off, h->n)))
            return rc;
        hipError_t e = launch_adj_coeffs<Real>(h->d_desc, h->desc.data(), h->n, b, a, h->stream);
        if (e != hipSuccess) return fail(VS_E_DEVICE, std::string("adjoint coefficients: ") + hipGetErrorString(e));
        h->adj_ready = true;
    }
    return VS_OK;
}

// Per-cell buffers plus the checkpoint / segment slabs of the per-step
// kernels for checkpoint spacing S (again when S changes).
template <typename Real>
int adj_prepare(vs_handle* h, int S) {
    AdjBufs<Real>& a = adj_of<Real>(h);
    int rc = adj_cells<Real>(h);
    if (rc != VS_OK) return rc;
    if (h->adj_S != S) {
        VS_HIP(hipStreamSynchronize(h->stream));
        for (void* p : h->adj_slabs) (void)hipFree(p);
        h->adj_slabs.clear();
        a.ck = a.seg = nullptr;
        h->adj_S = 0;
        h->adj_forward = false;
        std::vector<long long> cko(h->n), sgo(h->n);
        long long ck = 0, sg = 0;
        for (int k = 0; k < h->n; ++k) {
            const TubeDesc& d = h->desc[k];
            const long long st = 3LL * d.H * d.W;
            cko[k] = ck;
            sgo[k] = sg;
            ck += (((long long)d.num_steps + S - 1) / S) * st;
            sg += ((long long)S + 1) * st;
        }
        void* p1 = nullptr;
        void* p2 = nullptr;
        hipError_t e = hipMalloc(&p1, sizeof(Real) * (size_t)std::max(ck, 1LL));
        if (e != hipSuccess)
            return fail(VS_E_DEVICE, "checkpoint slab of " + std::to_string(sizeof(Real) * ck) + " bytes: " +
                                         hipGetErrorString(e));
        h->adj_slabs.push_back(p1);
        e = hipMalloc(&p2, sizeof(Real) * (size_t)std::max(sg, 1LL));
        if (e != hipSuccess)
            return fail(VS_E_DEVICE, "segment slab of " + std::to_string(sizeof(Real) * sg) + " bytes: " +
                                         hipGetErrorString(e));
        h->adj_slabs.push_back(p2);
        a.ck = static_cast<Real*>(p1);
        a.seg = static_cast<Real*>(p2);
        VS_HIP(hipMemcpyAsync(a.ck_off, cko.data(), h->n * sizeof(long long), hipMemcpyHostToDevice, h->stream));
        VS_HIP(hipMemcpyAsync(a.seg_off, sgo.data(), h->n * sizeof(long long), hipMemcpyHostToDevice, h->stream));
        VS_HIP(hipStreamSynchronize(h->stream));
        a.S = S;
        h->adj_S = S;
    }
    return VS_OK;
}

// ----- on-chip FP32 reverse pass -----
// Cell lanes per tube of the on-chip reverse kernels (one work-group of
// kAdjLanes + 64 threads per tube in the forward, kAdjLanes in the reverse).
constexpr int kAdjThreads = 1024;
constexpr int kAdjLanes = kAdjThreads - 64;
// Device memory allowed for the record slabs (two buffers of two slabs):
// half of the free memory, at most 64 GiB.
constexpr long long kAdjSlabCap = 64LL << 30;

bool adj_uses_onchip(const vs_handle* h) {
    return h->precision == VS_FP32 && h->adj_kernel == 0 && fast_fits(h, kAdjThreads, nullptr, nullptr);
}

// Slabs of the on-chip pass for checkpoint spacing S: checkpoints of the
// whole state, the record block of B <= S steps (rec, adj), the adjoint
// state between launches and, when B < S, the block starts of one segment.
int adj_fast_prepare(vs_handle* h, int S) {
    int K = 0, G = 0;
    if (!fast_fits(h, kAdjThreads, &K, &G)) return fail(VS_E_INVALID, "batch does not fit the on-chip kernels");
    int rc = prepare_fast(h, kAdjThreads);
    if (rc != VS_OK) return rc;
    if (h->adj_fast_S == S && h->adj_K == K) return VS_OK;
    release_adj_fast_slabs(h);
    h->adj_fast_S = 0;
    h->adj_forward = false;
    const long long KT = (long long)K * kAdjLanes;
    const long long per_step = 2LL * h->n * 3 * KT * (long long)sizeof(float);   // two buffers of rec (2) + adj (1)
    size_t mem_free = 0, mem_total = 0;
    VS_HIP(hipMemGetInfo(&mem_free, &mem_total));
    // slabs in the cache count as free: the requests below take them first
    const long long budget =
        std::min<long long>(kAdjSlabCap, (long long)((mem_free + slab_cached_bytes(h->device)) / 2));
    int B = (int)std::max(1LL, std::min<long long>(S, budget / per_step));
    if (h->adj_block > 0) B = std::min(B, h->adj_block);
    const long long C = h->total_cells;
    const long long nck = (h->max_steps + S - 1) / S;
    const long long nsub = B < S ? (S + B - 1) / B : 0;
    auto get = [&](void** p, long long bytes, const char* what) -> int {
        size_t got = 0;
        hipError_t e = slab_alloc(h->device, (size_t)std::max(bytes, 1LL), p, &got);
        if (e != hipSuccess)
            return fail(VS_E_DEVICE, std::string(what) + " of " + std::to_string(bytes) + " bytes: " + hipGetErrorString(e));
        h->adj_fast_slabs.emplace_back(*p, got);
        return VS_OK;
    };
    void *ck, *sub, *rec0, *rec1, *adj0, *adj1, *sa0, *sa1, *st, *sk, *tr, *pe, *tm;
    const long long slab = (long long)h->n * B * KT * 4;   // one field
    if ((rc = get(&ck, nck * 3 * C * 4, "checkpoints")) || (rc = get(&sub, nsub * 3 * C * 4, "block starts")) ||
        (rc = get(&rec0, 2 * slab, "forward record slab")) || (rc = get(&rec1, 2 * slab, "forward record slab")) ||
        (rc = get(&adj0, slab, "adjoint record slab")) || (rc = get(&adj1, slab, "adjoint record slab")) ||
        (rc = get(&sa0, (long long)h->n * 2 * KT * 4, "block-start adjoints")) ||
        (rc = get(&sa1, (long long)h->n * 2 * KT * 4, "block-start adjoints")) ||
        (rc = get(&st, (long long)h->n * 3 * KT * 4, "adjoint state")) || (rc = get(&sk, 3 * KT * 4, "store sink")) ||
        (rc = get(&tr, h->total_src * 8, "source trace")) || (rc = get(&pe, h->n * 8LL, "excitation pressure")) ||
        (rc = get(&tm, h->n * (long long)sizeof(TwoMassState), "two-mass state")))
        return rc;
    h->adj_ck = static_cast<float*>(ck);
    h->adj_sub = static_cast<float*>(sub);
    h->adj_strace = static_cast<double*>(tr);
    h->adj_pexc = static_cast<double*>(pe);
    h->adj_tm = static_cast<TwoMassState*>(tm);
    h->adj_rec[0] = static_cast<float*>(rec0);
    h->adj_rec[1] = static_cast<float*>(rec1);
    h->adj_adj[0] = static_cast<float*>(adj0);
    h->adj_adj[1] = static_cast<float*>(adj1);
    h->adj_start[0] = static_cast<float*>(sa0);
    h->adj_start[1] = static_cast<float*>(sa1);
    if (!h->adj_rev_stream) {
        VS_HIP(hipStreamCreateWithFlags(&h->adj_rec_stream, hipStreamNonBlocking));
        VS_HIP(hipStreamCreateWithFlags(&h->adj_rev_stream, hipStreamNonBlocking));
        VS_HIP(hipStreamCreateWithFlags(&h->adj_sum_stream, hipStreamNonBlocking));
        for (hipEvent_t& e : h->adj_ev) VS_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    // Recording and reverse launches hold one CU per tube for the whole
    // launch.  The sums run one 1024-thread work-group per 1024 slots of a
    // tube, so each thread walks one cell.  Work-groups queue for the CUs
    // the other two launches leave, and fill theirs as they finish: on the
    // 64-tube gradient workload this took the backward pass from 128-133 ms
    // to 101-104 ms, against a grid of two work-groups per CU on the CUs the
    // recording and reverse launches leave free (smaller grids and 1.3-2.7
    // times this one measured in between).
    h->adj_sum_groups = h->n * ((K * kAdjLanes + 1023) / 1024);
    FastAdj& fa = h->fa;
    fa.rec = fa.adj = fa.start = nullptr;   // per block: one of the two buffers
    fa.state = static_cast<float*>(st);
    fa.sink = static_cast<float*>(sk);
    fa.rec_tube = (long long)B * 2 * KT;
    fa.adj_tube = (long long)B * KT;
    fa.lanes = kAdjLanes;
    h->adj_K = K;
    h->adj_G = G;
    h->adj_B = B;
    h->adj_fast_S = S;
    return VS_OK;
}

// The forward state of the on-chip pass lives in the adjoint buffers, so the
// solver's own state, probe and source trace are left untouched.
Bufs<float> adj_fast_bufs(vs_handle* h) {
    Bufs<float> b = h->bf;
    AdjBufs<float>& a = h->af;
    b.P = a.P;
    b.Vx = a.LX;   // velocity write-back of a launch: scratch
    b.Vy = a.LY;
    b.probe = a.probe;
    b.strace = h->adj_strace;
    b.pexc = h->adj_pexc;
    b.tm = h->adj_tm;
    return b;
}
FastBufs adj_fast_flux(vs_handle* h) {
    FastBufs f = h->fb;
    f.Uxf = h->af.Ux;
    f.Uyf = h->af.Uy;
    return f;
}

// Copies the forward state (P, Ux, Uy) to / from a slab of 3 * total_cells.
int adj_state_copy(vs_handle* h, float* slab, bool to_slab, hipStream_t st) {
    AdjBufs<float>& a = h->af;
    const size_t cb = sizeof(float) * (size_t)h->total_cells;
    const long long C = h->total_cells;
    float* state[3] = {a.P, a.Ux, a.Uy};
    for (int i = 0; i < 3; ++i) {
        float* dst = to_slab ? slab + i * C : state[i];
        const float* src = to_slab ? state[i] : slab + i * C;
        VS_HIP(hipMemcpyAsync(dst, src, cb, hipMemcpyDeviceToDevice, st));
    }
    return VS_OK;
}

int adj_check(vs_handle* h) {
    for (int k = 0; k < h->n; ++k)
        if (h->desc[k].source_type == 2)
            return fail(VS_E_INVALID, "tube " + std::to_string(k) +
                                          ": the reverse-time pass needs a table source (source_type 0 or 1)");
    if (h->threads < 64 || h->threads > 1024 || h->threads % 64)
        return fail(VS_E_INVALID, "threads must be a multiple of 64 in [64, 1024]");
    return VS_OK;
}

// Shape of the checkpointing forward.  Its state is kept in cell layout, so
// it need not share the lanes shape of the recording and reverse launches:
// the fastest of the shapes that fit, odd-run or wide-access form (which give
// the same values), is timed once per handle (64 steps each, on the
// adjoint's own state, which the forward pass then clears).
// (batch size, largest padded interior and widest padded row in both layouts,
// wide form allowed) -> (threads, wide),
// shared by the handles of one process (a differentiable call builds a new
// handle per call).
std::vector<std::pair<std::array<int, 6>, std::array<int, 2>>> g_adj_shape_cache;
std::mutex g_adj_shape_mutex;

int adj_pick_forward_threads(vs_handle* h) {
    if (h->adj_fwd_threads) return VS_OK;
    const std::array<int, 6> key{h->n, h->max_padded, h->max_padded_w[2], h->max_padded4, h->max_wp4, h->fast_wide};
    {
        std::lock_guard<std::mutex> lk(g_adj_shape_mutex);
        for (auto& kv : g_adj_shape_cache)
            if (kv.first == key) {
                h->adj_fwd_threads = kv.second[0];
                h->adj_fwd_wide = kv.second[1];
                return VS_OK;
            }
    }
    const Bufs<float> b = adj_fast_bufs(h);
    const FastBufs f = adj_fast_flux(h);
    float best = 1e30f;
    int best_t = kAdjThreads, best_w = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    VS_HIP(hipEventCreate(&e0));
    VS_HIP(hipEventCreate(&e1));
    int rc = VS_OK;
    for (int wide = 0; wide < 2 && rc == VS_OK; ++wide)
        for (int t = kAdjThreads; t >= kAdjThreads - 4 * 64 && rc == VS_OK; t -= 64) {
            int K = 0, G = 0;
            WideShape w;
            if (wide ? !wide_shape(h, t, w) : !fast_fits(h, t, &K, &G)) continue;
            if (wide) {
                K = w.K;
                G = w.GW;
            }
            hipError_t e = wide ? configure_fast_wide(K, w.ms) : configure_fast(K);
            float ms = 1e30f;
            for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {
                e = hipEventRecord(e0, h->stream);
                if (e == hipSuccess)
                    e = wide ? launch_fast_wide(K, w.ms, h->d_desc, h->d_order, h->n, b, f, 0, 64, t, G, w.q, nullptr,
                                                h->stream)
                             : launch_fast_step(K, h->d_desc, h->d_order, h->n, b, f, 0, 64, t, G, nullptr, h->stream);
                if (e == hipSuccess) e = hipEventRecord(e1, h->stream);
                if (e == hipSuccess) e = hipEventSynchronize(e1);
                float m = 0.f;
                if (e == hipSuccess) e = hipEventElapsedTime(&m, e0, e1);
                if (e == hipSuccess) ms = std::min(ms, m);
            }
            if (e != hipSuccess) rc = fail(VS_E_DEVICE, std::string("adjoint forward shape: ") + hipGetErrorString(e));
            else if (ms < best) { best = ms; best_t = t; best_w = wide; }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc == VS_OK) {
        h->adj_fwd_threads = best_t;
        h->adj_fwd_wide = best_w;
        std::lock_guard<std::mutex> lk(g_adj_shape_mutex);
        g_adj_shape_cache.push_back({key, {best_t, best_w}});
    }
    return rc;
}

// Copies the forward pass's probe traces to `probe` (host or device memory:
// the direction is taken from the pointer) and waits for the stream, so the
// traces are complete when the entry point returns.
int adj_probe_out(vs_handle* h, const double* src, double* probe) {
    if (probe)
        VS_HIP(hipMemcpyAsync(probe, src, sizeof(double) * (size_t)h->total_src, hipMemcpyDefault, h->stream));
    VS_HIP(hipStreamSynchronize(h->stream));
    return VS_OK;
}

// Forward pass on chip: the FP32 flux-form step kernel (the form and shape
// adj_pick_forward_threads timed) in S-step launches, the whole state copied
// to a checkpoint before each.
int adj_forward_fast(vs_handle* h, int S, double* probe) {
    int rc;
    if ((rc = adj_cells<float>(h)) != VS_OK || (rc = adj_fast_prepare(h, S)) != VS_OK ||
        (rc = adj_pick_forward_threads(h)) != VS_OK)
        return rc;
    int FK = 0, FG = 0;
    const bool wide = h->adj_fwd_wide != 0;
    WideShape fw;
    if (wide ? !wide_shape(h, h->adj_fwd_threads, fw) : !fast_fits(h, h->adj_fwd_threads, &FK, &FG))
        return fail(VS_E_INVALID, "forward shape no longer fits");
    if (wide) {
        FK = fw.K;
        FG = fw.GW;
    }
    {
        const hipError_t e = wide ? configure_fast_wide(FK, fw.ms) : configure_fast(FK);
        if (e != hipSuccess) return fail(VS_E_DEVICE, std::string("configure: ") + hipGetErrorString(e));
    }
    AdjBufs<float>& a = h->af;
    const size_t cb = sizeof(float) * (size_t)h->total_cells;
    VS_HIP(hipMemsetAsync(a.P, 0, cb, h->stream));
    VS_HIP(hipMemsetAsync(a.Ux, 0, cb, h->stream));
    VS_HIP(hipMemsetAsync(a.Uy, 0, cb, h->stream));
    const Bufs<float> b = adj_fast_bufs(h);
    const FastBufs f = adj_fast_flux(h);
    for (int t0 = 0; t0 < h->max_steps; t0 += S) {
        if ((rc = adj_state_copy(h, h->adj_ck + (long long)(t0 / S) * 3 * h->total_cells, true, h->stream)) != VS_OK)
            return rc;
        hipError_t e = wide ? launch_fast_wide(FK, fw.ms, h->d_desc, h->d_order, h->n, b, f, t0, S,
                                               h->adj_fwd_threads, FG, fw.q, nullptr, h->stream)
                            : launch_fast_step(FK, h->d_desc, h->d_order, h->n, b, f, t0, S, h->adj_fwd_threads, FG,
                                               nullptr, h->stream);
        if (e != hipSuccess) return fail(VS_E_DEVICE, std::string("adjoint forward: ") + hipGetErrorString(e));
    }
    h->adj_forward = true;
    h->adj_onchip = true;
    return adj_probe_out(h, a.probe, probe);
}

// Reverse pass on chip.  Per segment (last first): the segment is recomputed
// from its checkpoint in record blocks of B steps (block starts kept when
// B < S); per block, last first: the recording forward launch (this stream),
// the reverse launch (its own stream: it reads only the adjoint state, so it
// overlaps the recording) and the per-cell sums of the block (a third stream,
// after both).  Record slabs alternate between two buffers; a buffer is
// written again only after the sums of the block before last have read it.
// Leaves d J / d log(gX, gY, kP) in a.Gx, a.Gy, a.GkP.
int adj_backward_fast(vs_handle* h) {
    AdjBufs<float>& a = h->af;
    const int S = h->adj_fast_S, B = h->adj_B, K = h->adj_K, G = h->adj_G;
    const long long C = h->total_cells;
    const Bufs<float> b = adj_fast_bufs(h);
    const FastBufs f = adj_fast_flux(h);
    FastAdj fa = h->fa;
    fa.gprobe = a.gprobe;
    fa.Lx = a.Gx;
    fa.Ly = a.Gy;
    fa.Lp = a.GkP;
    hipStream_t s_main = h->stream, sm = h->adj_rec_stream, sr = h->adj_rev_stream, ss = h->adj_sum_stream;
    hipEvent_t* ev = h->adj_ev;   // 0 init, 1 + i rec, 3 + i rev, 5 + i sums (buffer i)
    VS_HIP(hipMemsetAsync(fa.state, 0, sizeof(float) * (size_t)h->n * 3 * K * kAdjLanes, s_main));
    VS_HIP(hipEventRecord(ev[0], s_main));
    VS_HIP(hipStreamWaitEvent(sm, ev[0], 0));
    VS_HIP(hipStreamWaitEvent(sr, ev[0], 0));
    VS_HIP(hipStreamWaitEvent(ss, ev[0], 0));
    const int nseg = (h->max_steps + S - 1) / S;
    int rc, blocks = 0;
    for (int k = nseg - 1; k >= 0; --k) {
        const int t0 = k * S, len = std::min(S, h->max_steps - t0);
        const int nb = (len + B - 1) / B;
        if ((rc = adj_state_copy(h, h->adj_ck + (long long)k * 3 * C, false, sm)) != VS_OK) return rc;
        for (int j = 0; j + 1 < nb; ++j) {
            if ((rc = adj_state_copy(h, h->adj_sub + (long long)j * 3 * C, true, sm)) != VS_OK) return rc;
            hipError_t e = launch_fast_step(K, h->d_desc, h->d_order, h->n, b, f, t0 + j * B, B, kAdjThreads, G,
                                            nullptr, sm);
            if (e != hipSuccess) return fail(VS_E_DEVICE, std::string("adjoint block starts: ") + hipGetErrorString(e));
        }
        for (int j = nb - 1; j >= 0; --j, ++blocks) {
            if (j + 1 < nb && (rc = adj_state_copy(h, h->adj_sub + (long long)j * 3 * C, false, sm)) != VS_OK)
                return rc;
            const int i = blocks & 1;
            const int s0 = t0 + j * B, ns = std::min(B, len - j * B);
            fa.rec = h->adj_rec[i];
            fa.adj = h->adj_adj[i];
            fa.start = h->adj_start[i];
            if (blocks >= 2) VS_HIP(hipStreamWaitEvent(sm, ev[5 + i], 0));
            hipError_t e = launch_fast_record(K, h->d_desc, h->d_order, h->n, b, f, fa, s0, ns, kAdjThreads, G, sm);
            if (e == hipSuccess) e = hipEventRecord(ev[1 + i], sm);
            if (e == hipSuccess && blocks >= 2) e = hipStreamWaitEvent(sr, ev[5 + i], 0);
            if (e == hipSuccess) e = launch_fast_reverse(K, h->d_desc, h->d_order, h->n, f, fa, s0, ns, G, sr);
            if (e == hipSuccess) e = hipEventRecord(ev[3 + i], sr);
            if (e == hipSuccess) e = hipStreamWaitEvent(ss, ev[1 + i], 0);
            if (e == hipSuccess) e = hipStreamWaitEvent(ss, ev[3 + i], 0);
            if (e == hipSuccess)
                e = launch_fast_adj_sums(K, h->d_desc, h->d_order, h->n, f, fa, h->bf.source, s0, ns,
                                         h->adj_sum_groups, ss);
            if (e == hipSuccess) e = hipEventRecord(ev[5 + i], ss);
            if (e != hipSuccess) return fail(VS_E_DEVICE, std::string("adjoint reverse: ") + hipGetErrorString(e));
        }
    }
    // the sums stream has waited for every recording and reverse launch
    VS_HIP(hipEventRecord(ev[0], ss));
    VS_HIP(hipStreamWaitEvent(s_main, ev[0], 0));
    return VS_OK;
}

template <typename Real>
int adj_forward(vs_handle* h, int S, double* probe, size_t len) {
    int rc = adj_check(h);
    if (rc != VS_OK) return rc;
    if (probe && len < (size_t)h->total_src)
        return fail(VS_E_RANGE, "probe buffer needs " + std::to_string(h->total_src) + " values");
    if (adj_uses_onchip(h)) return adj_forward_fast(h, S, probe);
    if ((rc = adj_prepare<Real>(h, S)) != VS_OK) return rc;
    AdjBufs<Real>& a = adj_of<Real>(h);
    Bufs<Real>& b = bufs_of<Real>(h);
    const size_t cb = sizeof(Real) * (size_t)h->total_cells;
    VS_HIP(hipMemsetAsync(a.P, 0, cb, h->stream));
    VS_HIP(hipMemsetAsync(a.Ux, 0, cb, h->stream));
    VS_HIP(hipMemsetAsync(a.Uy, 0, cb, h->stream));
    for (int t0 = 0; t0 < h->max_steps; t0 += S) {
        hipError_t e = launch_adj_fwd<Real>(h->d_desc, h->d_order, h->n, b, a, t0, S, 1, -1, h->threads, h->stream);
        if (e != hipSuccess) return fail(VS_E_DEVICE, std::string("adjoint forward: ") + hipGetErrorString(e));
    }
    h->adj_forward = true;
    h->adj_onchip = false;
    return adj_probe_out(h, a.probe, probe);
}

template <typename Real>
int adj_backward(vs_handle* h, const double* gprobe, size_t glen, double* g_area, double* g_scale, double* g_mu,
                 size_t nsec) {
    if (!h->adj_forward) return fail(VS_E_STATE, "vs_adjoint_backward needs a completed vs_adjoint_forward");
    if (!gprobe || glen != (size_t)h->total_src)
        return fail(VS_E_RANGE, "gprobe needs " + std::to_string(h->total_src) + " values (tube-major)");
    if (nsec != (size_t)h->total_secs)
        return fail(VS_E_RANGE, "gradient buffers need " + std::to_string(h->total_secs) + " values");
    AdjBufs<Real>& a = adj_of<Real>(h);
    Bufs<Real>& b = bufs_of<Real>(h);
    const int S = h->adj_S;
    const auto t_start = std::chrono::steady_clock::now();
    // gprobe may be host or device memory (the direction is taken from the pointer)
    VS_HIP(hipMemcpyAsync(a.gprobe, gprobe, sizeof(double) * glen, hipMemcpyDefault, h->stream));
    const size_t cr = sizeof(Real) * (size_t)h->total_cells, cd = sizeof(double) * (size_t)h->total_cells;
    VS_HIP(hipMemsetAsync(a.LP, 0, cr, h->stream));
    VS_HIP(hipMemsetAsync(a.LX, 0, cr, h->stream));
    VS_HIP(hipMemsetAsync(a.LY, 0, cr, h->stream));
    VS_HIP(hipMemsetAsync(a.M, 0, cr, h->stream));
    VS_HIP(hipMemsetAsync(a.GkP, 0, cd, h->stream));
    VS_HIP(hipMemsetAsync(a.Gx, 0, cd, h->stream));
    VS_HIP(hipMemsetAsync(a.Gy, 0, cd, h->stream));
    int rc = VS_OK;
    if (h->adj_onchip) {
        if ((rc = adj_backward_fast(h)) != VS_OK) return rc;
    } else {
        const int nseg = (h->max_steps + S - 1) / S;
        for (int k = nseg - 1; k >= 0; --k) {
            hipError_t e = launch_adj_fwd<Real>(h->d_desc, h->d_order, h->n, b, a, k * S, S, 2, k, h->threads, h->stream);
            if (e == hipSuccess) e = launch_adj_rev<Real>(h->d_desc, h->d_order, h->n, b, a, k * S, S, h->threads, h->stream);
            if (e != hipSuccess) return fail(VS_E_DEVICE, std::string("adjoint reverse: ") + hipGetErrorString(e));
        }
    }
    // raw depth maps for the chain through smoothing and clamping
    long long maxHW = 0;
    for (auto& d : h->desc) maxHW = std::max(maxHW, (long long)d.H * d.W);
    void* raw_v = nullptr;
    size_t raw_got = 0;
    VS_HIP(slab_alloc(h->device, sizeof(Real) * (size_t)maxHW * h->n * 2, &raw_v, &raw_got));
    Real* raw = static_cast<Real*>(raw_v);
    hipError_t e = launch_depth_raw<Real>(h->d_desc, h->desc.data(), h->n, b, raw, raw + maxHW * h->n, maxHW, h->stream);
    hipError_t e2 = hipStreamSynchronize(h->stream);
    if (e != hipSuccess || e2 != hipSuccess) slab_give(h->device, raw_v, raw_got);
    if (e != hipSuccess) return fail(VS_E_DEVICE, std::string("raw depth: ") + hipGetErrorString(e));
    if (e2 != hipSuccess) return fail(VS_E_DEVICE, std::string("reverse pass: ") + hipGetErrorString(e2));
    const auto t_device = std::chrono::steady_clock::now();
    // everything the host chain reads, copied back into one pinned block:
    // Lp, Lx, Ly (d J / d log of each cell's kP, x-face and y-face
    // coefficient), the raw depth maps and the set-up fields it chains through
    const long long C = h->total_cells;
    size_t total = 0;
    auto place = [&](size_t bytes) {
        const size_t o = total;
        total += (bytes + 255) & ~size_t(255);
        return o;
    };
    const size_t R = sizeof(Real);
    const size_t oLp = place(8 * C), oLx = place(8 * C), oLy = place(8 * C);
    const size_t oRaw = place(R * (size_t)maxHW * h->n * 2);
    const size_t odP = place(R * C), odX = place(R * C), odY = place(R * C), ocls = place((size_t)C);
    const size_t ozc = place(R * h->total_cols), oseg = place(4 * (size_t)h->total_walls),
                 oup = place(4 * (size_t)h->total_walls), olo = place(4 * (size_t)h->total_walls),
                 omin = place(R * h->total_secs), odsc = place(R * h->total_secs);
    const bool coef = !h->adj_onchip;   // the per-step kernels accumulate d J / d (kP, cx, cy)
    const size_t okP = coef ? place(R * C) : 0, ocx = coef ? place(R * C) : 0, ocy = coef ? place(R * C) : 0;
    PinnedBlock pin;
    e = pinned_take(total, &pin.p, &pin.bytes);
    if (e != hipSuccess) {
        pin.p = nullptr;
        slab_give(h->device, raw_v, raw_got);
        return fail(VS_E_DEVICE, std::string("pinned staging: ") + hipGetErrorString(e));
    }
    char* const P = static_cast<char*>(pin.p);
    struct Copy {
        size_t off;
        const void* src;
        size_t bytes;
    };
    std::vector<Copy> copies = {{oLp, a.GkP, 8 * (size_t)C}, {oLx, a.Gx, 8 * (size_t)C}, {oLy, a.Gy, 8 * (size_t)C},
                                {oRaw, raw, R * (size_t)maxHW * h->n * 2},
                                {odP, b.dP, R * C}, {odX, b.dX, R * C}, {odY, b.dY, R * C}, {ocls, b.cls, (size_t)C},
                                {ozc, b.zcol, R * h->total_cols}, {oseg, b.seg, 4 * (size_t)h->total_walls},
                                {oup, b.upper, 4 * (size_t)h->total_walls}, {olo, b.lower, 4 * (size_t)h->total_walls},
                                {omin, b.minor, R * h->total_secs}, {odsc, b.dscale, R * h->total_secs}};
    if (coef) {
        copies.push_back({okP, a.kP, R * C});
        copies.push_back({ocx, a.cx, R * C});
        copies.push_back({ocy, a.cy, R * C});
    }
    for (const Copy& c : copies)
        if (e == hipSuccess && c.bytes) e = hipMemcpyAsync(P + c.off, c.src, c.bytes, hipMemcpyDeviceToHost, h->stream);
    e2 = hipStreamSynchronize(h->stream);
    slab_give(h->device, raw_v, raw_got);
    if (e != hipSuccess) return fail(VS_E_DEVICE, std::string("copy back: ") + hipGetErrorString(e));
    if (e2 != hipSuccess) return fail(VS_E_DEVICE, std::string("copy back: ") + hipGetErrorString(e2));
    double* const Lp = reinterpret_cast<double*>(P + oLp);
    double* const Lx = reinterpret_cast<double*>(P + oLx);
    double* const Ly = reinterpret_cast<double*>(P + oLy);
    const Real* const rawh = reinterpret_cast<const Real*>(P + oRaw);
    const Real* const dP = reinterpret_cast<const Real*>(P + odP);
    const Real* const dX = reinterpret_cast<const Real*>(P + odX);
    const Real* const dY = reinterpret_cast<const Real*>(P + odY);
    const uint8_t* const cls = reinterpret_cast<const uint8_t*>(P + ocls);
    const Real* const zcol = reinterpret_cast<const Real*>(P + ozc);
    const int* const seg = reinterpret_cast<const int*>(P + oseg);
    const int* const upper = reinterpret_cast<const int*>(P + oup);
    const int* const lower = reinterpret_cast<const int*>(P + olo);
    const Real* const minor = reinterpret_cast<const Real*>(P + omin);
    const Real* const dscale = reinterpret_cast<const Real*>(P + odsc);
    if (coef) {
        const Real* const kP = reinterpret_cast<const Real*>(P + okP);
        const Real* const cx = reinterpret_cast<const Real*>(P + ocx);
        const Real* const cy = reinterpret_cast<const Real*>(P + ocy);
        for (long long q = 0; q < C; ++q) {
            Lp[q] *= (double)kP[q];
            Lx[q] *= (double)cx[q];
            Ly[q] *= (double)cy[q];
        }
    }
    // tubes are independent: each reads the shared host copies and writes
    // only its own sections, so they are chained on up to 16 host threads
    auto chain_tube = [&](int k) {
        const TubeDesc& d = h->desc[k];
        const int H = d.H, W = d.W, HW = H * W;
        const long long base = d.cell_off;
        const Real* X = rawh + (size_t)k * maxHW;
        const Real* Y = rawh + (size_t)(maxHW * h->n) + (size_t)k * maxHW;
        const int mid = d.midY * W + d.te;
        const Real open = (Real)0.50f, mnR = (Real)0.001f;
        const bool depth3d = d.solver_type == 1;
        std::vector<double> gX(HW, 0.0), gY(HW, 0.0), gXs(HW, 0.0), gYs(HW, 0.0), gzc(W, 0.0);
        // smoothed maps, in the arithmetic of the setup kernel
        std::vector<Real> Xs(HW), Ys(HW);
        for (int i = 0; i < HW; ++i) {
            const int r = i / W, c = i - r * W;
            const Real mx = X[mid], my = Y[mid];
            Real xc = X[i]; if (xc == open) xc = mx;
            Real yc = Y[i]; if (yc == open) yc = my;
            Real xr = mx;
            if (c < W - 1) { xr = X[i + 1]; if (xr == open) xr = mx; }
            Real ya = my;
            if (r > 0) { ya = Y[i - W]; if (ya == open) ya = my; }
            Xs[i] = (Real)0.5f * (xc + xr);
            Ys[i] = (Real)0.5f * (yc + ya);
        }
        for (int r = 1; r < H - 1; ++r)
            for (int c = 1; c < W - 1; ++c) {
                const int i = r * W + c;
                const long long q = base + i;
                if (depth3d) {
                    // kP ~ 1 / dP, cx ~ dX, cy ~ dY
                    const double gdp = -Lp[q] / (double)dP[q];
                    const double gdx = Lx[q] / (double)dX[q];
                    const double gdy = Ly[q] / (double)dY[q];
                    if (!(Xs[i] < mnR)) gXs[i] += gdx;
                    if (!(Ys[i] < mnR)) gYs[i] += gdy;
                    const Real lX = Xs[i - 1], dYv = Ys[i + W];
                    const Real p = (Real)0.25f * (Xs[i] + Ys[i] + lX + dYv);
                    if (!(p < mnR)) {
                        const double g4 = 0.25 * gdp;
                        gXs[i] += g4; gYs[i] += g4; gXs[i - 1] += g4; gYs[i + W] += g4;
                    }
                }
                const int tc = cls[q] & kTypeMask;
                if ((cls[q] & kSegWallBit) && tc != kExcitation) {
                    const double z = (double)zcol[d.col_off + c];
                    // admittance faces: cx, cy ~ z
                    if (z != 0.0) gzc[c] += (Lx[q] + Ly[q]) / z;
                }
            }
        // half-cell smoothing, transposed (open-space cells read the mouth-end cell)
        for (int i = 0; i < HW && depth3d; ++i) {
            const int r = i / W, c = i - r * W;
            const double tx = 0.5 * gXs[i], ty = 0.5 * gYs[i];
            gX[X[i] == open ? mid : i] += tx;
            gX[(c < W - 1 && X[i + 1] != open) ? i + 1 : mid] += tx;
            gY[Y[i] == open ? mid : i] += ty;
            gY[(r > 0 && Y[i - W] != open) ? i - W : mid] += ty;
        }
        const long long so = d.sec_off;
        std::vector<double> gminor(d.nsec, 0.0), gsc(d.nsec, 0.0), gz(d.nsec, 0.0);
        for (int j = 0; j < d.L; ++j) {
            const int c = d.ts + j, sc = seg[d.wall_off + j];
            gz[sc] += gzc[c];
            if (!depth3d) continue;
            for (int r = upper[d.wall_off + j] + 1; r < lower[d.wall_off + j]; ++r) {
                const int i = r * W + c;
                const double v = gX[i] * (double)X[i] + gY[i] * (double)Y[i];
                gsc[sc] += v / (double)dscale[so + sc];
                gminor[sc] += v / (double)minor[so + sc];
            }
        }
        const double rho_c = d.ph.rho * d.ph.c;
        for (int s = 0; s < d.nsec; ++s) {
            // minor axis = 2 sqrt(A / (q pi)): d minor / d A = minor / (2 A); the
            // circular shape's depth does not read the minor axis
            g_area[so + s] = (depth3d && d.shape == 2) ? gminor[s] * (double)minor[so + s] / (2.0 * h->sec_area[so + s]) : 0.0;
            g_scale[so + s] = depth3d ? gsc[s] : 0.0;
            g_mu[so + s] = gz[s] * dz_dmu(h->sec_mu[so + s], rho_c);
        }
    };
    const int nthr = std::max(1, std::min({h->n, 16, (int)std::thread::hardware_concurrency()}));
    std::atomic<int> next{0};
    std::atomic<bool> failed{false};
    auto worker = [&] {
        try {
            for (int k; (k = next.fetch_add(1)) < h->n && !failed.load();) chain_tube(k);
        } catch (...) {
            failed.store(true);
        }
    };
    if (nthr == 1) {
        worker();
    } else {
        std::vector<std::thread> pool;
        for (int t = 0; t < nthr; ++t) pool.emplace_back(worker);
        for (std::thread& t : pool) t.join();
    }
    if (failed.load()) return fail(VS_E_STATE, "gradient chain: host memory allocation failed");
    const auto t_end = std::chrono::steady_clock::now();
    h->adj_reverse_ms = std::chrono::duration<double, std::milli>(t_device - t_start).count();
    h->adj_chain_ms = std::chrono::duration<double, std::milli>(t_end - t_device).count();
    return VS_OK;
}

}  // namespace

extern "C" {

int vs_adjoint_timing(const vs_handle* h, double* reverse_ms, double* chain_ms) {
    return guarded("vs_adjoint_timing", [&]() -> int {
        if (check_handle(h)) return VS_E_INVALID;
        if (reverse_ms) *reverse_ms = h->adj_reverse_ms;
        if (chain_ms) *chain_ms = h->adj_chain_ms;
        return VS_OK;
    });
}

int vs_adjoint_forward(vs_handle* h, int checkpoint_every, double* probe, size_t probe_len) {
    return guarded("vs_adjoint_forward", [&]() -> int {
        if (int rc = check_handle(h)) return rc;
        if (checkpoint_every < 1 || checkpoint_every > (1 << 20))
            return fail(VS_E_INVALID, "checkpoint_every must be in [1, 2^20]");
        return h->precision == VS_FP64 ? adj_forward<double>(h, checkpoint_every, probe, probe_len)
                                       : adj_forward<float>(h, checkpoint_every, probe, probe_len);
    });
}

int vs_adjoint_backward(vs_handle* h, const double* gprobe, size_t gprobe_len, double* g_area, double* g_depth_scale,
                        double* g_mu3d, size_t n_sections_total) {
    return guarded("vs_adjoint_backward", [&]() -> int {
        if (int rc = check_handle(h)) return rc;
        if (!g_area || !g_depth_scale || !g_mu3d) return fail(VS_E_INVALID, "null gradient buffer");
        return h->precision == VS_FP64
                   ? adj_backward<double>(h, gprobe, gprobe_len, g_area, g_depth_scale, g_mu3d, n_sections_total)
                   : adj_backward<float>(h, gprobe, gprobe_len, g_area, g_depth_scale, g_mu3d, n_sections_total);
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Audio export (results/generate_audio.py:9-42)
// ---------------------------------------------------------------------------
extern "C" {

int vs_butter2_lowpass(double cutoff_hz, double fs, double* b3, double* a3) {
    return guarded("vs_butter2_lowpass", [&]() -> int {
        if (!b3 || !a3) return fail(VS_E_INVALID, "null coefficient buffer");
        if (!(cutoff_hz > 0.0) || !(cutoff_hz < 0.5 * fs)) return fail(VS_E_INVALID, "cutoff must lie in (0, fs / 2)");
        const AudioFilter f = butter2_lowpass(cutoff_hz, fs);
        for (int k = 0; k < 3; ++k) { b3[k] = f.b[k]; a3[k] = f.a[k]; }
        return VS_OK;
    });
}

int vs_probe_audio(vs_handle* h, int tube, float* out, size_t cap, size_t* n_out) {
    return guarded("vs_probe_audio", [&]() -> int {
        if (int rc = check_handle(h)) return rc;
        if (tube < -1 || tube >= h->n) return fail(VS_E_RANGE, "tube index out of range");
        const int t0 = tube < 0 ? 0 : tube, t1 = tube < 0 ? h->n : tube + 1;
        std::vector<AudioJob> jobs;
        long long total = 0;
        for (int k = t0; k < t1; ++k) {
            const TubeDesc& d = h->desc[k];
            AudioJob J{};
            J.in_off = d.src_off;
            J.n = std::min(h->steps_done, d.num_steps);
            J.decim = d.srate_mul;   // fs = 44100 * srate_mul
            J.flt = butter2_lowpass(5050.0, 44100.0 * d.srate_mul);
            J.out_off = total;
            total += J.decim == 1 ? J.n : (J.n + J.decim - 1) / J.decim;
            jobs.push_back(J);
        }
        if (n_out) *n_out = (size_t)total;
        if (!out) return VS_OK;
        if (cap < (size_t)total) return fail(VS_E_RANGE, "audio buffer needs " + std::to_string(total) + " samples");
        if (total == 0) return VS_OK;
        AudioJob* dj = nullptr;
        float* dout = nullptr;
        VS_HIP(hipMalloc(&dj, jobs.size() * sizeof(AudioJob)));
        hipError_t e = hipMalloc(&dout, (size_t)total * sizeof(float));
        if (e == hipSuccess) e = hipMemcpyAsync(dj, jobs.data(), jobs.size() * sizeof(AudioJob), hipMemcpyHostToDevice, h->stream);
        const double* probe = h->precision == 64 ? h->bd.probe : h->bf.probe;
        if (e == hipSuccess) e = launch_audio(probe, dj, (int)jobs.size(), dout, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e == hipSuccess) e = copy_on(h->stream, out, dout, (size_t)total * sizeof(float), hipMemcpyDeviceToHost);
        (void)hipFree(dj);
        if (dout) (void)hipFree(dout);
        if (e != hipSuccess) return fail(VS_E_DEVICE, std::string("audio: ") + hipGetErrorString(e));
        return VS_OK;
    });
}

int vs_write_wav(const char* path, const float* samples, size_t n, int rate) {
    return guarded("vs_write_wav", [&]() -> int {
        if (!path || (!samples && n)) return fail(VS_E_INVALID, "vs_write_wav: bad argument");
        if (rate <= 0) return fail(VS_E_INVALID, "sample rate must be positive");
        FILE* f = std::fopen(path, "wb");
        if (!f) return fail(VS_E_INVALID, std::string("cannot open ") + path);
        auto u32 = [&](uint32_t v) { unsigned char b[4] = {(unsigned char)v, (unsigned char)(v >> 8), (unsigned char)(v >> 16), (unsigned char)(v >> 24)}; std::fwrite(b, 1, 4, f); };
        auto u16 = [&](uint16_t v) { unsigned char b[2] = {(unsigned char)v, (unsigned char)(v >> 8)}; std::fwrite(b, 1, 2, f); };
        const uint32_t data_bytes = (uint32_t)(n * 4);
        // RIFF / WAVE with an 18-byte fmt chunk (IEEE float, cbSize 0) and a fact
        // chunk, the layout scipy.io.wavfile.write uses for float32 data
        std::fwrite("RIFF", 1, 4, f);
        u32(4 + (8 + 18) + (8 + 4) + (8 + data_bytes));
        std::fwrite("WAVE", 1, 4, f);
        std::fwrite("fmt ", 1, 4, f);
        u32(18);
        u16(3);                    // WAVE_FORMAT_IEEE_FLOAT
        u16(1);                    // channels
        u32((uint32_t)rate);
        u32((uint32_t)rate * 4);   // bytes per second
        u16(4);                    // block align
        u16(32);                   // bits per sample
        u16(0);                    // cbSize
        std::fwrite("fact", 1, 4, f);
        u32(4);
        u32((uint32_t)n);
        std::fwrite("data", 1, 4, f);
        u32(data_bytes);
        for (size_t i = 0; i < n; ++i) {
            uint32_t v;
            std::memcpy(&v, &samples[i], 4);
            u32(v);
        }
        const bool ok = std::fclose(f) == 0;
        return ok ? VS_OK : fail(VS_E_INVALID, std::string("write failed: ") + path);
    });
}

}  // extern "C"
