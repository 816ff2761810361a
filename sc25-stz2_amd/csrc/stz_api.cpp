// stz_api.cpp -- C ABI of the STZ device path (include/stz.h): context,
// memory pool, the compress / decompress pipelines, the host zstd container
// and the stage entry points.
//
// F5 = src/zero_all_line_dynamic_random_access_optfast3_optcr_lb_fastdec5.cpp
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <chrono>
#include <sched.h>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <unordered_map>
#include <vector>

#include "host_codec.h"
#include "kernels.h"
#include "layout.h"
#include "stz.h"

extern "C" {  // libzstd (host), the five entry points used
size_t ZSTD_compress(void* dst, size_t cap, const void* src, size_t n, int level);
size_t ZSTD_decompress(void* dst, size_t cap, const void* src, size_t n);
size_t ZSTD_compressBound(size_t n);
unsigned ZSTD_isError(size_t code);
unsigned long long ZSTD_getFrameContentSize(const void* src, size_t n);
}

using namespace stz;

namespace {

struct StzError : std::runtime_error {
    int code;
    StzError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                       \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
        if (e_ != hipSuccess) throw StzError(STZ_E_HIP, std::string(#x) + ": " +       \
                                                            hipGetErrorString(e_));     \
    } while (0)

inline uint64_t align8(uint64_t v) { return (v + 7) & ~7ull; }
inline uint64_t align_up(uint64_t v, uint64_t a) { return (v + a - 1) / a * a; }

// ---------------------------------------------------------------- memory pool
struct Pool {
    std::mutex mu;
    std::multimap<size_t, void*> free_;
    std::unordered_map<void*, size_t> size_;
    int device = 0;
    ~Pool()
    {
        int cur = 0;
        (void)hipGetDevice(&cur);
        (void)hipSetDevice(device);
        for (auto& kv : free_) (void)hipFree(kv.second);
        (void)hipSetDevice(cur);
    }
    void* get(size_t n)
    {
        if (n == 0) n = 256;
        n = align_up(n, 256);
        {
            std::lock_guard<std::mutex> g(mu);
            auto it = free_.lower_bound(n);
            if (it != free_.end() && it->first <= 2 * n + (64u << 20)) {
                void* p = it->second;
                free_.erase(it);
                return p;
            }
        }
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            release_all();
            e = hipMalloc(&p, n);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                throw StzError(STZ_E_OOM, "hipMalloc failed");
            }
        }
        std::lock_guard<std::mutex> g(mu);
        size_[p] = n;
        return p;
    }
    void put(void* p)
    {
        if (!p) return;
        std::lock_guard<std::mutex> g(mu);
        free_.insert({size_[p], p});
    }
    void release_all()
    {
        std::lock_guard<std::mutex> g(mu);
        for (auto& kv : free_) {
            size_.erase(kv.second);
            (void)hipFree(kv.second);
        }
        free_.clear();
    }
};

struct DBuf {
    std::shared_ptr<Pool> pool;
    void* p = nullptr;
    DBuf() = default;
    DBuf(std::shared_ptr<Pool> pl, size_t n) : pool(std::move(pl)), p(pool->get(n)) {}
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : pool(std::move(o.pool)), p(o.p) { o.p = nullptr; }
    DBuf& operator=(DBuf&& o) noexcept
    {
        reset();
        pool = std::move(o.pool);
        p = o.p;
        o.p = nullptr;
        return *this;
    }
    ~DBuf() { reset(); }
    void reset()
    {
        if (p && pool) pool->put(p);
        p = nullptr;
    }
    template <typename U> U* as() const { return (U*)p; }
};

struct ProfRec { const char* name; hipEvent_t a, b; };
struct BatchPlan;

// Persistent host workers for the per-stream host work of a call (codebook
// construction): run(n, f) calls f(0..n-1) on the workers and the calling
// thread and returns when the n calls are done (not when every worker has
// woken up: a worker that wakes late finds the tasks taken), rethrowing the
// first exception.  Each run has its own record, so a late worker never
// takes a task of a later run.
class TaskPool {
public:
    explicit TaskPool(int nworkers)
    {
        for (int i = 0; i < nworkers; ++i) th_.emplace_back([this] { loop(); });
    }
    ~TaskPool()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
            stop_flag_.store(true);
        }
        cv_.notify_all();
        for (auto& t : th_) t.join();
    }
    void run(int n, const std::function<void(int)>& f)
    {
        if (n <= 1 || th_.empty()) {
            for (int i = 0; i < n; ++i) f(i);
            return;
        }
        auto job = std::make_shared<Job>();
        job->f = &f;
        job->n = n;
        {
            std::lock_guard<std::mutex> lk(mu_);
            cur_ = job;
            gen_.store(gen_.load() + 1);
        }
        cv_.notify_all();
        work(*job);
        {
            std::unique_lock<std::mutex> lk(mu_);
            done_.wait(lk, [&] { return job->finished.load() == n; });
            if (cur_ == job) cur_.reset();
        }
        if (job->err) std::rethrow_exception(job->err);
    }

private:
    struct Job {
        const std::function<void(int)>* f = nullptr;  // valid while tasks remain unclaimed
        int n = 0;
        std::atomic<int> next{0}, finished{0};
        std::exception_ptr err;  // written under the pool's mutex
    };
    void work(Job& j)
    {
        for (;;) {
            const int i = j.next.fetch_add(1);
            if (i >= j.n) return;
            try {
                (*j.f)(i);
            } catch (...) {
                std::lock_guard<std::mutex> lk(mu_);
                if (!j.err) j.err = std::current_exception();
            }
            if (j.finished.fetch_add(1) + 1 == j.n) {
                std::lock_guard<std::mutex> lk(mu_);
                done_.notify_all();
            }
        }
    }
    void loop()
    {
        uint64_t seen = 0;
        for (;;) {
            // a compress calls run() twice about half a millisecond apart
            // (base/low codebooks, then high): spin up to 1 ms on the
            // generation before sleeping, so the second call finds the
            // workers awake
            const auto t0 = std::chrono::steady_clock::now();
            while (gen_.load(std::memory_order_acquire) == seen && !stop_flag_.load(std::memory_order_relaxed) &&
                   std::chrono::steady_clock::now() - t0 < std::chrono::milliseconds(1))
                std::this_thread::yield();
            std::shared_ptr<Job> j;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || gen_.load() != seen; });
                if (stop_) return;
                seen = gen_.load();
                j = cur_;
            }
            if (j) work(*j);
        }
    }
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    std::shared_ptr<Job> cur_;
    std::atomic<uint64_t> gen_{0};  // written under mu_, read without it while spinning
    bool stop_ = false;
    std::atomic<bool> stop_flag_{false};
};

}  // namespace

// Waits for an event by polling it: the blocking wait sleeps and its wake-up
// costs tens of microseconds on a path where the host's next launches are
// what the GPU waits for (a compress waits twice for small readbacks); after
// about 20 ms of polling it blocks instead.
static void wait_event(hipEvent_t e)
{
    const auto t0 = std::chrono::steady_clock::now();
    for (int k = 0;; ++k) {
        const hipError_t r = hipEventQuery(e);
        if (r == hipSuccess) return;
        if (r != hipErrorNotReady) HIPCHK(r);
        if ((k & 255) == 255 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(20)) {
            HIPCHK(hipEventSynchronize(e));
            return;
        }
    }
}

// A stream whose CU mask covers every CU of the device (current): HIP gives
// a CU-masked stream a hardware queue of its own instead of sharing one of
// the GPU_MAX_HW_QUEUES queues with other streams, where work runs in order.
static hipStream_t full_mask_stream(int device)
{
    int n_cu = 0;
    HIPCHK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, device));
    std::vector<uint32_t> mask((size_t)(n_cu + 31) / 32, 0xFFFFFFFFu);
    if (n_cu % 32) mask.back() = (1u << (n_cu % 32)) - 1u;
    hipStream_t s = nullptr;
    HIPCHK(hipExtStreamCreateWithCUMask(&s, (uint32_t)mask.size(), mask.data()));
    return s;
}

struct stz_ctx {
    int device = 0;
    // this device's copy of the allocation-free entry point's scratch array
    void* scratch = nullptr;
    hipStream_t own = nullptr, stream = nullptr;
    // marks a mid-compress readback, so host codebook builds start while
    // the stream still runs the later kernels
    hipEvent_t ev_fork = nullptr;
    // the value range of a compress has reached the host
    hipEvent_t ev_range = nullptr;
    // a compress's histogram readback has landed
    hipEvent_t ev_hist = nullptr;
    // second stream of a full decode (created on first use): the high
    // streams' Huffman decode runs there while the base and low levels
    // decode on `stream`; ev_items / ev_high order it
    hipStream_t side_stream = nullptr;
    hipEvent_t ev_items = nullptr, ev_high = nullptr;
    // STZ_CTX_OWN_QUEUE: the caller-facing stream is created with a CU mask
    // over every CU, which gives it its own hardware queue.  The second
    // stream always gets one: with HIP's default of 4 queues it otherwise
    // shares a queue with the caller's stream, and the work meant to
    // overlap runs in order behind it (512^3 decompress 0.77 -> 0.98 ms)
    bool own_queue = false;
    hipStream_t second()
    {
        if (!side_stream) {
            side_stream = full_mask_stream(device);
            HIPCHK(hipEventCreateWithFlags(&ev_items, hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&ev_high, hipEventDisableTiming));
        }
        return side_stream;
    }
    std::shared_ptr<Pool> pool;
    bool prof = false;
    std::vector<ProfRec> recs;
    std::map<std::string, std::pair<long, double>> agg;

    struct Arena {
        void* pinned = nullptr;
        size_t cap = 0, off = 0;
        std::vector<void*> retired;  // outgrown staging buffers, freed when the arena is reused
        hipEvent_t ev = nullptr;     // the stream position after the call that used it
        bool pending = false;
        hipStream_t stream = nullptr;  // the stream its copies went to
    };
    Arena arena[2];
    int slot = 0;
    // launch geometry of the multi-field level kernels, per (dtype, mode,
    // size class), chosen by timing the candidates once
    struct LaunchCfg { int waves = 4; uint32_t grid = 0; float ms = 0.f; std::string trials; };
    std::map<std::string, LaunchCfg> tuned;
    bool autotune = true;
    int n_cu = 0;
    // fraction of the resident workgroup slots the persistent full-level
    // decode kernels take (stz_set_grid_share)
    float grid_share = 1.0f;
    // captured batch decode (graph + the buffers it reads and writes)
    std::unique_ptr<BatchPlan> plan;
    // learned tuner model (fp32 MLP parameters on the device)
    DBuf tuner_w;
    bool have_tuner = false;
    float tune_min_conf = 0.f;
    // host workers (created on first use): three for a call's few streams
    // (more only cost wake-ups there), and a wide pool for a batch's many
    // (96 fields: 1440 codebooks, 11-15 ms on four threads), sized to the
    // CPUs this process may run on, at most 16
    std::unique_ptr<TaskPool> workers, workers_wide;
    TaskPool& tasks()
    {
        if (!workers) {
            const unsigned hc = std::thread::hardware_concurrency();
            workers.reset(new TaskPool((int)std::min<unsigned>(hc > 1 ? hc - 1 : 0, 3)));
        }
        return *workers;
    }
    TaskPool& tasks_wide()
    {
        if (!workers_wide) {
            cpu_set_t set;
            int cpus = 0;
            if (sched_getaffinity(0, sizeof(set), &set) == 0) cpus = CPU_COUNT(&set);
            if (cpus < 1) cpus = (int)std::thread::hardware_concurrency();
            workers_wide.reset(new TaskPool(std::max(0, std::min(cpus, 16) - 1)));
        }
        return *workers_wide;
    }

    // Pinned staging for async copies: each request gets a region of its own
    // for the rest of the API call (a bump arena reset when a call starts).
    // Two arenas alternate between calls: a call starts by marking the
    // stream position after the previous call's copies and waits only for
    // the call before that one, so the host may run one call ahead of the
    // GPU (a compress returns while its bit pack still runs, and the next
    // call's host work overlaps it).
    void* host_stage(size_t n)
    {
        Arena& A = arena[slot];
        n = align_up(std::max<size_t>(n, 1), 256);
        if (A.off + n > A.cap) {
            if (A.pinned) A.retired.push_back(A.pinned);
            A.pinned = nullptr;
            A.cap = 0;
            A.off = 0;
            const size_t cap = align_up(std::max<size_t>(2 * n, 8u << 20), 1u << 20);
            HIPCHK(hipHostMalloc(&A.pinned, cap, hipHostMallocDefault));
            A.cap = cap;
        }
        void* r = (uint8_t*)A.pinned + A.off;
        A.off += n;
        return r;
    }
    // start of an API call
    void stage_reset()
    {
        Arena& cur = arena[slot];
        if (cur.off == 0 && cur.retired.empty()) return;
        if (!cur.ev) HIPCHK(hipEventCreateWithFlags(&cur.ev, hipEventDisableTiming));
        HIPCHK(hipEventRecord(cur.ev, cur.stream ? cur.stream : stream));
        cur.pending = true;
        slot ^= 1;
        Arena& nxt = arena[slot];
        if (nxt.pending) {
            wait_event(nxt.ev);
            nxt.pending = false;
        }
        for (void* q : nxt.retired) (void)hipHostFree(q);
        nxt.retired.clear();
        nxt.off = 0;
        nxt.stream = stream;
    }
    // a call that moves to another stream first lets the staged copies of
    // the calls before it finish on the old one
    void stage_drain()
    {
        for (Arena& A : arena)
            if (A.pending) {
                wait_event(A.ev);
                A.pending = false;
            }
        if (arena[slot].off || !arena[slot].retired.empty()) HIPCHK(hipStreamSynchronize(arena[slot].stream ? arena[slot].stream : stream));
        for (Arena& A : arena) {
            for (void* q : A.retired) (void)hipHostFree(q);
            A.retired.clear();
        }
        arena[slot].off = 0;
    }
};

// host-side metadata of a payload
struct Meta {
    FileHeader fh{};
    StreamHeader sh[kStreams]{};
    Codebook cb[kStreams];
};

struct stz_dstream {
    int device = 0;
    std::shared_ptr<Pool> pool;
    DBuf payload;
    size_t bytes = 0;
    Meta meta;
    // decode tables (built on first decode)
    DBuf tabs, sorted;
    bool have_tabs = false;
    // decoded base level kept for random access (stz_dstream_prepare)
    DBuf base;
    bool have_base = false;
    // how the configuration was chosen (0 given, 1 learned, 2 trial) and the
    // model's confidence when it was asked
    int tuned = 0;
    float tune_conf = 0.f;
    // identity for cached launch plans (addresses can be reused)
    uint64_t gen = next_gen();
    static uint64_t next_gen()
    {
        static std::atomic<uint64_t> g{1};
        return g.fetch_add(1);
    }
};

namespace {

template <typename F>
void K(stz_ctx* c, const char* name, F&& f)
{
    if (!c->prof) {
        f();
    } else {
        ProfRec r{name, nullptr, nullptr};
        HIPCHK(hipEventCreate(&r.a));
        HIPCHK(hipEventCreate(&r.b));
        HIPCHK(hipEventRecord(r.a, c->stream));
        f();
        HIPCHK(hipEventRecord(r.b, c->stream));
        c->recs.push_back(r);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw StzError(STZ_E_HIP, std::string(name) + ": " + hipGetErrorString(e));
}

void resolve_prof(stz_ctx* c)
{
    if (c->recs.empty()) return;
    HIPCHK(hipStreamSynchronize(c->stream));
    for (auto& r : c->recs) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
        auto& a = c->agg[r.name];
        a.first += 1;
        a.second += ms;
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    c->recs.clear();
}

bool is_device_ptr(const void* p)
{
    if (!p) return false;
    hipPointerAttribute_t a;
    const hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged || a.isManaged;
}

size_t tsize_of(int dtype) { return dtype == STZ_F32 ? 4 : 8; }
size_t pad4(size_t n) { size_t p = (n + 3) & ~(size_t)3; return p < 4 ? 4 : p; }

// ------------------------------------------------------------ geometry
// sub-chunk index of every chunk of one stream grid (kSub symbols each, a
// chunk's last one partial): chunk_sub[c] .. chunk_sub[c+1]-1
uint32_t chunk_subs(const Geom& g, uint64_t n, std::vector<uint32_t>& chunk_sub)
{
    chunk_sub.resize(g.n_chunks + 1);
    uint32_t sub = 0;
    for (uint32_t c = 0; c < g.n_chunks; ++c) {
        chunk_sub[c] = sub;
        sub += (uint32_t)((chunk_volume(g, c, n) + kSub - 1) / kSub);
    }
    chunk_sub[g.n_chunks] = sub;
    return sub;
}

struct HostLayout {
    size_t nx, ny, nz, px, py, pz;
    int chunk_high, chunk_low;
    Geom g[kStreams];
    uint64_t n[kStreams];
    uint32_t n_sub[kStreams];
    std::vector<uint32_t> chunk_sub[kStreams];
    uint64_t code_off[kStreams];
    uint64_t total_codes;
};

// interp: the base codec (a Lorenzo/regression base, 3, adds its tiles'
// header symbols to the base stream)
void build_layout(HostLayout& L, size_t nx, size_t ny, size_t nz, size_t px, size_t py, size_t pz, int ch, int cl,
                  int interp)
{
    L.nx = nx; L.ny = ny; L.nz = nz;
    L.px = px; L.py = py; L.pz = pz;
    L.chunk_high = ch; L.chunk_low = cl;
    const int lx = (int)(px / 4), ly = (int)(py / 4), lz = (int)(pz / 4);
    const int mx = (int)(px / 2), my = (int)(py / 2), mz = (int)(pz / 2);
    uint64_t off = 0;
    for (int s = 0; s < kStreams; ++s) {
        if (s == 0) L.g[s] = make_geom(lx, ly, lz, 0, base_extra_symbols(interp, lx, ly, lz));
        else if (s < 8) L.g[s] = make_geom(lx, ly, lz, cl);
        else L.g[s] = make_geom(mx, my, mz, ch);
        const Geom& g = L.g[s];
        L.n[s] = (uint64_t)g.dx * g.dy * g.dz + (s == 0 ? base_extra_symbols(interp, g.dx, g.dy, g.dz) : 0);
        L.n_sub[s] = chunk_subs(g, L.n[s], L.chunk_sub[s]);
        L.code_off[s] = off;
        off += align_up(L.n[s], 64);
    }
    L.total_codes = off;
}

// scratch bytes of the staged predictors' reconstructions of blocks 0..5 of
// either residual level (the high level's block grid is the larger)
inline size_t staged_scratch(const HostLayout& L, size_t elem)
{
    return 6 * (size_t)L.n[8] * elem;
}

// Layouts by shape: building one walks every chunk of every stream (tens of
// microseconds at 512^3), and fields of one shape come back again and again
// (steps, serving, batches), so the process keeps the recent ones.
void make_layout(HostLayout& L, size_t nx, size_t ny, size_t nz, size_t px, size_t py, size_t pz, int ch, int cl,
                 int interp)
{
    typedef std::array<uint64_t, 9> Key;
    static std::mutex mu;
    static std::map<Key, std::shared_ptr<const HostLayout>> cache;
    const uint64_t ib = interp == 3 ? 1 : 0;  // the only base codec with a layout of its own
    const Key k = {nx, ny, nz, px, py, pz, (uint64_t)(uint32_t)ch, (uint64_t)(uint32_t)cl, ib};
    std::shared_ptr<const HostLayout> hit;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(k);
        if (it != cache.end()) hit = it->second;
    }
    if (hit) {
        L = *hit;
        return;
    }
    auto fresh = std::make_shared<HostLayout>();
    build_layout(*fresh, nx, ny, nz, px, py, pz, ch, cl, interp);
    L = *fresh;
    std::lock_guard<std::mutex> lk(mu);
    if (cache.size() >= 64) cache.clear();
    cache.emplace(k, std::move(fresh));
}

void check_dims(size_t nx, size_t ny, size_t nz, int ch, int cl)
{
    if (nx == 0 || ny == 0 || nz == 0) throw StzError(STZ_E_DIMS, "zero dimension");
    if (ch <= 0 || cl <= 0 || ch > 64 || cl > 64) throw StzError(STZ_E_ARG, "chunk dims must be in 1..64");
    const size_t px = pad4(nx), py = pad4(ny), pz = pad4(nz);
    if ((px / 2) * (py / 2) * (pz / 2) >= (1ull << 31) || px * py >= (1ull << 31))
        throw StzError(STZ_E_DIMS, "field too large for one stream grid");
}

// ABS / REL: the requested bound E is the maximum pointwise error.  The
// reconstruction rounds twice in the field's type (the residual
// fl(f - pred) and the value fl(dec + pred)), and the quantizer's check
// |dec - d| <= eb sees a difference rounded once more: together at most
// 3.25 u (M + E), u the unit roundoff, M = max |value| (predictions are
// weighted means of reconstructed values with weights summing to 1 and
// absolute sum <= 1.25).  The levels get E - 4 u (M + E), so the error is
// strictly at most E; a bound below that slack leaves 0 (every value
// stored).  A Lorenzo prediction sums seven reconstructed values
// (|prediction| <= 7 (M + E)), so the Lorenzo bases (interp 2 and 3, whose
// regression predictions are clamped to the value range) take
// E - 10 u (M + E).  REF keeps F5's meaning: level bounds are mult_* x eb.
void level_ebs(double eb, int mode, double vmin, double vmax, int dtype, const stz_params& p, double ebl[3],
               double& unit)
{
    if (mode == STZ_EB_REF) {
        unit = eb;
    } else {
        const double E = (mode == STZ_EB_REL) ? eb * (vmax - vmin) : eb;
        const double M = std::max(std::fabs(vmin), std::fabs(vmax));
        const double u = dtype == STZ_F32 ? 0x1p-24 : 0x1p-53;
        double Es = E - (p.interp >= 2 ? 10.0 : 4.0) * u * (M + E);
        if (!(Es > 0)) Es = 0;
        double mm = p.mult_base;
        if (p.mult_low > mm) mm = p.mult_low;
        if (p.mult_high > mm) mm = p.mult_high;
        unit = Es / mm;
    }
    ebl[0] = p.mult_base * unit;
    ebl[1] = p.mult_low * unit;
    ebl[2] = p.mult_high * unit;
}

// section offsets of one stream, the same formula the payload reader uses;
// kind 0 base, 1 low, 2 high (payload streams), 3 a standalone block stream
void fill_stream_header_g(StreamHeader& sh, uint32_t kind, uint32_t block, const Geom& g, uint64_t n, uint32_t n_sub,
                          int dtype, double eb, uint32_t n_used, uint64_t n_unpred, uint64_t total_bits)
{
    std::memset(&sh, 0, sizeof(sh));
    sh.magic = kStreamMagic;
    sh.kind = kind;
    sh.block = block;
    sh.dtype = (uint32_t)dtype;
    sh.n_symbols = n;
    sh.dx = (uint32_t)g.dx; sh.dy = (uint32_t)g.dy; sh.dz = (uint32_t)g.dz;
    sh.chunk = g.chunk > 0 ? (uint32_t)g.chunk : (uint32_t)kBaseChunk;
    sh.n_chunks = g.n_chunks;
    sh.n_sub = n_sub;
    sh.n_used = n_used;
    sh.lmax = kLmax;
    sh.n_unpred = n_unpred;
    sh.total_bits = total_bits;
    sh.eb = eb;
    sh.radius = kRadius;
    sh.sub_size = kSub;
    const uint64_t ts = tsize_of(dtype);
    uint64_t off = sizeof(StreamHeader);
    sh.off_codebook = off;     off += align8(4ull * n_used);
    sh.off_chunk_bit = off;    off += 8ull * (sh.n_chunks + 1);
    sh.off_chunk_unpred = off; off += 8ull * (sh.n_chunks + 1);
    sh.off_chunk_sub = off;    off += align8(4ull * (sh.n_chunks + 1));
    sh.off_sub_bits = off;     off += align8(2ull * sh.n_sub);
    sh.off_sub_unpred = off;   off += align8(2ull * sh.n_sub);
    sh.off_unpred = off;       off += align8(ts * n_unpred);
    sh.off_bits = off;         off += align8(4ull * ((total_bits + 31) / 32 + 2));
    sh.stream_bytes = off;
}

void fill_stream_header(StreamHeader& sh, int s, const HostLayout& L, int dtype, double eb, uint32_t n_used,
                        uint64_t n_unpred, uint64_t total_bits)
{
    fill_stream_header_g(sh, s == 0 ? 0u : s < 8 ? 1u : 2u, s == 0 ? 0u : (uint32_t)((s - 1) % 7), L.g[s], L.n[s],
                         L.n_sub[s], dtype, eb, n_used, n_unpred, total_bits);
}

// One stream of `avail` bytes at p checked against what its geometry
// implies (header fields, every section offset, the codebook's order and
// Kraft sum, the chunk and sub-chunk tables' monotone totals); the header
// and codebook are returned.  Every read the decode kernels make goes
// through offsets checked here.
void check_stream(const uint8_t* p, uint64_t avail, uint32_t kind, uint32_t block, const Geom& g, uint64_t n,
                  const std::vector<uint32_t>& chunk_sub, int dtype, StreamHeader& sh, Codebook& b)
{
    if (avail < sizeof(StreamHeader)) throw StzError(STZ_E_STREAM, "short stream");
    std::memcpy(&sh, p, sizeof(StreamHeader));
    const uint32_t n_sub = chunk_sub[g.n_chunks];
    // counts bounded before they size the sections (no wrap in the offsets)
    if (sh.n_used > 65536 || sh.n_unpred > n || sh.total_bits > n * kLmax)
        throw StzError(STZ_E_STREAM, "bad stream header");
    StreamHeader chk;
    fill_stream_header_g(chk, kind, block, g, n, n_sub, dtype, sh.eb, sh.n_used, sh.n_unpred, sh.total_bits);
    if (sh.magic != kStreamMagic || sh.n_symbols != n || sh.n_chunks != g.n_chunks || sh.n_sub != chk.n_sub ||
        sh.stream_bytes != chk.stream_bytes || sh.off_codebook != chk.off_codebook ||
        sh.off_chunk_bit != chk.off_chunk_bit || sh.off_chunk_unpred != chk.off_chunk_unpred ||
        sh.off_chunk_sub != chk.off_chunk_sub || sh.off_sub_bits != chk.off_sub_bits ||
        sh.off_sub_unpred != chk.off_sub_unpred || sh.off_unpred != chk.off_unpred || sh.off_bits != chk.off_bits ||
        sh.kind != chk.kind || sh.block != chk.block || sh.dtype != chk.dtype || sh.dx != chk.dx ||
        sh.dy != chk.dy || sh.dz != chk.dz || sh.chunk != chk.chunk || sh.lmax != (uint32_t)kLmax ||
        sh.radius != (uint32_t)kRadius || sh.sub_size != (uint32_t)kSub || sh.stream_bytes > avail)
        throw StzError(STZ_E_STREAM, "bad stream header");
    const uint32_t* cb = (const uint32_t*)(p + sh.off_codebook);
    b.sym.resize(sh.n_used);
    b.len.resize(sh.n_used);
    double kraft = 0;
    for (uint32_t i = 0; i < sh.n_used; ++i) {
        b.sym[i] = (uint16_t)(cb[i] & 0xFFFF);
        b.len[i] = (uint8_t)(cb[i] >> 16);
        if (b.len[i] == 0 || b.len[i] > kLmax || (i && b.sym[i] <= b.sym[i - 1]))
            throw StzError(STZ_E_STREAM, "bad codebook");
        kraft += std::ldexp(1.0, -b.len[i]);
    }
    if (kraft > 1.0 + 1e-12) throw StzError(STZ_E_STREAM, "codebook violates Kraft");
    // chunk tables: geometry-derived sub-chunk map, monotone offsets that
    // add up to the sub-chunk sizes (they steer every decode read)
    const uint64_t* cbit = (const uint64_t*)(p + sh.off_chunk_bit);
    const uint64_t* cun = (const uint64_t*)(p + sh.off_chunk_unpred);
    const uint32_t* csub = (const uint32_t*)(p + sh.off_chunk_sub);
    const uint16_t* sbits = (const uint16_t*)(p + sh.off_sub_bits);
    const uint16_t* sun = (const uint16_t*)(p + sh.off_sub_unpred);
    uint64_t ab = 0, au = 0;
    for (uint32_t c = 0; c < sh.n_chunks; ++c) {
        if (csub[c] != chunk_sub[c] || cbit[c] != ab || cun[c] != au) throw StzError(STZ_E_STREAM, "bad chunk table");
        for (uint32_t j = csub[c]; j < chunk_sub[c + 1]; ++j) {
            if (sbits[j] > (uint32_t)kSub * kLmax || sun[j] > (uint32_t)kSub)
                throw StzError(STZ_E_STREAM, "bad sub-chunk table");
            ab += sbits[j];
            au += sun[j];
        }
    }
    if (csub[sh.n_chunks] != chunk_sub[sh.n_chunks] || cbit[sh.n_chunks] != ab || ab != sh.total_bits ||
        cun[sh.n_chunks] != au || au != sh.n_unpred)
        throw StzError(STZ_E_STREAM, "bad chunk table totals");
}

StreamSet make_streamset(const HostLayout& L, const Meta& m, int dtype)
{
    StreamSet ss{};
    ss.nstreams = kStreams;
    uint32_t cb = 0, sb = 0;
    for (int s = 0; s < kStreams; ++s) {
        StreamDesc& d = ss.s[s];
        const StreamHeader& sh = m.sh[s];
        const uint64_t so = m.fh.stream_off[s];
        d.code_off = L.code_off[s];
        d.n_symbols = L.n[s];
        d.off_chunk_bit = so + sh.off_chunk_bit;
        d.off_chunk_unpred = so + sh.off_chunk_unpred;
        d.off_chunk_sub = so + sh.off_chunk_sub;
        d.off_sub_bits = so + sh.off_sub_bits;
        d.off_sub_unpred = so + sh.off_sub_unpred;
        d.off_unpred = so + sh.off_unpred;
        d.off_bits = so + sh.off_bits;
        d.bits_words = (sh.stream_bytes - sh.off_bits) / 4;
        d.g = L.g[s];
        d.n_chunks = L.g[s].n_chunks;
        d.n_sub = L.n_sub[s];
        d.chunk_base = cb;
        d.sub_base = sb;
        d.tsize = (uint32_t)tsize_of(dtype);
        d.n_unpred = sh.n_unpred;
        cb += d.n_chunks;
        sb += d.n_sub;
    }
    ss.total_chunks = cb;
    ss.total_sub = sb;
    return ss;
}

// base-codec schedule: the coarse passes (stride >= S0, S0 the smallest
// power of two whose lattice fits the LDS budget and kFusedPoints) run in the fused
// single-workgroup kernel on an LDS copy of the lattice; every finer pass gets
// a launch of its own.  Stream order is unchanged.
template <typename FF, typename FP>
void run_passes(const std::vector<InterpPass>& passes, size_t tsize, FF&& fused, FP&& single, bool many = false)
{
    if (passes.empty()) return;
    const int bx = passes[0].bx, by = passes[0].by, bz = passes[0].bz;
    // many fields: one fused workgroup per field runs side by side, so each
    // keeps the whole LDS budget; one field: at most kFusedPoints
    const size_t cap = many ? kLatticeBytes / tsize : std::min<size_t>(kLatticeBytes / tsize, (size_t)kFusedPoints);
    int S0 = 1;
    auto lat = [&](int S, int d) { return (d - 1) / S + 1; };
    while ((size_t)lat(S0, bx) * lat(S0, by) * lat(S0, bz) > cap) S0 <<= 1;
    InterpPassList fl{};
    fl.S0 = S0;
    fl.bx = bx; fl.by = by; fl.bz = bz;
    fl.lx = lat(S0, bx); fl.ly = lat(S0, by); fl.lz = lat(S0, bz);
    size_t i = 0;
    for (; i < passes.size(); ++i) {
        const InterpPass& p = passes[i];
        if (p.axis >= 0 && p.s < S0) break;
        if (fl.n == kMaxFusedPasses) break;
        InterpPass q = p;
        q.bx = fl.lx; q.by = fl.ly; q.bz = fl.lz;
        for (int a = 0; a < 3; ++a) {
            q.start[a] = p.start[a] / S0;
            q.step[a] = p.step[a] >= S0 ? p.step[a] / S0 : 1;
        }
        q.s = p.axis >= 0 ? p.s / S0 : 0;
        fl.p[fl.n++] = q;
    }
    fused(fl);
    for (; i < passes.size(); ++i) single(passes[i]);
}

// ------------------------------------------------------------ compress
// One field's compress state through the phases.  A single compress runs
// the phases back to back; a batch runs each phase over every field, with
// one host synchronisation between phases and the two residual levels of all
// fields in one persistent launch each.  Every phase enqueues on the
// context's stream without waiting, except where marked.
template <typename T>
struct FieldEnc {
    int dtype = 0, mode = 0;
    double eb = 0;
    stz_params prm{};
    size_t nx = 0, ny = 0, nz = 0, px = 0, py = 0, pz = 0;
    uint64_t nin = 0;
    HostLayout L;
    DBuf in_copy, padded, mm_part;
    const T* din = nullptr;
    const T* full = nullptr;
    double* mm_host = nullptr;  // pinned: min, max
    double ebl[3] = {0, 0, 0}, unit = 0, vmin = 0, vmax = 0;
    DBuf codes, side, hist, vbase, mid, compact, counts;
    uint8_t* hs = nullptr;      // pinned: stream symbol counts + first entries
    std::vector<std::vector<uint64_t>> lists;
    size_t enc_n = 0;
    stz_dstream* ds = nullptr;
    static constexpr uint32_t kFirst = 4096;
};

// checks, device input, padding; the value range is requested (no wait)
template <typename T>
void enc_phase0(stz_ctx* c, FieldEnc<T>& fe, const T* field)
{
    const stz_params& prm = fe.prm;
    check_dims(fe.nx, fe.ny, fe.nz, prm.chunk_high, prm.chunk_low);
    if (!(fe.eb >= 0) || prm.alpha < 1.0 || prm.beta < 1.0 || prm.interp < 0 || prm.interp > 3 || prm.order < 0 ||
        prm.order > 5 || !(prm.mult_base > 0) || !(prm.mult_low > 0) || !(prm.mult_high > 0) || prm.pred < 0 ||
        !pred_valid((uint64_t)prm.pred))
        throw StzError(STZ_E_ARG, "bad parameters");
    hipStream_t st = c->stream;
    auto pool = c->pool;
    fe.px = pad4(fe.nx);
    fe.py = pad4(fe.ny);
    fe.pz = pad4(fe.nz);
    fe.nin = (uint64_t)fe.nx * fe.ny * fe.nz;
    // input on device, 2-element aligned, padded to multiples of 4
    fe.din = field;
    const bool padded = fe.px != fe.nx || fe.py != fe.ny || fe.pz != fe.nz;
    if (!is_device_ptr(field)) {
        fe.in_copy = DBuf(pool, fe.nin * sizeof(T));
        HIPCHK(hipMemcpyAsync(fe.in_copy.p, field, fe.nin * sizeof(T), hipMemcpyHostToDevice, st));
        fe.din = fe.in_copy.template as<T>();
    } else if (((uintptr_t)field % (2 * sizeof(T))) != 0 && !padded) {
        fe.in_copy = DBuf(pool, fe.nin * sizeof(T));
        HIPCHK(hipMemcpyAsync(fe.in_copy.p, field, fe.nin * sizeof(T), hipMemcpyDeviceToDevice, st));
        fe.din = fe.in_copy.template as<T>();
    }
    fe.full = fe.din;
    if (padded) {
        fe.padded = DBuf(pool, fe.px * fe.py * fe.pz * sizeof(T));
        K(c, "pad", [&] {
            launch_pad<T>(st, fe.din, (int)fe.nx, (int)fe.ny, (int)fe.nz, fe.padded.template as<T>(), (int)fe.px, (int)fe.py,
                          (int)fe.pz);
        });
        fe.full = fe.padded.template as<T>();
    }
    // value range (computeRange, F5:92-104) -> level error bounds
    fe.mm_part = DBuf(pool, 2 * 2048 * sizeof(double) + 64);
    double* outp = fe.mm_part.template as<double>() + 2 * 2048;
    K(c, "minmax", [&] { launch_minmax<T>(st, fe.din, fe.nin, fe.mm_part.template as<double>(), outp); });
    fe.mm_host = (double*)c->host_stage(64);
    HIPCHK(hipMemcpyAsync(fe.mm_host, outp, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipEventRecord(c->ev_range, st));
    // the work that does not depend on the range is queued before the
    // caller waits for it (on ev_range): scratch, histograms cleared, base
    // grid gathered
    make_layout(fe.L, fe.nx, fe.ny, fe.nz, fe.px, fe.py, fe.pz, prm.chunk_high, prm.chunk_low, prm.interp);
    const HostLayout& L = fe.L;
    const int lx = (int)(fe.px / 4), ly = (int)(fe.py / 4), lz = (int)(fe.pz / 4);
    const int mx = (int)(fe.px / 2), my = (int)(fe.py / 2), mz = (int)(fe.pz / 2);
    fe.codes = DBuf(pool, L.total_codes * 2);
    fe.side = DBuf(pool, L.total_codes * sizeof(T));
    fe.hist = DBuf(pool, (size_t)kStreams * 65536 * 4);
    fe.vbase = DBuf(pool, (size_t)lx * ly * lz * sizeof(T));
    fe.mid = DBuf(pool, (size_t)mx * my * mz * sizeof(T));
    K(c, "hist_zero", [&] { launch_fill_u32(st, fe.hist.template as<uint32_t>(), (uint64_t)kStreams * 65536, 0u); });
    // base level (F5:2132-2145): gather low block 0; the pass schedule follows
    // once the bounds are known
    K(c, "gather_low0", [&] {
        launch_gather_low0<T>(st, fe.full, (int)fe.px, (int)fe.py, (int)fe.pz, fe.vbase.template as<T>());
    });
}

template <typename V>
DBuf upload_vec(stz_ctx* c, const std::vector<V>& v);

// enc_phase0 for many fields: the checks, input copies and buffers per
// field, then the padding, value ranges, histogram clear and base-grid
// gathers of all of them in one launch each (a batch of 96 fields otherwise
// issues some 700 small launches and copies here).  hist_all receives every
// field's histograms (15 x 65536 bins each, views into it) and must outlive
// their use.
template <typename T>
void enc_phase0_batch(stz_ctx* c, std::vector<FieldEnc<T>>& fe, const void* const* fields, DBuf& hist_all,
                      DBuf& mm_all)
{
    hipStream_t st = c->stream;
    auto pool = c->pool;
    const int n = (int)fe.size();
    const size_t hist_words = (size_t)kStreams * 65536;
    hist_all = DBuf(pool, (size_t)n * hist_words * 4);
    constexpr uint32_t kMaxWgs = 256;  // per field and kind
    std::vector<ProJob> jobs(n);
    uint32_t pad_wgs = 0, mm_wgs = 0, ga_wgs = 0;
    for (int i = 0; i < n; ++i) {
        FieldEnc<T>& f = fe[i];
        const stz_params& prm = f.prm;
        check_dims(f.nx, f.ny, f.nz, prm.chunk_high, prm.chunk_low);
        if (!(f.eb >= 0) || prm.alpha < 1.0 || prm.beta < 1.0 || prm.interp < 0 || prm.interp > 3 || prm.order < 0 ||
            prm.order > 5 || !(prm.mult_base > 0) || !(prm.mult_low > 0) || !(prm.mult_high > 0) || prm.pred < 0 ||
            !pred_valid((uint64_t)prm.pred))
            throw StzError(STZ_E_ARG, "bad parameters");
        f.px = pad4(f.nx);
        f.py = pad4(f.ny);
        f.pz = pad4(f.nz);
        f.nin = (uint64_t)f.nx * f.ny * f.nz;
        const T* field = (const T*)fields[i];
        f.din = field;
        const bool padded = f.px != f.nx || f.py != f.ny || f.pz != f.nz;
        if (!is_device_ptr(field)) {
            f.in_copy = DBuf(pool, f.nin * sizeof(T));
            HIPCHK(hipMemcpyAsync(f.in_copy.p, field, f.nin * sizeof(T), hipMemcpyHostToDevice, st));
            f.din = f.in_copy.template as<T>();
        } else if (((uintptr_t)field % (2 * sizeof(T))) != 0 && !padded) {
            f.in_copy = DBuf(pool, f.nin * sizeof(T));
            HIPCHK(hipMemcpyAsync(f.in_copy.p, field, f.nin * sizeof(T), hipMemcpyDeviceToDevice, st));
            f.din = f.in_copy.template as<T>();
        }
        f.full = f.din;
        if (padded) {
            f.padded = DBuf(pool, f.px * f.py * f.pz * sizeof(T));
            f.full = f.padded.template as<T>();
        }
        make_layout(f.L, f.nx, f.ny, f.nz, f.px, f.py, f.pz, prm.chunk_high, prm.chunk_low, prm.interp);
        const HostLayout& L = f.L;
        const size_t lb = (f.px / 4) * (f.py / 4) * (f.pz / 4);
        f.codes = DBuf(pool, L.total_codes * 2);
        f.side = DBuf(pool, L.total_codes * sizeof(T));
        f.hist.p = hist_all.as<uint32_t>() + (size_t)i * hist_words;  // a view (no pool: not freed)
        f.vbase = DBuf(pool, lb * sizeof(T));
        f.mid = DBuf(pool, (f.px / 2) * (f.py / 2) * (f.pz / 2) * sizeof(T));
        f.mm_part = DBuf(pool, 2 * kMaxWgs * sizeof(double));
        ProJob& j = jobs[i];
        j.in = f.din;
        j.padded = padded ? f.padded.p : nullptr;
        j.full = f.full;
        j.vbase = f.vbase.p;
        j.partial = f.mm_part.template as<double>();
        j.nx = (int)f.nx; j.ny = (int)f.ny; j.nz = (int)f.nz;
        j.px = (int)f.px; j.py = (int)f.py; j.pz = (int)f.pz;
        j.nin = f.nin;
        const uint64_t rows = (uint64_t)f.py * f.pz;
        j.pad_wg0 = pad_wgs;
        j.pad_nwg = padded ? (uint32_t)std::min<uint64_t>((rows + 7) / 8, kMaxWgs) : 0u;
        pad_wgs += j.pad_nwg;
        j.mm_wg0 = mm_wgs;
        j.mm_nwg = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(f.nin / (256 * 64), 1), kMaxWgs);
        mm_wgs += j.mm_nwg;
        j.ga_wg0 = ga_wgs;
        j.ga_nwg = (uint32_t)std::min<uint64_t>(std::max<uint64_t>((lb + 2047) / 2048, 1), kMaxWgs);
        ga_wgs += j.ga_nwg;
    }
    DBuf dj = upload_vec(c, jobs);
    mm_all = DBuf(pool, (size_t)n * 2 * sizeof(double) + 64);
    K(c, "hist_zero", [&] { launch_fill_u32(st, hist_all.as<uint32_t>(), (uint64_t)n * hist_words, 0u); });
    K(c, "minmax", [&] {
        launch_prologue_multi<T>(st, dj.as<ProJob>(), (uint32_t)n, pad_wgs, mm_wgs, ga_wgs, mm_all.as<double>());
    });
    double* mm_host = (double*)c->host_stage((size_t)n * 2 * sizeof(double));
    HIPCHK(hipMemcpyAsync(mm_host, mm_all.p, (size_t)n * 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipEventRecord(c->ev_range, st));
    for (int i = 0; i < n; ++i) fe[i].mm_host = mm_host + 2 * i;
}

// after the wait: level bounds from the range.  A learned choice made after
// the layout may have switched to or from the Lorenzo/regression base, whose
// base stream holds header symbols: then the layout and the code buffers
// (which no launch has touched yet) are made again.
template <typename T>
void enc_phase1(stz_ctx* c, FieldEnc<T>& fe)
{
    fe.vmin = fe.mm_host[0];
    fe.vmax = fe.mm_host[1];
    level_ebs(fe.eb, fe.mode, fe.vmin, fe.vmax, fe.dtype, fe.prm, fe.ebl, fe.unit);
    const Geom& g0 = fe.L.g[0];
    if (fe.L.n[0] != (uint64_t)g0.dx * g0.dy * g0.dz + base_extra_symbols(fe.prm.interp, g0.dx, g0.dy, g0.dz)) {
        make_layout(fe.L, fe.nx, fe.ny, fe.nz, fe.px, fe.py, fe.pz, fe.prm.chunk_high, fe.prm.chunk_low,
                    fe.prm.interp);
        fe.codes = DBuf(c->pool, fe.L.total_codes * 2);
        fe.side = DBuf(c->pool, fe.L.total_codes * sizeof(T));
    }
}

template <typename T>
LevelArgs enc_level_args(const FieldEnc<T>& fe, int level)
{
    LevelArgs a{};
    a.eb = fe.ebl[level];
    a.recip = 1.0 / fe.ebl[level];
    for (int b = 0; b < 7; ++b) a.code_off[b] = fe.L.code_off[(level == 1 ? 1 : 8) + b];
    a.opt = fe.prm.pred;
    return a;
}

// file and stream headers of a field from its codebooks (built before)
template <typename T>
void enc_phase3_headers(FieldEnc<T>& fe)
{
    Meta& m = fe.ds->meta;
    std::memset(&m.fh, 0, sizeof(m.fh));
    FileHeader& fh = m.fh;
    const stz_params& prm = fe.prm;
    fh.magic = kMagic;
    fh.version = kVersion;
    fh.dtype = (uint32_t)fe.dtype;
    fh.eb_mode = (uint32_t)fe.mode;
    fh.nx = fe.nx; fh.ny = fe.ny; fh.nz = fe.nz;
    fh.px = fe.px; fh.py = fe.py; fh.pz = fe.pz;
    fh.eb_user = fe.eb;
    fh.eb_unit = fe.unit;
    fh.mult_base = prm.mult_base; fh.mult_low = prm.mult_low; fh.mult_high = prm.mult_high;
    fh.eb_base = fe.ebl[0]; fh.eb_low = fe.ebl[1]; fh.eb_high = fe.ebl[2];
    fh.vmin = fe.vmin; fh.vmax = fe.vmax;
    fh.radius = kRadius;
    fh.chunk_high = (uint32_t)prm.chunk_high;
    fh.chunk_low = (uint32_t)prm.chunk_low;
    fh.sub_size = kSub;
    fh.interp = (uint32_t)prm.interp;
    fh.order = (uint32_t)prm.order;
    fh.pred = (uint64_t)prm.pred;
    fh.lmax = kLmax;
    fh.n_streams = kStreams;
    fh.alpha = prm.alpha;
    fh.beta = prm.beta;
    uint64_t total = sizeof(FileHeader);
    fe.enc_n = 0;
    for (int s = 0; s < kStreams; ++s) {
        const uint32_t n = (uint32_t)fe.lists[s].size();
        const uint64_t n_unpred = (n && (fe.lists[s][0] >> 32) == 0) ? (fe.lists[s][0] & 0xFFFFFFFFull) : 0;
        const double ebs = s == 0 ? fe.ebl[0] : s < 8 ? fe.ebl[1] : fe.ebl[2];
        fill_stream_header(m.sh[s], s, fe.L, fe.dtype, ebs, n, n_unpred, m.cb[s].total_bits);
        fh.stream_off[s] = total;
        fh.stream_bytes[s] = m.sh[s].stream_bytes;
        total += m.sh[s].stream_bytes;
        fe.enc_n += n;
    }
    fh.total_bytes = total;
}

template <typename T>
LevelArgs enc_level_args(const FieldEnc<T>& fe, int level);

template <typename T>
void compress_fields(stz_ctx* c, int dtype, int n, const void* const* fields, const size_t* dims, double eb,
                     int mode, const stz_params& prm, stz_dstream* const* out, std::vector<char>* low_conf);

template <typename T>
void compress_impl(stz_ctx* c, int dtype, const T* field, size_t nx, size_t ny, size_t nz, double eb, int mode,
                   const stz_params& prm, stz_dstream* ds)
{
    const void* f = field;
    const size_t dims[3] = {nx, ny, nz};
    compress_fields<T>(c, dtype, 1, &f, dims, eb, mode, prm, &ds, nullptr);
}

template <typename T>
int tune_trial(stz_ctx* c, int dtype, const T* d, size_t nx, size_t ny, size_t nz, double eb, int mode,
               const stz_params& base, stz_dstream* keep, uint64_t* sizes);

// ------------------------------------------------------------ decode helpers
void ensure_tabs(stz_ctx* c, stz_dstream* ds)
{
    if (ds->have_tabs) return;
    const Meta& m = ds->meta;
    // canonical descriptions built on the host workers, one stream each
    std::vector<DecHead> heads(kStreams);
    std::vector<std::vector<uint16_t>> sorted(kStreams);
    c->tasks().run(kStreams, [&](int s) {
        build_dechead(m.cb[s].sym.data(), m.cb[s].len.data(), (uint32_t)m.cb[s].sym.size(), heads[s], sorted[s]);
    });
    size_t ns = 0;
    for (int s = 0; s < kStreams; ++s) {
        heads[s].sorted_off = (uint32_t)ns;
        ns += sorted[s].size();
    }
    // one upload: the sorted symbols, then (16-byte aligned) the descriptions,
    // which stay behind them in the stream's buffer
    const size_t sb = align_up(ns * 2, 16), hb = sizeof(DecHead) * kStreams;
    ds->tabs = DBuf(ds->pool, sizeof(DecTab) * kStreams);
    ds->sorted = DBuf(ds->pool, sb + hb + 16);
    uint8_t* stage = (uint8_t*)c->host_stage(sb + hb + 64);
    size_t o = 0;
    for (int s = 0; s < kStreams; ++s) {
        std::memcpy(stage + o, sorted[s].data(), sorted[s].size() * 2);
        o += sorted[s].size() * 2;
    }
    std::memcpy(stage + sb, heads.data(), hb);
    HIPCHK(hipMemcpyAsync(ds->sorted.p, stage, sb + hb, hipMemcpyHostToDevice, c->stream));
    const DecHead* dh = (const DecHead*)(ds->sorted.as<uint8_t>() + sb);
    K(c, "dectabs", [&] { launch_build_dectabs(c->stream, dh, ds->sorted.as<uint16_t>(), ds->tabs.as<DecTab>(), kStreams); });
    ds->have_tabs = true;
}

// ensure_tabs for many streams at once: their canonical descriptions on the
// wide host pool, one upload per stream object, one table launch for all
void ensure_tabs_batch(stz_ctx* c, stz_dstream* const* dss, int n)
{
    std::vector<stz_dstream*> todo;
    for (int i = 0; i < n; ++i)
        if (!dss[i]->have_tabs) todo.push_back(dss[i]);
    const int m = (int)todo.size();
    if (m < 2) {
        for (stz_dstream* d : todo) ensure_tabs(c, d);
        return;
    }
    std::vector<DecHead> heads((size_t)m * kStreams);
    std::vector<std::vector<uint16_t>> sorted((size_t)m * kStreams);
    c->tasks_wide().run(m * kStreams, [&](int k) {
        const Meta& me = todo[k / kStreams]->meta;
        const int s = k % kStreams;
        build_dechead(me.cb[s].sym.data(), me.cb[s].len.data(), (uint32_t)me.cb[s].sym.size(), heads[k], sorted[k]);
    });
    std::vector<size_t> sb(m), off(m);
    size_t total = 0;
    for (int i = 0; i < m; ++i) {
        size_t ns = 0;
        for (int s = 0; s < kStreams; ++s) {
            heads[(size_t)i * kStreams + s].sorted_off = (uint32_t)ns;
            ns += sorted[(size_t)i * kStreams + s].size();
        }
        sb[i] = align_up(ns * 2, 16);
        off[i] = total;
        total += align_up(sb[i] + sizeof(DecHead) * kStreams, 16);
    }
    const size_t jobs_off = total;
    uint8_t* stage = (uint8_t*)c->host_stage(total + (size_t)m * sizeof(DecTabJob) + 64);
    std::vector<DecTabJob> jobs(m);
    for (int i = 0; i < m; ++i) {
        stz_dstream* d = todo[i];
        const size_t hb = sizeof(DecHead) * kStreams;
        d->tabs = DBuf(d->pool, sizeof(DecTab) * kStreams);
        d->sorted = DBuf(d->pool, sb[i] + hb + 16);
        uint8_t* b = stage + off[i];
        size_t o = 0;
        for (int s = 0; s < kStreams; ++s) {
            const std::vector<uint16_t>& v = sorted[(size_t)i * kStreams + s];
            std::memcpy(b + o, v.data(), v.size() * 2);
            o += v.size() * 2;
        }
        std::memcpy(b + sb[i], &heads[(size_t)i * kStreams], hb);
        HIPCHK(hipMemcpyAsync(d->sorted.p, b, sb[i] + hb, hipMemcpyHostToDevice, c->stream));
        jobs[i].heads = (const DecHead*)(d->sorted.as<uint8_t>() + sb[i]);
        jobs[i].sorted = d->sorted.as<uint16_t>();
        jobs[i].tabs = d->tabs.as<DecTab>();
        d->have_tabs = true;
    }
    std::memcpy(stage + jobs_off, jobs.data(), (size_t)m * sizeof(DecTabJob));
    DBuf dj(c->pool, (size_t)m * sizeof(DecTabJob));
    HIPCHK(hipMemcpyAsync(dj.p, stage + jobs_off, (size_t)m * sizeof(DecTabJob), hipMemcpyHostToDevice, c->stream));
    K(c, "dectabs", [&] { launch_build_dectabs_multi(c->stream, dj.as<DecTabJob>(), (uint32_t)m, kStreams); });
}

void layout_of(const Meta& m, HostLayout& L)
{
    make_layout(L, m.fh.nx, m.fh.ny, m.fh.nz, m.fh.px, m.fh.py, m.fh.pz, (int)m.fh.chunk_high, (int)m.fh.chunk_low,
                (int)m.fh.interp);
}

// Decoded symbols.  Full layout (every chunk of every stream): stream s
// starts at code_base[s] = its code_off.  Compact layout (random access):
// streams decoded in full keep their stream order from code_base[s]; listed
// chunks of the other streams are packed 32-symbol aligned, located through
// cmap[cm_base[s] + chunk].
struct Decoded {
    DBuf codes, side, cmap;
    DBuf items;             // decode work items (read by a launch on the second stream)
    bool split = false;     // streams kSplitStream.. decode on the second stream: join before use
    uint64_t code_base[kStreams] = {};
    uint32_t cm_base[kStreams] = {};
    bool compact = false;
};

// device error word: bit 0 corrupt bitstream, bit 1 a read outside the plan
struct ErrFlag {
    DBuf d;
    uint32_t* h = nullptr;
    explicit ErrFlag(stz_ctx* c) : d(c->pool, 64)
    {
        h = (uint32_t*)c->host_stage(64);
        *h = 0;
        K(c, "err_zero", [&] { launch_fill_u32(c->stream, d.as<uint32_t>(), 1, 0u); });
    }
    uint32_t* ptr() { return d.as<uint32_t>(); }
    void fetch(stz_ctx* c) { HIPCHK(hipMemcpyAsync(h, d.p, 4, hipMemcpyDeviceToHost, c->stream)); }
    void check() const
    {
        if (*h & 1u) throw StzError(STZ_E_STREAM, "corrupt Huffman bitstream");
        if (*h & 2u) throw StzError(STZ_E_STREAM, "random access read a chunk outside its plan");
    }
};

// Streams 0 .. kSplitStream-1 (base and low) feed the base and low levels;
// the high streams, about 7/8 of the symbols, only the high level.
constexpr int kSplitStream = 8;

// Decode form of a launch by the symbol-weighted bits per symbol of what it
// decodes: up to 2 the multi-symbol table, above it the single-symbol one
int dec_form(double bps)
{
    return bps <= 2.0 ? (int)kDecStaged : (int)kDecDense;
}

// sub-chunks per decode workgroup: 256, fewer for the staged form while a
// stream's expected bits with a 20% margin would overflow the LDS stage (on
// the 512^3 log-normal field the largest workgroup range of a stream is 1.2x
// its mean; lanes past the stage read global memory, correct but slower).
// Two staged workgroups share a CU, so 256 sub-chunks each hold the whole
// decode of such a field resident in one round.
uint32_t dec_ipw(int form, double bps)
{
    uint32_t ipw = 256;
    if (form == kDecStaged)
        while (ipw > 32 && ipw * (double)kSub * bps * 1.2 > (double)kDecStageWords * 32.0) ipw >>= 1;
    return ipw;
}

// Enqueue the Huffman decode of full[s] streams (every chunk) and of the
// chunks lists[s] of the others.  split: the high streams decode on the
// context's second stream, overlapping the base and low levels' decode on
// `stream`; join_high() orders `stream` after them.
void huff_decode(stz_ctx* c, stz_dstream* ds, const HostLayout& L, const bool* full,
                 const std::vector<int64_t>* lists, Decoded& out, uint32_t* err, bool split = false)
{
    ensure_tabs(c, ds);
    hipStream_t st = c->stream;
    const Meta& m = ds->meta;
    StreamSet ss = make_streamset(L, m, (int)m.fh.dtype);

    const size_t ts = tsize_of((int)m.fh.dtype);
    bool all_full = true;
    for (int s = 0; s < kStreams; ++s) all_full = all_full && full[s];
    out.compact = !all_full;
    std::vector<uint32_t> quads;  // (stream, chunk, first item within the stream, output position)
    uint64_t pos = 0;
    uint32_t cm = 0, item = 0, wg = 0;
    for (int s = 0; s < kStreams; ++s) {
        StreamDesc& d = ss.s[s];
        d.cm_base = cm;
        out.cm_base[s] = cm;
        cm += L.g[s].n_chunks;
        if (all_full) {
            out.code_base[s] = L.code_off[s];
            d.item_count = L.n_sub[s];
        } else {
            pos = align_up(pos, 64);
            out.code_base[s] = pos;
            d.code_off = 0;
            if (full[s]) {
                for (uint32_t ch = 0; ch < L.g[s].n_chunks; ++ch) {
                    quads.push_back((uint32_t)s);
                    quads.push_back(ch);
                    quads.push_back(L.chunk_sub[s][ch]);
                    quads.push_back((uint32_t)(pos + chunk_start(L.g[s], ch)));
                }
                pos += L.n[s];
                d.item_count = L.n_sub[s];
            } else {
                uint32_t k = 0;
                for (int64_t ch : lists[s]) {
                    quads.push_back((uint32_t)s);
                    quads.push_back((uint32_t)ch);
                    quads.push_back(k);
                    quads.push_back((uint32_t)pos);
                    k += L.chunk_sub[s][ch + 1] - L.chunk_sub[s][ch];
                    pos += align_up(chunk_volume(L.g[s], (uint32_t)ch, L.n[s]), 32);
                }
                d.item_count = k;
            }
            if (pos >= (1ull << 32)) throw StzError(STZ_E_DIMS, "random-access decode larger than 2^32 symbols");
        }
    }
    // Decode variant per launch (a split decode has two: streams below
    // kSplitStream, and the rest), by the symbol-weighted entropy of what it
    // decodes: up to 2 bits per symbol the LDS-staged kernel (256 sub-chunks
    // per workgroup, halved while a stream's expected bits with a 30% margin
    // would not fit the 44 KiB stage); above, the unstaged kernel, whose
    // smaller LDS footprint fits more workgroups per CU.
    const bool two = split;
    double wbits[2] = {0, 0}, wsym[2] = {0, 0};
    for (int s = 0; s < kStreams; ++s) {
        const int gi = two && s >= kSplitStream ? 1 : 0;
        const double bps = L.n[s] ? (double)m.sh[s].total_bits / (double)L.n[s] : 0.0;
        wbits[gi] += bps * ss.s[s].item_count;
        wsym[gi] += ss.s[s].item_count;
    }
    const int form_of[2] = {dec_form(wsym[0] > 0 ? wbits[0] / wsym[0] : 0.0),
                            dec_form(wsym[1] > 0 ? wbits[1] / wsym[1] : 0.0)};
    const int use_form = form_of[0];
    for (int s = 0; s < kStreams; ++s) {
        StreamDesc& d = ss.s[s];
        const double bps = L.n[s] ? (double)m.sh[s].total_bits / (double)L.n[s] : 0.0;
        d.ipw = dec_ipw(form_of[two && s >= kSplitStream ? 1 : 0], bps);
        d.item_base = item;
        d.wg_base = wg;
        const uint32_t nwg = (d.item_count + d.ipw - 1) / d.ipw;
        item += nwg * d.ipw;
        wg += nwg;
    }
    ss.total_items = item;
    ss.total_wgs = wg;
    const uint64_t ncodes = all_full ? L.total_codes : pos;
    out.codes = DBuf(c->pool, ncodes * 2 + 64);
    out.side = DBuf(c->pool, ncodes * ts + 64);
    out.cmap = DBuf(c->pool, (size_t)cm * 4 + 16);
    if (!all_full) K(c, "cmap_fill", [&] { launch_fill_u32(st, out.cmap.as<uint32_t>(), cm, kNoChunk); });
    if (!item) return;
    out.items = DBuf(c->pool, (size_t)item * sizeof(DecItem));
    DBuf& items = out.items;
    uint8_t* pay = ds->payload.as<uint8_t>();
    if (all_full) {
        K(c, "dec_prep", [&] { launch_dec_prep(st, ss, pay, nullptr, 0, items.as<DecItem>(), nullptr); });
    } else {
        DBuf ld(c->pool, quads.size() * 4 + 16);
        uint32_t* stage = (uint32_t*)c->host_stage(quads.size() * 4 + 16);
        std::memcpy(stage, quads.data(), quads.size() * 4);
        HIPCHK(hipMemcpyAsync(ld.p, stage, quads.size() * 4, hipMemcpyHostToDevice, st));
        K(c, "dec_prep", [&] {
            launch_dec_prep(st, ss, pay, ld.as<uint32_t>(), (uint32_t)(quads.size() / 4), items.as<DecItem>(),
                            out.cmap.as<uint32_t>());
        });
    }
    if (split && ss.s[kSplitStream].wg_base > 0 && ss.s[kSplitStream].wg_base < ss.total_wgs) {
        // two launches over one item array: streams 0 .. kSplitStream-1 here,
        // the rest (their workgroup ids rebased, their tables offset) on the
        // second stream once the items exist
        StreamSet lo = ss, hi = ss;
        const uint32_t w0 = ss.s[kSplitStream].wg_base;
        lo.nstreams = kSplitStream;
        lo.total_wgs = w0;
        hi.nstreams = kStreams - kSplitStream;
        for (int k = 0; k < (int)hi.nstreams; ++k) {
            hi.s[k] = ss.s[kSplitStream + k];
            hi.s[k].wg_base -= w0;
        }
        hi.total_wgs = ss.total_wgs - w0;
        hipStream_t s2 = c->second();
        HIPCHK(hipEventRecord(c->ev_items, st));
        HIPCHK(hipStreamWaitEvent(s2, c->ev_items, 0));
        K(c, "huff_decode", [&] {
            launch_huff_decode(st, lo, pay, ds->tabs.as<DecTab>(), ds->sorted.as<uint16_t>(), items.as<DecItem>(),
                               out.codes.as<uint16_t>(), out.side.p, err, form_of[0]);
            launch_huff_decode(s2, hi, pay, ds->tabs.as<DecTab>() + kSplitStream, ds->sorted.as<uint16_t>(),
                               items.as<DecItem>(), out.codes.as<uint16_t>(), out.side.p, err, form_of[1]);
        });
        HIPCHK(hipEventRecord(c->ev_high, s2));
        out.split = true;
        return;
    }
    K(c, "huff_decode", [&] {
        launch_huff_decode(st, ss, pay, ds->tabs.as<DecTab>(), ds->sorted.as<uint16_t>(), items.as<DecItem>(),
                           out.codes.as<uint16_t>(), out.side.p, err, use_form);
    });
}

// `stream` waits for the high streams' decode of a split huff_decode
void join_high(stz_ctx* c, Decoded& dd)
{
    if (!dd.split) return;
    HIPCHK(hipStreamWaitEvent(c->stream, c->ev_high, 0));
    dd.split = false;
}

// on an error between the split and the join: `stream` still waits for the
// second stream before the decode buffers go back to the pool
struct JoinGuard {
    stz_ctx* c;
    Decoded& dd;
    ~JoinGuard()
    {
        if (dd.split) {
            (void)hipStreamWaitEvent(c->stream, c->ev_high, 0);
            (void)hipStreamSynchronize(c->side_stream);
        }
    }
};

template <typename T>
void decode_base(stz_ctx* c, const Meta& m, const HostLayout& L, const Decoded& dd, T* v, int stop_level)
{
    if (m.fh.interp >= 2) {
        // Lorenzo tiles: the whole base grid whatever stop_level asks
        LorenzoJob j{};
        j.reg = m.fh.interp == 3;
        j.lo = m.fh.vmin;
        j.hi = m.fh.vmax;
        j.v = v;
        j.codes = dd.codes.as<uint16_t>() + dd.code_base[0];
        j.side = dd.side.as<T>() + dd.code_base[0];
        j.bx = L.g[0].dx; j.by = L.g[0].dy; j.bz = L.g[0].dz;
        j.eb = m.fh.eb_base;
        j.recip = 1.0 / m.fh.eb_base;
        K(c, "lorenzo_decode", [&] { launch_lorenzo<T>(c->stream, j, false); });
        return;
    }
    const auto passes = make_passes(L.g[0].dx, L.g[0].dy, L.g[0].dz, m.fh.eb_base, (int)m.fh.interp,
                                    (int)m.fh.order, m.fh.alpha, m.fh.beta, stop_level);
    run_passes(passes, sizeof(T), [&](const InterpPassList& fl) {
        K(c, "interp_decode_fused", [&] {
            launch_interp_decode_fused<T>(c->stream, fl, v, dd.codes.as<uint16_t>() + dd.code_base[0],
                                          dd.side.as<T>() + dd.code_base[0]);
        });
    }, [&](const InterpPass& p) {
        K(c, "interp_decode", [&] {
            launch_interp_decode<T>(c->stream, p, v, dd.codes.as<uint16_t>() + dd.code_base[0],
                                    dd.side.as<T>() + dd.code_base[0]);
        });
    });
}

// level 1 (low) or 2 (high) arguments; random access goes through the chunk map
LevelArgs level_args(const Meta& m, const Decoded& dd, int level, uint32_t* err)
{
    LevelArgs a{};
    a.eb = level == 1 ? m.fh.eb_low : m.fh.eb_high;
    a.recip = 1.0 / a.eb;
    const int s0 = level == 1 ? 1 : 8;
    for (int b = 0; b < 7; ++b) {
        a.code_off[b] = dd.code_base[s0 + b];
        a.cm_off[b] = dd.cm_base[s0 + b];
    }
    a.cmap = dd.compact ? dd.cmap.as<uint32_t>() : nullptr;
    a.err = err;
    a.opt = (int)m.fh.pred;
    return a;
}

// output buffer: the caller's device pointer, or a pool buffer copied back
struct OutBuf {
    void* user = nullptr;
    bool dev = false;
    DBuf tmp;
    void* ptr = nullptr;
    size_t bytes = 0;
    OutBuf() = default;
    OutBuf(stz_ctx* c, void* u, size_t n) : user(u), dev(is_device_ptr(u)), bytes(n)
    {
        if (dev) ptr = u;
        else { tmp = DBuf(c->pool, n + 16); ptr = tmp.p; }
    }
    void copy_back(stz_ctx* c)
    {
        if (!dev && bytes) HIPCHK(hipMemcpyAsync(user, ptr, bytes, hipMemcpyDeviceToHost, c->stream));
    }
};

template <typename T>
void decompress_full(stz_ctx* c, stz_dstream* ds, void* out)
{
    const Meta& m = ds->meta;
    HostLayout L;
    layout_of(m, L);
    bool all[kStreams];
    for (int s = 0; s < kStreams; ++s) all[s] = true;
    ErrFlag err(c);
    Decoded dd;
    // the high streams' Huffman decode overlaps the base and low levels
    // (one stream when profiling, so per-kernel events stay meaningful, and
    // for a context held to part of the GPU, which should not widen)
    JoinGuard guard{c, dd};
    huff_decode(c, ds, L, all, nullptr, dd, err.ptr(), !c->prof && c->grid_share >= 1.0f);
    hipStream_t st = c->stream;
    const int lx = L.g[0].dx, ly = L.g[0].dy, lz = L.g[0].dz;
    DBuf vb(c->pool, (size_t)lx * ly * lz * sizeof(T)), mid(c->pool, (size_t)L.g[8].dx * L.g[8].dy * L.g[8].dz * sizeof(T));
    decode_base<T>(c, m, L, dd, vb.as<T>(), 0);
    const LevelArgs la = level_args(m, dd, 1, err.ptr()), ha = level_args(m, dd, 2, err.ptr());
    OutBuf ob(c, out, m.fh.nx * m.fh.ny * m.fh.nz * sizeof(T));
    if (m.fh.pred & kPredStaged) {
        // dependency.cpp's staged predictors: three launches per level, the
        // high level's cells cropped to the field as they are written
        DBuf rec(c->pool, staged_scratch(L, sizeof(T)));
        K(c, "low_decode", [&] {
            launch_staged_decode<T>(st, vb.as<T>(), L.g[1], la, dd.codes.as<uint16_t>(), dd.side.as<T>(), mid.as<T>(),
                                    L.g[8].dx, L.g[8].dy, L.g[8].dz, rec.as<T>());
        });
        join_high(c, dd);
        K(c, "high_decode", [&] {
            launch_staged_decode<T>(st, mid.as<T>(), L.g[8], ha, dd.codes.as<uint16_t>(), dd.side.as<T>(), (T*)ob.ptr,
                                    (int)m.fh.nx, (int)m.fh.ny, (int)m.fh.nz, rec.as<T>());
        });
    } else {
        K(c, "low_decode", [&] {
            launch_low_decode<T>(st, vb.as<T>(), L.g[1], la, dd.codes.as<uint16_t>(), dd.side.as<T>(), mid.as<T>(),
                                 c->grid_share);
        });
        join_high(c, dd);
        K(c, "high_decode", [&] {
            launch_high_decode<T>(st, mid.as<T>(), L.g[8], ha, dd.codes.as<uint16_t>(), dd.side.as<T>(), (T*)ob.ptr,
                                  (int)m.fh.nx, (int)m.fh.ny, (int)m.fh.nz, c->grid_share);
        });
    }
    ob.copy_back(c);
    err.fetch(c);
    HIPCHK(hipStreamSynchronize(st));
    err.check();
}

// the normalised box of a box or slice request
Box request_box(const Meta& m, const stz_request& r)
{
    int a[3], b[3];
    if (r.kind == STZ_REQ_BOX) {
        for (int k = 0; k < 3; ++k) { a[k] = r.a[k]; b[k] = r.a[3 + k]; }
    } else {
        const int axis = r.a[0], index = r.a[1];
        if (axis < 0 || axis > 2) throw StzError(STZ_E_ARG, "slice axis");
        const uint64_t d[3] = {m.fh.nx, m.fh.ny, m.fh.nz};
        if (index < 0 || (uint64_t)index >= d[axis]) throw StzError(STZ_E_ARG, "slice index");
        for (int k = 0; k < 3; ++k) { a[k] = 0; b[k] = (int)d[k] - 1; }
        a[axis] = b[axis] = index;
    }
    return Box{a[0], b[0], a[1], b[1], a[2], b[2]};
}

// Decode of the kept base level into ds->base (stz_dstream_prepare).
template <typename T>
void prepare_base(stz_ctx* c, stz_dstream* ds)
{
    if (ds->have_base) return;
    const Meta& m = ds->meta;
    HostLayout L;
    layout_of(m, L);
    bool full[kStreams] = {};
    std::vector<int64_t> lists[kStreams];
    full[0] = true;
    ErrFlag err(c);
    Decoded dd;
    huff_decode(c, ds, L, full, lists, dd, err.ptr());
    DBuf b(ds->pool, (size_t)L.g[0].dx * L.g[0].dy * L.g[0].dz * sizeof(T));
    decode_base<T>(c, m, L, dd, b.as<T>(), 0);
    err.fetch(c);
    HIPCHK(hipStreamSynchronize(c->stream));
    err.check();
    ds->base = std::move(b);
    ds->have_base = true;
}

// Random-access requests served together: one Huffman decode of the union
// of their chunks, one launch for all low regions, one for all high boxes.
// Full-field requests run after, through the full pipeline.
template <typename T>
void decompress_batch(stz_ctx* c, stz_dstream* ds, int n, const stz_request* rq, void* const* outs, size_t* od)
{
    const Meta& m = ds->meta;
    HostLayout L;
    layout_of(m, L);
    hipStream_t st = c->stream;
    std::vector<Plan> plans(n);
    std::vector<std::vector<uint8_t>> want(kStreams);
    bool full[kStreams] = {};
    bool need_base = false, need_mid = false, need_lowfull = false;
    std::vector<int> full_reqs;
    for (int i = 0; i < n; ++i) {
        const stz_request& r = rq[i];
        size_t* o = od + 3 * i;
        if (r.kind == STZ_REQ_BOX || r.kind == STZ_REQ_SLICE) {
            const Box q = request_box(m, r);
            Plan& pl = plans[i];
            make_plan(pl, (int)m.fh.nx, (int)m.fh.ny, (int)m.fh.nz, (int)m.fh.px, (int)m.fh.py, (int)m.fh.pz,
                      (int)m.fh.chunk_high, (int)m.fh.chunk_low, q.x0, q.y0, q.z0, q.x1, q.y1, q.z1);
            // make_query keeps lo = min(a, b) when both corners lie past the
            // far end (F5:1629-1633), a box outside the field: refuse it
            if (pl.q.x0 >= (int)m.fh.nx || pl.q.y0 >= (int)m.fh.ny || pl.q.z0 >= (int)m.fh.nz)
                throw StzError(STZ_E_ARG, "query box outside the field");
            for (int b = 0; b < 7; ++b) {
                for (int lv = 0; lv < 2; ++lv) {
                    const BlockPlan& pb = lv == 0 ? pl.low[b] : pl.high[b];
                    if (!pb.valid) continue;
                    auto& w = want[(lv == 0 ? 1 : 8) + b];
                    if (w.empty()) w.assign(L.g[(lv == 0 ? 1 : 8) + b].n_chunks, 0);
                    for (int64_t ch : pb.chunks) w[(size_t)ch] = 1;
                }
            }
            o[0] = (size_t)(pl.q.x1 - pl.q.x0 + 1);
            o[1] = (size_t)(pl.q.y1 - pl.q.y0 + 1);
            o[2] = (size_t)(pl.q.z1 - pl.q.z0 + 1);
            need_base = need_mid = true;
        } else if (r.kind == STZ_REQ_LEVEL) {
            const int level = r.a[0];
            if (level < 0 || level > 3) throw StzError(STZ_E_ARG, "level");
            const int sh = 3 - level;
            o[0] = (m.fh.nx + (1u << sh) - 1) >> sh;
            o[1] = (m.fh.ny + (1u << sh) - 1) >> sh;
            o[2] = (m.fh.nz + (1u << sh) - 1) >> sh;
            if (level == 3) { full_reqs.push_back(i); continue; }
            need_base = true;
            if (level == 2) need_mid = need_lowfull = true;
        } else {
            throw StzError(STZ_E_ARG, "request kind");
        }
    }
    std::vector<OutBuf> ob(n);
    for (int i = 0; i < n; ++i) {
        if (!outs[i]) throw StzError(STZ_E_ARG, "null output");
        ob[i] = OutBuf(c, outs[i], od[3 * i] * od[3 * i + 1] * od[3 * i + 2] * sizeof(T));
    }
    ErrFlag err(c);
    if (m.fh.pred & kPredStaged) {
        // dependency.cpp's staged predictors: a block depends on reconstructed
        // neighbours of other blocks, past the planner's stencil footprint, so
        // the requests are crops of the levels decoded in full
        bool all[kStreams];
        for (int s = 0; s < kStreams; ++s) all[s] = true;
        Decoded dd;
        huff_decode(c, ds, L, all, nullptr, dd, err.ptr());
        const int lx = L.g[0].dx, ly = L.g[0].dy, lz = L.g[0].dz;
        const int mx = L.g[8].dx, my = L.g[8].dy, mz = L.g[8].dz;
        DBuf vbuf;
        const T* vb;
        if (ds->have_base) {
            vb = ds->base.as<T>();
        } else {
            vbuf = DBuf(c->pool, (size_t)lx * ly * lz * sizeof(T));
            decode_base<T>(c, m, L, dd, vbuf.as<T>(), 0);
            vb = vbuf.as<T>();
        }
        DBuf mid(c->pool, (size_t)mx * my * mz * sizeof(T)), pad, rec(c->pool, staged_scratch(L, sizeof(T)));
        const LevelArgs la = level_args(m, dd, 1, err.ptr()), ha = level_args(m, dd, 2, err.ptr());
        bool need_full = false, need_mid2 = false;
        for (int i = 0; i < n; ++i) {
            const bool lv = rq[i].kind == STZ_REQ_LEVEL;
            need_full = need_full || !lv || rq[i].a[0] == 3;
            need_mid2 = need_mid2 || !lv || rq[i].a[0] >= 2;
        }
        if (need_mid2)
            K(c, "low_decode", [&] {
                launch_staged_decode<T>(st, vb, L.g[1], la, dd.codes.as<uint16_t>(), dd.side.as<T>(), mid.as<T>(), mx,
                                        my, mz, rec.as<T>());
            });
        if (need_full) {
            pad = DBuf(c->pool, (size_t)m.fh.px * m.fh.py * m.fh.pz * sizeof(T));
            K(c, "high_decode", [&] {
                launch_staged_decode<T>(st, mid.as<T>(), L.g[8], ha, dd.codes.as<uint16_t>(), dd.side.as<T>(),
                                        pad.as<T>(), (int)m.fh.px, (int)m.fh.py, (int)m.fh.pz, rec.as<T>());
            });
        }
        for (int i = 0; i < n; ++i) {
            const size_t* o = od + 3 * i;
            K(c, "crop", [&] {
                if (rq[i].kind == STZ_REQ_LEVEL) {
                    const int level = rq[i].a[0];
                    if (level == 3)
                        launch_crop<T>(st, pad.as<T>(), (int)m.fh.px, (int)m.fh.py, (int)o[0], (int)o[1], (int)o[2], 0,
                                       0, 0, 1, (T*)ob[i].ptr);
                    else if (level == 2)
                        launch_crop<T>(st, mid.as<T>(), mx, my, (int)o[0], (int)o[1], (int)o[2], 0, 0, 0, 1,
                                       (T*)ob[i].ptr);
                    else
                        launch_crop<T>(st, vb, lx, ly, (int)o[0], (int)o[1], (int)o[2], 0, 0, 0, level == 0 ? 2 : 1,
                                       (T*)ob[i].ptr);
                } else {
                    const Box& q = plans[i].q;
                    launch_crop<T>(st, pad.as<T>(), (int)m.fh.px, (int)m.fh.py, (int)o[0], (int)o[1], (int)o[2], q.x0,
                                   q.y0, q.z0, 1, (T*)ob[i].ptr);
                }
            });
            ob[i].copy_back(c);
        }
        err.fetch(c);
        HIPCHK(hipStreamSynchronize(st));
        err.check();
        return;
    }
    if (need_base) {
        full[0] = !ds->have_base;
        for (int b = 0; b < 7; ++b) full[1 + b] = need_lowfull;
        std::vector<int64_t> lists[kStreams];
        for (int s = 1; s < kStreams; ++s)
            if (!full[s] && !want[s].empty())
                for (uint32_t ch = 0; ch < want[s].size(); ++ch)
                    if (want[s][ch]) lists[s].push_back(ch);
        Decoded dd;
        huff_decode(c, ds, L, full, lists, dd, err.ptr());
        const int lx = L.g[0].dx, ly = L.g[0].dy, lz = L.g[0].dz;
        const int mx = L.g[8].dx, my = L.g[8].dy, mz = L.g[8].dz;
        DBuf vbuf;
        const T* vb;
        if (ds->have_base) {
            vb = ds->base.as<T>();
        } else {
            vbuf = DBuf(c->pool, (size_t)lx * ly * lz * sizeof(T));
            decode_base<T>(c, m, L, dd, vbuf.as<T>(), 0);
            vb = vbuf.as<T>();
        }
        DBuf mid;
        if (need_mid) mid = DBuf(c->pool, (size_t)mx * my * mz * sizeof(T));
        const LevelArgs la = level_args(m, dd, 1, err.ptr()), ha = level_args(m, dd, 2, err.ptr());
        if (need_lowfull)
            K(c, "low_decode", [&] {
                launch_low_decode<T>(st, vb, L.g[1], la, dd.codes.as<uint16_t>(), dd.side.as<T>(), mid.as<T>(),
                                     c->grid_share);
            });
        // box jobs: low regions first, then the high boxes
        std::vector<BoxJob> lowj, highj;
        uint32_t lowb = 0, highb = 0;
        auto add = [](std::vector<BoxJob>& v, uint32_t& blocks, const Box& R, void* out) {
            const uint64_t vol = (uint64_t)(R.x1 - R.x0 + 1) * (R.y1 - R.y0 + 1) * (R.z1 - R.z0 + 1);
            BoxJob j{};
            j.x0 = R.x0; j.x1 = R.x1; j.y0 = R.y0; j.y1 = R.y1; j.z0 = R.z0; j.z1 = R.z1;
            j.blk0 = blocks;
            j.out = out;
            blocks += (uint32_t)std::min<uint64_t>(std::max<uint64_t>((vol + 1023) / 1024, 1), 2048);
            v.push_back(j);
        };
        for (int i = 0; i < n; ++i) {
            if (rq[i].kind != STZ_REQ_BOX && rq[i].kind != STZ_REQ_SLICE) continue;
            for (const Box& R : plans[i].regions)
                if (R.x0 <= R.x1 && R.y0 <= R.y1 && R.z0 <= R.z1 && !need_lowfull) add(lowj, lowb, R, nullptr);
            add(highj, highb, plans[i].q, ob[i].ptr);
        }
        const size_t jb = (lowj.size() + highj.size()) * sizeof(BoxJob);
        DBuf jobs(c->pool, jb + 16);
        if (jb) {
            uint8_t* stage = (uint8_t*)c->host_stage(jb);
            if (!lowj.empty()) std::memcpy(stage, lowj.data(), lowj.size() * sizeof(BoxJob));
            if (!highj.empty()) std::memcpy(stage + lowj.size() * sizeof(BoxJob), highj.data(), highj.size() * sizeof(BoxJob));
            HIPCHK(hipMemcpyAsync(jobs.p, stage, jb, hipMemcpyHostToDevice, st));
        }
        const BoxJob* jl = jobs.as<BoxJob>();
        if (!lowj.empty())
            K(c, "low_decode_box", [&] {
                launch_low_decode_boxes<T>(st, vb, L.g[1], la, dd.codes.as<uint16_t>(), dd.side.as<T>(), mid.as<T>(),
                                           mx, my, jl, (uint32_t)lowj.size(), lowb);
            });
        if (!highj.empty())
            K(c, "high_decode_box", [&] {
                launch_high_decode_boxes<T>(st, mid.as<T>(), L.g[8], ha, dd.codes.as<uint16_t>(), dd.side.as<T>(),
                                            jl + lowj.size(), (uint32_t)highj.size(), highb);
            });
        for (int i = 0; i < n; ++i) {
            if (rq[i].kind != STZ_REQ_LEVEL || rq[i].a[0] == 3) continue;
            const size_t* o = od + 3 * i;
            const int level = rq[i].a[0];
            if (level == 2)
                K(c, "crop", [&] { launch_crop<T>(st, mid.as<T>(), mx, my, (int)o[0], (int)o[1], (int)o[2], 0, 0, 0, 1, (T*)ob[i].ptr); });
            else
                K(c, "crop", [&] { launch_crop<T>(st, vb, lx, ly, (int)o[0], (int)o[1], (int)o[2], 0, 0, 0, level == 0 ? 2 : 1, (T*)ob[i].ptr); });
        }
        for (int i = 0; i < n; ++i)
            if (!(rq[i].kind == STZ_REQ_LEVEL && rq[i].a[0] == 3)) ob[i].copy_back(c);
        err.fetch(c);
        HIPCHK(hipStreamSynchronize(st));
        err.check();
    }
    for (int i : full_reqs) decompress_full<T>(c, ds, outs[i]);
}

// ---------------------------------------------------------------- tuner
// The configurations the tuner chooses among: predictor set (4) x base codec
// (interpolation: 2 interpolators x 6 axis orders x 4 level-eb scalings;
// the Lorenzo / regression base (interp 3, SZ3's configuration in
// sz.cpp:63-76): 3 scalings -- the fourth differs from the first only in the
// interpolation levels' alpha / beta), 51 per set; index
// k = 51 * set + j, j = 24 * interp + 4 * order + scaling below 48 and
// 48 + scaling for Lorenzo / regression, so 0..47 are F5's own predictors
// with interpolation.  Predictor sets: F5's (F5:690-748); F5 built with
// PRED_PLANE_AXIS and PRED_CUBE_AXIS (F5:579-622, F5:627-686); stz2.cpp's
// 12-point plane and 16-point cube predictors (stz2.cpp:122-196);
// dependency.cpp's staged inverse-variance predictors (dependency.cpp:162-292).
// Scalings: F5's multipliers
// (F5:2127-2129); float.cpp's 1 / 2.5 / 6.25 (float.cpp:598); one bound for
// every level; F5's multipliers with the base codec's level bounds shrinking
// by 1.5 per level up to 4x (SZ3-style).  In ABS / REL modes the largest
// multiplier maps to the requested bound, so no choice changes it.
struct TuneScale { double mb, ml, mh, alpha, beta; };
const TuneScale kScales[4] = {
    {0.7, 2.7, 8.2, 1.0, 1.0}, {1.0, 2.5, 6.25, 1.0, 1.0}, {1.0, 1.0, 1.0, 1.0, 1.0}, {0.7, 2.7, 8.2, 1.5, 4.0}};
const int kTunePred[4] = {0, kPredPlaneAxis | kPredCubeAxis, kPredBicubic12 | kPredDiag16, kPredStaged};
constexpr int kTunePerSet = 51;
constexpr int kTuneConfigs = 4 * kTunePerSet;
// one MLP output per configuration; the outputs past them (the MFMA tiles
// are 16 wide) are never chosen
static_assert(kTuneConfigs <= kMlpOut && kMlpOut - kTuneConfigs < 16, "MLP outputs");

void tune_config(int k, stz_params& p)
{
    p.pred = kTunePred[k / kTunePerSet];
    const int j = k % kTunePerSet;
    p.interp = j < 48 ? j / 24 : 3;
    p.order = j < 48 ? (j / 4) % 6 : 0;
    const TuneScale& t = kScales[j < 48 ? j % 4 : j - 48];
    p.mult_base = t.mb;
    p.mult_low = t.ml;
    p.mult_high = t.mh;
    p.alpha = t.alpha;
    p.beta = t.beta;
}

// 16 features of one 64^3 sub-block from its k_block_stats sums, relative to
// the absolute bound; logs in units of 8 octaves
void block_features(const double* S, double eb, float* f)
{
    auto lg = [](double v) { return (float)(std::log2(v + 1.0 / 1024.0) / 8.0); };
    const double n = S[0] > 0 ? S[0] : 1.0;
    const double mean = S[1] / n, var = std::max(S[2] / n - mean * mean, 0.0);
    f[0] = lg(std::sqrt(var) / eb);
    f[1] = lg((S[5] - S[4]) / eb);
    for (int a = 0; a < 3; ++a) {
        f[2 + a] = lg(S[6 + a] / n / eb);
        f[5 + a] = lg(S[9 + a] / n / eb);
        f[8 + a] = lg(S[12 + a] / n / eb);
    }
    f[11] = (float)(S[0] > 0 ? S[15] / n : 0.0);
    f[12] = (float)(S[0] > 0 ? S[16] / n : 0.0);
    f[13] = lg(S[17] / n / eb);
    const double m2 = S[2] / n, m4 = S[3] / n;
    f[14] = (float)(std::log2(m4 / (m2 * m2 + 1e-300) + 1e-30) / 8.0);
    const double lmax = std::max(S[6], std::max(S[7], S[8])), lmin = std::min(S[6], std::min(S[7], S[8]));
    f[15] = (float)(std::log2((lmax + 1e-300) / (lmin + 1e-300)) / 8.0);
    if (S[0] <= 0)
        for (int k = 2; k < 16; ++k) f[k] = 0.f;
}

// ------------------------------------------------------------ batches
// Several fields in one call.  Their residual levels run as one persistent
// launch per level fed by a device work queue (k_level_multi); its launch
// geometry (waves per workgroup: the tile height; workgroups per CU) is
// chosen per dtype, level mode and size class by timing every candidate once
// on the live batch (the runs are idempotent or their histograms are
// cleared), and kept in the context.  Results are bit-identical to
// single-field calls: every point runs the same arithmetic.
struct JobSet {
    std::vector<LevelJob> h[2];  // waves 4, waves 8
    uint32_t total[2] = {0, 0};
    DBuf d[2];
    DBuf counter;
};

void jobs_finish(stz_ctx* c, JobSet& js)
{
    for (int v = 0; v < 2; ++v) {
        const int waves = v == 0 ? 4 : 8;
        uint32_t item = 0;
        for (LevelJob& j : js.h[v]) {
            j.item0 = item;
            j.ntiles = j.g.dx >= 4 ? level_tiles(j.g, waves) : 0;
            j.nedge = level_edge_groups(j.g, waves);
            item += j.ntiles + j.nedge;
        }
        js.total[v] = item;
        const size_t bytes = js.h[v].size() * sizeof(LevelJob);
        js.d[v] = DBuf(c->pool, bytes + 64);
        void* hs = c->host_stage(bytes);
        std::memcpy(hs, js.h[v].data(), bytes);
        HIPCHK(hipMemcpyAsync(js.d[v].p, hs, bytes, hipMemcpyHostToDevice, c->stream));
    }
    js.counter = DBuf(c->pool, 64);
}

template <typename T>
void jobs_launch(stz_ctx* c, int mode, JobSet& js, const stz_ctx::LaunchCfg& cfg)
{
    const int v = cfg.waves == 8 ? 1 : 0;
    K(c, "counter_zero", [&] { launch_fill_u32(c->stream, js.counter.as<uint32_t>(), 1, 0u); });
    K(c, mode == kLowEnc ? "low_encode_multi" : mode == kHighEnc ? "high_encode_multi"
         : mode == kLowDec ? "low_decode_multi" : "high_decode_multi", [&] {
        bool var = false;
        for (const LevelJob& j : js.h[v]) var = var || j.a.opt != 0;
        launch_level_multi<T>(c->stream, mode, js.d[v].as<LevelJob>(), (uint32_t)js.h[v].size(), js.total[v],
                              js.counter.as<uint32_t>(), cfg.waves, cfg.grid, var);
    });
}

// the autotuner's key of a level launch: dtype, mode and size class
template <typename T>
std::string level_cfg_key(int mode, const JobSet& js)
{
    uint64_t pts = 0;
    for (const LevelJob& j : js.h[0]) pts += (uint64_t)j.g.dx * j.g.dy * j.g.dz;
    int lg = 0;
    while ((1ull << (lg + 1)) <= pts) ++lg;
    char key[64];
    std::snprintf(key, sizeof(key), "%s/%s/2^%d", sizeof(T) == 4 ? "f32" : "f64",
                  mode == kLowEnc ? "low_encode" : mode == kHighEnc ? "high_encode" : mode == kLowDec ? "low_decode"
                                                                                           : "high_decode", lg);
    return key;
}

// whether level_cfg would time candidates for this launch
template <typename T>
bool level_cfg_pending(stz_ctx* c, int mode, const JobSet& js)
{
    return c->autotune && c->tuned.find(level_cfg_key<T>(mode, js)) == c->tuned.end();
}

// the tuned geometry for this launch; times the candidates the first time
// (reset() restores the launch's inputs before each candidate)
template <typename T, typename R>
stz_ctx::LaunchCfg level_cfg(stz_ctx* c, int mode, JobSet& js, R&& reset)
{
    const std::string key = level_cfg_key<T>(mode, js);
    auto it = c->tuned.find(key);
    if (it != c->tuned.end()) return it->second;
    stz_ctx::LaunchCfg best;
    best.waves = 4;
    best.grid = (uint32_t)std::max(1, c->n_cu) * 2;
    if (!c->autotune) return best;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    best.ms = 1e30f;
    std::string trials;
    const bool prof = c->prof;
    c->prof = false;
    for (int waves : {4, 8})
        for (int per_cu : {1, 2, 4}) {
            stz_ctx::LaunchCfg cand;
            cand.waves = waves;
            cand.grid = (uint32_t)(std::max(1, c->n_cu) * per_cu * 4 / waves);
            reset();
            HIPCHK(hipEventRecord(e0, c->stream));
            jobs_launch<T>(c, mode, js, cand);
            HIPCHK(hipEventRecord(e1, c->stream));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            char t[96];
            std::snprintf(t, sizeof(t), "%s{\"waves\": %d, \"grid\": %u, \"ms\": %.4f}", trials.empty() ? "" : ", ",
                          waves, cand.grid, ms);
            trials += t;
            if (ms < best.ms) {
                best = cand;
                best.ms = ms;
            }
        }
    c->prof = prof;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    best.trials = "[" + trials + "]";
    reset();
    c->tuned[key] = best;
    return best;
}

// upload a host vector into a pool buffer (staged copy on the context stream)
template <typename V>
DBuf upload_vec(stz_ctx* c, const std::vector<V>& v)
{
    const size_t bytes = v.size() * sizeof(V);
    DBuf d(c->pool, bytes + 64);
    if (bytes) {
        void* hs = c->host_stage(bytes);
        std::memcpy(hs, v.data(), bytes);
        HIPCHK(hipMemcpyAsync(d.p, hs, bytes, hipMemcpyHostToDevice, c->stream));
    }
    return d;
}

// The learned tuner's choice for every field of a batch: the sub-block
// statistics of all fields, one MLP launch over all their blocks, then per
// field the same weighted vote as a single-field call.
template <typename T>
void learned_choice_batch(stz_ctx* c, std::vector<FieldEnc<T>>& fe, std::vector<float>* confs = nullptr)
{
    if (!c->have_tuner) throw StzError(STZ_E_ARG, "no tuner model set (stz_tune_set_model)");
    hipStream_t st = c->stream;
    const size_t n = fe.size();
    std::vector<size_t> off(n + 1, 0);
    for (size_t i = 0; i < n; ++i)
        off[i + 1] = off[i] + ((fe[i].nx + 63) / 64) * ((fe[i].ny + 63) / 64) * ((fe[i].nz + 63) / 64);
    const size_t nb = off[n];
    std::vector<double> ebabs(n);
    DBuf sums(c->pool, nb * kFeatSums * sizeof(double)), part(c->pool, nb * kStatParts * kFeatSums * sizeof(double));
    for (size_t i = 0; i < n; ++i) {
        const FieldEnc<T>& f = fe[i];
        ebabs[i] = f.mode == STZ_EB_REL ? f.eb * (f.mm_host[1] - f.mm_host[0]) : f.eb;
        if (!(ebabs[i] > 0)) ebabs[i] = 1e-300;
        K(c, "block_stats", [&] {
            launch_block_stats<T>(st, f.din, (int)f.nx, (int)f.ny, (int)f.nz, ebabs[i],
                                  sums.as<double>() + off[i] * kFeatSums,
                                  part.as<double>() + off[i] * kStatParts * kFeatSums);
        });
    }
    double* hs = (double*)c->host_stage(nb * kFeatSums * sizeof(double));
    HIPCHK(hipMemcpyAsync(hs, sums.p, nb * kFeatSums * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::vector<float> X(nb * kMlpIn);
    for (size_t i = 0; i < n; ++i)
        for (size_t b = off[i]; b < off[i + 1]; ++b) block_features(hs + b * kFeatSums, ebabs[i], &X[b * kMlpIn]);
    DBuf x(c->pool, X.size() * 4), y(c->pool, nb * kMlpOut * 4);
    float* hx = (float*)c->host_stage(X.size() * 4 + nb * kMlpOut * 4);
    std::memcpy(hx, X.data(), X.size() * 4);
    HIPCHK(hipMemcpyAsync(x.p, hx, X.size() * 4, hipMemcpyHostToDevice, st));
    K(c, "mlp_forward", [&] { launch_mlp_forward(st, c->tuner_w.as<float>(), x.as<float>(), (uint32_t)nb, y.as<float>()); });
    float* hy = hx + X.size();
    HIPCHK(hipMemcpyAsync(hy, y.p, nb * kMlpOut * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (confs) confs->assign(n, 0.f);
    for (size_t i = 0; i < n; ++i) {
        double wsum = 0;
        for (size_t b = off[i]; b < off[i + 1]; ++b) wsum += hs[b * kFeatSums];
        std::vector<double> lp(kTuneConfigs, 0.0), pr(kTuneConfigs, 0.0);
        double e[kTuneConfigs];
        for (size_t b = off[i]; b < off[i + 1]; ++b) {
            const double w = wsum > 0 ? hs[b * kFeatSums] / wsum : 1.0 / (double)(off[i + 1] - off[i]);
            const float* r = hy + b * kMlpOut;
            double mx = r[0];
            for (int k = 1; k < kTuneConfigs; ++k) mx = std::max(mx, (double)r[k]);
            double se = 0;
            for (int k = 0; k < kTuneConfigs; ++k) {
                e[k] = std::exp(r[k] - mx);
                se += e[k];
            }
            const double lse = mx + std::log(se);
            for (int k = 0; k < kTuneConfigs; ++k) {
                lp[k] += w * (r[k] - lse);
                pr[k] += w * e[k] / se;
            }
        }
        int best = 0;
        for (int k = 1; k < kTuneConfigs; ++k)
            if (lp[k] > lp[best]) best = k;
        tune_config(best, fe[i].prm);
        if (confs) (*confs)[i] = (float)pr[best];
    }
}

// Compress n fields of one dtype.  Phases over all fields, with one wait
// between them; within a phase each step is one launch for every field: the
// base-codec passes (coarse passes one workgroup per field, then one launch
// per fine-pass round), the two residual levels (one field: the tiled
// kernels; many: one persistent launch per level), the histogram compaction,
// the header/codebook upload (one copy + one segment-scatter launch) and the
// marked Huffman encode (lengths, chunk scan, pack).
template <typename T>
void compress_fields(stz_ctx* c, int dtype, int n, const void* const* fields, const size_t* dims, double eb,
                     int mode, const stz_params& prm, stz_dstream* const* out, std::vector<char>* low_conf)
{
    hipStream_t st = c->stream;
    auto pool = c->pool;
    DBuf hist_all, mm_all;  // a batch's histograms and value ranges (outlive fe's views)
    std::vector<FieldEnc<T>> fe(n);
    for (int i = 0; i < n; ++i) {
        if (!fields[i]) throw StzError(STZ_E_ARG, "null field");
        FieldEnc<T>& f = fe[i];
        f.dtype = dtype;
        f.mode = mode;
        f.eb = eb;
        f.prm = prm;
        f.nx = dims[3 * i];
        f.ny = dims[3 * i + 1];
        f.nz = dims[3 * i + 2];
        f.ds = out[i];
    }
    if (n == 1) enc_phase0(c, fe[0], (const T*)fields[0]);
    else enc_phase0_batch(c, fe, fields, hist_all, mm_all);
    // every field's range is on the host once the last one is (stream order);
    // the gathers queued after them keep the GPU busy meanwhile
    wait_event(c->ev_range);
    // fields the caller takes back for its exhaustive trial; kept alive to
    // the end, after every launch that may still read their buffers
    std::vector<FieldEnc<T>> dropped;
    if (prm.tuner == 1) {
        // below the context's confidence threshold a field is marked in
        // *low_conf and left to the caller's trial; the others go on
        std::vector<float> conf;
        learned_choice_batch<T>(c, fe, &conf);
        for (int i = 0; i < n; ++i) {
            fe[i].ds->tuned = 1;
            fe[i].ds->tune_conf = conf[i];
        }
        if (low_conf) {
            low_conf->assign(n, 0);
            std::vector<FieldEnc<T>> keep;
            for (int i = 0; i < n; ++i) {
                (*low_conf)[i] = conf[i] < c->tune_min_conf;
                ((*low_conf)[i] ? dropped : keep).push_back(std::move(fe[i]));
            }
            fe.swap(keep);
            n = (int)fe.size();
            if (n == 0) {
                HIPCHK(hipStreamSynchronize(st));
                return;
            }
        }
    }
    for (int i = 0; i < n; ++i) enc_phase1(c, fe[i]);
    // base level: coarse passes of every field, then the fine-pass rounds;
    // Lorenzo-coded bases (interp 2) in one tile launch
    {
        std::vector<FusedEncJob> fj;
        std::vector<std::vector<PassEncJob>> rounds;
        std::vector<LorenzoJob> lz;
        uint32_t lz_wgs = 0;
        for (int i = 0; i < n; ++i) {
            FieldEnc<T>& f = fe[i];
            const HostLayout& L = f.L;
            if (f.prm.interp >= 2) {
                LorenzoJob j{};
                j.reg = f.prm.interp == 3;
                j.lo = f.vmin;
                j.hi = f.vmax;
                j.v = f.vbase.p;
                j.codes = f.codes.template as<uint16_t>() + L.code_off[0];
                j.side = f.side.template as<T>() + L.code_off[0];
                j.hist = f.hist.template as<uint32_t>();
                j.bx = L.g[0].dx; j.by = L.g[0].dy; j.bz = L.g[0].dz;
                j.eb = f.ebl[0];
                j.recip = 1.0 / f.ebl[0];
                j.wg0 = lz_wgs;
                lz_wgs += lorenzo_tiles(j.bx, j.by, j.bz);
                lz.push_back(j);
                continue;
            }
            const auto passes = make_passes(L.g[0].dx, L.g[0].dy, L.g[0].dz, f.ebl[0], f.prm.interp, f.prm.order,
                                            f.prm.alpha, f.prm.beta, 0);
            uint16_t* c0 = f.codes.template as<uint16_t>() + L.code_off[0];
            T* s0 = f.side.template as<T>() + L.code_off[0];
            size_t r = 0;
            run_passes(passes, sizeof(T), [&](const InterpPassList& fl) {
                if (fl.n <= 0) return;
                FusedEncJob j{};
                j.l = fl;
                j.v = f.vbase.p;
                j.codes = c0;
                j.side = s0;
                j.hist = f.hist.template as<uint32_t>();
                fj.push_back(j);
            }, [&](const InterpPass& p) {
                if (!p.n) return;
                if (rounds.size() <= r) rounds.resize(r + 1);
                PassEncJob j{};
                j.p = p;
                j.v = f.vbase.p;
                j.codes = c0;
                j.side = s0;
                j.hist = f.hist.template as<uint32_t>();
                j.nwg = (uint32_t)std::min<uint64_t>((p.n + 255) / 256, kPassEncMaxWgs);
                rounds[r].push_back(j);
                ++r;
            }, n > 1);
        }
        if (lz.size() == 1) {
            K(c, "lorenzo_encode", [&] { launch_lorenzo<T>(st, lz[0], true); });
        } else if (!lz.empty()) {
            DBuf dlz = upload_vec(c, lz);
            K(c, "lorenzo_encode", [&] {
                launch_lorenzo_multi<T>(st, dlz.as<LorenzoJob>(), (uint32_t)lz.size(), lz_wgs, true);
            });
        }
        DBuf dfj = upload_vec(c, fj);
        K(c, "interp_encode_fused", [&] { launch_interp_encode_fused_multi<T>(st, dfj.as<FusedEncJob>(), (uint32_t)fj.size()); });
        for (auto& rj : rounds) {
            if (rj.size() == 1) {
                // one field: the pass goes by value (no job upload, whose copy
                // costs more GPU time than a small pass)
                const PassEncJob& q = rj[0];
                K(c, "interp_encode", [&] { launch_interp_encode<T>(st, q.p, (T*)q.v, q.codes, (T*)q.side, q.hist); });
                continue;
            }
            uint32_t wg = 0;
            for (PassEncJob& q : rj) {
                q.wg0 = wg;
                wg += q.nwg;
            }
            DBuf dr = upload_vec(c, rj);
            K(c, "interp_encode", [&] { launch_interp_encode_multi<T>(st, dr.as<PassEncJob>(), (uint32_t)rj.size(), wg); });
        }
    }
    // Histograms -> host (F5:1271 / F5:1402 preprocess_encode_fast).  With
    // few fields the base and low streams' histograms are final once the low
    // level has run: they are compacted and read back then, and their
    // codebooks are built on the host while the GPU runs the high level.
    constexpr uint32_t kFirst = FieldEnc<T>::kFirst;
    DBuf compact(pool, (size_t)n * kStreams * 65536 * 8), counts(pool, (size_t)n * 64 * 4);
    std::vector<EncJob> hj(n);
    for (int i = 0; i < n; ++i) {
        hj[i].hist = fe[i].hist.template as<uint32_t>();
        hj[i].compact = compact.as<uint64_t>() + (size_t)i * kStreams * 65536;
        hj[i].counts = counts.as<uint32_t>() + (size_t)i * 64;
        fe[i].lists.assign(kStreams, {});
    }
    DBuf dhj = upload_vec(c, hj);
    const bool early = n <= 4;
    const size_t hs_bytes = (size_t)n * 256 + (size_t)n * kStreams * kFirst * 8;
    uint8_t* hs_lo = early ? (uint8_t*)c->host_stage(hs_bytes) : nullptr;
    uint8_t* hs_hi = (uint8_t*)c->host_stage(hs_bytes);
    // compact streams [s0, s0+ns) of every field and copy counts + first entries into hs
    auto readback = [&](uint8_t* hs, int s0, int ns, hipStream_t rs) {
        K(c, "hist_compact", [&] { launch_hist_compact_multi(rs, dhj.as<EncJob>(), (uint32_t)n, s0, ns); });
        HIPCHK(hipMemcpyAsync(hs, counts.p, (size_t)n * 256, hipMemcpyDeviceToHost, rs));
        uint64_t* hf = (uint64_t*)(hs + (size_t)n * 256);
        for (int i = 0; i < n; ++i) {
            const size_t row = (size_t)i * kStreams + s0;
            HIPCHK(hipMemcpy2DAsync(hf + row * kFirst, kFirst * 8, compact.as<uint64_t>() + row * 65536, 65536 * 8,
                                    kFirst * 8, (size_t)ns, hipMemcpyDeviceToHost, rs));
        }
    };
    // after the copies have landed: symbol lists, then codebooks on the host workers
    auto books = [&](const uint8_t* hs, int s0, int ns) {
        const uint64_t* hf = (const uint64_t*)(hs + (size_t)n * 256);
        (n * ns >= 64 ? c->tasks_wide() : c->tasks()).run(n * ns, [&](int k) {
            const int i = k / ns, s2 = s0 + k % ns;
            FieldEnc<T>& f = fe[i];
            const uint32_t m = ((const uint32_t*)(hs + (size_t)i * 256))[s2];
            f.lists[s2].resize(m);
            const uint64_t* src = hf + ((size_t)i * kStreams + s2) * kFirst;
            if (m <= kFirst) std::memcpy(f.lists[s2].data(), src, (size_t)m * 8);
            else
                HIPCHK(hipMemcpy(f.lists[s2].data(), compact.as<uint64_t>() + ((size_t)i * kStreams + s2) * 65536,
                                 (size_t)m * 8, hipMemcpyDeviceToHost));
            build_codebook(f.lists[s2].data(), m, kLmax, f.ds->meta.cb[s2]);
        });
    };
    // on an error after the low readback forked to the second stream,
    // `stream` still waits for it before the buffers go back to the pool
    struct ForkGuard {
        stz_ctx* c;
        bool on = false;
        ~ForkGuard()
        {
            if (on) (void)hipStreamWaitEvent(c->stream, c->ev_fork, 0);
        }
    } fork_guard{c};
    // residual levels
    for (int level = 1; level <= 2; ++level) {
        if (early && level == 2) {
            // the base and low streams' histograms are final: compacted and
            // read back on the second stream (beside the high level; on
            // `stream` when profiling), their codebooks built meanwhile
            hipStream_t rs = st;
            if (!c->prof) {
                rs = c->second();
                HIPCHK(hipEventRecord(c->ev_items, st));
                HIPCHK(hipStreamWaitEvent(rs, c->ev_items, 0));
            }
            readback(hs_lo, 0, 8, rs);
            HIPCHK(hipEventRecord(c->ev_fork, rs));
            fork_guard.on = rs != st;
        }
        // one field, and fields under the staged predictors (a learned batch
        // mixes them with the others): launches of their own per field
        {
            for (int i = 0; i < n; ++i) {
                FieldEnc<T>& f = fe[i];
                if (n > 1 && !(f.prm.pred & kPredStaged)) continue;
                const LevelArgs a = enc_level_args(f, level);
                if (f.prm.pred & kPredStaged) {
                    // dependency.cpp's staged predictors: three launches per
                    // level; only the low level's 2x grid is written (the
                    // high level's coarser grid)
                    T* out = level == 1 ? f.mid.template as<T>() : nullptr;
                    DBuf rec(pool, staged_scratch(f.L, sizeof(T)));
                    K(c, level == 1 ? "low_encode" : "high_encode", [&] {
                        launch_staged_encode<T>(st, f.full, (int)f.px, (int)f.py, level == 1 ? 4 : 2,
                                                level == 1 ? f.vbase.template as<T>() : f.mid.template as<T>(),
                                                f.L.g[level == 1 ? 1 : 8], a, f.codes.template as<uint16_t>(),
                                                f.side.template as<T>(),
                                                f.hist.template as<uint32_t>() + (level == 1 ? 1 : 8) * 65536, out,
                                                rec.template as<T>());
                    });
                } else if (level == 1) {
                    K(c, "low_encode", [&] {
                        launch_low_encode<T>(st, f.full, (int)f.px, (int)f.py, (int)f.pz, f.vbase.template as<T>(),
                                             f.L.g[1], a, f.codes.template as<uint16_t>(), f.side.template as<T>(),
                                             f.hist.template as<uint32_t>() + 65536, f.mid.template as<T>());
                    });
                } else {
                    K(c, "high_encode", [&] {
                        launch_high_encode<T>(st, f.full, (int)f.px, (int)f.py, (int)f.pz, f.mid.template as<T>(),
                                              f.L.g[8], a, f.codes.template as<uint16_t>(), f.side.template as<T>(),
                                              f.hist.template as<uint32_t>() + 8 * 65536);
                    });
                }
            }
            if (n == 1) continue;
        }
        // the other fields: one persistent launch for all of them
        const int md = level == 1 ? kLowEnc : kHighEnc;
        JobSet js;
        for (int v = 0; v < 2; ++v)
            for (int i = 0; i < n; ++i) {
                FieldEnc<T>& f = fe[i];
                if (f.prm.pred & kPredStaged) continue;
                LevelJob j{};
                j.full = f.full;
                j.px = (int)f.px;
                j.py = (int)f.py;
                j.ref = level == 1 ? f.vbase.p : f.mid.p;
                j.mid = level == 1 ? f.mid.p : nullptr;
                j.codes = f.codes.template as<uint16_t>();
                j.side = f.side.p;
                j.hist = f.hist.template as<uint32_t>() + (level == 1 ? 1 : 8) * 65536;
                j.g = f.L.g[level == 1 ? 1 : 8];
                j.a = enc_level_args(f, level);
                js.h[v].push_back(j);
            }
        if (js.h[0].empty()) continue;
        jobs_finish(c, js);
        const stz_ctx::LaunchCfg cfg = level_cfg<T>(c, md, js, [&] {
            for (int i = 0; i < n; ++i)
                if (!(fe[i].prm.pred & kPredStaged))
                    launch_fill_u32(st, fe[i].hist.template as<uint32_t>() + (level == 1 ? 1 : 8) * 65536, 7 * 65536,
                                    0u);
        });
        jobs_launch<T>(c, md, js, cfg);
    }
    if (early) {
        HIPCHK(hipStreamWaitEvent(st, c->ev_fork, 0));  // the low readback's buffers, before they are reused
        fork_guard.on = false;
        readback(hs_hi, 8, kStreams - 8, st);
        HIPCHK(hipEventRecord(c->ev_hist, st));
        wait_event(c->ev_fork);
        books(hs_lo, 0, 8);  // overlaps the high level on the GPU
        wait_event(c->ev_hist);
        books(hs_hi, 8, kStreams - 8);
    } else {
        readback(hs_hi, 0, kStreams, st);
        HIPCHK(hipEventRecord(c->ev_hist, st));
        wait_event(c->ev_hist);
        books(hs_hi, 0, kStreams);
    }
    for (int i = 0; i < n; ++i) enc_phase3_headers(fe[i]);
    // payloads: zeroed, host-written sections in one upload + scatter, encoder tables, marked encode
    std::vector<DBuf> enc(n), pre_b(n), pre_u(n), tot_b(n), tot_u(n);
    std::vector<EncJob> ej(n);
    std::vector<Segment> segs;
    // one upload: the host-written payload sections, then the segment list,
    // then the encoder-table entries (field, stream<<16|symbol, code|len<<24;
    // (stream<<16|symbol, code|len<<24) pairs for one field)
    size_t img_bytes = 0, nsym = 0;
    for (int i = 0; i < n; ++i) {
        const Meta& m = fe[i].ds->meta;
        img_bytes += align_up(sizeof(FileHeader), 16);
        for (int s2 = 0; s2 < kStreams; ++s2) {
            img_bytes += align_up(m.sh[s2].off_sub_bits, 16);
            nsym += m.cb[s2].sym.size();
        }
    }
    const uint32_t ew = n == 1 ? 2u : 3u;
    const size_t seg_off = img_bytes, enc_off = align_up(seg_off + (size_t)n * (kStreams + 1) * sizeof(Segment), 16);
    const size_t up_bytes = enc_off + nsym * ew * 4;
    uint8_t* img = (uint8_t*)c->host_stage(up_bytes + 64);
    uint32_t* encl = (uint32_t*)(img + enc_off);
    size_t ne = 0;
    size_t io = 0;
    uint32_t chunk0 = 0;
    for (int i = 0; i < n; ++i) {
        FieldEnc<T>& f = fe[i];
        stz_dstream* ds = f.ds;
        Meta& m = ds->meta;
        const FileHeader& fh = m.fh;
        const uint64_t total = fh.total_bytes;
        ds->payload = DBuf(pool, align_up(total, 4) + 64);
        ds->bytes = total;
        ds->pool = pool;
        ds->device = c->device;
        ds->have_tabs = false;
        uint8_t* pay = ds->payload.as<uint8_t>();
        K(c, "payload_zero", [&] { launch_fill_u32(st, (uint32_t*)pay, align_up(total, 4) / 4, 0u); });
        std::memcpy(img + io, &fh, sizeof(fh));
        segs.push_back(Segment{io, pay, sizeof(fh)});
        io += align_up(sizeof(FileHeader), 16);
        for (int s2 = 0; s2 < kStreams; ++s2) {
            const StreamHeader& sh = m.sh[s2];
            uint8_t* b = img + io;
            std::memset(b, 0, sh.off_sub_bits);
            std::memcpy(b, &sh, sizeof(sh));
            uint32_t* cbw = (uint32_t*)(b + sh.off_codebook);
            for (size_t k = 0; k < m.cb[s2].sym.size(); ++k)
                cbw[k] = (uint32_t)m.cb[s2].sym[k] | ((uint32_t)m.cb[s2].len[k] << 16);
            std::memcpy(b + sh.off_chunk_sub, f.L.chunk_sub[s2].data(), 4 * f.L.chunk_sub[s2].size());
            segs.push_back(Segment{io, pay + fh.stream_off[s2], sh.off_sub_bits});
            io += align_up(sh.off_sub_bits, 16);
            for (size_t k = 0; k < m.cb[s2].sym.size(); ++k) {
                if (ew == 3) encl[ne++] = (uint32_t)i;
                encl[ne++] = ((uint32_t)s2 << 16) | m.cb[s2].sym[k];
                encl[ne++] = m.cb[s2].code[k] | ((uint32_t)m.cb[s2].len[k] << 24);
            }
        }
        StreamSet ss = make_streamset(f.L, m, f.dtype);
        // encoder-table window staged in LDS by the encode kernels: the used
        // symbols, or the kEncWindow of them nearest the centre
        for (int s2 = 0; s2 < kStreams; ++s2) {
            const std::vector<uint16_t>& sym = m.cb[s2].sym;  // ascending
            StreamDesc& d = ss.s[s2];
            d.enc_lo = 0;
            d.enc_n = 0;
            if (sym.empty()) continue;
            const uint32_t a = sym.front(), b = sym.back();
            if (b - a + 1 <= kEncWindow) {
                d.enc_lo = a;
                d.enc_n = b - a + 1;
            } else {
                // the centre window, moved inside [a, b]
                const uint32_t c0 = (uint32_t)kRadius - kEncWindow / 2;
                d.enc_lo = c0 < a ? a : (c0 > b + 1 - kEncWindow ? b + 1 - kEncWindow : c0);
                d.enc_n = kEncWindow;
            }
        }
        enc[i] = DBuf(pool, (size_t)kStreams * 65536 * 4);
        pre_b[i] = DBuf(pool, (size_t)ss.total_sub * 4 + 16);
        pre_u[i] = DBuf(pool, (size_t)ss.total_sub * 4 + 16);
        tot_b[i] = DBuf(pool, (size_t)ss.total_chunks * 8 + 16);
        tot_u[i] = DBuf(pool, (size_t)ss.total_chunks * 8 + 16);
        EncJob& j = ej[i];
        j.ss = ss;
        j.codes = f.codes.template as<uint16_t>();
        j.side = f.side.p;
        j.enc = enc[i].as<uint32_t>();
        j.payload = pay;
        j.sub_pre_bits = pre_b[i].as<uint32_t>();
        j.sub_pre_un = pre_u[i].as<uint32_t>();
        j.chunk_tot_bits = tot_b[i].as<uint64_t>();
        j.chunk_tot_un = tot_u[i].as<uint64_t>();
        j.chunk0 = chunk0;
        chunk0 += ss.total_chunks;
    }
    std::memcpy(img + seg_off, segs.data(), segs.size() * sizeof(Segment));
    DBuf dimg(pool, up_bytes + 64);
    HIPCHK(hipMemcpyAsync(dimg.p, img, up_bytes, hipMemcpyHostToDevice, st));
    uint8_t* dup = dimg.as<uint8_t>();
    K(c, "copy_segments", [&] { launch_copy_segments(st, dup, (const Segment*)(dup + seg_off), (uint32_t)segs.size()); });
    DBuf dej;
    if (n == 1) {
        K(c, "enc_scatter", [&] { launch_enc_scatter(st, (const uint32_t*)(dup + enc_off), (uint32_t)nsym, ej[0].enc); });
    } else {
        dej = upload_vec(c, ej);
        K(c, "enc_scatter", [&] {
            launch_enc_scatter_multi(st, (const uint32_t*)(dup + enc_off), (uint32_t)nsym, dej.as<EncJob>());
        });
    }
    // marked encode (F5:1412 encode_marked_fast): lengths -> chunk scan -> pack
    // one field: the single-stream-set kernels, whose buffers arrive as
    // kernel pointer arguments (global accesses rather than the generic ones
    // a pointer read out of a job record gets)
    if (n == 1) {
        const EncJob& j = ej[0];
        K(c, "huff_encode", [&] {
            launch_huff_len(st, j.ss, j.codes, j.enc, j.payload, j.sub_pre_bits, j.sub_pre_un, j.chunk_tot_bits,
                            j.chunk_tot_un);
            launch_huff_scan(st, j.ss, j.chunk_tot_bits, j.chunk_tot_un, j.payload);
            launch_huff_pack(st, j.ss, j.codes, j.side, j.enc, j.sub_pre_bits, j.sub_pre_un, j.payload);
        });
    } else {
        K(c, "huff_encode", [&] { launch_huff_encode_multi(st, dej.as<EncJob>(), (uint32_t)n, chunk0); });
    }
    // No wait here: the bit pack and the copies before it run on in stream
    // order.  The field was last read by the high level, which the histogram
    // wait above follows; the staging arena this call filled is reused only
    // once the stream has passed it (stage_reset); the pool's buffers freed
    // here are reused only by work ordered after this call on `stream`.
}

template <typename T>
void compress_batch_impl(stz_ctx* c, int dtype, int n, const void* const* fields, const size_t* dims, double eb,
                         int mode, const stz_params& prm, stz_dstream** out)
{
    std::vector<std::unique_ptr<stz_dstream>> dss(n);
    std::vector<stz_dstream*> raw(n);
    for (int i = 0; i < n; ++i) {
        dss[i].reset(new stz_dstream());
        raw[i] = dss[i].get();
    }
    if (prm.tuner == 2) {
        // exhaustive trial over the whole batch: every configuration, each
        // field keeps its smallest payload
        std::vector<std::unique_ptr<stz_dstream>> tmp(n);
        std::vector<stz_dstream*> traw(n);
        for (int k = 0; k < kTuneConfigs; ++k) {
            stz_params pk = prm;
            pk.tuner = 0;
            tune_config(k, pk);
            for (int i = 0; i < n; ++i) {
                tmp[i].reset(new stz_dstream());
                traw[i] = tmp[i].get();
            }
            compress_fields<T>(c, dtype, n, fields, dims, eb, mode, pk, traw.data(), nullptr);
            for (int i = 0; i < n; ++i)
                if (k == 0 || traw[i]->bytes < raw[i]->bytes) std::swap(dss[i], tmp[i]), raw[i] = dss[i].get();
        }
        for (int i = 0; i < n; ++i) raw[i]->tuned = 2;
    } else if (prm.tuner == 1) {
        // the model's choice for every field in one pass; the fields below
        // the confidence threshold get the exhaustive trial, as one batch per
        // configuration (each keeps its first smallest payload, as a
        // single-field trial does)
        std::vector<char> low;
        compress_fields<T>(c, dtype, n, fields, dims, eb, mode, prm, raw.data(), &low);
        std::vector<int> lows;
        for (int i = 0; i < n; ++i)
            if (low[i]) lows.push_back(i);
        const int m = (int)lows.size();
        if (m > 0) {
            std::vector<const void*> lf(m);
            std::vector<size_t> ld(3 * (size_t)m);
            std::vector<float> conf(m);
            std::vector<uint64_t> best(m, ~0ull);
            for (int j = 0; j < m; ++j) {
                lf[j] = fields[lows[j]];
                for (int a = 0; a < 3; ++a) ld[3 * j + a] = dims[3 * lows[j] + a];
                conf[j] = raw[lows[j]]->tune_conf;
            }
            for (int k = 0; k < kTuneConfigs; ++k) {
                stz_params pk = prm;
                pk.tuner = 0;
                tune_config(k, pk);
                std::vector<std::unique_ptr<stz_dstream>> tmp(m);
                std::vector<stz_dstream*> traw(m);
                for (int j = 0; j < m; ++j) {
                    tmp[j].reset(new stz_dstream());
                    traw[j] = tmp[j].get();
                }
                compress_fields<T>(c, dtype, m, lf.data(), ld.data(), eb, mode, pk, traw.data(), nullptr);
                for (int j = 0; j < m; ++j)
                    if (traw[j]->bytes < best[j]) {
                        best[j] = traw[j]->bytes;
                        std::swap(*raw[lows[j]], *traw[j]);
                    }
            }
            for (int j = 0; j < m; ++j) {
                raw[lows[j]]->tuned = 2;
                raw[lows[j]]->tune_conf = conf[j];
            }
        }
    } else {
        compress_fields<T>(c, dtype, n, fields, dims, eb, mode, prm, raw.data(), nullptr);
    }
    for (int i = 0; i < n; ++i) out[i] = dss[i].release();
}

// Batch decode: per field the Huffman decode and the base passes, then the
// low and high levels of all fields as one persistent launch each (fields
// under the staged predictors: their three stage launches per level, one
// field after another).  The launch sequence is captured as a HIP graph on
// the first call and replayed while the call names the same streams and
// outputs.
struct BatchPlan {
    std::vector<stz_dstream*> ds;
    std::vector<uint64_t> gens;
    std::vector<void*> outs;
    int dtype = 0;
    struct Field {
        HostLayout L;
        StreamSet ss;
        bool stage = false;
        Decoded dd;
        DBuf items, vb, mid;
    };
    std::vector<Field> f;
    DBuf err;
    // Huffman decode of all fields: prep over every chunk, then the staged and
    // the unstaged fields as one launch each; base passes: the fused coarse
    // passes one workgroup per field, then one launch per fine-pass round
    DBuf dec_all, dec_st, dec_un;
    uint32_t n_st = 0, n_un = 0, wg_st = 0, wg_un = 0, chunks = 0;
    DBuf fused;
    uint32_t n_fused = 0;
    DBuf lz;                   // Lorenzo-coded bases: tile jobs
    uint32_t n_lz = 0, lz_wgs = 0;
    std::vector<DBuf> rounds;
    std::vector<std::pair<uint32_t, uint32_t>> round_n;  // jobs, workgroups
    JobSet low, high;
    stz_ctx::LaunchCfg cfg_low, cfg_high;
    // fields under the staged predictors: their levels in three stage
    // launches each (one scratch for all: they run one after another)
    struct Staged {
        const void* vb;
        void* mid;
        Geom g1, g8;
        LevelArgs la, ha;
        const uint16_t* codes;
        const void* side;
        void* out;
        int nx, ny, nz;
    };
    std::vector<Staged> stg;
    DBuf rec;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~BatchPlan()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

// the full Huffman decode of one stream set into preallocated buffers
void plan_huff_setup(stz_ctx* c, stz_dstream* ds, BatchPlan::Field& F)
{
    ensure_tabs(c, ds);
    const Meta& m = ds->meta;
    layout_of(m, F.L);
    const HostLayout& L = F.L;
    F.ss = make_streamset(L, m, (int)m.fh.dtype);
    double wbits = 0, wsym = 0;
    for (int s = 0; s < kStreams; ++s) {
        F.ss.s[s].item_count = L.n_sub[s];
        F.dd.code_base[s] = L.code_off[s];
        const double bps = L.n[s] ? (double)m.sh[s].total_bits / (double)L.n[s] : 0.0;
        wbits += bps * L.n_sub[s];
        wsym += L.n_sub[s];
    }
    F.stage = wsym > 0 && wbits / wsym <= 2.0;
    uint32_t item = 0, wg = 0;
    for (int s = 0; s < kStreams; ++s) {
        StreamDesc& d = F.ss.s[s];
        const double bps = L.n[s] ? (double)m.sh[s].total_bits / (double)L.n[s] : 0.0;
        d.ipw = dec_ipw(F.stage ? kDecStaged : kDecDense, bps);
        d.item_base = item;
        d.wg_base = wg;
        const uint32_t nwg = (d.item_count + d.ipw - 1) / d.ipw;
        item += nwg * d.ipw;
        wg += nwg;
    }
    F.ss.total_items = item;
    F.ss.total_wgs = wg;
    const size_t ts = tsize_of((int)m.fh.dtype);
    F.dd.codes = DBuf(c->pool, L.total_codes * 2 + 64);
    F.dd.side = DBuf(c->pool, L.total_codes * ts + 64);
    F.dd.compact = false;
    F.items = DBuf(c->pool, (size_t)std::max<uint32_t>(item, 1) * sizeof(DecItem));
    F.vb = DBuf(c->pool, (size_t)L.g[0].dx * L.g[0].dy * L.g[0].dz * ts);
    F.mid = DBuf(c->pool, (size_t)L.g[8].dx * L.g[8].dy * L.g[8].dz * ts);
}

template <typename T>
void plan_staged(stz_ctx* c, BatchPlan& P)
{
    for (const BatchPlan::Staged& q : P.stg) {
        K(c, "low_decode", [&] {
            launch_staged_decode<T>(c->stream, (const T*)q.vb, q.g1, q.la, q.codes, (const T*)q.side, (T*)q.mid,
                                    q.g8.dx, q.g8.dy, q.g8.dz, P.rec.as<T>());
        });
        K(c, "high_decode", [&] {
            launch_staged_decode<T>(c->stream, (const T*)q.mid, q.g8, q.ha, q.codes, (const T*)q.side, (T*)q.out,
                                    q.nx, q.ny, q.nz, P.rec.as<T>());
        });
    }
}

template <typename T>
void plan_enqueue(stz_ctx* c, BatchPlan& P, bool levels)
{
    hipStream_t st = c->stream;
    const int n = (int)P.ds.size();
    K(c, "err_zero", [&] { launch_fill_u32(st, P.err.as<uint32_t>(), (uint64_t)n, 0u); });
    K(c, "dec_prep", [&] { launch_dec_prep_multi(st, P.dec_all.as<DecJob>(), (uint32_t)n, P.chunks); });
    K(c, "huff_decode", [&] {
        launch_huff_decode_multi(st, P.dec_st.as<DecJob>(), P.n_st, P.wg_st, kDecStaged);
        launch_huff_decode_multi(st, P.dec_un.as<DecJob>(), P.n_un, P.wg_un, kDecDense);
    });
    K(c, "interp_decode_fused", [&] { launch_interp_decode_fused_multi<T>(st, P.fused.as<FusedJob>(), P.n_fused); });
    K(c, "lorenzo_decode", [&] { launch_lorenzo_multi<T>(st, P.lz.as<LorenzoJob>(), P.n_lz, P.lz_wgs, false); });
    for (size_t r = 0; r < P.rounds.size(); ++r)
        K(c, "interp_decode", [&] {
            launch_interp_decode_multi<T>(st, P.rounds[r].as<PassJob>(), P.round_n[r].first, P.round_n[r].second);
        });
    if (!levels) return;
    if (!P.low.h[0].empty()) {
        jobs_launch<T>(c, kLowDec, P.low, P.cfg_low);
        jobs_launch<T>(c, kHighDec, P.high, P.cfg_high);
    }
    plan_staged<T>(c, P);
}

// upload a host vector into a pool buffer (staged copy on the context stream)

// the decode launches of all fields: Huffman jobs and base-pass jobs
template <typename T>
void plan_decode_jobs(stz_ctx* c, BatchPlan& P)
{
    const int n = (int)P.ds.size();
    std::vector<DecJob> all, stj, unj;
    std::vector<FusedJob> fj;
    std::vector<std::vector<PassJob>> rounds;
    std::vector<LorenzoJob> lz;
    uint32_t chunk = 0;
    for (int i = 0; i < n; ++i) {
        BatchPlan::Field& F = P.f[i];
        stz_dstream* ds = P.ds[i];
        DecJob j{};
        j.ss = F.ss;
        j.payload = ds->payload.as<uint8_t>();
        j.tabs = ds->tabs.as<DecTab>();
        j.sorted = ds->sorted.as<uint16_t>();
        j.items = F.items.as<DecItem>();
        j.codes = F.dd.codes.as<uint16_t>();
        j.side = F.dd.side.p;
        j.err = P.err.as<uint32_t>() + i;
        j.chunk0 = chunk;
        chunk += F.ss.total_chunks;
        all.push_back(j);
        if (F.ss.total_wgs) {
            j.wg0 = F.stage ? P.wg_st : P.wg_un;
            (F.stage ? P.wg_st : P.wg_un) += F.ss.total_wgs;
            (F.stage ? stj : unj).push_back(j);
        }
        // base passes (Lorenzo-coded bases: tile jobs)
        const Meta& m = ds->meta;
        const uint16_t* c0 = F.dd.codes.as<uint16_t>() + F.dd.code_base[0];
        const T* s0 = F.dd.side.as<T>() + F.dd.code_base[0];
        if (m.fh.interp >= 2) {
            LorenzoJob q{};
            q.reg = m.fh.interp == 3;
            q.lo = m.fh.vmin;
            q.hi = m.fh.vmax;
            q.v = F.vb.p;
            q.codes = (uint16_t*)c0;
            q.side = (void*)s0;
            q.bx = F.L.g[0].dx; q.by = F.L.g[0].dy; q.bz = F.L.g[0].dz;
            q.eb = m.fh.eb_base;
            q.recip = 1.0 / m.fh.eb_base;
            q.wg0 = P.lz_wgs;
            P.lz_wgs += lorenzo_tiles(q.bx, q.by, q.bz);
            lz.push_back(q);
            continue;
        }
        const auto passes = make_passes(F.L.g[0].dx, F.L.g[0].dy, F.L.g[0].dz, m.fh.eb_base, (int)m.fh.interp,
                                        (int)m.fh.order, m.fh.alpha, m.fh.beta, 0);
        size_t r = 0;
        run_passes(passes, sizeof(T), [&](const InterpPassList& fl) {
            if (fl.n <= 0) return;
            FusedJob f{};
            f.l = fl;
            f.v = F.vb.p;
            f.codes = c0;
            f.side = s0;
            fj.push_back(f);
        }, [&](const InterpPass& p) {
            if (!p.n) return;
            if (rounds.size() <= r) rounds.resize(r + 1);
            PassJob q{};
            q.p = p;
            q.v = F.vb.p;
            q.codes = c0;
            q.side = s0;
            q.nwg = (uint32_t)std::min<uint64_t>((p.n + 255) / 256, 1024);
            rounds[r].push_back(q);
            ++r;
        }, n > 1);
    }
    P.chunks = chunk;
    P.n_st = (uint32_t)stj.size();
    P.n_un = (uint32_t)unj.size();
    P.dec_all = upload_vec(c, all);
    P.dec_st = upload_vec(c, stj);
    P.dec_un = upload_vec(c, unj);
    P.n_fused = (uint32_t)fj.size();
    P.fused = upload_vec(c, fj);
    P.n_lz = (uint32_t)lz.size();
    P.lz = upload_vec(c, lz);
    for (auto& rj : rounds) {
        uint32_t wg = 0;
        for (PassJob& q : rj) {
            q.wg0 = wg;
            wg += q.nwg;
        }
        P.round_n.push_back({(uint32_t)rj.size(), wg});
        P.rounds.push_back(upload_vec(c, rj));
    }
}

template <typename T>
void decompress_batch_impl(stz_ctx* c, int n, stz_dstream* const* ds, void* const* outs)
{
    bool same = c->plan && (int)c->plan->ds.size() == n;
    for (int i = 0; same && i < n; ++i)
        same = c->plan->ds[i] == ds[i] && c->plan->gens[i] == ds[i]->gen && c->plan->outs[i] == outs[i];
    if (!same) {
        c->plan.reset();
        std::unique_ptr<BatchPlan> P(new BatchPlan());
        P->ds.assign(ds, ds + n);
        for (int i = 0; i < n; ++i) P->gens.push_back(ds[i]->gen);
        P->outs.assign(outs, outs + n);
        P->f.resize(n);
        P->err = DBuf(c->pool, (size_t)n * 4 + 64);
        for (int i = 0; i < n; ++i)
            if (!is_device_ptr(outs[i])) throw StzError(STZ_E_ARG, "batch outputs must be device memory");
        ensure_tabs_batch(c, ds, n);
        for (int i = 0; i < n; ++i) plan_huff_setup(c, ds[i], P->f[i]);
        size_t rec_bytes = 0;
        for (int i = 0; i < n; ++i) {
            BatchPlan::Field& F = P->f[i];
            const Meta& m = ds[i]->meta;
            if (m.fh.pred & kPredStaged) {
                // dependency.cpp's staged predictors: stage launches of their own
                BatchPlan::Staged q{};
                q.vb = F.vb.p;
                q.mid = F.mid.p;
                q.g1 = F.L.g[1];
                q.g8 = F.L.g[8];
                q.la = level_args(m, F.dd, 1, P->err.as<uint32_t>() + i);
                q.ha = level_args(m, F.dd, 2, P->err.as<uint32_t>() + i);
                q.codes = F.dd.codes.as<uint16_t>();
                q.side = F.dd.side.p;
                q.out = outs[i];
                q.nx = (int)m.fh.nx;
                q.ny = (int)m.fh.ny;
                q.nz = (int)m.fh.nz;
                P->stg.push_back(q);
                rec_bytes = std::max(rec_bytes, staged_scratch(F.L, sizeof(T)));
                continue;
            }
            for (int v = 0; v < 2; ++v) {
                LevelJob lo{};
                lo.ref = F.vb.p;
                lo.mid = F.mid.p;
                lo.codes = F.dd.codes.as<uint16_t>();
                lo.side = F.dd.side.p;
                lo.g = F.L.g[1];
                lo.a = level_args(m, F.dd, 1, P->err.as<uint32_t>() + i);
                P->low.h[v].push_back(lo);
                LevelJob hi{};
                hi.ref = F.mid.p;
                hi.out = outs[i];
                hi.codes = F.dd.codes.as<uint16_t>();
                hi.side = F.dd.side.p;
                hi.nx = (int)m.fh.nx;
                hi.ny = (int)m.fh.ny;
                hi.nz = (int)m.fh.nz;
                hi.vec = ((m.fh.nx & 1) == 0) && (((uintptr_t)outs[i] % (2 * sizeof(T))) == 0);
                hi.g = F.L.g[8];
                hi.a = level_args(m, F.dd, 2, P->err.as<uint32_t>() + i);
                P->high.h[v].push_back(hi);
            }
        }
        if (rec_bytes) P->rec = DBuf(c->pool, rec_bytes);
        plan_decode_jobs<T>(c, *P);
        const bool multi = !P->low.h[0].empty();
        if (multi) {
            jobs_finish(c, P->low);
            jobs_finish(c, P->high);
        }
        // A new set of streams runs eagerly: most batches are decoded once,
        // and a capture plus instantiation costs more host time than the
        // launches.  The first batch of a size class also times the level
        // geometries, the eager run's results being their inputs (a decode
        // is idempotent, so a candidate that runs again writes the same).
        if (multi && (level_cfg_pending<T>(c, kLowDec, P->low) || level_cfg_pending<T>(c, kHighDec, P->high))) {
            plan_enqueue<T>(c, *P, false);
            P->cfg_low = level_cfg<T>(c, kLowDec, P->low, [] {});
            jobs_launch<T>(c, kLowDec, P->low, P->cfg_low);
            P->cfg_high = level_cfg<T>(c, kHighDec, P->high, [] {});
            jobs_launch<T>(c, kHighDec, P->high, P->cfg_high);
            plan_staged<T>(c, *P);
        } else {
            if (multi) {
                P->cfg_low = level_cfg<T>(c, kLowDec, P->low, [] {});
                P->cfg_high = level_cfg<T>(c, kHighDec, P->high, [] {});
            }
            plan_enqueue<T>(c, *P, true);
        }
        c->plan = std::move(P);
    } else {
        // the same streams and outputs again: the sequence is captured as a
        // graph once and replayed from then on
        BatchPlan& P = *c->plan;
        if (!P.exec) {
            const bool prof = c->prof;
            c->prof = false;
            HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
            try {
                plan_enqueue<T>(c, P, true);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(c->stream, &g);
                if (g) (void)hipGraphDestroy(g);
                c->prof = prof;
                throw;
            }
            HIPCHK(hipStreamEndCapture(c->stream, &P.graph));
            c->prof = prof;
            HIPCHK(hipGraphInstantiate(&P.exec, P.graph, nullptr, nullptr, 0));
        }
        HIPCHK(hipGraphLaunch(P.exec, c->stream));
    }
    uint32_t* herr = (uint32_t*)c->host_stage((size_t)n * 4);
    HIPCHK(hipMemcpyAsync(herr, c->plan->err.p, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int i = 0; i < n; ++i)
        if (herr[i]) throw StzError(STZ_E_STREAM, "corrupt Huffman bitstream");
}

// ------------------------------------------------------------ payload parsing
void parse_meta(const uint8_t* p, size_t bytes, Meta& m)
{
    if (bytes < sizeof(FileHeader)) throw StzError(STZ_E_STREAM, "short payload");
    std::memcpy(&m.fh, p, sizeof(FileHeader));
    const FileHeader& fh = m.fh;
    if (fh.magic != kMagic || fh.version != kVersion || fh.n_streams != kStreams || fh.total_bytes > bytes ||
        !pred_valid(fh.pred) || fh.interp > 3 || fh.order > 5 || !(fh.alpha >= 1.0) || !(fh.beta >= 1.0) ||
        (fh.dtype != STZ_F32 && fh.dtype != STZ_F64) || fh.radius != (uint32_t)kRadius || fh.sub_size != (uint32_t)kSub)
        throw StzError(STZ_E_STREAM, "bad payload header");
    if (fh.nx == 0 || fh.ny == 0 || fh.nz == 0 || fh.px != pad4(fh.nx) || fh.py != pad4(fh.ny) || fh.pz != pad4(fh.nz))
        throw StzError(STZ_E_STREAM, "bad dims");
    check_dims(fh.nx, fh.ny, fh.nz, (int)fh.chunk_high, (int)fh.chunk_low);
    HostLayout L;
    make_layout(L, fh.nx, fh.ny, fh.nz, fh.px, fh.py, fh.pz, (int)fh.chunk_high, (int)fh.chunk_low, (int)fh.interp);
    // streams are packed back to back after the file header (the writer's
    // layout); every bound below is checked without overflow
    uint64_t next = sizeof(FileHeader);
    for (int s = 0; s < kStreams; ++s) {
        if (fh.stream_off[s] != next || fh.stream_off[s] > bytes || fh.stream_bytes[s] < sizeof(StreamHeader) ||
            fh.stream_bytes[s] > bytes - fh.stream_off[s])
            throw StzError(STZ_E_STREAM, "stream out of bounds");
        next = fh.stream_off[s] + fh.stream_bytes[s];
        const uint32_t kind = s == 0 ? 0u : s < 8 ? 1u : 2u, block = s == 0 ? 0u : (uint32_t)((s - 1) % 7);
        check_stream(p + fh.stream_off[s], fh.stream_bytes[s], kind, block, L.g[s], L.n[s], L.chunk_sub[s],
                     (int)fh.dtype, m.sh[s], m.cb[s]);
        if (m.sh[s].stream_bytes != fh.stream_bytes[s]) throw StzError(STZ_E_STREAM, "bad stream size");
    }
    if (next != fh.total_bytes) throw StzError(STZ_E_STREAM, "bad payload size");
}

// ------------------------------------------------------------ host container
constexpr uint32_t kContMagic = 0x435A5453u;  // "STZC"
struct ContHeader {
    uint32_t magic, version, n_frames, zstd_level;
    uint64_t payload_bytes;
    uint64_t raw[kStreams + 1], comp[kStreams + 1];
};

void to_container(const uint8_t* pay, size_t bytes, const Meta& m, int level, std::vector<uint8_t>& out)
{
    ContHeader ch{};
    ch.magic = kContMagic;
    ch.version = 1;
    ch.n_frames = kStreams + 1;
    ch.zstd_level = (uint32_t)level;
    ch.payload_bytes = bytes;
    std::vector<std::vector<uint8_t>> fr(kStreams + 1);
    fr[0].assign(pay, pay + sizeof(FileHeader));
    ch.raw[0] = ch.comp[0] = sizeof(FileHeader);
    std::vector<std::thread> th;
    std::vector<int> bad(kStreams + 1, 0);
    for (int s = 0; s < kStreams; ++s) {
        th.emplace_back([&, s] {
            const uint8_t* src = pay + m.fh.stream_off[s];
            const size_t n = m.fh.stream_bytes[s];
            fr[s + 1].resize(ZSTD_compressBound(n));
            const size_t r = ZSTD_compress(fr[s + 1].data(), fr[s + 1].size(), src, n, level);
            if (ZSTD_isError(r)) { bad[s + 1] = 1; return; }
            fr[s + 1].resize(r);
        });
    }
    for (auto& t : th) t.join();
    for (int s = 0; s <= kStreams; ++s)
        if (bad[s]) throw StzError(STZ_E_ZSTD, "zstd compress failed");
    for (int s = 0; s < kStreams; ++s) {
        ch.raw[s + 1] = m.fh.stream_bytes[s];
        ch.comp[s + 1] = fr[s + 1].size();
    }
    size_t total = sizeof(ch);
    for (auto& f : fr) total += f.size();
    out.resize(total);
    std::memcpy(out.data(), &ch, sizeof(ch));
    size_t o = sizeof(ch);
    for (auto& f : fr) {
        std::memcpy(out.data() + o, f.data(), f.size());
        o += f.size();
    }
}

// container -> payload bytes (host)
void from_container(const uint8_t* in, size_t n, std::vector<uint8_t>& pay)
{
    if (n >= 4 && *(const uint32_t*)in == kMagic) {  // a bare payload
        pay.assign(in, in + n);
        return;
    }
    if (n < sizeof(ContHeader)) throw StzError(STZ_E_STREAM, "short container");
    ContHeader ch;
    std::memcpy(&ch, in, sizeof(ch));
    if (ch.magic != kContMagic || ch.version != 1 || ch.n_frames != kStreams + 1 || ch.raw[0] != sizeof(FileHeader))
        throw StzError(STZ_E_STREAM, "bad container");
    size_t o = sizeof(ch), po = 0;
    if (ch.payload_bytes > (1ull << 40)) throw StzError(STZ_E_STREAM, "bad container sizes");
    // validate every frame before starting any thread: a throw with joinable
    // std::threads alive would terminate the process.  Each frame's raw size
    // must fit what is left of the payload (no wrapped sums) and equal the
    // content size its zstd frame header declares.
    {
        size_t oo = sizeof(ch);
        uint64_t left = ch.payload_bytes;
        for (int f = 0; f <= kStreams; ++f) {
            if (ch.comp[f] > n - oo) throw StzError(STZ_E_STREAM, "truncated container");
            if (ch.raw[f] > left) throw StzError(STZ_E_STREAM, "bad container sizes");
            if (f == 0 && ch.comp[f] != ch.raw[f]) throw StzError(STZ_E_STREAM, "bad header frame");
            if (f > 0 && ZSTD_getFrameContentSize(in + oo, ch.comp[f]) != ch.raw[f])
                throw StzError(STZ_E_STREAM, "bad zstd frame size");
            left -= ch.raw[f];
            oo += ch.comp[f];
        }
        if (left != 0) throw StzError(STZ_E_STREAM, "bad container sizes");
    }
    pay.resize(ch.payload_bytes);
    std::memcpy(pay.data(), in + o, ch.raw[0]);
    o += ch.comp[0];
    po += ch.raw[0];
    std::vector<std::thread> th;
    std::vector<int> bad(kStreams + 1, 0);
    for (int f = 1; f <= kStreams; ++f) {
        const uint8_t* src = in + o;
        uint8_t* dst = pay.data() + po;
        const size_t cn = ch.comp[f], rn = ch.raw[f];
        th.emplace_back([=, &bad] {
            const size_t r = ZSTD_decompress(dst, rn, src, cn);
            if (ZSTD_isError(r) || r != rn) bad[f] = 1;
        });
        o += cn;
        po += rn;
    }
    for (auto& t : th) t.join();
    for (int f = 0; f <= kStreams; ++f)
        if (bad[f]) throw StzError(STZ_E_STREAM, "zstd frame failed");
}

stz_dstream* upload(stz_ctx* c, const uint8_t* pay, size_t bytes)
{
    auto ds = std::unique_ptr<stz_dstream>(new stz_dstream());
    parse_meta(pay, bytes, ds->meta);
    ds->pool = c->pool;
    ds->device = c->device;
    ds->bytes = ds->meta.fh.total_bytes;
    ds->payload = DBuf(c->pool, align_up(ds->bytes, 4));
    uint8_t* stage = (uint8_t*)c->host_stage(ds->bytes);
    std::memcpy(stage, pay, ds->bytes);
    HIPCHK(hipMemcpyAsync(ds->payload.p, stage, ds->bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return ds.release();
}

template <typename F>
int guard(stz_ctx* c, F&& f)
{
    try {
        if (c) {
            HIPCHK(hipSetDevice(c->device));
            c->stage_reset();
        }
        f();
        return STZ_OK;
    } catch (const StzError& e) {
        return e.code;
    } catch (const std::bad_alloc&) {
        return STZ_E_OOM;
    } catch (...) {
        return STZ_E_HIP;
    }
}

stz_params defaults()
{
    stz_params p;
    stz_default_params(&p);
    return p;
}

// device view of a field (copied when it is host memory)
struct FieldView {
    DBuf copy;
    const void* p;
    FieldView(stz_ctx* c, const void* f, size_t bytes)
    {
        if (is_device_ptr(f)) {
            p = f;
        } else {
            copy = DBuf(c->pool, bytes);
            HIPCHK(hipMemcpyAsync(copy.p, f, bytes, hipMemcpyHostToDevice, c->stream));
            p = copy.p;
        }
    }
};

template <typename T>
double abs_bound(stz_ctx* c, const T* d, size_t n, double eb, int mode)
{
    if (mode != STZ_EB_REL) return eb;
    DBuf part(c->pool, 2 * 2048 * sizeof(double) + 64);
    double* outp = part.as<double>() + 2 * 2048;
    K(c, "minmax", [&] { launch_minmax<T>(c->stream, d, n, part.as<double>(), outp); });
    double* h = (double*)c->host_stage(64);
    HIPCHK(hipMemcpyAsync(h, outp, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return eb * (h[1] - h[0]);
}

template <typename T>
std::vector<float> tune_features(stz_ctx* c, const T* d, size_t nx, size_t ny, size_t nz, double eb_abs,
                                 std::vector<double>* counts = nullptr)
{
    const size_t nb = ((nx + 63) / 64) * ((ny + 63) / 64) * ((nz + 63) / 64);
    DBuf sums(c->pool, nb * kFeatSums * sizeof(double)), part(c->pool, nb * kStatParts * kFeatSums * sizeof(double));
    K(c, "block_stats", [&] {
        launch_block_stats<T>(c->stream, d, (int)nx, (int)ny, (int)nz, eb_abs, sums.as<double>(), part.as<double>());
    });
    double* h = (double*)c->host_stage(nb * kFeatSums * sizeof(double));
    HIPCHK(hipMemcpyAsync(h, sums.p, nb * kFeatSums * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<float> f(nb * kMlpIn);
    const double e = eb_abs > 0 ? eb_abs : 1e-300;
    for (size_t b = 0; b < nb; ++b) block_features(h + b * kFeatSums, e, &f[b * kMlpIn]);
    if (counts) {
        counts->resize(nb);
        for (size_t b = 0; b < nb; ++b) (*counts)[b] = h[b * kFeatSums];
    }
    return f;
}

// MLP over the sub-blocks; the field's choice maximises the mean log
// probability over blocks, weighted by each block's sampled points (blocks
// too thin for the stencils carry no vote); confidence = the weighted mean
// probability of that choice.
template <typename T>
int tune_learned(stz_ctx* c, const T* d, size_t nx, size_t ny, size_t nz, double eb, int mode, float& conf)
{
    if (!c->have_tuner) throw StzError(STZ_E_ARG, "no tuner model set (stz_tune_set_model)");
    const size_t n = nx * ny * nz;
    const double eb_abs = abs_bound<T>(c, d, n, eb, mode);
    std::vector<double> cnt;
    const std::vector<float> f = tune_features<T>(c, d, nx, ny, nz, eb_abs, &cnt);
    const size_t nb = f.size() / kMlpIn;
    double wsum = 0;
    for (double v : cnt) wsum += v;
    if (wsum <= 0) {
        for (double& v : cnt) v = 1.0;
        wsum = (double)nb;
    }
    DBuf x(c->pool, f.size() * 4), y(c->pool, nb * kMlpOut * 4);
    float* hs = (float*)c->host_stage(f.size() * 4 + nb * kMlpOut * 4);
    std::memcpy(hs, f.data(), f.size() * 4);
    HIPCHK(hipMemcpyAsync(x.p, hs, f.size() * 4, hipMemcpyHostToDevice, c->stream));
    K(c, "mlp_forward", [&] { launch_mlp_forward(c->stream, c->tuner_w.as<float>(), x.as<float>(), (uint32_t)nb, y.as<float>()); });
    float* hy = hs + f.size();
    HIPCHK(hipMemcpyAsync(hy, y.p, nb * kMlpOut * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<double> lp(kTuneConfigs, 0.0), pr(kTuneConfigs, 0.0);
    for (size_t b = 0; b < nb; ++b) {
        const float* r = hy + b * kMlpOut;
        double mx = r[0];
        for (int k = 1; k < kTuneConfigs; ++k) mx = std::max(mx, (double)r[k]);
        double se = 0, e[kTuneConfigs];
        for (int k = 0; k < kTuneConfigs; ++k) {
            e[k] = std::exp(r[k] - mx);
            se += e[k];
        }
        const double lse = mx + std::log(se);
        const double w = cnt[b] / wsum;
        for (int k = 0; k < kTuneConfigs; ++k) {
            lp[k] += w * (r[k] - lse);
            pr[k] += w * e[k] / se;
        }
    }
    int best = 0;
    for (int k = 1; k < kTuneConfigs; ++k)
        if (lp[k] > lp[best]) best = k;
    conf = (float)pr[best];
    return best;
}

// exhaustive trial: every configuration on the whole field; the smallest
// payload wins (kept in *keep when given); sizes[k] = payload bytes of k
template <typename T>
int tune_trial(stz_ctx* c, int dtype, const T* d, size_t nx, size_t ny, size_t nz, double eb, int mode,
               const stz_params& base, stz_dstream* keep, uint64_t* sizes)
{
    int best = -1;
    uint64_t bb = ~0ull;
    for (int k = 0; k < kTuneConfigs; ++k) {
        stz_params p = base;
        p.tuner = 0;
        tune_config(k, p);
        stz_dstream tmp;
        compress_impl<T>(c, dtype, d, nx, ny, nz, eb, mode, p, &tmp);
        if (sizes) sizes[k] = tmp.bytes;
        if (tmp.bytes < bb) {
            bb = tmp.bytes;
            best = k;
            if (keep) std::swap(*keep, tmp);
        }
    }
    return best;
}

// compress with the parameters the tuner picks (params.tuner 1: learned,
// falling back to the trial below the context's confidence threshold; 2: trial)
template <typename T>
void compress_tuned(stz_ctx* c, int dtype, const void* field, size_t nx, size_t ny, size_t nz, double eb, int mode,
                    stz_params p, stz_dstream* ds)
{
    if (p.tuner == 0) {
        compress_impl<T>(c, dtype, (const T*)field, nx, ny, nz, eb, mode, p, ds);
        return;
    }
    if (mode == STZ_EB_REF || p.tuner < 0 || p.tuner > 2)
        throw StzError(STZ_E_ARG, "tuning needs an ABS or REL bound");
    check_dims(nx, ny, nz, p.chunk_high, p.chunk_low);
    FieldView fv(c, field, nx * ny * nz * sizeof(T));
    const T* d = (const T*)fv.p;
    if (p.tuner == 1) {
        // the model's choice is made inside the compress, after its value
        // range pass (the same statistics, MLP and vote as stz_tune)
        std::vector<char> low;
        const void* f = d;
        const size_t dims[3] = {nx, ny, nz};
        compress_fields<T>(c, dtype, 1, &f, dims, eb, mode, p, &ds, &low);
        if (!low[0]) return;
    }
    const float conf = ds->tune_conf;
    tune_trial<T>(c, dtype, d, nx, ny, nz, eb, mode, p, ds, nullptr);
    ds->tuned = 2;
    ds->tune_conf = conf;
}

}  // namespace

// =================================================================== C ABI
extern "C" {

void stz_default_params(stz_params* p)
{
    p->mult_base = 0.7;   // F5:2127
    p->mult_low = 2.7;    // F5:2128
    p->mult_high = 8.2;   // F5:2129
    p->chunk_high = 32;   // F5:34
    p->chunk_low = 16;    // F5:35
    p->interp = 1;
    p->order = 0;
    p->alpha = 1.0;
    p->beta = 1.0;
    p->zstd_level = 3;    // F5:3048 (SZ3 default)
    p->tuner = 0;
    p->pred = 0;          // F5's default build: no PRED_* option
}

const char* stz_strerror(int code)
{
    switch (code) {
    case STZ_OK: return "ok";
    case STZ_E_ARG: return "bad argument";
    case STZ_E_DIMS: return "unsupported dimensions";
    case STZ_E_STREAM: return "corrupt or unsupported stream";
    case STZ_E_OOM: return "out of memory";
    case STZ_E_HIP: return "HIP runtime error";
    case STZ_E_ZSTD: return "zstd error";
    case STZ_E_NODEV: return "no GPU device";
    case STZ_E_SCRATCH: return "input exceeds the allocation-free entry point's scratch (STZ_NOALLOC_SCRATCH_BYTES)";
    default: return "unknown error";
    }
}

void stz_free(void* p) { std::free(p); }

int stz_create(stz_ctx** out, int device) { return stz_create_ex(out, device, 0u); }

int stz_create_ex(stz_ctx** out, int device, unsigned flags)
{
    if (!out || (flags & ~STZ_CTX_OWN_QUEUE)) return STZ_E_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return STZ_E_NODEV;
    }
    if (device < 0 || device >= n) return STZ_E_ARG;
    auto c = std::unique_ptr<stz_ctx>(new stz_ctx());
    c->device = device;
    return guard(c.get(), [&] {
        // a blocking stream: it orders after work on the legacy default
        // stream, where torch puts the caller's tensors by default
        c->own_queue = (flags & STZ_CTX_OWN_QUEUE) != 0;
        if (c->own_queue)
            c->own = full_mask_stream(device);
        else
            HIPCHK(hipStreamCreateWithFlags(&c->own, hipStreamDefault));
        HIPCHK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&c->ev_range, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&c->ev_hist, hipEventDisableTiming));
        c->stream = c->own;
        c->pool = std::make_shared<Pool>();
        c->pool->device = device;
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, device));
        c->n_cu = prop.multiProcessorCount;
        // the device is current (guard): its scratch address, taken once per device
        if (scratch_setup(device, &c->scratch) != 0) throw StzError(STZ_E_HIP, "scratch symbol address");
        *out = c.release();
    });
}

void stz_destroy(stz_ctx* c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& r : c->recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    c->plan.reset();
    for (auto& A : c->arena) {
        if (A.stream) (void)hipStreamSynchronize(A.stream);
        if (A.pinned) (void)hipHostFree(A.pinned);
        for (void* p : A.retired) (void)hipHostFree(p);
        if (A.ev) (void)hipEventDestroy(A.ev);
    }
    if (c->own) (void)hipStreamDestroy(c->own);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_range) (void)hipEventDestroy(c->ev_range);
    if (c->ev_hist) (void)hipEventDestroy(c->ev_hist);
    if (c->side_stream) {
        (void)hipStreamSynchronize(c->side_stream);
        (void)hipStreamDestroy(c->side_stream);
        (void)hipEventDestroy(c->ev_items);
        (void)hipEventDestroy(c->ev_high);
    }
    delete c;
}

int stz_set_stream(stz_ctx* c, void* s)
{
    if (!c) return STZ_E_ARG;
    const hipStream_t ns = s ? (hipStream_t)s : c->own;
    if (ns == c->stream) return STZ_OK;
    try {
        HIPCHK(hipSetDevice(c->device));
        c->stage_drain();
    } catch (const StzError& e) {
        return e.code;
    }
    c->stream = ns;
    return STZ_OK;
}

void* stz_get_stream(stz_ctx* c) { return c ? (void*)c->stream : nullptr; }

int stz_set_grid_share(stz_ctx* c, float share)
{
    if (!c || !(share > 0.0f && share <= 1.0f)) return STZ_E_ARG;
    c->grid_share = share;
    return STZ_OK;
}

int stz_profile(stz_ctx* c, int enable)
{
    if (!c) return STZ_E_ARG;
    return guard(c, [&] {
        resolve_prof(c);
        if (enable) c->agg.clear();
        c->prof = enable != 0;
    });
}

int stz_profile_report(stz_ctx* c, char* buf, size_t cap)
{
    if (!c || !buf || !cap) return STZ_E_ARG;
    return guard(c, [&] {
        resolve_prof(c);
        std::string s = "{";
        bool first = true;
        for (auto& kv : c->agg) {
            char tmp[256];
            std::snprintf(tmp, sizeof(tmp), "%s\"%s\": [%ld, %.6f]", first ? "" : ", ", kv.first.c_str(),
                          kv.second.first, kv.second.second);
            s += tmp;
            first = false;
        }
        s += "}";
        if (s.size() + 1 > cap) throw StzError(STZ_E_ARG, "report buffer too small");
        std::memcpy(buf, s.c_str(), s.size() + 1);
    });
}

int stz_compress_device(stz_ctx* c, stz_dtype dtype, const void* field, size_t nx, size_t ny, size_t nz, double eb,
                        stz_eb_mode mode, const stz_params* params, stz_dstream** out)
{
    if (!c || !field || !out || (dtype != STZ_F32 && dtype != STZ_F64) || mode < STZ_EB_ABS || mode > STZ_EB_REF)
        return STZ_E_ARG;
    *out = nullptr;
    const stz_params p = params ? *params : defaults();
    return guard(c, [&] {
        auto ds = std::unique_ptr<stz_dstream>(new stz_dstream());
        if (dtype == STZ_F32) compress_tuned<float>(c, dtype, field, nx, ny, nz, eb, mode, p, ds.get());
        else compress_tuned<double>(c, dtype, field, nx, ny, nz, eb, mode, p, ds.get());
        *out = ds.release();
    });
}

// ---- batches ------------------------------------------------------------
int stz_compress_batch(stz_ctx* c, stz_dtype dtype, int n, const void* const* fields, const size_t* dims,
                       double eb, stz_eb_mode mode, const stz_params* params, stz_dstream** out)
{
    if (!c || n < 0 || (n > 0 && (!fields || !dims || !out)) || (dtype != STZ_F32 && dtype != STZ_F64) ||
        mode < STZ_EB_ABS || mode > STZ_EB_REF)
        return STZ_E_ARG;
    const stz_params p = params ? *params : defaults();
    if (p.tuner < 0 || p.tuner > 2 || (p.tuner != 0 && mode == STZ_EB_REF)) return STZ_E_ARG;
    for (int i = 0; i < n; ++i) out[i] = nullptr;
    return guard(c, [&] {
        if (dtype == STZ_F32) compress_batch_impl<float>(c, dtype, n, fields, dims, eb, mode, p, out);
        else compress_batch_impl<double>(c, dtype, n, fields, dims, eb, mode, p, out);
    });
}

int stz_decompress_batch(stz_ctx* c, int n, stz_dstream* const* ds, void* const* outs)
{
    if (!c || n < 0 || (n > 0 && (!ds || !outs))) return STZ_E_ARG;
    if (n == 0) return STZ_OK;
    return guard(c, [&] {
        const uint32_t dt = ds[0] ? ds[0]->meta.fh.dtype : 0;
        for (int i = 0; i < n; ++i) {
            if (!ds[i] || !outs[i]) throw StzError(STZ_E_ARG, "null stream or output");
            if (ds[i]->meta.fh.dtype != dt) throw StzError(STZ_E_ARG, "mixed dtypes in one batch");
        }
        if (dt == STZ_F32) decompress_batch_impl<float>(c, n, ds, outs);
        else decompress_batch_impl<double>(c, n, ds, outs);
    });
}

int stz_autotune(stz_ctx* c, int enable, char* report, size_t cap)
{
    if (!c) return STZ_E_ARG;
    return guard(c, [&] {
        if (enable >= 0) c->autotune = enable != 0;
        if (!report) return;
        std::string s = "{";
        bool first = true;
        for (auto& kv : c->tuned) {
            char tmp[160];
            std::snprintf(tmp, sizeof(tmp), "%s\"%s\": {\"waves\": %d, \"grid\": %u, \"ms\": %.4f, \"trials\": ",
                          first ? "" : ", ", kv.first.c_str(), kv.second.waves, kv.second.grid, kv.second.ms);
            s += tmp;
            s += kv.second.trials.empty() ? "[]" : kv.second.trials;
            s += "}";
            first = false;
        }
        s += "}";
        if (s.size() + 1 > cap) throw StzError(STZ_E_ARG, "report buffer too small");
        std::memcpy(report, s.c_str(), s.size() + 1);
    });
}

// ---- learned tuner ------------------------------------------------------
int stz_tune_config(int index, stz_params* p)
{
    if (!p || index < 0 || index >= kTuneConfigs) return STZ_E_ARG;
    tune_config(index, *p);
    return STZ_OK;
}

int stz_tune_features(stz_ctx* c, stz_dtype dtype, const void* field, size_t nx, size_t ny, size_t nz,
                      double eb_abs, float* feats, size_t cap_blocks, size_t* n_blocks)
{
    if (!c || !field || !feats || !n_blocks || !(eb_abs > 0) || (dtype != STZ_F32 && dtype != STZ_F64))
        return STZ_E_ARG;
    return guard(c, [&] {
        if (!nx || !ny || !nz) throw StzError(STZ_E_DIMS, "empty field");
        const size_t ts = dtype == STZ_F32 ? 4 : 8;
        FieldView fv(c, field, nx * ny * nz * ts);
        const std::vector<float> f = dtype == STZ_F32
                                         ? tune_features<float>(c, (const float*)fv.p, nx, ny, nz, eb_abs)
                                         : tune_features<double>(c, (const double*)fv.p, nx, ny, nz, eb_abs);
        *n_blocks = f.size() / kMlpIn;
        if (*n_blocks > cap_blocks) throw StzError(STZ_E_ARG, "feature buffer too small");
        std::memcpy(feats, f.data(), f.size() * 4);
    });
}

int stz_tune_set_model(stz_ctx* c, const float* params, size_t n, float min_confidence)
{
    if (!c || !params || n != (size_t)kMlpParams) return STZ_E_ARG;
    return guard(c, [&] {
        c->tuner_w = DBuf(c->pool, n * 4);
        float* h = (float*)c->host_stage(n * 4);
        std::memcpy(h, params, n * 4);
        HIPCHK(hipMemcpyAsync(c->tuner_w.p, h, n * 4, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        c->have_tuner = true;
        c->tune_min_conf = min_confidence;
    });
}

int stz_tune(stz_ctx* c, stz_dtype dtype, const void* field, size_t nx, size_t ny, size_t nz, double eb,
             stz_eb_mode mode, stz_params* inout, float* confidence, int* config)
{
    if (!c || !field || !inout || (dtype != STZ_F32 && dtype != STZ_F64) || mode == STZ_EB_REF) return STZ_E_ARG;
    return guard(c, [&] {
        const size_t ts = dtype == STZ_F32 ? 4 : 8;
        FieldView fv(c, field, nx * ny * nz * ts);
        float conf = 0.f;
        const int k = dtype == STZ_F32 ? tune_learned<float>(c, (const float*)fv.p, nx, ny, nz, eb, mode, conf)
                                       : tune_learned<double>(c, (const double*)fv.p, nx, ny, nz, eb, mode, conf);
        tune_config(k, *inout);
        if (confidence) *confidence = conf;
        if (config) *config = k;
    });
}

int stz_tune_trial(stz_ctx* c, stz_dtype dtype, const void* field, size_t nx, size_t ny, size_t nz, double eb,
                   stz_eb_mode mode, const stz_params* base, uint64_t* sizes, int* best)
{
    if (!c || !field || (dtype != STZ_F32 && dtype != STZ_F64) || mode == STZ_EB_REF) return STZ_E_ARG;
    const stz_params p = base ? *base : defaults();
    return guard(c, [&] {
        check_dims(nx, ny, nz, p.chunk_high, p.chunk_low);
        const size_t ts = dtype == STZ_F32 ? 4 : 8;
        FieldView fv(c, field, nx * ny * nz * ts);
        const int k = dtype == STZ_F32
                          ? tune_trial<float>(c, dtype, (const float*)fv.p, nx, ny, nz, eb, mode, p, nullptr, sizes)
                          : tune_trial<double>(c, dtype, (const double*)fv.p, nx, ny, nz, eb, mode, p, nullptr, sizes);
        if (best) *best = k;
    });
}

int stz_mlp_forward(stz_ctx* c, const float* params, const float* x, int n, float* logits)
{
    if (!c || !params || !x || !logits || n < 0) return STZ_E_ARG;
    return guard(c, [&] {
        if (!is_device_ptr(params) || !is_device_ptr(x) || !is_device_ptr(logits))
            throw StzError(STZ_E_ARG, "device pointers expected");
        K(c, "mlp_forward", [&] { launch_mlp_forward(c->stream, params, x, (uint32_t)n, logits); });
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int stz_mlp_grad(stz_ctx* c, const float* params, const float* x, const float* target, int n, float* grad,
                 float* loss)
{
    if (!c || !params || !x || !target || !grad || !loss || n <= 0) return STZ_E_ARG;
    return guard(c, [&] {
        if (!is_device_ptr(params) || !is_device_ptr(x) || !is_device_ptr(target) || !is_device_ptr(grad) ||
            !is_device_ptr(loss))
            throw StzError(STZ_E_ARG, "device pointers expected");
        K(c, "grad_zero", [&] {
            launch_fill_u32(c->stream, (uint32_t*)grad, kMlpParams, 0u);
            launch_fill_u32(c->stream, (uint32_t*)loss, 1, 0u);
        });
        K(c, "mlp_grad", [&] { launch_mlp_grad(c->stream, params, x, target, (uint32_t)n, grad, loss); });
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int stz_mlp_adam(stz_ctx* c, float* params, float* grad, float* m, float* v, int step, float lr, float weight_decay)
{
    if (!c || !params || !grad || !m || !v || step < 1) return STZ_E_ARG;
    return guard(c, [&] {
        K(c, "adam", [&] { launch_adam(c->stream, params, grad, m, v, kMlpParams, lr, 0.9f, 0.999f, 1e-8f, step, weight_decay); });
    });
}

int stz_dstream_upload(stz_ctx* c, const void* stream, size_t bytes, stz_dstream** out)
{
    if (!c || !stream || !out) return STZ_E_ARG;
    *out = nullptr;
    return guard(c, [&] {
        std::vector<uint8_t> pay;
        from_container((const uint8_t*)stream, bytes, pay);
        *out = upload(c, pay.data(), pay.size());
    });
}

int stz_dstream_download(stz_ctx* c, const stz_dstream* ds, int zstd_level, void** stream, size_t* bytes)
{
    if (!c || !ds || !stream || !bytes) return STZ_E_ARG;
    *stream = nullptr;
    *bytes = 0;
    return guard(c, [&] {
        std::vector<uint8_t> pay(ds->bytes);
        HIPCHK(hipMemcpy(pay.data(), ds->payload.p, ds->bytes, hipMemcpyDeviceToHost));
        std::vector<uint8_t> out;
        if (zstd_level < 0) out.swap(pay);  // bare payload, no zstd
        else to_container(pay.data(), pay.size(), ds->meta, zstd_level, out);
        void* m = std::malloc(out.size());
        if (!m) throw StzError(STZ_E_OOM, "malloc");
        std::memcpy(m, out.data(), out.size());
        *stream = m;
        *bytes = out.size();
    });
}

int stz_dstream_info(const stz_dstream* ds, stz_info* info)
{
    if (!ds || !info) return STZ_E_ARG;
    const FileHeader& fh = ds->meta.fh;
    info->nx = fh.nx; info->ny = fh.ny; info->nz = fh.nz;
    info->px = fh.px; info->py = fh.py; info->pz = fh.pz;
    info->dtype = (stz_dtype)fh.dtype;
    info->eb_base = fh.eb_base; info->eb_low = fh.eb_low; info->eb_high = fh.eb_high;
    info->payload_bytes = ds->bytes;
    info->interp = (int)fh.interp; info->order = (int)fh.order;
    info->mult_base = fh.mult_base; info->mult_low = fh.mult_low; info->mult_high = fh.mult_high;
    info->alpha = fh.alpha; info->beta = fh.beta;
    info->pred = (int)fh.pred;
    info->tuner = ds->tuned;
    info->tune_confidence = ds->tune_conf;
    return STZ_OK;
}

int stz_dstream_payload(const stz_dstream* ds, const void** p, size_t* bytes)
{
    if (!ds || !p || !bytes) return STZ_E_ARG;
    *p = ds->payload.p;
    *bytes = ds->bytes;
    return STZ_OK;
}

void stz_dstream_free(stz_dstream* ds)
{
    if (!ds) return;
    (void)hipSetDevice(ds->device);
    delete ds;
}

int stz_dstream_decompress(stz_ctx* c, stz_dstream* ds, void* out)
{
    if (!c || !ds || !out) return STZ_E_ARG;
    return guard(c, [&] {
        if (ds->meta.fh.dtype == STZ_F32) decompress_full<float>(c, ds, out);
        else decompress_full<double>(c, ds, out);
    });
}

int stz_dstream_decompress_roi(stz_ctx* c, stz_dstream* ds, int x0, int y0, int z0, int x1, int y1, int z1, void* out,
                               size_t od[3])
{
    if (!c || !ds || !out || !od) return STZ_E_ARG;
    stz_request r{};
    r.kind = STZ_REQ_BOX;
    r.a[0] = x0; r.a[1] = y0; r.a[2] = z0; r.a[3] = x1; r.a[4] = y1; r.a[5] = z1;
    return stz_dstream_decompress_batch(c, ds, 1, &r, &out, od);
}

int stz_dstream_decompress_level(stz_ctx* c, stz_dstream* ds, int level, void* out, size_t od[3])
{
    if (!c || !ds || !out || !od || level < 0 || level > 3) return STZ_E_ARG;
    stz_request r{};
    r.kind = STZ_REQ_LEVEL;
    r.a[0] = level;
    return stz_dstream_decompress_batch(c, ds, 1, &r, &out, od);
}

int stz_dstream_prepare(stz_ctx* c, stz_dstream* ds, int keep_base)
{
    if (!c || !ds) return STZ_E_ARG;
    return guard(c, [&] {
        ensure_tabs(c, ds);
        if (keep_base) {
            if (ds->meta.fh.dtype == STZ_F32) prepare_base<float>(c, ds);
            else prepare_base<double>(c, ds);
        }
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int stz_dstream_decompress_batch(stz_ctx* c, stz_dstream* ds, int n, const stz_request* reqs, void* const* outs,
                                 size_t* out_dims)
{
    if (!c || !ds || n < 0 || (n > 0 && (!reqs || !outs || !out_dims))) return STZ_E_ARG;
    if (n == 0) return STZ_OK;
    return guard(c, [&] {
        if (ds->meta.fh.dtype == STZ_F32) decompress_batch<float>(c, ds, n, reqs, outs, out_dims);
        else decompress_batch<double>(c, ds, n, reqs, outs, out_dims);
    });
}

int stz_compress(stz_ctx* c, stz_dtype dtype, const void* field, size_t nx, size_t ny, size_t nz, double eb,
                 stz_eb_mode mode, const stz_params* params, void** stream, size_t* bytes)
{
    if (!stream || !bytes) return STZ_E_ARG;
    stz_dstream* ds = nullptr;
    int rc = stz_compress_device(c, dtype, field, nx, ny, nz, eb, mode, params, &ds);
    if (rc != STZ_OK) return rc;
    const int level = params ? params->zstd_level : 3;
    rc = stz_dstream_download(c, ds, level, stream, bytes);
    stz_dstream_free(ds);
    return rc;
}

int stz_decompress(stz_ctx* c, const void* stream, size_t bytes, void* out)
{
    stz_dstream* ds = nullptr;
    int rc = stz_dstream_upload(c, stream, bytes, &ds);
    if (rc != STZ_OK) return rc;
    rc = stz_dstream_decompress(c, ds, out);
    stz_dstream_free(ds);
    return rc;
}

int stz_decompress_roi(stz_ctx* c, const void* stream, size_t bytes, int x0, int y0, int z0, int x1, int y1, int z1,
                       void* out, size_t od[3])
{
    stz_dstream* ds = nullptr;
    int rc = stz_dstream_upload(c, stream, bytes, &ds);
    if (rc != STZ_OK) return rc;
    rc = stz_dstream_decompress_roi(c, ds, x0, y0, z0, x1, y1, z1, out, od);
    stz_dstream_free(ds);
    return rc;
}

int stz_decompress_slice(stz_ctx* c, const void* stream, size_t bytes, int axis, int index, void* out, size_t od[3])
{
    if (axis < 0 || axis > 2) return STZ_E_ARG;
    stz_info info;
    int rc = stz_query_info(stream, bytes, &info);
    if (rc != STZ_OK) return rc;
    if (index < 0 || (size_t)index >= (axis == 0 ? info.nx : axis == 1 ? info.ny : info.nz)) return STZ_E_ARG;
    int a[3] = {0, 0, 0}, b[3] = {(int)info.nx - 1, (int)info.ny - 1, (int)info.nz - 1};
    a[axis] = b[axis] = index;
    return stz_decompress_roi(c, stream, bytes, a[0], a[1], a[2], b[0], b[1], b[2], out, od);
}

int stz_decompress_level(stz_ctx* c, const void* stream, size_t bytes, int level, void* out, size_t od[3])
{
    stz_dstream* ds = nullptr;
    int rc = stz_dstream_upload(c, stream, bytes, &ds);
    if (rc != STZ_OK) return rc;
    rc = stz_dstream_decompress_level(c, ds, level, out, od);
    stz_dstream_free(ds);
    return rc;
}

int stz_stream_check(const void* stream, size_t bytes)
{
    if (!stream) return STZ_E_ARG;
    return guard(nullptr, [&] {
        std::vector<uint8_t> pay;
        from_container((const uint8_t*)stream, bytes, pay);
        Meta m;
        parse_meta(pay.data(), pay.size(), m);
    });
}

int stz_query_info(const void* stream, size_t bytes, stz_info* info)
{
    if (!stream || !info) return STZ_E_ARG;
    return guard(nullptr, [&] {
        const uint8_t* p = (const uint8_t*)stream;
        FileHeader fh;
        if (bytes >= sizeof(FileHeader) && *(const uint32_t*)p == kMagic) {
            std::memcpy(&fh, p, sizeof(fh));
        } else {
            if (bytes < sizeof(ContHeader) + sizeof(FileHeader)) throw StzError(STZ_E_STREAM, "short stream");
            ContHeader ch;
            std::memcpy(&ch, p, sizeof(ch));
            if (ch.magic != kContMagic) throw StzError(STZ_E_STREAM, "bad container");
            std::memcpy(&fh, p + sizeof(ContHeader), sizeof(fh));
        }
        if (fh.magic != kMagic || fh.version != kVersion) throw StzError(STZ_E_STREAM, "bad header");
        info->nx = fh.nx; info->ny = fh.ny; info->nz = fh.nz;
        info->px = fh.px; info->py = fh.py; info->pz = fh.pz;
        info->dtype = (stz_dtype)fh.dtype;
        info->eb_base = fh.eb_base; info->eb_low = fh.eb_low; info->eb_high = fh.eb_high;
        info->payload_bytes = fh.total_bytes;
        info->interp = (int)fh.interp; info->order = (int)fh.order;
        info->mult_base = fh.mult_base; info->mult_low = fh.mult_low; info->mult_high = fh.mult_high;
        info->alpha = fh.alpha; info->beta = fh.beta;
        info->pred = (int)fh.pred;
        info->tuner = 0;
        info->tune_confidence = 0.f;
    });
}

}  // extern "C"

// =================================================================== stage entry points
namespace {

// a device view of a caller buffer: copies in (and back) when it is host memory
struct DView {
    stz_ctx* c;
    const void* user;
    size_t bytes;
    bool dev;
    DBuf tmp;
    void* ptr;
    DView(stz_ctx* cc, const void* u, size_t n, bool copy_in) : c(cc), user(u), bytes(n), dev(is_device_ptr(u))
    {
        if (dev) {
            ptr = (void*)u;
        } else {
            tmp = DBuf(c->pool, n);
            ptr = tmp.p;
            if (copy_in) HIPCHK(hipMemcpyAsync(ptr, u, n, hipMemcpyHostToDevice, c->stream));
        }
    }
    void back()
    {
        if (!dev) HIPCHK(hipMemcpyAsync((void*)user, ptr, bytes, hipMemcpyDeviceToHost, c->stream));
    }
};

template <typename T>
void split_impl(stz_ctx* c, const void* full, size_t nx, size_t ny, size_t nz, void* const sub[8])
{
    if (nx % 2 || ny % 2 || nz % 2 || !nx || !ny || !nz) throw StzError(STZ_E_DIMS, "split needs even dims");
    const size_t hs = (nx / 2) * (ny / 2) * (nz / 2) * sizeof(T);
    DView in(c, full, nx * ny * nz * sizeof(T), true);
    std::vector<std::unique_ptr<DView>> outs;
    T* ptrs[8];
    for (int k = 0; k < 8; ++k) {
        outs.emplace_back(new DView(c, sub[k], hs, false));
        ptrs[k] = (T*)outs.back()->ptr;
    }
    K(c, "split", [&] { launch_split<T>(c->stream, (const T*)in.ptr, (int)nx, (int)ny, (int)nz, ptrs); });
    for (auto& o : outs) o->back();
    HIPCHK(hipStreamSynchronize(c->stream));
}

template <typename T>
void merge_impl(stz_ctx* c, const void* const sub[8], size_t sx, size_t sy, size_t sz, void* full)
{
    const size_t hs = sx * sy * sz * sizeof(T);
    std::vector<std::unique_ptr<DView>> ins;
    const T* ptrs[8];
    for (int k = 0; k < 8; ++k) {
        ins.emplace_back(new DView(c, sub[k], hs, true));
        ptrs[k] = (const T*)ins.back()->ptr;
    }
    DView out(c, full, 8 * hs, false);
    K(c, "merge", [&] { launch_merge<T>(c->stream, ptrs, (int)sx, (int)sy, (int)sz, (T*)out.ptr); });
    out.back();
    HIPCHK(hipStreamSynchronize(c->stream));
}

template <typename T>
void residual_impl(stz_ctx* c, int block, const void* a, const void* ref, size_t dx, size_t dy, size_t dz, void* out,
                   int sign)
{
    if (block < 0 || block > 6 || !dx || !dy || !dz || dx * dy * dz >= (1ull << 31))
        throw StzError(STZ_E_ARG, "bad block or dims");
    const size_t n = dx * dy * dz * sizeof(T);
    DView va(c, a, n, true), vr(c, ref, n, true), vo(c, out, n, false);
    K(c, "residual", [&] {
        launch_residual_predict<T>(c->stream, block, (const T*)va.ptr, (const T*)vr.ptr, (int)dx, (int)dy, (int)dz,
                                   (T*)vo.ptr, sign);
    });
    vo.back();
    HIPCHK(hipStreamSynchronize(c->stream));
}

}  // namespace

namespace stz {
// ordered compaction / expansion of unpredictable values (stage entry points)
template <typename T>
void launch_compact_unpred(hipStream_t s, const uint16_t* codes, const T* data, uint64_t n, T* out, uint64_t* count);
template <typename T>
void launch_dequant_list(hipStream_t s, const uint16_t* codes, const T* unpred, uint64_t n, double eb, T* out);
}  // namespace stz

namespace {

// ---- per-stream stage entry points ------------------------------------------

// number of zero codes of n codes the caller holds (host or device)
size_t zero_codes(stz_ctx* c, const uint16_t* codes, size_t n)
{
    if (!is_device_ptr(codes)) {
        size_t nz = 0;
        for (size_t i = 0; i < n; ++i) nz += codes[i] == 0;
        return nz;
    }
    std::vector<uint16_t> h(n);
    HIPCHK(hipMemcpyAsync(h.data(), codes, n * 2, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    size_t nz = 0;
    for (size_t i = 0; i < n; ++i) nz += h[i] == 0;
    return nz;
}

void check_block(size_t dx, size_t dy, size_t dz)
{
    if (!dx || !dy || !dz || dx * dy * dz >= (1ull << 31) || dx * dy >= (1ull << 31))
        throw StzError(STZ_E_DIMS, "block dims must be nonzero and hold < 2^31 points");
}

template <typename T>
void residual_encode_impl(stz_ctx* c, int block, const void* sub, const void* ref, size_t dx, size_t dy, size_t dz,
                          double eb, uint16_t* codes, void* unpred, size_t* n_unpred, void* rec)
{
    if (block < 0 || block > 6) throw StzError(STZ_E_ARG, "block must be 0..6");
    check_block(dx, dy, dz);
    const size_t n = dx * dy * dz, nb = n * sizeof(T);
    DView vs(c, sub, nb, true), vr(c, ref, nb, true), vc(c, codes, n * 2, false), vu(c, unpred, nb, false);
    std::unique_ptr<DView> vrec(rec ? new DView(c, rec, nb, false) : nullptr);
    DBuf diff(c->pool, nb + 64), cnt(c->pool, 64);
    K(c, "residual_encode", [&] {
        launch_residual_encode<T>(c->stream, block, (const T*)vs.ptr, (const T*)vr.ptr, (int)dx, (int)dy, (int)dz,
                                  eb, (uint16_t*)vc.ptr, diff.as<T>(), vrec ? (T*)vrec->ptr : nullptr);
    });
    K(c, "compact_unpred", [&] {
        launch_compact_unpred<T>(c->stream, (const uint16_t*)vc.ptr, diff.as<T>(), n, (T*)vu.ptr, cnt.as<uint64_t>());
    });
    uint64_t* h = (uint64_t*)c->host_stage(64);
    HIPCHK(hipMemcpyAsync(h, cnt.p, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    vc.back();
    if (!vu.dev) HIPCHK(hipMemcpyAsync(unpred, vu.ptr, h[0] * sizeof(T), hipMemcpyDeviceToHost, c->stream));
    if (vrec) vrec->back();
    HIPCHK(hipStreamSynchronize(c->stream));
    *n_unpred = (size_t)h[0];
}

template <typename T>
void residual_decode_impl(stz_ctx* c, int block, const uint16_t* codes, const void* unpred, const void* ref,
                          size_t dx, size_t dy, size_t dz, double eb, void* out)
{
    if (block < 0 || block > 6) throw StzError(STZ_E_ARG, "block must be 0..6");
    check_block(dx, dy, dz);
    const size_t n = dx * dy * dz, nb = n * sizeof(T);
    const size_t nz = zero_codes(c, codes, n);
    if (nz && !unpred) throw StzError(STZ_E_ARG, "unpredictable list missing");
    DView vc(c, codes, n * 2, true), vr(c, ref, nb, true), vo(c, out, nb, false);
    DView vu(c, unpred ? unpred : (const void*)codes, (nz ? nz : 1) * sizeof(T), unpred != nullptr && nz > 0);
    DBuf de(c->pool, nb + 64);
    K(c, "dequant_list", [&] {
        launch_dequant_list<T>(c->stream, (const uint16_t*)vc.ptr, (const T*)vu.ptr, n, eb, de.as<T>());
    });
    K(c, "residual", [&] {
        launch_residual_predict<T>(c->stream, block, de.as<T>(), (const T*)vr.ptr, (int)dx, (int)dy, (int)dz,
                                   (T*)vo.ptr, -1);
    });
    vo.back();
    HIPCHK(hipStreamSynchronize(c->stream));
}

// stream descriptor of a standalone block stream (its bytes start at 0)
StreamDesc block_desc(const StreamHeader& sh, const Geom& g, int dtype)
{
    StreamDesc d{};
    d.code_off = 0;
    d.n_symbols = sh.n_symbols;
    d.off_chunk_bit = sh.off_chunk_bit;
    d.off_chunk_unpred = sh.off_chunk_unpred;
    d.off_chunk_sub = sh.off_chunk_sub;
    d.off_sub_bits = sh.off_sub_bits;
    d.off_sub_unpred = sh.off_sub_unpred;
    d.off_unpred = sh.off_unpred;
    d.off_bits = sh.off_bits;
    d.bits_words = (sh.stream_bytes - sh.off_bits) / 4;
    d.g = g;
    d.n_chunks = g.n_chunks;
    d.n_sub = sh.n_sub;
    d.tsize = (uint32_t)tsize_of(dtype);
    d.n_unpred = sh.n_unpred;
    return d;
}

// SZ_compress_finish (F5:1258-1307) on the device: histogram, codebook,
// chunk-marked encode of one block's codes into a standalone stream
template <typename T>
void huff_encode_chunks_impl(stz_ctx* c, int dtype, const uint16_t* codes, const void* unpred, size_t n_unpred,
                             size_t dx, size_t dy, size_t dz, int chunk, double eb, std::vector<uint8_t>& outv)
{
    check_block(dx, dy, dz);
    if (chunk <= 0 || chunk > 64) throw StzError(STZ_E_ARG, "chunk dims must be in 1..64");
    if (!(eb >= 0)) throw StzError(STZ_E_ARG, "bad error bound");
    if (n_unpred && !unpred) throw StzError(STZ_E_ARG, "unpredictable list missing");
    hipStream_t st = c->stream;
    const size_t n = dx * dy * dz, ts = sizeof(T);
    const Geom g = make_geom((int)dx, (int)dy, (int)dz, chunk);
    std::vector<uint32_t> csub;
    const uint32_t n_sub = chunk_subs(g, n, csub);
    DView vc(c, codes, n * 2, true);
    // histogram -> host: the zero count must equal the caller's list length
    // before the list is read on the device
    DBuf hist(c->pool, 65536 * 4), compact(c->pool, 65536 * 8 + 64), counts(c->pool, 256);
    K(c, "hist_zero", [&] { launch_fill_u32(st, hist.as<uint32_t>(), 65536, 0u); });
    K(c, "hist_codes", [&] { launch_hist_codes(st, (const uint16_t*)vc.ptr, n, hist.as<uint32_t>()); });
    K(c, "hist_compact", [&] { launch_hist_compact(st, hist.as<uint32_t>(), 1, compact.as<uint64_t>(), counts.as<uint32_t>()); });
    uint32_t* hc = (uint32_t*)c->host_stage(64);
    HIPCHK(hipMemcpyAsync(hc, counts.p, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    const uint32_t m = hc[0];
    std::vector<uint64_t> list(m);
    if (m) HIPCHK(hipMemcpy(list.data(), compact.p, (size_t)m * 8, hipMemcpyDeviceToHost));
    const uint64_t zeros = (m && (list[0] >> 32) == 0) ? (list[0] & 0xFFFFFFFFull) : 0;
    if (zeros != n_unpred) throw StzError(STZ_E_ARG, "n_unpred differs from the number of zero codes");
    Codebook cb;
    build_codebook(list.data(), m, kLmax, cb);
    StreamHeader sh;
    fill_stream_header_g(sh, 3u, 0u, g, n, n_sub, dtype, eb, m, n_unpred, cb.total_bits);
    // stream buffer: zeroed, host-written sections (header, codebook, chunk
    // sub-chunk map) uploaded, then the marked encode fills the rest
    DBuf pay(c->pool, align_up(sh.stream_bytes, 4) + 64);
    K(c, "payload_zero", [&] { launch_fill_u32(st, pay.as<uint32_t>(), align_up(sh.stream_bytes, 4) / 4, 0u); });
    const size_t pre = sh.off_sub_bits;
    const size_t enc_off = align_up(pre, 16);
    uint8_t* img = (uint8_t*)c->host_stage(enc_off + (size_t)m * 8 + 64);
    std::memset(img, 0, pre);
    std::memcpy(img, &sh, sizeof(sh));
    uint32_t* cbw = (uint32_t*)(img + sh.off_codebook);
    uint32_t* encl = (uint32_t*)(img + enc_off);
    for (uint32_t k = 0; k < m; ++k) {
        cbw[k] = (uint32_t)cb.sym[k] | ((uint32_t)cb.len[k] << 16);
        encl[2 * k] = cb.sym[k];
        encl[2 * k + 1] = cb.code[k] | ((uint32_t)cb.len[k] << 24);
    }
    std::memcpy(img + sh.off_chunk_sub, csub.data(), 4 * csub.size());
    DBuf up(c->pool, enc_off + (size_t)m * 8 + 64);
    HIPCHK(hipMemcpyAsync(up.p, img, enc_off + (size_t)m * 8, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(pay.p, up.p, pre, hipMemcpyDeviceToDevice, st));
    DBuf enc(c->pool, 65536 * 4);
    K(c, "enc_zero", [&] { launch_fill_u32(st, enc.as<uint32_t>(), 65536, 0u); });
    if (m) K(c, "enc_scatter", [&] { launch_enc_scatter(st, (const uint32_t*)(up.as<uint8_t>() + enc_off), m, enc.as<uint32_t>()); });
    // codes and unpredictables into stream order
    DBuf cs(c->pool, n * 2 + 64), side(c->pool, n * ts + 64);
    DView vu(c, n_unpred ? unpred : (const void*)codes, (n_unpred ? n_unpred : 1) * ts, n_unpred > 0);
    K(c, "block_to_stream", [&] {
        launch_block_to_stream<T>(st, (const uint16_t*)vc.ptr, (const T*)vu.ptr, n_unpred, g, cs.as<uint16_t>(),
                                  side.as<T>());
    });
    StreamSet ss{};
    ss.nstreams = 1;
    ss.total_chunks = g.n_chunks;
    ss.total_sub = n_sub;
    StreamDesc& d = ss.s[0];
    d = block_desc(sh, g, dtype);
    if (m) {
        const uint32_t a = cb.sym.front(), b = cb.sym.back();
        if (b - a + 1 <= kEncWindow) {
            d.enc_lo = a;
            d.enc_n = b - a + 1;
        } else {
            const uint32_t c0 = (uint32_t)kRadius - kEncWindow / 2;
            d.enc_lo = c0 < a ? a : (c0 > b + 1 - kEncWindow ? b + 1 - kEncWindow : c0);
            d.enc_n = kEncWindow;
        }
    }
    DBuf pre_b(c->pool, (size_t)n_sub * 4 + 16), pre_u(c->pool, (size_t)n_sub * 4 + 16);
    DBuf tot_b(c->pool, (size_t)g.n_chunks * 8 + 16), tot_u(c->pool, (size_t)g.n_chunks * 8 + 16);
    K(c, "huff_encode", [&] {
        launch_huff_len(st, ss, cs.as<uint16_t>(), enc.as<uint32_t>(), pay.as<uint8_t>(), pre_b.as<uint32_t>(),
                        pre_u.as<uint32_t>(), tot_b.as<uint64_t>(), tot_u.as<uint64_t>());
        launch_huff_scan(st, ss, tot_b.as<uint64_t>(), tot_u.as<uint64_t>(), pay.as<uint8_t>());
        launch_huff_pack(st, ss, cs.as<uint16_t>(), side.p, enc.as<uint32_t>(), pre_b.as<uint32_t>(),
                         pre_u.as<uint32_t>(), pay.as<uint8_t>());
    });
    outv.resize(sh.stream_bytes);
    HIPCHK(hipMemcpyAsync(outv.data(), pay.p, sh.stream_bytes, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
}

// SZ_decompress_nopred_marked_chunks (F5:1841-1911) / SZ_decompress_separated
// (F5:1466-1576) on the device for a standalone block stream
void huff_decode_chunks_impl(stz_ctx* c, const void* stream, size_t bytes, const int64_t* ids, int64_t n_ids,
                             void* out, uint16_t* codes_out, size_t* dsym, size_t* uzero)
{
    hipStream_t st = c->stream;
    std::vector<uint8_t> hbuf;
    const uint8_t* p = (const uint8_t*)stream;
    if (is_device_ptr(stream)) {
        hbuf.resize(bytes);
        HIPCHK(hipMemcpy(hbuf.data(), stream, bytes, hipMemcpyDeviceToHost));
        p = hbuf.data();
    }
    if (bytes < sizeof(StreamHeader)) throw StzError(STZ_E_STREAM, "short stream");
    StreamHeader h0;
    std::memcpy(&h0, p, sizeof(h0));
    if (h0.magic != kStreamMagic || h0.kind != 3u || (h0.dtype != STZ_F32 && h0.dtype != STZ_F64) ||
        h0.chunk == 0 || h0.chunk > 64)
        throw StzError(STZ_E_STREAM, "not a block stream");
    check_block(h0.dx, h0.dy, h0.dz);
    const int dtype = (int)h0.dtype;
    const size_t ts = tsize_of(dtype);
    const Geom g = make_geom((int)h0.dx, (int)h0.dy, (int)h0.dz, (int)h0.chunk);
    const uint64_t n = (uint64_t)g.dx * g.dy * g.dz;
    std::vector<uint32_t> csub;
    chunk_subs(g, n, csub);
    StreamHeader sh;
    Codebook cb;
    check_stream(p, bytes, 3u, 0u, g, n, csub, dtype, sh, cb);
    // chunk list (row-major ids, F5:1578-1582); none given: every chunk
    std::vector<int64_t> all;
    if (!ids || n_ids < 0) {
        all.resize(g.n_chunks);
        for (uint32_t k = 0; k < g.n_chunks; ++k) all[k] = k;
        ids = all.data();
        n_ids = (int64_t)g.n_chunks;
    }
    const uint64_t* cun = (const uint64_t*)(p + sh.off_chunk_unpred);
    std::vector<uint32_t> quads, pairs;
    uint64_t pos = 0, syms = 0, zeros = 0;
    uint32_t item = 0;
    for (int64_t k = 0; k < n_ids; ++k) {
        const int64_t ch = ids[k];
        if (ch < 0 || ch >= (int64_t)g.n_chunks) throw StzError(STZ_E_ARG, "chunk id out of range");
        quads.push_back(0u);
        quads.push_back((uint32_t)ch);
        quads.push_back(item);
        quads.push_back((uint32_t)pos);
        pairs.push_back((uint32_t)ch);
        pairs.push_back((uint32_t)pos);
        item += csub[ch + 1] - csub[ch];
        const uint64_t vol = chunk_volume(g, (uint32_t)ch, n);
        syms += vol;
        zeros += cun[ch + 1] - cun[ch];
        pos += align_up(vol, 32);
        if (pos >= (1ull << 32)) throw StzError(STZ_E_DIMS, "chunk list decodes more than 2^32 symbols");
    }
    *dsym = (size_t)syms;
    *uzero = (size_t)zeros;
    if (!n_ids) return;
    // stream bytes and decode tables on the device
    DBuf pay(c->pool, align_up(bytes, 4) + 64);
    uint8_t* stg = (uint8_t*)c->host_stage(sh.stream_bytes);
    std::memcpy(stg, p, sh.stream_bytes);
    HIPCHK(hipMemcpyAsync(pay.p, stg, sh.stream_bytes, hipMemcpyHostToDevice, st));
    DecHead head;
    std::vector<uint16_t> sorted;
    build_dechead(cb.sym.data(), cb.len.data(), (uint32_t)cb.sym.size(), head, sorted);
    head.sorted_off = 0;
    const size_t sb = align_up(sorted.size() * 2, 16);
    DBuf srt(c->pool, sb + sizeof(DecHead) + 16), tabs(c->pool, sizeof(DecTab));
    uint8_t* hs = (uint8_t*)c->host_stage(sb + sizeof(DecHead) + 16);
    std::memcpy(hs, sorted.data(), sorted.size() * 2);
    std::memcpy(hs + sb, &head, sizeof(head));
    HIPCHK(hipMemcpyAsync(srt.p, hs, sb + sizeof(DecHead), hipMemcpyHostToDevice, st));
    K(c, "dectabs", [&] {
        launch_build_dectabs(st, (const DecHead*)(srt.as<uint8_t>() + sb), srt.as<uint16_t>(), tabs.as<DecTab>(), 1);
    });
    StreamSet ss{};
    ss.nstreams = 1;
    ss.total_chunks = g.n_chunks;
    ss.total_sub = sh.n_sub;
    StreamDesc& d = ss.s[0];
    d = block_desc(sh, g, dtype);
    d.item_count = item;
    // decode variant as for payload streams: up to 2 bits per symbol the
    // LDS-staged kernel (fewer items per workgroup while a workgroup's bits
    // would overflow the stage), above it the unstaged one
    const double bps = (double)sh.total_bits / (double)n;
    const bool stage = bps <= 2.0;
    d.ipw = dec_ipw(stage ? kDecStaged : kDecDense, bps);
    const uint32_t nwg = (item + d.ipw - 1) / d.ipw;
    ss.total_items = nwg * d.ipw;
    ss.total_wgs = nwg;
    DBuf dq(c->pool, quads.size() * 4 + 16), dp(c->pool, pairs.size() * 4 + 16);
    uint32_t* lst = (uint32_t*)c->host_stage((quads.size() + pairs.size()) * 4 + 16);
    std::memcpy(lst, quads.data(), quads.size() * 4);
    std::memcpy(lst + quads.size(), pairs.data(), pairs.size() * 4);
    HIPCHK(hipMemcpyAsync(dq.p, lst, quads.size() * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dp.p, lst + quads.size(), pairs.size() * 4, hipMemcpyHostToDevice, st));
    DBuf items(c->pool, (size_t)ss.total_items * sizeof(DecItem) + 16), cmap(c->pool, (size_t)g.n_chunks * 4 + 16);
    DBuf codes(c->pool, pos * 2 + 64), side(c->pool, pos * ts + 64);
    ErrFlag err(c);
    K(c, "dec_prep", [&] {
        launch_dec_prep(st, ss, pay.as<uint8_t>(), dq.as<uint32_t>(), (uint32_t)n_ids, items.as<DecItem>(),
                        cmap.as<uint32_t>());
    });
    K(c, "huff_decode", [&] {
        launch_huff_decode(st, ss, pay.as<uint8_t>(), tabs.as<DecTab>(), srt.as<uint16_t>(), items.as<DecItem>(),
                           codes.as<uint16_t>(), side.p, err.ptr(), stage ? kDecStaged : kDecDense);
    });
    // points outside the listed chunks keep the caller's values
    DView vo(c, out, n * ts, true);
    std::unique_ptr<DView> vco(codes_out ? new DView(c, codes_out, n * 2, true) : nullptr);
    K(c, "chunks_to_block", [&] {
        if (dtype == STZ_F32)
            launch_chunks_to_block<float>(st, dp.as<uint32_t>(), (uint32_t)n_ids, g, codes.as<uint16_t>(),
                                          side.as<float>(), sh.eb, (float*)vo.ptr,
                                          vco ? (uint16_t*)vco->ptr : nullptr);
        else
            launch_chunks_to_block<double>(st, dp.as<uint32_t>(), (uint32_t)n_ids, g, codes.as<uint16_t>(),
                                           side.as<double>(), sh.eb, (double*)vo.ptr,
                                           vco ? (uint16_t*)vco->ptr : nullptr);
    });
    err.fetch(c);
    HIPCHK(hipStreamSynchronize(st));
    err.check();
    vo.back();
    if (vco) vco->back();
    HIPCHK(hipStreamSynchronize(st));
}

}  // namespace

extern "C" {

int stz_split(stz_ctx* c, stz_dtype dtype, const void* full, size_t nx, size_t ny, size_t nz, void* const sub[8])
{
    if (!c || !full || !sub) return STZ_E_ARG;
    return guard(c, [&] {
        if (dtype == STZ_F32) split_impl<float>(c, full, nx, ny, nz, sub);
        else split_impl<double>(c, full, nx, ny, nz, sub);
    });
}

int stz_merge(stz_ctx* c, stz_dtype dtype, const void* const sub[8], size_t sx, size_t sy, size_t sz, void* full)
{
    if (!c || !full || !sub || !sx || !sy || !sz) return STZ_E_ARG;
    return guard(c, [&] {
        if (dtype == STZ_F32) merge_impl<float>(c, sub, sx, sy, sz, full);
        else merge_impl<double>(c, sub, sx, sy, sz, full);
    });
}

int stz_residual_predict(stz_ctx* c, stz_dtype dtype, int block, const void* sub, const void* ref, size_t dx,
                         size_t dy, size_t dz, void* diff)
{
    if (!c || !sub || !ref || !diff) return STZ_E_ARG;
    return guard(c, [&] {
        if (dtype == STZ_F32) residual_impl<float>(c, block, sub, ref, dx, dy, dz, diff, 1);
        else residual_impl<double>(c, block, sub, ref, dx, dy, dz, diff, 1);
    });
}

int stz_residual_reconstruct(stz_ctx* c, stz_dtype dtype, int block, const void* de, const void* ref, size_t dx,
                             size_t dy, size_t dz, void* out)
{
    if (!c || !de || !ref || !out) return STZ_E_ARG;
    return guard(c, [&] {
        if (dtype == STZ_F32) residual_impl<float>(c, block, de, ref, dx, dy, dz, out, -1);
        else residual_impl<double>(c, block, de, ref, dx, dy, dz, out, -1);
    });
}

int stz_quantize(stz_ctx* c, stz_dtype dtype, const void* data, size_t n, double eb, uint16_t* codes, void* unpred,
                 size_t* n_unpred)
{
    if (!c || !data || !codes || !unpred || !n_unpred || !(eb >= 0) || n == 0) return STZ_E_ARG;
    return guard(c, [&] {
        const size_t ts = tsize_of(dtype);
        DView vd(c, data, n * ts, true), vc(c, codes, n * 2, false), vu(c, unpred, n * ts, false);
        DBuf cnt(c->pool, 64);
        if (dtype == STZ_F32) {
            K(c, "quantize", [&] { launch_quantize<float>(c->stream, (const float*)vd.ptr, n, eb, (uint16_t*)vc.ptr); });
            K(c, "compact_unpred", [&] { launch_compact_unpred<float>(c->stream, (const uint16_t*)vc.ptr, (const float*)vd.ptr, n, (float*)vu.ptr, cnt.as<uint64_t>()); });
        } else {
            K(c, "quantize", [&] { launch_quantize<double>(c->stream, (const double*)vd.ptr, n, eb, (uint16_t*)vc.ptr); });
            K(c, "compact_unpred", [&] { launch_compact_unpred<double>(c->stream, (const uint16_t*)vc.ptr, (const double*)vd.ptr, n, (double*)vu.ptr, cnt.as<uint64_t>()); });
        }
        vc.back();
        vu.back();
        uint64_t* h = (uint64_t*)c->host_stage(64);
        HIPCHK(hipMemcpyAsync(h, cnt.p, 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        *n_unpred = (size_t)h[0];
    });
}

int stz_dequantize(stz_ctx* c, stz_dtype dtype, const uint16_t* codes, size_t n, double eb, const void* unpred, void* out)
{
    if (!c || !codes || !out || n == 0) return STZ_E_ARG;
    return guard(c, [&] {
        const size_t ts = tsize_of(dtype);
        size_t nz = 0;
        if (!is_device_ptr(codes)) {
            for (size_t i = 0; i < n; ++i) nz += codes[i] == 0;
        } else {
            // device codes: the list must be device memory too (its length is unknown here)
            if (unpred && !is_device_ptr(unpred)) throw StzError(STZ_E_ARG, "device codes need a device list");
            nz = 1;
        }
        if (nz && !unpred) throw StzError(STZ_E_ARG, "unpredictable list missing");
        DView vc(c, codes, n * 2, true), vo(c, out, n * ts, false);
        DView vu(c, unpred ? unpred : (const void*)codes, (nz ? nz : 1) * ts, unpred != nullptr);
        if (dtype == STZ_F32)
            K(c, "dequant_list", [&] { launch_dequant_list<float>(c->stream, (const uint16_t*)vc.ptr, (const float*)vu.ptr, n, eb, (float*)vo.ptr); });
        else
            K(c, "dequant_list", [&] { launch_dequant_list<double>(c->stream, (const uint16_t*)vc.ptr, (const double*)vu.ptr, n, eb, (double*)vo.ptr); });
        vo.back();
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int stz_residual_encode(stz_ctx* c, stz_dtype dtype, int block, const void* sub, const void* ref, size_t dx,
                        size_t dy, size_t dz, double eb, uint16_t* codes, void* unpred, size_t* n_unpred, void* rec)
{
    if (!c || !sub || !ref || !codes || !unpred || !n_unpred || !(eb >= 0)) return STZ_E_ARG;
    return guard(c, [&] {
        if (dtype == STZ_F32) residual_encode_impl<float>(c, block, sub, ref, dx, dy, dz, eb, codes, unpred, n_unpred, rec);
        else residual_encode_impl<double>(c, block, sub, ref, dx, dy, dz, eb, codes, unpred, n_unpred, rec);
    });
}

int stz_residual_decode(stz_ctx* c, stz_dtype dtype, int block, const uint16_t* codes, const void* unpred,
                        const void* ref, size_t dx, size_t dy, size_t dz, double eb, void* out)
{
    if (!c || !codes || !ref || !out) return STZ_E_ARG;
    return guard(c, [&] {
        if (dtype == STZ_F32) residual_decode_impl<float>(c, block, codes, unpred, ref, dx, dy, dz, eb, out);
        else residual_decode_impl<double>(c, block, codes, unpred, ref, dx, dy, dz, eb, out);
    });
}

int stz_huff_encode_chunks(stz_ctx* c, stz_dtype dtype, const uint16_t* codes, const void* unpred, size_t n_unpred,
                           size_t dx, size_t dy, size_t dz, int chunk, double eb, void** stream, size_t* bytes)
{
    if (!c || !codes || !stream || !bytes) return STZ_E_ARG;
    return guard(c, [&] {
        std::vector<uint8_t> v;
        if (dtype == STZ_F32) huff_encode_chunks_impl<float>(c, dtype, codes, unpred, n_unpred, dx, dy, dz, chunk, eb, v);
        else huff_encode_chunks_impl<double>(c, dtype, codes, unpred, n_unpred, dx, dy, dz, chunk, eb, v);
        void* o = std::malloc(v.size());
        if (!o) throw std::bad_alloc();
        std::memcpy(o, v.data(), v.size());
        *stream = o;
        *bytes = v.size();
    });
}

int stz_huff_decode_chunks(stz_ctx* c, const void* stream, size_t bytes, const int64_t* chunk_ids, int64_t n_ids,
                           void* out, uint16_t* codes_out, size_t* decoded_symbols, size_t* unpred_zeros)
{
    if (!c || !stream || !out || !decoded_symbols || !unpred_zeros) return STZ_E_ARG;
    return guard(c, [&] {
        huff_decode_chunks_impl(c, stream, bytes, chunk_ids, n_ids, out, codes_out, decoded_symbols, unpred_zeros);
    });
}

static_assert(kScratchBytes == STZ_NOALLOC_SCRATCH_BYTES, "scratch size in the header");

int stz_residual_encode_noalloc(stz_ctx* c, stz_dtype dtype, int block, const void* sub, const void* ref, size_t dx,
                                size_t dy, size_t dz, double eb, uint16_t* codes, void* unpred, uint64_t* n_unpred,
                                void* rec)
{
    if (!c || !sub || !ref || !codes || !unpred || !n_unpred || !(eb >= 0) || block < 0 || block > 6 ||
        (dtype != STZ_F32 && dtype != STZ_F64))
        return STZ_E_ARG;
    if (!dx || !dy || !dz) return STZ_E_DIMS;
    // points bounded by the scratch, checked without overflow
    const size_t cap = kScratchBytes / tsize_of(dtype);
    if (dx > cap || dy > cap || dz > cap || dx * dy > cap || dx * dy * dz > cap) return STZ_E_SCRATCH;
    if (hipSetDevice(c->device) != hipSuccess) {
        (void)hipGetLastError();
        return STZ_E_HIP;
    }
    if (!c->scratch) return STZ_E_HIP;
    // every buffer is device memory of this context's device: nothing is staged
    const void* ptrs[6] = {sub, ref, codes, unpred, n_unpred, rec};
    for (int k = 0; k < 6; ++k) {
        if (k == 5 && !rec) break;
        hipPointerAttribute_t a;
        if (hipPointerGetAttributes(&a, ptrs[k]) != hipSuccess) {
            (void)hipGetLastError();
            return STZ_E_ARG;
        }
        if (a.type != hipMemoryTypeDevice || a.device != c->device) return STZ_E_ARG;
    }
    (void)hipGetLastError();
    const int ix = (int)dx, iy = (int)dy, iz = (int)dz;
    if (dtype == STZ_F32)
        launch_residual_encode_noalloc<float>(c->stream, c->scratch, block, (const float*)sub, (const float*)ref, ix,
                                              iy, iz, eb, codes, (float*)unpred, n_unpred, (float*)rec);
    else
        launch_residual_encode_noalloc<double>(c->stream, c->scratch, block, (const double*)sub, (const double*)ref,
                                               ix, iy, iz, eb, codes, (double*)unpred, n_unpred, (double*)rec);
    if (hipGetLastError() != hipSuccess) return STZ_E_HIP;
    return STZ_OK;
}

int stz_huff_lengths(const uint64_t* freq, int nsym, int lmax, uint8_t* len)
{
    if (!freq || !len || nsym <= 0 || nsym > 65536 || lmax < 1 || lmax > kLmax) return STZ_E_ARG;
    int used = 0;
    for (int s = 0; s < nsym; ++s) used += freq[s] != 0;
    if (used > (1 << lmax) && used > 1) return STZ_E_ARG;
    huff_lengths_dense(freq, nsym, lmax, len);
    return STZ_OK;
}

int stz_plan_query(int fx, int fy, int fz, int chunk_high, int chunk_low, int ax, int ay, int az, int bx, int by,
                   int bz, int* query, int* high_plan, int64_t* high_chunks, int* low_plan, int64_t* low_chunks,
                   int64_t cap)
{
    if (fx <= 0 || fy <= 0 || fz <= 0 || chunk_high <= 0 || chunk_low <= 0 || !query || !high_plan || !low_plan ||
        !high_chunks || !low_chunks || cap < 0)
        return STZ_E_ARG;
    Plan pl;
    make_plan(pl, fx, fy, fz, (int)pad4((size_t)fx), (int)pad4((size_t)fy), (int)pad4((size_t)fz), chunk_high,
              chunk_low, ax, ay, az, bx, by, bz);
    query[0] = pl.q.x0; query[1] = pl.q.x1; query[2] = pl.q.y0; query[3] = pl.q.y1;
    query[4] = pl.q.z0; query[5] = pl.q.z1; query[6] = pl.is_slice ? 1 : 0;
    int rc = STZ_OK;
    for (int b = 0; b < 7; ++b) {
        const BlockPlan* ps[2] = {&pl.high[b], &pl.low[b]};
        int* dst[2] = {high_plan + 8 * b, low_plan + 8 * b};
        int64_t* cd[2] = {high_chunks + cap * b, low_chunks + cap * b};
        for (int t = 0; t < 2; ++t) {
            const BlockPlan& p = *ps[t];
            dst[t][0] = p.valid ? 1 : 0;
            dst[t][1] = p.r.x0; dst[t][2] = p.r.x1; dst[t][3] = p.r.y0;
            dst[t][4] = p.r.y1; dst[t][5] = p.r.z0; dst[t][6] = p.r.z1;
            dst[t][7] = (int)p.chunks.size();
            if ((int64_t)p.chunks.size() > cap) { rc = STZ_E_ARG; continue; }
            for (size_t i = 0; i < p.chunks.size(); ++i) cd[t][i] = p.chunks[i];
        }
    }
    return rc;
}

}  // extern "C"
