// C-ABI layer of liblofar_hip: per-device context, workspace, twiddle tables,
// the reference drop-ins (wstack, phase_correction, test) and the device entry
// points.  Host pointers only cross PCIe in the drop-ins; the lofar_*_dev entry
// points enqueue on the caller's stream and never synchronise except to read
// back the sample-validation flag and the bucket total (16 bytes) per gridding.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "common.h"
#include "fft_kernels.h"
#include "grid_kernels.h"

// A sort of fixed samples kept across calls (lofar_plan_*).  gmem holds the gridding
// bucket offsets and entries (each entry's sample index in its real part's slot),
// hmem the degridding bucket offsets, the home entries and the weights; for windows
// wider than the tile kernels take, hmem holds copies of uu, vv, ww and the weights.
struct lofar_plan_s {
    int device = -1, precision = 0, nw = 0, support = 0, nx = 0, ny = 0;
    long n = 0, F = 1, P = 1;
    double dx = 0.0, dw = 0.0;
    bool direct = false;
    lofar::TileMap gmap{}, hmap{};
    long long gtotal = 0, htotal = 0;
    long long* goff = nullptr;
    void* gent = nullptr;
    long long* hoff = nullptr;
    lofar::DEntry* hent = nullptr;
    unsigned long long* need = nullptr;  // tiles the degridding reads (powers of two)
    float* wt = nullptr;
    double *uu = nullptr, *vv = nullptr, *ww = nullptr;
    char* gmem = nullptr;
    char* hmem = nullptr;
};

namespace lofar {
namespace {

thread_local int t_last_error = LOFAR_OK;

struct Buffer {
    void* ptr = nullptr;
    size_t size = 0;
};

struct Prof {
    int stage;
    hipEvent_t a, b;
};

// A small persistent pool of host threads that splits a byte range between them: the
// drop-ins' copies between the caller's pageable arrays and pinned staging buffers, and
// the all-zero test of the caller's grid.
class HostPool {
  public:
    ~HostPool()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
            ++gen_;
        }
        cv_.notify_all();
        for (auto& t : th_) t.join();
    }
    // fn(lo, hi) over [0, n) in one slice per thread (the caller's thread takes one)
    template <typename F>
    void run(size_t n, F&& fn)
    {
        start();
        const size_t parts = th_.size() + 1;
        if (n < ((size_t)4 << 20) || parts == 1) {
            fn((size_t)0, n);
            return;
        }
        const size_t step = (n / parts + 4095) & ~(size_t)4095;
        std::function<void(size_t, size_t)> job = fn;
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = &job;
            n_ = n;
            step_ = step;
            pending_ = (int)th_.size();
            ++gen_;
        }
        cv_.notify_all();
        fn(0, std::min(step, n));
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [&] { return pending_ == 0; });
        job_ = nullptr;
    }

  private:
    void start()
    {
        std::call_once(once_, [&] {
            unsigned hw = std::thread::hardware_concurrency();
            const unsigned nt = std::min(16u, hw > 1 ? hw : 1u) - 1;
            for (unsigned i = 0; i < nt; i++) th_.emplace_back([this, i] { work(i + 1); });
        });
    }
    void work(size_t part)
    {
        unsigned long long seen = 0;
        for (;;) {
            std::function<void(size_t, size_t)>* job;
            size_t n, step;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
                if (stop_) return;
                job = job_;
                n = n_;
                step = step_;
            }
            const size_t lo = std::min(n, part * step), hi = std::min(n, lo + step);
            if (lo < hi) (*job)(lo, hi);
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (--pending_ == 0) done_cv_.notify_one();
            }
        }
    }
    std::once_flag once_;
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    std::function<void(size_t, size_t)>* job_ = nullptr;
    size_t n_ = 0, step_ = 0;
    int pending_ = 0;
    unsigned long long gen_ = 0;
    bool stop_ = false;
};

HostPool& host_pool()
{
    static HostPool pool;
    return pool;
}

struct DeviceCtx {
    bool ready = false;
    int device = -1;
    std::mutex mu;                  // serialises calls that use this device's workspace
    hipStream_t stream = nullptr;   // stream of the host-pointer drop-ins
    hipEvent_t last_done = nullptr; // end of the previous call's work
    Buffer meta;                    // counts, offsets, scan partials
    Buffer ents;                    // bucket entries and windows
    Buffer host_io;                 // device copies of drop-in host arrays
    static constexpr int kStage = 3;           // pinned staging buffers of the drop-ins
    static constexpr size_t kStageBytes = 64ull << 20;
    char* stage_buf[kStage] = {};
    hipEvent_t stage_done[kStage] = {};
    Buffer occ;                     // per-tile occupancy of the fused imaging path
    long long* status_host = nullptr;  // pinned: [total, err]
    char* scratch = nullptr;        // this device's copy of `scratch` (allocation-free calls)
    hipError_t scratch_err = hipSuccess;
    std::map<long long, void*> twiddles;
    std::vector<Prof> prof;
    std::vector<hipEvent_t> pool;
};

DeviceCtx g_ctx[64];
std::mutex g_init_mu;

// Workspace of the allocation-free entry point (lofar_image_dev_static): one static
// device array, of which every GPU the library runs on holds its own copy until the
// process ends.  Host code never uses the array's name as an address (on the host it is
// a stand-in symbol); each device's copy is located once, at the context's set-up, with
// hipGetSymbolAddress, and only that pointer reaches kernels.
constexpr size_t kScratchBytes = LOFAR_STATIC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
// stage timing switch (lofar_profile_enable); read by every call's launches
std::atomic<bool> g_profiling{false};

int fail(int code)
{
    t_last_error = code;
    return code;
}

#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) {                                                            \
            fprintf(stderr, "liblofar_hip: %s failed: %s\n", #x, hipGetErrorString(e_));   \
            return fail(e_ == hipErrorOutOfMemory ? LOFAR_ENOMEM : LOFAR_EHIP);             \
        }                                                                                  \
    } while (0)

int get_ctx(DeviceCtx** out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return LOFAR_ENODEV;
    if (dev < 0 || dev >= 64) return LOFAR_ENODEV;
    DeviceCtx& c = g_ctx[dev];
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (!c.ready) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return LOFAR_ENODEV;
        if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
            fprintf(stderr, "liblofar_hip: device %d is %s; the library is compiled for gfx950 only\n", dev,
                    prop.gcnArchName);
            return LOFAR_ENODEV;
        }
        if (hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess)
            return LOFAR_EHIP;
        if (hipEventCreateWithFlags(&c.last_done, hipEventDisableTiming) != hipSuccess)
            return LOFAR_EHIP;
        if (hipHostMalloc((void**)&c.status_host, 2 * sizeof(long long), hipHostMallocDefault) !=
            hipSuccess)
            return LOFAR_EHIP;
        // this device's copy of the allocation-free workspace (device `dev` is current)
        void* sp = nullptr;
        c.scratch_err = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
        if (c.scratch_err == hipSuccess) {
            c.scratch = (char*)sp;
        } else {
            (void)hipGetLastError();
            c.scratch = nullptr;
        }
        c.device = dev;
        c.ready = true;
    }
    *out = &c;
    return LOFAR_OK;
}

// grow-only device buffer; waits for the previous call before freeing
int ensure(DeviceCtx& c, Buffer& b, size_t bytes)
{
    if (b.size >= bytes) return LOFAR_OK;
    HIPCHK(hipEventSynchronize(c.last_done));
    if (b.ptr) HIPCHK(hipFree(b.ptr));
    b.ptr = nullptr;
    b.size = 0;
    size_t want = bytes + bytes / 8 + 4096;
    if (hipMalloc(&b.ptr, want) != hipSuccess) {
        b.ptr = nullptr;
        (void)hipGetLastError();
        return fail(LOFAR_ENOMEM);
    }
    b.size = want;
    return LOFAR_OK;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Pinned staging of the drop-ins (allocated on first use, kept with the context).
int stage_init(DeviceCtx& c)
{
    if (c.stage_buf[0]) return LOFAR_OK;
    for (int i = 0; i < DeviceCtx::kStage; i++) {
        HIPCHK(hipHostMalloc((void**)&c.stage_buf[i], DeviceCtx::kStageBytes, hipHostMallocDefault));
        HIPCHK(hipEventCreateWithFlags(&c.stage_done[i], hipEventDisableTiming));
    }
    return LOFAR_OK;
}

// Caller's pageable host array -> device, in chunks: the host threads copy chunk k into
// a pinned buffer while the DMA engine moves chunk k-1, so the two overlap.
int h2d_staged(DeviceCtx& c, void* dst, const void* src, size_t bytes, hipStream_t st)
{
    int r = stage_init(c);
    if (r) return r;
    const size_t cb = DeviceCtx::kStageBytes;
    for (size_t off = 0, k = 0; off < bytes; off += cb, k++) {
        const int b = (int)(k % DeviceCtx::kStage);
        const size_t len = std::min(cb, bytes - off);
        HIPCHK(hipEventSynchronize(c.stage_done[b]));  // the buffer's previous DMA is done
        char* pin = c.stage_buf[b];
        const char* from = (const char*)src + off;
        host_pool().run(len, [&](size_t lo, size_t hi) { memcpy(pin + lo, from + lo, hi - lo); });
        HIPCHK(hipMemcpyAsync((char*)dst + off, pin, len, hipMemcpyHostToDevice, st));
        HIPCHK(hipEventRecord(c.stage_done[b], st));
    }
    return LOFAR_OK;
}

// Device -> caller's pageable host array, in chunks: DMA of chunk k into a pinned buffer
// overlaps the host threads' copy of chunk k-1 out of another one.  Returns when the
// whole array is on the host.
int d2h_staged(DeviceCtx& c, void* dst, const void* src, size_t bytes, hipStream_t st)
{
    int r = stage_init(c);
    if (r) return r;
    const size_t cb = DeviceCtx::kStageBytes;
    const size_t nchunks = (bytes + cb - 1) / cb;
    auto drain = [&](size_t k) -> int {
        const int b = (int)(k % DeviceCtx::kStage);
        const size_t off = k * cb, len = std::min(cb, bytes - off);
        HIPCHK(hipEventSynchronize(c.stage_done[b]));
        const char* pin = c.stage_buf[b];
        char* to = (char*)dst + off;
        host_pool().run(len, [&](size_t lo, size_t hi) { memcpy(to + lo, pin + lo, hi - lo); });
        return LOFAR_OK;
    };
    for (size_t k = 0; k < nchunks; k++) {
        if (k >= (size_t)DeviceCtx::kStage - 1) {  // keep kStage-1 transfers in flight
            r = drain(k - (DeviceCtx::kStage - 1));
            if (r) return r;
        }
        const int b = (int)(k % DeviceCtx::kStage);
        const size_t off = k * cb, len = std::min(cb, bytes - off);
        HIPCHK(hipMemcpyAsync(c.stage_buf[b], (const char*)src + off, len, hipMemcpyDeviceToHost,
                              st));
        HIPCHK(hipEventRecord(c.stage_done[b], st));
    }
    for (size_t k = nchunks > (size_t)DeviceCtx::kStage - 1 ? nchunks - (DeviceCtx::kStage - 1) : 0;
         k < nchunks; k++) {
        r = drain(k);
        if (r) return r;
    }
    return LOFAR_OK;
}

// true when every double of the caller's array is +0.0 (scanned by the host threads)
bool host_all_zero(const double* a, size_t n)
{
    std::atomic<bool> nz{false};
    host_pool().run(n * sizeof(double), [&](size_t lo, size_t hi) {
        const unsigned long long* p = (const unsigned long long*)((const char*)a + lo);
        const size_t m = (hi - lo) / 8;
        for (size_t i = 0; i < m; i += 4096) {
            if (nz.load(std::memory_order_relaxed)) return;
            unsigned long long acc = 0;
            const size_t e = std::min(m, i + 4096);
            for (size_t j = i; j < e; j++) acc |= p[j];
            if (acc) {
                nz.store(true, std::memory_order_relaxed);
                return;
            }
        }
    });
    return !nz.load();
}

// Where a call's workspace comes from: the context's grow-only buffers (base ==
// nullptr), or consecutive 256-byte-aligned slices of the static `scratch` array
// (the allocation-free entry point), which never grows.
struct Arena {
    char* base = nullptr;
    size_t cap = 0, used = 0;
};

int take(DeviceCtx& c, Arena* ar, Buffer& b, size_t bytes, char** out)
{
    if (!ar || !ar->base) {
        int r = ensure(c, b, bytes);
        if (r) return r;
        *out = (char*)b.ptr;
        return LOFAR_OK;
    }
    const size_t at = align256(ar->used);
    if (at + bytes > ar->cap) {
        fprintf(stderr,
                "liblofar_hip: lofar_image_dev_static needs more than its %zu-byte static "
                "workspace (%zu bytes so far); use lofar_image_dev for this size\n",
                ar->cap, at + bytes);
        return fail(LOFAR_ESCRATCH);
    }
    *out = ar->base + at;
    ar->used = at + bytes;
    return LOFAR_OK;
}

// exp(+2*pi*i*m/N), m < N, from long double arithmetic, rounded once
int twiddles(DeviceCtx& c, int n, int precision, void** out)
{
    const long long key = (long long)n * 4 + precision;
    auto it = c.twiddles.find(key);
    if (it != c.twiddles.end()) {
        *out = it->second;
        return LOFAR_OK;
    }
    std::vector<double> tw(2 * (size_t)n);
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int m = 0; m < n; m++) {
        const long double a = two_pi * (long double)m / (long double)n;
        tw[2 * m] = (double)cosl(a);
        tw[2 * m + 1] = (double)sinl(a);
    }
    void* d = nullptr;
    size_t bytes = (size_t)n * (precision == LOFAR_FP64 ? 16 : 8);
    HIPCHK(hipMalloc(&d, bytes));
    if (precision == LOFAR_FP64) {
        HIPCHK(hipMemcpy(d, tw.data(), bytes, hipMemcpyHostToDevice));
    } else {
        std::vector<float> f(2 * (size_t)n);
        for (size_t i = 0; i < f.size(); i++) f[i] = (float)tw[i];
        HIPCHK(hipMemcpy(d, f.data(), bytes, hipMemcpyHostToDevice));
    }
    c.twiddles[key] = d;
    *out = d;
    return LOFAR_OK;
}

// Twiddles of the allocation-free entry point: a slice of the static workspace
// filled on the device (no allocation and no host copy).
int twiddles_for(DeviceCtx& c, Arena* ar, int n, int precision, hipStream_t st, void** out)
{
    if (!ar || !ar->base) return twiddles(c, n, precision, out);
    Buffer unused;
    char* p = nullptr;
    int r = take(c, ar, unused, (size_t)n * (precision == LOFAR_FP64 ? 16 : 8), &p);
    if (r) return r;
    HIPCHK(fill_twiddles(precision, n, p, st));
    *out = p;
    return LOFAR_OK;
}

hipEvent_t pool_event(DeviceCtx& c)
{
    if (!c.pool.empty()) {
        hipEvent_t e = c.pool.back();
        c.pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Runs one stage's launches, bracketed by timing events when profiling is on.
template <typename F>
hipError_t stage(DeviceCtx& c, int id, hipStream_t st, F&& fn)
{
    if (!g_profiling.load(std::memory_order_relaxed)) return fn();
    Prof p{id, pool_event(c), pool_event(c)};
    if (p.a) (void)hipEventRecord(p.a, st);
    hipError_t e = fn();
    if (p.b) (void)hipEventRecord(p.b, st);
    if (p.a && p.b) c.prof.push_back(p);
    return e;
}

enum Stage {
    kStCount = 0, kStScan, kStScatter, kStGrid, kStCols, kStRows, kStRowsPhase, kStPhase,
    kStRowsPhaseAdj, kStDegrid
};
const char* kStageNames[LOFAR_NSTAGES] = {
    "bucket_count", "bucket_scan", "bucket_scatter", "grid_tiles", "fft_cols",
    "fft_rows", "rows_phase", "phase_sum", "rows_phase_adj", "degrid_tiles"};

bool pow2_ok(int n) { return n >= 2 && n <= 8192 && (n & (n - 1)) == 0; }

// Mixed-radix transforms of every plane (sides that are not powers of two): columns
// (with `col_scale`), then rows (with `row_scale`); out may alias in.
int fft_generic_planes(DeviceCtx& c, Arena* ar, int precision, int sign, const void* in,
                       void* out, int nx, int ny, int nw, double row_scale, hipStream_t st,
                       bool rows_first = false)
{
    void *twx, *twy;
    int r = twiddles_for(c, ar, nx, precision, st, &twx);
    if (r) return r;
    r = twiddles_for(c, ar, ny, precision, st, &twy);
    if (r) return r;
    auto cols = [&](const void* src) {
        return precision == LOFAR_FP64
                   ? fft_generic<double>(sign, src, out, twy, ny, (long long)nw * nx, nx,
                                         (long long)nx * ny, 1, nx, rows_first ? row_scale : 1.0,
                                         st)
                   : fft_generic<float>(sign, src, out, twy, ny, (long long)nw * nx, nx,
                                        (long long)nx * ny, 1, nx, rows_first ? row_scale : 1.0,
                                        st);
    };
    auto rows = [&](const void* src) {
        const long long lines = (long long)nw * ny;
        return precision == LOFAR_FP64
                   ? fft_generic<double>(sign, src, out, twx, nx, lines, lines, 0, nx, 1,
                                         rows_first ? 1.0 : row_scale, st)
                   : fft_generic<float>(sign, src, out, twx, nx, lines, lines, 0, nx, 1,
                                        rows_first ? 1.0 : row_scale, st);
    };
    if (rows_first) {
        HIPCHK(stage(c, kStRows, st, [&] { return rows(in); }));
        HIPCHK(stage(c, kStCols, st, [&] { return cols(out); }));
    } else {
        HIPCHK(stage(c, kStCols, st, [&] { return cols(in); }));
        HIPCHK(stage(c, kStRows, st, [&] { return rows(out); }));
    }
    return LOFAR_OK;
}

int check_grid_args(int precision, int nw, long n, long F, long P, double dx, double dw,
                    int support, int nx, int ny)
{
    if (precision != LOFAR_FP64 && precision != LOFAR_FP32C) return LOFAR_EARG;
    if (nw < 1 || n < 0 || F < 1 || P < 1 || support < 1 || support > 65535) return LOFAR_EARG;
    if (nx < 1 || ny < 1 || nx > 65536 || ny > 65536) return LOFAR_EARG;
    if (!(dx > 0.0) || !(dw > 0.0) || !std::isfinite(dx) || !std::isfinite(dw)) return LOFAR_EARG;
    return LOFAR_OK;
}

Geometry make_geometry(int nw, int support, int nx, int ny, double dx, double dw)
{
    Geometry g;
    g.nx = nx;
    g.ny = ny;
    g.nw = nw;
    g.K = (support - 1) / 2;
    g.ntu = (nx + kTile - 1) / kTile;
    g.ntv = (ny + kTile - 1) / kTile;
    g.dx = dx;
    g.dw = dw;
    return g;
}

// Bucket-sort tuning: super-tile side (tiles), samples per level-1 workgroup,
// records per level-2 block, and which scatters are staged through LDS (kStageL1 |
// kStageL2).  LOFAR_BUCKET_CFG="sst,chunk,block" and LOFAR_BUCKET_STAGE=<bits>
// override, for measurement.  16x16-tile super-tiles keep a 4096-record round of the
// staged level 1 at 8 records per super-tile on a 4096^2 x 32 cube.
struct BucketCfg {
    int sst;
    long long chunk, block;
    int stage;
};

BucketCfg bucket_cfg()
{
    static const BucketCfg cfg = [] {
        int stage = kStageL1 | kStageL2;
        if (const char* e = getenv("LOFAR_BUCKET_STAGE")) stage = atoi(e) & (kStageL1 | kStageL2);
        int a = (stage & kStageL1) ? 16 : 8;
        long long b = 16384, l2 = 8192;
        if (const char* e = getenv("LOFAR_BUCKET_CFG")) sscanf(e, "%d,%lld,%lld", &a, &b, &l2);
        if (a < 1 || a > 16 || (a & (a - 1))) a = 16;
        if (b < 256) b = 256;
        if (l2 < 256) l2 = 256;
        return BucketCfg{a, b, l2, stage};
    }();
    return cfg;
}

// Launch order of the per-tile kernels (TileMap::launch_bucket): blocks of 2^olog x
// 2^olog tiles, 0 for bucket order.  LOFAR_TILE_ORDER=<olog> overrides, for measurement.
int tile_order_log(const TileMap& m)
{
    static const int want = [] {
        const char* e = getenv("LOFAR_TILE_ORDER");
        const int v = e ? atoi(e) : 3;
        return v < 0 || v > 6 ? 3 : v;
    }();
    const int side = 1 << want;
    if (want == 0 || ((m.nstu << m.lsst) % side) || ((m.nstv << m.lsst) % side)) return 0;
    return want;
}

// Level-2 block table and fine counts, placed after the records once the entry
// total is known: returns the bytes needed; with base set, assigns the pointers.
size_t level2_layout(Buckets& b, long long total, char* base)
{
    b.l2_max_blocks = (total + b.l2_block - 1) / b.l2_block + b.map.nk1;
    const size_t nf = (size_t)b.map.sst * b.map.sst;
    const size_t blk_bytes = align256((size_t)b.l2_max_blocks * sizeof(L2Block));
    const size_t cnt_bytes = align256((size_t)b.l2_max_blocks * nf * sizeof(unsigned));
    if (base) {
        b.l2_blk = (L2Block*)base;
        b.l2_cnt = (unsigned*)(base + blk_bytes);
    }
    return blk_bytes + cnt_bytes;
}

// Two-level counting sort of the samples into (w, tile) buckets: sizes the
// workspace, counts per super-tile, scans, and reads back the entry total and
// the validation flag (the only host synchronisation of a call).  The level-1
// histograms of nw*ceil(ntv/sst)*ceil(ntu/sst) super-tiles (records and
// entries) live in LDS up to kMaxSuperTiles (64 KiB), in global memory beyond.
constexpr long long kMaxSuperTiles = 8192;

// LDS histogram capacity in super-tiles; LOFAR_L1_LDS_MAX lowers it (tests of the
// global-histogram path on small grids)
long long l1_lds_max()
{
    static long long v = -1;
    if (v < 0) {
        const char* e = getenv("LOFAR_L1_LDS_MAX");
        v = e ? atoll(e) : kMaxSuperTiles;
        if (v < 1 || v > kMaxSuperTiles) v = kMaxSuperTiles;
    }
    return v;
}

int build_buckets(DeviceCtx& c, int mode, long n, const double* uu, const double* vv,
                  const double* ww, const Geometry& g, hipStream_t st, Buckets& b,
                  long long& total, Arena* ar = nullptr)
{
    const BucketCfg cfg = bucket_cfg();
    int sst = cfg.sst;
    const long long chunk = cfg.chunk;
    b.l2_block = cfg.block;
    const int span = g.ntu > g.ntv ? g.ntu : g.ntv;
    while (sst > 1 && sst / 2 >= span) sst /= 2;  // no super-tile wider than the grid
    auto super_tiles = [&](int s) {
        return (long long)g.nw * ((g.ntv + s - 1) / s) * ((g.ntu + s - 1) / s);
    };
    while (sst < 16 && super_tiles(sst) > kMaxSuperTiles) sst *= 2;  // large grids
    TileMap& m = b.map;
    m.sst = sst;
    m.lsst = 0;
    while ((1 << m.lsst) < sst) m.lsst++;
    m.nstu = (g.ntu + sst - 1) / sst;
    m.nstv = (g.ntv + sst - 1) / sst;
    m.rfb = rec_fraction_bits(sst);
    m.nk1 = super_tiles(sst);
    m.olog = tile_order_log(m);
    // beyond the LDS histograms' capacity, level 1 counts and places records with global
    // atomics on per-super-tile counters (slower, no size limit)
    b.global_l1 = m.nk1 > l1_lds_max();
    b.stage = cfg.stage & ((b.global_l1 || m.nk1 > kStageMaxBins) ? kStageL2 : ~0);
    b.nwg = n > 0 ? (n + chunk - 1) / chunk : 1;
    if (b.nwg > 4096) b.nwg = 4096;
    b.chunk = n > 0 ? (n + b.nwg - 1) / b.nwg : 1;
    const long long nblk = (m.nk1 + 4095) / 4096;
    const size_t off_status = 0;
    const size_t off_totrec = align256(off_status + 2 * sizeof(long long));
    const size_t off_count1 = align256(off_totrec + sizeof(long long));
    const size_t off_off1 = align256(off_count1 + (size_t)m.nk1 * sizeof(unsigned));
    const size_t off_count1e = align256(off_off1 + (size_t)(m.nk1 + 1) * sizeof(long long));
    const size_t off_eoff1 = align256(off_count1e + (size_t)m.nk1 * sizeof(unsigned));
    const size_t off_base1 = align256(off_eoff1 + (size_t)(m.nk1 + 1) * sizeof(long long));
    const size_t nbase1 = b.global_l1 ? (size_t)m.nk1 : (size_t)(b.nwg * m.nk1);
    const size_t off_offsets = align256(off_base1 + nbase1 * sizeof(unsigned));
    const size_t off_partials =
        align256(off_offsets + (size_t)(m.nbuckets() + 1) * sizeof(long long));
    const size_t off_l2nblk = align256(off_partials + (size_t)(nblk + 1) * sizeof(long long));
    const size_t off_l2boff = align256(off_l2nblk + (size_t)m.nk1 * sizeof(unsigned));
    const size_t bytes = align256(off_l2boff + (size_t)(m.nk1 + 1) * sizeof(long long));
    char* base = nullptr;
    int r = take(c, ar, c.meta, bytes, &base);
    if (r) return r;
    b.total = (long long*)(base + off_status);
    b.err = (int*)(base + off_status + sizeof(long long));
    b.total_rec = (long long*)(base + off_totrec);
    b.count1 = (unsigned*)(base + off_count1);
    b.off1 = (long long*)(base + off_off1);
    b.count1e = (unsigned*)(base + off_count1e);
    b.eoff1 = (long long*)(base + off_eoff1);
    b.base1 = (unsigned*)(base + off_base1);
    b.offsets = (long long*)(base + off_offsets);
    b.partials = (long long*)(base + off_partials);
    b.l2_nblk = (unsigned*)(base + off_l2nblk);
    b.l2_boff = (long long*)(base + off_l2boff);
    HIPCHK(stage(c, kStCount, st, [&] { return bucket_count(mode, n, uu, vv, ww, g, b, st); }));
    HIPCHK(stage(c, kStScan, st, [&] { return bucket_scan(b, st); }));
    HIPCHK(hipMemcpyAsync(c.status_host, b.total, 2 * sizeof(long long), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    const int err = (int)(c.status_host[1] & 0xffffffff);
    if (err) return fail(LOFAR_ECOORD);
    total = c.status_host[0];
    return LOFAR_OK;
}

// Sample validation of the bucket-free path (the checks of the bucket count): read back
// before anything is written, so that a bad sample leaves the outputs untouched.
int validate_direct(DeviceCtx& c, Arena* ar, long n, const double* uu, const double* vv,
                    const double* ww, const Geometry& g, hipStream_t st)
{
    char* p = nullptr;
    int r = take(c, ar, c.meta, 256, &p);
    if (r) return r;
    HIPCHK(validate_samples(n, uu, vv, ww, g, (int*)p, st));
    HIPCHK(hipMemcpyAsync(&c.status_host[1], p, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if ((int)(c.status_host[1] & 0xffffffff)) return fail(LOFAR_ECOORD);
    return LOFAR_OK;
}

// plan: the samples' sort from lofar_plan_create (uu, vv, ww unused): its entries get
// this call's visibilities, without validation, sorting or host synchronisation.
int grid_impl(DeviceCtx& c, int precision, int accumulate, int nw, long n, long F, long P,
              const double* uu, const double* vv, const double* ww, const float* vr,
              const float* vi, const float* wt, double dx, double dw, int support, int nx, int ny,
              void* cube, hipStream_t st, unsigned long long* occ = nullptr, int sm_log = -1,
              Arena* ar = nullptr, const lofar_plan_s* plan = nullptr)
{
    int r = check_grid_args(precision, nw, n, F, P, dx, dw, support, nx, ny);
    if (r) return fail(r);
    if (!cube || (n > 0 && ((!plan && (!uu || !vv || !ww)) || !vr || !vi || !wt)))
        return fail(LOFAR_EARG);
    const Geometry g = make_geometry(nw, support, nx, ny, dx, dw);
    if (plan && plan->direct) {
        if (occ || sm_log >= 0) return fail(LOFAR_EARG);
        HIPCHK(stage(c, kStGrid, st, [&] {
            return grid_direct(precision, n, F, P, plan->uu, plan->vv, plan->ww, vr, vi, wt, g,
                               cube, accumulate, st);
        }));
        return LOFAR_OK;
    }
    if (plan) {
        Buckets b{};
        b.map = plan->gmap;
        b.offsets = plan->goff;
        b.nent = plan->gtotal;
        if (grid_gathers(precision, accumulate)) {
            // the gridder gathers each entry's sample sum: no filled copy of the entries
            char* vb = nullptr;
            r = take(c, ar, c.ents, align256(plan_vis_bytes(precision, n)) + 256, &vb);
            if (r) return r;
            b.vis = vb;
            HIPCHK(stage(c, kStScatter, st, [&] {
                return plan_vis(precision, n, F, P, vr, vi, wt, vb, st);
            }));
            HIPCHK(stage(c, kStGrid, st, [&] {
                return grid_tiles(precision, g, b, plan->gent, cube, accumulate, occ, sm_log, st);
            }));
            return LOFAR_OK;
        }
        const size_t esz = precision == LOFAR_FP64 ? sizeof(Entry64) : sizeof(Entry32);
        const size_t ebytes = align256((size_t)plan->gtotal * esz);
        char* eb = nullptr;
        r = take(c, ar, c.ents, ebytes + align256(plan_vis_bytes(precision, n)) + 256, &eb);
        if (r) return r;
        HIPCHK(stage(c, kStScatter, st, [&] {
            return plan_fill(precision, n, plan->gtotal, plan->gent, eb, F, P, vr, vi, wt,
                             eb + ebytes, st);
        }));
        HIPCHK(stage(c, kStGrid, st, [&] {
            return grid_tiles(precision, g, b, eb, cube, accumulate, occ, sm_log, st);
        }));
        return LOFAR_OK;
    }
    if (2 * g.K + 1 > kMaxTileWin) {
        // windows wider than the tile kernels take: validated, then gridded with
        // global atomics, no buckets; the cube is written whole (no occupancy bits)
        r = validate_direct(c, ar, n, uu, vv, ww, g, st);
        if (r) return r;
        if (occ || sm_log >= 0) return fail(LOFAR_EARG);
        HIPCHK(stage(c, kStGrid, st, [&] {
            return grid_direct(precision, n, F, P, uu, vv, ww, vr, vi, wt, g, cube, accumulate,
                               st);
        }));
        return LOFAR_OK;
    }
    Buckets b;
    long long total = 0;
    r = build_buckets(c, kBucketFootprint, n, uu, vv, ww, g, st, b, total, ar);
    if (r) return r;
    // level-1 records (at most one per entry), the sorted entries, then the
    // level-2 block table
    const size_t esz = precision == LOFAR_FP64 ? sizeof(Entry64) : sizeof(Entry32);
    const size_t rbytes = align256((size_t)total * sizeof(Entry64));
    const size_t ebytes = align256((size_t)total * esz);
    const size_t l2 = level2_layout(b, total, nullptr);
    char* eb = nullptr;
    r = take(c, ar, c.ents, rbytes + ebytes + l2 + 256, &eb);
    if (r) return r;
    Entry64* rec = (Entry64*)eb;
    void* ents = eb + rbytes;
    level2_layout(b, total, eb + rbytes + ebytes);
    b.nent = total;
    HIPCHK(stage(c, kStScatter, st, [&] {
        return bucket_sort_grid(precision, n, F, P, uu, vv, ww, vr, vi, wt, g, b, rec, ents, st);
    }));
    HIPCHK(stage(c, kStGrid, st,
                 [&] { return grid_tiles(precision, g, b, ents, cube, accumulate, occ, sm_log, st); }));
    return LOFAR_OK;
}

int fft_impl(DeviceCtx& c, int precision, int sign, const void* in, void* out, int nx, int ny,
             int nw, double scale, hipStream_t st)
{
    if (precision != LOFAR_FP64 && precision != LOFAR_FP32C) return fail(LOFAR_EARG);
    if (sign != 1 && sign != -1) return fail(LOFAR_EARG);
    if (nw < 1 || !in || !out) return fail(LOFAR_EARG);
    if (!fft_size_ok(nx) || !fft_size_ok(ny)) return fail(LOFAR_ESIZE);
    if (!pow2_ok(nx) || !pow2_ok(ny))
        return fft_generic_planes(c, nullptr, precision, sign, in, out, nx, ny, nw, scale, st);
    void *twx, *twy;
    int r = twiddles(c, nx, precision, &twx);
    if (r) return r;
    r = twiddles(c, ny, precision, &twy);
    if (r) return r;
    const long long lines = (long long)nw * ny;
    if (precision == LOFAR_FP64) {
        HIPCHK(stage(c, kStCols, st, [&] {
            return fft_cols<double>(sign, in, out, twy, nx, ny, nw, 1.0, st);
        }));
        HIPCHK(stage(c, kStRows, st, [&] {
            return fft_rows<double>(sign, out, out, twx, nx, lines, scale, st);
        }));
    } else {
        HIPCHK(stage(c, kStCols, st, [&] {
            return fft_cols<float>(sign, in, out, twy, nx, ny, nw, 1.0, st);
        }));
        HIPCHK(stage(c, kStRows, st, [&] {
            return fft_rows<float>(sign, out, out, twx, nx, lines, scale, st);
        }));
    }
    return LOFAR_OK;
}

int check_phase_args(int precision, int xaxis, int yaxis, int nw, int xtot, int ytot)
{
    if (precision != LOFAR_FP64 && precision != LOFAR_FP32C) return LOFAR_EARG;
    if (xaxis < 1 || yaxis < 1 || nw < 1 || xtot < 1 || ytot < 1) return LOFAR_EARG;
    return LOFAR_OK;
}

// Calls on one device are serialised on its mutex, and each call's work is
// ordered after the previous call's (event wait), so the shared workspace is
// never used by two calls at once even when they use different streams.
struct CallGuard {
    DeviceCtx* c = nullptr;
    hipStream_t st = nullptr;
    std::unique_lock<std::mutex> lk;
    int status = LOFAR_OK;
    explicit CallGuard(void* stream)
    {
        status = get_ctx(&c);
        if (status) {
            t_last_error = status;
            return;
        }
        lk = std::unique_lock<std::mutex>(c->mu);
        st = (hipStream_t)stream;
        if (hipStreamWaitEvent(st, c->last_done, 0) != hipSuccess) status = fail(LOFAR_EHIP);
    }
    ~CallGuard()
    {
        if (c && lk.owns_lock()) (void)hipEventRecord(c->last_done, st);
    }
};

void report_once(const char* what, int code)
{
    static std::atomic<bool> printed{false};
    if (!printed.exchange(true))
        fprintf(stderr, "liblofar_hip: %s failed: %s (further errors not printed)\n", what,
                lofar_error_string(code));
}

// strip width (log2 columns) of the fused path's scratch cube; -1 = row-major
int sm_log_cfg(int precision)
{
    static int v = -2;
    if (v == -2) {
        const char* e = getenv("LOFAR_SM_LOG");
        v = e ? atoi(e) : -3;
        if (v < -1 || v > 6) v = -3;
    }
    if (v != -3) return v;
    return precision == LOFAR_FP64 ? 6 : -1;
}

// 8192-wide rows: split row passes (LOFAR_ROWS_SPLIT=0 keeps the fused ones, for
// measurement)
int rows_split()
{
    static int v = -1;
    if (v < 0) {
        const char* e = getenv("LOFAR_ROWS_SPLIT");
        v = e ? atoi(e) != 0 : 1;
    }
    return v;
}
}  // namespace
}  // namespace lofar

using namespace lofar;

extern "C" {

int lofar_last_error(void) { return t_last_error; }

const char* lofar_error_string(int code)
{
    switch (code) {
        case LOFAR_OK: return "ok";
        case LOFAR_EARG: return "bad argument";
        case LOFAR_ECOORD: return "sample not finite or outside the w-planes";
        case LOFAR_ESIZE: return "FFT side must be in [2, 8192] with no prime factor above 13";
        case LOFAR_EHIP: return "HIP runtime error";
        case LOFAR_ENOMEM: return "device allocation failed";
        case LOFAR_ENODEV: return "no usable gfx950 device";
        case LOFAR_ESCRATCH: return "workspace exceeds the allocation-free entry point's static array";
        default: return "unknown error";
    }
}

int lofar_version(void) { return 10000 * 0 + 100 * 1 + 0; }

int test(int nnn) { return nnn + 1; }

int lofar_profile_enable(int on)
{
    g_profiling.store(on != 0);
    return LOFAR_OK;
}

int lofar_profile_read(double* total_ms, int* launches)
{
    DeviceCtx* c = nullptr;
    int r = get_ctx(&c);
    if (r) return fail(r);
    std::lock_guard<std::mutex> lk(c->mu);
    for (int i = 0; i < LOFAR_NSTAGES; i++) {
        if (total_ms) total_ms[i] = 0.0;
        if (launches) launches[i] = 0;
    }
    for (auto& p : c->prof) {
        float ms = 0.f;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            if (total_ms) total_ms[p.stage] += ms;
            if (launches) launches[p.stage] += 1;
        }
        c->pool.push_back(p.a);
        c->pool.push_back(p.b);
    }
    c->prof.clear();
    return LOFAR_OK;
}

const char* lofar_stage_name(int stage)
{
    return (stage >= 0 && stage < LOFAR_NSTAGES) ? kStageNames[stage] : "unknown";
}

int lofar_grid_dev(int precision, int accumulate, int num_w_planes, long num_points,
                   long freq_per_chan, long polarizations, const double* uu, const double* vv,
                   const double* ww, const float* vis_real, const float* vis_img,
                   const float* weight, double dx, double dw, int w_support, int grid_size_x,
                   int grid_size_y, void* cube, void* stream)
{
    CallGuard cg(stream);
    if (cg.status) return cg.status;
    return grid_impl(*cg.c, precision, accumulate, num_w_planes, num_points, freq_per_chan,
                     polarizations, uu, vv, ww, vis_real, vis_img, weight, dx, dw, w_support,
                     grid_size_x, grid_size_y, cube, cg.st);
}

int lofar_fft_dev(int precision, int sign, const void* in, void* out, int nx, int ny, int nw,
                  double scale, void* stream)
{
    CallGuard cg(stream);
    if (cg.status) return cg.status;
    return fft_impl(*cg.c, precision, sign, in, out, nx, ny, nw, scale, cg.st);
}

int lofar_phase_dev(int precision, const void* cube, double* image_real, double* image_imag,
                    int xaxis, int yaxis, int num_w_planes, int xaxistot, int yaxistot,
                    double resolution, double wmin, double wmax, int accumulate, void* stream)
{
    int r = check_phase_args(precision, xaxis, yaxis, num_w_planes, xaxistot, yaxistot);
    if (r) return fail(r);
    if (!cube || !image_real || !image_imag) return fail(LOFAR_EARG);
    CallGuard cg(stream);
    if (cg.status) return cg.status;
    DeviceCtx& c = *cg.c;
    HIPCHK(stage(c, kStPhase, cg.st, [&] {
        return phase_sum(precision, cube, image_real, image_imag, xaxis, yaxis, num_w_planes,
                         xaxistot, yaxistot, resolution, wmin, wmax, accumulate, cg.st);
    }));
    return LOFAR_OK;
}

int lofar_phase_adjoint_dev(int precision, const double* image_real, const double* image_imag,
                            void* cube, int xaxis, int yaxis, int num_w_planes, int xaxistot,
                            int yaxistot, double resolution, double wmin, double wmax, void* stream)
{
    int r = check_phase_args(precision, xaxis, yaxis, num_w_planes, xaxistot, yaxistot);
    if (r) return fail(r);
    if (!cube || !image_real || !image_imag) return fail(LOFAR_EARG);
    CallGuard cg(stream);
    if (cg.status) return cg.status;
    DeviceCtx& c = *cg.c;
    HIPCHK(stage(c, kStPhase, cg.st, [&] {
        return phase_adjoint(precision, image_real, image_imag, cube, xaxis, yaxis, num_w_planes,
                             xaxistot, yaxistot, resolution, wmin, wmax, cg.st);
    }));
    return LOFAR_OK;
}

// The fused imaging pipeline of lofar_image_dev and lofar_image_dev_static; `ar`
// selects where the workspace comes from (nullptr: the context's buffers).
static int image_impl(DeviceCtx& c, hipStream_t st, Arena* ar, int precision, int num_w_planes,
                      long num_points, long freq_per_chan, long polarizations, const double* uu,
                      const double* vv, const double* ww, const float* vis_real,
                      const float* vis_img, const float* weight, double dx, double dw,
                      int w_support, int grid_size_x, int grid_size_y, double resolution,
                      double wmin, double wmax, void* cube, double* image_real,
                      double* image_imag, int accumulate, const lofar_plan_s* plan = nullptr)
{
    // per-tile occupancy bits: the gridder skips empty tiles and the column pass
    // reads them as zeros (through 32-bit plane offsets, so for planes below 2 GiB;
    // larger planes are written whole)
    int r = LOFAR_OK;
    unsigned long long* occ = nullptr;
    const long long plane_bytes = (long long)grid_size_x * grid_size_y *
                                  (precision == LOFAR_FP64 ? 16 : 8);
    // windows wider than the tile kernels take are gridded without buckets into a whole
    // row-major cube
    const bool direct = w_support >= 1 && 2 * ((w_support - 1) / 2) + 1 > kMaxTileWin;
    // sides that are not powers of two: the mixed-radix transforms and the standalone
    // phase sum over a whole row-major cube
    const bool generic = !pow2_ok(grid_size_x) || !pow2_ok(grid_size_y);
    if (plane_bytes < (1LL << 31) && !direct && !generic) {
        const long long ntv = (grid_size_y + kTile - 1) / kTile;
        const long long ntu = (grid_size_x + kTile - 1) / kTile;
        const long long nocc = (long long)num_w_planes * ntu * ((ntv + 63) / 64);
        char* op = nullptr;
        r = take(c, ar, c.occ, (size_t)nocc * 8 + 256, &op);
        if (r) return r;
        occ = (unsigned long long*)op;
        HIPCHK(fill_u32((unsigned*)occ, 2 * nocc, 0u, st));
    }
    // FP64: the scratch cube is laid out strip-major, in strips of 2^sm_log whole
    // columns (the 64-column tile width by default), so every tile flush is one
    // contiguous block and the column transforms read 64-byte runs 1 KiB apart instead
    // of a plane row apart; FP32C keeps row-major planes (no gain measured there).
    // LOFAR_SM_LOG overrides, for measurement (-1: row-major; at most the tile width).
    // rows of 8192 run the split row passes, which read row-major planes
    const bool split = grid_size_x >= 8192 && rows_split();
    const int sm_pref = (direct || generic || split) ? -1 : sm_log_cfg(precision);
    const int sm_log = (sm_pref >= 0 && (1 << sm_pref) <= kTile &&
                        grid_size_x % (1 << sm_pref) == 0 &&
                        (1 << sm_pref) % (precision == LOFAR_FP64 ? 4 : 8) == 0)
                           ? sm_pref
                           : -1;
    r = grid_impl(c, precision, 0, num_w_planes, num_points, freq_per_chan, polarizations, uu, vv,
                  ww, vis_real, vis_img, weight, dx, dw, w_support, grid_size_x, grid_size_y, cube,
                  st, occ, sm_log, ar, plan);
    if (r) return r;
    if (generic) {
        r = fft_generic_planes(c, ar, precision, 1, cube, cube, grid_size_x, grid_size_y,
                               num_w_planes, 1.0 / ((double)grid_size_x * grid_size_y), st);
        if (r) return r;
        HIPCHK(stage(c, kStPhase, st, [&] {
            return phase_sum(precision, cube, image_real, image_imag, grid_size_x, grid_size_y,
                             num_w_planes, grid_size_x, grid_size_y, resolution, wmin, wmax,
                             accumulate, st);
        }));
        return LOFAR_OK;
    }
    void *twx, *twy;
    r = twiddles_for(c, ar, grid_size_x, precision, st, &twx);
    if (r) return r;
    r = twiddles_for(c, ar, grid_size_y, precision, st, &twy);
    if (r) return r;
    const PhaseParams pp =
        phase_params(num_w_planes, grid_size_x, grid_size_y, resolution, wmin, wmax);
    const double scale = 1.0 / ((double)grid_size_x * (double)grid_size_y);
    HIPCHK(stage(c, kStCols, st, [&] {
        return precision == LOFAR_FP64
                   ? fft_cols<double>(1, cube, cube, twy, grid_size_x, grid_size_y, num_w_planes,
                                      1.0, st, occ, sm_log)
                   : fft_cols<float>(1, cube, cube, twy, grid_size_x, grid_size_y, num_w_planes,
                                     1.0, st, occ, sm_log);
    }));
    if (split) {
        // rows of 8192 run the fused pass at 1024 threads a workgroup, where its
        // per-pixel state does not fit the registers: the row transform (in place)
        // and the phase/plane sum run as two passes instead
        const long long lines = (long long)num_w_planes * grid_size_y;
        HIPCHK(stage(c, kStRows, st, [&] {
            return precision == LOFAR_FP64
                       ? fft_rows<double>(1, cube, cube, twx, grid_size_x, lines, scale, st)
                       : fft_rows<float>(1, cube, cube, twx, grid_size_x, lines, scale, st);
        }));
        HIPCHK(stage(c, kStPhase, st, [&] {
            return phase_sum(precision, cube, image_real, image_imag, grid_size_x, grid_size_y,
                             num_w_planes, grid_size_x, grid_size_y, resolution, wmin, wmax,
                             accumulate, st);
        }));
    } else {
        HIPCHK(stage(c, kStRowsPhase, st, [&] {
            return precision == LOFAR_FP64
                       ? rows_phase<double>(cube, twx, grid_size_x, grid_size_y, num_w_planes, pp,
                                            scale, image_real, image_imag, accumulate, st, sm_log)
                       : rows_phase<float>(cube, twx, grid_size_x, grid_size_y, num_w_planes, pp,
                                           scale, image_real, image_imag, accumulate, st, sm_log);
        }));
    }
    return LOFAR_OK;
}

int lofar_image_dev(int precision, int num_w_planes, long num_points, long freq_per_chan,
                    long polarizations, const double* uu, const double* vv, const double* ww,
                    const float* vis_real, const float* vis_img, const float* weight, double dx,
                    double dw, int w_support, int grid_size_x, int grid_size_y, double resolution,
                    double wmin, double wmax, void* cube, double* image_real, double* image_imag,
                    int accumulate, void* stream)
{
    if (!fft_size_ok(grid_size_x) || !fft_size_ok(grid_size_y)) return fail(LOFAR_ESIZE);
    if (!image_real || !image_imag) return fail(LOFAR_EARG);
    CallGuard cg(stream);
    if (cg.status) return cg.status;
    return image_impl(*cg.c, cg.st, nullptr, precision, num_w_planes, num_points, freq_per_chan,
                      polarizations, uu, vv, ww, vis_real, vis_img, weight, dx, dw, w_support,
                      grid_size_x, grid_size_y, resolution, wmin, wmax, cube, image_real,
                      image_imag, accumulate);
}

int lofar_image_dev_static(int precision, int num_w_planes, long num_points,
                           long freq_per_chan, long polarizations, const double* uu,
                           const double* vv, const double* ww, const float* vis_real,
                           const float* vis_img, const float* weight, double dx, double dw,
                           int w_support, int grid_size_x, int grid_size_y, double resolution,
                           double wmin, double wmax, void* cube, double* image_real,
                           double* image_imag, int accumulate, void* stream)
{
    if (!fft_size_ok(grid_size_x) || !fft_size_ok(grid_size_y)) return fail(LOFAR_ESIZE);
    if (!image_real || !image_imag) return fail(LOFAR_EARG);
    CallGuard cg(stream);
    if (cg.status) return cg.status;
    if (!cg.c->scratch) {
        fprintf(stderr, "liblofar_hip: static workspace unavailable on device %d: %s\n",
                cg.c->device, hipGetErrorString(cg.c->scratch_err));
        return fail(LOFAR_EHIP);
    }
    Arena ar;
    ar.base = cg.c->scratch;
    ar.cap = kScratchBytes;
    return image_impl(*cg.c, cg.st, &ar, precision, num_w_planes, num_points, freq_per_chan,
                      polarizations, uu, vv, ww, vis_real, vis_img, weight, dx, dw, w_support,
                      grid_size_x, grid_size_y, resolution, wmin, wmax, cube, image_real,
                      image_imag, accumulate);
}

int lofar_degrid_cube_dev(int precision, int num_w_planes, long num_points, long freq_per_chan,
                          long polarizations, const double* uu, const double* vv, const double* ww,
                          const float* weight, double dx, double dw, int w_support,
                          int grid_size_x, int grid_size_y, const void* cube, void* model_real,
                          void* model_imag, void* stream)
{
    int r = check_grid_args(precision, num_w_planes, num_points, freq_per_chan, polarizations, dx,
                            dw, w_support, grid_size_x, grid_size_y);
    if (r) return fail(r);
    if (!cube || (num_points > 0 && (!uu || !vv || !ww || !weight || !model_real || !model_imag)))
        return fail(LOFAR_EARG);
    CallGuard cg(stream);
    if (cg.status) return cg.status;
    DeviceCtx& c = *cg.c;
    const Geometry g =
        make_geometry(num_w_planes, w_support, grid_size_x, grid_size_y, dx, dw);
    if (2 * g.K + 1 > kMaxTileWin) {
        r = validate_direct(c, nullptr, num_points, uu, vv, ww, g, cg.st);
        if (r) return r;
        HIPCHK(stage(c, kStDegrid, cg.st, [&] {
            return degrid_direct(precision, num_points, freq_per_chan, polarizations, uu, vv, ww,
                                 weight, g, cube, model_real, model_imag, cg.st);
        }));
        return LOFAR_OK;
    }
    Buckets b;
    long long total = 0;
    r = build_buckets(c, kBucketHome, num_points, uu, vv, ww, g, cg.st, b, total);
    if (r) return r;
    const size_t ebytes = align256((size_t)total * sizeof(DEntry));
    const size_t l2 = level2_layout(b, total, nullptr);
    r = ensure(c, c.ents, 2 * ebytes + l2 + 256);
    if (r) return r;
    char* eb = (char*)c.ents.ptr;
    DEntry* rec = (DEntry*)eb;
    DEntry* ents = (DEntry*)(eb + ebytes);
    level2_layout(b, total, eb + 2 * ebytes);
    HIPCHK(stage(c, kStScatter, cg.st, [&] {
        return bucket_sort_home(num_points, uu, vv, ww, polarizations == 1 ? weight : nullptr, g,
                                b, rec, ents, cg.st);
    }));
    HIPCHK(stage(c, kStDegrid, cg.st, [&] {
        return degrid_tiles(precision, g, b, ents, cube, weight, freq_per_chan, polarizations,
                            model_real, model_imag, cg.st);
    }));
    return LOFAR_OK;
}

// The transforms of the adjoint (lofar_degrid_dev) up to the interpolation: conjugate
// phase screen, forward row and column transforms with the 1/(nx*ny), into `cube`.
// need (optional): the tiles the interpolation reads (need_tiles); the last column
// transform stores only those.
static int degrid_transform(DeviceCtx& c, hipStream_t st, int precision, int num_w_planes,
                            int grid_size_x, int grid_size_y, double resolution, double wmin,
                            double wmax, const double* image_real, const double* image_imag,
                            void* cube, const unsigned long long* need = nullptr)
{
    int r = LOFAR_OK;
    void *twx, *twy;
    r = twiddles(c, grid_size_x, precision, &twx);
    if (r) return r;
    r = twiddles(c, grid_size_y, precision, &twy);
    if (r) return r;
    const PhaseParams pp =
        phase_params(num_w_planes, grid_size_x, grid_size_y, resolution, wmin, wmax);
    const double scale = 1.0 / ((double)grid_size_x * (double)grid_size_y);
    if (!pow2_ok(grid_size_x) || !pow2_ok(grid_size_y)) {
        // sides that are not powers of two: conjugate screen, then the mixed-radix
        // forward transforms (rows, then columns with the 1/(nx*ny))
        HIPCHK(stage(c, kStRowsPhaseAdj, st, [&] {
            return phase_adjoint(precision, image_real, image_imag, cube, grid_size_x,
                                 grid_size_y, num_w_planes, grid_size_x, grid_size_y,
                                 resolution, wmin, wmax, st);
        }));
        r = fft_generic_planes(c, nullptr, precision, -1, cube, cube, grid_size_x,
                               grid_size_y, num_w_planes, scale, st, true);
        if (r) return r;
    } else if (grid_size_x >= 8192 && rows_split()) {
        // rows of 8192: the fused adjoint row pass does not fit its 1024-thread
        // workgroups' registers, so the conjugate screen (every plane written) and
        // the forward row transform (in place) run as two passes
        HIPCHK(stage(c, kStRowsPhaseAdj, st, [&] {
            const hipError_t e =
                phase_adjoint(precision, image_real, image_imag, cube, grid_size_x,
                              grid_size_y, num_w_planes, grid_size_x, grid_size_y,
                              resolution, wmin, wmax, st);
            if (e != hipSuccess) return e;
            const long long lines = (long long)num_w_planes * grid_size_y;
            return precision == LOFAR_FP64
                       ? fft_rows<double>(-1, cube, cube, twx, grid_size_x, lines, 1.0, st)
                       : fft_rows<float>(-1, cube, cube, twx, grid_size_x, lines, 1.0, st);
        }));
        HIPCHK(stage(c, kStCols, st, [&] {
            return precision == LOFAR_FP64
                       ? fft_cols<double>(-1, cube, cube, twy, grid_size_x, grid_size_y,
                                          num_w_planes, scale, st, need)
                       : fft_cols<float>(-1, cube, cube, twy, grid_size_x, grid_size_y,
                                         num_w_planes, scale, st, need);
        }));
    } else if (precision == LOFAR_FP64) {
        HIPCHK(stage(c, kStRowsPhaseAdj, st, [&] {
            return rows_phase_adj<double>(cube, twx, grid_size_x, grid_size_y, num_w_planes,
                                          pp, 1.0, image_real, image_imag, st);
        }));
        HIPCHK(stage(c, kStCols, st, [&] {
            return fft_cols<double>(-1, cube, cube, twy, grid_size_x, grid_size_y,
                                    num_w_planes, scale, st, need);
        }));
    } else {
        HIPCHK(stage(c, kStRowsPhaseAdj, st, [&] {
            return rows_phase_adj<float>(cube, twx, grid_size_x, grid_size_y, num_w_planes,
                                         pp, 1.0, image_real, image_imag, st);
        }));
        HIPCHK(stage(c, kStCols, st, [&] {
            return fft_cols<float>(-1, cube, cube, twy, grid_size_x, grid_size_y,
                                   num_w_planes, scale, st, need);
        }));
    }
    return LOFAR_OK;
}

int lofar_degrid_dev(int precision, int num_w_planes, long num_points, long freq_per_chan,
                     long polarizations, const double* uu, const double* vv, const double* ww,
                     const float* weight, double dx, double dw, int w_support, int grid_size_x,
                     int grid_size_y, double resolution, double wmin, double wmax,
                     const double* image_real, const double* image_imag, void* cube,
                     void* model_real, void* model_imag, void* stream)
{
    if (!fft_size_ok(grid_size_x) || !fft_size_ok(grid_size_y)) return fail(LOFAR_ESIZE);
    int r = check_grid_args(precision, num_w_planes, num_points, freq_per_chan, polarizations, dx,
                            dw, w_support, grid_size_x, grid_size_y);
    if (r) return fail(r);
    if (!cube || !image_real || !image_imag ||
        (num_points > 0 && (!uu || !vv || !ww || !weight || !model_real || !model_imag)))
        return fail(LOFAR_EARG);
    CallGuard cg(stream);
    if (cg.status) return cg.status;
    DeviceCtx& c = *cg.c;
    const hipStream_t st = cg.st;
    const Geometry g = make_geometry(num_w_planes, w_support, grid_size_x, grid_size_y, dx, dw);
    if (2 * g.K + 1 > kMaxTileWin) {
        // wide windows: validated, transformed whole, interpolated per sample
        r = validate_direct(c, nullptr, num_points, uu, vv, ww, g, st);
        if (r) return r;
        r = degrid_transform(c, st, precision, num_w_planes, grid_size_x, grid_size_y,
                             resolution, wmin, wmax, image_real, image_imag, cube);
        if (r) return r;
        HIPCHK(stage(c, kStDegrid, st, [&] {
            return degrid_direct(precision, num_points, freq_per_chan, polarizations, uu, vv, ww,
                                 weight, g, cube, model_real, model_imag, st);
        }));
        return LOFAR_OK;
    }
    // the home sort first (a bad sample is reported before anything is written), then
    // the tiles it reads, the transforms storing only those, and the interpolation
    Buckets b;
    long long total = 0;
    r = build_buckets(c, kBucketHome, num_points, uu, vv, ww, g, st, b, total);
    if (r) return r;
    const size_t ebytes = align256((size_t)total * sizeof(DEntry));
    const size_t l2 = level2_layout(b, total, nullptr);
    r = ensure(c, c.ents, 2 * ebytes + l2 + 256);
    if (r) return r;
    char* eb = (char*)c.ents.ptr;
    DEntry* rec = (DEntry*)eb;
    DEntry* ents = (DEntry*)(eb + ebytes);
    level2_layout(b, total, eb + 2 * ebytes);
    unsigned long long* need = nullptr;
    HIPCHK(stage(c, kStScatter, st, [&] {
        return bucket_sort_home(num_points, uu, vv, ww, polarizations == 1 ? weight : nullptr, g,
                                b, rec, ents, st);
    }));
    if (pow2_ok(grid_size_x) && pow2_ok(grid_size_y)) {
        r = ensure(c, c.occ, (size_t)need_words(g) * 8 + 256);
        if (r) return r;
        need = (unsigned long long*)c.occ.ptr;
        HIPCHK(stage(c, kStScatter, st, [&] { return need_tiles(g, b, need, st); }));
    }
    r = degrid_transform(c, st, precision, num_w_planes, grid_size_x, grid_size_y, resolution,
                         wmin, wmax, image_real, image_imag, cube, need);
    if (r) return r;
    HIPCHK(stage(c, kStDegrid, st, [&] {
        return degrid_tiles(precision, g, b, ents, cube, weight, freq_per_chan, polarizations,
                            model_real, model_imag, st);
    }));
    return LOFAR_OK;
}

// ---------------------------------------------------------------------------
// plans: one sort of fixed samples, kept for repeated imaging and degridding

int lofar_plan_create(int precision, int num_w_planes, long num_points, long freq_per_chan,
                      long polarizations, const double* uu, const double* vv, const double* ww,
                      const float* weight, double dx, double dw, int w_support, int grid_size_x,
                      int grid_size_y, void* stream, lofar_plan* plan)
{
    if (!plan) return fail(LOFAR_EARG);
    *plan = nullptr;
    if (!fft_size_ok(grid_size_x) || !fft_size_ok(grid_size_y)) return fail(LOFAR_ESIZE);
    int r = check_grid_args(precision, num_w_planes, num_points, freq_per_chan, polarizations, dx,
                            dw, w_support, grid_size_x, grid_size_y);
    if (r) return fail(r);
    if (num_points > 0 && (!uu || !vv || !ww || !weight)) return fail(LOFAR_EARG);
    // an FP32C entry carries its sample's index in the 32 bits of its real part's slot
    if (precision == LOFAR_FP32C && num_points > 0xffffffffL) return fail(LOFAR_ESIZE);
    CallGuard cg(stream);
    if (cg.status) return cg.status;
    DeviceCtx& c = *cg.c;
    const hipStream_t st = cg.st;
    std::unique_ptr<lofar_plan_s, int (*)(lofar_plan)> p(new (std::nothrow) lofar_plan_s(),
                                                         [](lofar_plan q) {
                                                             if (q) {
                                                                 (void)hipFree(q->gmem);
                                                                 (void)hipFree(q->hmem);
                                                             }
                                                             delete q;
                                                             return 0;
                                                         });
    if (!p) return fail(LOFAR_ENOMEM);
    p->device = c.device;
    p->precision = precision;
    p->nw = num_w_planes;
    p->support = w_support;
    p->nx = grid_size_x;
    p->ny = grid_size_y;
    p->n = num_points;
    p->F = freq_per_chan;
    p->P = polarizations;
    p->dx = dx;
    p->dw = dw;
    const Geometry g = make_geometry(num_w_planes, w_support, grid_size_x, grid_size_y, dx, dw);
    p->direct = 2 * g.K + 1 > kMaxTileWin;
    const size_t wbytes = align256((size_t)num_points * polarizations * sizeof(float));
    auto alloc = [&](char** mem, size_t bytes) {
        if (hipMalloc((void**)mem, bytes + 256) != hipSuccess) {
            *mem = nullptr;
            (void)hipGetLastError();
            return fail(LOFAR_ENOMEM);
        }
        return (int)LOFAR_OK;
    };
    if (p->direct) {
        r = validate_direct(c, nullptr, num_points, uu, vv, ww, g, st);
        if (r) return r;
        const size_t cb = align256((size_t)num_points * sizeof(double));
        r = alloc(&p->hmem, 3 * cb + wbytes);
        if (r) return r;
        p->uu = (double*)p->hmem;
        p->vv = (double*)(p->hmem + cb);
        p->ww = (double*)(p->hmem + 2 * cb);
        p->wt = (float*)(p->hmem + 3 * cb);
        if (num_points > 0) {
            HIPCHK(hipMemcpyAsync(p->uu, uu, num_points * sizeof(double), hipMemcpyDeviceToDevice, st));
            HIPCHK(hipMemcpyAsync(p->vv, vv, num_points * sizeof(double), hipMemcpyDeviceToDevice, st));
            HIPCHK(hipMemcpyAsync(p->ww, ww, num_points * sizeof(double), hipMemcpyDeviceToDevice, st));
            HIPCHK(hipMemcpyAsync(p->wt, weight, num_points * polarizations * sizeof(float),
                                  hipMemcpyDeviceToDevice, st));
        }
    } else {
        // gridding: the sort of lofar_grid_dev with sample indices for visibilities
        Buckets b;
        long long total = 0;
        r = build_buckets(c, kBucketFootprint, num_points, uu, vv, ww, g, st, b, total);
        if (r) return r;
        const size_t esz = precision == LOFAR_FP64 ? sizeof(Entry64) : sizeof(Entry32);
        const size_t rbytes = align256((size_t)total * sizeof(Entry64));
        const size_t ebytes = align256((size_t)total * esz);
        const size_t l2 = level2_layout(b, total, nullptr);
        r = ensure(c, c.ents, rbytes + ebytes + l2 + 256);
        if (r) return r;
        char* eb = (char*)c.ents.ptr;
        level2_layout(b, total, eb + rbytes + ebytes);
        HIPCHK(bucket_sort_grid(precision, num_points, freq_per_chan, polarizations, uu, vv, ww,
                                nullptr, nullptr, nullptr, g, b, (Entry64*)eb, eb + rbytes, st,
                                true));
        const long long nbg = b.map.nbuckets();
        const size_t obytes = align256((size_t)(nbg + 1) * sizeof(long long));
        r = alloc(&p->gmem, obytes + ebytes);
        if (r) return r;
        p->gmap = b.map;
        p->gtotal = total;
        p->goff = (long long*)p->gmem;
        p->gent = p->gmem + obytes;
        HIPCHK(hipMemcpyAsync(p->goff, b.offsets, (nbg + 1) * sizeof(long long),
                              hipMemcpyDeviceToDevice, st));
        if (total > 0)
            HIPCHK(hipMemcpyAsync(p->gent, eb + rbytes, total * esz, hipMemcpyDeviceToDevice, st));
        // the home sort below reuses (and may reallocate) the workspace copied from
        HIPCHK(hipStreamSynchronize(st));
        // degridding: the home sort of lofar_degrid_cube_dev
        Buckets hb;
        long long htotal = 0;
        r = build_buckets(c, kBucketHome, num_points, uu, vv, ww, g, st, hb, htotal);
        if (r) return r;
        const size_t hbytes = align256((size_t)htotal * sizeof(DEntry));
        const size_t hl2 = level2_layout(hb, htotal, nullptr);
        r = ensure(c, c.ents, 2 * hbytes + hl2 + 256);
        if (r) return r;
        char* hbuf = (char*)c.ents.ptr;
        level2_layout(hb, htotal, hbuf + 2 * hbytes);
        HIPCHK(bucket_sort_home(num_points, uu, vv, ww, polarizations == 1 ? weight : nullptr, g,
                                hb, (DEntry*)hbuf, (DEntry*)(hbuf + hbytes), st));
        const long long nbh = hb.map.nbuckets();
        const size_t hobytes = align256((size_t)(nbh + 1) * sizeof(long long));
        const bool pow2 = pow2_ok(grid_size_x) && pow2_ok(grid_size_y);
        const size_t nbytes = pow2 ? align256((size_t)need_words(g) * 8) : 0;
        r = alloc(&p->hmem, hobytes + hbytes + wbytes + nbytes);
        if (r) return r;
        p->hmap = hb.map;
        p->htotal = htotal;
        p->hoff = (long long*)p->hmem;
        p->hent = (DEntry*)(p->hmem + hobytes);
        p->wt = (float*)(p->hmem + hobytes + hbytes);
        if (pow2) {
            p->need = (unsigned long long*)(p->hmem + hobytes + hbytes + wbytes);
            HIPCHK(need_tiles(g, hb, p->need, st));
        }
        HIPCHK(hipMemcpyAsync(p->hoff, hb.offsets, (nbh + 1) * sizeof(long long),
                              hipMemcpyDeviceToDevice, st));
        if (htotal > 0)
            HIPCHK(hipMemcpyAsync(p->hent, hbuf + hbytes, htotal * sizeof(DEntry),
                                  hipMemcpyDeviceToDevice, st));
        if (num_points > 0)
            HIPCHK(hipMemcpyAsync(p->wt, weight, num_points * polarizations * sizeof(float),
                                  hipMemcpyDeviceToDevice, st));
    }
    // the caller may change or free its arrays once this returns
    HIPCHK(hipStreamSynchronize(st));
    *plan = p.release();
    return LOFAR_OK;
}

int lofar_plan_image(lofar_plan plan, const float* vis_real, const float* vis_img,
                     const float* weight, double resolution, double wmin, double wmax,
                     void* cube, double* image_real, double* image_imag, int accumulate,
                     void* stream)
{
    if (!plan || !image_real || !image_imag) return fail(LOFAR_EARG);
    CallGuard cg(stream);
    if (cg.status) return cg.status;
    if (cg.c->device != plan->device) return fail(LOFAR_EARG);
    return image_impl(*cg.c, cg.st, nullptr, plan->precision, plan->nw, plan->n, plan->F,
                      plan->P, nullptr, nullptr, nullptr, vis_real, vis_img, weight, plan->dx,
                      plan->dw, plan->support, plan->nx, plan->ny, resolution, wmin, wmax, cube,
                      image_real, image_imag, accumulate, plan);
}

int lofar_plan_degrid(lofar_plan plan, double resolution, double wmin, double wmax,
                      const double* image_real, const double* image_imag, void* cube,
                      void* model_real, void* model_imag, void* stream)
{
    if (!plan || !cube || !image_real || !image_imag) return fail(LOFAR_EARG);
    if (plan->n > 0 && (!model_real || !model_imag)) return fail(LOFAR_EARG);
    CallGuard cg(stream);
    if (cg.status) return cg.status;
    if (cg.c->device != plan->device) return fail(LOFAR_EARG);
    DeviceCtx& c = *cg.c;
    int r = degrid_transform(c, cg.st, plan->precision, plan->nw, plan->nx, plan->ny, resolution,
                             wmin, wmax, image_real, image_imag, cube, plan->need);
    if (r) return r;
    const Geometry g = make_geometry(plan->nw, plan->support, plan->nx, plan->ny, plan->dx,
                                     plan->dw);
    if (plan->direct) {
        HIPCHK(stage(c, kStDegrid, cg.st, [&] {
            return degrid_direct(plan->precision, plan->n, plan->F, plan->P, plan->uu, plan->vv,
                                 plan->ww, plan->wt, g, cube, model_real, model_imag, cg.st);
        }));
        return LOFAR_OK;
    }
    Buckets hb{};
    hb.map = plan->hmap;
    hb.offsets = plan->hoff;
    HIPCHK(stage(c, kStDegrid, cg.st, [&] {
        return degrid_tiles(plan->precision, g, hb, plan->hent, cube, plan->wt, plan->F, plan->P,
                            model_real, model_imag, cg.st);
    }));
    return LOFAR_OK;
}

int lofar_plan_destroy(lofar_plan plan)
{
    if (!plan) return LOFAR_OK;
    int cur = -1;
    if (hipGetDevice(&cur) == hipSuccess && cur != plan->device) (void)hipSetDevice(plan->device);
    // hipFree waits for the device's queued work, which may still read the plan
    (void)hipFree(plan->gmem);
    (void)hipFree(plan->hmem);
    if (cur >= 0 && cur != plan->device) (void)hipSetDevice(cur);
    delete plan;
    return LOFAR_OK;
}

// ---------------------------------------------------------------------------
// drop-ins: host pointers in, host results out, synchronous

void wstack(int num_w_planes, long num_points, long freq_per_chan, long polarizations,
            double* uu, double* vv, double* ww, float* vis_real, float* vis_img, float* weight,
            double dx, double dw, int w_support, int grid_size_x, int grid_size_y, double* grid,
            int num_threads)
{
    (void)num_threads;  // OpenMP thread count of the CPU reference; unused on the GPU
    t_last_error = LOFAR_OK;
    int r = check_grid_args(LOFAR_FP64, num_w_planes, num_points, freq_per_chan, polarizations, dx,
                            dw, w_support, grid_size_x, grid_size_y);
    if (r || !grid || (num_points > 0 && (!uu || !vv || !ww || !vis_real || !vis_img || !weight))) {
        fail(r ? r : LOFAR_EARG);
        report_once("wstack", t_last_error);
        return;
    }
    CallGuard cg(nullptr);
    if (cg.status) {
        report_once("wstack", cg.status);
        return;
    }
    DeviceCtx& c = *cg.c;
    cg.st = c.stream;
    (void)hipStreamWaitEvent(cg.st, c.last_done, 0);
    const size_t n = (size_t)num_points, nfp = n * freq_per_chan * polarizations;
    const size_t cube_bytes = (size_t)2 * num_w_planes * grid_size_x * grid_size_y * sizeof(double);
    const size_t o_uu = 0, o_vv = align256(o_uu + 8 * n), o_ww = align256(o_vv + 8 * n);
    const size_t o_vr = align256(o_ww + 8 * n), o_vi = align256(o_vr + 4 * nfp);
    const size_t o_wt = align256(o_vi + 4 * nfp), o_cube = align256(o_wt + 4 * n * polarizations);
    r = ensure(c, c.host_io, o_cube + cube_bytes);
    if (r) {
        report_once("wstack", r);
        return;
    }
    char* d = (char*)c.host_io.ptr;
    if (n > 0) {
        r = h2d_staged(c, d + o_uu, uu, 8 * n, cg.st);
        if (!r) r = h2d_staged(c, d + o_vv, vv, 8 * n, cg.st);
        if (!r) r = h2d_staged(c, d + o_ww, ww, 8 * n, cg.st);
        if (!r) r = h2d_staged(c, d + o_vr, vis_real, 4 * nfp, cg.st);
        if (!r) r = h2d_staged(c, d + o_vi, vis_img, 4 * nfp, cg.st);
        if (!r) r = h2d_staged(c, d + o_wt, weight, 4 * n * polarizations, cg.st);
    }
    // the reference accumulates into the caller's grid (w-stacking_omp.c:135-138); a grid
    // of zeros (the calloc'd grid of init.c:369) needs no copy to the device: it is gridded
    // fresh, with the exact fixed-point tiles, and copied back
    const bool zero = host_all_zero(grid, cube_bytes / sizeof(double));
    if (!r && !zero) r = h2d_staged(c, d + o_cube, grid, cube_bytes, cg.st);
    if (r) {
        report_once("wstack", r);
        return;
    }
    r = grid_impl(c, LOFAR_FP64, zero ? 0 : 1, num_w_planes, num_points, freq_per_chan,
                  polarizations, (const double*)(d + o_uu), (const double*)(d + o_vv),
                  (const double*)(d + o_ww), (const float*)(d + o_vr), (const float*)(d + o_vi),
                  (const float*)(d + o_wt), dx, dw, w_support, grid_size_x, grid_size_y,
                  d + o_cube, cg.st);
    if (!r) r = d2h_staged(c, grid, d + o_cube, cube_bytes, cg.st);
    if (r) report_once("wstack", r);
}

void phase_correction(double* gridss, double* image_real, double* image_imag, int xaxis, int yaxis,
                      int num_w_planes, int xaxistot, int yaxistot, double resolution, double wmin,
                      double wmax, int num_threads)
{
    (void)num_threads;
    t_last_error = LOFAR_OK;
    int r = check_phase_args(LOFAR_FP64, xaxis, yaxis, num_w_planes, xaxistot, yaxistot);
    if (r || !gridss || !image_real || !image_imag) {
        fail(r ? r : LOFAR_EARG);
        report_once("phase_correction", t_last_error);
        return;
    }
    CallGuard cg(nullptr);
    if (cg.status) {
        report_once("phase_correction", cg.status);
        return;
    }
    DeviceCtx& c = *cg.c;
    cg.st = c.stream;
    (void)hipStreamWaitEvent(cg.st, c.last_done, 0);
    const size_t npix = (size_t)xaxis * yaxis;
    const size_t cube_bytes = (size_t)num_w_planes * npix * 16;
    const size_t o_re = align256(cube_bytes), o_im = align256(o_re + 8 * npix);
    r = ensure(c, c.host_io, o_im + 8 * npix);
    if (r) {
        report_once("phase_correction", r);
        return;
    }
    char* d = (char*)c.host_io.ptr;
    r = h2d_staged(c, d, gridss, cube_bytes, cg.st);
    if (!r) r = h2d_staged(c, d + o_re, image_real, 8 * npix, cg.st);
    if (!r) r = h2d_staged(c, d + o_im, image_imag, 8 * npix, cg.st);
    if (!r) {
        const hipError_t e = stage(c, kStPhase, cg.st, [&] {
            return phase_sum(LOFAR_FP64, d, (double*)(d + o_re), (double*)(d + o_im), xaxis, yaxis,
                             num_w_planes, xaxistot, yaxistot, resolution, wmin, wmax, 1, cg.st);
        });
        if (e != hipSuccess) r = fail(LOFAR_EHIP);
    }
    if (!r) r = d2h_staged(c, image_real, d + o_re, 8 * npix, cg.st);
    if (!r) r = d2h_staged(c, image_imag, d + o_im, 8 * npix, cg.st);
    if (r) report_once("phase_correction", r);
}

int lofar_fft_planes(const double* grid, double* gridss, int nx, int ny, int nw)
{
    t_last_error = LOFAR_OK;
    if (!grid || !gridss || nw < 1) return fail(LOFAR_EARG);
    if (!fft_size_ok(nx) || !fft_size_ok(ny)) return fail(LOFAR_ESIZE);
    CallGuard cg(nullptr);
    if (cg.status) return cg.status;
    DeviceCtx& c = *cg.c;
    cg.st = c.stream;
    (void)hipStreamWaitEvent(cg.st, c.last_done, 0);
    const size_t bytes = (size_t)nw * nx * ny * 16;
    int r = ensure(c, c.host_io, bytes);
    if (r) return r;
    r = h2d_staged(c, c.host_io.ptr, grid, bytes, cg.st);
    if (r) return r;
    r = fft_impl(c, LOFAR_FP64, 1, c.host_io.ptr, c.host_io.ptr, nx, ny, nw,
                 1.0 / ((double)nx * (double)ny), cg.st);
    if (r) return r;
    return d2h_staged(c, gridss, c.host_io.ptr, bytes, cg.st);
}

}  // extern "C"
