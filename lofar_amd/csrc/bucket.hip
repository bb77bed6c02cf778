// Bucketing of the samples into (w-plane, uv tile) buckets for gfx950: a
// two-level counting sort whose histograms live in LDS.
//
// Level 1 groups the samples by super-tile (sst x sst tiles of one w-plane):
// one record (position and channel/polarisation-summed visibility,
// w-stacking_omp.c:85-90,116-133; 32 bytes in FP64 mode, 16 in FP32 mode, Rec32)
// per super-tile the sample's window touches,
// which is one record for almost every sample.  Every workgroup takes a
// contiguous chunk of samples, counts its records (and the tile entries they
// will expand into) per super-tile in LDS histograms, reserves its range in
// each super-tile with one global atomic per non-empty bin, and after the scan
// places its records with LDS cursors.
//
// Level 2 splits every super-tile's records into blocks of at most l2_block
// records, so that the work spreads over many workgroups even when the w
// distribution piles samples into a few planes: a per-block LDS count of the
// entries per tile, a per-super-tile scan over its blocks (which also yields
// the bucket offsets), then a per-block scatter with LDS cursors that expands
// every record into one entry per tile of its window (w-stacking_omp.c:93-96).
// Entries are self-contained (no side arrays), so every scattered write is one
// aligned 16- or 32-byte store.
//
// Both scatters run staged where their bins fit one per thread (level 1 up to
// kStageMaxBins super-tiles, level 2 always): a round's records are counted, scanned
// and placed in an LDS buffer in bin order, then stored with consecutive lanes on
// consecutive slots, so each bin's records leave as a run instead of one lane per bin
// (the scattered 16-byte stores cost two 32-byte writes each).
#include <type_traits>

#include "grid_kernels.h"

namespace lofar {

constexpr int kScanBlock = 4096;  // elements per block of the local scan (256 threads x 16)
constexpr int kMaxFine = 256;     // sst <= 16
constexpr int kUnroll = 4;        // samples or records per thread per round

__global__ __launch_bounds__(256) void k_fill_u32(unsigned* p, long long n, unsigned v)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Tiles [tu_a, tu_b] x [tv_a, tv_b] of a non-empty window.
struct TileSpan {
    int tu_a, tu_b, tv_a, tv_b;
};
__device__ __forceinline__ TileSpan tiles_of(const Window& s)
{
    return TileSpan{s.j0 / kTile, s.j1 / kTile, s.k0 / kTile, s.k1 / kTile};
}

// ---------------------------------------------------------------------------
// level 1: count

// GLOBAL: more super-tiles than the LDS histograms hold; the counts go straight to
// count1/count1e with global atomics (no per-workgroup bases).
template <bool GLOBAL>
__global__ __launch_bounds__(256) void k_l1_count(int mode, long long n, const double* uu,
                                                  const double* vv, const double* ww, Geometry g,
                                                  TileMap tm, long long chunk, unsigned* count1,
                                                  unsigned* count1e, unsigned* base1, int* err)
{
    extern __shared__ unsigned hist_lds[];  // [nk1] records, then [nk1] entries
    unsigned* hist = GLOBAL ? count1 : hist_lds;
    unsigned* hent = GLOBAL ? count1e : hist_lds + tm.nk1;
    const int tid = threadIdx.x;
    if constexpr (!GLOBAL) {
        for (long long k = tid; k < 2 * tm.nk1; k += 256) hist[k] = 0u;
        __syncthreads();
    }
    const long long beg = (long long)blockIdx.x * chunk;
    const long long end = min(n, beg + chunk);
    bool bad = false;
    for (long long i0 = beg + tid; i0 < end; i0 += 256 * kUnroll) {
#pragma unroll
      for (int u = 0; u < kUnroll; u++) {
        const long long i = i0 + (long long)u * 256;
        if (i >= end) break;
        SampleWin s;
        if (!sample_window(uu[i], vv[i], ww[i], g, s)) {
            bad = true;
            continue;
        }
        if (mode == kBucketHome) {
            int cu, cv;
            home_cell(s, g, cu, cv);
            const long long k = tm.key1(s.gw, cv / kTile, cu / kTile);
            atomicAdd(&hist[k], 1u);
            atomicAdd(&hent[k], 1u);
            continue;
        }
        if (s.j0 > s.j1 || s.k0 > s.k1) continue;
        const TileSpan t = tiles_of(s);
        for (int sv = t.tv_a >> tm.lsst; sv <= t.tv_b >> tm.lsst; sv++) {
            const int nv = min(t.tv_b, (sv << tm.lsst) + tm.sst - 1) - max(t.tv_a, sv << tm.lsst) + 1;
            for (int su = t.tu_a >> tm.lsst; su <= t.tu_b >> tm.lsst; su++) {
                const int nu =
                    min(t.tu_b, (su << tm.lsst) + tm.sst - 1) - max(t.tu_a, su << tm.lsst) + 1;
                const long long k = tm.key_st(s.gw, sv, su);
                atomicAdd(&hist[k], 1u);
                atomicAdd(&hent[k], (unsigned)(nu * nv));
            }
        }
      }
    }
    if (bad) atomicOr(err, 1);
    if constexpr (GLOBAL) return;
    __syncthreads();
    unsigned* my_base = base1 + (long long)blockIdx.x * tm.nk1;
    for (long long k = tid; k < tm.nk1; k += 256) {
        const unsigned c = hist[k], ce = hent[k];
        my_base[k] = c ? atomicAdd(&count1[k], c) : 0u;
        if (ce) atomicAdd(&count1e[k], ce);
    }
}

// ---------------------------------------------------------------------------
// exclusive scan (counts -> offsets; the total to two places)

__global__ __launch_bounds__(256) void k_scan_local(const unsigned* in, long long* out,
                                                    long long* partials, long long n)
{
    __shared__ long long sh[256];
    const int tid = threadIdx.x;
    const long long base = (long long)blockIdx.x * kScanBlock + tid * 16;
    long long v[16];
    long long sum = 0;
#pragma unroll
    for (int e = 0; e < 16; e++) {
        v[e] = (base + e < n) ? (long long)in[base + e] : 0;
        sum += v[e];
    }
    sh[tid] = sum;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const long long x = tid >= off ? sh[tid - off] : 0;
        __syncthreads();
        sh[tid] += x;
        __syncthreads();
    }
    long long run = sh[tid] - sum;
#pragma unroll
    for (int e = 0; e < 16; e++) {
        if (base + e < n) out[base + e] = run;
        run += v[e];
    }
    if (tid == 255) partials[blockIdx.x] = sh[255];
}

// single workgroup: exclusive scan of the block partials; writes the grand total
__global__ __launch_bounds__(1024) void k_scan_partials(long long* partials, long long nblk,
                                                        long long* total_a, long long* total_b)
{
    __shared__ long long sh[1024];
    const int tid = threadIdx.x;
    long long carry = 0;
    for (long long c0 = 0; c0 < nblk; c0 += 1024) {
        const long long idx = c0 + tid;
        const long long x = idx < nblk ? partials[idx] : 0;
        sh[tid] = x;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const long long y = tid >= off ? sh[tid - off] : 0;
            __syncthreads();
            sh[tid] += y;
            __syncthreads();
        }
        if (idx < nblk) partials[idx] = carry + sh[tid] - x;
        const long long chunk = sh[1023];
        __syncthreads();
        carry += chunk;
    }
    if (tid == 0) {
        *total_a = carry;
        *total_b = carry;
    }
}

__global__ __launch_bounds__(256) void k_scan_add(long long* out, const long long* partials,
                                                  long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += partials[i / kScanBlock];
}

// ---------------------------------------------------------------------------
// level 1: scatter records into super-tile order

// The record of sample i (position e.pu, e.pv and summed visibility e.vr, e.vi) in
// super-tile (sv, su); IDX: the sample's index in the real part's slot instead.
template <typename Rec, bool IDX>
__device__ __forceinline__ Rec make_record(const Entry64& e, const TileMap& tm, int sv, int su,
                                           long long i)
{
    if constexpr (sizeof(Rec) == sizeof(Rec32)) {
        Rec32 r;
        r.u = rec_axis(e.pu, (su << tm.lsst) * kTile, tm.rfb);
        r.v = rec_axis(e.pv, (sv << tm.lsst) * kTile, tm.rfb);
        r.vr = IDX ? __uint_as_float((unsigned)i) : (float)e.vr;
        r.vi = (float)e.vi;
        return r;
    } else {
        Entry64 r = e;
        if constexpr (IDX) r.vr = __longlong_as_double(i);
        return r;
    }
}

// Rec: Entry64 (FP64 mode) or Rec32 (FP32 mode, position relative to the
// super-tile, rec_axis)
// GLOBAL: base1 is one global cursor per super-tile (zeroed), taken with global atomics.
// IDX: the records carry the sample's index (bits of the real part's slot, entry_index)
// instead of its visibility, for a sort kept across calls (plan_fill).
template <typename Rec, bool GLOBAL, bool IDX>
__global__ __launch_bounds__(256) void k_l1_scatter_grid(long long n, long long F, long long P,
                                                         const double* uu, const double* vv,
                                                         const double* ww, const float* vr,
                                                         const float* vi, const float* wt,
                                                         Geometry g, TileMap tm, long long chunk,
                                                         const long long* off1,
                                                         unsigned* base1, Rec* rec)
{
    extern __shared__ unsigned cur_lds[];
    unsigned* cur = GLOBAL ? base1 : cur_lds;
    const int tid = threadIdx.x;
    if constexpr (!GLOBAL) {
        const unsigned* my_base = base1 + (long long)blockIdx.x * tm.nk1;
        for (long long k = tid; k < tm.nk1; k += 256) cur[k] = my_base[k];
        __syncthreads();
    }
    const long long beg = (long long)blockIdx.x * chunk;
    const long long end = min(n, beg + chunk);
    // kUnroll samples per thread per round, their loads issued together
    for (long long i0 = beg + tid; i0 < end; i0 += 256 * kUnroll) {
        SampleWin s[kUnroll];
        Entry64 e[kUnroll];
        bool ok[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            const long long i = min(i0 + (long long)u * 256, end - 1);
            ok[u] = i0 + (long long)u * 256 < end && sample_window(uu[i], vv[i], ww[i], g, s[u]) &&
                    s[u].j0 <= s[u].j1 && s[u].k0 <= s[u].k1;
            if constexpr (IDX) {
                e[u].vr = 0.0;
                e[u].vi = 0.0;
            } else {
                sample_vis(i, F, P, vr, vi, wt, e[u].vr, e[u].vi);
            }
            e[u].pu = s[u].pu;
            e[u].pv = s[u].pv;
        }
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            if (!ok[u]) continue;
            const TileSpan t = tiles_of(s[u]);
            for (int sv = t.tv_a >> tm.lsst; sv <= t.tv_b >> tm.lsst; sv++)
                for (int su = t.tu_a >> tm.lsst; su <= t.tu_b >> tm.lsst; su++) {
                    const long long k = tm.key_st(s[u].gw, sv, su);
                    const long long at = off1[k] + atomicAdd(&cur[k], 1u);
                    rec[at] = make_record<Rec, IDX>(e[u], tm, sv, su, i0 + u * 256);
                }
        }
    }
}

// ---------------------------------------------------------------------------
// staged scatters: a 512-thread workgroup takes its items in rounds, counts the
// round's records per bin in LDS, scans, places them in an LDS buffer in bin order
// and then stores the buffer with consecutive lanes on consecutive slots of a bin,
// so that a bin's records leave in runs instead of one lane per bin.  A round's
// records beyond the buffer are stored directly (same slots).  One bin per thread.

constexpr int kStageThreads = 512;  // two workgroups per CU overlap their rounds' phases
constexpr int kStageBytes = 65536;  // LDS buffer of one round's entries (level 2)
static_assert(kStageMaxBins <= kStageThreads && kStageMaxBins <= 65536, "one bin per thread");

// Level 1: buffer capacity (records) and samples per thread and round, so that
// buffer, keys and bins stay within half a CU's 160 KiB of LDS (two workgroups per
// CU) and a round's records (about 1.02 per sample on 16x16-tile super-tiles)
// mostly fit.
template <typename Rec>
struct L1Stage {
    static constexpr int CAP = sizeof(Rec) == sizeof(Rec32) ? 4032 : 2144;
    static constexpr int U = 4;
    static_assert(CAP * (sizeof(Rec) + 2) + kStageMaxBins * 16 + 64 <= 81920, "LDS");
};

// Exclusive scan of one value per thread over the workgroup; the sum to total.
// Ends with every wave past the barrier that publishes the wave sums.
template <int NT>
__device__ __forceinline__ unsigned block_exscan(unsigned v, unsigned* wsum, unsigned& total)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    unsigned x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    unsigned pre = 0u, tot = 0u;
#pragma unroll
    for (int w = 0; w < NT / 64; w++) {
        const unsigned s = wsum[w];
        pre += w < wid ? s : 0u;
        tot += s;
    }
    total = tot;
    return pre + x - v;
}

// Per-round bin bookkeeping of a staged scatter.  gnext (thread tid's register, bin
// tid): global slot of the bin's next record.  After open(): dst[k] + p is the
// global slot of staged position p of bin k, pos[k] the bin's first staged position.
template <int NT>
struct StageBins {
    long long* dst;
    unsigned *cnt, *pos, *wsum;
    // after the counting pass (and a barrier): the round's layout; zeroes the counts
    __device__ __forceinline__ unsigned open(long long nbins, long long& gnext)
    {
        const int tid = threadIdx.x;
        const unsigned c = tid < nbins ? cnt[tid] : 0u;
        unsigned total;
        const unsigned ex = block_exscan<NT>(c, wsum, total);
        if (tid < nbins) {
            dst[tid] = gnext - (long long)ex;
            pos[tid] = ex;
            cnt[tid] = 0u;
            gnext += c;
        }
        __syncthreads();
        return total;
    }
};

template <typename Rec, bool IDX>
__global__ __launch_bounds__(kStageThreads) __attribute__((amdgpu_waves_per_eu(4)))
void k_l1_scatter_staged(
    long long n, long long F, long long P, const double* uu, const double* vv, const double* ww,
    const float* vr, const float* vi, const float* wt, Geometry g, TileMap tm, long long chunk,
    const long long* off1, const unsigned* base1, Rec* rec)
{
    constexpr int NT = kStageThreads;
    constexpr int CAP = L1Stage<Rec>::CAP;
    constexpr int U = L1Stage<Rec>::U;  // samples per thread and round
    // visibility sum as the record keeps it (float for Rec32: the same bits as rounding later)
    using V = typename std::conditional<sizeof(Rec) == sizeof(Rec32), float, double>::type;
    __shared__ Rec stage[CAP];
    __shared__ unsigned short skey[CAP];
    __shared__ long long dst[kStageMaxBins];
    __shared__ unsigned cnt[kStageMaxBins], pos[kStageMaxBins], wsum[NT / 64];
    StageBins<NT> sb{dst, cnt, pos, wsum};
    const int tid = threadIdx.x;
    long long gnext = 0;
    if (tid < tm.nk1) {
        gnext = off1[tid] + base1[(long long)blockIdx.x * tm.nk1 + tid];
        cnt[tid] = 0u;
    }
    __syncthreads();
    const long long beg = (long long)blockIdx.x * chunk;
    const long long end = min(n, beg + chunk);
    for (long long r0 = beg; r0 < end; r0 += (long long)NT * U) {
        // per sample only the position, plane (-1: no record) and visibility; the
        // window is re-derived from the position in each pass
        double pu[U], pv[U];
        int gw[U];
        V sr[U], si[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const long long i0 = r0 + (long long)u * NT + tid;
            const long long i = min(i0, end - 1);
            SampleWin s;
            const bool ok = i0 < end && sample_window(uu[i], vv[i], ww[i], g, s) &&
                            s.j0 <= s.j1 && s.k0 <= s.k1;
            pu[u] = ok ? s.pu : 0.0;
            pv[u] = ok ? s.pv : 0.0;
            gw[u] = ok ? s.gw : -1;
            if constexpr (IDX) {
                sr[u] = si[u] = (V)0;
            } else {
                double a, b;
                sample_vis(i, F, P, vr, vi, wt, a, b);
                sr[u] = (V)a;
                si[u] = (V)b;
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (gw[u] < 0) continue;
            const TileSpan t = tiles_of(window_of(pu[u], pv[u], g));
            for (int sv = t.tv_a >> tm.lsst; sv <= t.tv_b >> tm.lsst; sv++)
                for (int su = t.tu_a >> tm.lsst; su <= t.tu_b >> tm.lsst; su++)
                    atomicAdd(&cnt[tm.key_st(gw[u], sv, su)], 1u);
        }
        __syncthreads();
        const unsigned total = sb.open(tm.nk1, gnext);
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (gw[u] < 0) continue;
            const TileSpan t = tiles_of(window_of(pu[u], pv[u], g));
            const Entry64 e{pu[u], pv[u], (double)sr[u], (double)si[u]};
            for (int sv = t.tv_a >> tm.lsst; sv <= t.tv_b >> tm.lsst; sv++)
                for (int su = t.tu_a >> tm.lsst; su <= t.tu_b >> tm.lsst; su++) {
                    const int k = (int)tm.key_st(gw[u], sv, su);
                    const unsigned p = atomicAdd(&pos[k], 1u);
                    const Rec r = make_record<Rec, IDX>(e, tm, sv, su,
                                                        r0 + (long long)u * NT + tid);
                    if (p < (unsigned)CAP) {
                        stage[p] = r;
                        skey[p] = (unsigned short)k;
                    } else {
                        rec[dst[k] + p] = r;
                    }
                }
        }
        __syncthreads();
        const unsigned m = min(total, (unsigned)CAP);
        for (unsigned j = tid; j < m; j += NT) rec[dst[skey[j]] + j] = stage[j];
        __syncthreads();
    }
}

// wt1: the samples' weights when there is one polarisation (each entry then carries its
// sample's weight, so the degridder reads no weight array in the caller's order), else null.
template <bool GLOBAL>
__global__ __launch_bounds__(256) void k_l1_scatter_home(long long n, const double* uu,
                                                         const double* vv, const double* ww,
                                                         const float* wt1, Geometry g,
                                                         TileMap tm, long long chunk,
                                                         const long long* off1,
                                                         unsigned* base1, DEntry* rec)
{
    extern __shared__ unsigned cur_lds[];
    unsigned* cur = GLOBAL ? base1 : cur_lds;
    const int tid = threadIdx.x;
    if constexpr (!GLOBAL) {
        const unsigned* my_base = base1 + (long long)blockIdx.x * tm.nk1;
        for (long long k = tid; k < tm.nk1; k += 256) cur[k] = my_base[k];
        __syncthreads();
    }
    const long long beg = (long long)blockIdx.x * chunk;
    const long long end = min(n, beg + chunk);
    for (long long i = beg + tid; i < end; i += 256) {
        SampleWin s;
        if (!sample_window(uu[i], vv[i], ww[i], g, s)) continue;
        int cu, cv;
        home_cell(s, g, cu, cv);
        const long long k = tm.key1(s.gw, cv / kTile, cu / kTile);
        DEntry e;
        e.pu = s.pu;
        e.pv = s.pv;
        e.idx = i;
        e.pad = wt1 ? (long long)__float_as_uint(wt1[i]) : 0;
        rec[off1[k] + atomicAdd(&cur[k], 1u)] = e;
    }
}

// Staged form of k_l1_scatter_home (see k_l1_scatter_staged): one record per valid
// sample, in its home tile's super-tile.
__global__ __launch_bounds__(kStageThreads) __attribute__((amdgpu_waves_per_eu(4)))
void k_l1_scatter_home_staged(long long n, const double* uu, const double* vv,
                              const double* ww, const float* wt1, Geometry g, TileMap tm,
                              long long chunk, const long long* off1, const unsigned* base1,
                              DEntry* rec)
{
    constexpr int NT = kStageThreads;
    constexpr int CAP = L1Stage<DEntry>::CAP;
    constexpr int U = L1Stage<DEntry>::U;
    static_assert(U * NT <= CAP, "one record per sample: a round always fits the buffer");
    __shared__ DEntry stage[CAP];
    __shared__ unsigned short skey[CAP];
    __shared__ long long dst[kStageMaxBins];
    __shared__ unsigned cnt[kStageMaxBins], pos[kStageMaxBins], wsum[NT / 64];
    StageBins<NT> sb{dst, cnt, pos, wsum};
    const int tid = threadIdx.x;
    long long gnext = 0;
    if (tid < tm.nk1) {
        gnext = off1[tid] + base1[(long long)blockIdx.x * tm.nk1 + tid];
        cnt[tid] = 0u;
    }
    __syncthreads();
    const long long beg = (long long)blockIdx.x * chunk;
    const long long end = min(n, beg + chunk);
    for (long long r0 = beg; r0 < end; r0 += (long long)NT * U) {
        DEntry e[U];
        int key[U];  // -1: no record
#pragma unroll
        for (int u = 0; u < U; u++) {
            const long long i0 = r0 + (long long)u * NT + tid;
            const long long i = min(i0, end - 1);
            SampleWin s;
            key[u] = -1;
            if (i0 < end && sample_window(uu[i], vv[i], ww[i], g, s)) {
                int cu, cv;
                home_cell(s, g, cu, cv);
                key[u] = (int)tm.key1(s.gw, cv / kTile, cu / kTile);
                e[u].pu = s.pu;
                e[u].pv = s.pv;
                e[u].idx = i;
                e[u].pad = wt1 ? (long long)__float_as_uint(wt1[i]) : 0;
                atomicAdd(&cnt[key[u]], 1u);
            }
        }
        __syncthreads();
        const unsigned total = sb.open(tm.nk1, gnext);
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (key[u] < 0) continue;
            const unsigned p = atomicAdd(&pos[key[u]], 1u);
            stage[p] = e[u];
            skey[p] = (unsigned short)key[u];
        }
        __syncthreads();
        for (unsigned j = tid; j < total; j += NT) rec[dst[skey[j]] + j] = stage[j];
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// level 2

enum { kOutFp64 = 0, kOutFp32 = 1, kOutHome = 2 };

// Tiles of the record's window inside super-tile (sv, su), as inclusive ranges
// of fine coordinates; for the home mode the single home tile.
template <typename Rec>
__device__ __forceinline__ void record_pos(const Rec& r, const TileMap&, int, int, double& pu,
                                           double& pv)
{
    pu = r.pu;
    pv = r.pv;
}
__device__ __forceinline__ void record_pos(const Rec32& r, const TileMap& tm, int sv, int su,
                                           double& pu, double& pv)
{
    pu = rec_pos(r.u, (su << tm.lsst) * kTile, tm.rfb);
    pv = rec_pos(r.v, (sv << tm.lsst) * kTile, tm.rfb);
}

template <int OUT, typename Rec>
__device__ __forceinline__ bool record_tiles(const Rec& r, const Geometry& g, const TileMap& tm,
                                             int sv, int su, int& fu_a, int& fu_b, int& fv_a,
                                             int& fv_b, Window& s)
{
    double pu, pv;
    record_pos(r, tm, sv, su, pu, pv);
    s = window_of(pu, pv, g);
    if constexpr (OUT == kOutHome) {
        int cu, cv;
        home_cell(s, g, cu, cv);
        fu_a = fu_b = (cu / kTile) - (su << tm.lsst);
        fv_a = fv_b = (cv / kTile) - (sv << tm.lsst);
        return true;
    } else {
        const TileSpan t = tiles_of(s);
        fu_a = max(t.tu_a - (su << tm.lsst), 0);
        fu_b = min(t.tu_b - (su << tm.lsst), tm.sst - 1);
        fv_a = max(t.tv_a - (sv << tm.lsst), 0);
        fv_b = min(t.tv_b - (sv << tm.lsst), tm.sst - 1);
        return fu_a <= fu_b && fv_a <= fv_b;
    }
}

// blocks per segment
__global__ __launch_bounds__(256) void k_l2_nblocks(long long nk1, const long long* off1,
                                                    long long bs, unsigned* nblk)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nk1) nblk[k] = (unsigned)((off1[k + 1] - off1[k] + bs - 1) / bs);
}

// block descriptors: segment, first record, end record
__global__ __launch_bounds__(256) void k_l2_blocks(long long nk1, const long long* off1,
                                                   long long bs, const long long* boff,
                                                   L2Block* blk)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk1) return;
    const long long beg = off1[k], end = off1[k + 1];
    long long j = boff[k];
    for (long long r = beg; r < end; r += bs, j++) {
        L2Block d;
        d.k1 = k;
        d.beg = r;
        d.end = min(end, r + bs);
        d.pad = 0;
        blk[j] = d;
    }
}

template <int OUT, typename Rec>
__global__ __launch_bounds__(256) void k_l2_count(Geometry g, TileMap tm, const long long* nblocks,
                                                  const L2Block* blk, const Rec* rec,
                                                  unsigned* cnt2)
{
    __shared__ unsigned cnt[kMaxFine];
    const long long b = blockIdx.x;
    if (b >= *nblocks) return;  // the grid is sized from an upper bound
    const int tid = threadIdx.x;
    const int nf = tm.sst * tm.sst;
    for (int f = tid; f < nf; f += 256) cnt[f] = 0u;
    __syncthreads();
    const L2Block d = blk[b];
    int sv, su;
    tm.super_of(d.k1, sv, su);
    for (long long r0 = d.beg + tid; r0 < d.end; r0 += 256 * kUnroll) {
        Rec xs[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) xs[u] = rec[min(r0 + (long long)u * 256, d.end - 1)];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            if (r0 + (long long)u * 256 >= d.end) break;
            int fu_a, fu_b, fv_a, fv_b;
            Window s;
            if (!record_tiles<OUT>(xs[u], g, tm, sv, su, fu_a, fu_b, fv_a, fv_b, s)) continue;
            for (int fv = fv_a; fv <= fv_b; fv++)
                for (int fu = fu_a; fu <= fu_b; fu++) atomicAdd(&cnt[(fv << tm.lsst) | fu], 1u);
        }
    }
    __syncthreads();
    for (int f = tid; f < nf; f += 256) cnt2[b * nf + f] = cnt[f];
}

// one workgroup per segment: per fine tile, running sums over the segment's
// blocks (replacing the counts by block bases relative to the segment's first
// entry), then the exclusive scan over fine tiles gives the bucket offsets
__global__ __launch_bounds__(256) void k_l2_scan(TileMap tm, const long long* eoff1,
                                                 const long long* boff, unsigned* cnt2,
                                                 long long* offsets)
{
    __shared__ unsigned tot[kMaxFine];
    const int tid = threadIdx.x;
    const int nf = tm.sst * tm.sst;
    const long long k1 = blockIdx.x;
    const long long b0 = boff[k1], b1 = boff[k1 + 1];
    for (int f = tid; f < nf; f += 256) {
        unsigned run = 0;
        for (long long j = b0; j < b1; j++) {
            const unsigned c = cnt2[j * nf + f];
            cnt2[j * nf + f] = run;
            run += c;
        }
        tot[f] = run;
    }
    __syncthreads();
    if (tid == 0) {
        unsigned run = 0;
        for (int f = 0; f < nf; f++) {
            const unsigned c = tot[f];
            tot[f] = run;
            run += c;
        }
    }
    __syncthreads();
    const long long beg = eoff1[k1];
    for (int f = tid; f < nf; f += 256) {
        offsets[k1 * nf + f] = beg + tot[f];
        for (long long j = b0; j < b1; j++) cnt2[j * nf + f] += tot[f];
    }
    if (k1 == 0 && tid == 0) offsets[tm.nk1 * nf] = eoff1[tm.nk1];
}

// The entry of record x (window s) for fine tile (fv, fu) of super-tile (sv, su).
template <int OUT, typename Rec, typename Out>
__device__ __forceinline__ Out make_entry(const Rec& x, const Window& s, const TileMap& tm, int sv,
                                          int su, int fv, int fu)
{
    if constexpr (OUT == kOutFp32) {
        const int u0 = ((su << tm.lsst) + fu) * kTile;
        const int v0 = ((sv << tm.lsst) + fv) * kTile;
        double pu, pv;
        record_pos(x, tm, sv, su, pu, pv);
        Entry32 e;
        e.wu = pack_cell(s.gu - u0, pu - (double)s.gu);
        e.wv = pack_cell(s.gv - v0, pv - (double)s.gv);
        e.vr = x.vr;  // bit for bit (a plan's index, see k_l1_scatter_grid)
        e.vi = x.vi;
        return e;
    } else {
        return x;
    }
}

template <int OUT, typename Rec, typename Out>
__global__ __launch_bounds__(256) void k_l2_scatter(Geometry g, TileMap tm, const long long* eoff1,
                                                    const long long* nblocks, const L2Block* blk,
                                                    const unsigned* cnt2, const Rec* rec,
                                                    Out* ent)
{
    __shared__ unsigned cur[kMaxFine];
    const long long b = blockIdx.x;
    if (b >= *nblocks) return;
    const int tid = threadIdx.x;
    const int nf = tm.sst * tm.sst;
    for (int f = tid; f < nf; f += 256) cur[f] = cnt2[b * nf + f];
    __syncthreads();
    const L2Block d = blk[b];
    int sv, su;
    tm.super_of(d.k1, sv, su);
    const long long beg = eoff1[d.k1];
    for (long long r0 = d.beg + tid; r0 < d.end; r0 += 256 * kUnroll) {
      Rec xs[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; u++) xs[u] = rec[min(r0 + (long long)u * 256, d.end - 1)];
#pragma unroll
      for (int u = 0; u < kUnroll; u++) {
        if (r0 + (long long)u * 256 >= d.end) break;
        const Rec& x = xs[u];
        int fu_a, fu_b, fv_a, fv_b;
        Window s;
        if (!record_tiles<OUT>(x, g, tm, sv, su, fu_a, fu_b, fv_a, fv_b, s)) continue;
        for (int fv = fv_a; fv <= fv_b; fv++) {
            for (int fu = fu_a; fu <= fu_b; fu++) {
                const long long p = beg + atomicAdd(&cur[(fv << tm.lsst) | fu], 1u);
                ent[p] = make_entry<OUT, Rec, Out>(x, s, tm, sv, su, fv, fu);
            }
        }
      }
    }
}

// Level-2 scatter of one block, staged (see k_l1_scatter_staged): rounds of U records
// per thread, the round's entries placed in LDS in fine-tile order, then stored in runs.
template <int OUT, typename Rec, typename Out>
__global__ __launch_bounds__(kStageThreads) void k_l2_scatter_staged(
    Geometry g, TileMap tm, const long long* eoff1, const long long* nblocks, const L2Block* blk,
    const unsigned* cnt2, const Rec* rec, Out* ent)
{
    constexpr int NT = kStageThreads;
    constexpr int CAP = kStageBytes / (int)sizeof(Out);
    constexpr int U = sizeof(Out) == sizeof(Entry32) ? 4 : 2;  // records per thread and round
    __shared__ Out stage[CAP];
    __shared__ unsigned char skey[CAP];
    __shared__ long long dst[kMaxFine];
    __shared__ unsigned cnt[kMaxFine], pos[kMaxFine], wsum[NT / 64];
    StageBins<NT> sb{dst, cnt, pos, wsum};
    const long long b = blockIdx.x;
    if (b >= *nblocks) return;  // the grid is sized from an upper bound
    const int tid = threadIdx.x;
    const int nf = tm.sst * tm.sst;
    const L2Block d = blk[b];
    int sv, su;
    tm.super_of(d.k1, sv, su);
    long long gnext = 0;
    if (tid < nf) {
        gnext = eoff1[d.k1] + cnt2[b * nf + tid];
        cnt[tid] = 0u;
    }
    __syncthreads();
    for (long long r0 = d.beg; r0 < d.end; r0 += (long long)NT * U) {
        Rec xs[U];
        Window s[U];
        int fu_a[U], fu_b[U], fv_a[U], fv_b[U];
        bool ok[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const long long r = r0 + (long long)u * NT + tid;
            xs[u] = rec[min(r, d.end - 1)];
            ok[u] = r < d.end && record_tiles<OUT>(xs[u], g, tm, sv, su, fu_a[u], fu_b[u],
                                                   fv_a[u], fv_b[u], s[u]);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (!ok[u]) continue;
            for (int fv = fv_a[u]; fv <= fv_b[u]; fv++)
                for (int fu = fu_a[u]; fu <= fu_b[u]; fu++)
                    atomicAdd(&cnt[(fv << tm.lsst) | fu], 1u);
        }
        __syncthreads();
        const unsigned total = sb.open(nf, gnext);
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (!ok[u]) continue;
            for (int fv = fv_a[u]; fv <= fv_b[u]; fv++)
                for (int fu = fu_a[u]; fu <= fu_b[u]; fu++) {
                    const int f = (fv << tm.lsst) | fu;
                    const unsigned p = atomicAdd(&pos[f], 1u);
                    const Out e = make_entry<OUT, Rec, Out>(xs[u], s[u], tm, sv, su, fv, fu);
                    if (p < (unsigned)CAP) {
                        stage[p] = e;
                        skey[p] = (unsigned char)f;
                    } else {
                        ent[dst[f] + p] = e;
                    }
                }
        }
        __syncthreads();
        const unsigned m = min(total, (unsigned)CAP);
        for (unsigned j = tid; j < m; j += NT) ent[dst[skey[j]] + j] = stage[j];
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// host dispatch

static inline unsigned blocks_for(long long n, int t) { return (unsigned)((n + t - 1) / t); }

static void scan(const unsigned* in, long long* out, long long n, long long* partials,
                 long long* total_b, hipStream_t st)
{
    const long long nblk = (n + kScanBlock - 1) / kScanBlock;
    hipLaunchKernelGGL(k_scan_local, dim3((unsigned)nblk), dim3(256), 0, st, in, out, partials, n);
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, st, partials, nblk, out + n,
                       total_b);
    hipLaunchKernelGGL(k_scan_add, dim3(blocks_for(n, 256)), dim3(256), 0, st, out, partials, n);
}

hipError_t fill_u32(unsigned* p, long long n, unsigned v, hipStream_t st)
{
    if (n > 0) hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(n, 256)), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t bucket_count(int mode, long long n, const double* uu, const double* vv, const double* ww,
                        const Geometry& g, const Buckets& b, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(b.map.nk1, 256)), dim3(256), 0, st, b.count1,
                       b.map.nk1, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(b.map.nk1, 256)), dim3(256), 0, st, b.count1e,
                       b.map.nk1, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, st, (unsigned*)b.err, 1LL, 0u);
    if (n > 0 && b.global_l1)
        hipLaunchKernelGGL(k_l1_count<true>, dim3((unsigned)b.nwg), dim3(256), 0, st, mode, n, uu,
                           vv, ww, g, b.map, b.chunk, b.count1, b.count1e, b.base1, b.err);
    else if (n > 0)
        hipLaunchKernelGGL(k_l1_count<false>, dim3((unsigned)b.nwg), dim3(256),
                           2 * (size_t)b.map.nk1 * sizeof(unsigned), st, mode, n, uu, vv, ww, g,
                           b.map, b.chunk, b.count1, b.count1e, b.base1, b.err);
    return hipGetLastError();
}

hipError_t bucket_scan(const Buckets& b, hipStream_t st)
{
    scan(b.count1, b.off1, b.map.nk1, b.partials, b.total_rec, st);
    scan(b.count1e, b.eoff1, b.map.nk1, b.partials, b.total, st);
    return hipGetLastError();
}

template <int OUT, typename Rec, typename Out>
static hipError_t level2(const Geometry& g, const Buckets& b, const Rec* rec, Out* ent,
                         hipStream_t st)
{
    const long long nk1 = b.map.nk1;
    hipLaunchKernelGGL(k_l2_nblocks, dim3(blocks_for(nk1, 256)), dim3(256), 0, st, nk1, b.off1,
                       b.l2_block, b.l2_nblk);
    scan(b.l2_nblk, b.l2_boff, nk1, b.partials, b.l2_boff + nk1, st);
    hipLaunchKernelGGL(k_l2_blocks, dim3(blocks_for(nk1, 256)), dim3(256), 0, st, nk1, b.off1,
                       b.l2_block, b.l2_boff, b.l2_blk);
    const unsigned grid = (unsigned)b.l2_max_blocks;
    const long long* nblocks = b.l2_boff + nk1;
    hipLaunchKernelGGL((k_l2_count<OUT, Rec>), dim3(grid), dim3(256), 0, st, g, b.map, nblocks,
                       b.l2_blk, rec, b.l2_cnt);
    hipLaunchKernelGGL(k_l2_scan, dim3((unsigned)nk1), dim3(256), 0, st, b.map, b.eoff1, b.l2_boff,
                       b.l2_cnt, b.offsets);
    if (b.stage & kStageL2)
        hipLaunchKernelGGL((k_l2_scatter_staged<OUT, Rec, Out>), dim3(grid), dim3(kStageThreads),
                           0, st, g, b.map, b.eoff1, nblocks, b.l2_blk, b.l2_cnt, rec, ent);
    else
        hipLaunchKernelGGL((k_l2_scatter<OUT, Rec, Out>), dim3(grid), dim3(256), 0, st, g, b.map,
                           b.eoff1, nblocks, b.l2_blk, b.l2_cnt, rec, ent);
    return hipGetLastError();
}

template <typename Rec>
static void launch_l1_scatter_grid(bool global_l1, bool index_vis, size_t lds, hipStream_t st,
                                   long long n, long long F, long long P, const double* uu,
                                   const double* vv, const double* ww, const float* vr,
                                   const float* vi, const float* wt, const Geometry& g,
                                   const Buckets& b, Rec* rec)
{
    if ((b.stage & kStageL1) && !global_l1) {
        auto k = index_vis ? k_l1_scatter_staged<Rec, true> : k_l1_scatter_staged<Rec, false>;
        hipLaunchKernelGGL(k, dim3((unsigned)b.nwg), dim3(kStageThreads), 0, st, n, F, P, uu, vv,
                           ww, vr, vi, wt, g, b.map, b.chunk, b.off1, b.base1, rec);
        return;
    }
    auto k = global_l1 ? (index_vis ? k_l1_scatter_grid<Rec, true, true>
                                    : k_l1_scatter_grid<Rec, true, false>)
                       : (index_vis ? k_l1_scatter_grid<Rec, false, true>
                                    : k_l1_scatter_grid<Rec, false, false>);
    hipLaunchKernelGGL(k, dim3((unsigned)b.nwg), dim3(256), lds, st, n, F, P, uu, vv, ww, vr, vi,
                       wt, g, b.map, b.chunk, b.off1, b.base1, rec);
}

hipError_t bucket_sort_grid(int precision, long long n, long long F, long long P,
                            const double* uu, const double* vv, const double* ww, const float* vr,
                            const float* vi, const float* wt, const Geometry& g, const Buckets& b,
                            Entry64* rec, void* entries, hipStream_t st, bool index_vis)
{
    const size_t lds = b.global_l1 ? 0 : (size_t)b.map.nk1 * sizeof(unsigned);
    if (n > 0 && b.global_l1) {
        hipError_t e = fill_u32(b.base1, b.map.nk1, 0u, st);
        if (e != hipSuccess) return e;
    }
    if (precision == LOFAR_FP64) {
        if (n > 0)
            launch_l1_scatter_grid<Entry64>(b.global_l1, index_vis, lds, st, n, F, P, uu, vv, ww,
                                            vr, vi, wt, g, b, rec);
        return level2<kOutFp64>(g, b, (const Entry64*)rec, (Entry64*)entries, st);
    }
    Rec32* rec32 = (Rec32*)rec;  // the record buffer holds total_rec 32-byte records
    if (n > 0)
        launch_l1_scatter_grid<Rec32>(b.global_l1, index_vis, lds, st, n, F, P, uu, vv, ww, vr,
                                      vi, wt, g, b, rec32);
    return level2<kOutFp32>(g, b, (const Rec32*)rec32, (Entry32*)entries, st);
}

// A plan's entries, with this call's visibilities, in two passes: every sample's
// channel/polarisation-summed weighted visibility as k_l1_scatter_grid forms it (in the
// samples' order, coalesced; rounded to float for Entry32, as the entries hold it), then
// each entry's sample index (left in its real part's slot by the sort) replaced by its
// sample's sum, one 8- or 16-byte gather per entry.
template <typename V2>
__global__ __launch_bounds__(256) void k_plan_vis(long long n, long long F, long long P,
                                                  const float* vr, const float* vi,
                                                  const float* wt, V2* vis)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double sr, si;
    sample_vis(i, F, P, vr, vi, wt, sr, si);
    V2 x;
    x.x = (decltype(x.x))sr;
    x.y = (decltype(x.y))si;
    vis[i] = x;
}

template <typename Ent, typename V2>
__global__ __launch_bounds__(256) void k_plan_fill(long long total, const Ent* pent, Ent* ent,
                                                   const V2* vis)
{
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    Ent x = pent[e];
    long long i;
    if constexpr (sizeof(Ent) == sizeof(Entry64))
        i = __double_as_longlong(x.vr);
    else
        i = (long long)__float_as_uint(x.vr);
    const V2 v = vis[i];
    x.vr = v.x;
    x.vi = v.y;
    ent[e] = x;
}

size_t plan_vis_bytes(int precision, long long n)
{
    return (size_t)n * (precision == LOFAR_FP64 ? sizeof(double2) : sizeof(float2));
}

hipError_t plan_vis(int precision, long long n, long long F, long long P, const float* vr,
                    const float* vi, const float* wt, void* vis, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    const dim3 gs(blocks_for(n, 256));
    if (precision == LOFAR_FP64)
        hipLaunchKernelGGL((k_plan_vis<double2>), gs, dim3(256), 0, st, n, F, P, vr, vi, wt,
                           (double2*)vis);
    else
        hipLaunchKernelGGL((k_plan_vis<float2>), gs, dim3(256), 0, st, n, F, P, vr, vi, wt,
                           (float2*)vis);
    return hipGetLastError();
}

hipError_t plan_fill(int precision, long long n, long long total, const void* pent, void* ent,
                     long long F, long long P, const float* vr, const float* vi, const float* wt,
                     void* vis, hipStream_t st)
{
    if (total <= 0) return hipSuccess;
    const dim3 gs(blocks_for(n, 256)), grid(blocks_for(total, 256));
    if (precision == LOFAR_FP64) {
        hipLaunchKernelGGL((k_plan_vis<double2>), gs, dim3(256), 0, st, n, F, P, vr, vi, wt,
                           (double2*)vis);
        hipLaunchKernelGGL((k_plan_fill<Entry64, double2>), grid, dim3(256), 0, st, total,
                           (const Entry64*)pent, (Entry64*)ent, (const double2*)vis);
    } else {
        hipLaunchKernelGGL((k_plan_vis<float2>), gs, dim3(256), 0, st, n, F, P, vr, vi, wt,
                           (float2*)vis);
        hipLaunchKernelGGL((k_plan_fill<Entry32, float2>), grid, dim3(256), 0, st, total,
                           (const Entry32*)pent, (Entry32*)ent, (const float2*)vis);
    }
    return hipGetLastError();
}

hipError_t bucket_sort_home(long long n, const double* uu, const double* vv, const double* ww,
                            const float* wt1, const Geometry& g, const Buckets& b, DEntry* rec,
                            DEntry* entries, hipStream_t st)
{
    if (n > 0 && b.global_l1) {
        hipError_t e = fill_u32(b.base1, b.map.nk1, 0u, st);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_l1_scatter_home<true>, dim3((unsigned)b.nwg), dim3(256), 0, st, n,
                           uu, vv, ww, wt1, g, b.map, b.chunk, b.off1, b.base1, rec);
    } else if (n > 0 && (b.stage & kStageL1)) {
        hipLaunchKernelGGL(k_l1_scatter_home_staged, dim3((unsigned)b.nwg), dim3(kStageThreads),
                           0, st, n, uu, vv, ww, wt1, g, b.map, b.chunk, b.off1, b.base1, rec);
    } else if (n > 0) {
        hipLaunchKernelGGL(k_l1_scatter_home<false>, dim3((unsigned)b.nwg), dim3(256),
                           (size_t)b.map.nk1 * sizeof(unsigned), st, n, uu, vv, ww, wt1, g,
                           b.map, b.chunk, b.off1, b.base1, rec);
    }
    return level2<kOutHome>(g, b, (const DEntry*)rec, entries, st);
}

}  // namespace lofar
