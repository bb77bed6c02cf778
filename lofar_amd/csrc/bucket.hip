// Bucketing of the samples into (w-plane, uv tile) buckets for gfx950: a
// two-level counting sort whose histograms live in LDS.
//
// Level 1 groups the (sample, tile) pairs by super-tile (sst x sst tiles of one
// w-plane).  Every workgroup takes a contiguous chunk of samples, counts its
// pairs per super-tile in an LDS histogram and reserves its range in each
// super-tile with one global atomic per non-empty bin (instead of one per pair),
// then, after the scan, places its pairs with LDS atomics, so each workgroup
// writes runs of consecutive entries.  Level 2 gives each super-tile one
// workgroup that counts its entries per fine tile in LDS and writes them in
// (w, tile) bucket order.  The entries carry what the gridder needs: the
// window clipped to the tile (w-stacking_omp.c:93-96) and the sample's
// position and channel/polarisation-summed visibility (:85-90,116-133).
#include "grid_kernels.h"

namespace lofar {

constexpr int kScanBlock = 4096;  // elements per block of the local scan (256 threads x 16)

__global__ __launch_bounds__(256) void k_fill_u32(unsigned* p, long long n, unsigned v)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// level 1: count

__global__ __launch_bounds__(256) void k_l1_count(int mode, long long n, const double* uu,
                                                  const double* vv, const double* ww, Geometry g,
                                                  TileMap tm, long long chunk, unsigned* count1,
                                                  unsigned* base1, int* err)
{
    extern __shared__ unsigned hist[];
    const int tid = threadIdx.x;
    for (long long k = tid; k < tm.nk1; k += 256) hist[k] = 0u;
    __syncthreads();
    const long long beg = (long long)blockIdx.x * chunk;
    const long long end = min(n, beg + chunk);
    bool bad = false;
    for (long long i = beg + tid; i < end; i += 256) {
        SampleWin s;
        if (!sample_window(uu[i], vv[i], ww[i], g, s)) {
            bad = true;
            continue;
        }
        if (mode == kBucketHome) {
            const int cu = min(max(s.gu, 0), g.nx - 1), cv = min(max(s.gv, 0), g.ny - 1);
            atomicAdd(&hist[tm.key1(s.gw, cv / kTile, cu / kTile)], 1u);
            continue;
        }
        if (s.j0 > s.j1 || s.k0 > s.k1) continue;
        for (int tv = s.k0 / kTile; tv <= s.k1 / kTile; tv++)
            for (int tu = s.j0 / kTile; tu <= s.j1 / kTile; tu++)
                atomicAdd(&hist[tm.key1(s.gw, tv, tu)], 1u);
    }
    if (bad) atomicOr(err, 1);
    __syncthreads();
    unsigned* my_base = base1 + (long long)blockIdx.x * tm.nk1;
    for (long long k = tid; k < tm.nk1; k += 256) {
        const unsigned c = hist[k];
        my_base[k] = c ? atomicAdd(&count1[k], c) : 0u;
    }
}

// ---------------------------------------------------------------------------
// exclusive scan (counts -> offsets; total to offsets[n] and to the status word)

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
// level 1: scatter into super-tile order

// Channel/polarisation sum of one sample with the reference's weighting:
// float product weight*vis, promoted to double, NaN real parts skipped
// (w-stacking_omp.c:116-133).
__device__ __forceinline__ void sample_vis(long long i, long long F, long long P, const float* vr,
                                           const float* vi, const float* wt, double& sr,
                                           double& si)
{
    sr = 0.0;
    si = 0.0;
    const long long vbase = i * F * P;
    for (long long f = 0; f < F; f++) {
        for (long long p = 0; p < P; p++) {
            const long long ifine = vbase + f * P + p;
            const float x = vr[ifine];
            if (!isnan(x)) {
                const float w = wt[i * P + p];
                const float a = w * x;
                const float b = w * vi[ifine];
                sr += (double)a;
                si += (double)b;
            }
        }
    }
}

template <bool FP64>
__global__ __launch_bounds__(256) void k_l1_scatter_grid(long long n, long long F, long long P,
                                                         const double* uu, const double* vv,
                                                         const double* ww, const float* vr,
                                                         const float* vi, const float* wt,
                                                         Geometry g, TileMap tm, long long chunk,
                                                         const long long* off1,
                                                         const unsigned* base1, void* rec,
                                                         uint32_t* recwin)
{
    extern __shared__ unsigned cur[];
    const int tid = threadIdx.x;
    const unsigned* my_base = base1 + (long long)blockIdx.x * tm.nk1;
    for (long long k = tid; k < tm.nk1; k += 256) cur[k] = my_base[k];
    __syncthreads();
    const long long beg = (long long)blockIdx.x * chunk;
    const long long end = min(n, beg + chunk);
    for (long long i = beg + tid; i < end; i += 256) {
        SampleWin s;
        if (!sample_window(uu[i], vv[i], ww[i], g, s)) continue;
        if (s.j0 > s.j1 || s.k0 > s.k1) continue;
        double sr, si;
        sample_vis(i, F, P, vr, vi, wt, sr, si);
        for (int tv = s.k0 / kTile; tv <= s.k1 / kTile; tv++) {
            const int v0 = tv * kTile;
            const int b0 = max(s.k0, v0) - v0, b1 = min(s.k1, v0 + kTile - 1) - v0;
            for (int tu = s.j0 / kTile; tu <= s.j1 / kTile; tu++) {
                const int u0 = tu * kTile;
                const int a0 = max(s.j0, u0) - u0, a1 = min(s.j1, u0 + kTile - 1) - u0;
                const long long k1 = tm.key1(s.gw, tv, tu);
                const long long pos = off1[k1] + atomicAdd(&cur[k1], 1u);
                recwin[pos] = pack_win(a0, a1, b0, b1, tm.fine(tv, tu));
                if constexpr (FP64) {
                    Entry64 e;
                    e.pu = s.pu;
                    e.pv = s.pv;
                    e.vr = sr;
                    e.vi = si;
                    ((Entry64*)rec)[pos] = e;
                } else {
                    Entry32 e;
                    e.du0 = (float)((double)(u0 + a0) + 0.5 - s.pu);
                    e.dv0 = (float)((double)(v0 + b0) + 0.5 - s.pv);
                    e.vr = (float)sr;
                    e.vi = (float)si;
                    ((Entry32*)rec)[pos] = e;
                }
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_l1_scatter_home(long long n, const double* uu,
                                                         const double* vv, const double* ww,
                                                         Geometry g, TileMap tm, long long chunk,
                                                         const long long* off1,
                                                         const unsigned* base1, DEntry* rec)
{
    extern __shared__ unsigned cur[];
    const int tid = threadIdx.x;
    const unsigned* my_base = base1 + (long long)blockIdx.x * tm.nk1;
    for (long long k = tid; k < tm.nk1; k += 256) cur[k] = my_base[k];
    __syncthreads();
    const long long beg = (long long)blockIdx.x * chunk;
    const long long end = min(n, beg + chunk);
    for (long long i = beg + tid; i < end; i += 256) {
        SampleWin s;
        if (!sample_window(uu[i], vv[i], ww[i], g, s)) continue;
        const int cu = min(max(s.gu, 0), g.nx - 1), cv = min(max(s.gv, 0), g.ny - 1);
        const int tv = cv / kTile, tu = cu / kTile;
        const long long k1 = tm.key1(s.gw, tv, tu);
        const long long pos = off1[k1] + atomicAdd(&cur[k1], 1u);
        DEntry e;
        e.pu = s.pu;
        e.pv = s.pv;
        e.idx = i;
        e.pad = tm.fine(tv, tu);
        rec[pos] = e;
    }
}

// ---------------------------------------------------------------------------
// level 2: one workgroup per super-tile, fine-tile order

constexpr int kMaxFine = 256;  // sst <= 16: the fine index fills the top byte of a packed window

// fine-tile index of a level-1 record: in DEntry.pad (home) or in the top byte
// of the packed window (footprint)
__device__ __forceinline__ int fine_of(const DEntry* rec, const uint32_t*, long long r)
{
    return (int)rec[r].pad;
}
template <typename Rec>
__device__ __forceinline__ int fine_of(const Rec*, const uint32_t* recwin, long long r)
{
    return (int)(recwin[r] >> 24);
}

template <typename Rec, bool HOME>
__global__ __launch_bounds__(256) void k_l2_sort(TileMap tm, const long long* off1,
                                                 const Rec* rec, const uint32_t* recwin, Rec* ent,
                                                 uint32_t* entwin, long long* offsets)
{
    __shared__ unsigned cnt[kMaxFine];
    const int tid = threadIdx.x;
    const int nf = tm.sst * tm.sst;
    const long long k1 = blockIdx.x;
    const long long beg = off1[k1], end = off1[k1 + 1];
    for (int f = tid; f < nf; f += 256) cnt[f] = 0u;
    __syncthreads();
    for (long long r = beg + tid; r < end; r += 256) {
        const int f = fine_of(rec, recwin, r);
        atomicAdd(&cnt[f], 1u);
    }
    __syncthreads();
    if (tid == 0) {
        unsigned run = 0;
        for (int f = 0; f < nf; f++) {
            const unsigned c = cnt[f];
            cnt[f] = run;
            run += c;
        }
    }
    __syncthreads();
    for (int f = tid; f < nf; f += 256) offsets[k1 * nf + f] = beg + cnt[f];
    if (k1 == 0 && tid == 0) offsets[tm.nk1 * nf] = off1[tm.nk1];
    __syncthreads();
    for (long long r = beg + tid; r < end; r += 256) {
        const int f = fine_of(rec, recwin, r);
        const long long p = beg + atomicAdd(&cnt[f], 1u);
        ent[p] = rec[r];
        if constexpr (!HOME) entwin[p] = recwin[r] & 0xffffffu;
    }
}

// ---------------------------------------------------------------------------
// host dispatch

static inline unsigned blocks_for(long long n, int t) { return (unsigned)((n + t - 1) / t); }

hipError_t bucket_count(int mode, long long n, const double* uu, const double* vv, const double* ww,
                        const Geometry& g, const Buckets& b, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(b.map.nk1, 256)), dim3(256), 0, st, b.count1,
                       b.map.nk1, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, st, (unsigned*)b.err, 1LL, 0u);
    if (n > 0)
        hipLaunchKernelGGL(k_l1_count, dim3((unsigned)b.nwg), dim3(256),
                           (size_t)b.map.nk1 * sizeof(unsigned), st, mode, n, uu, vv, ww, g, b.map,
                           b.chunk, b.count1, b.base1, b.err);
    return hipGetLastError();
}

hipError_t bucket_scan(const Buckets& b, hipStream_t st)
{
    const long long nb = b.map.nk1;
    const long long nblk = (nb + kScanBlock - 1) / kScanBlock;
    hipLaunchKernelGGL(k_scan_local, dim3((unsigned)nblk), dim3(256), 0, st, b.count1, b.off1,
                       b.partials, nb);
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, st, b.partials, nblk, b.off1 + nb,
                       b.total);
    hipLaunchKernelGGL(k_scan_add, dim3(blocks_for(nb, 256)), dim3(256), 0, st, b.off1, b.partials,
                       nb);
    return hipGetLastError();
}

hipError_t bucket_sort_grid(int precision, long long n, long long F, long long P,
                            const double* uu, const double* vv, const double* ww, const float* vr,
                            const float* vi, const float* wt, const Geometry& g, const Buckets& b,
                            void* rec, uint32_t* recwin, void* entries, uint32_t* wins,
                            hipStream_t st)
{
    const size_t lds = (size_t)b.map.nk1 * sizeof(unsigned);
    if (n > 0) {
        if (precision == LOFAR_FP64)
            hipLaunchKernelGGL((k_l1_scatter_grid<true>), dim3((unsigned)b.nwg), dim3(256), lds, st,
                               n, F, P, uu, vv, ww, vr, vi, wt, g, b.map, b.chunk, b.off1, b.base1,
                               rec, recwin);
        else
            hipLaunchKernelGGL((k_l1_scatter_grid<false>), dim3((unsigned)b.nwg), dim3(256), lds,
                               st, n, F, P, uu, vv, ww, vr, vi, wt, g, b.map, b.chunk, b.off1,
                               b.base1, rec, recwin);
    }
    if (precision == LOFAR_FP64)
        hipLaunchKernelGGL((k_l2_sort<Entry64, false>), dim3((unsigned)b.map.nk1), dim3(256), 0, st,
                           b.map, b.off1, (const Entry64*)rec, recwin, (Entry64*)entries, wins,
                           b.offsets);
    else
        hipLaunchKernelGGL((k_l2_sort<Entry32, false>), dim3((unsigned)b.map.nk1), dim3(256), 0, st,
                           b.map, b.off1, (const Entry32*)rec, recwin, (Entry32*)entries, wins,
                           b.offsets);
    return hipGetLastError();
}

hipError_t bucket_sort_home(long long n, const double* uu, const double* vv, const double* ww,
                            const Geometry& g, const Buckets& b, DEntry* rec, DEntry* entries,
                            hipStream_t st)
{
    const size_t lds = (size_t)b.map.nk1 * sizeof(unsigned);
    if (n > 0)
        hipLaunchKernelGGL(k_l1_scatter_home, dim3((unsigned)b.nwg), dim3(256), lds, st, n, uu, vv,
                           ww, g, b.map, b.chunk, b.off1, b.base1, rec);
    hipLaunchKernelGGL((k_l2_sort<DEntry, true>), dim3((unsigned)b.map.nk1), dim3(256), 0, st,
                       b.map, b.off1, (const DEntry*)rec, (const uint32_t*)nullptr, entries,
                       (uint32_t*)nullptr, b.offsets);
    return hipGetLastError();
}

}  // namespace lofar
