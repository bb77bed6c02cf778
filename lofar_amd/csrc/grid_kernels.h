// Host dispatch prototypes of the bucketing, gridding and degridding kernels.
#pragma once
#include "common.h"

namespace lofar {

// Window of a sample at grid position (pu, pv) = (uu/dx, vv/dx), exactly as
// w-stacking_omp.c:85-96 derives it: (int) truncations and the clipping rule
// at the grid edges.  Empty when j0 > j1 or k0 > k1.
struct Window {
    int gu, gv;          // (int) truncations
    int j0, j1, k0, k1;  // clipped window
};

__host__ __device__ inline Window window_of(double pu, double pv, const Geometry& g)
{
    Window s;
    s.gu = (int)pu;
    s.gv = (int)pv;
    s.j0 = (s.gu > g.K - 1) ? s.gu - g.K : 0;
    s.j1 = (s.gu < g.nx - g.K) ? s.gu + g.K : g.nx - 1;
    s.k0 = (s.gv > g.K - 1) ? s.gv - g.K : 0;
    s.k1 = (s.gv < g.ny - g.K) ? s.gv + g.K : g.ny - 1;
    return s;
}

// Validated geometry of one sample.
struct SampleWin : Window {
    double pu, pv;  // uu/dx, vv/dx
    int gw;         // (int)(ww/dw)
};

__device__ __forceinline__ bool sample_window(double u, double v, double w, const Geometry& g,
                                              SampleWin& s)
{
    if (!isfinite(u) || !isfinite(v) || !isfinite(w)) return false;
    const double pu = u / g.dx, pv = v / g.dx, pw = w / g.dw;
    const double lim = 1073741824.0;  // 2^30 keeps every (int) cast defined
    if (!(fabs(pu) < lim) || !(fabs(pv) < lim)) return false;
    if (!(pw > -1.0) || !(pw < (double)g.nw)) return false;
    static_cast<Window&>(s) = window_of(pu, pv, g);
    s.pu = pu;
    s.pv = pv;
    s.gw = (int)pw;
    return true;
}

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

// Home cell of a sample for degridding: its truncated position clamped to the grid.
__device__ __forceinline__ void home_cell(const Window& s, const Geometry& g, int& cu, int& cv)
{
    cu = min(max(s.gu, 0), g.nx - 1);
    cv = min(max(s.gv, 0), g.ny - 1);
}

// Bucket order of the (w-plane, tile) buckets: super-tiles of sst x sst tiles,
// (w, super-row, super-column) major, then the fine tile inside the super-tile.
struct TileMap {
    int sst, lsst;   // super-tile side in tiles (power of two) and its log2
    int nstu, nstv;  // super-tiles along u and v
    int rfb;         // fraction bits of an Rec32 position (rec_fraction_bits(sst))
    int olog;        // launch order of the per-tile kernels (launch_bucket)
    long long nk1;   // super-tile buckets, nw*nstv*nstu
    __host__ __device__ long long key1(int w, int tv, int tu) const
    {
        return ((long long)w * nstv + (tv >> lsst)) * nstu + (tu >> lsst);
    }
    __host__ __device__ long long key_st(int w, int sv, int su) const
    {
        return ((long long)w * nstv + sv) * nstu + su;
    }
    // super-tile coordinates of level-1 key k1
    __host__ __device__ void super_of(long long k1, int& sv, int& su) const
    {
        su = (int)(k1 % nstu);
        sv = (int)((k1 / nstu) % nstv);
    }
    __host__ __device__ int fine(int tv, int tu) const
    {
        return ((tv & (sst - 1)) << lsst) | (tu & (sst - 1));
    }
    __host__ __device__ long long nbuckets() const { return nk1 << (2 * lsst); }
    // tile of bucket bkt; false for the slots of partial super-tiles beyond the grid
    __device__ bool decode(long long bkt, const Geometry& g, long long& w, int& tv, int& tu) const
    {
        const long long k1 = bkt >> (2 * lsst);
        const int f = (int)(bkt & ((1 << (2 * lsst)) - 1));
        const int stu = (int)(k1 % nstu);
        const long long r = k1 / nstu;
        const int stv = (int)(r % nstv);
        w = r / nstv;
        tv = (stv << lsst) + (f >> lsst);
        tu = (stu << lsst) + (f & (sst - 1));
        return tv < g.ntv && tu < g.ntu;
    }
    // Launch order of the per-tile kernels (one workgroup per bucket slot): with olog > 0,
    // workgroup x takes the tile at x of a walk through each plane in blocks of
    // 2^olog x 2^olog tiles (row-major blocks, row-major tiles inside), so that the
    // tiles in flight together do not depend on the super-tile side; olog = 0: bucket
    // order.  Returns that tile's bucket, or -1 for the slots of partial super-tiles
    // beyond the grid.
    __device__ long long launch_bucket(long long x, const Geometry& g, long long& w, int& tv,
                                       int& tu) const
    {
        if (olog == 0) return decode(x, g, w, tv, tu) ? x : -1;
        const long long per_plane = (long long)nstv * nstu << (2 * lsst);
        w = x / per_plane;
        const long long r = x - w * per_plane;
        const long long ob = r >> (2 * olog);  // block of the padded plane
        const int nou = (nstu << lsst) >> olog;
        const int f = (int)(r & ((1 << (2 * olog)) - 1));
        tv = (int)(ob / nou << olog) + (f >> olog);
        tu = (int)(ob % nou << olog) + (f & ((1 << olog) - 1));
        if (tv >= g.ntv || tu >= g.ntu) return -1;
        return (key1((int)w, tv, tu) << (2 * lsst)) | fine(tv, tu);
    }
};

// Level-1 record of the FP32 mode, 16 bytes: per axis the sample's position p
// relative to its super-tile's origin s0 (first cell), as floor(p) - s0 + 16 in
// the high 32 - rfb bits and the fraction p - floor(p) in the low rfb bits.  A
// record only exists for a super-tile its window touches, so floor(p) - s0 lies
// in [-K-1, sst*64+K), K <= 15.  The fraction is rounded to rfb bits but kept
// nonzero when it is nonzero and below 1, so the decoded position has the
// sample's floor and (int) truncation, hence its exact window and tiles; only
// the offset inside the cell is quantised (2^-rfb cells; rfb = 22 for 8x8-tile
// super-tiles).  The visibility sums are rounded to float as the FP32 entries
// hold them.
struct alignas(16) Rec32 {
    unsigned u, v;
    float vr, vi;
};

// bits left for the fraction when floor(p) - s0 + 16 must fit in [0, sst*64 + 32)
__host__ __device__ inline int rec_fraction_bits(int sst)
{
    int cb = 0;
    while ((1 << cb) < sst * kTile + 32) cb++;
    return 32 - cb;
}

__device__ __forceinline__ unsigned rec_axis(double p, int s0, int fb)
{
    const double fl = floor(p);
    const double f = p - fl;  // [0, 1), exact
    const long long one = 1LL << fb;
    long long q = (long long)(f * (double)one + 0.5);
    if (f > 0.0 && q == 0) q = 1;
    if (q >= one) q = one - 1;
    return ((unsigned)((int)fl - s0 + kCellBias) << fb) | (unsigned)q;
}

__device__ __forceinline__ double rec_pos(unsigned x, int s0, int fb)
{
    return (double)((int)(x >> fb) - kCellBias + s0) +
           (double)(x & ((1u << fb) - 1u)) * ldexp(1.0, -fb);
}

// A level-2 block: records [beg, end) of super-tile segment k1.
struct alignas(32) L2Block {
    long long k1, beg, end, pad;
};

// Workspace of one bucketing of n samples (two-level counting sort).  Level 1
// writes one record per (sample, super-tile its window touches), level 2
// expands every record into one entry per tile of its window in the super-tile.
struct Buckets {
    TileMap map;
    long long nwg, chunk;  // level-1 workgroups and samples per workgroup
    bool global_l1;        // level-1 histograms and cursors in global memory (more
                           // super-tiles than the LDS histograms hold)
    int stage;             // kStageL1 | kStageL2: scatters staged through LDS
    unsigned* count1;      // [nk1] records per super-tile
    long long* off1;       // [nk1+1] record offsets
    unsigned* count1e;     // [nk1] entries per super-tile
    long long* eoff1;      // [nk1+1] entry offsets
    unsigned* base1;       // [nwg*nk1] each workgroup's first slot in each super-tile
                           // ([nk1] global cursors when global_l1)
    long long* offsets;    // [nbuckets+1] entries of every (w, tile) bucket, bucket order
    long long nent = 0;    // entries in all buckets, when the host knows it (else 0)
    const void* vis = nullptr;  // index entries: each sample's summed visibility (float2),
                                // gathered by the gridder (grid_gathers)
    long long* partials;   // scan partials
    long long* total;      // host-visible total entry count (status[0])
    long long* total_rec;  // total record count (device)
    int* err;              // host-visible validation flag (status[1])
    // level 2 (set once the entry total is known)
    long long l2_block;       // records per level-2 block
    long long l2_max_blocks;  // grid of the per-block kernels: ceil(total/l2_block) + nk1
    unsigned* l2_nblk;        // [nk1] blocks per segment
    long long* l2_boff;       // [nk1+1] first block of each segment; [nk1] = block count
    L2Block* l2_blk;          // [l2_max_blocks]
    unsigned* l2_cnt;         // [l2_max_blocks * sst^2] fine counts, then block bases
};

enum { kBucketFootprint = 0, kBucketHome = 1 };
// Staged scatters (LDS buffer, records stored in runs): level 1 takes up to
// kStageMaxBins super-tiles (LDS histograms), level 2 any block.
enum { kStageL1 = 1, kStageL2 = 2 };
constexpr long long kStageMaxBins = 512;

// Level 1 (bucket.hip): count records and entries per super-tile, validate.
hipError_t bucket_count(int mode, long long n, const double* uu, const double* vv, const double* ww,
                        const Geometry& g, const Buckets& b, hipStream_t st);
hipError_t bucket_scan(const Buckets& b, hipStream_t st);
// Level 1 scatter into super-tile order (rec: one Entry64 per record), then level
// 2 into (w, tile) order (entries: Entry64 for FP64, Entry32 for FP32C).
// index_vis: the entries carry each sample's index in the real part's slot instead
// of its visibility (a sort kept across calls; plan_fill puts visibilities in).
hipError_t bucket_sort_grid(int precision, long long n, long long F, long long P,
                            const double* uu, const double* vv, const double* ww, const float* vr,
                            const float* vi, const float* wt, const Geometry& g, const Buckets& b,
                            Entry64* rec, void* entries, hipStream_t st, bool index_vis = false);
// total entries of index_vis sorted entries (pent) with this call's visibilities of the n
// samples, into ent; vis: plan_vis_bytes(precision, n) bytes of scratch
size_t plan_vis_bytes(int precision, long long n);
// each sample's summed visibility into vis (plan_fill's first pass alone)
hipError_t plan_vis(int precision, long long n, long long F, long long P, const float* vr,
                    const float* vi, const float* wt, void* vis, hipStream_t st);
hipError_t plan_fill(int precision, long long n, long long total, const void* pent, void* ent,
                     long long F, long long P, const float* vr, const float* vi, const float* wt,
                     void* vis, hipStream_t st);
// wt1 (one polarisation): the samples' weights, stored in the entries' spare word
hipError_t bucket_sort_home(long long n, const double* uu, const double* vv, const double* ww,
                            const float* wt1, const Geometry& g, const Buckets& b, DEntry* rec,
                            DEntry* entries, hipStream_t st);
// occ (optional, nw*ntu*ceil(ntv/64) words, zeroed by the caller): bit tv%64 of
// word (w, tu, tv/64) is set for every tile holding samples; empty tiles are then
// left unwritten (for a consumer that reads them as zeros).
// sm_log >= 0 (without accumulate): the cube is written strip-major, each plane
// [nx/L][ny][L] with L = 2^sm_log columns per strip.
// whether grid_tiles takes index entries with b.vis (the packed FP32C kernel)
bool grid_gathers(int precision, int accumulate);
hipError_t grid_tiles(int precision, const Geometry& g, const Buckets& b, const void* entries,
                      void* cube, int accumulate, unsigned long long* occ, int sm_log,
                      hipStream_t st);
// fills n words with v (stream-ordered, no memset node)
hipError_t fill_u32(unsigned* p, long long n, unsigned v, hipStream_t st);
// Bucket-free kernels for windows wider than kMaxTileWin: validation of every sample
// (*err set when one fails sample_window), gridding with global atomic adds (the cube is
// zeroed first unless accumulate), and a per-sample window gather for degridding.
hipError_t validate_samples(long long n, const double* uu, const double* vv, const double* ww,
                            const Geometry& g, int* err, hipStream_t st);
hipError_t grid_direct(int precision, long long n, long long F, long long P, const double* uu,
                       const double* vv, const double* ww, const float* vr, const float* vi,
                       const float* wt, const Geometry& g, void* cube, int accumulate,
                       hipStream_t st);
hipError_t degrid_direct(int precision, long long n, long long F, long long P, const double* uu,
                         const double* vv, const double* ww, const float* wt, const Geometry& g,
                         const void* cube, void* model_re, void* model_im, hipStream_t st);
// The tiles a degridding from home buckets b reads (each home tile and its 8
// neighbours): need_words(g) words, grid_tiles' occupancy indexing; for the adjoint's
// inverse column transform (fft_cols with sign -1), which then stores only those.
long long need_words(const Geometry& g);
hipError_t need_tiles(const Geometry& g, const Buckets& b, unsigned long long* need,
                      hipStream_t st);
hipError_t degrid_tiles(int precision, const Geometry& g, const Buckets& b, const DEntry* entries,
                        const void* cube, const float* weight, long long F, long long P,
                        void* model_re, void* model_im, hipStream_t st);
hipError_t phase_sum(int precision, const void* cube, double* re, double* im, int xaxis, int yaxis,
                     int nw, int xtot, int ytot, double resolution, double wmin, double wmax,
                     int accumulate, hipStream_t st);
hipError_t phase_adjoint(int precision, const double* re, const double* im, void* cube, int xaxis,
                         int yaxis, int nw, int xtot, int ytot, double resolution, double wmin,
                         double wmax, hipStream_t st);



}  // namespace lofar
