// Host dispatch prototypes of the bucketing, gridding and degridding kernels.
#pragma once
#include "common.h"

namespace lofar {

// Geometry of one sample, exactly as w-stacking_omp.c:85-96 derives it.
struct SampleWin {
    double pu, pv;       // uu/dx, vv/dx
    int gw, gu, gv;      // (int) truncations
    int j0, j1, k0, k1;  // clipped window (may be empty)
};

__device__ __forceinline__ bool sample_window(double u, double v, double w, const Geometry& g,
                                              SampleWin& s)
{
    if (!isfinite(u) || !isfinite(v) || !isfinite(w)) return false;
    const double pu = u / g.dx, pv = v / g.dx, pw = w / g.dw;
    const double lim = 1073741824.0;  // 2^30 keeps every (int) cast defined
    if (!(fabs(pu) < lim) || !(fabs(pv) < lim)) return false;
    if (!(pw > -1.0) || !(pw < (double)g.nw)) return false;
    s.pu = pu;
    s.pv = pv;
    s.gw = (int)pw;
    s.gu = (int)pu;
    s.gv = (int)pv;
    s.j0 = (s.gu > g.K - 1) ? s.gu - g.K : 0;
    s.j1 = (s.gu < g.nx - g.K) ? s.gu + g.K : g.nx - 1;
    s.k0 = (s.gv > g.K - 1) ? s.gv - g.K : 0;
    s.k1 = (s.gv < g.ny - g.K) ? s.gv + g.K : g.ny - 1;
    return true;
}

// Bucket order of the (w-plane, tile) buckets: super-tiles of sst x sst tiles,
// (w, super-row, super-column) major, then the fine tile inside the super-tile.
struct TileMap {
    int sst, lsst;   // super-tile side in tiles (power of two) and its log2
    int nstu, nstv;  // super-tiles along u and v
    long long nk1;   // super-tile buckets, nw*nstv*nstu
    __host__ __device__ long long key1(int w, int tv, int tu) const
    {
        return ((long long)w * nstv + (tv >> lsst)) * nstu + (tu >> lsst);
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
};

// Workspace of one bucketing of n samples (two-level counting sort).
struct Buckets {
    TileMap map;
    long long nwg, chunk;  // level-1 workgroups and samples per workgroup
    unsigned* count1;      // [nk1] entries per super-tile
    long long* off1;       // [nk1+1]
    unsigned* base1;       // [nwg*nk1] each workgroup's first slot in each super-tile
    long long* offsets;    // [nbuckets+1] entries of every (w, tile) bucket, bucket order
    long long* partials;   // scan partials
    long long* total;      // host-visible total entry count (status[0])
    int* err;              // host-visible validation flag (status[1])
};

enum { kBucketFootprint = 0, kBucketHome = 1 };

// Level 1 (bucket.hip): count the (sample, tile) pairs per super-tile.
hipError_t bucket_count(int mode, long long n, const double* uu, const double* vv, const double* ww,
                        const Geometry& g, const Buckets& b, hipStream_t st);
hipError_t bucket_scan(const Buckets& b, hipStream_t st);
// Level 1 scatter into super-tile order, then level 2 into (w, tile) order.
hipError_t bucket_sort_grid(int precision, long long n, long long F, long long P,
                            const double* uu, const double* vv, const double* ww, const float* vr,
                            const float* vi, const float* wt, const Geometry& g, const Buckets& b,
                            void* rec, uint32_t* recwin, void* entries, uint32_t* wins,
                            hipStream_t st);
hipError_t bucket_sort_home(long long n, const double* uu, const double* vv, const double* ww,
                            const Geometry& g, const Buckets& b, DEntry* rec, DEntry* entries,
                            hipStream_t st);
hipError_t grid_tiles(int precision, const Geometry& g, const Buckets& b, const void* entries,
                      const uint32_t* wins, void* cube, int accumulate, hipStream_t st);
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
