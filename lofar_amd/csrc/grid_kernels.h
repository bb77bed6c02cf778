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

// Workspace layout for one bucketing of n samples into nb tiles.
struct Buckets {
    long long nb;          // number of (w, tile_v, tile_u) buckets
    unsigned* counts;      // [nb]
    unsigned* slots;       // [4n]: position of each (sample, tile) pair within its bucket
    long long* offsets;    // [nb+1]
    long long* partials;   // scan partials
    long long* total;      // host-visible total entry count (status[0])
    int* err;              // host-visible validation flag (status[1])
};

hipError_t bucket_count(int mode, long long n, const double* uu, const double* vv, const double* ww,
                        const Geometry& g, const Buckets& b, hipStream_t st);
hipError_t bucket_scan(const Buckets& b, hipStream_t st);
hipError_t bucket_zero(const Buckets& b, hipStream_t st);
hipError_t bucket_scatter_grid(int precision, long long n, long long F, long long P,
                               const double* uu, const double* vv, const double* ww,
                               const float* vr, const float* vi, const float* wt,
                               const Geometry& g, const Buckets& b, void* entries,
                               uint32_t* wins, hipStream_t st);
hipError_t bucket_scatter_home(long long n, const double* uu, const double* vv, const double* ww,
                               const Geometry& g, const Buckets& b, DEntry* entries,
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

enum { kBucketFootprint = 0, kBucketHome = 1 };

}  // namespace lofar
