// Bucketing, tile gridding and tile degridding for gfx950.
//
// Gridding (wstack(), w-stacking_omp.c:25-145) is done owner-computes: every
// (w-plane, 64x64 uv tile) is one workgroup that accumulates all samples whose
// Gaussian window touches the tile in an LDS tile of FP64 accumulators
// (ds_add_f64), then writes the tile to HBM once with plain coalesced stores.
// A sample whose window straddles tiles is listed in each of them with the
// window clipped to that tile, so no two workgroups write the same cell and no
// global atomics or zero-fill pass are needed.  The uv-tile buckets are built
// by a counting sort: count -> exclusive scan -> scatter.
#include <cstdlib>

#include "grid_kernels.h"

namespace lofar {

constexpr int kScanBlock = 4096;  // elements per block of the local scan (256 threads x 16)

// ---------------------------------------------------------------------------
// bucketing

__global__ __launch_bounds__(256) void k_fill_u32(unsigned* p, long long n, unsigned v)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Counts every (sample, tile) pair and keeps the position the counting atomic
// returned, so the scatter places entries without atomics.  A window spans at
// most 2x2 tiles (2K+1 <= 17 < kTile): slot t = 2*(tv-tv0) + (tu-tu0).
__global__ __launch_bounds__(256) void k_count(int mode, long long n, const double* uu,
                                               const double* vv, const double* ww, Geometry g,
                                               unsigned* counts, unsigned* slots, int* err)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    SampleWin s;
    if (!sample_window(uu[i], vv[i], ww[i], g, s)) {
        atomicOr(err, 1);
        return;
    }
    const long long pbase = (long long)s.gw * g.ntv;
    if (mode == kBucketHome) {
        const int cu = min(max(s.gu, 0), g.nx - 1), cv = min(max(s.gv, 0), g.ny - 1);
        slots[4 * i] = atomicAdd(&counts[(pbase + cv / kTile) * g.ntu + cu / kTile], 1u);
        return;
    }
    if (s.j0 > s.j1 || s.k0 > s.k1) return;
    const int tv0 = s.k0 / kTile, tu0 = s.j0 / kTile;
    for (int tv = tv0; tv <= s.k1 / kTile; tv++)
        for (int tu = tu0; tu <= s.j1 / kTile; tu++)
            slots[4 * i + 2 * (tv - tv0) + (tu - tu0)] =
                atomicAdd(&counts[(pbase + tv) * g.ntu + tu], 1u);
}

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
__global__ __launch_bounds__(256) void k_scatter_grid(long long n, long long F, long long P,
                                                      const double* uu, const double* vv,
                                                      const double* ww, const float* vr,
                                                      const float* vi, const float* wt, Geometry g,
                                                      const long long* offsets,
                                                      const unsigned* slots, void* entries,
                                                      uint32_t* wins)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    SampleWin s;
    if (!sample_window(uu[i], vv[i], ww[i], g, s)) return;
    if (s.j0 > s.j1 || s.k0 > s.k1) return;
    double sr, si;
    sample_vis(i, F, P, vr, vi, wt, sr, si);
    const long long pbase = (long long)s.gw * g.ntv;
    const int tv0 = s.k0 / kTile, tu0 = s.j0 / kTile;
    for (int tv = tv0; tv <= s.k1 / kTile; tv++) {
        const int v0 = tv * kTile;
        const int b0 = max(s.k0, v0) - v0, b1 = min(s.k1, v0 + kTile - 1) - v0;
        for (int tu = tu0; tu <= s.j1 / kTile; tu++) {
            const int u0 = tu * kTile;
            const int a0 = max(s.j0, u0) - u0, a1 = min(s.j1, u0 + kTile - 1) - u0;
            const long long key = (pbase + tv) * g.ntu + tu;
            const long long pos = offsets[key] + slots[4 * i + 2 * (tv - tv0) + (tu - tu0)];
            wins[pos] = pack_win(a0, a1, b0, b1);
            if constexpr (FP64) {
                Entry64 e;
                e.pu = s.pu;
                e.pv = s.pv;
                e.vr = sr;
                e.vi = si;
                ((Entry64*)entries)[pos] = e;
            } else {
                Entry32 e;
                e.du0 = (float)((double)(u0 + a0) + 0.5 - s.pu);
                e.dv0 = (float)((double)(v0 + b0) + 0.5 - s.pv);
                e.vr = (float)sr;
                e.vi = (float)si;
                ((Entry32*)entries)[pos] = e;
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_scatter_home(long long n, const double* uu,
                                                      const double* vv, const double* ww,
                                                      Geometry g, const long long* offsets,
                                                      const unsigned* slots, DEntry* entries)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    SampleWin s;
    if (!sample_window(uu[i], vv[i], ww[i], g, s)) return;
    const int cu = min(max(s.gu, 0), g.nx - 1), cv = min(max(s.gv, 0), g.ny - 1);
    const long long key = ((long long)s.gw * g.ntv + cv / kTile) * g.ntu + cu / kTile;
    const long long pos = offsets[key] + slots[4 * i];
    DEntry e;
    e.pu = s.pu;
    e.pv = s.pv;
    e.idx = i;
    e.pad = 0;
    entries[pos] = e;
}

// ---------------------------------------------------------------------------
// gridding

// One workgroup per (w, tile_v, tile_u) bucket, one lane per entry.  The
// Gaussian weight norm*exp(-(du^2+dv^2)/2) (w-stacking_omp.c:13-18,50-52) is
// evaluated as the separable product norm*exp(-du^2/2)*exp(-dv^2/2): 2*WIN
// exponentials per entry instead of WIN^2.
template <int WIN, bool FP64>
__global__ __launch_bounds__(256) void k_grid_tiles(Geometry g, const long long* offsets,
                                                    const void* entries, const uint32_t* wins,
                                                    void* cube, int accumulate)
{
    __shared__ double acc[2 * kTile * kTile];
    const long long bkt = blockIdx.x;
    const int tu = (int)(bkt % g.ntu);
    const long long r = bkt / g.ntu;
    const int tv = (int)(r % g.ntv);
    const long long w = r / g.ntv;
    const int u0 = tu * kTile, v0 = tv * kTile;
    const int tcols = min(kTile, g.nx - u0), trows = min(kTile, g.ny - v0);
    const long long beg = offsets[bkt], end = offsets[bkt + 1];
    const long long cell0 = (w * g.ny + v0) * (long long)g.nx + u0;
    const int tid = threadIdx.x;

    if (beg == end) {
        if (!accumulate) {
            for (int c = tid; c < kTile * kTile; c += 256) {
                const int row = c / kTile, col = c % kTile;
                if (row < trows && col < tcols) {
                    const long long o = cell0 + (long long)row * g.nx + col;
                    if constexpr (FP64)
                        ((double2*)cube)[o] = double2{0.0, 0.0};
                    else
                        ((float2*)cube)[o] = float2{0.0f, 0.0f};
                }
            }
        }
        return;
    }
    for (int c = tid; c < kTile * kTile; c += 256) {
        const int row = c / kTile, col = c % kTile;
        double re = 0.0, im = 0.0;
        if (accumulate && row < trows && col < tcols) {
            const long long o = cell0 + (long long)row * g.nx + col;
            if constexpr (FP64) {
                const double2 x = ((const double2*)cube)[o];
                re = x.x;
                im = x.y;
            } else {
                const float2 x = ((const float2*)cube)[o];
                re = x.x;
                im = x.y;
            }
        }
        acc[2 * c] = re;
        acc[2 * c + 1] = im;
    }
    __syncthreads();

    const double norm = 0.5 / LOFAR_REF_PI;  // std22/PI with std = 1 cell
    for (long long e = beg + tid; e < end; e += 256) {
        const uint32_t wn = wins[e];
        const int a0 = wn & 255, na = (int)((wn >> 8) & 255) - a0;
        const int b0 = (wn >> 16) & 255, nb = (int)(wn >> 24) - b0;
        double* base = &acc[2 * (b0 * kTile + a0)];
        if constexpr (FP64) {
            const Entry64 en = ((const Entry64*)entries)[e];
            double eu[WIN];
#pragma unroll
            for (int q = 0; q < WIN; q++) {
                const double du = (double)(u0 + a0 + q) + 0.5 - en.pu;
                eu[q] = (q <= na) ? exp(-(du * du) * 0.5) : 0.0;
            }
#pragma unroll
            for (int kq = 0; kq < WIN; kq++) {
                if (kq <= nb) {
                    const double dv = (double)(v0 + b0 + kq) + 0.5 - en.pv;
                    const double f = norm * exp(-(dv * dv) * 0.5);
                    const double fr = f * en.vr, fi = f * en.vi;
                    double* row = base + 2 * kq * kTile;
#pragma unroll
                    for (int q = 0; q < WIN; q++) {
                        if (q <= na) {
                            atomicAdd(row + 2 * q, eu[q] * fr);
                            atomicAdd(row + 2 * q + 1, eu[q] * fi);
                        }
                    }
                }
            }
        } else {
            const Entry32 en = ((const Entry32*)entries)[e];
            float eu[WIN];
#pragma unroll
            for (int q = 0; q < WIN; q++) {
                const float du = en.du0 + (float)q;
                eu[q] = (q <= na) ? expf(-(du * du) * 0.5f) : 0.0f;
            }
#pragma unroll
            for (int kq = 0; kq < WIN; kq++) {
                if (kq <= nb) {
                    const float dv = en.dv0 + (float)kq;
                    const float f = (float)norm * expf(-(dv * dv) * 0.5f);
                    const float fr = f * en.vr, fi = f * en.vi;
                    double* row = base + 2 * kq * kTile;
#pragma unroll
                    for (int q = 0; q < WIN; q++) {
                        if (q <= na) {
                            atomicAdd(row + 2 * q, (double)(eu[q] * fr));
                            atomicAdd(row + 2 * q + 1, (double)(eu[q] * fi));
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    for (int c = tid; c < kTile * kTile; c += 256) {
        const int row = c / kTile, col = c % kTile;
        if (row < trows && col < tcols) {
            const long long o = cell0 + (long long)row * g.nx + col;
            if constexpr (FP64)
                ((double2*)cube)[o] = double2{acc[2 * c], acc[2 * c + 1]};
            else
                ((float2*)cube)[o] = float2{(float)acc[2 * c], (float)acc[2 * c + 1]};
        }
    }
}

// Same tile ownership, but WIN lanes share one entry: lane dj of a group owns
// window column dj, evaluates exp(-du^2/2) for it and exp(-dv^2/2) for window
// row dj, and walks the rows taking each row's factor from the group lane that
// computed it.  The atomics of one instruction then hit runs of WIN adjacent
// cells (few LDS bank conflicts) instead of 64 scattered ones.
template <int WIN, bool FP64>
__global__ __launch_bounds__(256) void k_grid_tiles_cols(Geometry g, const long long* offsets,
                                                         const void* entries,
                                                         const uint32_t* wins, void* cube,
                                                         int accumulate)
{
    constexpr int GPW = 64 / WIN;  // entries per wave per step
    __shared__ double acc[2 * kTile * kTile];
    const long long bkt = blockIdx.x;
    const int tu = (int)(bkt % g.ntu);
    const long long r = bkt / g.ntu;
    const int tv = (int)(r % g.ntv);
    const long long w = r / g.ntv;
    const int u0 = tu * kTile, v0 = tv * kTile;
    const int tcols = min(kTile, g.nx - u0), trows = min(kTile, g.ny - v0);
    const long long beg = offsets[bkt], end = offsets[bkt + 1];
    const long long cell0 = (w * g.ny + v0) * (long long)g.nx + u0;
    const int tid = threadIdx.x;

    if (beg == end) {
        if (!accumulate) {
            for (int c = tid; c < kTile * kTile; c += 256) {
                const int row = c / kTile, col = c % kTile;
                if (row < trows && col < tcols) {
                    const long long o = cell0 + (long long)row * g.nx + col;
                    if constexpr (FP64)
                        ((double2*)cube)[o] = double2{0.0, 0.0};
                    else
                        ((float2*)cube)[o] = float2{0.0f, 0.0f};
                }
            }
        }
        return;
    }
    for (int c = tid; c < kTile * kTile; c += 256) {
        const int row = c / kTile, col = c % kTile;
        double re = 0.0, im = 0.0;
        if (accumulate && row < trows && col < tcols) {
            const long long o = cell0 + (long long)row * g.nx + col;
            if constexpr (FP64) {
                const double2 x = ((const double2*)cube)[o];
                re = x.x;
                im = x.y;
            } else {
                const float2 x = ((const float2*)cube)[o];
                re = x.x;
                im = x.y;
            }
        }
        acc[2 * c] = re;
        acc[2 * c + 1] = im;
    }
    __syncthreads();

    const double norm = 0.5 / LOFAR_REF_PI;
    const int lane = tid & 63, wave = tid >> 6;
    const int grp = lane / WIN, dj = lane % WIN;
    const int src0 = (grp < GPW ? grp : 0) * WIN;  // first lane of this group
    for (long long e0 = beg + (long long)wave * GPW; e0 < end; e0 += 4 * GPW) {
        const long long e = e0 + grp;
        const bool live = grp < GPW && e < end;
        const long long ee = live ? e : beg;  // dead lanes read a valid entry and add nothing
        const uint32_t wn = wins[ee];
        const int a0 = wn & 255, na = (int)((wn >> 8) & 255) - a0;
        const int b0 = (wn >> 16) & 255, nb = (int)(wn >> 24) - b0;
        const bool col_ok = live && dj <= na;
        double* colp = &acc[2 * (b0 * kTile + a0 + dj)];
        if constexpr (FP64) {
            const Entry64 en = ((const Entry64*)entries)[ee];
            const double du = (double)(u0 + a0 + dj) + 0.5 - en.pu;
            const double eu = exp(-(du * du) * 0.5);
            const double dvm = (double)(v0 + b0 + dj) + 0.5 - en.pv;
            const double evm = norm * exp(-(dvm * dvm) * 0.5);
            const double fr = eu * en.vr, fi = eu * en.vi;
#pragma unroll
            for (int kq = 0; kq < WIN; kq++) {
                const double ev = __shfl(evm, src0 + kq);
                if (col_ok && kq <= nb) {
                    atomicAdd(colp + 2 * kq * kTile, ev * fr);
                    atomicAdd(colp + 2 * kq * kTile + 1, ev * fi);
                }
            }
        } else {
            const Entry32 en = ((const Entry32*)entries)[ee];
            const float du = en.du0 + (float)dj;
            const float eu = expf(-(du * du) * 0.5f);
            const float dvm = en.dv0 + (float)dj;
            const float evm = (float)norm * expf(-(dvm * dvm) * 0.5f);
            const float fr = eu * en.vr, fi = eu * en.vi;
#pragma unroll
            for (int kq = 0; kq < WIN; kq++) {
                const float ev = __shfl(evm, src0 + kq);
                if (col_ok && kq <= nb) {
                    atomicAdd(colp + 2 * kq * kTile, (double)(ev * fr));
                    atomicAdd(colp + 2 * kq * kTile + 1, (double)(ev * fi));
                }
            }
        }
    }
    __syncthreads();
    for (int c = tid; c < kTile * kTile; c += 256) {
        const int row = c / kTile, col = c % kTile;
        if (row < trows && col < tcols) {
            const long long o = cell0 + (long long)row * g.nx + col;
            if constexpr (FP64)
                ((double2*)cube)[o] = double2{acc[2 * c], acc[2 * c + 1]};
            else
                ((float2*)cube)[o] = float2{(float)acc[2 * c], (float)acc[2 * c + 1]};
        }
    }
}

// ---------------------------------------------------------------------------
// degridding (adjoint of gridding)

// One workgroup per home tile: the tile plus a halo of (WIN-1)/2 cells is staged
// in LDS; one lane per sample gathers its window (same clipping and weights as
// the gridder) and writes weight*sum for every channel and polarisation.
template <int WIN, bool FP64>
__global__ __launch_bounds__(256) void k_degrid_tiles(Geometry g, const long long* offsets,
                                                      const DEntry* ents, const void* cube,
                                                      const float* weight, long long F,
                                                      long long P, void* mre, void* mim)
{
    using Real = typename std::conditional<FP64, double, float>::type;
    using T2 = typename cplx<Real>::type;
    constexpr int H = (WIN - 1) / 2;
    constexpr int RS = kTile + 2 * H;
    __shared__ T2 region[RS * RS];
    const long long bkt = blockIdx.x;
    const int tu = (int)(bkt % g.ntu);
    const long long r = bkt / g.ntu;
    const int tv = (int)(r % g.ntv);
    const long long w = r / g.ntv;
    const int u0 = tu * kTile, v0 = tv * kTile;
    const long long beg = offsets[bkt], end = offsets[bkt + 1];
    if (beg == end) return;
    const T2* plane = (const T2*)cube + w * (long long)g.ny * g.nx;
    const int tid = threadIdx.x;
    for (int c = tid; c < RS * RS; c += 256) {
        const int rr = c / RS, cc = c % RS;
        const int gv = v0 - H + rr, gu = u0 - H + cc;
        T2 x{0, 0};
        if (gv >= 0 && gv < g.ny && gu >= 0 && gu < g.nx) x = plane[(long long)gv * g.nx + gu];
        region[c] = x;
    }
    __syncthreads();
    const double norm = 0.5 / LOFAR_REF_PI;
    for (long long e = beg + tid; e < end; e += 256) {
        const DEntry en = ents[e];
        const int gu = (int)en.pu, gv = (int)en.pv;
        const int j0 = (gu > g.K - 1) ? gu - g.K : 0;
        const int j1 = (gu < g.nx - g.K) ? gu + g.K : g.nx - 1;
        const int k0 = (gv > g.K - 1) ? gv - g.K : 0;
        const int k1 = (gv < g.ny - g.K) ? gv + g.K : g.ny - 1;
        const int na = j1 - j0, nb = k1 - k0;
        Real sr = 0, si = 0;
        if (na >= 0 && nb >= 0) {
            Real eu[WIN];
#pragma unroll
            for (int q = 0; q < WIN; q++) {
                const double du = (double)(j0 + q) + 0.5 - en.pu;
                if constexpr (FP64)
                    eu[q] = (q <= na) ? exp(-(du * du) * 0.5) : 0.0;
                else
                    eu[q] = (q <= na) ? expf(-(float)(du * du) * 0.5f) : 0.0f;
            }
            const T2* base = &region[(k0 - (v0 - H)) * RS + (j0 - (u0 - H))];
#pragma unroll
            for (int kq = 0; kq < WIN; kq++) {
                if (kq <= nb) {
                    const double dv = (double)(k0 + kq) + 0.5 - en.pv;
                    Real f;
                    if constexpr (FP64)
                        f = norm * exp(-(dv * dv) * 0.5);
                    else
                        f = (float)norm * expf(-(float)(dv * dv) * 0.5f);
                    Real rr = 0, ri = 0;
                    const T2* row = base + kq * RS;
#pragma unroll
                    for (int q = 0; q < WIN; q++) {
                        if (q <= na) {
                            const T2 x = row[q];
                            rr += eu[q] * x.x;
                            ri += eu[q] * x.y;
                        }
                    }
                    sr += f * rr;
                    si += f * ri;
                }
            }
        }
        const long long i = en.idx;
        for (long long p = 0; p < P; p++) {
            const Real wgt = (Real)weight[i * P + p];
            for (long long f = 0; f < F; f++) {
                const long long o = i * F * P + f * P + p;
                ((Real*)mre)[o] = wgt * sr;
                ((Real*)mim)[o] = wgt * si;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// host dispatch

static inline unsigned blocks_for(long long n, int t) { return (unsigned)((n + t - 1) / t); }

hipError_t bucket_zero(const Buckets& b, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(b.nb, 256)), dim3(256), 0, st, b.counts, b.nb,
                       0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, st, (unsigned*)b.err, 1LL, 0u);
    return hipGetLastError();
}

hipError_t bucket_count(int mode, long long n, const double* uu, const double* vv, const double* ww,
                        const Geometry& g, const Buckets& b, hipStream_t st)
{
    if (n > 0)
        hipLaunchKernelGGL(k_count, dim3(blocks_for(n, 256)), dim3(256), 0, st, mode, n, uu, vv,
                           ww, g, b.counts, b.slots, b.err);
    return hipGetLastError();
}

hipError_t bucket_scan(const Buckets& b, hipStream_t st)
{
    const long long nblk = (b.nb + kScanBlock - 1) / kScanBlock;
    hipLaunchKernelGGL(k_scan_local, dim3((unsigned)nblk), dim3(256), 0, st, b.counts, b.offsets,
                       b.partials, b.nb);
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, st, b.partials, nblk,
                       b.offsets + b.nb, b.total);
    hipLaunchKernelGGL(k_scan_add, dim3(blocks_for(b.nb, 256)), dim3(256), 0, st, b.offsets,
                       b.partials, b.nb);
    return hipGetLastError();
}

hipError_t bucket_scatter_grid(int precision, long long n, long long F, long long P,
                               const double* uu, const double* vv, const double* ww,
                               const float* vr, const float* vi, const float* wt,
                               const Geometry& g, const Buckets& b, void* entries, uint32_t* wins,
                               hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    if (precision == LOFAR_FP64)
        hipLaunchKernelGGL((k_scatter_grid<true>), dim3(blocks_for(n, 256)), dim3(256), 0, st, n, F,
                           P, uu, vv, ww, vr, vi, wt, g, b.offsets, b.slots, entries, wins);
    else
        hipLaunchKernelGGL((k_scatter_grid<false>), dim3(blocks_for(n, 256)), dim3(256), 0, st, n,
                           F, P, uu, vv, ww, vr, vi, wt, g, b.offsets, b.slots, entries, wins);
    return hipGetLastError();
}

hipError_t bucket_scatter_home(long long n, const double* uu, const double* vv, const double* ww,
                               const Geometry& g, const Buckets& b, DEntry* entries,
                               hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter_home, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, uu, vv, ww, g,
                       b.offsets, b.slots, entries);
    return hipGetLastError();
}

// gridding kernel variant: 0 = one lane per entry, 1 = WIN lanes per entry
// (LOFAR_GRID_CFG overrides, for tuning)
static int grid_cfg()
{
    static int cfg = -1;
    if (cfg < 0) {
        const char* s = getenv("LOFAR_GRID_CFG");
        cfg = s ? atoi(s) : 0;
    }
    return cfg;
}

template <int WIN>
static void launch_grid(int precision, const Geometry& g, const Buckets& b, const void* entries,
                        const uint32_t* wins, void* cube, int accumulate, hipStream_t st)
{
    const bool cols = grid_cfg() == 1;
    if (precision == LOFAR_FP64) {
        if (cols)
            hipLaunchKernelGGL((k_grid_tiles_cols<WIN, true>), dim3((unsigned)b.nb), dim3(256), 0,
                               st, g, b.offsets, entries, wins, cube, accumulate);
        else
            hipLaunchKernelGGL((k_grid_tiles<WIN, true>), dim3((unsigned)b.nb), dim3(256), 0, st,
                               g, b.offsets, entries, wins, cube, accumulate);
    } else {
        if (cols)
            hipLaunchKernelGGL((k_grid_tiles_cols<WIN, false>), dim3((unsigned)b.nb), dim3(256), 0,
                               st, g, b.offsets, entries, wins, cube, accumulate);
        else
            hipLaunchKernelGGL((k_grid_tiles<WIN, false>), dim3((unsigned)b.nb), dim3(256), 0, st,
                               g, b.offsets, entries, wins, cube, accumulate);
    }
}

hipError_t grid_tiles(int precision, const Geometry& g, const Buckets& b, const void* entries,
                      const uint32_t* wins, void* cube, int accumulate, hipStream_t st)
{
    const int win = 2 * g.K + 1;
    if (win <= 7)
        launch_grid<7>(precision, g, b, entries, wins, cube, accumulate, st);
    else if (win <= 9)
        launch_grid<9>(precision, g, b, entries, wins, cube, accumulate, st);
    else if (win <= 17)
        launch_grid<17>(precision, g, b, entries, wins, cube, accumulate, st);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

template <int WIN>
static void launch_degrid(int precision, const Geometry& g, const Buckets& b, const DEntry* ents,
                          const void* cube, const float* weight, long long F, long long P,
                          void* mre, void* mim, hipStream_t st)
{
    if (precision == LOFAR_FP64)
        hipLaunchKernelGGL((k_degrid_tiles<WIN, true>), dim3((unsigned)b.nb), dim3(256), 0, st, g,
                           b.offsets, ents, cube, weight, F, P, mre, mim);
    else
        hipLaunchKernelGGL((k_degrid_tiles<WIN, false>), dim3((unsigned)b.nb), dim3(256), 0, st, g,
                           b.offsets, ents, cube, weight, F, P, mre, mim);
}

hipError_t degrid_tiles(int precision, const Geometry& g, const Buckets& b, const DEntry* ents,
                        const void* cube, const float* weight, long long F, long long P,
                        void* mre, void* mim, hipStream_t st)
{
    const int win = 2 * g.K + 1;
    if (win <= 7)
        launch_degrid<7>(precision, g, b, ents, cube, weight, F, P, mre, mim, st);
    else if (win <= 9)
        launch_degrid<9>(precision, g, b, ents, cube, weight, F, P, mre, mim, st);
    else if (win <= 17)
        launch_degrid<17>(precision, g, b, ents, cube, weight, F, P, mre, mim, st);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace lofar
