// Bucketing, tile gridding and tile degridding for gfx950.
//
// Gridding (wstack(), w-stacking_omp.c:25-145) is done owner-computes: every
// (w-plane, 64x64 uv tile) is one workgroup that accumulates all samples whose
// Gaussian window touches the tile in an LDS tile (64-bit fixed point with
// ds_add_u64 in both precision modes, ds_add_f64 for tiles holding non-finite
// values and for accumulate mode), then writes the tile to HBM once with plain
// coalesced stores.
// A sample whose window straddles tiles is listed in each of them with the
// window clipped to that tile, so no two workgroups write the same cell and no
// global atomics or zero-fill pass are needed.  The (w-plane, tile) buckets
// come from the two-level counting sort in bucket.hip; a bucket index that
// names a tile beyond the grid (a partial super-tile) ends its workgroup at once.
#include <algorithm>
#include <cstdlib>

#include "grid_kernels.h"

namespace lofar {

// ---------------------------------------------------------------------------
// gridding

__global__ __launch_bounds__(64) void k_fill_u32_1(unsigned* p)
{
    if (threadIdx.x == 0) *p = 0u;
}

// Window of an entry at truncated cell (gu, gv), clipped to the grid as
// w-stacking_omp.c:93-96 does and then to the tile at (u0, v0): first cell
// (a0, b0) and inclusive extents (na, nb), tile-local.
struct TileWin {
    int a0, na, b0, nb;
};
__device__ __forceinline__ TileWin clip_window(int gu, int gv, const Geometry& g, int u0, int v0)
{
    const int j0 = (gu > g.K - 1) ? gu - g.K : 0;
    const int j1 = (gu < g.nx - g.K) ? gu + g.K : g.nx - 1;
    const int k0 = (gv > g.K - 1) ? gv - g.K : 0;
    const int k1 = (gv < g.ny - g.K) ? gv + g.K : g.ny - 1;
    TileWin t;
    t.a0 = max(j0, u0) - u0;
    t.na = min(j1, u0 + kTile - 1) - u0 - t.a0;
    t.b0 = max(k0, v0) - v0;
    t.nb = min(k1, v0 + kTile - 1) - v0 - t.b0;
    return t;
}

// FP32 entry decode: the window in the tile, and the centre of the window's
// first cell relative to the sample (j0 + 0.5 - pu) as an exact integer part
// plus the cell offset (0.5 - (pu - gu)).
struct Win32 {
    TileWin t;
    int iu, iv;      // u0 + a0 - gu, v0 + b0 - gv
    float hu, hv;    // 0.5 - (pu - gu), 0.5 - (pv - gv), exact in float
};
__device__ __forceinline__ Win32 decode32(const Entry32& en, const Geometry& g, int u0, int v0)
{
    Win32 w;
    const int gu = u0 + cell_of(en.wu), gv = v0 + cell_of(en.wv);
    w.t = clip_window(gu, gv, g, u0, v0);
    w.iu = u0 + w.t.a0 - gu;
    w.iv = v0 + w.t.b0 - gv;
    w.hu = 0.5f - frac_of(en.wu);
    w.hv = 0.5f - frac_of(en.wv);
    return w;
}

// Cell c of a tile's flush, ordered so that consecutive c are stored at
// consecutive addresses: row-major planes, row c/64 and column c%64;
// strip-major planes ([nx/L][ny][L], L = 2^sm_log), strip by strip, each strip's
// 64 rows of L columns.  a is the cell's index in the LDS tile.
struct FlushCell {
    int row, col, a;
    long long off;
};
__device__ __forceinline__ FlushCell flush_cell(int c, int sm_log, const Geometry& g, long long w,
                                                int u0, int v0)
{
    FlushCell f;
    if (sm_log < 0) {
        f.row = c / kTile;
        f.col = c % kTile;
        f.off = (w * g.ny + v0 + f.row) * (long long)g.nx + u0 + f.col;
    } else {
        const int L = 1 << sm_log;
        static_assert(kTile == 64, "strip indexing assumes 64-cell tiles");
        const int s = c >> (6 + sm_log), rem = c & ((kTile << sm_log) - 1);
        f.row = rem >> sm_log;
        f.col = s * L + (rem & (L - 1));
        const int u = u0 + f.col, v = v0 + f.row;
        f.off = w * (long long)g.ny * g.nx + ((((long long)(u >> sm_log)) * g.ny + v) << sm_log) +
                (u & (L - 1));
    }
    f.a = f.row * kTile + f.col;
    return f;
}

// LDS tiles hold the real parts of the 64x64 cells, then the imaginary parts:
// an atomic instruction's lanes (random cells) then spread over all banks.
constexpr int kPlane = kTile * kTile;

// One workgroup per (w, tile_v, tile_u) bucket, one lane per entry.  The
// Gaussian weight norm*exp(-(du^2+dv^2)/2) (w-stacking_omp.c:13-18,50-52) is
// evaluated as the separable product norm*exp(-du^2/2)*exp(-dv^2/2): 2*WIN
// exponentials per entry instead of WIN^2.
template <int WIN, bool FP64, int NT>
__global__ __launch_bounds__(NT) void k_grid_tiles(Geometry g, TileMap tm, const long long* offsets,
                                                    const void* entries, void* cube,
                                                    int accumulate, unsigned long long* occ,
                                                    int sm_log)
{
    __shared__ double acc[2 * kTile * kTile];
    long long w;
    int tv, tu;
    const long long bkt = tm.launch_bucket(blockIdx.x, g, w, tv, tu);
    if (bkt < 0) return;
    const int u0 = tu * kTile, v0 = tv * kTile;
    const int tcols = min(kTile, g.nx - u0), trows = min(kTile, g.ny - v0);
    const long long beg = offsets[bkt], end = offsets[bkt + 1];
    const long long cell0 = (w * g.ny + v0) * (long long)g.nx + u0;
    const int tid = threadIdx.x;

    if (occ) {
        // fused imaging: set the tile's occupancy bit (word (w, tu, tv/64), bit
        // tv%64, cleared beforehand); the column transform reads an empty tile as
        // zeros, so its cells are left unwritten
        if (beg == end) return;
        if (tid == 0)
            atomicOr(&occ[(w * g.ntu + tu) * ((g.ntv + 63) / 64) + tv / 64], 1ull << (tv % 64));
    }
    if (beg == end) {
        if (!accumulate) {
            for (int c = tid; c < kPlane; c += NT) {
                const FlushCell f = flush_cell(c, sm_log, g, w, u0, v0);
                if (f.row < trows && f.col < tcols) {
                    const long long o = f.off;
                    if constexpr (FP64)
                        ((double2*)cube)[o] = double2{0.0, 0.0};
                    else
                        ((float2*)cube)[o] = float2{0.0f, 0.0f};
                }
            }
        }
        return;
    }
    for (int c = tid; c < kTile * kTile; c += NT) {
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
        acc[c] = re;
        acc[kPlane + c] = im;
    }
    __syncthreads();

    const double norm = 0.5 / LOFAR_REF_PI;  // std22/PI with std = 1 cell
    for (long long e = beg + tid; e < end; e += NT) {
        if constexpr (FP64) {
            const Entry64 en = ((const Entry64*)entries)[e];
            const TileWin tw = clip_window((int)en.pu, (int)en.pv, g, u0, v0);
            const int a0 = tw.a0, na = tw.na, b0 = tw.b0, nb = tw.nb;
            double* base = &acc[b0 * kTile + a0];
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
                    double* row = base + kq * kTile;
#pragma unroll
                    for (int q = 0; q < WIN; q++) {
                        if (q <= na) {
                            atomicAdd(row + q, eu[q] * fr);
                            atomicAdd(row + kPlane + q, eu[q] * fi);
                        }
                    }
                }
            }
        } else {
            const Entry32 en = ((const Entry32*)entries)[e];
            const Win32 w32 = decode32(en, g, u0, v0);
            const int a0 = w32.t.a0, na = w32.t.na, b0 = w32.t.b0, nb = w32.t.nb;
            double* base = &acc[b0 * kTile + a0];
            float eu[WIN];
#pragma unroll
            for (int q = 0; q < WIN; q++) {
                const float du = (float)(w32.iu + q) + w32.hu;
                eu[q] = (q <= na) ? expf(-(du * du) * 0.5f) : 0.0f;
            }
#pragma unroll
            for (int kq = 0; kq < WIN; kq++) {
                if (kq <= nb) {
                    const float dv = (float)(w32.iv + kq) + w32.hv;
                    const float f = (float)norm * expf(-(dv * dv) * 0.5f);
                    const float fr = f * en.vr, fi = f * en.vi;
                    double* row = base + kq * kTile;
#pragma unroll
                    for (int q = 0; q < WIN; q++) {
                        if (q <= na) {
                            atomicAdd(row + q, (double)(eu[q] * fr));
                            atomicAdd(row + kPlane + q, (double)(eu[q] * fi));
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    for (int c = tid; c < kPlane; c += NT) {
        const FlushCell f = flush_cell(c, sm_log, g, w, u0, v0);
        if (f.row < trows && f.col < tcols) {
            if constexpr (FP64)
                ((double2*)cube)[f.off] = double2{acc[f.a], acc[kPlane + f.a]};
            else
                ((float2*)cube)[f.off] = float2{(float)acc[f.a], (float)acc[kPlane + f.a]};
        }
    }
}

// FP32C gridding with exact fixed-point accumulation.  Every contribution
// eu*(norm*ev*vis) is formed exactly as the double product of two floats,
// scaled by a per-tile power of two 2^S and rounded to the nearest 64-bit
// integer; the LDS tile accumulates those integers with ds_add_u64 (about twice
// the rate of ds_add_f64 and order-independent, hence bitwise reproducible).
// S leaves log2(entries)+2 bits of headroom below 2^51 for the largest possible
// cell sum (every entry of the tile at its maximum |vis|*norm), so no sum can
// overflow, and the quantum is ~2^-(49-log2(entries)) of the tile's largest
// contribution (far below the float rounding of the output).  The flush
// converts each integer sum back to double (exact below 2^53) times 2^-S and
// rounds once to float.  A tile holding a non-finite visibility takes the
// FP64-atomic path of k_grid_tiles instead (the reference propagates inf/NaN
// into the grid).
constexpr int kFixedBits = 51;

// Largest |vis| component of a tile over the mean of its other components above which the
// tile is summed in floating point instead of fixed point: the fixed-point quantum is
// ~2^-(58-log2(entries)) of the largest contribution in FP64 mode (2^-(47-...) in FP32C),
// so below this ratio a faint cell's relative error stays under ~4e-12 (FP64) and ~1e-8
// (FP32C, under the float rounding of the cube), and a bright outlier takes the double sums.
constexpr double kHdrRatio = 1024.0;

// Round |y| < 2^51 to the nearest integer as a two's-complement 64-bit value:
// adding 1.5*2^52 leaves round(y) in the low 52 bits of the sum's encoding,
// whose low word is 0 for the constant itself, so only the high word needs the
// constant's high word (0x43380000) taken off.
__device__ __forceinline__ unsigned long long to_fixed(double y)
{
    const double t = y + 0x1.8p52;
    const unsigned long long b = (unsigned long long)__double_as_longlong(t);
    const unsigned lo = (unsigned)b;
    const unsigned hi = (unsigned)(b >> 32) - 0x43380000u;
    return ((unsigned long long)hi << 32) | lo;
}

// One FP32C entry into the fixed-point tile (real parts, then imaginary parts
// kPlane words on): the cells of its clipped window.
template <int WIN>
__device__ __forceinline__ void accumulate_fx32(unsigned long long* acc, const Geometry& g, int u0,
                                                int v0, double norm, double scale,
                                                const Entry32& en)
{
    const Win32 w32 = decode32(en, g, u0, v0);
    const int a0 = w32.t.a0, na = w32.t.na, b0 = w32.t.b0, nb = w32.t.nb;
    unsigned long long* base = &acc[b0 * kTile + a0];
    double eu[WIN];
#pragma unroll
    for (int q = 0; q < WIN; q++) {
        const float du = (float)(w32.iu + q) + w32.hu;
        eu[q] = (q <= na) ? (double)expf(-(du * du) * 0.5f) : 0.0;
    }
#pragma unroll
    for (int kq = 0; kq < WIN; kq++) {
        if (kq <= nb) {
            const float dv = (float)(w32.iv + kq) + w32.hv;
            const float f = (float)norm * expf(-(dv * dv) * 0.5f);
            const double fr = (double)(f * en.vr) * scale, fi = (double)(f * en.vi) * scale;
            unsigned long long* row = base + kq * kTile;
#pragma unroll
            for (int q = 0; q < WIN; q++) {
                if (q <= na) {
                    atomicAdd(row + q, to_fixed(eu[q] * fr));
                    atomicAdd(row + kPlane + q, to_fixed(eu[q] * fi));
                }
            }
        }
    }
}

template <int WIN, int NT>
__global__ __launch_bounds__(NT) void k_grid_tiles_fx(Geometry g, TileMap tm, const long long* offsets,
                                                       const Entry32* entries, float2* cube,
                                                       unsigned long long* occ, int sm_log)
{
    __shared__ __align__(16) unsigned long long acc[2 * kPlane];
    __shared__ float wred[NT / 64];
    __shared__ double wsum[NT / 64];
    __shared__ int wfin[NT / 64];
    long long w;
    int tv, tu;
    const long long bkt = tm.launch_bucket(blockIdx.x, g, w, tv, tu);
    if (bkt < 0) return;
    const int u0 = tu * kTile, v0 = tv * kTile;
    const int tcols = min(kTile, g.nx - u0), trows = min(kTile, g.ny - v0);
    const long long beg = offsets[bkt], end = offsets[bkt + 1];
    const int tid = threadIdx.x;
    const double norm = 0.5 / LOFAR_REF_PI;

    if (occ) {
        // fused imaging: set the tile's occupancy bit (word (w, tu, tv/64), bit
        // tv%64, cleared beforehand); the column transform reads an empty tile as
        // zeros, so its cells are left unwritten
        if (beg == end) return;
        if (tid == 0)
            atomicOr(&occ[(w * g.ntu + tu) * ((g.ntv + 63) / 64) + tv / 64], 1ull << (tv % 64));
    }
    if (beg == end) {
        for (int c = tid; c < kPlane; c += NT) {
            const FlushCell f = flush_cell(c, sm_log, g, w, u0, v0);
            if (f.row < trows && f.col < tcols) cube[f.off] = float2{0.0f, 0.0f};
        }
        return;
    }
    // largest |vis| of the tile (NaN real parts are skipped upstream; any other
    // non-finite value sends the tile down the floating-point path).  The first
    // NT entries are loaded once, before the LDS clear, and kept in registers.
    const long long n = end - beg;
    const int lane = tid & 63, wave = tid >> 6;
    Entry32 en0 = {0u, 0u, 0.0f, 0.0f};
    if (tid < n) en0 = entries[beg + tid];
    for (int c = tid; c < kPlane; c += NT)  // 2*kPlane words, 16 bytes a store
        ((ulonglong2*)acc)[c] = ulonglong2{0ull, 0ull};  // also 0.0 as double
    float m = fmaxf(fabsf(en0.vr), fabsf(en0.vi));
    double sa = (double)fabsf(en0.vr) + (double)fabsf(en0.vi);
    bool finite = isfinite(en0.vr) && isfinite(en0.vi);
    for (long long e = beg + NT + tid; e < end; e += NT) {
        const Entry32 en = entries[e];
        finite = finite && isfinite(en.vr) && isfinite(en.vi);
        m = fmaxf(m, fmaxf(fabsf(en.vr), fabsf(en.vi)));
        sa += (double)fabsf(en.vr) + (double)fabsf(en.vi);
    }
    for (int off = 32; off > 0; off >>= 1) {
        m = fmaxf(m, __shfl_xor(m, off));
        sa += __shfl_xor(sa, off);
    }
    const bool wave_finite = __all(finite);
    if (lane == 0) {
        wred[wave] = m;
        wsum[wave] = sa;
        wfin[wave] = wave_finite ? 1 : 0;
    }
    __syncthreads();
    m = wred[0];
    sa = wsum[0];
    bool all_finite = wfin[0] != 0;
#pragma unroll
    for (int i = 1; i < NT / 64; i++) {
        m = fmaxf(m, wred[i]);
        sa += wsum[i];
        all_finite = all_finite && wfin[i] != 0;
    }
    // a tile whose largest |vis| is far above its mean (a bright outlier) takes the
    // floating-point path: the fixed-point quantum follows the largest contribution
    if (!all_finite || (double)m > kHdrRatio * (sa - (double)m) / (double)(2 * n - 1)) {
        // floating-point accumulation (ds_add_f64), as k_grid_tiles<WIN, false>
        double* dacc = (double*)acc;
        for (long long e = beg + tid; e < end; e += NT) {
            const Entry32 en = entries[e];
            const Win32 w32 = decode32(en, g, u0, v0);
            const int a0 = w32.t.a0, na = w32.t.na, b0 = w32.t.b0, nb = w32.t.nb;
            double* base = &dacc[b0 * kTile + a0];
            float eu[WIN];
#pragma unroll
            for (int q = 0; q < WIN; q++) {
                const float du = (float)(w32.iu + q) + w32.hu;
                eu[q] = (q <= na) ? expf(-(du * du) * 0.5f) : 0.0f;
            }
#pragma unroll
            for (int kq = 0; kq < WIN; kq++) {
                if (kq <= nb) {
                    const float dv = (float)(w32.iv + kq) + w32.hv;
                    const float f = (float)norm * expf(-(dv * dv) * 0.5f);
                    const float fr = f * en.vr, fi = f * en.vi;
                    double* row = base + kq * kTile;
#pragma unroll
                    for (int q = 0; q < WIN; q++) {
                        if (q <= na) {
                            atomicAdd(row + q, (double)(eu[q] * fr));
                            atomicAdd(row + kPlane + q, (double)(eu[q] * fi));
                        }
                    }
                }
            }
        }
        __syncthreads();
        for (int c = tid; c < kPlane; c += NT) {
            const FlushCell f = flush_cell(c, sm_log, g, w, u0, v0);
            if (f.row < trows && f.col < tcols)
                cube[f.off] = float2{(float)dacc[f.a], (float)dacc[kPlane + f.a]};
        }
        return;
    }
    int ex = 0;
    frexp(norm * (double)(m > 0.0f ? m : 1.0f), &ex);  // norm*m < 2^ex
    int hb = 2;
    for (long long cnt = n; cnt > 1; cnt >>= 1) hb++;
    const int S = kFixedBits - hb - ex;
    const double scale = ldexp(1.0, S);

    if (tid < n) accumulate_fx32<WIN>(acc, g, u0, v0, norm, scale, en0);
    for (long long e = beg + NT + tid; e < end; e += NT)
        accumulate_fx32<WIN>(acc, g, u0, v0, norm, scale, entries[e]);
    __syncthreads();
    const double inv = ldexp(1.0, -S);
    for (int c = tid; c < kPlane; c += NT) {
        const FlushCell f = flush_cell(c, sm_log, g, w, u0, v0);
        if (f.row < trows && f.col < tcols) {
            const double re = (double)(long long)acc[f.a] * inv;
            const double im = (double)(long long)acc[kPlane + f.a] * inv;
            cube[f.off] = float2{(float)re, (float)im};
        }
    }
}

// FP32C gridding with one ds_add_u64 per cell.  A cell's real and imaginary
// fixed-point sums share one 64-bit word as re*2^32 + im (two's complement,
// modulo 2^64): the sum of packed contributions is the packed pair of sums, and
// while each final sum lies in (-2^31, 2^31) the flush recovers the imaginary sum
// as the sign-extended low half and the real sum as (word - im) >> 32, whatever
// the partial sums did on the way.  The quantum 2^-S comes from exact integer
// bounds on the cell sums of each 4x4 block of the tile: the sum, over the
// entries whose clipped window touches the block, of norm*max(|vr|,|vi|) times
// the entry's largest Gaussian factor in the block, counted in units of 2^-20 of
// the tile's largest component and rounded up; the largest block bound, below
// 2^ex, gives S = 30 - ex (a factor 2 of headroom).  Contributions are exact
// double products (float factors) rounded once to the quantum; the integer sums
// make the cube independent of entry order (bitwise reproducible) and the LDS
// tile is 32 KiB.  On 64M samples over 4096x4096x32 the image error against
// the FP64 mode is 2.5e-7 relative RMS (2.1e-7 with two 64-bit words per cell)
// and the gridder takes 24% less time.  The quantum grows with the number of
// entries near the tile's densest block: a tile with a block touched by more
// than kPkMaxBlk entries is summed with one 64-bit word per component instead
// (the two-word fixed point of k_grid_tiles_fx, real parts and then imaginary
// parts through the same 32 KiB), and a tile holding a non-finite value or a
// bright outlier (kHdrRatio) in double (ds_add_f64), likewise in two halves.
constexpr int kBlk = 4;              // bound block side (cells)
constexpr int kNBlk = kTile / kBlk;  // blocks per tile side
constexpr int kPkHead = 30;          // block bound * 2^S < 2^30
constexpr float kBndUnit = 1048576.0f;  // bound units per the tile's largest |component|
constexpr int kPkMaxBlk = 256;       // entries touching a block, at most, for the packed sums
constexpr long long kPkMaxEntries = 65536;  // entries of a tile, below which the counts fit
constexpr int kCntShift = 48;        // block word: entry count above bit 48, units below

// round(y) modulo 2^32 for |y| < 2^51: the low word of the encoding of
// y + 1.5*2^52, whose unit in the last place is 1
__device__ __forceinline__ unsigned round_u32(double y)
{
    return (unsigned)__double_as_longlong(y + 0x1.8p52);
}

// One FP32C entry into the blocks its clipped window touches: one count (above
// kCntShift) and its bound units (below: fewer than 2^16 entries of at most
// 2^20 + 1 units stay below 2^37).  An entry's bound in a block is its largest
// component times the largest Gaussian factor of a cell of the block inside its
// window, per axis exp(-d^2/2) for the distance d from the sample to the span of
// those cells' centres (no cell centre is nearer); the hardware exponential's
// error is far inside the bound's factor 2 of headroom.  rm = 2^20 / (the tile's
// largest component).
template <int WIN>
__device__ __forceinline__ void add_block_bound(unsigned long long* bsum, const Geometry& g,
                                                int u0, int v0, float rm, const Entry32& en)
{
    constexpr int NB = (WIN - 1 + kBlk - 1) / kBlk + 1;  // blocks a window spans per axis
    const int cu = cell_of(en.wu), cv = cell_of(en.wv);
    const TileWin t = clip_window(u0 + cu, v0 + cv, g, u0, v0);
    const float pu = (float)cu + frac_of(en.wu), pv = (float)cv + frac_of(en.wv);
    const float x = fmaxf(fabsf(en.vr), fabsf(en.vi)) * rm;
    const int bc0 = t.a0 / kBlk, br0 = t.b0 / kBlk;
    const int bc1 = (t.a0 + t.na) / kBlk, br1 = (t.b0 + t.nb) / kBlk;
    float gu[NB], gv[NB];
#pragma unroll
    for (int j = 0; j < NB; j++) {
        const int bc = bc0 + j, br = br0 + j;
        const float lu = (float)max(bc * kBlk, t.a0) + 0.5f;
        const float hu = (float)min(bc * kBlk + kBlk - 1, t.a0 + t.na) + 0.5f;
        const float du = fmaxf(0.0f, fmaxf(lu - pu, pu - hu));
        gu[j] = __expf(-0.5f * du * du);
        const float lv = (float)max(br * kBlk, t.b0) + 0.5f;
        const float hv = (float)min(br * kBlk + kBlk - 1, t.b0 + t.nb) + 0.5f;
        const float dv = fmaxf(0.0f, fmaxf(lv - pv, pv - hv));
        gv[j] = x * __expf(-0.5f * dv * dv);
    }
#pragma unroll
    for (int i = 0; i < NB; i++) {
        if (br0 + i > br1) break;
#pragma unroll
        for (int j = 0; j < NB; j++) {
            if (bc0 + j > bc1) break;
            const unsigned long long xi =
                (unsigned long long)(unsigned)(gv[i] * gu[j]) + 1ull + (1ull << kCntShift);
            atomicAdd(&bsum[(br0 + i) * kNBlk + bc0 + j], xi);
        }
    }
}

// One FP32C entry into the packed fixed-point tile at 2^S: the cells of its clipped
// window.
template <int WIN>
__device__ __forceinline__ void accumulate_pk32(unsigned long long* acc, const Geometry& g, int u0,
                                                int v0, double norm, double scale,
                                                const Entry32& en)
{
    const Win32 w32 = decode32(en, g, u0, v0);
    const int a0 = w32.t.a0, na = w32.t.na, b0 = w32.t.b0, nb = w32.t.nb;
    unsigned long long* base = &acc[b0 * kTile + a0];
    float eu[WIN];
#pragma unroll
    for (int q = 0; q < WIN; q++) {
        const float du = (float)(w32.iu + q) + w32.hu;
        eu[q] = (q <= na) ? expf(-(du * du) * 0.5f) : 0.0f;
    }
#pragma unroll
    for (int kq = 0; kq < WIN; kq++) {
        if (kq <= nb) {
            const float dv = (float)(w32.iv + kq) + w32.hv;
            const float f = (float)norm * expf(-(dv * dv) * 0.5f);
            const double fr = (double)(f * en.vr) * scale, fi = (double)(f * en.vi) * scale;
            unsigned long long* row = base + kq * kTile;
#pragma unroll
            for (int q = 0; q < WIN; q++) {
                if (q <= na) {
                    const unsigned re = round_u32((double)eu[q] * fr);
                    const unsigned im = round_u32((double)eu[q] * fi);
                    atomicAdd(row + q, ((unsigned long long)re << 32) +
                                           (unsigned long long)(long long)(int)im);
                }
            }
        }
    }
}

// One FP32C entry's real (HALF 0) or imaginary (HALF 1) parts into a one-word-per-
// cell tile: 64-bit fixed point at 2^S (FX, the arithmetic of accumulate_fx32) or
// double.
template <int WIN, int HALF, bool FX>
__device__ __forceinline__ void accumulate_half(unsigned long long* acc, const Geometry& g, int u0,
                                                int v0, double norm, double scale,
                                                const Entry32& en)
{
    const Win32 w32 = decode32(en, g, u0, v0);
    const int a0 = w32.t.a0, na = w32.t.na, b0 = w32.t.b0, nb = w32.t.nb;
    unsigned long long* base = &acc[b0 * kTile + a0];
    float eu[WIN];
#pragma unroll
    for (int q = 0; q < WIN; q++) {
        const float du = (float)(w32.iu + q) + w32.hu;
        eu[q] = (q <= na) ? expf(-(du * du) * 0.5f) : 0.0f;
    }
#pragma unroll
    for (int kq = 0; kq < WIN; kq++) {
        if (kq <= nb) {
            const float dv = (float)(w32.iv + kq) + w32.hv;
            const float f = (float)norm * expf(-(dv * dv) * 0.5f);
            const float fv = f * (HALF ? en.vi : en.vr);
            const double fs = (double)fv * scale;
            unsigned long long* row = base + kq * kTile;
#pragma unroll
            for (int q = 0; q < WIN; q++) {
                if (q <= na) {
                    if constexpr (FX)
                        atomicAdd(row + q, to_fixed((double)eu[q] * fs));
                    else
                        atomicAdd((double*)row + q, (double)(eu[q] * fv));
                }
            }
        }
    }
}

template <bool FX>
__device__ __forceinline__ float half_value(unsigned long long word, double inv)
{
    if constexpr (FX)
        return (float)((double)(long long)word * inv);
    else
        return (float)__longlong_as_double((long long)word);
}

// Entry e; with vis (a plan's entries, whose real part's slot holds the sample index),
// its visibility is the sample's summed visibility in vis.
__device__ __forceinline__ Entry32 entry_at(const Entry32* entries, const float2* vis, long long e)
{
    Entry32 x = entries[e];
    if (vis) {
        const float2 v = vis[__float_as_uint(x.vr)];
        x.vr = v.x;
        x.vi = v.y;
    }
    return x;
}

// A tile in two passes through one 32 KiB LDS tile (cleared on entry): real parts,
// stored to the cube's real halves, then imaginary parts; S is the fixed-point
// scale (FX).
template <int WIN, int NT, bool FX>
__device__ __forceinline__ void grid_halves(unsigned long long* acc, const Geometry& g, long long w,
                                            int u0, int v0, int trows, int tcols, int sm_log,
                                            double norm, int S, const Entry32* entries,
                                            const float2* vis, long long beg, long long end,
                                            float2* cube)
{
    const int tid = threadIdx.x;
    const double scale = ldexp(1.0, S), inv = ldexp(1.0, -S);
    float* out = (float*)cube;
    for (long long e = beg + tid; e < end; e += NT)
        accumulate_half<WIN, 0, FX>(acc, g, u0, v0, norm, scale, entry_at(entries, vis, e));
    __syncthreads();
    for (int c = tid; c < kPlane; c += NT) {
        const FlushCell f = flush_cell(c, sm_log, g, w, u0, v0);
        if (f.row < trows && f.col < tcols) out[2 * f.off] = half_value<FX>(acc[f.a], inv);
    }
    __syncthreads();
    for (int c = tid; c < kPlane / 2; c += NT) ((ulonglong2*)acc)[c] = ulonglong2{0ull, 0ull};
    __syncthreads();
    for (long long e = beg + tid; e < end; e += NT)
        accumulate_half<WIN, 1, FX>(acc, g, u0, v0, norm, scale, entry_at(entries, vis, e));
    __syncthreads();
    for (int c = tid; c < kPlane; c += NT) {
        const FlushCell f = flush_cell(c, sm_log, g, w, u0, v0);
        if (f.row < trows && f.col < tcols) out[2 * f.off + 1] = half_value<FX>(acc[f.a], inv);
    }
}

template <int WIN, int NT>
__global__ __launch_bounds__(NT) void k_grid_tiles_pk(Geometry g, TileMap tm, const long long* offsets,
                                                       const Entry32* entries, const float2* vis,
                                                       float2* cube, unsigned long long* occ,
                                                       int sm_log)
{
    __shared__ __align__(16) unsigned long long acc[kPlane];
    __shared__ unsigned long long bsum[kNBlk * kNBlk];
    __shared__ float wred[NT / 64];
    __shared__ double wsum[NT / 64];
    __shared__ int wfin[NT / 64];
    __shared__ int dense_blk, tile_ex;
    long long w;
    int tv, tu;
    const long long bkt = tm.launch_bucket(blockIdx.x, g, w, tv, tu);
    if (bkt < 0) return;
    const int u0 = tu * kTile, v0 = tv * kTile;
    const int tcols = min(kTile, g.nx - u0), trows = min(kTile, g.ny - v0);
    const long long beg = offsets[bkt], end = offsets[bkt + 1];
    const int tid = threadIdx.x;
    const double norm = 0.5 / LOFAR_REF_PI;

    if (occ) {
        // fused imaging: set the tile's occupancy bit (as k_grid_tiles_fx); an
        // empty tile is left unwritten
        if (beg == end) return;
        if (tid == 0)
            atomicOr(&occ[(w * g.ntu + tu) * ((g.ntv + 63) / 64) + tv / 64], 1ull << (tv % 64));
    }
    if (beg == end) {
        for (int c = tid; c < kPlane; c += NT) {
            const FlushCell f = flush_cell(c, sm_log, g, w, u0, v0);
            if (f.row < trows && f.col < tcols) cube[f.off] = float2{0.0f, 0.0f};
        }
        return;
    }
    // the first 2*NT entries are loaded once and kept in registers
    const long long n = end - beg;
    const int lane = tid & 63, wave = tid >> 6;
    Entry32 en0 = {0u, 0u, 0.0f, 0.0f}, en1 = {0u, 0u, 0.0f, 0.0f};
    if (tid < n) en0 = entry_at(entries, vis, beg + tid);
    if (NT + tid < n) en1 = entry_at(entries, vis, beg + NT + tid);
    for (int c = tid; c < kPlane / 2; c += NT)  // 16 bytes a store
        ((ulonglong2*)acc)[c] = ulonglong2{0ull, 0ull};  // also 0.0 as double
    for (int blk = tid; blk < kNBlk * kNBlk; blk += NT) bsum[blk] = 0ull;
    if (tid == 0) {
        dense_blk = 0;
        tile_ex = -1074;
    }
    float m = fmaxf(fmaxf(fabsf(en0.vr), fabsf(en0.vi)), fmaxf(fabsf(en1.vr), fabsf(en1.vi)));
    double sa = ((double)fabsf(en0.vr) + (double)fabsf(en0.vi)) +
                ((double)fabsf(en1.vr) + (double)fabsf(en1.vi));
    bool finite = isfinite(en0.vr) && isfinite(en0.vi) && isfinite(en1.vr) && isfinite(en1.vi);
    for (long long e = beg + 2 * NT + tid; e < end; e += NT) {
        const Entry32 en = entry_at(entries, vis, e);
        finite = finite && isfinite(en.vr) && isfinite(en.vi);
        m = fmaxf(m, fmaxf(fabsf(en.vr), fabsf(en.vi)));
        sa += (double)fabsf(en.vr) + (double)fabsf(en.vi);
    }
    for (int off = 32; off > 0; off >>= 1) {
        m = fmaxf(m, __shfl_xor(m, off));
        sa += __shfl_xor(sa, off);
    }
    const bool wave_finite = __all(finite);
    if (lane == 0) {
        wred[wave] = m;
        wsum[wave] = sa;
        wfin[wave] = wave_finite ? 1 : 0;
    }
    __syncthreads();
    m = wred[0];
    sa = wsum[0];
    bool all_finite = wfin[0] != 0;
#pragma unroll
    for (int i = 1; i < NT / 64; i++) {
        m = fmaxf(m, wred[i]);
        sa += wsum[i];
        all_finite = all_finite && wfin[i] != 0;
    }
    if (!all_finite || (double)m > kHdrRatio * (sa - (double)m) / (double)(2 * n - 1)) {
        grid_halves<WIN, NT, false>(acc, g, w, u0, v0, trows, tcols, sm_log, norm, 0, entries,
                                    vis, beg, end, cube);
        return;
    }
    // the two-word sums also take tiles too large for the block counts and tiles
    // whose largest component is too small for 2^20/m in float
    bool dense = n >= kPkMaxEntries || m < 0x1p-100f;
    if (!dense) {
        // block counts and bounds; the largest bound sets the tile's exponent
        const float rm = kBndUnit / m;
        if (tid < n) add_block_bound<WIN>(bsum, g, u0, v0, rm, en0);
        if (NT + tid < n) add_block_bound<WIN>(bsum, g, u0, v0, rm, en1);
        for (long long e = beg + 2 * NT + tid; e < end; e += NT)
            add_block_bound<WIN>(bsum, g, u0, v0, rm, entry_at(entries, vis, e));
        __syncthreads();
        for (int blk = tid; blk < kNBlk * kNBlk; blk += NT) {
            const unsigned long long b = bsum[blk];
            if (b == 0ull) continue;  // no entry touches the block: it bounds nothing
            int ex = 0;
            frexp((double)(b & ((1ull << kCntShift) - 1)) * (norm * (double)m / (double)kBndUnit),
                  &ex);  // bound < 2^ex
            atomicMax(&tile_ex, ex);
            if ((b >> kCntShift) > (unsigned long long)kPkMaxBlk) dense_blk = 1;
        }
        __syncthreads();
        dense = dense_blk != 0;
    }
    if (dense) {
        int ex = 0;
        frexp(norm * (double)(m > 0.0f ? m : 1.0f), &ex);  // norm*m < 2^ex
        int hb = 2;
        for (long long cnt = n; cnt > 1; cnt >>= 1) hb++;
        grid_halves<WIN, NT, true>(acc, g, w, u0, v0, trows, tcols, sm_log, norm,
                                   kFixedBits - hb - ex, entries, vis, beg, end, cube);
        return;
    }
    const int S = kPkHead - tile_ex;
    const double scale = ldexp(1.0, S), inv = ldexp(1.0, -S);
    if (tid < n) accumulate_pk32<WIN>(acc, g, u0, v0, norm, scale, en0);
    if (NT + tid < n) accumulate_pk32<WIN>(acc, g, u0, v0, norm, scale, en1);
    for (long long e = beg + 2 * NT + tid; e < end; e += NT)
        accumulate_pk32<WIN>(acc, g, u0, v0, norm, scale, entry_at(entries, vis, e));
    __syncthreads();
    for (int c = tid; c < kPlane; c += NT) {
        const FlushCell f = flush_cell(c, sm_log, g, w, u0, v0);
        if (f.row < trows && f.col < tcols) {
            const unsigned long long t = acc[f.a];
            const int im = (int)(unsigned)t;
            const long long re = (long long)(t - (unsigned long long)(long long)im) >> 32;
            cube[f.off] = float2{(float)((double)re * inv), (float)((double)im * inv)};
        }
    }
}

// FP64 gridding with exact fixed-point accumulation: as k_grid_tiles_fx, with
// double visibilities and a 62-bit range (S = 62 - hb - ex); every contribution
// eu * (norm*ev*vis*2^S) is rounded once, from the exact product, to the nearest
// integer (fixed62_of_product), and the integer sums are exact and independent
// of order.  The quantum is ~2^-(60-log2(entries)) of the tile's largest
// contribution and the rounding is unbiased, so the cells agree with a
// double-precision sum to ~1e-15 of the tile's largest contribution.  The
// Gaussian factors come from gauss_row (two exponentials per axis).
constexpr int kFixedBits64 = 62;

// round(a*b) (nearest, ties to even, of the exact product) as a two's-complement
// 64-bit value, |a*b| < 2^62, in three full-rate double operations and one 64-bit
// shift-add:
//   t1 = fma(a, b, C1), C1 = 1.5*2^84 - 0x43380000*2^32: t1 stays in [2^84, 2^85),
//        so ulp(t1) = 2^32 and t1 = C1 + k*2^32 with k = round(a*b / 2^32); the low
//        word of t1 is k - 0x43380000 (mod 2^32);
//   c  = (C1 + C2) - t1 = C2 - k*2^32, exact (Sterbenz), C2 = 1.5*2^52;
//   t2 = fma(a, b, c): ulp(t2) = 1, so t2 = C2 + round(a*b - k*2^32), one rounding;
//        t2's bits are C2's bits (high word 0x43380000, low word 0) plus that
//        remainder as a 64-bit integer.
// (t1 bits << 32) + (t2 bits) = remainder + k*2^32, the high words' 0x43380000
// cancelling; as k*2^32 has a zero low word this is one 32-bit add to t2's high
// word (written on a two-word vector so that it stays a 32-bit add).  C1 and
// C1 + C2 are exact doubles.
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ unsigned long long fixed62_of_product(double a, double b)
{
    const double t1 = fma(a, b, 0x1.7ffffbcc80000p+84);
    const double t2 = fma(a, b, 0x1.7ffffbce00000p+84 - t1);
    u32x2 r = __builtin_bit_cast(u32x2, t2);
    r.y += __builtin_bit_cast(u32x2, t1).x;
    return __builtin_bit_cast(unsigned long long, r);
}

// A fixed-point sum back to double, times 2^-S (exact scaling, one rounding)
__device__ __forceinline__ double from_fixed(unsigned long long x, double inv)
{
    const double hi = (double)(int)(x >> 32), lo = (double)(unsigned)x;
    return fma(hi, 0x1p32 * inv, lo * inv);
}

// exp(-(q + 1/2)), q = 0..15, correctly rounded
__constant__ double kGaussStep[16] = {
    0x1.368b2fc6f960ap-1, 0x1.c8f87724b5c1dp-3, 0x1.50385c094f425p-4, 0x1.eec1018e4ff66p-6,
    0x1.6c0504695c417p-7, 0x1.0bd4a5aca7728p-8, 0x1.8a1e18a954ffap-10, 0x1.21f9ba40f31d5p-11,
    0x1.aab43fb3940a1p-13, 0x1.39f38a5498fe2p-14, 0x1.cdfc263f6a0bap-16, 0x1.53e8d76780b04p-17,
    0x1.f42ed3f68e690p-19, 0x1.7003974e945dfp-20, 0x1.0ec503d2696e8p-21, 0x1.8e7138077168dp-23};

// Gaussian factors g[q] = exp(-(d0 + q)^2 / 2), q < WIN, from two exponentials:
// g[q+1] = g[q] * exp(-d0) * exp(-(q + 1/2)); each step adds two roundings, so
// g[q] is within ~2q ulp of the directly evaluated exponential.
template <int WIN>
__device__ __forceinline__ void gauss_row(double d0, double (&gq)[WIN])
{
    double x = exp(-(d0 * d0) * 0.5);
    const double b = exp(-d0);
    gq[0] = x;
#pragma unroll
    for (int q = 1; q < WIN; q++) {
        x *= b * kGaussStep[q - 1];
        gq[q] = x;
    }
}

// One entry into the fixed-point tile (real parts, then imaginary parts kPlane
// words on): cell (a, b) of the entry's clipped window gets
// round(eu[a] * norm * ev[b] * vis * 2^S).  Branch-free: cells of the WIN x WIN
// square at the window's first cell that lie outside the clipped window get a
// zero factor, whose fixed-point value is exactly 0, so every lane adds the same
// cells; the LDS tile is padded so that those addresses stay inside it.
template <int WIN>
__device__ __forceinline__ void accumulate_fx64(unsigned long long* acc, const Geometry& g, int u0,
                                                int v0, double norm, double scale,
                                                const Entry64& en)
{
    const TileWin tw = clip_window((int)en.pu, (int)en.pv, g, u0, v0);
    double eu[WIN], ev[WIN];
    gauss_row<WIN>((double)(u0 + tw.a0) + 0.5 - en.pu, eu);
    gauss_row<WIN>((double)(v0 + tw.b0) + 0.5 - en.pv, ev);
#pragma unroll
    for (int q = 0; q < WIN; q++) {
        eu[q] = (q <= tw.na) ? eu[q] : 0.0;
        ev[q] = (q <= tw.nb) ? ev[q] : 0.0;
    }
    unsigned long long* base = &acc[tw.b0 * kTile + tw.a0];
#pragma unroll
    for (int kq = 0; kq < WIN; kq++) {
        const double f = norm * ev[kq];
        const double fr = f * en.vr * scale, fi = f * en.vi * scale;
        unsigned long long* row = base + kq * kTile;
#pragma unroll
        for (int q = 0; q < WIN; q++) {
            atomicAdd(row + q, fixed62_of_product(eu[q], fr));
            atomicAdd(row + kPlane + q, fixed62_of_product(eu[q], fi));
        }
    }
}

// The same entry in floating point (ds_add_f64), for tiles holding a non-finite
// visibility: direct exponentials, as k_grid_tiles<WIN, true>.
template <int WIN>
__device__ __forceinline__ void accumulate_f64(double* acc, const Geometry& g, int u0, int v0,
                                               double norm, const Entry64& en)
{
    const TileWin tw = clip_window((int)en.pu, (int)en.pv, g, u0, v0);
    double eu[WIN];
#pragma unroll
    for (int q = 0; q < WIN; q++) {
        const double du = (double)(u0 + tw.a0 + q) + 0.5 - en.pu;
        eu[q] = (q <= tw.na) ? exp(-(du * du) * 0.5) : 0.0;
    }
    double* base = &acc[tw.b0 * kTile + tw.a0];
#pragma unroll
    for (int kq = 0; kq < WIN; kq++) {
        if (kq <= tw.nb) {
            const double dv = (double)(v0 + tw.b0 + kq) + 0.5 - en.pv;
            const double f = norm * exp(-(dv * dv) * 0.5);
            const double fr = f * en.vr, fi = f * en.vi;
            double* row = base + kq * kTile;
#pragma unroll
            for (int q = 0; q < WIN; q++) {
                if (q <= tw.na) {
                    atomicAdd(row + q, eu[q] * fr);
                    atomicAdd(row + kPlane + q, eu[q] * fi);
                }
            }
        }
    }
}

// One 512-thread workgroup per (w-plane, tile), as k_grid_tiles_fx.
template <int WIN, int NT>
__global__ __launch_bounds__(NT) void k_grid_tiles_fx64(Geometry g, TileMap tm,
                                                         const long long* offsets,
                                                         const Entry64* entries, double2* cube,
                                                         unsigned long long* occ, int sm_log)
{
    // the WIN x WIN square at a window's first cell reaches (WIN-1)*(kTile+1) words
    // past the tile's last cell (accumulate_fx64)
    constexpr int kPad = ((WIN - 1) * (kTile + 1) + 1) / 2 * 2;
    __shared__ __align__(16) unsigned long long acc[2 * kPlane + kPad];
    __shared__ double wred[NT / 64];
    __shared__ double wsum[NT / 64];
    __shared__ int wfin[NT / 64];
    long long w;
    int tv, tu;
    const long long bkt = tm.launch_bucket(blockIdx.x, g, w, tv, tu);
    if (bkt < 0) return;
    const int u0 = tu * kTile, v0 = tv * kTile;
    const int tcols = min(kTile, g.nx - u0), trows = min(kTile, g.ny - v0);
    const long long beg = offsets[bkt], end = offsets[bkt + 1];
    const int tid = threadIdx.x;
    const double norm = 0.5 / LOFAR_REF_PI;

    if (occ) {
        if (beg == end) return;
        if (tid == 0)
            atomicOr(&occ[(w * g.ntu + tu) * ((g.ntv + 63) / 64) + tv / 64], 1ull << (tv % 64));
    }
    if (beg == end) {
        for (int c = tid; c < kPlane; c += NT) {
            const FlushCell f = flush_cell(c, sm_log, g, w, u0, v0);
            if (f.row < trows && f.col < tcols) cube[f.off] = double2{0.0, 0.0};
        }
        return;
    }
    // Entries are taken in rounds of NT; a round of r entries goes to ceil(r/32)
    // waves (at most all of them), so that a small tile still spreads over the
    // SIMDs.  The first round's entry is loaded once, before the LDS clear, and
    // kept in registers.
    const long long n = end - beg;
    const int lane = tid & 63, wave = tid >> 6;
    auto slot = [&](long long q0) -> long long {
        const long long r = min((long long)NT, n - q0);
        const int wu = (int)min((long long)(NT / 64), (r + 31) / 32);
        const int j = lane * wu + wave;
        return (wave < wu && j < r) ? q0 + j : -1;
    };
    const long long j0 = slot(0);
    Entry64 en0 = {0.0, 0.0, 0.0, 0.0};
    if (j0 >= 0) en0 = entries[beg + j0];
    for (int c = tid; c < kPlane; c += NT)  // 2*kPlane words, 16 bytes a store
        ((ulonglong2*)acc)[c] = ulonglong2{0ull, 0ull};  // also 0.0 as double
    double m = fmax(fabs(en0.vr), fabs(en0.vi));
    double sa = fabs(en0.vr) + fabs(en0.vi);
    bool finite = isfinite(en0.vr) && isfinite(en0.vi);
    for (long long q0 = NT; q0 < n; q0 += NT) {
        const long long j = slot(q0);
        if (j >= 0) {
            const Entry64 en = entries[beg + j];
            finite = finite && isfinite(en.vr) && isfinite(en.vi);
            m = fmax(m, fmax(fabs(en.vr), fabs(en.vi)));
            sa += fabs(en.vr) + fabs(en.vi);
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        m = fmax(m, __shfl_xor(m, off));
        sa += __shfl_xor(sa, off);
    }
    const bool wave_finite = __all(finite);
    if (lane == 0) {
        wred[wave] = m;
        wsum[wave] = sa;
        wfin[wave] = wave_finite ? 1 : 0;
    }
    __syncthreads();
    m = wred[0];
    sa = wsum[0];
    bool all_finite = wfin[0] != 0;
#pragma unroll
    for (int i = 1; i < NT / 64; i++) {
        m = fmax(m, wred[i]);
        sa += wsum[i];
        all_finite = all_finite && wfin[i] != 0;
    }
    // a bright outlier (largest |vis| far above the tile's mean): floating-point path
    if (!all_finite || m > kHdrRatio * (sa - m) / (double)(2 * n - 1)) {
        double* dacc = (double*)acc;
        if (j0 >= 0) accumulate_f64<WIN>(dacc, g, u0, v0, norm, en0);
        for (long long q0 = NT; q0 < n; q0 += NT) {
            const long long j = slot(q0);
            if (j >= 0) accumulate_f64<WIN>(dacc, g, u0, v0, norm, entries[beg + j]);
        }
        __syncthreads();
        for (int c = tid; c < kPlane; c += NT) {
            const FlushCell f = flush_cell(c, sm_log, g, w, u0, v0);
            if (f.row < trows && f.col < tcols)
                cube[f.off] = double2{dacc[f.a], dacc[kPlane + f.a]};
        }
        return;
    }
    int ex = 0;
    frexp(norm * (m > 0.0 ? m : 1.0), &ex);  // norm*m < 2^ex
    int hb = 2;
    for (long long cnt = n; cnt > 1; cnt >>= 1) hb++;
    const int S = kFixedBits64 - hb - ex;
    const double scale = ldexp(1.0, S);

    if (j0 >= 0) accumulate_fx64<WIN>(acc, g, u0, v0, norm, scale, en0);
    for (long long q0 = NT; q0 < n; q0 += NT) {
        const long long j = slot(q0);
        if (j >= 0) accumulate_fx64<WIN>(acc, g, u0, v0, norm, scale, entries[beg + j]);
    }
    __syncthreads();
    const double inv = ldexp(1.0, -S);
    for (int c = tid; c < kPlane; c += NT) {
        const FlushCell f = flush_cell(c, sm_log, g, w, u0, v0);
        if (f.row < trows && f.col < tcols)
            cube[f.off] = double2{from_fixed(acc[f.a], inv), from_fixed(acc[kPlane + f.a], inv)};
    }
}

// ---------------------------------------------------------------------------
// degridding (adjoint of gridding)

// One workgroup per home tile: the tile plus a halo of (WIN-1)/2 cells is staged
// in LDS; one lane per sample gathers its window (same clipping and weights as
// the gridder) and writes weight*sum for every channel and polarisation.
template <int WIN, bool FP64>
__global__ __launch_bounds__(256) void k_degrid_tiles(Geometry g, TileMap tm, const long long* offsets,
                                                      const DEntry* ents, const void* cube,
                                                      const float* weight, long long F,
                                                      long long P, void* mre, void* mim)
{
    using Real = typename std::conditional<FP64, double, float>::type;
    using T2 = typename cplx<Real>::type;
    constexpr int H = (WIN - 1) / 2;
    constexpr int RS = kTile + 2 * H;
    __shared__ T2 region[RS * RS];
    long long w;
    int tv, tu;
    const long long bkt = tm.launch_bucket(blockIdx.x, g, w, tv, tu);
    if (bkt < 0) return;
    const int u0 = tu * kTile, v0 = tv * kTile;
    const long long beg = offsets[bkt], end = offsets[bkt + 1];
    if (beg == end) return;
    const T2* plane = (const T2*)cube + w * (long long)g.ny * g.nx;
    const int tid = threadIdx.x;
    // every load of the thread's part of the region is issued before the first LDS
    // store (a load-store loop keeps one load in flight per thread)
    constexpr int NLD = (RS * RS + 255) / 256;
    T2 xs[NLD];
#pragma unroll
    for (int k = 0; k < NLD; k++) {
        const int c = tid + k * 256;
        const int rr = c / RS, cc = c % RS;
        const int gv = v0 - H + rr, gu = u0 - H + cc;
        xs[k] = T2{0, 0};
        if (c < RS * RS && gv >= 0 && gv < g.ny && gu >= 0 && gu < g.nx)
            xs[k] = plane[(long long)gv * g.nx + gu];
    }
#pragma unroll
    for (int k = 0; k < NLD; k++) {
        const int c = tid + k * 256;
        if (c < RS * RS) region[c] = xs[k];
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
        // sums in double in both modes: the window's terms cancel heavily for
        // noise-like cubes, and float sums would cost ~1e-6 of the result
        double sr = 0, si = 0;
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
                    double rr = 0, ri = 0;
                    const T2* row = base + kq * RS;
#pragma unroll
                    for (int q = 0; q < WIN; q++) {
                        if (q <= na) {
                            const T2 x = row[q];
                            rr += (double)eu[q] * (double)x.x;
                            ri += (double)eu[q] * (double)x.y;
                        }
                    }
                    sr += (double)f * rr;
                    si += (double)f * ri;
                }
            }
        }
        const long long i = en.idx;
        for (long long p = 0; p < P; p++) {
            // one polarisation: the weight rides in the entry (no read in caller order)
            const Real wgt = (Real)(P == 1 ? __uint_as_float((unsigned)en.pad) : weight[i * P + p]);
            for (long long f = 0; f < F; f++) {
                const long long o = i * F * P + f * P + p;
                ((Real*)mre)[o] = wgt * sr;
                ((Real*)mim)[o] = wgt * si;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// bucket-free kernels for windows wider than the tile kernels take

__global__ __launch_bounds__(256) void k_validate(long long n, const double* uu, const double* vv,
                                                  const double* ww, Geometry g, int* err)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    SampleWin s;
    if (!sample_window(uu[i], vv[i], ww[i], g, s)) atomicOr(err, 1);
}

// One thread per sample adds weight*vis*norm*exp(-(du^2+dv^2)/2) to every cell of its
// clipped window with global atomic adds, as the reference's OpenMP loop does
// (w-stacking.cu:222-307, `omp atomic` on each cell); the order of the sums is not fixed.
template <typename Real>
__global__ __launch_bounds__(256) void k_grid_direct(long long n, long long F, long long P,
                                                     const double* uu, const double* vv,
                                                     const double* ww, const float* vr,
                                                     const float* vi, const float* wt,
                                                     Geometry g, Real* cube)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    SampleWin s;
    if (!sample_window(uu[i], vv[i], ww[i], g, s) || s.j0 > s.j1 || s.k0 > s.k1) return;
    double sr, si;
    sample_vis(i, F, P, vr, vi, wt, sr, si);
    const double norm = 0.5 / LOFAR_REF_PI;
    Real* plane = cube + 2 * (long long)s.gw * g.ny * g.nx;
    for (int k = s.k0; k <= s.k1; k++) {
        const double dv = (double)k + 0.5 - s.pv;
        for (int j = s.j0; j <= s.j1; j++) {
            const double du = (double)j + 0.5 - s.pu;
            const double cw = norm * exp(-(du * du + dv * dv) * 0.5);
            Real* cell = plane + 2 * ((long long)k * g.nx + j);
            atomicAdd(cell, (Real)(sr * cw));
            atomicAdd(cell + 1, (Real)(si * cw));
        }
    }
}

// One thread per sample: weight * sum over its clipped window of cw * cube (in double),
// for every channel and polarisation (the adjoint of k_grid_direct).
template <typename Real>
__global__ __launch_bounds__(256) void k_degrid_direct(long long n, long long F, long long P,
                                                       const double* uu, const double* vv,
                                                       const double* ww, const float* weight,
                                                       Geometry g, const Real* cube, Real* mre,
                                                       Real* mim)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    SampleWin s;
    if (!sample_window(uu[i], vv[i], ww[i], g, s)) return;  // rejected before the launch
    const double norm = 0.5 / LOFAR_REF_PI;
    const Real* plane = cube + 2 * (long long)s.gw * g.ny * g.nx;
    double sr = 0.0, si = 0.0;
    for (int k = s.k0; k <= s.k1; k++) {
        const double dv = (double)k + 0.5 - s.pv;
        for (int j = s.j0; j <= s.j1; j++) {
            const double du = (double)j + 0.5 - s.pu;
            const double cw = norm * exp(-(du * du + dv * dv) * 0.5);
            const Real* cell = plane + 2 * ((long long)k * g.nx + j);
            sr += cw * (double)cell[0];
            si += cw * (double)cell[1];
        }
    }
    for (long long p = 0; p < P; p++) {
        const double wgt = (double)weight[i * P + p];
        for (long long f = 0; f < F; f++) {
            const long long o = i * F * P + f * P + p;
            mre[o] = (Real)(wgt * sr);
            mim[o] = (Real)(wgt * si);
        }
    }
}

// Tiles a degridding reads: every tile holding home entries and its 8 neighbours (the
// window's halo, K <= 8 cells, stays inside them), as bit tv%64 of word (w, tu, tv/64),
// the indexing of grid_tiles' occupancy bits.  The words are cleared beforehand.
__global__ __launch_bounds__(256) void k_need_tiles(Geometry g, TileMap tm,
                                                    const long long* offsets,
                                                    unsigned long long* need)
{
    const long long bkt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (bkt >= tm.nbuckets()) return;
    long long w;
    int tv, tu;
    if (!tm.decode(bkt, g, w, tv, tu) || offsets[bkt] == offsets[bkt + 1]) return;
    const int nwd = (g.ntv + 63) / 64;
    for (int u = max(tu - 1, 0); u <= min(tu + 1, g.ntu - 1); u++)
        for (int v = max(tv - 1, 0); v <= min(tv + 1, g.ntv - 1); v++)
            atomicOr(&need[(w * g.ntu + u) * nwd + v / 64], 1ull << (v % 64));
}

// ---------------------------------------------------------------------------
// host dispatch

// Accumulation without accumulate: 0 = exact fixed point (default: FP64 two
// 64-bit words per cell, FP32C one packed word), 1 = ds_add_f64, 2 = FP32C with
// two 64-bit words per cell (LOFAR_GRID_CFG overrides, for measurement)
static int grid_cfg()
{
    static int cfg = -1;
    if (cfg < 0) {
        const char* s = getenv("LOFAR_GRID_CFG");
        cfg = s ? atoi(s) : 0;
    }
    return cfg;
}

// threads per gridding workgroup: LOFAR_GRID_NT (256, 512 or 1024) overrides, for
// measurement.  Otherwise 512, and 256 for FP32C tiles averaging fewer than 256 entries
// (their per-tile phases, not the LDS atomics, set the time: 8192x8192x128 on 256M samples
// grids ~7 ms faster per image; 590 entries a tile ran 2.61 -> 2.83 ms slower at 256, and the
// FP64 gridder at ~134 entries a tile 2.02 -> 2.12 ms).
static int grid_nt(int precision, const Buckets& b)
{
    static const int forced = [] {
        const char* s = getenv("LOFAR_GRID_NT");
        const int v = s ? atoi(s) : 0;
        return (v == 256 || v == 512 || v == 1024) ? v : 0;
    }();
    if (forced) return forced;
    if (precision != LOFAR_FP64 && b.nent > 0 && b.nent < 256 * b.map.nbuckets()) return 256;
    return 512;
}

template <int WIN, int NT>
static void launch_grid_nt(int precision, const Geometry& g, const Buckets& b, const void* entries,
                           void* cube, int accumulate, unsigned long long* occ, int sm_log,
                           hipStream_t st)
{
    const dim3 grid((unsigned)b.map.nbuckets());
    const int cfg = grid_cfg();
    if (precision == LOFAR_FP64 && !accumulate && cfg != 1)
        hipLaunchKernelGGL((k_grid_tiles_fx64<WIN, NT>), grid, dim3(NT), 0, st, g, b.map,
                           b.offsets, (const Entry64*)entries, (double2*)cube, occ, sm_log);
    else if (precision == LOFAR_FP64)
        hipLaunchKernelGGL((k_grid_tiles<WIN, true, NT>), grid, dim3(NT), 0, st, g, b.map,
                           b.offsets, entries, cube, accumulate, occ, sm_log);
    else if (!accumulate && cfg == 0)
        hipLaunchKernelGGL((k_grid_tiles_pk<WIN, NT>), grid, dim3(NT), 0, st, g, b.map, b.offsets,
                           (const Entry32*)entries, (const float2*)b.vis, (float2*)cube, occ,
                           sm_log);
    else if (!accumulate && cfg == 2)
        hipLaunchKernelGGL((k_grid_tiles_fx<WIN, NT>), grid, dim3(NT), 0, st, g, b.map, b.offsets,
                           (const Entry32*)entries, (float2*)cube, occ, sm_log);
    else
        hipLaunchKernelGGL((k_grid_tiles<WIN, false, NT>), grid, dim3(NT), 0, st, g, b.map,
                           b.offsets, entries, cube, accumulate, occ, sm_log);
}

template <int WIN>
static void launch_grid(int precision, const Geometry& g, const Buckets& b, const void* entries,
                        void* cube, int accumulate, unsigned long long* occ, int sm_log,
                        hipStream_t st)
{
    switch (grid_nt(precision, b)) {
        case 1024: return launch_grid_nt<WIN, 1024>(precision, g, b, entries, cube, accumulate, occ, sm_log, st);
        case 256: return launch_grid_nt<WIN, 256>(precision, g, b, entries, cube, accumulate, occ, sm_log, st);
        default: return launch_grid_nt<WIN, 512>(precision, g, b, entries, cube, accumulate, occ, sm_log, st);
    }
}

long long need_words(const Geometry& g) { return (long long)g.nw * g.ntu * ((g.ntv + 63) / 64); }

hipError_t need_tiles(const Geometry& g, const Buckets& b, unsigned long long* need,
                      hipStream_t st)
{
    hipError_t e = fill_u32((unsigned*)need, 2 * need_words(g), 0u, st);
    if (e != hipSuccess) return e;
    const long long nb = b.map.nbuckets();
    if (nb > 0)
        hipLaunchKernelGGL(k_need_tiles, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, g,
                           b.map, b.offsets, need);
    return hipGetLastError();
}

hipError_t validate_samples(long long n, const double* uu, const double* vv, const double* ww,
                            const Geometry& g, int* err, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_u32_1, dim3(1), dim3(64), 0, st, (unsigned*)err);
    if (n > 0)
        hipLaunchKernelGGL(k_validate, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, uu,
                           vv, ww, g, err);
    return hipGetLastError();
}

hipError_t grid_direct(int precision, long long n, long long F, long long P, const double* uu,
                       const double* vv, const double* ww, const float* vr, const float* vi,
                       const float* wt, const Geometry& g, void* cube, int accumulate,
                       hipStream_t st)
{
    const long long words = 2LL * g.nw * g.ny * g.nx * (precision == LOFAR_FP64 ? 2 : 1);
    if (!accumulate) {
        hipError_t e = fill_u32((unsigned*)cube, words, 0u, st);
        if (e != hipSuccess) return e;
    }
    if (n <= 0) return hipGetLastError();
    const dim3 blocks((unsigned)((n + 255) / 256));
    if (precision == LOFAR_FP64)
        hipLaunchKernelGGL((k_grid_direct<double>), blocks, dim3(256), 0, st, n, F, P, uu, vv, ww,
                           vr, vi, wt, g, (double*)cube);
    else
        hipLaunchKernelGGL((k_grid_direct<float>), blocks, dim3(256), 0, st, n, F, P, uu, vv, ww,
                           vr, vi, wt, g, (float*)cube);
    return hipGetLastError();
}

hipError_t degrid_direct(int precision, long long n, long long F, long long P, const double* uu,
                         const double* vv, const double* ww, const float* wt, const Geometry& g,
                         const void* cube, void* mre, void* mim, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    const dim3 blocks((unsigned)((n + 255) / 256));
    if (precision == LOFAR_FP64)
        hipLaunchKernelGGL((k_degrid_direct<double>), blocks, dim3(256), 0, st, n, F, P, uu, vv,
                           ww, wt, g, (const double*)cube, (double*)mre, (double*)mim);
    else
        hipLaunchKernelGGL((k_degrid_direct<float>), blocks, dim3(256), 0, st, n, F, P, uu, vv,
                           ww, wt, g, (const float*)cube, (float*)mre, (float*)mim);
    return hipGetLastError();
}

bool grid_gathers(int precision, int accumulate)
{
    return precision != LOFAR_FP64 && !accumulate && grid_cfg() == 0;
}

hipError_t grid_tiles(int precision, const Geometry& g, const Buckets& b, const void* entries,
                      void* cube, int accumulate, unsigned long long* occ, int sm_log,
                      hipStream_t st)
{
    // strip-major flushes need whole strips no wider than a tile
    if (sm_log >= 0 && (accumulate || (1 << sm_log) > kTile || g.nx % (1 << sm_log) != 0))
        return hipErrorInvalidValue;
    // only the packed FP32C kernel gathers visibilities
    if (b.vis && !grid_gathers(precision, accumulate)) return hipErrorInvalidValue;
    const int win = 2 * g.K + 1;
    if (win <= 7)
        launch_grid<7>(precision, g, b, entries, cube, accumulate, occ, sm_log, st);
    else if (win <= 9)
        launch_grid<9>(precision, g, b, entries, cube, accumulate, occ, sm_log, st);
    else if (win <= kMaxTileWin)
        launch_grid<17>(precision, g, b, entries, cube, accumulate, occ, sm_log, st);
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
        hipLaunchKernelGGL((k_degrid_tiles<WIN, true>), dim3((unsigned)b.map.nbuckets()), dim3(256), 0, st, g,
                           b.map, b.offsets, ents, cube, weight, F, P, mre, mim);
    else
        hipLaunchKernelGGL((k_degrid_tiles<WIN, false>), dim3((unsigned)b.map.nbuckets()), dim3(256), 0, st, g,
                           b.map, b.offsets, ents, cube, weight, F, P, mre, mim);
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
    else if (win <= kMaxTileWin)
        launch_degrid<17>(precision, g, b, ents, cube, weight, F, P, mre, mim, st);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace lofar
