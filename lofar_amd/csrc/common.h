// Shared definitions for the gfx950 w-stacking kernels and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/lofar_hip.h"

// w-stacking.h:6 -- the reference's value of pi, used by the Gaussian norm and
// the w-phase argument.
#define LOFAR_REF_PI 3.14159265359

namespace lofar {

// Side of the square uv tile one workgroup owns while gridding.
constexpr int kTile = 64;
// Largest Gaussian window (2K+1) the tile kernels take; wider windows use the
// bucket-free direct kernels (grid_direct, degrid_direct).
constexpr int kMaxTileWin = 17;

template <typename T> struct cplx;
template <> struct cplx<double> { using type = double2; };
template <> struct cplx<float> { using type = float2; };

struct Geometry {
    int nx, ny, nw;     // grid and w-plane counts
    int K;              // kernel half-width, (w_support-1)/2
    int ntu, ntv;       // tiles along u and v
    double dx, dw;      // cell size in normalised uv units, plane width in w units
};

// Gridding entry, FP64 mode (and the level-1 record of both modes): absolute
// grid position (uu/dx, vv/dx) and the visibility summed over channels and
// polarisations with the reference's weighting (w-stacking_omp.c:118-133).
// The gridder derives the window and its clipping to the tile from the position.
struct alignas(32) Entry64 {
    double pu, pv;
    double vr, vi;
};

// Gridding entry, FP32 mode: per axis, the sample's cell relative to the tile
// origin (gu-u0+16, 8 bits) and its offset within the cell (pu-gu+1 in
// (0, 2), 23 fraction bits) packed in one word; the weighted visibility as float.
struct alignas(16) Entry32 {
    unsigned wu, wv;
    float vr, vi;
};

constexpr int kCellBias = 16;  // gu-u0 lies in [-K, 63+K] for every entry of a tile, K <= 8

__host__ __device__ inline unsigned pack_cell(int cell_rel, double frac)
{
    double x = (frac + 1.0) * 8388608.0;  // 2^23
    x = x < 0.0 ? 0.0 : (x > 16777215.0 ? 16777215.0 : x);
    return ((unsigned)(cell_rel + kCellBias) << 24) | (unsigned)(long long)(x + 0.5);
}
__host__ __device__ inline int cell_of(unsigned w) { return (int)(w >> 24) - kCellBias; }
__host__ __device__ inline float frac_of(unsigned w)
{
    return (float)(w & 0xffffffu) * 0x1p-23f - 1.0f;
}

// Degridding entry: position, the sample's index in the caller's order and (with one
// polarisation) its weight's bits in `pad`.
struct alignas(32) DEntry {
    double pu, pv;
    long long idx;
    long long pad;
};

// Workgroup barrier that orders LDS accesses only: it waits for this wave's LDS
// operations (lgkmcnt 0) but not for its global memory operations, so global
// loads (direct-to-LDS loads into another buffer among them) and stores stay in
// flight across it.  The empty asm statements keep
// the compiler from moving memory accesses across the barrier.
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xC07F);  // gfx9 encoding: lgkmcnt(0), vmcnt/expcnt at max
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// Wait until this wave's global loads (including direct-to-LDS loads) are done.
__device__ __forceinline__ void wait_loads()
{
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0), lgkmcnt/expcnt at max
    asm volatile("" ::: "memory");
}

// number of compute units of the current device (cached; persistent-grid sizing)
int device_cus();

}  // namespace lofar
