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
// Largest Gaussian window (2K+1) handled by the unrolled gridding kernels.
constexpr int kMaxWin = 9;

template <typename T> struct cplx;
template <> struct cplx<double> { using type = double2; };
template <> struct cplx<float> { using type = float2; };

struct Geometry {
    int nx, ny, nw;     // grid and w-plane counts
    int K;              // kernel half-width, (w_support-1)/2
    int ntu, ntv;       // tiles along u and v
    double dx, dw;      // cell size in normalised uv units, plane width in w units
};

// Packs a per-tile clipped window [a0,a1] x [b0,b1] (tile-local, each < 64)
// and, in the top byte, the tile's index inside its super-tile (bucket sort).
__host__ __device__ inline uint32_t pack_win(int a0, int a1, int b0, int b1, int fine = 0)
{
    return (uint32_t)a0 | ((uint32_t)a1 << 6) | ((uint32_t)b0 << 12) | ((uint32_t)b1 << 18) |
           ((uint32_t)fine << 24);
}

// Gridding entry, FP64 mode: absolute grid position (uu/dx, vv/dx) and the
// visibility summed over channels and polarisations with the reference's
// weighting (w-stacking_omp.c:118-133).
struct alignas(16) Entry64 {
    double pu, pv;
    double vr, vi;
};

// Gridding entry, FP32 mode: distances from the first window cell of the tile
// (j0+0.5-pos_u, k0+0.5-pos_v) and the weighted visibility.
struct alignas(16) Entry32 {
    float du0, dv0;
    float vr, vi;
};

// Degridding entry: position and the sample's index in the caller's order.
struct alignas(16) DEntry {
    double pu, pv;
    long long idx;
    long long pad;
};

}  // namespace lofar
