// Batched per-plane 2D FFT for gfx950: Stockham autosort, radix-16 stages with
// a radix-2/4/8 remainder, one or several lines per workgroup staged in LDS.
//
// Semantics follow fftw_data() (fourier_transform.c:7-74): per w-plane an
// unnormalised FFTW_BACKWARD transform over dims (n0=ny, n1=nx) of the [v][u]
// plane, scaled by 1/(nx*ny).  Column transforms run first, then rows, so the
// last pass (rows, contiguous) can fuse the w-phase screen and the plane sum of
// phase_correction() (phase_correction.cu:151-210) and the image is written once.
#include "fft_kernels.h"

namespace lofar {

// ---------------------------------------------------------------------------
// standalone passes

// Row transforms: one line (row) per workgroup, N/E threads.
template <typename Real, int N, int SIGN>
__global__ __launch_bounds__(Plan<N>::TPL) void k_fft_rows(const typename cplx<Real>::type* in,
                                                           typename cplx<Real>::type* out,
                                                           const typename cplx<Real>::type* tw,
                                                           Real scale)
{
    using T2 = typename cplx<Real>::type;
    constexpr int E = Plan<N>::E;
    __shared__ T2 lds[N];
    const long long line = blockIdx.x;
    const T2* src = in + line * N;
    T2* dst = out + line * N;
    const int t = threadIdx.x;
    T2 v[E];
#pragma unroll
    for (int q = 0; q < E; q++) v[q] = src[Plan<N>::in_index(t, q)];
    stages<T2, N, SIGN, 1>(v, lds, 1, t, tw);
#pragma unroll
    for (int e = 0; e < E; e++) {
        T2 x = v[e];
        x.x *= scale;
        x.y *= scale;
        dst[Plan<N>::out_index(t, e)] = x;
    }
}

// Column transforms: W adjacent columns per workgroup (W*16 B or W*8 B contiguous
// per row access), LDS laid out [N][W].  Workgroups are mapped so that the
// strips one XCD runs at a time are adjacent in memory and share its L2.
template <typename Real, int N, int W, int SIGN>
__global__ __launch_bounds__(W* Plan<N>::TPL) void k_fft_cols(const typename cplx<Real>::type* in,
                                                              typename cplx<Real>::type* out,
                                                              const typename cplx<Real>::type* tw,
                                                              int nx, int strips_per_plane,
                                                              Real scale)
{
    using T2 = typename cplx<Real>::type;
    constexpr int E = Plan<N>::E;
    __shared__ T2 lds[N * W];
    const long long s = xcd_remap(blockIdx.x, gridDim.x);
    const long long plane = s / strips_per_plane;
    const int u0 = (int)(s % strips_per_plane) * W;
    const int c = threadIdx.x % W;
    const int t = threadIdx.x / W;
    const bool active = u0 + c < nx;  // the last strip may be partial
    const long long base = plane * (long long)N * nx + (active ? u0 + c : 0);
    const T2* src = in + base;
    T2* dst = out + base;
    T2 v[E];
#pragma unroll
    for (int q = 0; q < E; q++)
        v[q] = active ? src[(long long)Plan<N>::in_index(t, q) * nx] : T2{0, 0};
    stages<T2, N, SIGN, 1>(v, lds + c, W, t, tw);
    if (active) {
#pragma unroll
        for (int e = 0; e < E; e++) {
            T2 x = v[e];
            x.x *= scale;
            x.y *= scale;
            dst[(long long)Plan<N>::out_index(t, e) * nx] = x;
        }
    }
}

// ---------------------------------------------------------------------------
// fused passes

// Row transforms of every plane of one image row, each output multiplied by the
// w-phase screen and summed over planes in registers (phase_correction.cu:151-210);
// the image row is written once.  FP64 phase argument and FP64 accumulation in
// both precision modes.
template <typename Real, int N>
__global__ __launch_bounds__(Plan<N>::TPL) void k_rows_phase(
    const typename cplx<Real>::type* cube, const typename cplx<Real>::type* tw, int ny, int nw,
    PhaseParams pp, double scale, double* img_re, double* img_im, int accumulate)
{
    using T2 = typename cplx<Real>::type;
    constexpr int E = Plan<N>::E;
    __shared__ T2 lds[N];
    const int iv = blockIdx.x;
    const int t = threadIdx.x;
    const long long img_row = (long long)iv * N;

    // per-pixel factor sqrt(1-l^2-m^2) for this thread's output pixels
    const double ycoord = pp.coord(iv, pp.ytot);
    double fac[E];
    double sr[E], si[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int iu = Plan<N>::out_index(t, e);
        const double xcoord = pp.coord(iu, pp.xtot);
        const double radius2 = (xcoord * xcoord + ycoord * ycoord);
        fac[e] = sqrt(1 - radius2);
        sr[e] = 0.0;
        si[e] = 0.0;
    }
    for (int iw = 0; iw < nw; iw++) {
        const T2* src = cube + ((long long)iw * ny + iv) * N;
        T2 v[E];
#pragma unroll
        for (int q = 0; q < E; q++) v[q] = src[Plan<N>::in_index(t, q)];
        stages<T2, N, 1, 1>(v, lds, 1, t, tw);
        const double wterm = pp.wmin + iw * pp.dw;
        const double c2w = 2.0 * LOFAR_REF_PI * wterm;
#pragma unroll
        for (int e = 0; e < E; e++) {
            const double p = (double)v[e].x * scale;
            const double q = (double)v[e].y * scale;
            if (pp.phase_on) {
                double s, r;
                sincos(c2w * (fac[e] - 1.0), &s, &r);
                sr[e] += (p * r - q * s) * pp.dwnorm * fac[e];
                si[e] += (p * s + q * r) * pp.dwnorm * fac[e];
            } else {
                sr[e] += p;
                si[e] += q;
            }
        }
    }
#pragma unroll
    for (int e = 0; e < E; e++) {
        const long long o = img_row + Plan<N>::out_index(t, e);
        if (accumulate) {
            img_re[o] += sr[e];
            img_im[o] += si[e];
        } else {
            img_re[o] = sr[e];
            img_im[o] = si[e];
        }
    }
}

// Adjoint of k_rows_phase: for every plane, the image row times conj(phase_w)
// times the same factor, forward row transform, plane row written out.
template <typename Real, int N>
__global__ __launch_bounds__(Plan<N>::TPL) void k_rows_phase_adj(
    typename cplx<Real>::type* cube, const typename cplx<Real>::type* tw, int ny, int nw,
    PhaseParams pp, double scale, const double* img_re, const double* img_im)
{
    using T2 = typename cplx<Real>::type;
    constexpr int E = Plan<N>::E;
    __shared__ T2 lds[N];
    const int iv = blockIdx.x;
    const int t = threadIdx.x;
    const long long img_row = (long long)iv * N;
    const double ycoord = pp.coord(iv, pp.ytot);
    double fac[E], a[E], b[E];
#pragma unroll
    for (int q = 0; q < E; q++) {
        const int iu = Plan<N>::in_index(t, q);
        const double xcoord = pp.coord(iu, pp.xtot);
        const double radius2 = (xcoord * xcoord + ycoord * ycoord);
        fac[q] = sqrt(1 - radius2);
        a[q] = img_re[img_row + iu];
        b[q] = img_im[img_row + iu];
    }
    for (int iw = 0; iw < nw; iw++) {
        const double wterm = pp.wmin + iw * pp.dw;
        const double c2w = 2.0 * LOFAR_REF_PI * wterm;
        T2 v[E];
#pragma unroll
        for (int q = 0; q < E; q++) {
            double xr = a[q], xi = b[q];
            if (pp.phase_on) {
                double s, r;
                sincos(c2w * (fac[q] - 1.0), &s, &r);
                const double f = pp.dwnorm * fac[q];
                xr = (a[q] * r + b[q] * s) * f;
                xi = (b[q] * r - a[q] * s) * f;
            }
            v[q].x = (Real)xr;
            v[q].y = (Real)xi;
        }
        stages<T2, N, -1, 1>(v, lds, 1, t, tw);
        T2* dst = cube + ((long long)iw * ny + iv) * N;
#pragma unroll
        for (int e = 0; e < E; e++) {
            T2 x = v[e];
            x.x = (Real)(x.x * scale);
            x.y = (Real)(x.y * scale);
            dst[Plan<N>::out_index(t, e)] = x;
        }
    }
}

// ---------------------------------------------------------------------------
// host-side dispatch

template <typename Real, int N>
static hipError_t launch_rows_n(int sign, const void* in, void* out, const void* tw,
                                long long lines, double scale, hipStream_t st)
{
    using T2 = typename cplx<Real>::type;
    if (lines <= 0) return hipSuccess;
    if (sign > 0)
        hipLaunchKernelGGL((k_fft_rows<Real, N, 1>), dim3((unsigned)lines), dim3(Plan<N>::TPL), 0,
                           st, (const T2*)in, (T2*)out, (const T2*)tw, (Real)scale);
    else
        hipLaunchKernelGGL((k_fft_rows<Real, N, -1>), dim3((unsigned)lines), dim3(Plan<N>::TPL), 0,
                           st, (const T2*)in, (T2*)out, (const T2*)tw, (Real)scale);
    return hipGetLastError();
}

template <typename Real, int N>
static hipError_t launch_cols_n(int sign, const void* in, void* out, const void* tw, int nx, int nw,
                                double scale, hipStream_t st)
{
    using T2 = typename cplx<Real>::type;
    constexpr int W = ColW<Real, N>::W;
    const int strips = (nx + W - 1) / W;
    const long long blocks = (long long)strips * nw;
    if (sign > 0)
        hipLaunchKernelGGL((k_fft_cols<Real, N, W, 1>), dim3((unsigned)blocks),
                           dim3(W * Plan<N>::TPL), 0, st, (const T2*)in, (T2*)out, (const T2*)tw,
                           nx, strips, (Real)scale);
    else
        hipLaunchKernelGGL((k_fft_cols<Real, N, W, -1>), dim3((unsigned)blocks),
                           dim3(W * Plan<N>::TPL), 0, st, (const T2*)in, (T2*)out, (const T2*)tw,
                           nx, strips, (Real)scale);
    return hipGetLastError();
}

template <typename Real, int N>
static hipError_t launch_rows_phase_n(const void* cube, const void* tw, int ny, int nw,
                                      const PhaseParams& pp, double scale, double* re, double* im,
                                      int accumulate, hipStream_t st)
{
    using T2 = typename cplx<Real>::type;
    hipLaunchKernelGGL((k_rows_phase<Real, N>), dim3(ny), dim3(Plan<N>::TPL), 0, st,
                       (const T2*)cube, (const T2*)tw, ny, nw, pp, scale, re, im, accumulate);
    return hipGetLastError();
}

template <typename Real, int N>
static hipError_t launch_rows_phase_adj_n(void* cube, const void* tw, int ny, int nw,
                                          const PhaseParams& pp, double scale, const double* re,
                                          const double* im, hipStream_t st)
{
    using T2 = typename cplx<Real>::type;
    hipLaunchKernelGGL((k_rows_phase_adj<Real, N>), dim3(ny), dim3(Plan<N>::TPL), 0, st, (T2*)cube,
                       (const T2*)tw, ny, nw, pp, scale, re, im);
    return hipGetLastError();
}

#define LOFAR_FFT_SWITCH(N_, CALL)        \
    switch (N_) {                         \
        case 2: return CALL(2);           \
        case 4: return CALL(4);           \
        case 8: return CALL(8);           \
        case 16: return CALL(16);         \
        case 32: return CALL(32);         \
        case 64: return CALL(64);         \
        case 128: return CALL(128);       \
        case 256: return CALL(256);       \
        case 512: return CALL(512);       \
        case 1024: return CALL(1024);     \
        case 2048: return CALL(2048);     \
        case 4096: return CALL(4096);     \
        case 8192: return CALL(8192);     \
        default: return hipErrorInvalidValue; \
    }

template <typename Real>
hipError_t fft_rows(int sign, const void* in, void* out, const void* tw, int nx,
                    long long lines, double scale, hipStream_t st)
{
#define CALL_(n) launch_rows_n<Real, n>(sign, in, out, tw, lines, scale, st)
    LOFAR_FFT_SWITCH(nx, CALL_)
#undef CALL_
}

template <typename Real>
hipError_t fft_cols(int sign, const void* in, void* out, const void* tw, int nx, int ny, int nw,
                    double scale, hipStream_t st)
{
#define CALL_(n) launch_cols_n<Real, n>(sign, in, out, tw, nx, nw, scale, st)
    LOFAR_FFT_SWITCH(ny, CALL_)
#undef CALL_
}

template <typename Real>
hipError_t rows_phase(const void* cube, const void* tw, int nx, int ny, int nw,
                      const PhaseParams& pp, double scale, double* re, double* im, int accumulate,
                      hipStream_t st)
{
#define CALL_(n) launch_rows_phase_n<Real, n>(cube, tw, ny, nw, pp, scale, re, im, accumulate, st)
    LOFAR_FFT_SWITCH(nx, CALL_)
#undef CALL_
}

template <typename Real>
hipError_t rows_phase_adj(void* cube, const void* tw, int nx, int ny, int nw,
                          const PhaseParams& pp, double scale, const double* re, const double* im,
                          hipStream_t st)
{
#define CALL_(n) launch_rows_phase_adj_n<Real, n>(cube, tw, ny, nw, pp, scale, re, im, st)
    LOFAR_FFT_SWITCH(nx, CALL_)
#undef CALL_
}

template hipError_t fft_rows<double>(int, const void*, void*, const void*, int, long long, double,
                                     hipStream_t);
template hipError_t fft_rows<float>(int, const void*, void*, const void*, int, long long, double,
                                    hipStream_t);
template hipError_t fft_cols<double>(int, const void*, void*, const void*, int, int, int, double,
                                     hipStream_t);
template hipError_t fft_cols<float>(int, const void*, void*, const void*, int, int, int, double,
                                    hipStream_t);
template hipError_t rows_phase<double>(const void*, const void*, int, int, int, const PhaseParams&,
                                       double, double*, double*, int, hipStream_t);
template hipError_t rows_phase<float>(const void*, const void*, int, int, int, const PhaseParams&,
                                      double, double*, double*, int, hipStream_t);
template hipError_t rows_phase_adj<double>(void*, const void*, int, int, int, const PhaseParams&,
                                           double, const double*, const double*, hipStream_t);
template hipError_t rows_phase_adj<float>(void*, const void*, int, int, int, const PhaseParams&,
                                          double, const double*, const double*, hipStream_t);

}  // namespace lofar
