// Batched per-plane 2D FFT for gfx950: Stockham autosort, radix-E stages
// (E = 8 or 16 elements per thread) with a smaller remainder stage, lines
// staged in LDS between stages.
//
// Semantics follow fftw_data() (fourier_transform.c:7-74): per w-plane an
// unnormalised FFTW_BACKWARD transform over dims (n0=ny, n1=nx) of the [v][u]
// plane, scaled by 1/(nx*ny).  Column transforms run first, then rows, so the
// last pass (rows, contiguous) fuses the w-phase screen and the plane sum of
// phase_correction() (phase_correction.cu:151-210) and the image is written once.
#include <cstdlib>

#include "fft_kernels.h"

namespace lofar {

template <typename Real>
using C2 = typename cplx<Real>::type;

// Raw buffer load of one complex value at byte offset voff + soff (soff
// wave-uniform, held in a scalar register).
template <typename T2>
__device__ __forceinline__ T2 load_soff(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    if constexpr (sizeof(T2) == 16)
        return __builtin_bit_cast(T2, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
    else
        return __builtin_bit_cast(T2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

// Raw buffer store of one complex value at byte offset voff + soff.  The whole
// offset goes in the vector register and the scalar offset field is 0: with a
// scalar register there, the compiler does not keep the next VALU instruction
// from overwriting the data registers of a 128-bit store before the store has
// read them, and complex-double stores wrote the following element's value.
template <typename T2>
__device__ __forceinline__ void store_at(T2 x, __amdgpu_buffer_rsrc_t r, unsigned voff,
                                           unsigned soff)
{
    using U4 = unsigned __attribute__((ext_vector_type(4)));
    using U2 = unsigned __attribute__((ext_vector_type(2)));
    if constexpr (sizeof(T2) == 16)
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(U4, x), r, voff + soff, 0, 0);
    else
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, x), r, voff + soff, 0, 0);
}

// ---------------------------------------------------------------------------
// standalone passes

// Row transforms: one line (row) per workgroup.
template <typename Real, int N, int E, int SIGN>
__global__ __launch_bounds__((Plan<N, E>::TPL)) void k_fft_rows(const C2<Real>* in, C2<Real>* out,
                                                              const C2<Real>* tw, Real scale)
{
    using T2 = C2<Real>;
    using P = Plan<N, E>;
    __shared__ T2 lds[P::LDS];
    const long long line = blockIdx.x;
    const T2* src = in + line * N;
    T2* dst = out + line * N;
    const int t = threadIdx.x;
    T2 v[P::E];
#pragma unroll
    for (int q = 0; q < P::E; q++) v[q] = src[P::in_index(t, q)];
    stages<T2, N, E, SIGN, 1>(v, lds, 1, t, tw);
#pragma unroll
    for (int e = 0; e < P::E; e++) {
        T2 x = v[e];
        x.x *= scale;
        x.y *= scale;
        dst[P::out_index(t, e)] = x;
    }
}

// Column transforms: one strip of W adjacent columns per workgroup (W*16 or
// W*8 bytes contiguous per row access), LDS laid out [slot][W].  Workgroups
// are numbered so that the strips one XCD runs at a time are adjacent in
// memory and share its L2 lines.
// OCC: `occ` holds the forward path's occupancy bits (rows of empty tiles read as
// zeros).  NEED: `occ` holds the tiles a degridding reads (same indexing); rows of
// other tiles are not stored, and a strip whose tile column holds none is skipped.
template <typename Real, int N, int E, int W, int ROUNDS, int SIGN, bool OCC, bool NEED = false>
__global__ __launch_bounds__((W * Plan<N, E>::TPL)) void k_fft_cols(const C2<Real>* in,
                                                                 C2<Real>* out,
                                                                 const C2<Real>* tw, int nx,
                                                                 int strips_per_plane,
                                                                 long long nstrips, Real scale,
                                                                 const unsigned long long* occ,
                                                                 int sm_log)
{
    using T2 = C2<Real>;
    using P = Plan<N, E>;
    // complex doubles are stored through a buffer resource (one vector offset per
    // thread, no 64-bit address arithmetic per element), complex floats with
    // global stores (measured as fast, with fewer instructions around them)
    constexpr bool kBufStore = sizeof(T2) >= 16;
    constexpr int WR = W / ROUNDS;  // columns sharing the LDS buffer in one round
    __shared__ T2 lds[P::LDS * WR];
    const long long s = xcd_remap(blockIdx.x, gridDim.x);
    const long long plane = s / strips_per_plane;
    const int u0 = (int)(s % strips_per_plane) * W;
    const int c = threadIdx.x % W;
    const int t = threadIdx.x / W;
    const bool active = u0 + c < nx;  // the last strip of a plane may be partial
    // row-major planes: rows nx apart; strip-major planes ([nx/L][N][L], L = 2^sm_log,
    // a multiple of W): rows L apart inside the strip holding these columns
    const long long rs = sm_log >= 0 ? (1LL << sm_log) : nx;
    const long long plane0 = plane * (long long)N * nx;
    const long long base =
        plane0 +
        (sm_log >= 0 ? ((long long)(u0 >> sm_log) * N << sm_log) + (u0 & ((1 << sm_log) - 1)) + c
                     : (active ? u0 + c : 0));
    // Element q sits dq = in_index(t, q) - t rows below the thread's first row, and
    // output e at out_index(t, e) - t, the same for every thread: buffer accesses
    // are one per-thread byte offset into the plane plus a per-element offset
    // (planes are below 2 GiB, launch_cols_cfg checks: 32-bit offsets).
    const int plane_bytes = (int)((long long)N * nx * (long long)sizeof(T2));
    const unsigned row_step = (unsigned)(rs * (long long)sizeof(T2));
    const unsigned vbase =
        (unsigned)((base - plane0) * (long long)sizeof(T2)) + (unsigned)t * row_step;
    // complex doubles: the stages' twiddles as powers of roots loaded here, ahead
    // of the column data
    constexpr int NWD = ((N + kTile - 1) / kTile + 63) / 64;  // mask words per tile column
    static_assert(NWD <= 2, "tile masks cover at most 128 tile rows");
    unsigned long long need0 = ~0ull, need1 = ~0ull;
    if constexpr (NEED) {
        const int ntu = (nx + kTile - 1) / kTile;
        const unsigned long long* nm = occ + (plane * ntu + u0 / kTile) * NWD;
        need0 = nm[0];
        need1 = NWD > 1 ? nm[1] : 0ull;
        if ((need0 | need1) == 0ull) return;  // nothing of this tile column is read
    }
    constexpr bool kRoots = sizeof(T2) >= 16 && W * P::TPL <= 512;  // not at the 128-VGPR cap
    constexpr int NRT = kRoots ? root_count<N, E>(1) : 0;
    T2 rt[NRT > 0 ? NRT : 1];
    if constexpr (NRT > 0) load_roots<T2, N, E, 1>(rt, t, tw);
    T2 v[P::E];
    if constexpr (OCC) {
        // An offset past the plane (rows of tiles without samples, columns past the
        // grid) reads zeros without a memory access.
        const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(in + plane0), (short)0, plane_bytes, 0x00020000);  // raw, range-checked
        // bit tv of word tv/64: tile row tv of this strip's tile column holds samples
        const int ntu = (nx + kTile - 1) / kTile;
        const unsigned long long* om = occ + (plane * ntu + u0 / kTile) * NWD;
        const unsigned long long live0 = om[0], live1 = NWD > 1 ? om[1] : 0ull;
        if ((live0 | live1) == 0ull) {
            // no samples anywhere in this strip's tile column: its transform is
            // zero, stored without loads or arithmetic
            if (active) {
                const __amdgpu_buffer_rsrc_t wsrc = __builtin_amdgcn_make_buffer_rsrc(
                    (void*)(out + plane0), (short)0, plane_bytes, 0x00020000);
#pragma unroll
                for (int e = 0; e < P::E; e++)
                    store_at(T2{0, 0}, wsrc, vbase,
                               (unsigned)(P::out_index(0, e)) * row_step);
            }
            return;
        }
#pragma unroll
        for (int q = 0; q < P::E; q++) {
            const int dq = P::in_index(0, q);
            const int tv = (t + dq) / kTile;
            const bool live = active && ((((tv < 64) ? live0 : live1) >> (tv & 63)) & 1ull);
            // offsets of live rows are below 2^31; dead rows read past the plane
            const unsigned voff = live ? vbase : 0x80000000u;
            v[q] = load_soff<T2>(rsrc, voff, (unsigned)dq * row_step);
        }
    } else {
#pragma unroll
        for (int q = 0; q < P::E; q++)
            v[q] = active ? in[base + (long long)P::in_index(t, q) * rs] : T2{0, 0};
    }
    stages<T2, N, E, SIGN, 1, ROUNDS, (NRT > 0)>(v, lds + c % WR, WR, t, tw, c / WR, rt);
    if (active) {
        const __amdgpu_buffer_rsrc_t wsrc = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(out + plane0), (short)0, plane_bytes, 0x00020000);
#pragma unroll
        for (int e = 0; e < P::E; e++) {
            if constexpr (NEED) {
                const int tv = (t + P::out_index(0, e)) / kTile;
                if (!((((tv < 64) ? need0 : need1) >> (tv & 63)) & 1ull)) continue;
            }
            T2 x = v[e];
            x.x *= scale;
            x.y *= scale;
            if constexpr (kBufStore)
                store_at(x, wsrc, vbase, (unsigned)(P::out_index(0, e)) * row_step);
            else
                out[base + (long long)P::out_index(t, e) * rs] = x;
        }
    }
}

// ---------------------------------------------------------------------------
// fused passes

// Per-pixel factor sqrt(1-l^2-m^2) (phase_correction.cu:164-172); 1 without the phase.
__device__ __forceinline__ double pixel_fac(const PhaseParams& pp, int iu, double ycoord)
{
    if (!pp.phase_on) return 1.0;
    const double xcoord = pp.coord(iu, pp.xtot);
    const double radius2 = (xcoord * xcoord + ycoord * ycoord);
    return sqrt(1 - radius2);
}

// Per-pixel phase state of the fused row passes.  RECUR: plane factors by the
// complex recurrence z_{w+1} = z_w*step (two complex numbers per pixel);
// otherwise each plane's factor is evaluated directly with sincos_w (one double
// per pixel).
template <bool RECUR>
struct PixelPhase;

template <>
struct PixelPhase<true> {
    double2 z, stp;
    __device__ __forceinline__ void init(const PhaseParams& pp, double fac)
    {
        phase_rotation(pp, fac, z, stp);
    }
    __device__ __forceinline__ double2 next(const PhaseParams&, double)
    {
        const double2 f = z;
        z = cmul(z, stp);
        return f;
    }
};

template <>
struct PixelPhase<false> {
    double fac;
    __device__ __forceinline__ void init(const PhaseParams&, double f) { fac = f; }
    // factor of the plane whose 2*PI*wterm is c2w
    __device__ __forceinline__ double2 next(const PhaseParams& pp, double c2w)
    {
        if (!pp.phase_on) return double2{1.0, 0.0};
        double sn, cs;
        sincos_w(c2w * (fac - 1.0), &sn, &cs);
        const double f = pp.dwnorm * fac;
        return double2{cs * f, sn * f};
    }
};

// Row transforms of every plane of one image row; each output is multiplied by
// its w-phase factor exp(i*phi_w)*dwnorm*sqrt(1-r^2) (phase_correction.cu:160,
// 182-183,193-195) and summed over planes in registers, and the image row is
// written once.  The thread's twiddle factors are loaded into registers once
// (TWREG) since every plane's row uses the same ones.  FP64 phase and
// accumulation in both precision modes.
// Row prefetch for k_rows_phase: row iv of plane psrc goes straight into
// pbuf[0, N) with 16-byte direct-to-LDS loads (global_load_lds_dwordx4); wave
// instruction c of the row writes 1 KiB, elements [c*K, (c+1)*K) with K = 1024 /
// sizeof(T2) (one run of a row-major row, or of one strip of a strip-major plane
// when its strips are at least K elements wide).
template <typename T2, int N, int TPL>
__device__ __forceinline__ void prefetch_row(const T2* psrc, T2* pbuf, int t, int iv, int ny,
                                             int sm_log)
{
    constexpr int PER = 16 / (int)sizeof(T2);  // elements per lane
    constexpr int K = 64 * PER;                 // elements per wave instruction
    constexpr int NWAVE = TPL / 64;
    static_assert(NWAVE >= 1 && N % K == 0, "whole waves, whole 1 KiB runs");
    const int lane = t & 63, wave = t >> 6;
    constexpr int ITER = (N / K + NWAVE - 1) / NWAVE;
    // wave-uniform run start (scalar registers) plus a per-lane offset
    const int wave_u = __builtin_amdgcn_readfirstlane(wave);
    const T2* lane_src = psrc + lane * PER;
#pragma unroll
    for (int i = 0; i < ITER; i++) {
        const int c = wave_u + i * NWAVE;
        if (c >= N / K) break;
        const int u = c * K;  // the run never crosses a strip: strips are >= K wide
        const long long o = sm_log < 0 ? (long long)iv * N + u
                                        : ((((long long)(u >> sm_log) * ny + iv) << sm_log) |
                                           (u & ((1 << sm_log) - 1)));
        __builtin_amdgcn_global_load_lds((const void*)(lane_src + o),
                                         (__attribute__((address_space(3))) void*)(pbuf + u), 16,
                                         0, 0);
    }
}

// LDS bytes of k_rows_phase with the row prefetch buffer
template <typename Real, int N, int E>
constexpr bool rows_prefetch_fits()
{
    return (Plan<N, E>::LDS + N) * (int)sizeof(C2<Real>) <= 160 * 1024;
}

template <typename Real, int N, int E, bool RECUR, bool TWREG, bool PF>
__global__ __launch_bounds__((Plan<N, E>::TPL)) void k_rows_phase(const C2<Real>* cube,
                                                                const C2<Real>* tw, int ny, int nw,
                                                                PhaseParams pp, double scale,
                                                                double* img_re, double* img_im,
                                                                int accumulate, int sm_log)
{
    using T2 = C2<Real>;
    using P = Plan<N, E>;
    constexpr int NTW = tw_count<N, E>(1);
    __shared__ T2 lds[P::LDS];
    // PF: plane iw+1's row is loaded into pbuf while plane iw is transformed
    __shared__ __align__(16) T2 pbuf[PF ? N : 1];
    const int iv = blockIdx.x;
    const int t = threadIdx.x;
    const long long img_row = (long long)iv * N;
    const double ycoord = pp.phase_on ? pp.coord(iv, pp.ytot) : 0.0;
    PixelPhase<RECUR> ph[P::E];
    double2 acc[P::E];
#pragma unroll
    for (int e = 0; e < P::E; e++) {
        ph[e].init(pp, pixel_fac(pp, P::out_index(t, e), ycoord));
        if constexpr (RECUR) {
            // the transform's scale 1/(nx*ny) (a power of two: exact) rides in the
            // recurrence's first factor
            ph[e].z.x *= scale;
            ph[e].z.y *= scale;
        }
        acc[e] = double2{0.0, 0.0};
    }
    T2 twr[NTW > 0 ? NTW : 1];
    if constexpr (TWREG && NTW > 0) load_twiddles<T2, N, E, 1>(twr, t, tw);
    const long long plane_stride = (long long)ny * N;
    // element u of the row: at u (row-major planes) or, in strip-major planes
    // ([N/L][ny][L], L = 2^sm_log), at (u/L)*ny*L + iv*L + u%L
    long long off[P::E];
#pragma unroll
    for (int q = 0; q < P::E; q++) {
        const int u = P::in_index(t, q);
        off[q] = sm_log < 0 ? img_row + u
                            : ((((long long)(u >> sm_log) * ny + iv) << sm_log) |
                               (u & ((1 << sm_log) - 1)));
    }
    if constexpr (PF) prefetch_row<T2, N, P::TPL>(cube, pbuf, t, iv, ny, sm_log);
    T2 v[P::E];
    for (int iw = 0; iw < nw; iw++) {
        if constexpr (PF) {
            wait_loads();
            lds_barrier();  // every wave's part of the row is in pbuf
#pragma unroll
            for (int q = 0; q < P::E; q++) v[q] = pbuf[P::in_index(t, q)];
            lds_barrier();  // pbuf read by every wave: free for the next plane
            if (iw + 1 < nw)
                prefetch_row<T2, N, P::TPL>(cube + (long long)(iw + 1) * plane_stride, pbuf, t, iv,
                                            ny, sm_log);
        } else {
            const T2* psrc = cube + (long long)iw * plane_stride;
#pragma unroll
            for (int q = 0; q < P::E; q++) v[q] = psrc[off[q]];
        }
        if constexpr (TWREG)
            stages_rt<T2, N, E, 1, 1, PF>(v, lds, t, twr);
        else
            stages<T2, N, E, 1, 1>(v, lds, 1, t, tw);
        const double c2w = 2.0 * LOFAR_REF_PI * (pp.wmin + iw * pp.dw);
#pragma unroll
        for (int e = 0; e < P::E; e++) {
            if constexpr (RECUR) {
                // acc += v*z as four fused multiply-adds
                const double2 z = ph[e].next(pp, c2w);
                const double vx = (double)v[e].x, vy = (double)v[e].y;
                acc[e].x = fma(vx, z.x, fma(-vy, z.y, acc[e].x));
                acc[e].y = fma(vx, z.y, fma(vy, z.x, acc[e].y));
            } else {
                const double2 pq = double2{(double)v[e].x * scale, (double)v[e].y * scale};
                const double2 term = cmul(pq, ph[e].next(pp, c2w));
                acc[e].x += term.x;
                acc[e].y += term.y;
            }
        }
    }
#pragma unroll
    for (int e = 0; e < P::E; e++) {
        const long long o = img_row + P::out_index(t, e);
        if (accumulate) {
            img_re[o] += acc[e].x;
            img_im[o] += acc[e].y;
        } else {
            img_re[o] = acc[e].x;
            img_im[o] = acc[e].y;
        }
    }
}

// Adjoint of k_rows_phase: for every plane, the image row times the conjugate
// w-phase factor, forward row transform, plane row written out.
template <typename Real, int N, int E, bool RECUR, bool TWREG>
__global__ __launch_bounds__((Plan<N, E>::TPL)) void k_rows_phase_adj(C2<Real>* cube,
                                                                    const C2<Real>* tw, int ny,
                                                                    int nw, PhaseParams pp,
                                                                    double scale,
                                                                    const double* img_re,
                                                                    const double* img_im)
{
    using T2 = C2<Real>;
    using P = Plan<N, E>;
    constexpr int NTW = tw_count<N, E>(1);
    __shared__ T2 lds[P::LDS];
    const int iv = blockIdx.x;
    const int t = threadIdx.x;
    const long long img_row = (long long)iv * N;
    const double ycoord = pp.phase_on ? pp.coord(iv, pp.ytot) : 0.0;
    PixelPhase<RECUR> ph[P::E];
    double2 a[P::E];
#pragma unroll
    for (int q = 0; q < P::E; q++) {
        const int iu = P::in_index(t, q);
        ph[q].init(pp, pixel_fac(pp, iu, ycoord));
        a[q] = double2{img_re[img_row + iu], img_im[img_row + iu]};
    }
    T2 twr[NTW > 0 ? NTW : 1];
    if constexpr (TWREG && NTW > 0) load_twiddles<T2, N, E, 1>(twr, t, tw);
    const long long plane_stride = (long long)ny * N;
    for (int iw = 0; iw < nw; iw++) {
        const double c2w = 2.0 * LOFAR_REF_PI * (pp.wmin + iw * pp.dw);
        T2 v[P::E];
#pragma unroll
        for (int q = 0; q < P::E; q++) {
            const double2 z = ph[q].next(pp, c2w);
            // a * conj(z)
            v[q].x = (Real)(a[q].x * z.x + a[q].y * z.y);
            v[q].y = (Real)(a[q].y * z.x - a[q].x * z.y);
        }
        if constexpr (TWREG)
            stages_rt<T2, N, E, -1, 1>(v, lds, t, twr);
        else
            stages<T2, N, E, -1, 1>(v, lds, 1, t, tw);
        T2* dst = cube + (long long)iw * plane_stride + img_row;
#pragma unroll
        for (int e = 0; e < P::E; e++) {
            T2 x = v[e];
            x.x = (Real)(x.x * scale);
            x.y = (Real)(x.y * scale);
            dst[P::out_index(t, e)] = x;
        }
    }
}

// ---------------------------------------------------------------------------
// host-side dispatch

// elements per thread of the fused row passes and of the column pass, and
// columns per strip; LOFAR_FFT_CFG overrides the column configuration (tuning)
constexpr int kRowE = 16;
constexpr int kPhaseE = 8;

static int col_cfg()
{
    static int cfg = -1;
    if (cfg < 0) {
        const char* s = getenv("LOFAR_FFT_CFG");
        cfg = s ? atoi(s) : 0;
    }
    return cfg;
}

template <typename Real, int N>
static hipError_t launch_rows_n(int sign, const void* in, void* out, const void* tw,
                                long long lines, double scale, hipStream_t st)
{
    using T2 = C2<Real>;
    constexpr int TPL = Plan<N, kRowE>::TPL;
    if (lines <= 0) return hipSuccess;
    if (sign > 0)
        hipLaunchKernelGGL((k_fft_rows<Real, N, kRowE, 1>), dim3((unsigned)lines), dim3(TPL), 0,
                           st, (const T2*)in, (T2*)out, (const T2*)tw, (Real)scale);
    else
        hipLaunchKernelGGL((k_fft_rows<Real, N, kRowE, -1>), dim3((unsigned)lines), dim3(TPL), 0,
                           st, (const T2*)in, (T2*)out, (const T2*)tw, (Real)scale);
    return hipGetLastError();
}

template <typename Real, int N, int E, int WREQ, int RREQ>
static hipError_t launch_cols_cfg(int sign, const void* in, void* out, const void* tw, int nx,
                                  int nw, double scale, const unsigned long long* occ, int sm_log,
                                  hipStream_t st)
{
    using T2 = C2<Real>;
    constexpr int TPL = Plan<N, E>::TPL;
    // at most 1024 threads per workgroup; ROUNDS divides W (both powers of two)
    // and is large enough that one round's columns fit 140 KiB of LDS
    constexpr int W = WREQ * TPL <= 1024 ? WREQ : 1024 / TPL;
    constexpr int COL = Plan<N, E>::LDS * 2 * (int)sizeof(Real);
    constexpr int RMIN = (W * COL + 143359) / 143360;
    constexpr int RMIN2 = RMIN <= 1 ? 1 : RMIN <= 2 ? 2 : RMIN <= 4 ? 4 : RMIN <= 8 ? 8 : 16;
    constexpr int R1 = RREQ > RMIN2 ? RREQ : RMIN2;
    constexpr int ROUNDS = R1 < W ? R1 : W;
    const int strips = (nx + W - 1) / W;
    const long long nstrips = (long long)strips * nw;
    // the kernel addresses a plane with 32-bit byte offsets
    if ((long long)N * nx * (long long)sizeof(T2) >= (1LL << 31)) return hipErrorInvalidValue;
    // forward transforms read `occ` as occupancy, inverse ones as the tiles to store
    auto k = sign > 0 ? (occ ? k_fft_cols<Real, N, E, W, ROUNDS, 1, true>
                             : k_fft_cols<Real, N, E, W, ROUNDS, 1, false>)
                      : (occ ? k_fft_cols<Real, N, E, W, ROUNDS, -1, false, true>
                             : k_fft_cols<Real, N, E, W, ROUNDS, -1, false>);
    // strip-major planes: whole strips of a power-of-two width that W divides
    if (sm_log >= 0 && ((1 << sm_log) % W != 0 || nx % (1 << sm_log) != 0))
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k, dim3((unsigned)nstrips), dim3(W * TPL), 0, st, (const T2*)in, (T2*)out,
                       (const T2*)tw, nx, strips, nstrips, (Real)scale, occ, sm_log);
    return hipGetLastError();
}

// Column-pass configurations; entry 0 is the default.  W2: strip width whose
// LDS fits ~128 KiB in one round; W64: 64-byte row segments.
template <typename Real, int N>
static hipError_t launch_cols_n(int sign, const void* in, void* out, const void* tw, int nx,
                                int nw, double scale, const unsigned long long* occ, int sm_log,
                                hipStream_t st)
{
    constexpr int B = Plan<N, 16>::LDS * 2 * (int)sizeof(Real);
    constexpr int W2 = 140000 / B >= 1 ? (140000 / B > 16 ? 16 : 140000 / B) : 1;
    constexpr int W64 = 64 / (2 * (int)sizeof(Real));
    // 8192-long complex-double columns at 32 elements a thread would need 128 VGPRs
    // of data alone at 1024 threads (they spilled: 0.74 TB/s); 16 elements, 2 columns
    // and 2 LDS rounds run without spills (2.5 TB/s)
    if constexpr (sizeof(Real) == 8 && N >= 8192)
        if (col_cfg() == 0 || sm_log >= 0)
            return launch_cols_cfg<Real, N, 16, W64, 2>(sign, in, out, tw, nx, nw, scale, occ, sm_log, st);
    if (sm_log >= 0)
        return launch_cols_cfg<Real, N, 32, W64, 2>(sign, in, out, tw, nx, nw, scale, occ, sm_log, st);
    switch (col_cfg()) {
        case 1: return launch_cols_cfg<Real, N, 16, W2, 1>(sign, in, out, tw, nx, nw, scale, occ, sm_log, st);
        case 2: return launch_cols_cfg<Real, N, 32, W64, 4>(sign, in, out, tw, nx, nw, scale, occ, sm_log, st);
        case 3: return launch_cols_cfg<Real, N, 16, W64, 2>(sign, in, out, tw, nx, nw, scale, occ, sm_log, st);
        default: return launch_cols_cfg<Real, N, 32, W64, 2>(sign, in, out, tw, nx, nw, scale, occ, sm_log, st);
    }
}

static int phase_cfg()
{
    static int cfg = -1;
    if (cfg < 0) {
        const char* s = getenv("LOFAR_PHASE_CFG");
        cfg = s ? atoi(s) : 0;
    }
    return cfg;
}

template <typename Real, int N, int E, bool RECUR, bool TWREG, bool PF = false>
static hipError_t launch_rows_phase_cfg(const void* cube, const void* tw, int ny, int nw,
                                        const PhaseParams& pp, double scale, double* re,
                                        double* im, int accumulate, int sm_log, hipStream_t st)
{
    using T2 = C2<Real>;
    if constexpr (PF) {
        // the prefetch reads whole 1 KiB runs: row-major rows, or strips at least
        // that wide
        constexpr int K = 1024 / (int)sizeof(T2);
        if (sm_log >= 0 && (1 << sm_log) < K)
            return launch_rows_phase_cfg<Real, N, E, RECUR, TWREG, false>(
                cube, tw, ny, nw, pp, scale, re, im, accumulate, sm_log, st);
    }
    hipLaunchKernelGGL((k_rows_phase<Real, N, E, RECUR, TWREG, PF>), dim3(ny),
                       dim3(Plan<N, E>::TPL), 0, st, (const T2*)cube, (const T2*)tw, ny, nw, pp,
                       scale, re, im, accumulate, sm_log);
    return hipGetLastError();
}

// fused row-pass configurations (elements per thread, phase recurrence or direct
// sincos, twiddles in registers, next row prefetched into LDS); entry 0 is the
// default
template <typename Real, int N>
static hipError_t launch_rows_phase_n(const void* cube, const void* tw, int ny, int nw,
                                      const PhaseParams& pp, double scale, double* re, double* im,
                                      int accumulate, int sm_log, hipStream_t st)
{
#define ARGS_ cube, tw, ny, nw, pp, scale, re, im, accumulate, sm_log, st
    switch (phase_cfg()) {
        case 1: return launch_rows_phase_cfg<Real, N, 8, false, true>(ARGS_);
        case 2: return launch_rows_phase_cfg<Real, N, 8, true, false>(ARGS_);
        case 3: return launch_rows_phase_cfg<Real, N, 16, true, true>(ARGS_);
        case 4: return launch_rows_phase_cfg<Real, N, 4, true, true>(ARGS_);
        case 5: return launch_rows_phase_cfg<Real, N, 8, true, true>(ARGS_);
        default:
            return launch_rows_phase_cfg<Real, N, 8, true, true,
                                         (Plan<N, 8>::TPL >= 64 &&
                                          rows_prefetch_fits<Real, N, 8>())>(ARGS_);
    }
#undef ARGS_
}

template <typename Real, int N>
static hipError_t launch_rows_phase_adj_n(void* cube, const void* tw, int ny, int nw,
                                          const PhaseParams& pp, double scale, const double* re,
                                          const double* im, hipStream_t st)
{
    using T2 = C2<Real>;
    hipLaunchKernelGGL((k_rows_phase_adj<Real, N, kPhaseE, true, true>), dim3(ny),
                       dim3(Plan<N, kPhaseE>::TPL), 0, st, (T2*)cube, (const T2*)tw, ny, nw, pp,
                       scale, re, im);
    return hipGetLastError();
}

#define LOFAR_FFT_SWITCH(N_, CALL)            \
    switch (N_) {                             \
        case 2: return CALL(2);               \
        case 4: return CALL(4);               \
        case 8: return CALL(8);               \
        case 16: return CALL(16);             \
        case 32: return CALL(32);             \
        case 64: return CALL(64);             \
        case 128: return CALL(128);           \
        case 256: return CALL(256);           \
        case 512: return CALL(512);           \
        case 1024: return CALL(1024);         \
        case 2048: return CALL(2048);         \
        case 4096: return CALL(4096);         \
        case 8192: return CALL(8192);         \
        default: return hipErrorInvalidValue; \
    }

template <typename Real>
hipError_t fft_rows(int sign, const void* in, void* out, const void* tw, int nx, long long lines,
                    double scale, hipStream_t st)
{
#define CALL_(n) launch_rows_n<Real, n>(sign, in, out, tw, lines, scale, st)
    LOFAR_FFT_SWITCH(nx, CALL_)
#undef CALL_
}

template <typename Real>
hipError_t fft_cols(int sign, const void* in, void* out, const void* tw, int nx, int ny, int nw,
                    double scale, hipStream_t st, const unsigned long long* occ, int sm_log)
{
#define CALL_(n) launch_cols_n<Real, n>(sign, in, out, tw, nx, nw, scale, occ, sm_log, st)
    LOFAR_FFT_SWITCH(ny, CALL_)
#undef CALL_
}

template <typename Real>
hipError_t rows_phase(const void* cube, const void* tw, int nx, int ny, int nw,
                      const PhaseParams& pp, double scale, double* re, double* im, int accumulate,
                      hipStream_t st, int sm_log)
{
#define CALL_(n) launch_rows_phase_n<Real, n>(cube, tw, ny, nw, pp, scale, re, im, accumulate, sm_log, st)
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

// ---------------------------------------------------------------------------
// Mixed-radix transforms of any line length whose prime factors are at most 13 (the
// sides fftw_data() plans that are not powers of two, fourier_transform.c:25-27).
// One 256-thread workgroup per line; the line sits in LDS and each Stockham stage of
// radix P reads all of a thread's butterfly inputs into registers, then (after a
// barrier) writes the outputs in place: butterfly j (k = j mod Ns, Ns the product of
// the radices before this stage) takes x[j + q*N/P], q < P, times exp(+-2 pi i q k /
// (Ns P)), DFT over q, and stores y[(j/Ns)*Ns*P + k + r*Ns].  Twiddles and the
// radix-P roots come from the N-entry table exp(+2 pi i m/N).

constexpr int kGenT = 256;       // threads per line
constexpr int kGenMaxN = 8192;   // longest line
constexpr int kGenMaxStages = 16;

struct GenPlan {
    int n, nstages;
    int radix[kGenMaxStages];
};

template <int P, int SIGN, typename T2>
__device__ __forceinline__ void gen_stage(T2* buf, int n, int ns, int t,
                                          const T2* __restrict__ tw)
{
    constexpr int NB = (kGenMaxN / P + kGenT - 1) / kGenT;  // butterflies per thread, at most
    const int nb = n / P;
    const int step = n / (ns * P);  // table stride of exp(2 pi i / (Ns P))
    T2 a[NB][P];
#pragma unroll
    for (int i = 0; i < NB; i++) {
        const int j = t + i * kGenT;
        if (j < nb) {
#pragma unroll
            for (int q = 0; q < P; q++) a[i][q] = buf[j + q * nb];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NB; i++) {
        const int j = t + i * kGenT;
        if (j < nb) {
            const int k = j % ns;
            if (ns > 1) {
#pragma unroll
                for (int q = 1; q < P; q++) {
                    const T2 w = tw[(long long)q * k * step % n];
                    a[i][q] = c_mul_tw<SIGN>(a[i][q], w);
                }
            }
            T2 b[P];
            if constexpr (P == 2 || P == 4 || P == 8) {
#pragma unroll
                for (int q = 0; q < P; q++) b[q] = a[i][q];
                dft<P, SIGN>(b);
            } else {
                const int root = n / P;  // table stride of exp(2 pi i / P)
#pragma unroll
                for (int r = 0; r < P; r++) {
                    T2 acc = a[i][0];
#pragma unroll
                    for (int q = 1; q < P; q++) {
                        const T2 w = tw[(long long)((q * r) % P) * root];
                        const T2 x = c_mul_tw<SIGN>(a[i][q], w);
                        acc.x += x.x;
                        acc.y += x.y;
                    }
                    b[r] = acc;
                }
            }
            const int base = (j / ns) * ns * P + k;
#pragma unroll
            for (int r = 0; r < P; r++) buf[base + r * ns] = b[r];
        }
    }
    __syncthreads();
}

// lines: line l (l < nlines) starts at in/out + (l / per) * plane + (l % per) * lstride and
// its elements are estride apart (rows: per = lines, lstride = N, estride = 1; columns of
// [nw][ny][nx] planes: per = nx, lstride = 1, estride = nx, plane = nx*ny)
template <typename Real, int SIGN>
__global__ __launch_bounds__(kGenT) void k_fft_generic(const C2<Real>* in, C2<Real>* out,
                                                       const C2<Real>* tw, GenPlan plan,
                                                       long long per, long long plane,
                                                       long long lstride, long long estride,
                                                       Real scale)
{
    using T2 = C2<Real>;
    __shared__ T2 buf[kGenMaxN];
    const long long l = blockIdx.x;
    const long long base = (l / per) * plane + (l % per) * lstride;
    const int n = plan.n, t = threadIdx.x;
    for (int e = t; e < n; e += kGenT) buf[e] = in[base + e * estride];
    __syncthreads();
    int ns = 1;
    for (int s = 0; s < plan.nstages; s++) {
        const int p = plan.radix[s];
        switch (p) {
            case 2: gen_stage<2, SIGN>(buf, n, ns, t, tw); break;
            case 3: gen_stage<3, SIGN>(buf, n, ns, t, tw); break;
            case 4: gen_stage<4, SIGN>(buf, n, ns, t, tw); break;
            case 5: gen_stage<5, SIGN>(buf, n, ns, t, tw); break;
            case 7: gen_stage<7, SIGN>(buf, n, ns, t, tw); break;
            case 8: gen_stage<8, SIGN>(buf, n, ns, t, tw); break;
            case 11: gen_stage<11, SIGN>(buf, n, ns, t, tw); break;
            default: gen_stage<13, SIGN>(buf, n, ns, t, tw); break;
        }
        ns *= p;
    }
    for (int e = t; e < n; e += kGenT) {
        T2 x = buf[e];
        x.x *= scale;
        x.y *= scale;
        out[base + e * estride] = x;
    }
}

// radices of n (8s, then 4, 2, 3, 5, 7, 11, 13); false when another prime divides n
bool generic_plan(int n, GenPlan& p)
{
    if (n < 2 || n > kGenMaxN) return false;
    p.n = n;
    p.nstages = 0;
    int m = n;
    const int order[] = {8, 4, 2, 3, 5, 7, 11, 13};
    for (int r : order) {
        while (m % r == 0) {
            if (p.nstages == kGenMaxStages) return false;
            p.radix[p.nstages++] = r;
            m /= r;
        }
    }
    return m == 1;
}

bool fft_size_ok(int n)
{
    if (n >= 2 && n <= 8192 && (n & (n - 1)) == 0) return true;
    GenPlan p;
    return generic_plan(n, p);
}

template <typename Real>
hipError_t fft_generic(int sign, const void* in, void* out, const void* tw, int n, long long lines,
                       long long per, long long plane, long long lstride, long long estride,
                       double scale, hipStream_t st)
{
    GenPlan p;
    if (!generic_plan(n, p)) return hipErrorInvalidValue;
    if (lines <= 0) return hipSuccess;
    using T2 = C2<Real>;
    if (sign > 0)
        hipLaunchKernelGGL((k_fft_generic<Real, 1>), dim3((unsigned)lines), dim3(kGenT), 0, st,
                           (const T2*)in, (T2*)out, (const T2*)tw, p, per, plane, lstride,
                           estride, (Real)scale);
    else
        hipLaunchKernelGGL((k_fft_generic<Real, -1>), dim3((unsigned)lines), dim3(kGenT), 0, st,
                           (const T2*)in, (T2*)out, (const T2*)tw, p, per, plane, lstride,
                           estride, (Real)scale);
    return hipGetLastError();
}

template hipError_t fft_generic<double>(int, const void*, void*, const void*, int, long long,
                                        long long, long long, long long, long long, double,
                                        hipStream_t);
template hipError_t fft_generic<float>(int, const void*, void*, const void*, int, long long,
                                       long long, long long, long long, long long, double,
                                       hipStream_t);

template <typename Real>
__global__ __launch_bounds__(256) void k_twiddles(C2<Real>* tw, int n)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    double sn, cs;
    sincospi(2.0 * (double)m / (double)n, &sn, &cs);
    tw[m] = C2<Real>{(Real)cs, (Real)sn};
}

hipError_t fill_twiddles(int precision, int n, void* tw, hipStream_t st)
{
    const unsigned blocks = (unsigned)((n + 255) / 256);
    if (precision == LOFAR_FP64)
        hipLaunchKernelGGL((k_twiddles<double>), dim3(blocks), dim3(256), 0, st, (double2*)tw, n);
    else
        hipLaunchKernelGGL((k_twiddles<float>), dim3(blocks), dim3(256), 0, st, (float2*)tw, n);
    return hipGetLastError();
}

int device_cus()
{
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        hipDeviceProp_t p;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess)
            cus = p.multiProcessorCount;
        if (cus <= 0) cus = 256;
    }
    return cus;
}

template hipError_t fft_rows<double>(int, const void*, void*, const void*, int, long long, double,
                                     hipStream_t);
template hipError_t fft_rows<float>(int, const void*, void*, const void*, int, long long, double,
                                    hipStream_t);
template hipError_t fft_cols<double>(int, const void*, void*, const void*, int, int, int, double,
                                     hipStream_t, const unsigned long long*, int);
template hipError_t fft_cols<float>(int, const void*, void*, const void*, int, int, int, double,
                                    hipStream_t, const unsigned long long*, int);
template hipError_t rows_phase<double>(const void*, const void*, int, int, int, const PhaseParams&,
                                       double, double*, double*, int, hipStream_t, int);
template hipError_t rows_phase<float>(const void*, const void*, int, int, int, const PhaseParams&,
                                      double, double*, double*, int, hipStream_t, int);
template hipError_t rows_phase_adj<double>(void*, const void*, int, int, int, const PhaseParams&,
                                           double, const double*, const double*, hipStream_t);
template hipError_t rows_phase_adj<float>(void*, const void*, int, int, int, const PhaseParams&,
                                          double, const double*, const double*, hipStream_t);

}  // namespace lofar
