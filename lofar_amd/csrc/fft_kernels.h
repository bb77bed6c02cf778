// Device building blocks of the Stockham FFT and the host dispatch prototypes.
#pragma once
#include "common.h"

namespace lofar {

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n >> 1); }

// Per-line plan: E elements per thread, TPL threads per line; stages of radix
// RX = min(16, E) (E/RX butterflies per thread per stage) and one smaller
// remainder stage (last) when log2(N) is not a multiple of log2(RX).
template <int N, int EREQ>
struct Plan {
    static constexpr int E = N < EREQ ? N : EREQ;
    static constexpr int TPL = N / E;
    static constexpr int RX = E < 16 ? E : 16;
    static constexpr int LOG = ilog2(N);
    static constexpr int LOGR = ilog2(RX);
    static constexpr int R0 = N < RX ? N : RX;  // first stage
    static constexpr int RL = (N <= RX) ? N : ((LOG % LOGR == 0) ? RX : (1 << (LOG % LOGR)));
    // element of the line thread t loads into v[e] before the first stage
    __device__ static __forceinline__ int in_index(int t, int e)
    {
        return t + (e / R0) * TPL + (e % R0) * (N / R0);
    }
    // element of the line thread t holds in v[e] after the last stage
    __device__ static __forceinline__ int out_index(int t, int e)
    {
        return t + (e / RL) * TPL + (e % RL) * (N / RL);
    }
    // LDS slot of element i: one pad slot every R0 elements, so the first
    // exchange (butterfly b writes b*R0..b*R0+R0-1) does not put every lane of
    // a wave on the same banks
    __device__ static __forceinline__ int pad(int i) { return i + (i >> ilog2(R0)); }
    // pad(x + c) = pad(x) + padc(c) when c is a multiple of R0, or when x is and
    // c < R0: the exchanges address their elements as one runtime slot plus
    // compile-time offsets (immediate offsets of the LDS instructions)
    static constexpr int padc(int c) { return c + (c >> ilog2(R0)); }
    static constexpr int LDS = N + N / R0;  // slots per line
};

// Spread consecutive workgroups of one XCD (blockIdx % 8) over a contiguous range.
__device__ __forceinline__ long long xcd_remap(long long b, long long g)
{
    if (g % 8 != 0) return b;
    return (b % 8) * (g / 8) + b / 8;
}

// w-phase geometry of phase_correction.cu:92-94,164-172
struct PhaseParams {
    int xtot, ytot;
    double resolution, wmin, dw, dwnorm;
    int phase_on;
    __device__ __forceinline__ double coord(int i, int tot) const
    {
        double c = (double)(i - tot / 2);
        if (c < 0.0) c = (double)(i + tot / 2);
        return sin(c * resolution);
    }
};

// ---------------------------------------------------------------------------
// complex arithmetic

template <typename T2>
__device__ __forceinline__ T2 c_add(T2 a, T2 b)
{
    return T2{a.x + b.x, a.y + b.y};
}
template <typename T2>
__device__ __forceinline__ T2 c_sub(T2 a, T2 b)
{
    return T2{a.x - b.x, a.y - b.y};
}
// a * w, or a * conj(w) when SIGN < 0
template <int SIGN, typename T2>
__device__ __forceinline__ T2 c_mul_tw(T2 a, T2 w)
{
    if constexpr (SIGN > 0)
        return T2{a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x};
    else
        return T2{a.x * w.x + a.y * w.y, a.y * w.x - a.x * w.y};
}
// a * (S*i)
template <int S, typename T2>
__device__ __forceinline__ T2 c_muli(T2 a)
{
    if constexpr (S > 0)
        return T2{-a.y, a.x};
    else
        return T2{a.y, -a.x};
}

// exp(2*pi*i*m/16) components; values of cos(pi/8), cos(pi/4), cos(3*pi/8)
constexpr double kC1 = 0.92387953251128675613;
constexpr double kC2 = 0.70710678118654752440;
constexpr double kC3 = 0.38268343236508977173;
constexpr double cos16(int m)
{
    m &= 15;
    return m == 0 ? 1.0 : m == 1 ? kC1 : m == 2 ? kC2 : m == 3 ? kC3 : m == 4 ? 0.0
         : m == 5 ? -kC3 : m == 6 ? -kC2 : m == 7 ? -kC1 : m == 8 ? -1.0 : m == 9 ? -kC1
         : m == 10 ? -kC2 : m == 11 ? -kC3 : m == 12 ? 0.0 : m == 13 ? kC3 : m == 14 ? kC2 : kC1;
}
constexpr double sin16(int m) { return cos16(m - 4); }

// a * exp(S*2*pi*i*M/16)
template <int M, int S, typename T2>
__device__ __forceinline__ T2 c_w16(T2 a)
{
    constexpr int m = M & 15;
    if constexpr (m == 0)
        return a;
    else if constexpr (m == 4)
        return c_muli<S>(a);
    else if constexpr (m == 8)
        return T2{-a.x, -a.y};
    else if constexpr (m == 12)
        return c_muli<-S>(a);
    else {
        using R = decltype(a.x);
        constexpr R c = (R)cos16(m);
        constexpr R s = (R)(S * sin16(m));
        return T2{a.x * c - a.y * s, a.x * s + a.y * c};
    }
}

template <int S, typename T2>
__device__ __forceinline__ void dft4(T2& x0, T2& x1, T2& x2, T2& x3)
{
    const T2 t0 = c_add(x0, x2), t1 = c_sub(x0, x2), t2 = c_add(x1, x3);
    const T2 t3 = c_muli<S>(c_sub(x1, x3));
    x0 = c_add(t0, t2);
    x2 = c_sub(t0, t2);
    x1 = c_add(t1, t3);
    x3 = c_sub(t1, t3);
}

template <int N1, int S, typename T2>
__device__ __forceinline__ void tw16_col(T2* v)
{
    // v[4*k1 + N1] *= W16^(N1*k1), k1 = 1..3
    v[4 + N1] = c_w16<N1 * 1, S>(v[4 + N1]);
    v[8 + N1] = c_w16<N1 * 2, S>(v[8 + N1]);
    v[12 + N1] = c_w16<N1 * 3, S>(v[12 + N1]);
}

// In-register DFT of size R (natural order in and out), X_k = sum x_n exp(S*2*pi*i*n*k/R).
template <int R, int S, typename T2>
__device__ __forceinline__ void dft(T2* v)
{
    if constexpr (R == 2) {
        const T2 a = v[0];
        v[0] = c_add(a, v[1]);
        v[1] = c_sub(a, v[1]);
    } else if constexpr (R == 4) {
        dft4<S>(v[0], v[1], v[2], v[3]);
    } else if constexpr (R == 8) {
        // n = 2*n1 + n2: radix-4 over n1, twiddle W8^(n2*k1), radix-2 over n2
        dft4<S>(v[0], v[2], v[4], v[6]);
        dft4<S>(v[1], v[3], v[5], v[7]);
        v[3] = c_w16<2, S>(v[3]);
        v[5] = c_w16<4, S>(v[5]);
        v[7] = c_w16<6, S>(v[7]);
        T2 o[8];
#pragma unroll
        for (int k1 = 0; k1 < 4; k1++) {
            o[k1] = c_add(v[2 * k1], v[2 * k1 + 1]);
            o[k1 + 4] = c_sub(v[2 * k1], v[2 * k1 + 1]);
        }
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = o[k];
    } else if constexpr (R == 16) {
        // n = 4*n1 + n2: radix-4 over n1, twiddle W16^(n2*k1), radix-4 over n2
        dft4<S>(v[0], v[4], v[8], v[12]);
        dft4<S>(v[1], v[5], v[9], v[13]);
        dft4<S>(v[2], v[6], v[10], v[14]);
        dft4<S>(v[3], v[7], v[11], v[15]);
        tw16_col<1, S>(v);
        tw16_col<2, S>(v);
        tw16_col<3, S>(v);
        dft4<S>(v[0], v[1], v[2], v[3]);
        dft4<S>(v[4], v[5], v[6], v[7]);
        dft4<S>(v[8], v[9], v[10], v[11]);
        dft4<S>(v[12], v[13], v[14], v[15]);
        T2 o[16];
#pragma unroll
        for (int k1 = 0; k1 < 4; k1++)
#pragma unroll
            for (int k2 = 0; k2 < 4; k2++) o[k1 + 4 * k2] = v[4 * k1 + k2];
#pragma unroll
        for (int k = 0; k < 16; k++) v[k] = o[k];
    }
}

// Stockham stages from span NS on.  On entry v[m*R+q] holds element b+q*(N/R)
// of butterfly b = t + m*TPL; on return v holds the outputs in the order of
// Plan<N,E>::out_index.  `lds` addresses element i of this line at
// lds[Plan<N,E>::pad(i)*ls] (Plan<N,E>::LDS slots per line).  With ROUNDS > 1
// the lines of a workgroup take turns on a shared LDS buffer: a thread
// exchanges in round my_round only (all threads pass every barrier).
// Twiddles: tw[m] = exp(+2*pi*i*m/N), conjugated for SIGN < 0.
// RT: the twiddles of butterfly m are the powers w^q of one root per butterfly,
// rt[m] = tw[k*N/(NS*R)] (load_roots), taken from registers loaded when the
// kernel starts instead of R-1 table loads at the stage (whose latency a stage
// otherwise waits out); q-1 complex products per power, a few ulp in FP64.
template <typename T2, int N, int EREQ, int SIGN, int NS, int ROUNDS = 1, bool RT = false>
__device__ __forceinline__ void stages(T2* v, T2* lds, int ls, int t, const T2* __restrict__ tw,
                                       int my_round = 0, const T2* rt = nullptr)
{
    using P = Plan<N, EREQ>;
    constexpr int E = P::E;
    constexpr int TPL = P::TPL;
    constexpr int R = (N / NS >= P::RX) ? P::RX : N / NS;
    constexpr int M = E / R;
#pragma unroll
    for (int m = 0; m < M; m++) {
        const int b = t + m * TPL;
        if constexpr (NS > 1) {
            if constexpr (RT) {
                const T2 r = rt[m];
                T2 w = r;
#pragma unroll
                for (int q = 1; q < R; q++) {
                    v[m * R + q] = c_mul_tw<SIGN>(v[m * R + q], w);
                    if (q + 1 < R) w = c_mul_tw<1>(w, r);
                }
            } else {
                const int k = b & (NS - 1);
#pragma unroll
                for (int q = 1; q < R; q++)
                    v[m * R + q] = c_mul_tw<SIGN>(v[m * R + q], tw[q * k * (N / (NS * R))]);
            }
        }
        dft<R, SIGN>(&v[m * R]);
    }
    if constexpr (NS * R < N) {
        constexpr int NS2 = NS * R;
        constexpr int R2 = (N / NS2 >= P::RX) ? P::RX : N / NS2;
        constexpr int M2 = E / R2;
#pragma unroll
        for (int rr = 0; rr < ROUNDS; rr++) {
            __syncthreads();
            static_assert(NS == 1 ? R == P::R0 : NS % P::R0 == 0, "linear write offsets");
            static_assert((N / R2) % P::R0 == 0, "linear read offsets");
            // complex doubles: one runtime slot per butterfly plus compile-time
            // offsets (fewer integer instructions, measured faster); complex floats
            // keep a slot computed per element (the linear form showed no gain there)
            constexpr bool LIN = sizeof(T2) >= 16;
            if (rr == my_round) {
#pragma unroll
                for (int m = 0; m < M; m++) {
                    const int b = t + m * TPL;
                    const int k = b & (NS - 1);
                    const int base = (b - k) * R + k;
                    T2* row = lds + P::pad(base) * ls;
#pragma unroll
                    for (int q = 0; q < R; q++) {
                        if constexpr (LIN)
                            row[P::padc(q * NS) * ls] = v[m * R + q];
                        else
                            lds[P::pad(base + q * NS) * ls] = v[m * R + q];
                    }
                }
            }
            __syncthreads();
            if (rr == my_round) {
#pragma unroll
                for (int m = 0; m < M2; m++) {
                    const int b = t + m * TPL;
                    const T2* row = lds + P::pad(b) * ls;
#pragma unroll
                    for (int q = 0; q < R2; q++) {
                        if constexpr (LIN)
                            v[m * R2 + q] = row[P::padc(q * (N / R2)) * ls];
                        else
                            v[m * R2 + q] = lds[P::pad(b + q * (N / R2)) * ls];
                    }
                }
            }
        }
        stages<T2, N, EREQ, SIGN, NS2, ROUNDS, RT>(v, lds, ls, t, tw, my_round,
                                                  RT ? rt + (NS > 1 ? M : 0) : rt);
    }
}

// Number of twiddle roots stages<..., RT> uses from span NS on (one per butterfly
// of every stage after the first).
template <int N, int EREQ>
constexpr int root_count(int ns)
{
    using P = Plan<N, EREQ>;
    int total = 0;
    while (ns < N) {
        const int r = (N / ns >= P::RX) ? P::RX : N / ns;
        if (ns > 1) total += P::E / r;
        ns *= r;
    }
    return total;
}

// Loads the roots stages<..., RT> consumes, in its order.
template <typename T2, int N, int EREQ, int NS>
__device__ __forceinline__ void load_roots(T2* rt, int t, const T2* __restrict__ tw)
{
    using P = Plan<N, EREQ>;
    constexpr int R = (N / NS >= P::RX) ? P::RX : N / NS;
    constexpr int M = P::E / R;
    if constexpr (NS > 1) {
#pragma unroll
        for (int m = 0; m < M; m++) {
            const int k = (t + m * P::TPL) & (NS - 1);
            rt[m] = tw[k * (N / (NS * R))];
        }
    }
    if constexpr (NS * R < N) load_roots<T2, N, EREQ, NS * R>(rt + (NS > 1 ? M : 0), t, tw);
}

// Number of twiddle factors one thread applies from span NS to the last stage.
template <int N, int EREQ>
constexpr int tw_count(int ns)
{
    using P = Plan<N, EREQ>;
    int total = 0;
    while (ns < N) {
        const int r = (N / ns >= P::RX) ? P::RX : N / ns;
        if (ns > 1) total += (P::E / r) * (r - 1);
        ns *= r;
    }
    return total;
}

// Loads, once, every twiddle factor thread t applies in stages(), in the order
// stages_rt() consumes them, so loops over many lines of the same length keep
// them in registers instead of re-reading the table at every stage.
template <typename T2, int N, int EREQ, int NS>
__device__ __forceinline__ void load_twiddles(T2* twr, int t, const T2* __restrict__ tw)
{
    using P = Plan<N, EREQ>;
    constexpr int R = (N / NS >= P::RX) ? P::RX : N / NS;
    constexpr int M = P::E / R;
    if constexpr (NS > 1) {
#pragma unroll
        for (int m = 0; m < M; m++) {
            const int k = (t + m * P::TPL) & (NS - 1);
#pragma unroll
            for (int q = 1; q < R; q++) twr[m * (R - 1) + q - 1] = tw[q * k * (N / (NS * R))];
        }
    }
    if constexpr (NS * R < N)
        load_twiddles<T2, N, EREQ, NS * R>(twr + (NS > 1 ? M * (R - 1) : 0), t, tw);
}

template <bool LB>
__device__ __forceinline__ void stage_barrier()
{
    if constexpr (LB)
        lds_barrier();
    else
        __syncthreads();
}

// stages() with the twiddle factors taken from registers filled by load_twiddles().
// LB: exchange barriers wait for LDS operations only (lds_barrier).
template <typename T2, int N, int EREQ, int SIGN, int NS, bool LB = false>
__device__ __forceinline__ void stages_rt(T2* v, T2* lds, int t, const T2* twr)
{
    using P = Plan<N, EREQ>;
    constexpr int E = P::E;
    constexpr int TPL = P::TPL;
    constexpr int R = (N / NS >= P::RX) ? P::RX : N / NS;
    constexpr int M = E / R;
#pragma unroll
    for (int m = 0; m < M; m++) {
        if constexpr (NS > 1) {
#pragma unroll
            for (int q = 1; q < R; q++)
                v[m * R + q] = c_mul_tw<SIGN>(v[m * R + q], twr[m * (R - 1) + q - 1]);
        }
        dft<R, SIGN>(&v[m * R]);
    }
    if constexpr (NS * R < N) {
        constexpr int NS2 = NS * R;
        constexpr int R2 = (N / NS2 >= P::RX) ? P::RX : N / NS2;
        constexpr int M2 = E / R2;
        stage_barrier<LB>();
#pragma unroll
        for (int m = 0; m < M; m++) {
            const int b = t + m * TPL;
            const int k = b & (NS - 1);
            const int base = (b - k) * R + k;
#pragma unroll
            for (int q = 0; q < R; q++) lds[P::pad(base + q * NS)] = v[m * R + q];
        }
        stage_barrier<LB>();
#pragma unroll
        for (int m = 0; m < M2; m++) {
            const int b = t + m * TPL;
#pragma unroll
            for (int q = 0; q < R2; q++) v[m * R2 + q] = lds[P::pad(b + q * (N / R2))];
        }
        stages_rt<T2, N, EREQ, SIGN, NS2, LB>(v, lds, t, twr + (NS > 1 ? M * (R - 1) : 0));
    }
}

// sin and cos of x in double precision, from the mathematics: x = n*pi/2 + t
// with a three-part Cody-Waite pi/2 (the first two parts have 33 significant
// bits, so n*part is exact for |n| < 2^20), then Taylor series to t^17 / t^16 on
// |t| <= pi/4 (truncation below 1e-18).  |x| beyond 1e5 falls back to sincos().
constexpr double kPio2Hi = 1.5707963267341256;
constexpr double kPio2Mid = 6.077100506303966e-11;
constexpr double kPio2Lo = 2.0222662487959506e-21;
constexpr double kTwoOverPi = 0.6366197723675814;
__device__ __forceinline__ void sincos_w(double x, double* sn, double* cs)
{
    if (!(fabs(x) < 1.0e5)) {
        sincos(x, sn, cs);
        return;
    }
    const double n = rint(x * kTwoOverPi);
    double t = fma(-n, kPio2Hi, x);
    t = fma(-n, kPio2Mid, t);
    t = fma(-n, kPio2Lo, t);
    const double t2 = t * t;
    // 1/k! for k = 3..17 (odd) and 2..16 (even), Horner in t^2
    double ps = 2.8114572543455206e-15;
    ps = fma(ps, t2, -7.647163731819816e-13);
    ps = fma(ps, t2, 1.6059043836821613e-10);
    ps = fma(ps, t2, -2.505210838544172e-08);
    ps = fma(ps, t2, 2.7557319223985893e-06);
    ps = fma(ps, t2, -0.0001984126984126984);
    ps = fma(ps, t2, 0.008333333333333333);
    ps = fma(ps, t2, -0.16666666666666666);
    const double st = fma(ps * t2, t, t);
    double pc = 4.779477332387385e-14;
    pc = fma(pc, t2, -1.1470745597729725e-11);
    pc = fma(pc, t2, 2.08767569878681e-09);
    pc = fma(pc, t2, -2.755731922398589e-07);
    pc = fma(pc, t2, 2.48015873015873e-05);
    pc = fma(pc, t2, -0.001388888888888889);
    pc = fma(pc, t2, 0.041666666666666664);
    pc = fma(pc, t2, -0.5);
    const double ct = fma(pc, t2, 1.0);
    const int q = ((int)n) & 3;
    const double s_ = (q & 1) ? ct : st;
    const double c_ = (q & 1) ? st : ct;
    *sn = (q & 2) ? -s_ : s_;
    *cs = ((q + 1) & 2) ? -c_ : c_;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return double2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// w-phase factor of one pixel for the plane recurrence: z0 = dwnorm*fac*exp(i*phi_0)
// and step = exp(i*(phi_1-phi_0)), with phi_w = 2*PI*(wmin + w*dw)*(fac-1)
// (phase_correction.cu:160,182-183,193-195), so plane w's factor is z0*step^w.
// Without the phase (a single plane) both are 1.  FP64 throughout.
__device__ __forceinline__ void phase_rotation(const PhaseParams& pp, double fac, double2& z0,
                                               double2& step)
{
    if (!pp.phase_on) {
        z0 = double2{1.0, 0.0};
        step = double2{1.0, 0.0};
        return;
    }
    const double n1 = fac - 1.0;
    double sn, cs;
    sincos_w(2.0 * LOFAR_REF_PI * pp.wmin * n1, &sn, &cs);
    const double f = pp.dwnorm * fac;
    z0 = double2{cs * f, sn * f};
    sincos_w(2.0 * LOFAR_REF_PI * pp.dw * n1, &sn, &cs);
    step = double2{cs, sn};
}

// phase constants of phase_correction.cu:92-94 (phase.hip)
PhaseParams phase_params(int nw, int xtot, int ytot, double resolution, double wmin, double wmax);

// tw[m] = exp(+2*pi*i*m/n), m < n, computed on the device (sincospi in double, then
// rounded to float for LOFAR_FP32C): the allocation-free entry point's twiddle tables
hipError_t fill_twiddles(int precision, int n, void* tw, hipStream_t st);

// Mixed-radix line transforms for sides that are not powers of two (prime factors at most
// 13, side at most 8192): line l of `lines` starts at (l / per) * plane + (l % per) *
// lstride, its elements estride apart; out may alias in.  fft_size_ok: a side either the
// power-of-two kernels or these take.
bool fft_size_ok(int n);
template <typename Real>
hipError_t fft_generic(int sign, const void* in, void* out, const void* tw, int n, long long lines,
                       long long per, long long plane, long long lstride, long long estride,
                       double scale, hipStream_t st);

// host dispatch (fft.hip); all return hipErrorInvalidValue for unsupported sizes
template <typename Real>
hipError_t fft_rows(int sign, const void* in, void* out, const void* tw, int nx, long long lines,
                    double scale, hipStream_t st);
template <typename Real>
// occ (optional): for forward transforms (sign +1), per-tile occupancy bits from
// grid_tiles, rows of empty tiles read as zeros; for inverse transforms (sign -1),
// the tiles to store (need_tiles), other rows left unwritten
// sm_log >= 0: the planes are strip-major, [nx/L][ny][L] with L = 2^sm_log columns
hipError_t fft_cols(int sign, const void* in, void* out, const void* tw, int nx, int ny, int nw,
                    double scale, hipStream_t st, const unsigned long long* occ = nullptr,
                    int sm_log = -1);
template <typename Real>
// sm_log >= 0: the planes are strip-major with strips of 2^sm_log columns
hipError_t rows_phase(const void* cube, const void* tw, int nx, int ny, int nw,
                      const PhaseParams& pp, double scale, double* re, double* im, int accumulate,
                      hipStream_t st, int sm_log = -1);
template <typename Real>
hipError_t rows_phase_adj(void* cube, const void* tw, int nx, int ny, int nw,
                          const PhaseParams& pp, double scale, const double* re, const double* im,
                          hipStream_t st);


}  // namespace lofar
