// Standalone w-phase correction and plane sum (the phase_correction() drop-in)
// and its adjoint.  One thread per image pixel loops over the planes, so the
// sum is formed in registers in plane order: no atomics, deterministic.
#include "fft_kernels.h"
#include "grid_kernels.h"

namespace lofar {

// phase_correction.cu:151-210 (PHASE_ON): image += sum_w plane_w * exp(i*phi_w) *
// dwnorm * sqrt(1-r^2), phi_w = 2*PI*(wmin + w*dw)*(sqrt(1-r^2)-1); plain sum
// when there is a single plane.  FP64 cubes (the phase_correction() drop-in) evaluate
// every plane's factor directly; complex-float cubes (the 8192-wide split row passes,
// where the direct FP64 sincos per pixel and plane would bound the pass) advance it by
// the FP64 rotation recurrence of the fused row pass (phase_rotation).
template <typename Real>
__global__ __launch_bounds__(256) void k_phase_sum(const typename cplx<Real>::type* cube,
                                                   double* img_re, double* img_im, int xaxis,
                                                   int yaxis, int nw, PhaseParams pp,
                                                   int accumulate)
{
    using T2 = typename cplx<Real>::type;
    const long long npix = (long long)xaxis * yaxis;
    const long long pix = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= npix) return;
    const int iu = (int)(pix % xaxis), iv = (int)(pix / xaxis);
    double fac = 1.0;
    if (pp.phase_on) {
        const double xcoord = pp.coord(iu, pp.xtot);
        const double ycoord = pp.coord(iv, pp.ytot);
        const double radius2 = (xcoord * xcoord + ycoord * ycoord);
        fac = sqrt(1 - radius2);
    }
    double sr = 0.0, si = 0.0;
    if constexpr (sizeof(Real) == 4) {
        double2 z, stp;
        phase_rotation(pp, fac, z, stp);
        for (int iw = 0; iw < nw; iw++) {
            const T2 x = cube[(long long)iw * npix + pix];
            const double p = (double)x.x, q = (double)x.y;
            sr = fma(p, z.x, fma(-q, z.y, sr));
            si = fma(p, z.y, fma(q, z.x, si));
            z = cmul(z, stp);
        }
    } else
    for (int iw = 0; iw < nw; iw++) {
        const T2 x = cube[(long long)iw * npix + pix];
        const double p = (double)x.x, q = (double)x.y;
        if (pp.phase_on) {
            const double wterm = pp.wmin + iw * pp.dw;
            double s, r;
            sincos_w(2.0 * LOFAR_REF_PI * wterm * (fac - 1.0), &s, &r);
            sr += (p * r - q * s) * pp.dwnorm * fac;
            si += (p * s + q * r) * pp.dwnorm * fac;
        } else {
            sr += p;
            si += q;
        }
    }
    if (accumulate) {
        img_re[pix] += sr;
        img_im[pix] += si;
    } else {
        img_re[pix] = sr;
        img_im[pix] = si;
    }
}

// adjoint: plane_w = image * conj(exp(i*phi_w)) * dwnorm * sqrt(1-r^2)
template <typename Real>
__global__ __launch_bounds__(256) void k_phase_adjoint(const double* img_re, const double* img_im,
                                                       typename cplx<Real>::type* cube, int xaxis,
                                                       int yaxis, int nw, PhaseParams pp)
{
    using T2 = typename cplx<Real>::type;
    const long long npix = (long long)xaxis * yaxis;
    const long long pix = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= npix) return;
    const int iu = (int)(pix % xaxis), iv = (int)(pix / xaxis);
    double fac = 1.0;
    if (pp.phase_on) {
        const double xcoord = pp.coord(iu, pp.xtot);
        const double ycoord = pp.coord(iv, pp.ytot);
        const double radius2 = (xcoord * xcoord + ycoord * ycoord);
        fac = sqrt(1 - radius2);
    }
    const double a = img_re[pix], b = img_im[pix];
    if constexpr (sizeof(Real) == 4) {
        double2 z, stp;
        phase_rotation(pp, fac, z, stp);
        for (int iw = 0; iw < nw; iw++) {
            // image * conj(z)
            cube[(long long)iw * npix + pix] =
                T2{(Real)(a * z.x + b * z.y), (Real)(b * z.x - a * z.y)};
            z = cmul(z, stp);
        }
        return;
    }
    for (int iw = 0; iw < nw; iw++) {
        double xr = a, xi = b;
        if (pp.phase_on) {
            const double wterm = pp.wmin + iw * pp.dw;
            double s, r;
            sincos_w(2.0 * LOFAR_REF_PI * wterm * (fac - 1.0), &s, &r);
            const double f = pp.dwnorm * fac;
            xr = (a * r + b * s) * f;
            xi = (b * r - a * s) * f;
        }
        cube[(long long)iw * npix + pix] = T2{(Real)xr, (Real)xi};
    }
}

static PhaseParams make_pp(int nw, int xtot, int ytot, double resolution, double wmin, double wmax)
{
    PhaseParams pp;
    pp.xtot = xtot;
    pp.ytot = ytot;
    pp.resolution = resolution;
    pp.wmin = wmin;
    pp.dw = (wmax - wmin) / (double)nw;  // phase_correction.cu:92
    pp.dwnorm = pp.dw / (wmax - wmin);   // phase_correction.cu:94
    pp.phase_on = nw > 1;                // phase_correction.cu:162
    return pp;
}

hipError_t phase_sum(int precision, const void* cube, double* re, double* im, int xaxis, int yaxis,
                     int nw, int xtot, int ytot, double resolution, double wmin, double wmax,
                     int accumulate, hipStream_t st)
{
    const PhaseParams pp = make_pp(nw, xtot, ytot, resolution, wmin, wmax);
    const long long npix = (long long)xaxis * yaxis;
    const unsigned blocks = (unsigned)((npix + 255) / 256);
    if (npix == 0) return hipSuccess;
    if (precision == LOFAR_FP64)
        hipLaunchKernelGGL(k_phase_sum<double>, dim3(blocks), dim3(256), 0, st,
                           (const double2*)cube, re, im, xaxis, yaxis, nw, pp, accumulate);
    else
        hipLaunchKernelGGL(k_phase_sum<float>, dim3(blocks), dim3(256), 0, st, (const float2*)cube,
                           re, im, xaxis, yaxis, nw, pp, accumulate);
    return hipGetLastError();
}

hipError_t phase_adjoint(int precision, const double* re, const double* im, void* cube, int xaxis,
                         int yaxis, int nw, int xtot, int ytot, double resolution, double wmin,
                         double wmax, hipStream_t st)
{
    const PhaseParams pp = make_pp(nw, xtot, ytot, resolution, wmin, wmax);
    const long long npix = (long long)xaxis * yaxis;
    const unsigned blocks = (unsigned)((npix + 255) / 256);
    if (npix == 0) return hipSuccess;
    if (precision == LOFAR_FP64)
        hipLaunchKernelGGL(k_phase_adjoint<double>, dim3(blocks), dim3(256), 0, st, re, im,
                           (double2*)cube, xaxis, yaxis, nw, pp);
    else
        hipLaunchKernelGGL(k_phase_adjoint<float>, dim3(blocks), dim3(256), 0, st, re, im,
                           (float2*)cube, xaxis, yaxis, nw, pp);
    return hipGetLastError();
}

PhaseParams phase_params(int nw, int xtot, int ytot, double resolution, double wmin, double wmax)
{
    return make_pp(nw, xtot, ytot, resolution, wmin, wmax);
}

}  // namespace lofar
