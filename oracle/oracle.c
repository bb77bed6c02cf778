/*
 * CPU oracle for the w-stacking imaging hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
 * file, and only as the checker; the product library never links or calls it.
 * Every function restates the reference's arithmetic in double precision with
 * 64-bit indices; the citations name the reference lines each one follows.
 * Build: oracle/Makefile (gcc, -ffp-contract=off so that the expression order
 * written here is the order evaluated).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* w-stacking.h:6 -- the reference's own value of pi (not M_PI) */
#define REF_PI 3.14159265359

/* Gaussian convolution weight, w-stacking_omp.c:13-18 */
static double gauss_kernel_norm(double norm, double std22, double u_dist, double v_dist)
{
    return norm * exp(-((u_dist * u_dist) + (v_dist * v_dist)) * std22);
}

/* Validation shared with the product library: the reference performs none and
 * writes outside the cube for out-of-range w (w-stacking_omp.c:88,109). */
static int check_samples(int nw, long n, long F, long P, const double* uu, const double* vv,
                         const double* ww, double dx, double dw, int support, int nx, int ny)
{
    if (nw < 1 || n < 0 || F < 1 || P < 1 || support < 1 || nx < 1 || ny < 1) return ORACLE_EARG;
    if (!(dx > 0.0) || !(dw > 0.0) || !isfinite(dx) || !isfinite(dw)) return ORACLE_EARG;
    const double lim = 1073741824.0; /* 2^30: keeps the reference's (int) casts defined */
    for (long i = 0; i < n; i++) {
        if (!isfinite(uu[i]) || !isfinite(vv[i]) || !isfinite(ww[i])) return ORACLE_ECOORD;
        double pu = uu[i] / dx, pv = vv[i] / dx, pw = ww[i] / dw;
        if (!(fabs(pu) < lim) || !(fabs(pv) < lim)) return ORACLE_ECOORD;
        if (!(pw > -1.0) || !(pw < (double)nw)) return ORACLE_ECOORD;
    }
    return ORACLE_OK;
}

/* w-stacking_omp.c:25-145 (serial CPU loop). */
int oracle_wstack(int num_w_planes, long num_points, long freq_per_chan, long polarizations,
                  const double* uu, const double* vv, const double* ww,
                  const float* vis_real, const float* vis_img, const float* weight,
                  double dx, double dw, int w_support, int grid_size_x, int grid_size_y,
                  double* grid)
{
    int st = check_samples(num_w_planes, num_points, freq_per_chan, polarizations, uu, vv, ww,
                           dx, dw, w_support, grid_size_x, grid_size_y);
    if (st != ORACLE_OK) return st;

    /* kernel constants, w-stacking_omp.c:49-52 */
    const int KernelLen = (w_support - 1) / 2;
    const double std = 1.0;
    const double std22 = 1.0 / (2.0 * std * std);
    const double norm = std22 / REF_PI;
    const long nx = grid_size_x, ny = grid_size_y;

    for (long i = 0; i < num_points; i++) {
        const long visindex = i * freq_per_chan * polarizations;
        /* grid coordinates, :85-90 (v is divided by dx too) */
        const double pos_u = uu[i] / dx;
        const double pos_v = vv[i] / dx;
        const int grid_w = (int)(ww[i] / dw);
        const int grid_u = (int)pos_u;
        const int grid_v = (int)pos_v;
        /* clipped window, :93-96 */
        const long jmin = (grid_u > KernelLen - 1) ? grid_u - KernelLen : 0;
        const long jmax = (grid_u < grid_size_x - KernelLen) ? grid_u + KernelLen : grid_size_x - 1;
        const long kmin = (grid_v > KernelLen - 1) ? grid_v - KernelLen : 0;
        const long kmax = (grid_v < grid_size_y - KernelLen) ? grid_v + KernelLen : grid_size_y - 1;

        for (long k = kmin; k <= kmax; k++) {
            const double v_dist = (double)k + 0.5 - pos_v; /* :104 */
            for (long j = jmin; j <= jmax; j++) {
                const double u_dist = (double)j + 0.5 - pos_u; /* :108 */
                const long iKer = 2 * (j + k * nx + (long)grid_w * nx * ny); /* :109, 64-bit */
                const double conv_weight = gauss_kernel_norm(norm, std22, u_dist, v_dist);
                double add_term_real = 0.0, add_term_img = 0.0;
                long ifine = visindex;
                for (long ifreq = 0; ifreq < freq_per_chan; ifreq++) {
                    long iweight = visindex / freq_per_chan; /* :120 -> i*P */
                    for (long ipol = 0; ipol < polarizations; ipol++) {
                        if (!isnan(vis_real[ifine])) { /* :123, tests the real part only */
                            /* weight*vis is a float product, then promoted (:126-127) */
                            const float wr = weight[iweight] * vis_real[ifine];
                            const float wi = weight[iweight] * vis_img[ifine];
                            add_term_real += (double)wr * conv_weight;
                            add_term_img += (double)wi * conv_weight;
                        }
                        ifine++;
                        iweight++;
                    }
                }
                grid[iKer] += add_term_real; /* :135-138 accumulate */
                grid[iKer + 1] += add_term_img;
            }
        }
    }
    return ORACLE_OK;
}

/* ------------------------------------------------------------------------- */
/* FFT: the transform fourier_transform.c:25-27,47,56-57 asks FFTW for.       */

static int is_pow2(long n) { return n > 0 && (n & (n - 1)) == 0; }

/* tw[k] = exp(sign*2*pi*i*k/n), k in [0,n) */
static void make_twiddles(long n, int sign, double* tw)
{
    for (long k = 0; k < n; k++) {
        double a = 2.0 * M_PI * (double)k / (double)n;
        tw[2 * k] = cos(a);
        tw[2 * k + 1] = (double)sign * sin(a);
    }
}

/* in-place 1D transform of x[2n] using the table tw of length n */
static void fft1d(double* x, long n, const double* tw, double* tmp)
{
    if (is_pow2(n)) {
        /* bit reversal, then iterative radix-2 decimation in time */
        for (long i = 1, j = 0; i < n; i++) {
            long bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) {
                double t0 = x[2 * i], t1 = x[2 * i + 1];
                x[2 * i] = x[2 * j]; x[2 * i + 1] = x[2 * j + 1];
                x[2 * j] = t0; x[2 * j + 1] = t1;
            }
        }
        for (long len = 2; len <= n; len <<= 1) {
            long half = len >> 1, step = n / len;
            for (long b = 0; b < n; b += len) {
                for (long k = 0; k < half; k++) {
                    const double wr = tw[2 * k * step], wi = tw[2 * k * step + 1];
                    double* a = x + 2 * (b + k);
                    double* c = x + 2 * (b + k + half);
                    double tr = c[0] * wr - c[1] * wi;
                    double ti = c[0] * wi + c[1] * wr;
                    c[0] = a[0] - tr; c[1] = a[1] - ti;
                    a[0] = a[0] + tr; a[1] = a[1] + ti;
                }
            }
        }
    } else {
        /* direct DFT with exact index reduction */
        for (long y = 0; y < n; y++) {
            double sr = 0.0, si = 0.0;
            for (long u = 0; u < n; u++) {
                long m = (u * y) % n;
                sr += x[2 * u] * tw[2 * m] - x[2 * u + 1] * tw[2 * m + 1];
                si += x[2 * u] * tw[2 * m + 1] + x[2 * u + 1] * tw[2 * m];
            }
            tmp[2 * y] = sr; tmp[2 * y + 1] = si;
        }
        memcpy(x, tmp, (size_t)(2 * n) * sizeof(double));
    }
}

int oracle_fft_planes(const double* in, double* out, int nx, int ny, int nw, int sign,
                      double scale)
{
    if (nx < 1 || ny < 1 || nw < 1 || (sign != 1 && sign != -1)) return ORACLE_EARG;
    const long NX = nx, NY = ny, plane = NX * NY;
    double* twx = malloc(sizeof(double) * 2 * NX);
    double* twy = malloc(sizeof(double) * 2 * NY);
    long nmax = NX > NY ? NX : NY;
    double* line = malloc(sizeof(double) * 2 * nmax);
    double* tmp = malloc(sizeof(double) * 2 * nmax);
    if (!twx || !twy || !line || !tmp) { free(twx); free(twy); free(line); free(tmp); return ORACLE_EARG; }
    make_twiddles(NX, sign, twx);
    make_twiddles(NY, sign, twy);
    for (long w = 0; w < nw; w++) {
        const double* src = in + 2 * w * plane;
        double* dst = out + 2 * w * plane;
        if (dst != src) memcpy(dst, src, sizeof(double) * 2 * plane);
        /* rows: along u (n1 = Nx, contiguous) */
        for (long v = 0; v < NY; v++) fft1d(dst + 2 * v * NX, NX, twx, tmp);
        /* columns: along v (n0 = Ny) */
        for (long u = 0; u < NX; u++) {
            for (long v = 0; v < NY; v++) {
                line[2 * v] = dst[2 * (v * NX + u)];
                line[2 * v + 1] = dst[2 * (v * NX + u) + 1];
            }
            fft1d(line, NY, twy, tmp);
            for (long v = 0; v < NY; v++) {
                dst[2 * (v * NX + u)] = scale * line[2 * v];
                dst[2 * (v * NX + u) + 1] = scale * line[2 * v + 1];
            }
        }
    }
    free(twx); free(twy); free(line); free(tmp);
    return ORACLE_OK;
}

/* ------------------------------------------------------------------------- */
/* phase_correction.cu:89-95 (constants) and :151-210 (CPU loop, PHASE_ON).  */

/* per-pixel direction cosines, phase_correction.cu:164-172 */
static void pixel_geometry(long iu, long iv, int xaxistot, int yaxistot, double resolution,
                           double* radius2)
{
    double xcoord = (double)(iu - xaxistot / 2);
    if (xcoord < 0.0) xcoord = (double)(iu + xaxistot / 2);
    xcoord = sin(xcoord * resolution);
    double ycoord = (double)(iv - yaxistot / 2);
    if (ycoord < 0.0) ycoord = (double)(iv + yaxistot / 2);
    ycoord = sin(ycoord * resolution);
    *radius2 = (xcoord * xcoord + ycoord * ycoord);
}

void oracle_phase_correction(const double* gridss, double* image_real, double* image_imag,
                             int xaxis, int yaxis, int num_w_planes, int xaxistot, int yaxistot,
                             double resolution, double wmin, double wmax, int phase_on)
{
    const double dw = (wmax - wmin) / (double)num_w_planes; /* :92 */
    const double dwnorm = dw / (wmax - wmin);               /* :94 */
    const long X = xaxis, Y = yaxis;
    for (long iw = 0; iw < num_w_planes; iw++) {
        for (long iv = 0; iv < Y; iv++) {
            for (long iu = 0; iu < X; iu++) {
                const long index = 2 * (iu + iv * X + X * Y * iw);
                const long img_index = iu + iv * X;
                const double wterm = wmin + iw * dw; /* :160, lower plane edge */
                if (phase_on && num_w_planes > 1) {
                    double radius2;
                    pixel_geometry(iu, iv, xaxistot, yaxistot, resolution, &radius2);
                    /* :182-183 (the branch at :173-180 is overwritten) */
                    const double preal = cos(2.0 * REF_PI * wterm * (sqrt(1 - radius2) - 1.0));
                    const double pimag = sin(2.0 * REF_PI * wterm * (sqrt(1 - radius2) - 1.0));
                    const double p = gridss[index], q = gridss[index + 1];
                    const double r = preal, s = pimag;
                    image_real[img_index] += (p * r - q * s) * dwnorm * sqrt(1 - radius2); /* :193 */
                    image_imag[img_index] += (p * s + q * r) * dwnorm * sqrt(1 - radius2); /* :195 */
                } else {
                    image_real[img_index] += gridss[index]; /* :198-200 */
                    image_imag[img_index] += gridss[index + 1];
                }
            }
        }
    }
}

void oracle_phase_adjoint(const double* image_real, const double* image_imag, double* planes,
                          int xaxis, int yaxis, int num_w_planes, int xaxistot, int yaxistot,
                          double resolution, double wmin, double wmax, int phase_on)
{
    const double dw = (wmax - wmin) / (double)num_w_planes;
    const double dwnorm = dw / (wmax - wmin);
    const long X = xaxis, Y = yaxis;
    for (long iw = 0; iw < num_w_planes; iw++) {
        for (long iv = 0; iv < Y; iv++) {
            for (long iu = 0; iu < X; iu++) {
                const long index = 2 * (iu + iv * X + X * Y * iw);
                const long img_index = iu + iv * X;
                const double a = image_real[img_index], b = image_imag[img_index];
                if (phase_on && num_w_planes > 1) {
                    const double wterm = wmin + iw * dw;
                    double radius2;
                    pixel_geometry(iu, iv, xaxistot, yaxistot, resolution, &radius2);
                    const double r = cos(2.0 * REF_PI * wterm * (sqrt(1 - radius2) - 1.0));
                    const double s = sin(2.0 * REF_PI * wterm * (sqrt(1 - radius2) - 1.0));
                    const double f = dwnorm * sqrt(1 - radius2);
                    /* (a+ib)*(r-is)*f */
                    planes[index] = (a * r + b * s) * f;
                    planes[index + 1] = (b * r - a * s) * f;
                } else {
                    planes[index] = a;
                    planes[index + 1] = b;
                }
            }
        }
    }
}

/* Transpose of oracle_wstack's scatter: the same window, weights and sample
 * validation, gathering instead of accumulating. */
int oracle_degrid(int num_w_planes, long num_points, long freq_per_chan, long polarizations,
                  const double* uu, const double* vv, const double* ww, const float* weight,
                  double dx, double dw, int w_support, int grid_size_x, int grid_size_y,
                  const double* grid, double* model_real, double* model_imag)
{
    int st = check_samples(num_w_planes, num_points, freq_per_chan, polarizations, uu, vv, ww,
                           dx, dw, w_support, grid_size_x, grid_size_y);
    if (st != ORACLE_OK) return st;
    const int KernelLen = (w_support - 1) / 2;
    const double std22 = 0.5;
    const double norm = std22 / REF_PI;
    const long nx = grid_size_x, ny = grid_size_y;
    for (long i = 0; i < num_points; i++) {
        const double pos_u = uu[i] / dx;
        const double pos_v = vv[i] / dx;
        const int grid_w = (int)(ww[i] / dw);
        const int grid_u = (int)pos_u;
        const int grid_v = (int)pos_v;
        const long jmin = (grid_u > KernelLen - 1) ? grid_u - KernelLen : 0;
        const long jmax = (grid_u < grid_size_x - KernelLen) ? grid_u + KernelLen : grid_size_x - 1;
        const long kmin = (grid_v > KernelLen - 1) ? grid_v - KernelLen : 0;
        const long kmax = (grid_v < grid_size_y - KernelLen) ? grid_v + KernelLen : grid_size_y - 1;
        double sr = 0.0, si = 0.0;
        for (long k = kmin; k <= kmax; k++) {
            const double v_dist = (double)k + 0.5 - pos_v;
            for (long j = jmin; j <= jmax; j++) {
                const double u_dist = (double)j + 0.5 - pos_u;
                const long iKer = 2 * (j + k * nx + (long)grid_w * nx * ny);
                const double cw = gauss_kernel_norm(norm, std22, u_dist, v_dist);
                sr += cw * grid[iKer];
                si += cw * grid[iKer + 1];
            }
        }
        for (long f = 0; f < freq_per_chan; f++) {
            for (long p = 0; p < polarizations; p++) {
                const long o = i * freq_per_chan * polarizations + f * polarizations + p;
                const double wgt = (double)weight[i * polarizations + p];
                model_real[o] = wgt * sr;
                model_imag[o] = wgt * si;
            }
        }
    }
    return ORACLE_OK;
}

/* gridding.c:160: resolution = 1.0/MAX(abs(uvmin),abs(uvmax)) where abs() is the
 * C integer function, so each bound is first truncated to int. */
double oracle_resolution(double uvmin, double uvmax)
{
    int a = abs((int)uvmin), b = abs((int)uvmax);
    return 1.0 / (double)(a > b ? a : b);
}
