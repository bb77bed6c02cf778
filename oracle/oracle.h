/*
 * CPU oracle for the w-stacking imaging hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code: only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py use it, and only as the checker.
 *
 * A clean-room restatement, in plain C and double precision, of
 *   - wstack()            w-stacking_omp.c:25-145 (CPU loop; same body as w-stacking.cu:222-307)
 *   - fftw_data()         fourier_transform.c:7-74 (per-plane FFTW_BACKWARD, scaled by 1/(Nx*Ny))
 *   - phase_correction()  phase_correction.cu:89-95,145-210 (PHASE_ON build)
 * plus the exact adjoint of their composition (degridding), which has no
 * counterpart in the reference.  All indices are 64-bit.
 */
#ifndef LOFAR_ORACLE_H
#define LOFAR_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes, mirrored by the product library's LOFAR_E* values */
#define ORACLE_OK 0
#define ORACLE_EARG (-1)   /* bad scalar argument */
#define ORACLE_ECOORD (-2) /* a (u,v,w) sample outside the cube or not finite */

/* Validates every sample the way the product library does, then grids with the
 * reference arithmetic and accumulates (+=) into grid[2*nw*ny*nx] (interleaved
 * re,im; layout [w][v][u], u fastest).  On error nothing is written. */
int oracle_wstack(int num_w_planes, long num_points, long freq_per_chan, long polarizations,
                  const double* uu, const double* vv, const double* ww,
                  const float* vis_real, const float* vis_img, const float* weight,
                  double dx, double dw, int w_support, int grid_size_x, int grid_size_y,
                  double* grid);

/* out[plane] = scale * sum_{v,u} in[plane][v][u] * exp(sign*2*pi*i*(u*x/nx + v*y/ny)).
 * sign=+1 is FFTW_BACKWARD, sign=-1 FFTW_FORWARD.  Any nx, ny >= 1 (radix-2 for
 * powers of two, direct DFT otherwise).  in and out may alias. */
int oracle_fft_planes(const double* in, double* out, int nx, int ny, int nw, int sign,
                      double scale);

/* phase_correction.cu:89-210 with PHASE_ON when phase_on != 0; accumulates (+=). */
void oracle_phase_correction(const double* gridss, double* image_real, double* image_imag,
                             int xaxis, int yaxis, int num_w_planes, int xaxistot, int yaxistot,
                             double resolution, double wmin, double wmax, int phase_on);

/* Adjoint of oracle_phase_correction: planes[w] = image * conj(phase_w) * factor,
 * with the same per-pixel factor; overwrites planes[2*nw*yaxis*xaxis]. */
void oracle_phase_adjoint(const double* image_real, const double* image_imag, double* planes,
                          int xaxis, int yaxis, int num_w_planes, int xaxistot, int yaxistot,
                          double resolution, double wmin, double wmax, int phase_on);

/* Adjoint of oracle_wstack (Gaussian interpolation): for every sample i, channel f
 * and polarisation p, model[i*F*P+f*P+p] = weight[i*P+p] * sum_window cw * plane[w_i].
 * Overwrites model_real/model_imag.  Same validation as oracle_wstack. */
int oracle_degrid(int num_w_planes, long num_points, long freq_per_chan, long polarizations,
                  const double* uu, const double* vv, const double* ww, const float* weight,
                  double dx, double dw, int w_support, int grid_size_x, int grid_size_y,
                  const double* grid, double* model_real, double* model_imag);

/* resolution as computed at gridding.c:160, including C's integer abs(). */
double oracle_resolution(double uvmin, double uvmax);

#ifdef __cplusplus
}
#endif
#endif
