/*
 * liblofar_hip -- MI355X (gfx950) w-stacking imaging hot path behind a C ABI.
 *
 * Three groups of entry points:
 *   1. Drop-ins with the reference's exact signatures (host pointers, void
 *      return): wstack(), phase_correction(), test().
 *   2. Host-pointer extensions: lofar_fft_planes() gives the per-plane FFT of
 *      fftw_data() a function boundary; error and profiling queries.
 *   3. Device-resident entry points (lofar_*_dev): every array argument is
 *      device memory, `stream` is a hipStream_t (NULL = the null stream), work
 *      is enqueued on that stream and the call returns an int status.
 *
 * Cube layout everywhere: [w][v][u], u fastest; FP64 mode stores interleaved
 * (re,im) doubles (16 B per cell, exactly the reference's `grid`), FP32 mode
 * stores interleaved floats (8 B per cell).  Images are two separate double
 * arrays [v][u] (real, imaginary), as phase_correction() writes them.
 */
#ifndef LOFAR_HIP_H
#define LOFAR_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define LOFAR_OK 0
#define LOFAR_EARG (-1)    /* bad scalar argument or NULL pointer */
#define LOFAR_ECOORD (-2)  /* a (u,v,w) sample is not finite or lies outside the w-planes */
#define LOFAR_ESIZE (-3)   /* FFT side outside [2, 8192] or with a prime factor above 13 */
#define LOFAR_EHIP (-4)    /* HIP runtime error */
#define LOFAR_ENOMEM (-5)  /* device allocation failed */
#define LOFAR_ENODEV (-6)  /* no usable gfx950 device */
#define LOFAR_ESCRATCH (-7) /* the call needs more workspace than lofar_image_dev_static's
                               static array holds */

/* precision modes of the device entry points */
#define LOFAR_FP64 0  /* double cube, double FFT, double phase and plane sum */
#define LOFAR_FP32C 1 /* float cube and FFT; FP64 LDS accumulation, w-phase and plane sum */

/* ------------------------------------------------------------------ */
/* 1. drop-ins                                                         */

/* Replaces wstack() of w-stacking.h:12-29 (CPU loop w-stacking_omp.c:25-145,
 * CUDA w-stacking.cu:123-220).  Accumulates (+=) into the host `grid`
 * [2*num_w_planes*grid_size_y*grid_size_x] like the CPU reference.  num_threads
 * is accepted and ignored.  On a bad sample nothing is written, an error is
 * printed once and lofar_last_error() returns LOFAR_ECOORD. */
void wstack(int num_w_planes, long num_points, long freq_per_chan, long polarizations,
            double* uu, double* vv, double* ww, float* vis_real, float* vis_img, float* weight,
            double dx, double dw, int w_support, int grid_size_x, int grid_size_y,
            double* grid, int num_threads);

/* Replaces phase_correction() of w-stacking.h:39-51 (phase_correction.cu:89-215,
 * PHASE_ON build).  Accumulates (+=) into the host image arrays [yaxis*xaxis]. */
void phase_correction(double* gridss, double* image_real, double* image_imag,
                      int xaxis, int yaxis, int num_w_planes, int xaxistot, int yaxistot,
                      double resolution, double wmin, double wmax, int num_threads);

/* Replaces test() of w-stacking.h:34 (w-stacking_omp.c:241-247): returns nnn+1. */
int test(int nnn);

/* ------------------------------------------------------------------ */
/* 2. host-pointer extensions                                          */

/* The per-plane transform of fftw_data() (fourier_transform.c:7-74), which has
 * no function boundary in the reference: for each of nw planes,
 * gridss = FFTW_BACKWARD(grid) / (nx*ny) over dims (n0=ny, n1=nx).  nx and ny
 * are in [2, 8192] with no prime factor above 13 (powers of two take the radix-8/16
 * kernels, other sides mixed-radix 2/3/4/5/7/8/11/13 ones).  grid and gridss may
 * alias. */
int lofar_fft_planes(const double* grid, double* gridss, int nx, int ny, int nw);

/* Status of the calling thread's last drop-in call (LOFAR_OK if it succeeded). */
int lofar_last_error(void);
const char* lofar_error_string(int code);
/* Library version as 10000*major + 100*minor + patch. */
int lofar_version(void);

/* Per-stage kernel timing with HIP events on the stream the kernels run on.
 * lofar_profile_enable(1) starts recording; lofar_profile_read() waits for the
 * recorded work, writes total milliseconds and launch counts per stage id
 * (arrays of at least LOFAR_NSTAGES) and clears the record. */
#define LOFAR_NSTAGES 10
int lofar_profile_enable(int on);
int lofar_profile_read(double* total_ms, int* launches);
const char* lofar_stage_name(int stage);

/* ------------------------------------------------------------------ */
/* 3. device-resident entry points                                     */

/* Gridding (wstack semantics) into a device cube of num_w_planes*grid_size_y*
 * grid_size_x cells (double2 for LOFAR_FP64, float2 for LOFAR_FP32C).  With
 * accumulate=0 the whole cube is overwritten, otherwise the gridded values are
 * added to it. */
int lofar_grid_dev(int precision, int accumulate, int num_w_planes, long num_points,
                   long freq_per_chan, long polarizations, const double* uu, const double* vv,
                   const double* ww, const float* vis_real, const float* vis_img,
                   const float* weight, double dx, double dw, int w_support, int grid_size_x,
                   int grid_size_y, void* cube, void* stream);

/* out[w] = scale * DFT(in[w]) with exp(sign*2*pi*i*...) per plane; sign=+1 is
 * FFTW_BACKWARD.  in and out may alias. */
int lofar_fft_dev(int precision, int sign, const void* in, void* out, int nx, int ny, int nw,
                  double scale, void* stream);

/* phase_correction() on a device cube; accumulate=0 overwrites the image. */
int lofar_phase_dev(int precision, const void* cube, double* image_real, double* image_imag,
                    int xaxis, int yaxis, int num_w_planes, int xaxistot, int yaxistot,
                    double resolution, double wmin, double wmax, int accumulate, void* stream);

/* Whole hot path: grid -> per-plane FFTW_BACKWARD/(nx*ny) -> phase correction
 * and plane sum.  `cube` is caller-provided scratch of the gridding size; the
 * last pass fuses the row transforms with the phase screen and the plane sum. */
int lofar_image_dev(int precision, int num_w_planes, long num_points, long freq_per_chan,
                    long polarizations, const double* uu, const double* vv, const double* ww,
                    const float* vis_real, const float* vis_img, const float* weight, double dx,
                    double dw, int w_support, int grid_size_x, int grid_size_y, double resolution,
                    double wmin, double wmax, void* cube, double* image_real, double* image_imag,
                    int accumulate, void* stream);

/* lofar_image_dev without allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, kernels and results
 * as lofar_image_dev; the library's workspace (bucket tables, sorted entries,
 * per-tile occupancy bits, twiddle tables) comes from one static device array of
 * LOFAR_STATIC_SCRATCH_BYTES per GPU (held until the process ends) instead of
 * hipMalloc, and twiddles are computed on the device.  A call that needs more returns
 * LOFAR_ESCRATCH before writing the image (about 3M samples on 4096x4096 x 32 planes in
 * FP64 mode).  The caller still provides `cube` and the image.  Every call on one GPU
 * shares that GPU's array and leaves the last call's contents in it: calls on one device
 * are serialised by the library (its per-device lock and stream event), and overlapping
 * work that bypassed that ordering would race on the array. */
#define LOFAR_STATIC_SCRATCH_BYTES (256ull << 20)
int lofar_image_dev_static(int precision, int num_w_planes, long num_points,
                           long freq_per_chan, long polarizations, const double* uu,
                           const double* vv, const double* ww, const float* vis_real,
                           const float* vis_img, const float* weight, double dx, double dw,
                           int w_support, int grid_size_x, int grid_size_y, double resolution,
                           double wmin, double wmax, void* cube, double* image_real,
                           double* image_imag, int accumulate, void* stream);

/* Adjoint of phase_correction(): cube[w] = image * conj(phase_w) * dwnorm*sqrt(1-l^2-m^2). */
int lofar_phase_adjoint_dev(int precision, const double* image_real, const double* image_imag,
                            void* cube, int xaxis, int yaxis, int num_w_planes, int xaxistot,
                            int yaxistot, double resolution, double wmin, double wmax,
                            void* stream);

/* Adjoint of gridding (Gaussian interpolation at every sample): model[i*F*P+f*P+p]
 * = weight[i*P+p] * sum_window cw * cube[w_i]; model arrays are double for
 * LOFAR_FP64 and float for LOFAR_FP32C. */
int lofar_degrid_cube_dev(int precision, int num_w_planes, long num_points, long freq_per_chan,
                          long polarizations, const double* uu, const double* vv, const double* ww,
                          const float* weight, double dx, double dw, int w_support,
                          int grid_size_x, int grid_size_y, const void* cube, void* model_real,
                          void* model_imag, void* stream);

/* Exact adjoint of lofar_image_dev: phase adjoint fused with the row transforms
 * (FFTW_FORWARD/(nx*ny)), column transforms, then interpolation. */
int lofar_degrid_dev(int precision, int num_w_planes, long num_points, long freq_per_chan,
                     long polarizations, const double* uu, const double* vv, const double* ww,
                     const float* weight, double dx, double dw, int w_support, int grid_size_x,
                     int grid_size_y, double resolution, double wmin, double wmax,
                     const double* image_real, const double* image_imag, void* cube,
                     void* model_real, void* model_imag, void* stream);

/* ------------------------------------------------------------------ */
/* 4. plans: one sort of fixed samples, kept for repeated calls        */

/* No counterpart in the reference, which grids once per run
 * (gridding.c:245-261): iterative callers (chi^2 fits, major cycles) grid and
 * degrid the same (u, v, w) many times with new visibilities or a new image.
 * lofar_plan_create validates and sorts the samples once, for gridding (each
 * sample's index in (w, tile) order) and for degridding (home tiles), into device
 * memory the plan owns (about 20 bytes per sample in FP32C mode, 32 in FP64, plus
 * 32 per sample for degridding), and copies the weights it needs; the caller's
 * arrays may change or be freed afterwards.  lofar_plan_image and
 * lofar_plan_degrid then run lofar_image_dev and lofar_degrid_dev on the planned
 * samples without sorting and without host synchronisation: the same results, bit
 * for bit.  The degridding weights are those given to lofar_plan_create; the
 * imaging weights come with each call.  Windows wider than 17 cells keep copies of
 * uu, vv, ww and take the unsorted paths.  A plan belongs to the device current at
 * its creation (LOFAR_EARG elsewhere).  FP32C plans keep sample indices in 32 bits:
 * 2^32 samples or more return LOFAR_ESIZE. */
typedef struct lofar_plan_s* lofar_plan;
int lofar_plan_create(int precision, int num_w_planes, long num_points, long freq_per_chan,
                      long polarizations, const double* uu, const double* vv, const double* ww,
                      const float* weight, double dx, double dw, int w_support, int grid_size_x,
                      int grid_size_y, void* stream, lofar_plan* plan);
int lofar_plan_image(lofar_plan plan, const float* vis_real, const float* vis_img,
                     const float* weight, double resolution, double wmin, double wmax,
                     void* cube, double* image_real, double* image_imag, int accumulate,
                     void* stream);
int lofar_plan_degrid(lofar_plan plan, double resolution, double wmin, double wmax,
                      const double* image_real, const double* image_imag, void* cube,
                      void* model_real, void* model_imag, void* stream);
/* Frees the plan's device memory after the work queued on it has finished. */
int lofar_plan_destroy(lofar_plan plan);

#ifdef __cplusplus
}
#endif
#endif
