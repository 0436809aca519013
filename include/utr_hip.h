/*
 * utr_hip.h -- C ABI of the MI355X (gfx950) UTR reconstruction library
 * (libutr_hip.so).  Plain pointers and sizes only; no exceptions cross it.
 *
 * The reference (timg021/UTR_public) has no library API for this path: the
 * seam it replaces is the FFTW plan wrappers of the absent XArrayLibrary and
 * the OpenMP loop over orientations inside UTR.cpp:main.  Each entry point
 * names the reference code it stands in for.
 *
 * Life cycle:  utr_create -> utr_add_images (any number of calls, any row
 * order) -> utr_finalize -> utr_destroy.  Every function returns 0 on success
 * and a negative UTR_E_* code on failure; utr_last_error() then holds the
 * message (thread-local).  A context may be used by one host thread at a
 * time.
 */
#ifndef UTR_HIP_H
#define UTR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UTR_ABI_VERSION 1

/* error codes */
#define UTR_OK 0
#define UTR_E_INVALID (-1)   /* invalid parameter (the reference throws)   */
#define UTR_E_DEVICE (-2)    /* HIP runtime / device failure               */
#define UTR_E_NOMEM (-3)     /* device or host allocation failed           */
#define UTR_E_STATE (-4)     /* call out of order (e.g. rows missing)      */
#define UTR_E_UNSUPPORTED (-5)

/* precision forms */
#define UTR_FP32 0         /* batched FP32 path: brick scatter, exact fixed-point sums  */
#define UTR_FP64 1         /* plain FP64 path: per-row launches, atomic scatter         */
#define UTR_FP32_GATHER 2  /* batched FP32 path: per-voxel gather, Kahan-compensated    */

/*
 * Numeric parameters of the UTR parameter file (ReadmeUTR.txt:33-165, read at
 * UTR.cpp:77-782), in the file's units.  Field order matches the parameter
 * numbers.  For .xri files the caller sets modality = 1, zlo/zhi to the image
 * x-extent and asigma = -1/(delta/beta) (UTR.cpp:308-310,446-451).
 */
typedef struct utr_params {
    int nx, ny;            /* 8: image width/height in pixels (even)        */
    int modality;          /* 1: 0 = TEM/Angstrom, 1 = X-ray/micron         */
    double xst;            /* 9: pixel size (UL)                            */
    double zlo, zhi;       /* 10: output defocus range (UL)                 */
    double dzextra;        /* 10: defocus offset (UL)                       */
    double dxc, dyc, dzc;  /* 11: centre-of-rotation shifts (UL)            */
    double wl;             /* 12: wavelength (UL)                           */
    double aobj_mrad;      /* 13: objective aperture half-angle (mrad)      */
    double cs3_mm, cs5_mm; /* 14: spherical aberrations (mm)                */
    int ctf_mode;          /* 15: 0 none, 1 TIE-Hom-2D, 2 TIE-Hom-3D, 3 CTF-2D, 4 CTF-3D */
    int escc;              /* 16: Ewald sphere curvature correction 0/1     */
    double asigma;         /* 17: beta/delta                                */
    double nsr;            /* 18: average noise-to-signal ratio             */
    double epsilon;        /* 19: Tikhonov regularisation                   */
    int inv_laplace;       /* 22: inverse 3D Laplacian 0/1                  */
    double norm_f1, norm_f2; /* 7: input -> input/f1 + f2                   */
} utr_params;

typedef struct utr_ctx utr_ctx;

/* Phase timers in milliseconds, named as UTR.cpp:1554-1559 prints them. */
typedef struct utr_timings {
    double fft2d_ctf_ms;   /* "2D FFT and CTF correction"                   */
    double backprop_ms;    /* "backpropagation" (gridding + 3D filters)     */
    double fft3d_ms;       /* "3D FFT" (+ mask and scaling)                 */
    double total_ms;       /* add_images + finalize, device time            */
} utr_timings;

/*
 * Creates a reconstruction context on the current HIP device.
 * Replaces the set-up at UTR.cpp:840-867 (Fftwd2fc plan, V3/Samp3
 * allocation) together with the validation of UTR.cpp:55-784.
 * rows: nrows orientation rows (UTR.cpp:122 ReadDefocusParamsFile output):
 * angZ_deg, angY_deg (columns 1-2), angZ2_deg (column 3, may be NULL: the
 * Z'' angle, undone on each image before its 2D stage by a rotation about the
 * rotation centre, UTR.cpp:960-977), defocus (column 4, UL), weights
 * (subseries weights, UTR.cpp:1019; may be NULL).  precision: UTR_FP32 or
 * UTR_FP64.
 */
int utr_create(const utr_params* params, int nrows, const double* angZ_deg,
               const double* angY_deg, const double* angZ2_deg, const double* defocus,
               const float* weights, int precision, utr_ctx** out);

/*
 * Pinned host memory.  utr_add_images, utr_finalize and the other host-
 * memory entry points move data between memory from utr_host_alloc and the
 * device by DMA directly; pageable memory goes through the context's pinned
 * staging slots, filled and drained by a few host threads (UTR_HOST_THREADS,
 * default min(8, hardware threads)).  No counterpart in the reference, which
 * keeps its images in host memory; for callers that read their inputs
 * straight into pinned buffers.
 */
int utr_host_alloc(size_t bytes, void** out);
int utr_host_free(void* p);

/*
 * Supplies the flat-field corrected images of rows [first_row,
 * first_row+count) as count*ny*nx floats (x fastest) in host memory.
 * Replaces the per-row body of the OpenMP loop UTR.cpp:877-1040 up to and
 * including the Fourier transform and 2D filter (Log0/Shuffle/ForwardFFT,
 * UTR.cpp:951-1019).  Images must already be padded/trimmed to ny x nx.
 */
int utr_add_images(utr_ctx* ctx, int first_row, int count, const float* images);

/* Same, with images already resident in device memory (HBM). */
int utr_add_images_device(utr_ctx* ctx, int first_row, int count, const float* d_images);

/*
 * Runs the gridding onto the Ewald sphere (UTR.h:188-307, called at
 * UTR.cpp:1027), the sampling normalisation and noise filter
 * (UTR.cpp:1050-1172), the 3D CTF / TIE-Hom / inverse-Laplacian filters
 * (UTR.cpp:1177-1201), the Hermitian symmetrisation (UTR.cpp:1203-1217), the
 * inverse 3D FFT (Fftwd3frc, UTR.cpp:1219-1233), the sphere mask and final
 * scaling (UTR.cpp:1235-1279).  Writes nz*ny*nx floats (x fastest, centred as
 * K3out) to host memory.  All rows must have been added.  May be called again
 * after further utr_reset/utr_add_images cycles.
 */
int utr_finalize(utr_ctx* ctx, float* vol);

/* Same, writing to device memory. */
int utr_finalize_device(utr_ctx* ctx, float* d_vol);

/* Same, writing double precision to host memory (UTR_FP64 contexts only);
 * for precision studies against an FP64 CPU computation. */
int utr_finalize_f64(utr_ctx* ctx, double* vol);

/* RELION image shifts (UL, per orientation row, as read from a .relionnew
 * file), undone on each image before its 2D stage: the image content moves
 * by (floor(-shift_x/xst + 0.5), floor(-shift_y/xst + 0.5)) pixels, the
 * edge mean filling what is uncovered.  Call before any image is added.
 * Replaces the XArray2DMove::Move call of UTR.cpp:961-969. */
int utr_set_image_shifts(utr_ctx* ctx, const double* shift_x, const double* shift_y);

/* Astigmatism per orientation row (defocus difference dz in UL and angle in
 * degrees, as read from a .relionnew file into vastigm), used by the 2D CTF
 * inversion of ctf_mode 3 (UTR.cpp:1011-1016; the other modes ignore it, as
 * the reference does): the defocus at azimuth phi of the spatial frequency
 * is zout + dz cos 2(phi - angle).  Restated (InvertCTF_DT2 is in the absent
 * XArrayLibrary).  Call before any image is added. */
int utr_set_astigmatism(utr_ctx* ctx, const double* dz, const double* angle_deg);

/* Keep the sampling matrix Samp3 of the next utr_finalize in device memory
 * (parameter 32, UTR.cpp:706-712; off by default). */
int utr_keep_sampling(utr_ctx* ctx, int on);

/* Copies the sampling matrix Samp3 (nz*ny*nx, the weights the gridding
 * accumulated) to host memory; valid after utr_finalize (FP32 contexts need
 * utr_keep_sampling first).  Replaces the sampling-matrix output of
 * UTR.cpp:1043-1048 (parameter 32). */
int utr_get_sampling(utr_ctx* ctx, float* samp);

/* Discards the images added so far so the context can be reused for another
 * set of images with the same geometry and rows. */
int utr_reset(utr_ctx* ctx);

/* Volume dimensions derived from the parameters (nz from UTR.cpp:321). */
int utr_dims(const utr_ctx* ctx, int* nx, int* ny, int* nz);

/* Device-time phase timers of the last finalize. */
int utr_get_timings(const utr_ctx* ctx, utr_timings* t);

int utr_destroy(utr_ctx* ctx);

/* Message of the last failure on this thread ("" if none). */
const char* utr_last_error(void);

/* Library / ABI version. */
int utr_abi_version(void);

/*
 * Diagnostics (no counterpart in the reference): copy intermediate results
 * to host memory so a CPU restatement can be compared stage by stage.
 * utr_get_spectrum: the centred 2D spectrum K2four of one row (UTR.cpp:995/
 * 1017-1019), ny*nx interleaved complex, FP32 contexts after the row was
 * added.  utr_get_fourier_volume: V3 after normalisation, noise filter and 3D
 * filters (UTR.cpp:1201), nz*ny*nx interleaved complex as float; valid after
 * utr_finalize.
 */
int utr_get_spectrum(utr_ctx* ctx, int row, float* spec);
int utr_get_fourier_volume(utr_ctx* ctx, float* vol_c);

/*
 * Grid-unit coefficients of the gridding coordinate map for one orientation
 * (3x3 row-major, see UTR.h:243-256).  Exposed so a CPU restatement can check
 * that both sides evaluate identical coordinates.
 */
int utr_angle_coeffs(const utr_params* params, double angZ_deg, double angY_deg, double* m9);

/* ------------------------------------------------------- post-processing
 * High-pass filter of a reconstructed volume (UTR.cpp:1357-1378, parameter
 * 23): vol -= Gaussian(vol) with sigma = width / 2.355, the Gaussian applied
 * in Fourier space on the periodic grid (the reference's GaussFilter, whose
 * code is in the absent XArrayLibrary, restated as exp(-2 pi^2 sigma^2 q^2)).
 * vol is nz*ny*nx floats (x fastest) with steps xst, yst, zst and width in
 * the same length unit; on the device (d_vol) or in host memory.  Edges must
 * be FFT lengths the library supports. */
int utr_highpass3d_device(float* d_vol, int nx, int ny, int nz, double xst, double yst, double zst,
                          double width);
int utr_highpass3d(float* vol, int nx, int ny, int nz, double xst, double yst, double zst, double width);

/* ThresholdLow of parameter 23 (UTR.cpp:1383): v = v - background, then
 * values below threshold set to threshold; n floats in place. */
int utr_threshold_device(float* d_vol, long long n, float background, float threshold);
int utr_threshold(float* vol, long long n, float background, float threshold);

/* Reslice and axis flips of parameters 29-30 (UTR.cpp:1449-1453): output
 * slices in XY (mode 0), XZ (1), YZ (2), YX (3), ZX (4) or ZY (5) planes --
 * the first letter along a slice's rows, the second down the slice, slices
 * stacked along the remaining axis -- after reversing the flipped input
 * axes.  out holds nx*ny*nz floats, distinct from in; the output edges are
 * the input edges in the mode's order. */
int utr_reslice_device(const float* d_in, float* d_out, int nx, int ny, int nz, int mode, int flipx, int flipy,
                       int flipz);
int utr_reslice(const float* in, float* out, int nx, int ny, int nz, int mode, int flipx, int flipy, int flipz);

/* Thick slices of parameter 31 (UTR.cpp:1456-1490): nz / step output slices,
 * each the MIP (mode 1) or the average (mode 2) of step consecutive input
 * slices; out holds nx*ny*(nz/step) floats. */
int utr_thick_slices_device(const float* d_in, float* d_out, int nx, int ny, int nz, int mode, int step);
int utr_thick_slices(const float* in, float* out, int nx, int ny, int nz, int mode, int step);

/* 16-bit output of parameter 27 (UTR.cpp:1494-1498; Convert is in the absent
 * XArrayLibrary, restated): values clamped to [minin, maxin] and mapped
 * linearly onto 0 .. maxout, rounded half away from zero. */
int utr_convert16_device(const float* d_in, unsigned short* d_out, long long n, float minin, float maxin,
                         long long maxout);
int utr_convert16(const float* in, unsigned short* out, long long n, float minin, float maxin, long long maxout);

/* Allocation-free form of utr_highpass3d_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and result as utr_highpass3d_device; it calls
 * no hipMalloc/hipFree and creates no stream, using instead a fixed work area
 * that the library holds on every GPU it runs on for the life of the process
 * (16 MiB + 24 KiB: nx*ny*nz at most 2^21 voxels, e.g. 128^3; larger volumes
 * are rejected with UTR_E_INVALID).  Runs on the null stream and returns
 * after the result is in d_vol.  All calls on one GPU share that work area:
 * calls that overlap on one GPU (from several host threads) race. */
int utr_highpass3d_device_noalloc(float* d_vol, int nx, int ny, int nz, double xst, double yst,
                                  double zst, double width);

/* Regularised inverse (-Laplacian) of a real volume, replacing
 * fft3f.InverseMLaplacian(K3out, epsilon, epsilon) of UTR.cpp:1363-1367
 * (parameter 22 with parameter 33): the 3D FFT times
 * norm / (alpha + 4 pi^2 (kx^2/Lx^2 + ky^2/Ly^2 + kz^2/Lz^2)), L = n * step,
 * as InverseMLaplacian3D (UTR.h:135-185) does in Fourier space, then the
 * inverse FFT (Fftwd3frc's member is in the absent library; restated).
 * alpha must be positive.  Edges must be supported FFT lengths. */
int utr_invlaplace3d_device(float* d_vol, int nx, int ny, int nz, double xst, double yst, double zst,
                            double alpha, double norm);
int utr_invlaplace3d(float* vol, int nx, int ny, int nz, double xst, double yst, double zst, double alpha,
                     double norm);

/* 3D rotation of a volume, replacing XArray3DSpln::Rotate3(K3out, Z, Y', Z'')
 * of UTR.cpp:1387-1394 (parameter 28; the reference passes the angles
 * negated on reading, in radians -- here in degrees).  The content turns by
 * R = Rz(angZ2) Ry(angY) Rz(angZ) about the array centre (nx/2, ny/2, nz/2),
 * the rotation pdb.exe applies to atoms (pdb_run.cpp:419-432); values are
 * tricubic (Keys) interpolations, 0 outside.  The spline of the absent
 * library is restated.  _device: d_out must not alias d_in. */
int utr_rotate3d_device(const float* d_in, float* d_out, int nx, int ny, int nz, double angZ_deg,
                        double angY_deg, double angZ2_deg);
int utr_rotate3d(float* vol, int nx, int ny, int nz, double angZ_deg, double angY_deg, double angZ2_deg);

/* Symmetry average of UTR.cpp:1341-1355 (parameter 20 = 2): vol becomes the
 * mean of itself and its rotations turning back symmetry rows 1 .. nsym-1
 * (Rotate3 with -Z'', -Y', -Z of each row of the parameter-21 file, whose
 * columns 1-3 are angZ, angY, angZ2 in degrees). */
int utr_symmetrize3d_device(float* d_vol, int nx, int ny, int nz, int nsym, const double* angZ_deg,
                            const double* angY_deg, const double* angZ2_deg);
int utr_symmetrize3d(float* vol, int nx, int ny, int nz, int nsym, const double* angZ_deg,
                     const double* angY_deg, const double* angZ2_deg);

/* ----------------------------------------------------------- batched form
 * Several independent reconstructions of uneven size run as one HIP graph:
 * each job's launches (2D stage of all its rows, gridding, 3D inverse FFT)
 * form one branch, the branches run concurrently, and the graph is captured
 * once at utr_batch_create and replayed by every utr_batch_run.  No
 * counterpart in the reference, which runs one reconstruction per process
 * (UTR.cpp:main); it is the batched form of the project's target.
 */
typedef struct utr_job {
    utr_params params;
    int nrows;
    const double* angZ_deg;  /* per row, as utr_create                      */
    const double* angY_deg;
    const double* defocus;
    const float* weights;    /* may be NULL                                 */
    const float* images;     /* nrows*ny*nx                                 */
    float* vol;              /* nz*ny*nx result                             */
    int images_on_device;    /* 0: host memory (copied before each run)     */
    int vol_on_device;       /* 0: host memory (copied after each run)      */
    int precision;           /* UTR_FP32 or UTR_FP32_GATHER; ignored when autotuned */
} utr_job;

typedef struct utr_batch utr_batch;

/* Creates the contexts and captures the graph.  Device pointers in jobs must
 * stay valid while the batch exists; host buffers are read / written at each
 * utr_batch_run.  autotune != 0: for every distinct job shape (nx, ny, nz,
 * nrows) the launch configuration is tuned once on that job's images, each
 * candidate timed as a whole reconstruction (best of two runs), by
 * coordinate descent: y planes per plane-gridding workgroup (1, 2, 4),
 * resident workgroups per CU of the persistent row FFT (4, 8) and of the
 * column FFT (2, 4), then the gather form against the scatter; the choice is
 * cached for the process and captured.  Without autotune each job uses the
 * defaults (planes by size, row FFT 8 per CU, column FFT as LDS allows). */
int utr_batch_create(const utr_job* jobs, int njobs, int autotune, utr_batch** out);

/* Runs all jobs: host images in, one graph launch, host volumes out.
 * graph_ms (may be NULL): device time of the graph launch. */
int utr_batch_run(utr_batch* b, double* graph_ms);

/* Replays the graph `repeats` times back to back with no host work between
 * the launches (every job's images and volume on the device), then waits.
 * ms (may be NULL): device time of all repeats.  No counterpart in the
 * reference: the launch-free form of a job stream for callers that keep
 * their inputs resident. */
int utr_batch_launch(utr_batch* b, int repeats, double* ms);

/* Form chosen for job j (UTR_FP32 or UTR_FP32_GATHER). */
int utr_batch_job_form(const utr_batch* b, int j);

/* The launch configuration captured for job j: form, planes per gridding
 * workgroup and the FFT workgroups per CU (0: the default).  No counterpart
 * in the reference (its jobs are separate runs of UTR.exe). */
int utr_batch_job_tuning(const utr_batch* b, int j, int* form, int* planes, int* rows_wpc, int* cols_wpc);

int utr_batch_destroy(utr_batch* b);

/* ------------------------------------------------------------------ DHT
 * The single-defocus reconstructions of the DHT program
 * (MultisliceCppTest/PhaseRetrieval/DHT.cpp): method 3 (phase estimate
 * -Fact*0.5*(I-1), Fresnel back-propagation to every output plane,
 * DHT.cpp:919-1006), method 4 (variable-distance weak-object CTF inversion,
 * DHT.cpp:1010-1055), both followed by the real-space rotate-and-splat of
 * DHT.cpp:1081-1164 and the scale of DHT.cpp:1325-1340; method 5 (sCTF on the
 * Ewald sphere with the Filt3 filter, DHT.cpp:1056-1067, 1195-1416).
 * Methods 1 and 2 (IWFR and CTFL2, DHT.cpp:889-917) retrieve the complex
 * amplitude of each orientation from its defocus series and then follow
 * method 3's back-propagation: create them with dht_create_series.
 */
typedef struct dht_params {
    int nx, ny;              /* 6: image width/height (powers of 2; even for method 5) */
    double xst;              /* 7: pixel size (Angstrom)                     */
    double zlo, zhi;         /* 19: output defocus range                     */
    double dzextra;          /* 20: extra defocus (method 3 only, non-zero)  */
    double dxc, dyc, dzc;    /* 8: centre-of-rotation shifts                 */
    double wl;               /* 10: wavelength                               */
    double aobj_mrad;        /* 11: objective aperture                       */
    double cs3_mm, cs5_mm;   /* 12: spherical aberrations                    */
    double asigma;           /* 13: beta/delta (methods 4, 5)                */
    int method;              /* 14: 1 to 5                                   */
    int escc;                /* 15: Ewald sphere curvature (per-plane back-propagation) */
    double critical_snr;     /* 17: method 5                                 */
    double epsilon;          /* 18: IWFR stopping error (method 1), Fact
                                (method 3) or Tikhonov parameter (2, 4, 5)   */
    int inv_laplace;         /* 23: inverse 3D Laplacian (method 5)          */
    double alpha;            /* 24: its regularisation parameter             */
    double norm_f1, norm_f2; /* 9: input -> input/f1 + f2                    */
    int itermax;             /* 17: method 1, maximal number of iterations   */
} dht_params;

typedef struct dht_ctx dht_ctx;

/* Replaces the set-up of DHT.cpp:36-720 (validation, geometry) and the 3D
 * array allocation of DHT.cpp:716-740.  Rows as for utr_create (one defocus
 * distance per orientation; Z'' angles through dht_set_image_correction);
 * weights may be NULL. */
int dht_create(const dht_params* params, int nrows, const double* angZ_deg,
               const double* angY_deg, const double* defocus, const float* weights,
               dht_ctx** out);

/* Methods 1 and 2: nangles orientations with counts[a] >= 2 images each
 * (the rows of the defocus/orientations file with several defocus entries,
 * DHT.cpp:142-152, 272-283).  defocus holds one distance per image, in
 * orientation order (sum of counts entries); after this call the row index
 * of dht_set_image_correction and dht_add_images[_device] is the image
 * index in that order, so each image has its own Z'' angle (repeat an
 * orientation's RELION shift for its images).  IWFR retrieves the wave at
 * the mean distance of the series (DHT.cpp:895-897), CTFL2 at its first
 * distance (DHT.cpp:908).  Methods 3-5 with every count 1 behave as
 * dht_create. */
int dht_create_series(const dht_params* params, int nangles, const double* angZ_deg,
                      const double* angY_deg, const int* counts, const double* defocus,
                      const float* weights, dht_ctx** out);

/* Method 1: the IWFR iterations each orientation took in the last
 * dht_finalize (nangles ints; 0 for method 2).  No counterpart in the
 * reference, which prints them when verbose. */
int dht_get_iterations(const dht_ctx* ctx, int* iters);

/* Methods 1 and 2, after dht_finalize: the retrieved complex amplitudes of
 * orientations [first, first+count) at their zout, count*ny*nx (re, im)
 * float pairs -- the data of the GRC files DHT.cpp:902, 916 write when
 * parameter 16 is 1. */
int dht_get_waves(dht_ctx* ctx, int first, int count, float* psi);

/* Z'' angles (degrees) and RELION image shifts (Angstrom) per row, undone on
 * each renormalised input image before its 2D stage (the Move / Rotate block
 * of DHT.cpp:862-880, restated as for utr_create's angZ2 and
 * utr_set_image_shifts).  Any pointer may be NULL (zeros).  Call before any
 * image is added. */
int dht_set_image_correction(dht_ctx* ctx, const double* angZ2_deg, const double* shift_x,
                             const double* shift_y);

/* The per-orientation phase retrieval input (DHT.cpp:787-947): images of rows
 * [first_row, first_row+count), count*ny*nx floats in host / device memory. */
int dht_add_images(dht_ctx* ctx, int first_row, int count, const float* images);
int dht_add_images_device(dht_ctx* ctx, int first_row, int count, const float* d_images);

/* Back-propagation or CTF correction per plane, splat and scale
 * (DHT.cpp:957-1416): nz*ny*nx floats to host / device memory. */
int dht_finalize(dht_ctx* ctx, float* vol);
int dht_finalize_device(dht_ctx* ctx, float* d_vol);

int dht_reset(dht_ctx* ctx);
int dht_dims(const dht_ctx* ctx, int* nx, int* ny, int* nz);
int dht_get_timings(const dht_ctx* ctx, utr_timings* t);
int dht_destroy(dht_ctx* ctx);

/* ---------------------------------------------------------------------------
 * TEM multislice forward model: the image simulator of TemsimLibrary
 * (autosliccmd(), autosliccmd.cpp:177-1300, with autoslic::calculate,
 * autoslic.cpp:209-869), coherent branch: plane incident wave, no ice, no
 * crystal tilt, Z'' = 0 and no XY shifts of the images.  Thermal vibration
 * follows msct_params.nwobble: 0 one static configuration, 1 the Debye-Waller
 * smoothing of every slice's potential (trlayerDW, autoslic.cpp:1040-1104;
 * its 2D variant for output 3 is not built), > 1 the average over nwobble
 * displaced configurations (autoslic.cpp:700-711; the displacements come
 * from a seeded generator).  One context holds one atom set; each call
 * simulates any number of orientations and defocus distances.
 * ------------------------------------------------------------------------- */
typedef struct msct_params {
    int nx, ny;             /* pixels (autosliccmd line 12 "nx ny")                 */
    int mode;               /* 0 multislice, 1 projection, 2 first Born (the slice
                               loop of both skips propagation, autoslic.cpp:772)   */
    int output;             /* autosliccmd noutput: 0 intensity |psi|^2, 1 phase,
                               2 complex psi (re, im pairs), 3 the 3D potential on
                               the plane z = ctblength + defocus (no multislice)    */
    double ctblength;       /* cube side in Angstrom (the largest cell edge of the
                               XYZ file, autosliccmd.cpp:629-633)                   */
    double deltaz;          /* slice thickness, Angstrom (line 14)                 */
    double kev;             /* beam energy, keV (line 11)                          */
    double aobj_mrad;       /* objective aperture, mrad; 0 = none (line 8)         */
    double cs3_mm, cs5_mm;  /* spherical aberrations, mm (line 6)                  */
    double astig_dz;        /* two-fold astigmatism amplitude, Angstrom            */
    double astig_angle_deg; /* and its angle, degrees (autosliccmd's astigm pair)  */
    double dxc, dyc, dzc;   /* rotation centre minus the cube centre, A (line 13)  */
    int nwobble;            /* thermal vibrations (MsctKirkland lines 12-14): 0 none,
                               1 Debye-Waller smoothing of every slice's potential
                               (trlayerDW), > 1 configurations averaged            */
    int reserved;
    double temperature;     /* K (line 13): displacements scale by sqrt(T / 300)   */
    unsigned long long seed; /* of the configurations' displacements             */
} msct_params;

typedef struct msct_ctx msct_ctx;

/* One atom set (ReadXYZcoord's Z, x, y, z, occ and the rms thermal
 * displacement at 300 K, wobble, which may be NULL; natom may be 0) and the
 * electron scattering-factor parameters fe_params[(Z-1)*12 + i] of Kirkland
 * (2010) App. C for Z = 1..103 (slicelib.cpp:958-2209, shipped as
 * utr_public_amd/data/fe_params.txt).  Replaces the set-up half of
 * autosliccmd() (autosliccmd.cpp:600-690, 870-910). */
int msct_create(const msct_params* p, const double* fe_params, int natom, const int* Z, const float* x,
                const float* y, const float* z, const float* occ, const float* wobble, msct_ctx** out);

/* Images of nang orientations (angZ, angY in radians, the autosliccmd line
 * 25 pair: rotation about Z, then about Y', autosliccmd.cpp:733-748) at ndef
 * defocus distances (Angstrom, autosliccmd.cpp:1065): out holds
 * nang*ndef*ny*nx floats (x fastest), twice that for output 2,
 * orientation-major.  An atom leaving the cube after rotation is an error,
 * as in autosliccmd.cpp:857-860.  Host / device output. */
int msct_images(msct_ctx* ctx, int nang, const double* angZ_rad, const double* angY_rad, int ndef,
                const double* defocus, float* out);
int msct_images_device(msct_ctx* ctx, int nang, const double* angZ_rad, const double* angY_rad, int ndef,
                       const double* defocus, float* d_out);

/* Slices per orientation: k * deltaz for k = 1, 2, ... while below the cube
 * side + deltaz (autoslic.cpp:727-744). */
int msct_slices(const msct_ctx* ctx);
/* Orientations simulated together per batch: as many as fit the
 * UTR_MSCT_BATCH_BYTES budget (default 6 GiB) of device memory held per
 * orientation (three complex planes, the rotated atom records and the slice
 * table).  No counterpart in the reference (it runs one orientation at a
 * time per thread, MultisliceT.cpp:398-450). */
int msct_batch(const msct_ctx* ctx);
int msct_destroy(msct_ctx* ctx);

/* ---------------------------------------------------------------------------
 * Evaluation against the forward model.
 * ------------------------------------------------------------------------- */

/* Ground-truth potential of an atom set on an nz*ny*nx grid (device memory,
 * x fastest): voxel (i, j, k) at cube position (i, j, k) * xst, atoms (Z, x,
 * y, z in Angstrom, cube coordinates) on the periodic cube of sides n * xst.
 * It is the potential whose point samples autosliccmd's noutput 3 sums
 * (vatom, slicelib.cpp:2699-2737; autosliccmd.cpp:1146-1151), restricted to
 * spatial frequencies |q| < qmax (1/A; qmax <= 0: all up to Nyquist) and
 * evaluated exactly by Fourier synthesis from its transform
 * (al/pi) sum a_i/(q^2+b_i) + ag pi^-1.5 sum c_i exp(-d_i q^2), times scale.
 * fe_params as msct_create.  Edges must be supported FFT lengths. */
int utr_ground_truth_device(int nx, int ny, int nz, double xst, double qmax, double scale, const double* fe_params,
                            int natom, const int* Z, const float* x, const float* y, const float* z, float* d_vol);

/* Multislice reprojection of a reconstructed volume (device memory, UTR
 * units, nz*ny*nx with nx = ny = nz), replacing the imodeReproj block of
 * DHT.cpp:1509-1731: the volume times pi/(lambda E) zst (DHT.cpp:1514) is
 * turned to each row's orientation (about Z, then Y', as the forward model
 * turns atoms; trilinear sampling, the absent Multislice_eV restated),
 * propagated plane by plane with the multislice kernels (bandwidth limit,
 * propagator of thickness xst), then to the row's defocus distance from the
 * cube centre with the objective aperture and Cs of p.  d_out (may be NULL)
 * receives the nrows*ny*nx intensities; err (may be NULL, then d_images may
 * be NULL) the error of DHT.cpp:1686 per row,
 * || |psi| - sqrt(I0) ||^2 / || sqrt(I0) ||^2, against d_images renormalised
 * with p's f1, f2 and clamped at 0 (DHT.cpp:1641-1647).  TEM only. */
int utr_reproject_device(const utr_params* p, const float* d_vol, int nrows, const double* angZ_deg,
                         const double* angY_deg, const double* defocus, const float* d_images, float* d_out,
                         double* err);

/* ---------------------------------------------------------------------------
 * Learned residual corrector.  No counterpart in the reference (which has no
 * learned component): a 3-level 3D U-Net (3x3x3 convolutions with 16/32/64
 * channels, ReLU, 2x average pooling, nearest upsampling, skip connections by
 * channel concatenation; head: alpha * input + a 1x1x1 convolution of the last
 * 16 channels) mapping a UTR volume to a corrected one.  BF16 weights and
 * activations on the matrix cores with FP32 accumulation; trained on the GPU
 * by the library's own forward, backward and Adam kernels.  Values enter the
 * network multiplied by `scale` (utr_corr_set_params) and leave divided by it.
 * Crops are cubes of T = tile + 2 halo voxels (FP32, x fastest); the loss is
 * the mean squared error over the central tile^3 of each crop.
 * ------------------------------------------------------------------------- */
typedef struct utr_corr utr_corr;

/* Network with seeded He-initialised weights (the head starts as the
 * identity: alpha = 1), for crops of tile + 2 halo voxels (a multiple of 4,
 * at most 512) and up to `batch` crops per call. */
int utr_corr_create(int tile, int halo, int batch, unsigned long long seed, utr_corr** out);
/* Number of FP32 parameters: per layer W[27][cin][cout] then b[cout] for the
 * ten convolutions (cin, cout) = (16,16) (16,16) (16,32) (32,32) (32,64)
 * (64,64) (96,32) (32,32) (48,16) (16,16) -- the first layer's input is the
 * volume in channel 0 of 16 -- then the head w1[16], b1, alpha. */
int utr_corr_nparams(const utr_corr* c);
int utr_corr_tile(const utr_corr* c, int* tile, int* halo, int* batch);
/* Copies the parameters (nparams floats, may be NULL) and the input scale. */
int utr_corr_get_params(utr_corr* c, float* params, float* scale);
/* Sets the input scale and (params may be NULL) the parameters; resets the
 * Adam moments. */
int utr_corr_set_params(utr_corr* c, const float* params, float scale);
/* Network output for nb <= batch crops in device memory (nb * T^3 floats). */
int utr_corr_forward(utr_corr* c, const float* d_crops, int nb, float* d_out);
/* Loss and gradient (nparams floats to host memory, may be NULL) of nb crops
 * against targets of the same shape; no update. */
int utr_corr_gradient(utr_corr* c, const float* d_crops, const float* d_targets, int nb, float* loss,
                      float* grad);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) with learning rate lr;
 * loss (may be NULL) is the batch loss before the step.  Like every
 * utr_corr_* call, it returns after the device work is done. */
int utr_corr_train_step(utr_corr* c, const float* d_crops, const float* d_targets, int nb, float lr, float* loss);
/* Corrected volume (nz*ny*nx floats in device memory) from tiles of the
 * input volume with their halos (zeros outside the volume). */
int utr_corr_apply(utr_corr* c, const float* d_vol, float* d_out, int nx, int ny, int nz);
int utr_corr_destroy(utr_corr* c);

#ifdef __cplusplus
}
#endif
#endif /* UTR_HIP_H */
