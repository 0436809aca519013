/*
 * utr_oracle.h -- TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.
 *
 * FP64 CPU restatement of the UTR reconstruction path of
 * MultisliceCppTest/PhaseRetrieval/UTR.cpp:840-1279 and UTR.h:15-307
 * (timg021/UTR_public).  It is the checker the HIP library is compared
 * against.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py load it; the product library never links or calls it.
 *
 * Parity status per function (what pins it; see the tests named):
 *   ora_fft            pinned: brute-force DFT, Parseval, numpy.fft
 *                      (tests/test_oracle.py::test_fft_*)
 *   ora_spectrum       pinned for ctf_mode 0/2/4 (centred DFT of ln I against
 *                      a brute-force centred DFT and a closed-form Gaussian);
 *                      modes 1 and 3 restate the absent XArrayLibrary 2D
 *                      filters (InverseMLaplacian1, InvertCTF_DT2) from the
 *                      3D analogues in UTR.h; pinned by physics and by the 3D
 *                      filters (tests/test_filter_physics.py: Paganin's
 *                      homogeneous-object thickness; the multislice phase of
 *                      a thin specimen; equality on the q_z = 0 plane).
 *   ora_grid           pinned: brute-force direct splat, weight conservation,
 *                      axis-aligned exact mapping, linearity.
 *   ora_normalize_filter, ora_hermitian, ora_mask_scale
 *                      pinned: element-wise restatement checked against
 *                      independent numpy formulas written from UTR.cpp /
 *                      UTR.h line by line.
 *   ora_c2r            pinned: numpy.fft.irfftn on the same half spectrum.
 *   ora_reconstruct    pinned end to end by the blob round trip
 *                      (tests/test_oracle.py::test_blob_position).
 */
#ifndef UTR_ORACLE_H
#define UTR_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Raw numeric parameters of the UTR parameter file (ReadmeUTR.txt:33-165),
 * in the units the file uses. */
typedef struct ora_params {
    int nx, ny;          /* param 8: image width/height (even)             */
    int modality;        /* param 1: 0 = TEM (Angstrom), 1 = X-ray (micron) */
    double xst;          /* param 9: pixel size (UL)                        */
    double zlo, zhi;     /* param 10: output defocus range (UL)             */
    double dzextra;      /* param 10: defocus offset (UL)                   */
    double dxc, dyc, dzc;/* param 11: centre-of-rotation shifts (UL)        */
    double wl;           /* param 12: wavelength (UL)                       */
    double aobj_mrad;    /* param 13: objective aperture (mrad)             */
    double cs3_mm, cs5_mm; /* param 14: Cs3, Cs5 (mm)                       */
    int ctf_mode;        /* param 15: 0..4                                  */
    int escc;            /* param 16: Ewald sphere curvature correction     */
    double asigma;       /* param 17: beta/delta (already -1/(delta/beta) for .xri) */
    double nsr;          /* param 18: average NSR of the input images       */
    double epsilon;      /* param 19: Tikhonov regularisation               */
    int inv_laplace;     /* param 22: inverse 3D Laplacian                  */
    double norm_f1, norm_f2; /* param 7: input normalisation I/f1 + f2      */
} ora_params;

/* Geometry and constants derived from the parameters and the orientation
 * rows exactly as UTR.cpp:284-388 derives them. */
typedef struct ora_geom {
    int nx, ny, nz, nxd2, nyd2, nzd2;
    int modality, ctf_mode, escc, inv_laplace;
    double xst, yst, zst;
    double xlo, xhi, ylo, yhi, zlo, zhi;
    double fxst, fyst, fzst, fxlo, fylo, fzlo;
    double dxc, dyc, dzc;
    int rot_centre_shift;
    double xc, yc, zc;
    long ic, jc, kc;
    double wl, EE, k2maxo, cs3, cs5;
    double asigma, nsr, epsilon;
    double norm_f1, norm_f2;
    double zout_aver;     /* mean defocus minus dzextra (UTR.cpp:232-238,330) */
    double dzextra;
    int rot_y_only, rot_z_only, rot_z2_only;  /* UTR.cpp:211-222 */
    long nR2;             /* sphere-mask radius^2 (UTR.cpp:1236-1242) */
    double mask_dtemp;    /* 2/(kc^2+jc^2+ic^2) (UTR.cpp:1243) */
    double fnorm;         /* final scale (UTR.cpp:1273-1277) */
} ora_geom;

const char* ora_last_error(void);

/* Validates parameters (the numeric-path subset of UTR.cpp:55-784) and fills
 * *g.  angZ2_deg may be NULL (all zero).  Returns 0 or -1. */
int ora_derive(const ora_params* p, int nrows, const double* angZ_deg,
               const double* angY_deg, const double* angZ2_deg,
               const double* defocus, ora_geom* g);

/* 1D in-place complex FFT of interleaved (re,im) data, unnormalised,
 * exponent sign*2*pi*i*k*m/n.  Any n >= 1. */
int ora_fft(double* data, int n, int sign);

/* Grid-unit forward-map coefficients of one orientation, row-major 3x3:
 * [d(dx1)/da d(dx1)/db d(dx1)/dzz ; d(dy1)/.. ; d(dz1)/..] where a=i-nx/2,
 * b=j-ny/2 and zz is the Ewald height in physical units (UTR.h:243-256). */
void ora_angle_coeffs(const ora_geom* g, double angZ_deg, double angY_deg,
                      double* m9);

/* Per-row 2D stage (UTR.cpp:951-1019): renormalise, clamp, ln, centred 2D
 * DFT, optional 2D filter (modes 1/3), subseries weight.  K is ny*nx
 * interleaved complex, centred (zero frequency at (ny/2, nx/2)). */
/* Per-row astigmatism (defocus difference dz in UL, angle in degrees) used
 * by the mode-3 filter of the following ora_reconstruct calls; n = 0 or NULL
 * arrays clear it.  The arrays must stay valid while it is set. */
void ora_set_astigmatism(int n, const double* dz, const double* angle_deg);

/* ora_spectrum with astigmatism (dz UL, angle degrees) in the mode-3 filter. */
int ora_spectrum_astig(const ora_geom* g, const float* img, double zout, double weight,
                       double adz, double aang_deg, double* K);

int ora_spectrum(const ora_geom* g, const float* img, double zout,
                 double weight, double* K);

/* Ewald-sphere gridding of one spectrum (UTR.h:188-307), accumulating into
 * V (nz*ny*nx interleaved complex) and S (nz*ny*nx). */
int ora_grid(const ora_geom* g, const double* K, double angZ_deg,
             double angY_deg, double* V, double* S);

/* Sampling normalisation + noise filter (UTR.cpp:1050-1172) followed by the
 * 3D TIE-Hom / CTF filters and the inverse Laplacian (UTR.cpp:1177-1201). */
int ora_normalize_filter(const ora_geom* g, double* V, const double* S);

/* Hermitian symmetrisation in serial loop order (UTR.cpp:1203-1217). */
int ora_hermitian(const ora_geom* g, double* V);

/* 3D complex-to-real inverse DFT with the half consumed by a c2r plan after
 * un-centring, normalised 1/(nx*ny*nz), output centred (UTR.cpp:1219-1233). */
int ora_c2r(const ora_geom* g, const double* V, double* out);

/* Sphere mask, reverse-cupping compensation and final scaling
 * (UTR.cpp:1235-1279). */
int ora_mask_scale(const ora_geom* g, double* out);

/* Whole reconstruction: rows are processed in input order (nthreads == 1)
 * or angle-parallel with private accumulators summed in thread order
 * (nthreads > 1, used only for CPU timing).  images: nrows*ny*nx floats.
 * weights may be NULL.  vol: nz*ny*nx doubles.  samp may be NULL. */
int ora_reconstruct(const ora_params* p, int nrows, const double* angZ_deg,
                    const double* angY_deg, const double* angZ2_deg,
                    const double* defocus, const float* weights,
                    const float* images, double* vol, double* samp,
                    int nthreads);

/* Timings of the last ora_reconstruct call in milliseconds, named after
 * UTR.cpp:1554-1559: [0] 2D FFT+CTF, [1] backpropagation, [2] 3D FFT,
 * [3] total. */
void ora_last_timings(double* t4);

/* ------------------------------------------------------------------ DHT
 * FP64 restatement of the single-defocus reconstructions of the DHT program
 * (MultisliceCppTest/PhaseRetrieval/DHT.cpp:700-1416), in oracle/dht_oracle.c.
 * Test infrastructure like the rest of this header.  Parity status:
 *   ora_dht_splat      in-tree source restated line by line (DHT.cpp:1081-1164,
 *                      including the (int) truncation); pinned by a brute-force
 *                      Python transcription and weight conservation
 *                      (tests/test_dht_oracle.py).
 *   ora_dht_planes     method 3: the phase estimate of ReadmePhaseRetrieval.txt
 *                      (Phase = Fact * -0.5 * (I/I0 - 1)) with the Fresnel
 *                      propagator of the decision register (absent
 *                      XArray2DFFT::Fresnel) -- propagator PARITY UNPINNED,
 *                      pinned against closed forms (free-space unitarity,
 *                      propagation of a Gaussian, zero-distance identity).
 *                      method 4: weak-object CTF inversion (absent
 *                      XA_IWFR::InvertPhaseCTF), pinned by the forward
 *                      weak-object model it inverts and by the multislice
 *                      phase of a thin specimen it recovers in the plane
 *                      through the specimen (tests/test_filter_physics.py).
 *   ora_dht_retrieve   methods 1 and 2: IWFR and the multi-distance CTF
 *                      least squares, restated from their publications (the
 *                      absent XA_IWFR::Iwfr and ::CTFL2) -- PARITY UNPINNED,
 *                      pinned by the physics: both recover the exit wave of
 *                      a specimen simulated by the pinned multislice at
 *                      several defocus distances (tests/test_series_oracle.py).
 *   ora_dht_reconstruct method 5 reuses ora_spectrum (the mode 3 restatement,
 *                      pinned as above, for the absent InvertCTF_DT1) and ora_grid,
 *                      with the Filt3 filter, Hermitian fix, c2r and scale of
 *                      DHT.cpp:1341-1416 restated line by line.
 */
typedef struct ora_dht_params {
    int nx, ny;            /* param 6 */
    double xst;            /* param 7 (Angstrom) */
    double zlo, zhi;       /* param 19 */
    double dzextra;        /* param 20 */
    double dxc, dyc, dzc;  /* param 8 */
    double wl;             /* param 10 */
    double aobj_mrad;      /* param 11 */
    double cs3_mm, cs5_mm; /* param 12 */
    double asigma;         /* param 13 */
    int method;            /* param 14: 1 to 5 */
    int escc;              /* param 15 */
    double critical_snr;   /* param 17 (method 5) */
    double epsilon;        /* param 18 */
    int inv_laplace;       /* param 23 (method 5) */
    double alpha;          /* param 24 */
    double norm_f1, norm_f2; /* param 9 */
    int itermax;           /* param 17 (method 1) */
} ora_dht_params;

/* Output plane stack of one orientation (DHT.cpp:957-1055): method 3 or 4,
 * nplanes = nz with escc, 1 (the central plane, copied by the splat) without.
 * planes: nplanes*ny*nx doubles, already multiplied by `weight`. */
int ora_dht_planes(const ora_dht_params* p, const float* img, double zout, double weight,
                   double* planes, int* nplanes);

/* Real-space rotate-and-splat of a plane stack into K3out (nz*ny*nx),
 * DHT.cpp:1081-1164. */
int ora_dht_splat(const ora_dht_params* p, const double* planes, int nplanes,
                  double angZ_deg, double angY_deg, double* K3out);

/* Whole reconstruction (methods 3, 4, 5) including the final normalisation
 * (DHT.cpp:1325-1416).  vol: nz*ny*nx doubles.  nthreads > 1: orientation-
 * parallel with private accumulators (timing only). */
int ora_dht_reconstruct(const ora_dht_params* p, int nrows, const double* angZ_deg,
                        const double* angY_deg, const double* defocus, const float* weights,
                        const float* images, double* vol, int nthreads);

/* Methods 1 and 2 (DHT.cpp:889-917): the complex amplitude psi (2*ny*nx
 * doubles, re/im) at zout retrieved from the cnt images of one orientation
 * at distances defocus[0..cnt): IWFR at zout = the mean distance (iters: the
 * iterations done), CTFL2 at zout = defocus[0]. */
int ora_dht_retrieve(const ora_dht_params* p, int cnt, const float* imgs, const double* defocus,
                     double* psi, double* zout, int* iters);

/* Methods 1-3: output planes of one orientation from a complex amplitude
 * psi at zout (the back-propagation of DHT.cpp:957-1006). */
int ora_dht_planes_wave(const ora_dht_params* p, const double* psi, double zout, double weight,
                        double* planes, int* nplanes);

/* Whole reconstruction with counts[a] images per orientation (NULL: one);
 * defocus and images run over all images in orientation order. */
int ora_dht_reconstruct_series(const ora_dht_params* p, int nangles, const double* angZ_deg,
                               const double* angY_deg, const int* counts, const double* defocus,
                               const float* weights, const float* images, double* vol, int nthreads);

/* Volume depth of a DHT reconstruction (DHT.cpp:379). */
int ora_dht_nz(const ora_dht_params* p);

#ifdef __cplusplus
}
#endif
#endif
