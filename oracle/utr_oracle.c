/*
 * utr_oracle.c -- TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.
 *
 * FP64 CPU restatement of the UTR reconstruction path (timg021/UTR_public,
 * MultisliceCppTest/PhaseRetrieval/UTR.cpp and UTR.h).  Every stage cites the
 * reference lines it restates.  The reference itself cannot be compiled here
 * (its XArrayLibrary dependency is absent), so conventions that library hid
 * are defined as follows and pinned by closed-form tests:
 *   - forward 2D DFT: exponent -2*pi*i, unnormalised, centred on input and
 *     output (Shuffle/FFT/Shuffle, UTR.cpp:992-995);
 *   - 3D inverse: exponent +2*pi*i, 1/(nx*ny*nz), c2r semantics on the half
 *     x-frequency range {nx/2..nx-1, 0} in centred indices (UTR.cpp:1207-1232);
 *   - Log0(x) = ln x for x > 0, else 0.
 * Coordinates of the gridding are evaluated in grid units relative to the
 * grid centre with explicit fused multiply-adds; this is algebraically the
 * expression of UTR.h:243-256 (with fxlo = -fxst*nx/2, UTR.cpp:306) and makes
 * axis-aligned orientations map grid points exactly onto grid points.
 */
#include "utr_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORA_PI 3.14159265358979323846
#define ORA_PI180 (ORA_PI / 180.0)
#define ORA_TARGET_SNR 1.414213562373 /* UTR.h:12 */

static __thread char g_err[512];
static double g_timings[4];

const char* ora_last_error(void) { return g_err; }

static int fail(const char* msg) {
    snprintf(g_err, sizeof g_err, "%s", msg);
    return -1;
}

void ora_last_timings(double* t4) { memcpy(t4, g_timings, sizeof g_timings); }

int ora_set_error(const char* msg) { return fail(msg); }

static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

/* ------------------------------------------------------------------ FFT */

typedef struct {
    int n, nf;
    int fac[64];
    double* tw; /* tw[2t], tw[2t+1] = cos, sin of 2*pi*t/n */
} fft_plan;

static void plan_init(fft_plan* pl, int n) {
    pl->n = n;
    pl->nf = 0;
    int m = n;
    static const int pref[4] = {4, 2, 3, 5};
    for (int q = 0; q < 4; q++)
        while (m % pref[q] == 0 && m > 1) { pl->fac[pl->nf++] = pref[q]; m /= pref[q]; }
    for (int p = 7; m > 1; p += 2)
        while (m % p == 0) { pl->fac[pl->nf++] = p; m /= p; }
    pl->tw = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    for (int t = 0; t < n; t++) {
        double a = 2.0 * ORA_PI * (double)t / (double)n;
        pl->tw[2 * t] = cos(a);
        pl->tw[2 * t + 1] = sin(a);
    }
}

static void plan_free(fft_plan* pl) { free(pl->tw); }

/* Stockham autosort: a radix-R pass with p = product of the earlier radices
 * maps x[i + r*n/R] (twiddled by w^(r*(i mod p)), w = e^{s 2 pi i/(pR)}) to
 * y[(i/p)*p*R + (i mod p) + q*p]. */
static void fft_exec(const fft_plan* pl, double* x, double* y, int sign) {
    const int n = pl->n;
    int p = 1;
    double u[2 * 64], v[2 * 64];
    double* src = x;
    double* dst = y;
    for (int s = 0; s < pl->nf; s++) {
        const int R = pl->fac[s];
        const int m = n / R;
        const int step = n / (p * R); /* w^(t) = tw[t*step] */
        for (int i = 0; i < m; i++) {
            const int k = i % p;
            for (int r = 0; r < R; r++) {
                double a = src[2 * (i + r * m)], b = src[2 * (i + r * m) + 1];
                int t = (int)(((long)r * k * step) % n);
                double c = pl->tw[2 * t], sn = sign * pl->tw[2 * t + 1];
                u[2 * r] = a * c - b * sn;
                u[2 * r + 1] = a * sn + b * c;
            }
            for (int q = 0; q < R; q++) {
                double sr = 0, si = 0;
                for (int r = 0; r < R; r++) {
                    int t = (int)(((long)r * q * (n / R)) % n);
                    double c = pl->tw[2 * t], sn = sign * pl->tw[2 * t + 1];
                    sr += u[2 * r] * c - u[2 * r + 1] * sn;
                    si += u[2 * r] * sn + u[2 * r + 1] * c;
                }
                v[2 * q] = sr;
                v[2 * q + 1] = si;
            }
            const int j = (i / p) * p * R + k;
            for (int q = 0; q < R; q++) {
                dst[2 * (j + q * p)] = v[2 * q];
                dst[2 * (j + q * p) + 1] = v[2 * q + 1];
            }
        }
        double* t = src; src = dst; dst = t;
        p *= R;
    }
    if (src != x) memcpy(x, src, sizeof(double) * 2 * (size_t)n);
}

int ora_fft(double* data, int n, int sign) {
    if (n < 1) return fail("ora_fft: n must be positive");
    fft_plan pl;
    plan_init(&pl, n);
    for (int s = 0; s < pl.nf; s++)
        if (pl.fac[s] > 64) { plan_free(&pl); return fail("ora_fft: prime factor > 64"); }
    double* y = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    fft_exec(&pl, data, y, sign);
    free(y);
    plan_free(&pl);
    return 0;
}

/* Strided batch of 1D transforms along one axis of a contiguous array. */
static void fft_axis(double* a, int n, long stride, long count_outer, long inner,
                     int sign, const fft_plan* pl) {
    /* element (o, t, in) at ((o*n + t)*inner + in) with stride == inner */
    (void)stride;
    double* buf = (double*)malloc(sizeof(double) * 4 * (size_t)n);
    double* tmp = buf + 2 * n;
    for (long o = 0; o < count_outer; o++)
        for (long in = 0; in < inner; in++) {
            double* base = a + 2 * (o * n * inner + in);
            for (int t = 0; t < n; t++) {
                buf[2 * t] = base[2 * t * inner];
                buf[2 * t + 1] = base[2 * t * inner + 1];
            }
            fft_exec(pl, buf, tmp, sign);
            for (int t = 0; t < n; t++) {
                base[2 * t * inner] = buf[2 * t];
                base[2 * t * inner + 1] = buf[2 * t + 1];
            }
        }
    free(buf);
}

/* ------------------------------------------------------------ geometry */

int ora_derive(const ora_params* p, int nrows, const double* angZ_deg,
               const double* angY_deg, const double* angZ2_deg,
               const double* defocus, ora_geom* g) {
    memset(g, 0, sizeof *g);
    if (nrows < 1) return fail("no orientation rows");
    if (p->modality != 0 && p->modality != 1)
        return fail("Error: modality and units of length parameter in input parameter file must be 0 for TEM/Angstroms or 1 for X-ray CT/microns.");
    if (p->norm_f1 == 0)
        return fail("The first normalization factor cannot be zero (leads to division by zero).");
    if (p->nx <= 0 || p->ny <= 0 || p->nx % 2 || p->ny % 2)
        return fail("Error: width and height parameters in input parameter file must be even.");
    if (!(p->xst > 0)) return fail("Error: pixel size must be positive in input parameter file.");
    if (!(p->wl > 0) || p->wl > 1) return fail("Error: wavelength value appears to be wrong.");
    if (p->aobj_mrad < 0 || p->aobj_mrad > 1000)
        return fail("Error: objective aperture value appears to be wrong.");
    if (p->ctf_mode < 0 || p->ctf_mode > 4)
        return fail("Error: unknown value for phase retrieval parameter in input parameter file.");
    if (p->escc != 0 && p->escc != 1)
        return fail("Error: unknown value for Ewald sphere curvature correction switch in input parameter file.");
    if (p->modality == 0 && p->asigma < 0) return fail("Error: beta/delta must be non-negative in TEM mode.");
    if (p->modality == 1 && p->asigma >= 0)
        return fail("Error: beta/delta must be negative in hard X-ray mode (unless .xri parameter file is used).");
    if (p->asigma >= 0 && (p->ctf_mode == 1 || p->ctf_mode == 2))
        return fail("Error: beta/delta must be negative in TIE-Hom phase retrieval.");
    if (p->nsr < 0) return fail("Error: NSR in input images must be non-negative.");
    if (p->epsilon < 0) return fail("Error: the regulalization parameter must be non-negative.");
    if (p->epsilon == 0 && p->asigma == 0) return fail("Error: when beta/delta == 0, epsilon must be positive.");
    if (p->inv_laplace != 0 && p->inv_laplace != 1)
        return fail("Error: unknown value for 3D Laplacian filter mode in input parameter file.");

    g->nx = p->nx; g->ny = p->ny;
    g->nxd2 = p->nx / 2; g->nyd2 = p->ny / 2;
    g->modality = p->modality; g->ctf_mode = p->ctf_mode; g->escc = p->escc;
    g->inv_laplace = p->inv_laplace;
    /* UTR.cpp:294-306 */
    g->xst = p->xst; g->yst = p->xst;
    g->xlo = -0.5 * g->xst * g->nx; g->xhi = g->xlo + g->xst * g->nx;
    g->ylo = -0.5 * g->yst * g->ny; g->yhi = g->ylo + g->yst * g->ny;
    g->fxst = 1.0 / (g->xhi - g->xlo); g->fyst = 1.0 / (g->yhi - g->ylo);
    g->fxlo = -g->fxst * g->nxd2; g->fylo = -g->fyst * g->nyd2;
    /* UTR.cpp:308-329 */
    double zlo = p->zlo, zhi = p->zhi;
    if (zlo > zhi) { double t = zlo; zlo = zhi; zhi = t; }
    g->zlo = zlo; g->zhi = zhi; g->zst = g->xst;
    long nz = (long)((zhi - zlo) / g->zst + 0.5);
    if (nz <= 0) return fail("Error: number of z steps is not positive.");
    if (nz % 2) return fail("Error: number of z points must be even.");
    g->nz = (int)nz; g->nzd2 = g->nz / 2;
    g->fzst = 1.0 / (zhi - zlo);
    g->fzlo = -g->fzst * g->nzd2;
    g->dzextra = p->dzextra;
    /* UTR.cpp:232-238, 330 */
    double za = defocus[0];
    for (int i = 1; i < nrows; i++) za += defocus[i];
    za /= nrows;
    g->zout_aver = za - p->dzextra;
    /* UTR.cpp:335-345 */
    g->dxc = p->dxc; g->dyc = p->dyc; g->dzc = p->dzc;
    g->rot_centre_shift = (p->dxc != 0 || p->dyc != 0 || p->dzc != 0);
    g->xc = (g->xhi + g->xlo) / 2.0 + p->dxc;
    g->yc = (g->yhi + g->ylo) / 2.0 + p->dyc;
    g->zc = (g->zhi + g->zlo) / 2.0 + p->dzc;
    g->ic = (long)((g->xc - g->xlo) / g->xst + 0.5);
    g->jc = (long)((g->yc - g->ylo) / g->yst + 0.5);
    g->kc = (long)((g->zc - g->zlo) / g->zst + 0.5);
    /* UTR.cpp:349-370 */
    g->wl = p->wl;
    if (p->modality == 0) {
        const double hp = 6.62607004e-34, cc = 299792458, ee = 1.602176634e-19,
                     m0 = 9.1093837015e-31;
        const long double mc2 = m0 * cc * cc;
        long double chl2 = (long double)(cc * cc * hp * hp) / (long double)(p->wl * p->wl * 1.e-20);
        long double abra = sqrtl(mc2 * mc2 + chl2);
        g->EE = (double)((-mc2 + abra) / (long double)(ee));
    } else {
        g->EE = 0.001239841662513396 / p->wl;
    }
    /* UTR.cpp:388 (the 0.001f literal is single precision in the reference) */
    g->k2maxo = pow(p->aobj_mrad * (double)0.001f / p->wl, 2.0);
    g->cs3 = p->cs3_mm * 1.e+7; g->cs5 = p->cs5_mm * 1.e+7;
    g->asigma = p->asigma; g->nsr = p->nsr; g->epsilon = p->epsilon;
    g->norm_f1 = p->norm_f1; g->norm_f2 = p->norm_f2;
    /* UTR.cpp:211-222: the loop starts at row 1 */
    g->rot_y_only = g->rot_z_only = g->rot_z2_only = 1;
    for (int i = 1; i < nrows; i++) {
        if (angZ_deg[i] != 0) { g->rot_y_only = 0; g->rot_z2_only = 0; }
        if (angY_deg[i] != 0) { g->rot_z_only = 0; g->rot_z2_only = 0; }
        if (angZ2_deg && angZ2_deg[i] != 0) { g->rot_y_only = 0; g->rot_z_only = 0; }
    }
    /* UTR.cpp:1236-1243 */
    long nxR2 = (g->ic <= g->nxd2) ? g->ic * g->ic : (g->nx - g->ic) * (g->nx - g->ic);
    long nyR2 = (g->jc <= g->nyd2) ? g->jc * g->jc : (g->ny - g->jc) * (g->ny - g->jc);
    long nzR2 = (g->kc <= g->nzd2) ? g->kc * g->kc : (g->nz - g->kc) * (g->nz - g->kc);
    if (g->rot_y_only) g->nR2 = nxR2 < nzR2 ? nxR2 : nzR2;
    else if (g->rot_z_only || g->rot_z2_only) g->nR2 = nxR2 < nyR2 ? nxR2 : nyR2;
    else {
        long t = nxR2 < nyR2 ? nxR2 : nyR2;
        g->nR2 = t < nzR2 ? t : nzR2;
    }
    g->mask_dtemp = 2.0 / (double)(g->kc * g->kc + g->jc * g->jc + g->ic * g->ic);
    /* UTR.cpp:1273-1277 */
    double s2 = sqrt(1.0 + p->asigma * p->asigma);
    if (p->modality == 0) g->fnorm = g->EE * p->wl / (2.0 * ORA_PI * s2 * g->zst);
    else g->fnorm = p->asigma * p->wl / (4.0 * ORA_PI * s2 * g->zst);
    /* filter preconditions (UTR.h:67-68,148-152) */
    if (g->zout_aver != 0 && g->ctf_mode == 2) {
        double alpha = -4.0 * ORA_PI * g->asigma / (g->zout_aver * g->wl);
        if (alpha <= 0) return fail("InverseMLaplacian3D() cannot be called with alpha <= 0");
    }
    if (g->inv_laplace && g->ctf_mode != 1 && g->ctf_mode != 2 && g->epsilon <= 0)
        return fail("InverseMLaplacian3D() cannot be called with alpha <= 0");
    if (g->zout_aver != 0 && g->ctf_mode == 4) {
        double xs = (g->xhi - g->xlo) / g->nx, ys = (g->yhi - g->ylo) / g->ny,
               zs = (g->zhi - g->zlo) / g->ny; /* GetZStep(ny), UTR.h:54 */
        double srad = sqrt(0.25 / (xs * xs) + 0.25 / (ys * ys) + 0.25 / (zs * zs));
        if (srad * g->wl > 1.0)
            return fail("runtime_error in CTFcorrection3D() (evanescent waves present)");
    }
    return 0;
}

void ora_angle_coeffs(const ora_geom* g, double angZ_deg, double angY_deg, double* m) {
    /* UTR.cpp:883-888 angles; UTR.h:243-256 rotation, rewritten in grid units */
    double aZ = angZ_deg * ORA_PI180, aY = angY_deg * ORA_PI180;
    double cY = cos(aY), sY = sin(aY), cZ = cos(aZ), sZ = sin(aZ);
    m[0] = cY * cZ;                      /* dx1/da  */
    m[1] = sZ * (g->fyst / g->fxst);     /* dx1/db  */
    m[2] = -(sY * cZ) / g->fxst;         /* dx1/dzz */
    m[3] = -(cY * sZ) * (g->fxst / g->fyst);
    m[4] = cZ;
    m[5] = (sY * sZ) / g->fyst;
    m[6] = sY * (g->fxst / g->fzst);
    m[7] = 0.0;
    m[8] = cY / g->fzst;
}

/* ------------------------------------------------------------ 2D stage */

/* per-row astigmatism for the mode-3 filter (ora_set_astigmatism) */
static const double* g_astig_dz = NULL;
static const double* g_astig_ang = NULL;
static int g_astig_n = 0;

void ora_set_astigmatism(int n, const double* dz, const double* angle_deg) {
    g_astig_n = (dz && angle_deg) ? n : 0;
    g_astig_dz = dz;
    g_astig_ang = angle_deg;
}

int ora_spectrum_astig(const ora_geom* g, const float* img, double zout, double weight,
                       double adz, double aang_deg, double* K);

int ora_spectrum(const ora_geom* g, const float* img, double zout, double weight,
                 double* K) {
    return ora_spectrum_astig(g, img, zout, weight, 0.0, 0.0, K);
}

int ora_spectrum_astig(const ora_geom* g, const float* img, double zout, double weight,
                       double adz, double aang_deg, double* K) {
    const int nx = g->nx, ny = g->ny;
    /* UTR.cpp:951-958 and Log0 (UTR.cpp:991) */
    const float f1 = (float)g->norm_f1, f2 = (float)g->norm_f2;
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            float x = img[(size_t)j * nx + i];
            if (g->norm_f1 != 1.0) x = x / f1;
            if (g->norm_f2 != 0.0) x = x + f2;
            if (x < 0) x = 0.0f;
            double f = (x > 0) ? log((double)x) : 0.0;
            double sgn = ((i + j) & 1) ? -1.0 : 1.0;
            K[2 * ((size_t)j * nx + i)] = f * sgn;
            K[2 * ((size_t)j * nx + i) + 1] = 0.0;
        }
    fft_plan px, py;
    plan_init(&px, nx);
    plan_init(&py, ny);
    fft_axis(K, nx, 1, ny, 1, -1, &px);   /* along x */
    fft_axis(K, ny, nx, 1, nx, -1, &py);  /* along y */
    plan_free(&px);
    plan_free(&py);
    const double gsgn = ((g->nxd2 + g->nyd2) & 1) ? -1.0 : 1.0;
    for (int ky = 0; ky < ny; ky++)
        for (int kx = 0; kx < nx; kx++) {
            double s = (((kx + ky) & 1) ? -1.0 : 1.0) * gsgn;
            K[2 * ((size_t)ky * nx + kx)] *= s;
            K[2 * ((size_t)ky * nx + kx) + 1] *= s;
        }
    const double xap = g->xhi - g->xlo, yap = g->yhi - g->ylo;
    if (g->ctf_mode == 1 && zout != 0) {
        /* 2D TIE-Hom (UTR.cpp:1001-1006) restated from InverseMLaplacian3D
         * (UTR.h:159-181) in 2D (the function is in the absent XArrayLibrary).
         * Pinned in tests/test_filter_physics.py: it returns the thickness of
         * Paganin's homogeneous-object forward model to 0.5 % and equals
         * InverseMLaplacian3D on the q_z = 0 plane. */
        double alpha = -4.0 * ORA_PI * g->asigma / (g->zout_aver * g->wl);
        double norm = -alpha * sqrt(1.0 + g->asigma * g->asigma) / g->asigma;
        double fact = 4.0 * ORA_PI * ORA_PI;
        for (int ky = 0; ky < ny; ky++)
            for (int kx = 0; kx < nx; kx++) {
                double j1 = ky - g->nyd2, i1 = kx - g->nxd2;
                double q2 = fact * (i1 * i1 / (xap * xap) + j1 * j1 / (yap * yap));
                double f = norm / (alpha + q2);
                K[2 * ((size_t)ky * nx + kx)] *= f;
                K[2 * ((size_t)ky * nx + kx) + 1] *= f;
            }
    } else if (g->ctf_mode == 3) {
        /* 2D CTF inversion (UTR.cpp:1011-1016) restated from CTFcorrection3D
         * (UTR.h:70-128) in 2D (the function is in the absent XArrayLibrary).
         * Pinned in tests/test_filter_physics.py: it recovers the multislice
         * exit-wave phase of a thin specimen where the CTF is strong and
         * equals CTFcorrection3D on the q_z = 0 plane. */
        double omega = atan(g->asigma);
        double fac2 = ORA_PI * g->wl * zout;
        double dk2 = 1.0 / (xap * xap), de2 = 1.0 / (yap * yap);
        double eps = g->epsilon;
        if (eps <= 0) {
            double mn = dk2 < de2 ? dk2 : de2;
            eps = 0.1 * pow(fabs(fac2) * mn, 2);
        }
        double fac3 = ORA_PI * pow(g->wl, 3) / 2.0 * g->cs3;
        double fac5 = ORA_PI * pow(g->wl, 5) / 3.0 * g->cs5;
        int bC35 = (g->cs3 != 0) || (g->cs5 != 0);
        int aper = g->k2maxo > 0;
        /* astigmatism (UTR.cpp:1013, restated): defocus zout + dz cos 2(phi - phi0)
         * at azimuth phi of (qx, qy), i.e. chi += pi wl dz ((qx^2 - qy^2) cos 2phi0
         * + 2 qx qy sin 2phi0) */
        double a2 = 2.0 * aang_deg * ORA_PI / 180.0;
        double adc = ORA_PI * g->wl * adz * cos(a2), ads = ORA_PI * g->wl * adz * sin(a2);
        double dxy = 1.0 / (xap * yap);
        for (int ky = 0; ky < ny; ky++)
            for (int kx = 0; kx < nx; kx++) {
                double j1 = ky - g->nyd2, i1 = kx - g->nxd2;
                double q2 = dk2 * i1 * i1 + de2 * j1 * j1;
                if (aper && !(q2 < g->k2maxo)) continue;
                double chi = fac2 * q2 - omega;
                if (adz != 0) chi += adc * (dk2 * i1 * i1 - de2 * j1 * j1) + 2.0 * ads * dxy * i1 * j1;
                if (bC35) chi += fac3 * q2 * q2 + fac5 * q2 * q2 * q2;
                double sn = sin(chi), cs = cos(chi);
                double sr = sn / (sn * sn + eps);
                double cr = g->escc ? cs / (cs * cs + eps) : 0.0;
                double* k = &K[2 * ((size_t)ky * nx + kx)];
                double re = k[0] * sr + k[1] * cr, im = k[1] * sr - k[0] * cr;
                k[0] = re; k[1] = im;
            }
    }
    if (weight != 1.0) /* UTR.cpp:1019 */
        for (size_t t = 0; t < 2 * (size_t)nx * ny; t++) K[t] *= weight;
    return 0;
}

/* ------------------------------------------------------------ gridding */

int ora_grid(const ora_geom* g, const double* K, double angZ_deg, double angY_deg,
             double* V, double* S) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const long nxny = (long)nx * ny;
    double m[9];
    ora_angle_coeffs(g, angZ_deg, angY_deg, m);
    const double wl2 = 0.5 * g->wl;
    /* centre-of-rotation phase (UTR.cpp:893-899, UTR.h:264) */
    double Ax = 0, Ay = 0, Az = 0;
    if (g->rot_centre_shift) {
        double aZ = angZ_deg * ORA_PI180, aY = angY_deg * ORA_PI180;
        double cY = cos(aY), sY = sin(aY), cZ = cos(aZ), sZ = sin(aZ);
        double tPI = 2.0 * ORA_PI;
        Ax = tPI * (g->dxc * (1.0 - cY * cZ) - g->dyc * sZ + g->dzc * sY * cZ);
        Ay = tPI * (g->dxc * cY * sZ + g->dyc * (1.0 - cZ) - g->dzc * sY * sZ);
        Az = tPI * (-g->dxc * sY + g->dzc * (1.0 - cY));
    }
    for (int j = 0; j < ny; j++) {
        const double b = (double)(j - g->nyd2);
        const double qy = g->fyst * b;
        for (int i = 0; i < nx; i++) {
            const double a = (double)(i - g->nxd2);
            const double qx = g->fxst * a;
            const double zz = g->escc ? -wl2 * fma(qx, qx, qy * qy) : 0.0;
            const double ox = fma(a, m[0], fma(b, m[1], zz * m[2]));
            const double oy = fma(a, m[3], fma(b, m[4], zz * m[5]));
            const double oz = fma(a, m[6], zz * m[8]);
            const double dx1 = (double)g->nxd2 + ox;
            const double dy1 = (double)g->nyd2 + oy;
            const double dz1 = (double)g->nzd2 + oz;
            const double fnn = floor(dz1);
            if (fnn < 0 || fnn > nz - 2) continue;
            const double fii = floor(dx1);
            if (fii < 0 || fii > nx - 2) continue;
            const double fjj = floor(dy1);
            if (fjj < 0 || fjj > ny - 2) continue;
            const long ii = (long)fii, jj = (long)fjj, nn = (long)fnn;
            const double fx = dx1 - fii, fy = dy1 - fjj, fz = dz1 - fnn;
            const double wx[2] = {1.0 - fx, fx}, wy[2] = {1.0 - fy, fy}, wz[2] = {1.0 - fz, fz};
            double cr = K[2 * ((size_t)j * nx + i)], ci = K[2 * ((size_t)j * nx + i) + 1];
            if (g->rot_centre_shift) {
                double ph = -(Ax * (g->fxst * ox) + Ay * (g->fyst * oy) + Az * (g->fzst * oz));
                double c = cos(ph), s = sin(ph);
                double r2 = cr * c - ci * s, i2 = cr * s + ci * c;
                cr = r2; ci = i2;
            }
            for (int cz = 0; cz < 2; cz++)
                for (int cyy = 0; cyy < 2; cyy++)
                    for (int cx = 0; cx < 2; cx++) {
                        double w = wx[cx] * wy[cyy] * wz[cz];
                        long idx = (nn + cz) * nxny + (jj + cyy) * nx + (ii + cx);
                        S[idx] += w;
                        V[2 * idx] += w * cr;
                        V[2 * idx + 1] += w * ci;
                    }
        }
    }
    return 0;
}

/* ------------------------------------------------------ 3D filters */

static void inv_mlaplacian3d(const ora_geom* g, double* V, double zaper, double yaper,
                             double xaper, double alpha, double norm) {
    /* UTR.h:135-185 */
    const double fact = 4.0 * ORA_PI * ORA_PI;
    const double dksi2 = fact / (xaper * xaper), deta2 = fact / (yaper * yaper),
                 dzeta2 = fact / (zaper * zaper);
    for (int k = 0; k < g->nz; k++) {
        int k1 = k - g->nzd2;
        double zeta2 = k1 * k1 * dzeta2 + alpha;
        for (int j = 0; j < g->ny; j++) {
            int j1 = j - g->nyd2;
            double eta2 = j1 * j1 * deta2 + zeta2;
            for (int i = 0; i < g->nx; i++) {
                int i1 = i - g->nxd2;
                double ksi2 = dksi2 * i1 * i1 + eta2;
                double f = norm / ksi2;
                size_t idx = ((size_t)k * g->ny + j) * g->nx + i;
                V[2 * idx] *= f;
                V[2 * idx + 1] *= f;
            }
        }
    }
}

static void ctf_correction3d(const ora_geom* g, double* V) {
    /* UTR.h:15-133 with head3D = (wl, zlo, zhi, ylo, yhi, xlo, xhi),
     * defocdist = zoutAver, q2max = k2maxo (UTR.cpp:1191) */
    const int aper = g->k2maxo > 0;
    const double xap2 = (g->xhi - g->xlo) * (g->xhi - g->xlo);
    const double yap2 = (g->yhi - g->ylo) * (g->yhi - g->ylo);
    const double zap2 = (g->zhi - g->zlo) * (g->zhi - g->zlo);
    const double omega = atan(g->asigma);
    const int bC35 = (g->cs3 != 0) || (g->cs5 != 0);
    const double dksi2 = 1.0 / xap2, deta2 = 1.0 / yap2, dzeta2 = 1.0 / zap2;
    const double fac2 = ORA_PI * g->wl * g->zout_aver;
    double eps = g->epsilon;
    if (eps <= 0) {
        double mn = dksi2 < deta2 ? dksi2 : deta2;
        mn = mn < dzeta2 ? mn : dzeta2;
        eps = 0.1 * pow(fabs(fac2) * mn, 2);
    }
    const double fac3 = ORA_PI * pow(g->wl, 3) / 2.0 * g->cs3;
    const double fac5 = ORA_PI * pow(g->wl, 5) / 3.0 * g->cs5;
    for (int k = 0; k < g->nz; k++) {
        int k1 = k - g->nzd2;
        double zeta2 = dzeta2 * k1 * k1;
        double zeta2fac2 = fac2 * zeta2 - omega;
        for (int j = 0; j < g->ny; j++) {
            int j1 = j - g->nyd2;
            double eta2 = deta2 * j1 * j1;
            double eta2fac2 = fac2 * eta2 + zeta2fac2;
            eta2 += zeta2;
            for (int i = 0; i < g->nx; i++) {
                int i1 = i - g->nxd2;
                double ksi2 = dksi2 * i1 * i1;
                double ksi2fac2 = fac2 * ksi2 + eta2fac2;
                ksi2 += eta2;
                if (!aper || ksi2 < g->k2maxo) {
                    double cst = bC35 ? fac3 * ksi2 * ksi2 + fac5 * pow(ksi2, 3) : 0.0;
                    double dt = ksi2fac2 + cst;
                    double sn = sin(dt);
                    double sr = sn / (sn * sn + eps);
                    double cr = 0.0;
                    if (g->escc) {
                        double cs = cos(dt);
                        cr = cs / (cs * cs + eps);
                    }
                    size_t idx = ((size_t)k * g->ny + j) * g->nx + i;
                    double re = V[2 * idx], im = V[2 * idx + 1];
                    /* (re + i im) * (sr - i cr) */
                    V[2 * idx] = re * sr + im * cr;
                    V[2 * idx + 1] = im * sr - re * cr;
                }
            }
        }
    }
}

int ora_normalize_filter(const ora_geom* g, double* V, const double* S) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t N = (size_t)nx * ny * nz;
    /* UTR.cpp:1052 */
    const double nsr2 = g->nsr * sqrt((double)((size_t)nx * ny));
    if (nsr2 == 0) {
        for (size_t t = 0; t < N; t++) {
            double x = S[t];
            if (x == 0) { V[2 * t] = 0; V[2 * t + 1] = 0; }
            else { V[2 * t] *= 1.0 / x; V[2 * t + 1] *= 1.0 / x; }
        }
    } else if (g->ctf_mode != 1) {
        /* UTR.cpp:1072-1091 */
        for (size_t t = 0; t < N; t++) {
            double x = S[t];
            if (x == 0) { V[2 * t] = 0; V[2 * t + 1] = 0; continue; }
            double y = hypot(V[2 * t], V[2 * t + 1]) / (nsr2 * sqrt(x));
            double f = (1.0 / x) * (0.5 + atan(100.0 * (y - ORA_TARGET_SNR)) / ORA_PI);
            V[2 * t] *= f; V[2 * t + 1] *= f;
        }
    } else {
        /* UTR.cpp:1094-1128 */
        double norm = sqrt(1.0 + g->asigma * g->asigma) * nsr2;
        double fac2 = ORA_PI * g->wl * g->zout_aver;
        double dksi2 = fac2 / ((g->xhi - g->xlo) * (g->xhi - g->xlo));
        double deta2 = fac2 / ((g->yhi - g->ylo) * (g->yhi - g->ylo));
        double dzeta2 = fac2 / ((g->zhi - g->zlo) * (g->zhi - g->zlo));
        for (int k = 0; k < nz; k++) {
            int k1 = k - g->nzd2;
            double zeta2 = k1 * k1 * dzeta2 - g->asigma;
            for (int j = 0; j < ny; j++) {
                int j1 = j - g->nyd2;
                double eta2 = j1 * j1 * deta2 + zeta2;
                for (int i = 0; i < nx; i++) {
                    size_t t = ((size_t)k * ny + j) * nx + i;
                    double x = S[t];
                    if (x == 0) { V[2 * t] = 0; V[2 * t + 1] = 0; continue; }
                    int i1 = i - g->nxd2;
                    double fn = norm * sqrt(x) / (dksi2 * i1 * i1 + eta2 + g->epsilon);
                    double y = hypot(V[2 * t], V[2 * t + 1]) / fn;
                    double f = (1.0 / x) * (0.5 + atan(100.0 * (y - ORA_TARGET_SNR)) / ORA_PI);
                    V[2 * t] *= f; V[2 * t + 1] *= f;
                }
            }
        }
    }
    /* UTR.cpp:1178-1201 (the x-aperture argument xhi - ylo is as written) */
    if (g->zout_aver != 0) {
        if (g->ctf_mode == 2) {
            double alpha = -4.0 * ORA_PI * g->asigma / (g->zout_aver * g->wl);
            double norm = -alpha * sqrt(1.0 + g->asigma * g->asigma) / g->asigma;
            inv_mlaplacian3d(g, V, g->zhi - g->zlo, g->yhi - g->ylo, g->xhi - g->ylo, alpha, norm);
        } else if (g->ctf_mode == 4) {
            ctf_correction3d(g, V);
        }
    }
    if (g->inv_laplace && g->ctf_mode != 1 && g->ctf_mode != 2)
        inv_mlaplacian3d(g, V, g->zhi - g->zlo, g->yhi - g->ylo, g->xhi - g->ylo, g->epsilon,
                         g->epsilon);
    return 0;
}

int ora_hermitian(const ora_geom* g, double* V) {
    /* UTR.cpp:1203-1217, executed in serial loop order */
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    for (int k = 0; k < nz; k++) {
        int ka = (k == 0) ? 0 : nz - k;
        for (int j = 0; j < ny; j++) {
            int ja = (j == 0) ? 0 : ny - j;
            for (int i = 0; i < nx / 2; i++) {
                int ia = (i == 0) ? 0 : nx - i;
                size_t d = ((size_t)k * ny + j) * nx + ia;
                size_t s = ((size_t)ka * ny + ja) * nx + i;
                double re = 0.5 * (V[2 * d] + V[2 * s]);
                double im = 0.5 * (V[2 * d + 1] - V[2 * s + 1]);
                V[2 * d] = re;
                V[2 * d + 1] = im;
            }
        }
    }
    return 0;
}

int ora_c2r(const ora_geom* g, const double* V, double* out) {
    /* Fftwd3frc SetComplexXArray3D(V3, true) / InverseFFT / GetRealXArray3D
     * (K3out, true) (UTR.cpp:1223-1232): un-centre, keep wrapped x-frequency
     * 0..nx/2 (centred i = nx/2..nx-1, 0), Hermitian-extend, inverse DFT,
     * real part, 1/N, re-centre. */
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t N = (size_t)nx * ny * nz;
    double* E = (double*)malloc(sizeof(double) * 2 * N);
    if (!E) return fail("out of memory");
    for (int wz = 0; wz < nz; wz++)
        for (int wy = 0; wy < ny; wy++)
            for (int wx = 0; wx < nx; wx++) {
                int sz, sy, sx, cj;
                if (wx <= nx / 2) { sz = wz; sy = wy; sx = wx; cj = 0; }
                else { sz = (nz - wz) % nz; sy = (ny - wy) % ny; sx = nx - wx; cj = 1; }
                int ck = (sz + g->nzd2) % nz, cy = (sy + g->nyd2) % ny, cx = (sx + g->nxd2) % nx;
                size_t s = ((size_t)ck * ny + cy) * nx + cx;
                size_t d = ((size_t)wz * ny + wy) * nx + wx;
                E[2 * d] = V[2 * s];
                E[2 * d + 1] = cj ? -V[2 * s + 1] : V[2 * s + 1];
            }
    fft_plan px, py, pz;
    plan_init(&px, nx); plan_init(&py, ny); plan_init(&pz, nz);
    fft_axis(E, nx, 1, (long)ny * nz, 1, +1, &px);
    fft_axis(E, ny, nx, nz, nx, +1, &py);
    fft_axis(E, nz, (long)nx * ny, 1, (long)nx * ny, +1, &pz);
    plan_free(&px); plan_free(&py); plan_free(&pz);
    const double inv = 1.0 / (double)N;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                int wz = (k + g->nzd2) % nz, wy = (j + g->nyd2) % ny, wx = (i + g->nxd2) % nx;
                out[((size_t)k * ny + j) * nx + i] = E[2 * (((size_t)wz * ny + wy) * nx + wx)] * inv;
            }
    free(E);
    return 0;
}

int ora_mask_scale(const ora_geom* g, double* out) {
    /* UTR.cpp:1244-1262 and 1279 */
    for (int k = 0; k < g->nz; k++) {
        long kk2 = 0;
        if (!g->rot_z_only && !g->rot_z2_only) kk2 = (k - g->kc) * (k - g->kc);
        for (int j = 0; j < g->ny; j++) {
            long jj2 = g->rot_y_only ? kk2 : (j - g->jc) * (j - g->jc) + kk2;
            for (int i = 0; i < g->nx; i++) {
                long ii2 = (i - g->ic) * (i - g->ic) + jj2;
                size_t t = ((size_t)k * g->ny + j) * g->nx + i;
                if (ii2 > g->nR2) out[t] = 0.0;
                else out[t] *= exp(g->mask_dtemp * ii2) * g->fnorm;
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------ driver */

int ora_reconstruct(const ora_params* p, int nrows, const double* angZ_deg,
                    const double* angY_deg, const double* angZ2_deg,
                    const double* defocus, const float* weights, const float* images,
                    double* vol, double* samp, int nthreads) {
    ora_geom g;
    if (ora_derive(p, nrows, angZ_deg, angY_deg, angZ2_deg, defocus, &g)) return -1;
    const size_t N = (size_t)g.nx * g.ny * g.nz, n2 = (size_t)g.nx * g.ny;
    double t0 = now_ms();
    double t_fft = 0, t_grid = 0;
    if (nthreads < 1) nthreads = 1;
    double* V = (double*)calloc(2 * N, sizeof(double));
    double* S = (double*)calloc(N, sizeof(double));
    if (!V || !S) { free(V); free(S); return fail("out of memory"); }
    if (nthreads == 1) {
        double* K = (double*)malloc(sizeof(double) * 2 * n2);
        for (int r = 0; r < nrows; r++) {
            double a = now_ms();
            ora_spectrum_astig(&g, images + (size_t)r * n2, defocus[r] - g.dzextra,
                           weights ? (double)weights[r] : 1.0, r < g_astig_n ? g_astig_dz[r] : 0.0,
                           r < g_astig_n ? g_astig_ang[r] : 0.0, K);
            double b = now_ms();
            ora_grid(&g, K, angZ_deg[r], angY_deg[r], V, S);
            t_grid += now_ms() - b;
            t_fft += b - a;
        }
        free(K);
    } else {
#ifdef _OPENMP
        /* angle-parallel with private accumulators, summed in thread order */
        double** Vp = (double**)calloc(nthreads, sizeof(double*));
        double** Sp = (double**)calloc(nthreads, sizeof(double*));
        double a0 = now_ms();
#pragma omp parallel num_threads(nthreads)
        {
            int t = omp_get_thread_num();
            double* Vt = t == 0 ? V : (double*)calloc(2 * N, sizeof(double));
            double* St = t == 0 ? S : (double*)calloc(N, sizeof(double));
            Vp[t] = Vt; Sp[t] = St;
            double* K = (double*)malloc(sizeof(double) * 2 * n2);
#pragma omp for schedule(static)
            for (int r = 0; r < nrows; r++) {
                ora_spectrum_astig(&g, images + (size_t)r * n2, defocus[r] - g.dzextra,
                               weights ? (double)weights[r] : 1.0, r < g_astig_n ? g_astig_dz[r] : 0.0,
                               r < g_astig_n ? g_astig_ang[r] : 0.0, K);
                ora_grid(&g, K, angZ_deg[r], angY_deg[r], Vt, St);
            }
            free(K);
#pragma omp barrier
#pragma omp for schedule(static)
            for (long v = 0; v < (long)N; v++)
                for (int q = 1; q < nthreads; q++) {
                    V[2 * v] += Vp[q][2 * v];
                    V[2 * v + 1] += Vp[q][2 * v + 1];
                    S[v] += Sp[q][v];
                }
        }
        for (int q = 1; q < nthreads; q++) { free(Vp[q]); free(Sp[q]); }
        free(Vp); free(Sp);
        t_grid = now_ms() - a0; /* combined: the loop interleaves both phases */
#else
        free(V); free(S);
        return fail("oracle built without OpenMP");
#endif
    }
    if (samp) memcpy(samp, S, sizeof(double) * N);
    ora_normalize_filter(&g, V, S);
    free(S);
    ora_hermitian(&g, V);
    double c = now_ms();
    int rc = ora_c2r(&g, V, vol);
    double t3 = now_ms() - c;
    free(V);
    if (rc) return rc;
    ora_mask_scale(&g, vol);
    g_timings[0] = t_fft;
    g_timings[1] = t_grid;
    g_timings[2] = t3;
    g_timings[3] = now_ms() - t0;
    return 0;
}
