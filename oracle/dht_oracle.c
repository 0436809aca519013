/*
 * dht_oracle.c -- TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.
 *
 * FP64 CPU restatement of the single-defocus reconstructions of the DHT
 * program (timg021/UTR_public, MultisliceCppTest/PhaseRetrieval/DHT.cpp):
 *   method 3  phase estimate Fact*(-0.5)*(I/I0 - 1) (ReadmePhaseRetrieval.txt:
 *             "3" and parameter 18), Fresnel back-propagation to every output
 *             plane, contrast 1 - |psi|^2 (DHT.cpp:919-1006);
 *   method 4  variable-distance weak-object CTF inversion per plane
 *             (DHT.cpp:1010-1055);
 *   splat     real-space inverse rotation and trilinear splat with (int)
 *             truncation (DHT.cpp:1081-1164);
 *   method 5  sCTF on the Ewald sphere (DHT.cpp:1056-1067, 1195-1308), Filt3
 *             filter, Hermitian fix, c2r and scale (DHT.cpp:1341-1416);
 *   methods 1, 2  the complex amplitude at zout retrieved from a defocus
 *             series per orientation (IWFR, CTFL2; DHT.cpp:889-917), then
 *             method 3's back-propagation and splat;
 *   scales    DHT.cpp:1325-1340.
 * The absent XArrayLibrary functions are defined as follows (decision
 * register B11 and the weak-object model):
 *   Fresnel(d, Cs3, Cs5): spectrum * exp(-i*chi), chi = pi*wl*d*q^2
 *     + pi*wl^3*Cs3*q^4/2 + pi*wl^5*Cs5*q^6/3, zero for q^2 >= k2maxo when
 *     an aperture is set; 2D DFT exponent -1 forward, +1 and 1/(nx*ny) back.
 *   InvertPhaseCTF (plane at distance D = zout - z before the image plane):
 *     phi^ = (I-1)^ * T/(2(T^2 + eps)), T = sin(chi) - sigma*cos(chi), chi as
 *     above with d = D and the microscope's Cs (the inverse of the weak
 *     phase-amplitude object model I^ = 2 T phi^).
 *   InvertCTF_DT1: the same restated 2D CTF inversion as UTR mode 3.
 *   XA_IWFR::Iwfr and XA_IWFR::CTFL2: restated from the publications that
 *     ReadmePhaseRetrieval.txt names for them (see dht_iwfr, dht_ctfl2);
 *     the propagation between two planes of a series is the free-space
 *     propagator with the aperture, since the objective's Cs phases are the
 *     same at every defocus and cancel between two planes.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "utr_oracle.h"

#define DHT_PI 3.14159265358979323846
#define DHT_PI180 (DHT_PI / 180.0)

/* errors are reported through ora_last_error() like the rest of the oracle */
static int dfail(const char* m);

typedef struct {
    int nx, ny, nz;
    double xst, xlo, xhi, ylo, yhi, zlo, zhi, zlodz, zhidz, xc, yc, zc;
    double EE, k2maxo, cs3, cs5;
} dht_geom;

static double tem_energy(double wl) {
    /* UTR.cpp:352-366 / DHT.cpp:197-205 */
    const double hp = 6.62607004e-34, cc = 299792458, ee = 1.602176634e-19, m0 = 9.1093837015e-31;
    const long double mc2 = m0 * cc * cc;
    long double chl2 = (long double)(cc * cc * hp * hp) / (long double)(wl * wl * 1.e-20);
    long double abra = sqrtl(mc2 * mc2 + chl2);
    return (double)((-mc2 + abra) / (long double)(ee));
}

static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

static int dht_derive(const ora_dht_params* p, dht_geom* g) {
    memset(g, 0, sizeof *g);
    if (p->method < 1 || p->method > 5) return dfail("Error: unknown phase retrieval method in the input parameter file.");
    if (p->asigma != 0 && !(p->method == 4 || p->method == 5))
        return dfail("Error: selected phase retrieval method is incompatibe non-zero absorption coefficient in the input parameter file.");
    if (p->method != 5) {
        if (!is_pow2(p->nx) || !is_pow2(p->ny))
            return dfail("Error: Width and Height parameters in input parameter file must be integer powers of 2.");
    } else if (p->nx <= 0 || p->ny <= 0 || p->nx % 2 || p->ny % 2) {
        return dfail("Error: Width and Height parameters in input parameter file must be even.");
    }
    if (!(p->xst > 0)) return dfail("Error: pixel size must be positive in input parameter file.");
    if (!(p->wl > 0) || p->wl > 1) return dfail("Error: wavelength value appears to be wrong.");
    if (p->aobj_mrad < 0 || p->aobj_mrad > 1000) return dfail("Error: objective aperture value appears to be wrong.");
    if (p->escc != 0 && p->escc != 1)
        return dfail("Error: unknown value for Ewald sphere curvature correction switch in input parameter file.");
    if (p->norm_f1 == 0) return dfail("The first normalization factor cannot be zero (leads to division by zero).");
    if (p->method == 1 && p->itermax < 1) return dfail("Error: the maximal number of iterations must be >= 1.");
    if (p->method == 5 && p->critical_snr < 0) return dfail("Error: the critical SNR value must be non-negative.");
    if (p->epsilon < 0 && p->method != 3)
        return dfail("Error: minimal phase error / regulalization parameter must be non-negative, unless -0.5LogAmp phase retrieval method is selected.");
    if ((p->method == 2 || p->method == 4) && p->epsilon == 0)
        return dfail("Error: the estimated regularization parameter (epsilon = 0) is not supported; give epsilon > 0.");
    if (p->inv_laplace && p->alpha < 0)
        return dfail("Error: regularization parameter for 3D Laplacian filter must be non-negative.");
    if (p->dzextra == 0 && p->method <= 3)
        return dfail("Error: extra defocus = 0, which is unexpected for the selected reconstruction (phase retrieval) method.");
    if (p->dzextra != 0 && (p->method == 4 || p->method == 5))
        return dfail("ERROR: extra defocus != 0, which is unexpected for the phase retrieval method selected above.");
    g->nx = p->nx; g->ny = p->ny; g->xst = p->xst;
    g->xlo = -0.5 * p->xst * p->nx; g->xhi = g->xlo + p->xst * p->nx;
    g->ylo = -0.5 * p->xst * p->ny; g->yhi = g->ylo + p->xst * p->ny;
    double zlo = p->zlo, zhi = p->zhi;
    if (zlo > zhi) { double t = zlo; zlo = zhi; zhi = t; }
    g->zlo = zlo; g->zhi = zhi;
    g->zlodz = zlo + p->dzextra; g->zhidz = zhi + p->dzextra;  /* DHT.cpp:376-377 */
    long nz = (long)((g->zhidz - g->zlodz) / p->xst + 0.5);
    if (nz <= 0) return dfail("Error: number of output defocus planes must be positive.");
    g->nz = (int)nz;
    g->xc = (g->xhi + g->xlo) / 2.0 + p->dxc;
    g->yc = (g->yhi + g->ylo) / 2.0 + p->dyc;
    g->zc = (g->zhidz + g->zlodz) / 2.0 + p->dzc;  /* DHT.cpp:1075 */
    g->EE = tem_energy(p->wl);
    g->k2maxo = pow(p->aobj_mrad * (double)0.001f / p->wl, 2.0);  /* DHT.cpp:250 */
    g->cs3 = p->cs3_mm * 1.e+7; g->cs5 = p->cs5_mm * 1.e+7;
    return 0;
}

int ora_dht_nz(const ora_dht_params* p) {
    dht_geom g;
    if (dht_derive(p, &g)) return -1;
    return g.nz;
}

static double freq(int k, int n, double xst) {
    return (double)(k <= n / 2 ? k : k - n) / (n * xst);
}

/* 2D DFT of ny*nx interleaved complex data, natural order, exponent sign. */
static void fft2(double* a, int nx, int ny, int sign) {
    double* col = (double*)malloc(sizeof(double) * 2 * ny);
    for (int j = 0; j < ny; j++) ora_fft(a + 2 * (size_t)j * nx, nx, sign);
    for (int i = 0; i < nx; i++) {
        for (int j = 0; j < ny; j++) {
            col[2 * j] = a[2 * ((size_t)j * nx + i)];
            col[2 * j + 1] = a[2 * ((size_t)j * nx + i) + 1];
        }
        ora_fft(col, ny, sign);
        for (int j = 0; j < ny; j++) {
            a[2 * ((size_t)j * nx + i)] = col[2 * j];
            a[2 * ((size_t)j * nx + i) + 1] = col[2 * j + 1];
        }
    }
    free(col);
}

/* Output planes of one orientation from the 2D spectrum S of its field at
 * zout (DHT.cpp:957-1055): methods 1-3 back-propagate the complex amplitude
 * with the reverted aberrations (iSign = -1) and store 1 - |psi|^2; method 4
 * divides the spectrum of I - 1 by the weak-object CTF of each plane's
 * distance and stores the real part.  W: 2*ny*nx doubles of scratch. */
static void dht_planes_from_spectrum(const ora_dht_params* p, const dht_geom* g, const double* S,
                                     double zout, double weight, double* planes, double* W) {
    const int nx = g->nx, ny = g->ny;
    const size_t n2 = (size_t)nx * ny;
    const int np = p->escc ? g->nz : 1;
    const int back = p->method != 4;
    const double wl = p->wl;
    for (int q = 0; q < np; q++) {
        const int n = p->escc ? q : g->nz / 2;
        const double zn = g->zlodz + g->xst * n;  /* voutdefocus[n], DHT.cpp:388 */
        double d, c3, c5;
        if (back) { d = zn - zout; c3 = -g->cs3; c5 = -g->cs5; }  /* iSign = -1 */
        else { d = zout - zn; c3 = g->cs3; c5 = g->cs5; }
        for (int ky = 0; ky < ny; ky++) {
            const double fy = freq(ky, ny, g->xst);
            for (int kx = 0; kx < nx; kx++) {
                const double fx = freq(kx, nx, g->xst);
                const double q2 = fx * fx + fy * fy;
                const size_t t = (size_t)ky * nx + kx;
                double fr = 0, fi = 0;
                if (!(g->k2maxo > 0) || q2 < g->k2maxo) {
                    const double chi = DHT_PI * wl * d * q2 + DHT_PI * wl * wl * wl * c3 * q2 * q2 / 2.0 +
                                       DHT_PI * pow(wl, 5) * c5 * q2 * q2 * q2 / 3.0;
                    if (back) {
                        fr = cos(chi); fi = -sin(chi);
                    } else {
                        const double T = sin(chi) - p->asigma * cos(chi);
                        fr = T / (2.0 * (T * T + p->epsilon));
                    }
                }
                W[2 * t] = S[2 * t] * fr - S[2 * t + 1] * fi;
                W[2 * t + 1] = S[2 * t] * fi + S[2 * t + 1] * fr;
            }
        }
        fft2(W, nx, ny, +1);
        double* P = planes + (size_t)q * n2;
        const double inv = 1.0 / (double)n2;
        for (size_t t = 0; t < n2; t++) {
            const double re = W[2 * t] * inv, im = W[2 * t + 1] * inv;
            const double v = back ? 1.0 - (re * re + im * im) : re;
            P[t] = v * weight;
        }
    }
}

/* renormalise and clamp one input value (DHT.cpp:850-858), in double */
static double dht_intensity(const ora_dht_params* p, float v) {
    double x = v;
    if (p->norm_f1 != 1.0) x /= p->norm_f1;
    if (p->norm_f2 != 0.0) x += p->norm_f2;
    return x < 0 ? 0 : x;
}

int ora_dht_planes(const ora_dht_params* p, const float* img, double zout, double weight,
                   double* planes, int* nplanes) {
    dht_geom g;
    if (dht_derive(p, &g)) return -1;
    if (p->method != 3 && p->method != 4) return dfail("ora_dht_planes: method 3 or 4 only");
    const int nx = g.nx, ny = g.ny;
    const size_t n2 = (size_t)nx * ny;
    double* S = (double*)malloc(sizeof(double) * 2 * n2);
    double* W = (double*)malloc(sizeof(double) * 2 * n2);
    if (!S || !W) { free(S); free(W); return dfail("out of memory"); }
    for (size_t t = 0; t < n2; t++) {
        const double x = dht_intensity(p, img[t]);
        if (p->method == 3) {
            /* DHT.cpp:919-938: amplitude sqrt(I), phase Fact*(-0.5)*(I - 1) */
            const double amp = sqrt(x), phi = p->epsilon * (-0.5) * (x - 1.0);
            S[2 * t] = amp * cos(phi);
            S[2 * t + 1] = amp * sin(phi);
        } else {
            S[2 * t] = x - 1.0;
            S[2 * t + 1] = 0.0;
        }
    }
    fft2(S, nx, ny, -1);
    dht_planes_from_spectrum(p, &g, S, zout, weight, planes, W);
    free(S);
    free(W);
    *nplanes = p->escc ? g.nz : 1;
    return 0;
}

int ora_dht_planes_wave(const ora_dht_params* p, const double* psi, double zout, double weight,
                        double* planes, int* nplanes) {
    dht_geom g;
    if (dht_derive(p, &g)) return -1;
    if (p->method < 1 || p->method > 3) return dfail("ora_dht_planes_wave: methods 1, 2 or 3 only");
    const size_t n2 = (size_t)g.nx * g.ny;
    double* S = (double*)malloc(sizeof(double) * 2 * n2);
    double* W = (double*)malloc(sizeof(double) * 2 * n2);
    if (!S || !W) { free(S); free(W); return dfail("out of memory"); }
    memcpy(S, psi, sizeof(double) * 2 * n2);
    fft2(S, g.nx, g.ny, -1);
    dht_planes_from_spectrum(p, &g, S, zout, weight, planes, W);
    free(S);
    free(W);
    *nplanes = p->escc ? g.nz : 1;
    return 0;
}

/* S <- S * exp(-i pi wl d q^2), zero outside the objective aperture: free-
 * space propagation by d between two planes imaged through the same
 * objective (its Cs phases cancel between the planes). */
static void dht_fresnel_rel(const dht_geom* g, double wl, double d, double* S) {
    for (int ky = 0; ky < g->ny; ky++) {
        const double fy = freq(ky, g->ny, g->xst);
        for (int kx = 0; kx < g->nx; kx++) {
            const double fx = freq(kx, g->nx, g->xst), q2 = fx * fx + fy * fy;
            double* v = S + 2 * ((size_t)ky * g->nx + kx);
            if (g->k2maxo > 0 && !(q2 < g->k2maxo)) {
                v[0] = v[1] = 0;
                continue;
            }
            const double chi = DHT_PI * wl * d * q2, c = cos(chi), s = -sin(chi);
            const double re = v[0] * c - v[1] * s, im = v[0] * s + v[1] * c;
            v[0] = re;
            v[1] = im;
        }
    }
}

/* IWFR (Allen, McBride, O'Leary and Oxley, Ultramicroscopy 100 (2004)
 * 91-104), the call of DHT.cpp:889-901: psi_0 is the mean of the
 * amplitudes sqrt(I_n) propagated from their planes z_n to zout = mean z_n;
 * iteration k propagates psi_{k-1} to every z_n, replaces the modulus by
 * sqrt(I_n) (keeping the phase; sqrt(I_n) where the modulus is 0) and
 * averages the results propagated back to zout.  It stops after itermax
 * iterations or when ||psi_k - psi_{k-1}||^2 < epsilon ||psi_{k-1}||^2. */
static int dht_iwfr(const ora_dht_params* p, const dht_geom* g, int cnt, const float* imgs,
                    const double* z, double zout, double* Psi, int* iters) {
    const size_t n2 = (size_t)g->nx * g->ny;
    double* A = (double*)malloc(sizeof(double) * n2 * cnt);
    double* F = (double*)malloc(sizeof(double) * 2 * n2);
    double* New = (double*)malloc(sizeof(double) * 2 * n2);
    if (!A || !F || !New) { free(A); free(F); free(New); return dfail("out of memory"); }
    for (int n = 0; n < cnt; n++)
        for (size_t t = 0; t < n2; t++) A[n * n2 + t] = sqrt(dht_intensity(p, imgs[n * n2 + t]));
    memset(Psi, 0, sizeof(double) * 2 * n2);
    for (int n = 0; n < cnt; n++) {
        for (size_t t = 0; t < n2; t++) { F[2 * t] = A[n * n2 + t]; F[2 * t + 1] = 0; }
        fft2(F, g->nx, g->ny, -1);
        dht_fresnel_rel(g, p->wl, zout - z[n], F);
        for (size_t t = 0; t < 2 * n2; t++) Psi[t] += F[t];
    }
    for (size_t t = 0; t < 2 * n2; t++) Psi[t] /= cnt;
    const double inv = 1.0 / (double)n2;
    int k = 0;
    while (k < p->itermax) {
        k++;
        memset(New, 0, sizeof(double) * 2 * n2);
        for (int n = 0; n < cnt; n++) {
            memcpy(F, Psi, sizeof(double) * 2 * n2);
            dht_fresnel_rel(g, p->wl, z[n] - zout, F);
            fft2(F, g->nx, g->ny, +1);
            for (size_t t = 0; t < n2; t++) {
                const double re = F[2 * t] * inv, im = F[2 * t + 1] * inv, m = sqrt(re * re + im * im);
                const double a = A[n * n2 + t];
                if (m > 0) { F[2 * t] = a * re / m; F[2 * t + 1] = a * im / m; }
                else { F[2 * t] = a; F[2 * t + 1] = 0; }
            }
            fft2(F, g->nx, g->ny, -1);
            dht_fresnel_rel(g, p->wl, zout - z[n], F);
            for (size_t t = 0; t < 2 * n2; t++) New[t] += F[t];
        }
        double num = 0, den = 0;
        for (size_t t = 0; t < 2 * n2; t++) {
            New[t] /= cnt;
            num += (New[t] - Psi[t]) * (New[t] - Psi[t]);
            den += Psi[t] * Psi[t];
        }
        memcpy(Psi, New, sizeof(double) * 2 * n2);
        if (num / den < p->epsilon) break;
    }
    *iters = k;
    free(A); free(F); free(New);
    return 0;
}

/* CTFL2, the call of DHT.cpp:903-917: the phase at the first plane z_0 from
 * the Tikhonov-regularised least-squares solution of the weak phase-
 * amplitude object model over the series (Paganin, Barty, McMahon and
 * Nugent, J. Microsc. 214 (2004) 51-61, with several distances): a wave
 * 1 - mu + i phi at z_0 imaged at z_n has (I_n - 1)^ = 2 s_n phi^ - 2 c_n mu^,
 * s_n, c_n = sin, cos of chi_n = pi wl (z_n - z_0) q^2, so
 *   phi^ = (Sum c^2 Sum s I^ - Sum s c Sum c I^) / (2 (Sum s^2 Sum c^2 - (Sum s c)^2) + epsilon),
 * zero outside the objective aperture.  Output sqrt(I_0) exp(i phi). */
static int dht_ctfl2(const ora_dht_params* p, const dht_geom* g, int cnt, const float* imgs,
                     const double* z, double* Psi) {
    const size_t n2 = (size_t)g->nx * g->ny;
    double* C = (double*)malloc(sizeof(double) * 2 * n2 * cnt);
    if (!C) return dfail("out of memory");
    for (int n = 0; n < cnt; n++) {
        double* c = C + 2 * n2 * n;
        for (size_t t = 0; t < n2; t++) { c[2 * t] = dht_intensity(p, imgs[n * n2 + t]) - 1.0; c[2 * t + 1] = 0; }
        fft2(c, g->nx, g->ny, -1);
    }
    for (int ky = 0; ky < g->ny; ky++) {
        const double fy = freq(ky, g->ny, g->xst);
        for (int kx = 0; kx < g->nx; kx++) {
            const double fx = freq(kx, g->nx, g->xst), q2 = fx * fx + fy * fy;
            const size_t t = (size_t)ky * g->nx + kx;
            double ss = 0, cc = 0, sc = 0, sir = 0, sii = 0, cir = 0, cii = 0;
            if (!(g->k2maxo > 0) || q2 < g->k2maxo)
                for (int n = 0; n < cnt; n++) {
                    const double chi = DHT_PI * p->wl * (z[n] - z[0]) * q2, s = sin(chi), c = cos(chi);
                    const double* v = C + 2 * n2 * n + 2 * t;
                    ss += s * s; cc += c * c; sc += s * c;
                    sir += s * v[0]; sii += s * v[1]; cir += c * v[0]; cii += c * v[1];
                }
            const double den = 2.0 * (ss * cc - sc * sc) + p->epsilon;
            Psi[2 * t] = (cc * sir - sc * cir) / den;
            Psi[2 * t + 1] = (cc * sii - sc * cii) / den;
        }
    }
    fft2(Psi, g->nx, g->ny, +1);
    const double inv = 1.0 / (double)n2;
    for (size_t t = 0; t < n2; t++) {
        const double phi = Psi[2 * t] * inv, amp = sqrt(dht_intensity(p, imgs[t]));
        Psi[2 * t] = amp * cos(phi);
        Psi[2 * t + 1] = amp * sin(phi);
    }
    free(C);
    return 0;
}

int ora_dht_retrieve(const ora_dht_params* p, int cnt, const float* imgs, const double* defocus,
                     double* psi, double* zout, int* iters) {
    dht_geom g;
    if (dht_derive(p, &g)) return -1;
    if (p->method != 1 && p->method != 2) return dfail("ora_dht_retrieve: method 1 or 2 only");
    if (cnt < 2)
        return dfail("Error: selected phase retrieval method is incompatibe with a single defocus distance appearing in some lines of the defocus/orientations parameter file.");
    *iters = 0;
    if (p->method == 1) {
        double zo = 0;
        for (int n = 0; n < cnt; n++) zo += defocus[n];
        zo /= (double)cnt;  /* DHT.cpp:895-897 */
        *zout = zo;
        if (dht_iwfr(p, &g, cnt, imgs, defocus, zo, psi, iters)) return -1;
        fft2(psi, g.nx, g.ny, +1);
        const double inv = 1.0 / ((double)g.nx * g.ny);
        for (size_t t = 0; t < 2 * (size_t)g.nx * g.ny; t++) psi[t] *= inv;
        return 0;
    }
    *zout = defocus[0];  /* DHT.cpp:908 */
    return dht_ctfl2(p, &g, cnt, imgs, defocus, psi);
}

int ora_dht_splat(const ora_dht_params* p, const double* planes, int nplanes, double angZ_deg,
                  double angY_deg, double* K3out) {
    dht_geom g;
    if (dht_derive(p, &g)) return -1;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const double xst = g.xst, yst = g.xst, zst = g.xst;
    const double angleZ = angZ_deg * DHT_PI180, angleY = angY_deg * DHT_PI180;
    const double cosY = cos(angleY), sinY = sin(angleY), cosZ = cos(angleZ), sinZ = sin(angleZ);
    double* xs = (double*)malloc(sizeof(double) * 2 * nx);
    double* ys = (double*)malloc(sizeof(double) * 2 * ny);
    double* zs = (double*)malloc(sizeof(double) * 2 * nz);
    /* DHT.cpp:1083-1107 */
    for (int i = 0; i < nx; i++) {
        double xxx = g.xlo + xst * i - g.xc;
        xs[2 * i] = xxx * sinY;
        xs[2 * i + 1] = xxx * cosY;
    }
    for (int j = 0; j < ny; j++) {
        double yyy = g.ylo + yst * j - g.yc;
        ys[2 * j] = yyy * sinZ;
        ys[2 * j + 1] = yyy * cosZ;
    }
    for (int n = 0; n < nz; n++) {
        double zzz = (g.zlodz + zst * n) - g.zc;
        zs[2 * n] = zzz * sinY;
        zs[2 * n + 1] = zzz * cosY;
    }
    const int nx2 = nx - 2, ny2 = ny - 2, nz2 = nz - 2;
    const size_t nxny = (size_t)nx * ny;
    /* DHT.cpp:1115-1153 */
    for (int n = 0; n < nz; n++) {
        const double* K = planes + (size_t)(nplanes == 1 ? 0 : n) * nxny;
        for (int i = 0; i < nx; i++) {
            double zzz = g.zc + zs[2 * n + 1] + xs[2 * i];
            double dz1 = (zzz - g.zlodz) / zst;
            int nn = (int)dz1;
            if (nn < 0 || nn > nz2) continue;
            dz1 -= nn;
            double dz0 = 1.0 - dz1;
            double xx = -zs[2 * n] + xs[2 * i + 1];
            double xx_sinZ = xx * sinZ, xx_cosZ = xx * cosZ;
            for (int j = 0; j < ny; j++) {
                double xxx = g.xc + ys[2 * j] + xx_cosZ;
                double dx1 = (xxx - g.xlo) / xst;
                int ii = (int)dx1;
                if (ii < 0 || ii > nx2) continue;
                dx1 -= ii;
                double dx0 = 1.0 - dx1;
                double yyy = g.yc + ys[2 * j + 1] - xx_sinZ;
                double dy1 = (yyy - g.ylo) / yst;
                int jj = (int)dy1;
                if (jj < 0 || jj > ny2) continue;
                dy1 -= jj;
                double dy0 = 1.0 - dy1;
                const double k = K[(size_t)j * nx + i];
                double dz0K = dz0 * k, dz1K = dz1 * k;
                double* o = K3out + (size_t)nn * nxny + (size_t)jj * nx + ii;
                o[0] += dx0 * dy0 * dz0K;
                o[1] += dx1 * dy0 * dz0K;
                o[nxny] += dx0 * dy0 * dz1K;
                o[nxny + 1] += dx1 * dy0 * dz1K;
                o[nx] += dx0 * dy1 * dz0K;
                o[nx + 1] += dx1 * dy1 * dz0K;
                o[nxny + nx] += dx0 * dy1 * dz1K;
                o[nxny + nx + 1] += dx1 * dy1 * dz1K;
            }
        }
    }
    free(xs); free(ys); free(zs);
    return 0;
}

static int dht_sctf(const ora_dht_params* p, const dht_geom* dg, int nrows, const double* angZ,
                    const double* angY, const double* defocus, const float* weights,
                    const float* images, double* vol) {
    /* sCTF: the UTR Ewald-sphere machinery with the DHT filter and scale */
    ora_params q;
    memset(&q, 0, sizeof q);
    q.nx = p->nx; q.ny = p->ny; q.modality = 0; q.xst = p->xst;
    q.zlo = dg->zlo; q.zhi = dg->zhi; q.dzextra = 0;
    q.dxc = p->dxc; q.dyc = p->dyc; q.dzc = p->dzc; q.wl = p->wl;
    q.aobj_mrad = p->aobj_mrad; q.cs3_mm = p->cs3_mm; q.cs5_mm = p->cs5_mm;
    q.ctf_mode = 3; q.escc = p->escc; q.asigma = 0; q.nsr = 0; q.epsilon = 1;
    q.inv_laplace = 0; q.norm_f1 = p->norm_f1; q.norm_f2 = p->norm_f2;
    ora_geom g;
    if (ora_derive(&q, nrows, angZ, angY, NULL, defocus, &g)) return -1;
    g.asigma = p->asigma;
    g.epsilon = p->epsilon;
    const size_t N = (size_t)g.nx * g.ny * g.nz, n2 = (size_t)g.nx * g.ny;
    double* V = (double*)calloc(2 * N, sizeof(double));
    double* S = (double*)calloc(N, sizeof(double));
    double* K = (double*)malloc(sizeof(double) * 2 * n2);
    if (!V || !S || !K) { free(V); free(S); free(K); return dfail("out of memory"); }
    for (int r = 0; r < nrows; r++) {
        /* Log0 + InvertCTF_DT1 (DHT.cpp:1056-1067), weight (DHT.cpp:1070) */
        ora_spectrum(&g, images + (size_t)r * n2, defocus[r], weights ? (double)weights[r] : 1.0, K);
        ora_grid(&g, K, angZ[r], angY[r], V, S);  /* DHT.cpp:1195-1308 */
    }
    /* DHT.cpp:1350-1365 */
    for (size_t v = 0; v < N; v++) {
        double x = S[v];
        if (x == 0) continue;
        x = (1 / x) * exp((x - p->critical_snr) / (x * x * x * x));
        V[2 * v] *= x;
        V[2 * v + 1] *= x;
    }
    ora_hermitian(&g, V);  /* DHT.cpp:1378-1391, the same loop as UTR.cpp:1203-1217 */
    if (p->inv_laplace) {
        /* InverseMLaplacian1(zhi - zlo, yhi - ylo, xhi - ylo, alpha, alpha), DHT.cpp:1402 */
        const double fact = 4.0 * DHT_PI * DHT_PI;
        const double xaper = g.xhi - g.ylo, yaper = g.yhi - g.ylo, zaper = g.zhi - g.zlo;
        for (int k = 0; k < g.nz; k++)
            for (int j = 0; j < g.ny; j++)
                for (int i = 0; i < g.nx; i++) {
                    const int i1 = i - g.nxd2, j1 = j - g.nyd2, k1 = k - g.nzd2;
                    const double ksi2 = fact / (xaper * xaper) * i1 * i1 +
                                        (j1 * j1 * (fact / (yaper * yaper)) +
                                         (k1 * k1 * (fact / (zaper * zaper)) + p->alpha));
                    const double f = p->alpha / ksi2;
                    const size_t t = ((size_t)k * g.ny + j) * g.nx + i;
                    V[2 * t] *= f;
                    V[2 * t + 1] *= f;
                }
    }
    ora_c2r(&g, V, vol);
    const double dnorm = p->wl * dg->EE / (2.0 * DHT_PI * sqrt(1.0 + p->asigma * p->asigma));
    for (size_t v = 0; v < N; v++) vol[v] *= dnorm;
    free(V); free(S); free(K);
    return 0;
}

int ora_dht_reconstruct_series(const ora_dht_params* p, int nangles, const double* angZ_deg,
                               const double* angY_deg, const int* counts, const double* defocus,
                               const float* weights, const float* images, double* vol, int nthreads) {
    dht_geom g;
    if (dht_derive(p, &g)) return -1;
    if (nangles < 1) return dfail("no orientation rows");
    /* DHT.cpp:272-283 */
    int cmin = 1, cmax = 1;
    if (counts) {
        cmin = cmax = counts[0];
        for (int a = 1; a < nangles; a++) {
            if (counts[a] < cmin) cmin = counts[a];
            if (counts[a] > cmax) cmax = counts[a];
        }
    }
    if ((p->method == 1 || p->method == 2) && cmin < 2)
        return dfail("Error: selected phase retrieval method is incompatibe with a single defocus distance appearing in some lines of the defocus/orientations parameter file.");
    if (p->method >= 3 && cmax != 1)
        return dfail("Error: selected phase retrieval method is incompatibe with multiple defocus distances in some lines of the defocus/orientations parameter file.");
    if (p->method == 5)
        return dht_sctf(p, &g, nangles, angZ_deg, angY_deg, defocus, weights, images, vol);
    const size_t n2 = (size_t)g.nx * g.ny, N = n2 * g.nz;
    int* first = (int*)malloc(sizeof(int) * nangles);
    if (!first) return dfail("out of memory");
    for (int a = 0, f = 0; a < nangles; a++) { first[a] = f; f += counts ? counts[a] : 1; }
    if (nthreads < 1) nthreads = 1;
    memset(vol, 0, sizeof(double) * N);
    int err = 0;
#pragma omp parallel num_threads(nthreads)
    {
        double* acc = vol;
#ifdef _OPENMP
        if (omp_get_thread_num() > 0) acc = (double*)calloc(N, sizeof(double));
#endif
        double* planes = (double*)malloc(sizeof(double) * N);
        double* psi = (double*)malloc(sizeof(double) * 2 * n2);
#pragma omp for schedule(static)
        for (int a = 0; a < nangles; a++) {
            int np = 0, it = 0;
            const double w = weights ? (double)weights[a] : 1.0;
            const int f = first[a], cnt = counts ? counts[a] : 1;
            double zout;
            if (p->method <= 2) {
                if (ora_dht_retrieve(p, cnt, images + (size_t)f * n2, defocus + f, psi, &zout, &it) ||
                    ora_dht_planes_wave(p, psi, zout, w, planes, &np))
                    err = 1;
            } else if (ora_dht_planes(p, images + (size_t)f * n2, defocus[f], w, planes, &np)) {
                err = 1;
            }
            if (!err && ora_dht_splat(p, planes, np, angZ_deg[a], angY_deg[a], acc)) err = 1;
        }
        free(planes);
        free(psi);
#pragma omp barrier
#ifdef _OPENMP
        if (omp_get_thread_num() > 0) {
#pragma omp critical
            for (size_t v = 0; v < N; v++) vol[v] += acc[v];
            free(acc);
        }
#endif
    }
    free(first);
    if (err) return -1;
    /* DHT.cpp:1325-1340 */
    double total = 0;
    if (weights) for (int a = 0; a < nangles; a++) total += weights[a];
    else total = nangles;
    const double dAtomWidth = 1.0, dGamma = 0.1;
    double dnorm;
    if (p->method != 4) dnorm = g.EE * dGamma / dAtomWidth * (-p->wl / DHT_PI) / total;
    else dnorm = p->wl * g.EE / DHT_PI / (dAtomWidth * 8.0) / total;
    for (size_t v = 0; v < N; v++) vol[v] *= dnorm;
    return 0;
}

int ora_dht_reconstruct(const ora_dht_params* p, int nrows, const double* angZ_deg,
                        const double* angY_deg, const double* defocus, const float* weights,
                        const float* images, double* vol, int nthreads) {
    return ora_dht_reconstruct_series(p, nrows, angZ_deg, angY_deg, NULL, defocus, weights, images,
                                      vol, nthreads);
}

/* the oracle shares one error slot with utr_oracle.c */
int ora_set_error(const char* m);
static int dfail(const char* m) { return ora_set_error(m); }
