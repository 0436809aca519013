// utr_cpu32.cpp -- CPU BASELINE AND TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.
//
// FP32 C++ restatement of the reference's CPU path for the UTR reconstruction
// in CTF mode 0 (MultisliceCppTest/PhaseRetrieval/UTR.cpp:840-1279,
// UTR.h:188-307), in the reference's own structure and precision:
//   - an OpenMP loop over the orientation rows (UTR.cpp:877); each thread
//     takes a row's image through renormalisation, clamp, Log0, centring
//     signs and a 2D FFT of its own (UTR.cpp:851, 951-993), then grids the
//     spectrum with the reference's float coordinate tables and float
//     arithmetic (CT_3Dgridding<float>, UTR.h:188-307).  The reference adds
//     into the shared V3/Samp3 from every thread without synchronisation;
//     here each thread owns accumulators, summed after the loop;
//   - sampling normalisation and noise gate in float, parallel over planes
//     (UTR.cpp:1050-1091);
//   - Hermitian symmetrisation (UTR.cpp:1203-1217);
//   - a threaded 3D complex-to-real inverse FFT of the half spectrum
//     (Fftwd3frc with nThreads, UTR.cpp:1219-1233): inverse transforms
//     along z and y of the nx/2+1 kept columns, then one c2r per x line;
//   - sphere mask, cupping compensation and scale (UTR.cpp:1235-1279).
// FFTW is not installed in this environment, so the transforms are this
// file's radix-2 FFT in float (power-of-two sizes).  The bench reports which
// FFT ran.  Geometry comes from ora_derive (utr_oracle.c), the same
// derivation the FP64 oracle uses.  Only the cpu_baseline leg of bench.py and
// tests/ load this library.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <string>
#include <vector>
#include <omp.h>

#include "utr_oracle.h"

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr float kTargetSnr = 1.414213562373f;  // UTR.h:12

std::string g_err;

int fail(const std::string& m) {
    g_err = m;
    return -1;
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// In-place radix-2 FFT of n = 2^k interleaved float pairs, unnormalised,
// exponent sign * 2 pi i k m / n.
struct Fft {
    int n = 0;
    std::vector<int> rev;
    std::vector<float> tw;  // e^{-2 pi i k / n}, k < n/2

    explicit Fft(int len) : n(len), rev(len), tw(len) {
        int lg = 0;
        while ((1 << lg) < n) lg++;
        for (int i = 0; i < n; i++) {
            int r = 0;
            for (int b = 0; b < lg; b++) r |= ((i >> b) & 1) << (lg - 1 - b);
            rev[i] = r;
        }
        for (int k = 0; k < n / 2; k++) {
            tw[2 * k] = (float)std::cos(2.0 * kPi * k / n);
            tw[2 * k + 1] = (float)-std::sin(2.0 * kPi * k / n);
        }
    }

    void exec(float* x, int sign) const {
        for (int i = 0; i < n; i++) {
            const int r = rev[i];
            if (r > i) {
                std::swap(x[2 * i], x[2 * r]);
                std::swap(x[2 * i + 1], x[2 * r + 1]);
            }
        }
        const float sg = sign < 0 ? 1.f : -1.f;  // conjugate the table for +1
        for (int len = 2; len <= n; len <<= 1) {
            const int half = len >> 1, step = n / len;
            for (int s = 0; s < n; s += len) {
                for (int k = 0; k < half; k++) {
                    const float wr = tw[2 * k * step], wi = sg * tw[2 * k * step + 1];
                    float* a = x + 2 * (s + k);
                    float* b = x + 2 * (s + k + half);
                    const float tr = b[0] * wr - b[1] * wi, ti = b[0] * wi + b[1] * wr;
                    b[0] = a[0] - tr;
                    b[1] = a[1] - ti;
                    a[0] += tr;
                    a[1] += ti;
                }
            }
        }
    }
};

bool pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

// One row's 2D stage (UTR.cpp:951-993, 1019) into K (ny*nx float pairs,
// centred spectrum).
void row_spectrum(const ora_geom& g, const float* img, float weight, const Fft& fx, const Fft& fy,
                  float* K, float* col) {
    const int nx = g.nx, ny = g.ny;
    const float f1 = (float)g.norm_f1, f2 = (float)g.norm_f2;
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            float x = img[(size_t)j * nx + i];
            if (g.norm_f1 != 1.0) x = x / f1;
            if (g.norm_f2 != 0.0) x = x + f2;
            if (x < 0.f) x = 0.f;
            const float v = x > 0.f ? std::log(x) : 0.f;
            K[2 * ((size_t)j * nx + i)] = ((i + j) & 1) ? -v : v;
            K[2 * ((size_t)j * nx + i) + 1] = 0.f;
        }
    for (int j = 0; j < ny; j++) fx.exec(K + 2 * (size_t)j * nx, -1);
    for (int i = 0; i < nx; i++) {
        for (int j = 0; j < ny; j++) {
            col[2 * j] = K[2 * ((size_t)j * nx + i)];
            col[2 * j + 1] = K[2 * ((size_t)j * nx + i) + 1];
        }
        fy.exec(col, -1);
        for (int j = 0; j < ny; j++) {
            K[2 * ((size_t)j * nx + i)] = col[2 * j];
            K[2 * ((size_t)j * nx + i) + 1] = col[2 * j + 1];
        }
    }
    const float gs = ((g.nxd2 + g.nyd2) & 1) ? -1.f : 1.f;
    for (int ky = 0; ky < ny; ky++)
        for (int kx = 0; kx < nx; kx++) {
            const float s = (((kx + ky) & 1) ? -gs : gs) * weight;
            K[2 * ((size_t)ky * nx + kx)] *= s;
            K[2 * ((size_t)ky * nx + kx) + 1] *= s;
        }
}

// The float gridding of UTR.h:188-307 (no centre shift): per-row coordinate
// tables in float, the Ewald height, the inverse rotation about Y' then Z,
// floor and bounds per axis, eight trilinear corners into V (float pairs)
// and S.
void grid_row(const ora_geom& g, const float* K, double angZ_deg, double angY_deg, float* V, float* S,
              std::vector<float>& tab) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int nx2 = nx - 2, ny2 = ny - 2, nz2 = nz - 2;
    const size_t nxny = (size_t)nx * ny;
    const double wl2 = 0.5 * g.wl;
    const double aY = angY_deg * kPi / 180.0, aZ = angZ_deg * kPi / 180.0;
    const float sinY = (float)std::sin(aY), cosY = (float)std::cos(aY);
    const float sinZ = (float)std::sin(aZ), cosZ = (float)std::cos(aZ);
    tab.resize(3 * (size_t)nx + 3 * (size_t)ny);
    float* xq2 = tab.data();
    float* xsY = xq2 + nx;
    float* xcY = xsY + nx;
    float* yq2 = xcY + nx;
    float* ysZ = yq2 + ny;
    float* ycZ = ysZ + ny;
    for (int i = 0; i < nx; i++) {
        const double q = g.fxlo + g.fxst * i;
        xq2[i] = (float)(-wl2 * q * q);
        xsY[i] = (float)(q * std::sin(aY));
        xcY[i] = (float)(q * std::cos(aY));
    }
    for (int j = 0; j < ny; j++) {
        const double q = g.fylo + g.fyst * j;
        yq2[j] = (float)(-wl2 * q * q);
        ysZ[j] = (float)(q * std::sin(aZ));
        ycZ[j] = (float)(q * std::cos(aZ));
    }
    const float fxlo = (float)g.fxlo, fylo = (float)g.fylo, fzlo = (float)g.fzlo;
    const float fxst = (float)g.fxst, fyst = (float)g.fyst, fzst = (float)g.fzst;
    for (int j = 0; j < ny; j++) {
        for (int i = 0; i < nx; i++) {
            const float zz = g.escc ? xq2[i] + yq2[j] : 0.f;
            const float xx = xcY[i] - zz * sinY;
            const float zzz = xsY[i] + zz * cosY;
            float dz1 = (zzz - fzlo) / fzst;
            const int nn = (int)std::floor(dz1);
            if (nn < 0 || nn > nz2) continue;
            const float xxx = xx * cosZ + ysZ[j];
            float dx1 = (xxx - fxlo) / fxst;
            const int ii = (int)std::floor(dx1);
            if (ii < 0 || ii > nx2) continue;
            const float yyy = -xx * sinZ + ycZ[j];
            float dy1 = (yyy - fylo) / fyst;
            const int jj = (int)std::floor(dy1);
            if (jj < 0 || jj > ny2) continue;
            dx1 -= ii;
            dy1 -= jj;
            dz1 -= nn;
            const float wx[2] = {1.f - dx1, dx1}, wy[2] = {1.f - dy1, dy1}, wz[2] = {1.f - dz1, dz1};
            const float cr = K[2 * ((size_t)j * nx + i)], ci = K[2 * ((size_t)j * nx + i) + 1];
            const size_t base = nn * nxny + (size_t)jj * nx + ii;
            for (int cz = 0; cz < 2; cz++)
                for (int cy = 0; cy < 2; cy++)
                    for (int cx = 0; cx < 2; cx++) {
                        const float w = wx[cx] * wy[cy] * wz[cz];
                        const size_t t = base + cz * nxny + (size_t)cy * nx + cx;
                        S[t] += w;
                        V[2 * t] += w * cr;
                        V[2 * t + 1] += w * ci;
                    }
        }
    }
}

}  // namespace

extern "C" {

const char* cpu32_last_error(void) { return g_err.c_str(); }

const char* cpu32_fft_name(void) { return "own radix-2 FP32 FFT (FFTW not installed)"; }

// Whole reconstruction (CTF mode 0, no centre shift, no Z''/shifts).  vol:
// nz*ny*nx floats.  t4 (may be NULL): ms of [0] the row loop (2D stage and
// gridding), [1] accumulator sum, normalisation and Hermitian symmetry,
// [2] the 3D inverse FFT, [3] total.
int cpu32_reconstruct(const ora_params* p, int nrows, const double* angZ_deg, const double* angY_deg,
                      const double* defocus, const float* weights, const float* images, float* vol,
                      int nthreads, double* t4) {
    ora_geom g;
    if (ora_derive(p, nrows, angZ_deg, angY_deg, nullptr, defocus, &g)) return fail(ora_last_error());
    if (g.ctf_mode != 0 || g.inv_laplace) return fail("the FP32 port covers CTF mode 0 without the inverse Laplacian");
    if (g.rot_centre_shift) return fail("the FP32 port has no centre-of-rotation shift");
    if (!pow2(g.nx) || !pow2(g.ny) || !pow2(g.nz)) return fail("the FP32 port's FFT needs power-of-two sizes");
    if (nthreads < 1) nthreads = omp_get_max_threads();
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const size_t n2 = (size_t)nx * ny, N = n2 * nz;
    const double t0 = now_ms();
    std::vector<float*> Vp(nthreads, nullptr), Sp(nthreads, nullptr);
    bool oom = false;
    for (int t = 0; t < nthreads; t++) {
        Vp[t] = (float*)std::calloc(2 * N, sizeof(float));
        Sp[t] = (float*)std::calloc(N, sizeof(float));
        if (!Vp[t] || !Sp[t]) oom = true;
    }
    auto release = [&]() {
        for (int t = 0; t < nthreads; t++) {
            std::free(Vp[t]);
            std::free(Sp[t]);
        }
    };
    if (oom) {
        release();
        return fail("out of memory for the per-thread accumulators");
    }
    const Fft fx(nx), fy(ny), fz(nz);
    // rows in parallel, one set of accumulators per thread (UTR.cpp:877-1027)
#pragma omp parallel num_threads(nthreads)
    {
        const int t = omp_get_thread_num();
        std::vector<float> K(2 * n2), col(2 * (size_t)std::max(nx, std::max(ny, nz))), tab;
#pragma omp for schedule(dynamic, 1)
        for (int r = 0; r < nrows; r++) {
            row_spectrum(g, images + (size_t)r * n2, weights ? weights[r] : 1.f, fx, fy, K.data(), col.data());
            grid_row(g, K.data(), angZ_deg[r], angY_deg[r], Vp[t], Sp[t], tab);
        }
    }
    const double t1 = now_ms();
    float* V = Vp[0];
    float* S = Sp[0];
    // accumulator sum and the normalisation with the noise gate (UTR.cpp:1050-1091)
    const float nsr2 = (float)(g.nsr * std::sqrt((double)n2));
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (long v = 0; v < (long)N; v++) {
        float vr = V[2 * v], vi = V[2 * v + 1], s = S[v];
        for (int q = 1; q < nthreads; q++) {
            vr += Vp[q][2 * v];
            vi += Vp[q][2 * v + 1];
            s += Sp[q][v];
        }
        float f = 0.f;
        if (s != 0.f) {
            if (nsr2 == 0.f) {
                f = 1.f / s;
            } else {
                const float y = std::sqrt(vr * vr + vi * vi) / (nsr2 * std::sqrt(s));
                f = (1.f / s) * (0.5f + std::atan(100.f * (y - kTargetSnr)) / (float)kPi);
            }
        }
        V[2 * v] = vr * f;
        V[2 * v + 1] = vi * f;
    }
    for (int q = 1; q < nthreads; q++) {
        std::free(Vp[q]);
        std::free(Sp[q]);
        Vp[q] = Sp[q] = nullptr;
    }
    // Hermitian symmetrisation in the reference's loop order (UTR.cpp:1203-1217)
    for (int k = 0; k < nz; k++) {
        const int ka = k == 0 ? 0 : nz - k;
        for (int j = 0; j < ny; j++) {
            const int ja = j == 0 ? 0 : ny - j;
            for (int i = 0; i < nx / 2; i++) {
                const int ia = i == 0 ? 0 : nx - i;
                const size_t d = ((size_t)k * ny + j) * nx + ia, s = ((size_t)ka * ny + ja) * nx + i;
                V[2 * d] = 0.5f * (V[2 * d] + V[2 * s]);
                V[2 * d + 1] = 0.5f * (V[2 * d + 1] - V[2 * s + 1]);
            }
        }
    }
    const double t2 = now_ms();
    // 3D c2r of the un-centred half spectrum (x frequencies 0..nx/2)
    const int hx = nx / 2 + 1;
    std::vector<float> H(2 * (size_t)hx * ny * nz);
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int wz = 0; wz < nz; wz++)
        for (int wy = 0; wy < ny; wy++)
            for (int wx = 0; wx < hx; wx++) {
                const int cz = (wz + g.nzd2) % nz, cy = (wy + g.nyd2) % ny, cx = (wx + g.nxd2) % nx;
                const size_t s = ((size_t)cz * ny + cy) * nx + cx, d = ((size_t)wz * ny + wy) * hx + wx;
                H[2 * d] = V[2 * s];
                H[2 * d + 1] = V[2 * s + 1];
            }
    std::free(Vp[0]);
    std::free(Sp[0]);
    Vp[0] = Sp[0] = nullptr;
#pragma omp parallel num_threads(nthreads)
    {
        std::vector<float> line(2 * (size_t)std::max(nx, std::max(ny, nz)));
        // along z, for every (y, kept x)
#pragma omp for schedule(static)
        for (long l = 0; l < (long)ny * hx; l++) {
            const size_t off = (size_t)l;  // (wy, wx) in the plane
            for (int z = 0; z < nz; z++) {
                line[2 * z] = H[2 * ((size_t)z * ny * hx + off)];
                line[2 * z + 1] = H[2 * ((size_t)z * ny * hx + off) + 1];
            }
            fz.exec(line.data(), +1);
            for (int z = 0; z < nz; z++) {
                H[2 * ((size_t)z * ny * hx + off)] = line[2 * z];
                H[2 * ((size_t)z * ny * hx + off) + 1] = line[2 * z + 1];
            }
        }
        // along y, for every (z, kept x)
#pragma omp for schedule(static)
        for (long l = 0; l < (long)nz * hx; l++) {
            const size_t z = (size_t)(l / hx), x = (size_t)(l % hx);
            for (int y = 0; y < ny; y++) {
                line[2 * y] = H[2 * ((z * ny + y) * hx + x)];
                line[2 * y + 1] = H[2 * ((z * ny + y) * hx + x) + 1];
            }
            fy.exec(line.data(), +1);
            for (int y = 0; y < ny; y++) {
                H[2 * ((z * ny + y) * hx + x)] = line[2 * y];
                H[2 * ((z * ny + y) * hx + x) + 1] = line[2 * y + 1];
            }
        }
        // c2r along x per (z, y) line, normalised, re-centred into vol
        const float inv = (float)(1.0 / (double)N);
#pragma omp for schedule(static)
        for (long l = 0; l < (long)nz * ny; l++) {
            const int wz = (int)(l / ny), wy = (int)(l % ny);
            const float* h = H.data() + 2 * (size_t)l * hx;
            for (int x = 0; x < nx; x++) {
                if (x < hx) {
                    line[2 * x] = h[2 * x];
                    line[2 * x + 1] = h[2 * x + 1];
                } else {
                    line[2 * x] = h[2 * (nx - x)];
                    line[2 * x + 1] = -h[2 * (nx - x) + 1];
                }
            }
            fx.exec(line.data(), +1);
            const int k = (wz + nz - g.nzd2) % nz, j = (wy + ny - g.nyd2) % ny;
            for (int wx = 0; wx < nx; wx++) {
                const int i = (wx + nx - g.nxd2) % nx;
                vol[((size_t)k * ny + j) * nx + i] = line[2 * wx] * inv;
            }
        }
    }
    const double t3 = now_ms();
    // sphere mask, cupping compensation and scale (UTR.cpp:1235-1279)
    const float fnorm = (float)g.fnorm;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int k = 0; k < nz; k++) {
        long kk2 = 0;
        if (!g.rot_z_only && !g.rot_z2_only) kk2 = (long)(k - g.kc) * (k - g.kc);
        for (int j = 0; j < ny; j++) {
            const long jj2 = g.rot_y_only ? kk2 : (long)(j - g.jc) * (j - g.jc) + kk2;
            for (int i = 0; i < nx; i++) {
                const long ii2 = (long)(i - g.ic) * (i - g.ic) + jj2;
                float& o = vol[((size_t)k * ny + j) * nx + i];
                if (ii2 > g.nR2) o = 0.f;
                else o *= (float)std::exp(g.mask_dtemp * ii2) * fnorm;
            }
        }
    }
    release();
    if (t4) {
        t4[0] = t1 - t0;
        t4[1] = t2 - t1;
        t4[2] = t3 - t2;
        t4[3] = now_ms() - t0;
    }
    return 0;
}

}  // extern "C"
