// synth.cpp -- seeded synthetic inputs for UTR: random atom sets and their
// per-orientation projection images (flat-field corrected intensities).
//
// The atom rotation follows the forward simulator's convention
// (TemsimLibrary/autosliccmd.cpp:734-748): rotate about Z by +angleZ, then
// about Y' by +angleY, around the box centre, then project along z.  Each atom
// is a 3D Gaussian, so its projection is an analytic 2D Gaussian; the image is
// I = exp(-contrast * sum), optionally with Poisson noise at `dose` counts per
// pixel (I_noisy = Poisson(dose*I)/dose).  Only images of the same statistics
// as multislice output are claimed (ln I of order 0.1), not multislice physics.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include <omp.h>

namespace {

// SplitMix64 (Steele, Lea, Flood 2014): seeds and uniform doubles.
struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

// Poisson sampling: Knuth's product method for small means, Hormann's PTRS
// transformed rejection (Insurance: Mathematics and Economics 12, 1993) for
// means >= 10.
int64_t poisson(SplitMix64& r, double mu) {
    if (mu <= 0) return 0;
    if (mu < 10) {
        double L = std::exp(-mu), p = 1.0;
        int64_t k = 0;
        do { k++; p *= r.uniform(); } while (p > L);
        return k - 1;
    }
    const double smu = std::sqrt(mu), b = 0.931 + 2.53 * smu, a = -0.059 + 0.02483 * b;
    const double inv_alpha = 1.1239 + 1.1328 / (b - 3.4), vr = 0.9277 - 3.6224 / (b - 2);
    const double lmu = std::log(mu);
    for (;;) {
        double U = r.uniform() - 0.5, V = r.uniform();
        double us = 0.5 - std::fabs(U);
        int64_t k = (int64_t)std::floor((2 * a / us + b) * U + mu + 0.43);
        if (us >= 0.07 && V <= vr) return k;
        if (k < 0 || (us < 0.013 && V > us)) continue;
        if (std::log(V) + std::log(inv_alpha) - std::log(a / (us * us) + b) <=
            -mu + k * lmu - std::lgamma((double)k + 1))
            return k;
    }
}

}  // namespace

extern "C" {

// Seeded uniform atom positions (UL, relative to the volume centre).
// shape 0: cube of edge `size`; shape 1: sphere of radius `size`.
// Species are drawn with the given fractions; min_sep > 0 rejects atoms closer
// than min_sep to an earlier atom (bounded number of attempts).
int utr_synth_atoms(uint64_t seed, int natoms, int shape, double size, double min_sep,
                    const int* species_z, const double* species_frac, int nspecies,
                    double* xyz, int* zout) {
    if (natoms < 0 || nspecies < 1 || size <= 0) return -1;
    SplitMix64 r(seed);
    double tot = 0;
    for (int s = 0; s < nspecies; s++) tot += species_frac[s];
    const double ms2 = min_sep * min_sep;
    int placed = 0;
    long attempts = 0;
    while (placed < natoms) {
        if (++attempts > 1000L * (natoms + 10)) return -2;
        double x, y, z;
        if (shape == 0) {
            x = (r.uniform() - 0.5) * size;
            y = (r.uniform() - 0.5) * size;
            z = (r.uniform() - 0.5) * size;
        } else {
            do {
                x = (2 * r.uniform() - 1) * size;
                y = (2 * r.uniform() - 1) * size;
                z = (2 * r.uniform() - 1) * size;
            } while (x * x + y * y + z * z > size * size);
        }
        if (ms2 > 0) {
            bool ok = true;
            for (int k = 0; k < placed && ok; k++) {
                double dx = x - xyz[3 * k], dy = y - xyz[3 * k + 1], dz = z - xyz[3 * k + 2];
                ok = dx * dx + dy * dy + dz * dz >= ms2;
            }
            if (!ok) continue;
        }
        double u = r.uniform() * tot, c = 0;
        int sp = nspecies - 1;
        for (int s = 0; s < nspecies; s++) {
            c += species_frac[s];
            if (u < c) { sp = s; break; }
        }
        xyz[3 * placed] = x;
        xyz[3 * placed + 1] = y;
        xyz[3 * placed + 2] = z;
        zout[placed] = species_z[sp];
        placed++;
    }
    return 0;
}

// Projection images, nrows x ny x nx floats, row-major with x fastest.
// Pixel (i, j) sits at x = (i - nx/2)*xst, y = (j - ny/2)*xst.  An atom of
// atomic number Z contributes (Z/6)^0.7 * exp(-r^2 / (2 width^2)).
int utr_synth_project(const double* xyz, const int* zn, int natoms, int nx, int ny, double xst,
                      int nrows, const double* angZ_deg, const double* angY_deg, double width,
                      double contrast, double dose, uint64_t seed, float* images, int nthreads) {
    if (nx <= 0 || ny <= 0 || nrows < 0 || width <= 0 || xst <= 0) return -1;
    const double pi180 = 3.14159265358979323846 / 180.0;
    const int rad = (int)std::ceil(5.0 * width / xst) + 1;
    const double inv2w2 = 1.0 / (2.0 * width * width);
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
    for (int row = 0; row < nrows; row++) {
        std::vector<double> acc((size_t)nx * ny, 0.0);
        const double aZ = angZ_deg[row] * pi180, aY = angY_deg[row] * pi180;
        const double cZ = std::cos(aZ), sZ = std::sin(aZ), cY = std::cos(aY), sY = std::sin(aY);
        for (int k = 0; k < natoms; k++) {
            const double x = xyz[3 * k], y = xyz[3 * k + 1], z = xyz[3 * k + 2];
            const double x1 = x * cZ - y * sZ, y1 = x * sZ + y * cZ;  // about Z
            const double x2 = x1 * cY + z * sY;                      // about Y'
            const double amp = std::pow(zn[k] / 6.0, 0.7);
            const double pi = x2 / xst + nx / 2, pj = y1 / xst + ny / 2;
            const int i0 = (int)std::floor(pi) - rad, j0 = (int)std::floor(pj) - rad;
            for (int j = std::max(0, j0); j <= std::min(ny - 1, j0 + 2 * rad + 1); j++) {
                const double dy = (j - pj) * xst;
                for (int i = std::max(0, i0); i <= std::min(nx - 1, i0 + 2 * rad + 1); i++) {
                    const double dx = (i - pi) * xst;
                    acc[(size_t)j * nx + i] += amp * std::exp(-(dx * dx + dy * dy) * inv2w2);
                }
            }
        }
        float* img = images + (size_t)row * nx * ny;
        SplitMix64 r(seed * 0x100000001B3ull + (uint64_t)row * 0x9E3779B97F4A7C15ull + 1);
        for (size_t t = 0; t < (size_t)nx * ny; t++) {
            double I = std::exp(-contrast * acc[t]);
            if (dose > 0) I = (double)poisson(r, dose * I) / dose;
            img[t] = (float)I;
        }
    }
    return 0;
}

}  // extern "C"
