// utr_kernels.hip -- gfx950 kernels of the UTR reconstruction path.
//
// Data layout in HBM (all row-major, x fastest, "centred" as the reference's
// shuffled arrays: zero frequency / zero coordinate at index n/2):
//   images   [row][ny][nx]            float      input intensities
//   spectra  [row][ny][nx]            cplx<T>    K2four of every row
//   V        [nz][ny][nx]             cplx<T>    normalised, filtered V3
//   S        [nz][ny][nx]             T          Samp3 (optional output)
//   H        [nz][ny][nx/2+1]         cplx<T>    half spectrum during the 3D
//                                                inverse FFT (x frequency
//                                                wrapped 0..nx/2)
//   vol      [nz][ny][nx]             float/f64  K3out
//
// Reference stages (MultisliceCppTest/PhaseRetrieval):
//   k_rows / k_cols        UTR.cpp:951-1019  (Log0, Shuffle, FFT, 2D filters)
//   k_grid_yp, k_grid_fx,  UTR.h:188-307 gridding + UTR.cpp:1050-1201
//   k_gather               (y-plane tiles, bricks, per-voxel gather)
//   k_scatter, k_normalize plain FP64 form of the same
//   k_pack_z / k_fft_y / k_c2r_x   UTR.cpp:1203-1279
#include <hip/hip_runtime.h>

#include <cmath>

#include "utr_fft.hpp"
#include "utr_plan.hpp"

namespace utr {

// ------------------------------------------------------------ helpers

template <typename T>
__device__ __forceinline__ T dev_log(float x);
template <>
__device__ __forceinline__ float dev_log<float>(float x) { return logf(x); }
template <>
__device__ __forceinline__ double dev_log<double>(float x) { return log((double)x); }

// LDS bytes per FFT buffer, per kernel (measured on gfx950 at 256^3: small
// batches, and so more resident workgroups, win for the row and 3D passes;
// the persistent column pass prefers 64-byte column segments, as does the
// Hermitian pack, which reads mirrored column pairs)
constexpr int LDS_ROWS = 8192, LDS_COLS = 16384, LDS_PACKZ = 16384, LDS_FFTY = 8192, LDS_C2R = 8192,
              LDS_DHT = 16384, LDS_DHT_ROWS = 8192;

// sequences per LDS buffer of at most `budget` bytes, between 1 and 16
__host__ __device__ constexpr int lds_batch(int n, int elem_bytes, int budget) {
    return (budget / (n * elem_bytes)) < 1 ? 1 : ((budget / (n * elem_bytes)) > 16 ? 16 : (budget / (n * elem_bytes)));
}

// Sign (-1)^v for integer v.
__device__ __forceinline__ int par_sign(int v) { return (v & 1) ? -1 : 1; }

// XCD-aware tiling of a 1D launch: workgroups are dealt to the 8 XCDs round
// robin (workgroup L runs on XCD L % 8), so each XCD is given one contiguous
// range of tiles (x fastest).  Neighbouring column groups, which share
// 128-byte lines, are then fetched through the same L2.  Launch
// xcd_grid(gx * gy) workgroups; xcd_tile returns false for the padding.
__host__ __device__ constexpr unsigned xcd_grid(unsigned tiles) { return (tiles + 7u) & ~7u; }

__device__ __forceinline__ bool xcd_tile(int gx, int gy, int& tx, int& ty) {
    const unsigned L = blockIdx.x, T = gridDim.x;
    const unsigned w = (L & 7u) * (T >> 3) + (L >> 3);
    if (w >= (unsigned)gx * (unsigned)gy) return false;
    tx = (int)(w % (unsigned)gx);
    ty = (int)(w / (unsigned)gx);
    return true;
}

// ------------------------------------------------------------- 2D stage

// Rows: renormalise, clamp, Log0 (UTR.cpp:951-958, 991), checkerboard
// (-1)^(i+j) (the input Shuffle of UTR.cpp:993) and forward FFT along x,
// two real rows per complex transform.
// Persistent: each workgroup walks batches of NB rows (nbt batches in all,
// nbr per image) and loads the next batch's pixels into registers while it
// transforms the current one.
template <typename T, int N>
__global__ void __launch_bounds__(256) k_rows(const float* __restrict__ img,
                                              cplx<T>* __restrict__ out, int ny, GeomDev g,
                                              const cplx<T>* __restrict__ tw, int nbr, int nbt) {
    constexpr int NB = lds_batch(N, sizeof(cplx<T>), LDS_ROWS);
    constexpr int PER = (NB * N + 255) / 256;  // pixels per thread and batch
    __shared__ cplx<T> sm[2 * lds_len(NB * N)];
    float xn[PER];
    auto fetch = [&](int bt) {
        const int row = bt / nbr, j0 = (bt - row * nbr) * NB, nb = min(NB, ny - j0);
        const float* src = img + ((size_t)row * ny + j0) * N;
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int t = threadIdx.x + 256 * k;
            xn[k] = t < nb * N ? src[t] : 0.f;
        }
    };
    if ((int)blockIdx.x < nbt) fetch(blockIdx.x);
    for (int bt = blockIdx.x; bt < nbt; bt += gridDim.x) {
        float xc[PER];
#pragma unroll
        for (int k = 0; k < PER; k++) xc[k] = xn[k];
        if (bt + (int)gridDim.x < nbt) fetch(bt + gridDim.x);
        const int row = bt / nbr, j0 = (bt - row * nbr) * NB, nb = min(NB, ny - j0);
        cplx<T>* a = sm;
        cplx<T>* b = sm + lds_len(NB * N);
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int t = threadIdx.x + 256 * k;
            if (t >= nb * N) break;
            const int jj = t / N, i = t - jj * N;
            float x = xc[k];
            if (g.use_f1) x = x / g.f1;
            if (g.use_f2) x = x + g.f2;
            if (x < 0.f) x = 0.f;
            T f = (x > 0.f) ? dev_log<T>(x) : T(0);
            if ((i + j0 + jj) & 1) f = -f;
            // rows 2s and 2s+1 are the real and imaginary parts of sequence s
            T* e = reinterpret_cast<T*>(&a[lx((jj >> 1) * N + i)]);
            e[jj & 1] = f;
            if ((nb & 1) && jj == nb - 1) e[1] = T(0);
        }
        __syncthreads();
        const int np = (nb + 1) >> 1;
        fft_lds<T, N>(a, b, np, tw, -1, threadIdx.x, blockDim.x);
        // the two real rows' spectra from Z = X1 + i X2:
        // X1(k) = (Z(k) + conj Z(N-k)) / 2, X2(k) = (Z(k) - conj Z(N-k)) / 2i;
        // columns 0 .. N/2 only: k_cols forms the others from the Hermitian symmetry
        cplx<T>* dst = out + ((size_t)row * ny + j0) * N;
        constexpr int W = N / 2 + 1;
        for (int t = threadIdx.x; t < np * W; t += blockDim.x) {
            const int sq = t / W, k = t - sq * W;
            const cplx<T> z = a[lx(sq * N + k)], zm = a[lx(sq * N + (k == 0 ? 0 : N - k))];
            dst[(size_t)(2 * sq) * N + k] = {T(0.5) * (z.x + zm.x), T(0.5) * (z.y - zm.y)};
            if (2 * sq + 1 < nb)
                dst[(size_t)(2 * sq + 1) * N + k] = {T(0.5) * (z.y + zm.y), T(-0.5) * (z.x - zm.x)};
        }
        __syncthreads();  // the next batch overwrites both buffers
    }
}

// ----------------------------------------- input image correction (a6)
//
// UTR.cpp:960-977 undoes, before the 2D stage, the Z'' rotation of an input
// image (XArray2DSpln::Rotate(int0, -Z'', yc, xc, aver)) and the shift of a
// RELION row (XArray2DMove::Move(dy, dx, aver), applied first), with aver =
// int0.NormAverEdge(5) of the renormalised, clamped image.  Both routines
// live in the absent XArrayLibrary; restated here:
//   aver   the mean of the pixels within 5 of the image border;
//   Move   content shifted by (dx, dy) pixels, aver where nothing lands;
//   Rotate content turned by the angle about the rotation centre (px, py)
//          in pixel units, counter-clockwise in (x = column, y = row) as the
//          forward model turns atoms about Z (autosliccmd.cpp:734-740); the
//          value at a source position is the cubic convolution of Keys
//          (1981, a = -1/2) of its 4 x 4 neighbours, aver outside the image.
// The output goes back through the inverse of the renormalisation, since
// k_rows renormalises (and clamps) every pixel it reads.

__device__ __forceinline__ float renorm_clamp(float x, const GeomDev& g) {
    if (g.use_f1) x = x / g.f1;
    if (g.use_f2) x = x + g.f2;
    return x < 0.f ? 0.f : x;
}

// one workgroup per image of the batch: aver of image r
__global__ void __launch_bounds__(256) k_prep_aver(const float* __restrict__ img,
                                                   const PrepRow* __restrict__ prep, int row0,
                                                   GeomDev g, float* __restrict__ aver) {
    __shared__ double s_red[4];
    const int r = blockIdx.x, tid = threadIdx.x;
    if (!prep[row0 + r].on) return;  // uniform
    const int nx = g.nx, ny = g.ny, e = min(5, min(nx, ny) / 2);
    const float* src = img + (size_t)r * nx * ny;
    // frame: e full rows at the top and bottom, e pixels at both ends of the rows between
    const int nrow = 2 * e * nx, nside = 2 * e * (ny - 2 * e), total = nrow + nside;
    double sum = 0.0;
    for (int t = tid; t < total; t += 256) {
        int i, j;
        if (t < nrow) {
            const int k = t / nx;
            i = t - k * nx;
            j = k < e ? k : ny - 2 * e + k;
        } else {
            const int u = t - nrow, k = u / (2 * e), c = u - k * 2 * e;
            j = e + k;
            i = c < e ? c : nx - 2 * e + c;
        }
        sum += (double)renorm_clamp(src[(size_t)j * nx + i], g);
    }
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
    if ((tid & 63) == 0) s_red[tid >> 6] = sum;
    __syncthreads();
    if (tid == 0) aver[r] = (float)((s_red[0] + s_red[1] + s_red[2] + s_red[3]) / (double)total);
}

// Keys cubic convolution weights of the neighbours at offsets -1, 0, 1, 2
// for fractional position t in [0, 1), FP64, each expression rounded in the
// order written (no contraction), as the oracle evaluates it
__device__ __forceinline__ void keys_weights(double t, double w[4]) {
    const double t2 = __dmul_rn(t, t), t3 = __dmul_rn(t2, t);
    w[0] = __dsub_rn(__dadd_rn(__dmul_rn(-0.5, t3), t2), __dmul_rn(0.5, t));
    w[1] = __dadd_rn(__dsub_rn(__dmul_rn(1.5, t3), __dmul_rn(2.5, t2)), 1.0);
    w[2] = __dadd_rn(__dadd_rn(__dmul_rn(-1.5, t3), __dmul_rn(2.0, t2)), __dmul_rn(0.5, t));
    w[3] = __dsub_rn(__dmul_rn(0.5, t3), __dmul_rn(0.5, t2));
}

// grid (pixels / 256, images of the batch); (px, py): rotation centre in
// pixels.  FP64 throughout, one rounding to float at the end.
__global__ void __launch_bounds__(256) k_prep_img(const float* __restrict__ img, float* __restrict__ out,
                                                  const PrepRow* __restrict__ prep, int row0, GeomDev g,
                                                  double px, double py, const float* __restrict__ aver) {
    const int r = blockIdx.y;
    const int nx = g.nx, ny = g.ny;
    const size_t n2 = (size_t)nx * ny;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n2) return;
    const float* src = img + (size_t)r * n2;
    const PrepRow P = prep[row0 + r];
    if (!P.on) {
        out[(size_t)r * n2 + t] = src[t];
        return;
    }
    const int i = (int)(t % nx), j = (int)(t / nx);
    const double bg = (double)aver[r];
    // source position in the shifted image
    const double dxp = __dsub_rn((double)i, px), dyp = __dsub_rn((double)j, py);
    const double sx = __dadd_rn(__dadd_rn(px, __dmul_rn(dxp, P.cs)), __dmul_rn(dyp, P.sn));
    const double sy = __dadd_rn(__dsub_rn(py, __dmul_rn(dxp, P.sn)), __dmul_rn(dyp, P.cs));
    const double fx0 = floor(sx), fy0 = floor(sy);
    double wx[4], wy[4];
    keys_weights(__dsub_rn(sx, fx0), wx);
    keys_weights(__dsub_rn(sy, fy0), wy);
    const int ix = (int)fx0 - 1, iy = (int)fy0 - 1;
    double acc = 0.0;
#pragma unroll
    for (int b = 0; b < 4; b++) {
        const int yy = iy + b, ys = yy - P.dy;
#pragma unroll
        for (int a = 0; a < 4; a++) {
            const int xx = ix + a, xs = xx - P.dx;
            double v = bg;
            if ((unsigned)xx < (unsigned)nx && (unsigned)yy < (unsigned)ny && (unsigned)xs < (unsigned)nx &&
                (unsigned)ys < (unsigned)ny)
                v = (double)renorm_clamp(src[(size_t)ys * nx + xs], g);
            acc = __dadd_rn(acc, __dmul_rn(__dmul_rn(wx[a], wy[b]), v));
        }
    }
    // back through the inverse renormalisation (k_rows applies it again)
    if (g.use_f2) acc = __dsub_rn(acc, (double)g.f2);
    if (g.use_f1) acc = __dmul_rn(acc, (double)g.f1);
    out[(size_t)r * n2 + t] = (float)acc;
}

// Columns: forward FFT along y, output Shuffle as (-1)^(kx+ky+nx/2+ny/2),
// optional 2D filter (modes 1/3, restated) and subseries weight.
// The image is real, so its (centred) spectrum before the filter is
// Hermitian: K(nx - kx, (ny - ky) mod ny) = conj K(kx, ky); the sign,
// filters and weight are even in the frequency.  Only columns 0 .. ncol-1 =
// nx/2 are transformed (k_rows writes no others); the conjugate of each
// value goes through the same filter into its mirror in columns
// nx/2+1 .. nx-1 (the mode-3 filter with ESCC is complex, so the filtered
// spectrum itself need not be Hermitian).
// rowmax (optional): per-row max(|re|, |im|) of the result as float bits
// (non-negative floats order like their bit patterns); it sets the
// fixed-point scale of the brick gridding.
template <typename T, int N>
__global__ void __launch_bounds__(256) k_cols(cplx<T>* __restrict__ spec, int nx, int ncol, GeomDev g,
                                              const RowParams* __restrict__ rows, int row0,
                                              const cplx<T>* __restrict__ tw,
                                              unsigned* __restrict__ rowmax, int gx, int gy) {
    constexpr int NB = lds_batch(N, sizeof(cplx<T>), LDS_COLS);
    constexpr int PER = (NB * N + 255) / 256;  // elements per thread and tile
    __shared__ cplx<T> sm[2 * lds_len(NB * N)];
    const int ny = N;
    const int total = gx * gy;
    // Persistent: in round k this workgroup takes tile tile_of(k); within a
    // round the tiles of one XCD are contiguous, so neighbouring column groups
    // share that XCD's L2.  The next tile's columns are loaded into registers
    // while the current one is transformed.  gridDim.x is a multiple of 8.
    const unsigned L = blockIdx.x, G = gridDim.x;
    auto tile_of = [&](int k) { return (int)(k * G + (L & 7u) * (G >> 3) + (L >> 3)); };
    cplx<T> vn[PER];
    auto fetch = [&](int w) {
        const int row = w / gx, x0 = (w - row * gx) * NB, nb = min(NB, ncol - x0);
        const cplx<T>* src = spec + (size_t)row * ny * nx + x0;
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int t = threadIdx.x + 256 * q;
            if (t < nb * N) {
                const int y = t / nb, c = t - y * nb;
                vn[q] = src[(size_t)y * nx + c];
            }
        }
    };
    const int rounds = (total + (int)G - 1) / (int)G;
    if (tile_of(0) < total) fetch(tile_of(0));
    for (int rk = 0; rk < rounds; rk++) {
        const int wt = tile_of(rk);
        cplx<T> vc[PER];
#pragma unroll
        for (int q = 0; q < PER; q++) vc[q] = vn[q];
        if (rk + 1 < rounds && tile_of(rk + 1) < total) fetch(tile_of(rk + 1));
        if (wt >= total) continue;  // uniform over the workgroup
        cplx<T>* a = sm;
        cplx<T>* b = sm + lds_len(NB * N);
        const int row = wt / gx;
        const int x0 = (wt - row * gx) * NB;
        const int nb = min(NB, ncol - x0);
        cplx<T>* const rowbase = spec + (size_t)row * ny * nx;
        cplx<T>* base = rowbase + x0;
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int t = threadIdx.x + 256 * q;
            if (t < nb * N) {
                const int y = t / nb, c = t - y * nb;
                a[lx(c * N + y)] = vc[q];
            }
        }
        __syncthreads();
        fft_lds<T, N>(a, b, nb, tw, -1, threadIdx.x, blockDim.x);
        const RowParams& R = rows[row0 + row];
        const int gsgn = par_sign(g.nxd2 + g.nyd2);
        const double xap = g.xhi - g.xlo, yap = g.yhi - g.ylo;
        const double pi = 3.14159265358979323846;
        float vmax = 0.f;
        for (int t = threadIdx.x; t < nb * N; t += blockDim.x) {
            const int ky = t / nb, c = t - ky * nb;
            const int kx = x0 + c;
            cplx<T> v = a[lx(c * N + ky)];
            const T s = T(par_sign(kx + ky) * gsgn);
            v.x *= s;
            v.y *= s;
            // the mirror's value before the filter: the image spectrum is
            // Hermitian and the sign is even; the filter below is even in the
            // frequency but complex in mode 3 with ESCC, so it is applied to
            // both values
            const bool mirror = kx > 0 && kx < g.nxd2;
            cplx<T> vm = {v.x, -v.y};
            if (g.ctf_mode == 1 && R.zout != 0) {
                // 2D TIE-Hom restated from UTR.h:159-181 in 2D (unpinned)
                const double i1 = kx - g.nxd2, j1 = ky - g.nyd2;
                const double q2 = 4.0 * pi * pi * (i1 * i1 / (xap * xap) + j1 * j1 / (yap * yap));
                const T f = T(g.two_d_norm / (g.two_d_alpha + q2));
                v.x *= f;
                v.y *= f;
                vm.x *= f;
                vm.y *= f;
            } else if (g.ctf_mode == 3) {
                // 2D CTF inversion restated from UTR.h:70-128 in 2D (unpinned):
                // at centred frequency (i1, j1), inside the aperture, u -> u (sr - i cr)
                const double dk2 = 1.0 / (xap * xap), de2 = 1.0 / (yap * yap);
                auto coef = [&](double i1, double j1, T& sr, T& cr) -> bool {
                    const double q2 = dk2 * i1 * i1 + de2 * j1 * j1;
                    if (g.k2maxo > 0 && !(q2 < g.k2maxo)) return false;
                    const double fac2 = pi * g.wl * R.zout;
                    double eps = g.epsilon;
                    if (eps <= 0) {
                        const double mn = dk2 < de2 ? dk2 : de2;
                        eps = 0.1 * (fabs(fac2) * mn) * (fabs(fac2) * mn);
                    }
                    double chi = fac2 * q2 - atan(g.asigma);
                    // astigmatism: defocus zout + dz cos 2(phi - phi0) at azimuth phi
                    if (R.astig_c != 0 || R.astig_s != 0)
                        chi += R.astig_c * (dk2 * i1 * i1 - de2 * j1 * j1) +
                               2.0 * R.astig_s * (1.0 / (xap * yap)) * i1 * j1;
                    if (g.cs3 != 0 || g.cs5 != 0)
                        chi += pi * g.wl * g.wl * g.wl / 2.0 * g.cs3 * q2 * q2 +
                               pi * g.wl * g.wl * g.wl * g.wl * g.wl / 3.0 * g.cs5 * q2 * q2 * q2;
                    double sn, cs;
                    sincos(chi, &sn, &cs);
                    sr = T(sn / (sn * sn + eps));
                    cr = g.escc ? T(cs / (cs * cs + eps)) : T(0);
                    return true;
                };
                auto mul = [](cplx<T> u, T sr, T cr) { return cplx<T>{u.x * sr + u.y * cr, u.y * sr - u.x * cr}; };
                const double i1 = kx - g.nxd2, j1 = ky - g.nyd2;
                T sr, cr;
                const bool in = coef(i1, j1, sr, cr);
                if (in) v = mul(v, sr, cr);
                // the filter is even, so the mirror (-i1, -j1) takes the same
                // factor -- except on row ky = 0, which is its own mirror row
                // (j1 = -ny/2 on both sides): there the astigmatic cross term
                // i1 j1 changes sign
                if (mirror && ky == 0 && R.astig_s != 0) {
                    T sr2, cr2;
                    if (coef(-i1, j1, sr2, cr2)) vm = mul(vm, sr2, cr2);
                } else if (in) {
                    vm = mul(vm, sr, cr);
                }
            }
            if (R.weight != 1.0f) {
                v.x *= T(R.weight);
                v.y *= T(R.weight);
                vm.x *= T(R.weight);
                vm.y *= T(R.weight);
            }
            base[(size_t)ky * nx + c] = v;
            if (mirror) {
                rowbase[(size_t)(ky == 0 ? 0 : ny - ky) * nx + (nx - kx)] = vm;
                vmax = fmaxf(vmax, fmaxf(fabsf((float)vm.x), fabsf((float)vm.y)));
            }
            vmax = fmaxf(vmax, fmaxf(fabsf((float)v.x), fabsf((float)v.y)));
        }
        if (rowmax) {
            // one atomic per workgroup (rowmax is a kernel argument: uniform branch)
            __shared__ float s_wmax[4];
            for (int off = 32; off > 0; off >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, off));
            if ((threadIdx.x & 63) == 0) s_wmax[threadIdx.x >> 6] = vmax;
            __syncthreads();
            if (threadIdx.x == 0) {
                const float m = fmaxf(fmaxf(s_wmax[0], s_wmax[1]), fmaxf(s_wmax[2], s_wmax[3]));
                atomicMax(&rowmax[row0 + row], __float_as_uint(m));
            }
        }
        __syncthreads();  // the next tile overwrites both buffers
    }
}

// ------------------------------------------------- gridding coordinates

// Grid position of spectrum point (a, b) = (i - nx/2, j - ny/2) of one row:
// the rotation of UTR.h:243-256 in grid units, with explicit fma so that the
// CPU restatement evaluates bit-identical coordinates.
struct GridPos {
    double dx1, dy1, dz1, ox, oy, oz;
};

__device__ __forceinline__ GridPos grid_pos(const GeomDev& g, const double* m, int ia, int ib) {
    const double a = (double)ia, b = (double)ib;
    const double qx = __dmul_rn(g.fxst, a), qy = __dmul_rn(g.fyst, b);
    const double zz = g.escc ? __dmul_rn(-g.wl2, __fma_rn(qx, qx, __dmul_rn(qy, qy))) : 0.0;
    GridPos p;
    p.ox = __fma_rn(a, m[0], __fma_rn(b, m[1], __dmul_rn(zz, m[2])));
    p.oy = __fma_rn(a, m[3], __fma_rn(b, m[4], __dmul_rn(zz, m[5])));
    p.oz = __fma_rn(a, m[6], __dmul_rn(zz, m[8]));
    p.dx1 = __dadd_rn((double)g.nxd2, p.ox);
    p.dy1 = __dadd_rn((double)g.nyd2, p.oy);
    p.dz1 = __dadd_rn((double)g.nzd2, p.oz);
    return p;
}

// --------------------------------------------------- voxel epilogue

// 1/b: in FP32 from the hardware reciprocal and one Newton step (within an
// ulp of the correctly rounded quotient), in FP64 the IEEE quotient
__device__ __forceinline__ float recip_nr(float b) {
    const float r = __builtin_amdgcn_rcpf(b);
    return fmaf(fmaf(-b, r, 1.f), r, r);
}
__device__ __forceinline__ double recip_nr(double b) { return 1.0 / b; }
// square root: in FP32 the hardware instruction (within 1 ulp), in FP64 the
// IEEE one
__device__ __forceinline__ float vsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double vsqrt(double x) { return sqrt(x); }

// Sampling normalisation + noise filter (UTR.cpp:1050-1128) and the 3D
// filters (UTR.cpp:1177-1201; UTR.h:15-185) for voxel (i, j, k).  Factors
// that depend on the geometry only come precomputed in g.ep_* (FP64).
template <typename T>
__device__ __forceinline__ cplx<T> voxel_epilogue(const GeomDev& g, int i, int j, int k,
                                                  cplx<T> v, T x) {
    const T inv_pi = T(0.318309886183790671538);
    const T target = T(kTargetSNR);
    if (g.dht_sctf) {
        // DHT.cpp:1350-1365: x -> (1/x) exp((x - SNRc)/x^4) where sampled
        if (x != T(0)) {
            const double xd = (double)x;
            const T f = T((1.0 / xd) * exp((xd - g.critical_snr) / (xd * xd * xd * xd)));
            v.x *= f;
            v.y *= f;
        }
    } else if (g.nsr2 == 0) {
        if (x == T(0)) return {T(0), T(0)};
        const T r = T(1) / x;
        v.x *= r;
        v.y *= r;
    } else {
        if (x == T(0)) return {T(0), T(0)};
        T fnsr;
        if (g.ctf_mode != 1) {
            fnsr = T(g.nsr2 * (double)vsqrt(x));
        } else {
            const int k1 = k - g.nzd2, j1 = j - g.nyd2, i1 = i - g.nxd2;
            const double zeta2 = k1 * k1 * g.ep_dzeta2 - g.asigma;
            const double eta2 = j1 * j1 * g.ep_deta2 + zeta2;
            fnsr = T(g.ep_nsr1 * (double)vsqrt(x) / (g.ep_dksi2 * i1 * i1 + eta2 + g.epsilon));
        }
        const T y = vsqrt(v.x * v.x + v.y * v.y) * recip_nr(fnsr);
        const T f = recip_nr(x) * (T(0.5) + atan(T(100) * (y - target)) * inv_pi);
        v.x *= f;
        v.y *= f;
    }
    const int k1 = k - g.nzd2, j1 = j - g.nyd2, i1 = i - g.nxd2;
    if (g.do_tie3d || g.do_invlap) {
        // InverseMLaplacian3D(V3, zhi-zlo, yhi-ylo, xhi-ylo, alpha, norm)
        const double base = g.ep_lapx * ((double)i1 * i1) + g.ep_lapy * ((double)j1 * j1) +
                            g.ep_lapz * ((double)k1 * k1);
        if (g.do_tie3d) {
            const T f = T(g.tie_norm / (base + g.tie_alpha));
            v.x *= f;
            v.y *= f;
        }
        if (g.do_invlap) {
            const T f = T(g.invlap_eps / (base + g.invlap_eps));
            v.x *= f;
            v.y *= f;
        }
    }
    if (g.do_ctf3d) {
        // CTFcorrection3D (UTR.h:86-131); phase argument in FP64
        const double zeta2 = g.ep_iz2 * k1 * k1;
        const double zeta2fac2 = g.ep_fac2 * zeta2 - g.ep_atans;
        double eta2 = g.ep_iy2 * j1 * j1;
        const double eta2fac2 = g.ep_fac2 * eta2 + zeta2fac2;
        eta2 += zeta2;
        double ksi2 = g.ep_ix2 * i1 * i1;
        const double ksi2fac2 = g.ep_fac2 * ksi2 + eta2fac2;
        ksi2 += eta2;
        if (!(g.k2maxo > 0) || ksi2 < g.k2maxo) {
            double dt = ksi2fac2;
            if (g.cs3 != 0 || g.cs5 != 0)
                dt += g.ep_c3 * ksi2 * ksi2 + g.ep_c5 * ksi2 * ksi2 * ksi2;
            double sn, cs;
            sincos(dt, &sn, &cs);
            const T sr = T(sn / (sn * sn + g.ctf_eps));
            const T cr = g.escc ? T(cs / (cs * cs + g.ctf_eps)) : T(0);
            const cplx<T> w = {v.x * sr + v.y * cr, v.y * sr - v.x * cr};
            v = w;
        }
    }
    return v;
}

// ------------------------------------------------ gather gridding (FP32)

constexpr int GBX = 8, GBY = 8, GBZ = 4;  // brick of 256 voxels, one per thread
constexpr int LIST_CAP = 1024;

// Conservative test: may any point of this row touch a voxel of the brick
// centred at grid offset (cx, cy, cz) with half extents (ex, ey, ez)?
__device__ __forceinline__ bool brick_touch(const GeomDev& g, const RowParams& R, float cx,
                                            float cy, float cz, float ex, float ey, float ez,
                                            float fx2, float fy2, float wl2f) {
    const float* mi = R.minv;
    const float s = mi[0] * cx + mi[1] * cy + mi[2] * cz;
    const float t = mi[3] * cx + mi[4] * cy + mi[5] * cz;
    const float h = mi[6] * cx + mi[7] * cy + mi[8] * cz;
    const float Wa = fabsf(mi[0]) * ex + fabsf(mi[1]) * ey + fabsf(mi[2]) * ez + 0.01f;
    const float Wb = fabsf(mi[3]) * ex + fabsf(mi[4]) * ey + fabsf(mi[5]) * ez + 0.01f;
    const float Wz = (fabsf(mi[6]) * ex + fabsf(mi[7]) * ey + fabsf(mi[8]) * ez) * 1.01f + 0.01f * R.wz;
    float alo = fmaxf(s - Wa, (float)(-g.nxd2)), ahi = fminf(s + Wa, (float)(g.nx - 1 - g.nxd2));
    float blo = fmaxf(t - Wb, (float)(-g.nyd2)), bhi = fminf(t + Wb, (float)(g.ny - 1 - g.nyd2));
    if (alo > ahi || blo > bhi) return false;
    if (!g.escc) return fabsf(h) <= Wz;
    const float a2min = (alo <= 0.f && ahi >= 0.f) ? 0.f : fminf(alo * alo, ahi * ahi);
    const float a2max = fmaxf(alo * alo, ahi * ahi);
    const float b2min = (blo <= 0.f && bhi >= 0.f) ? 0.f : fminf(blo * blo, bhi * bhi);
    const float b2max = fmaxf(blo * blo, bhi * bhi);
    const float zz_hi = -wl2f * (fx2 * a2min + fy2 * b2min);
    const float zz_lo = -wl2f * (fx2 * a2max + fy2 * b2max);
    const float tol = 1e-3f * (fabsf(zz_lo) + Wz);
    return h >= zz_lo - Wz - tol && h <= zz_hi + Wz + tol;
}

__device__ __forceinline__ void kahan_add(float& sum, float& c, float v) {
    const float y = v - c;
    const float t = sum + y;
    c = (t - sum) - y;
    sum = t;
}

// One workgroup per brick; every thread owns one voxel and accumulates the
// contributions of all rows in row order (then j, then i, as UTR.h:230-240
// visits them) with compensated FP32 sums.  Deterministic, no atomics.
__global__ void __launch_bounds__(256) k_gather(const cplx<float>* __restrict__ spec,
                                                const RowParams* __restrict__ rows, int nrows,
                                                GeomDev g, cplx<float>* __restrict__ Vout,
                                                float* __restrict__ Sout) {
    __shared__ int s_list[LIST_CAP];
    __shared__ int s_cnt[4];
    __shared__ int s_n;
    const int tid = threadIdx.x;
    const int nbx = (g.nx + GBX - 1) / GBX, nby = (g.ny + GBY - 1) / GBY;
    const int bid = blockIdx.x;
    const int bx = bid % nbx, by = (bid / nbx) % nby, bz = bid / (nbx * nby);
    const int vx = bx * GBX + (tid % GBX);
    const int vy = by * GBY + ((tid / GBX) % GBY);
    const int vz = bz * GBZ + tid / (GBX * GBY);
    const float fx2 = (float)(g.fxst * g.fxst), fy2 = (float)(g.fyst * g.fyst);
    const float wl2f = (float)g.wl2;
    // brick centre and half extents in grid offsets, grown by one voxel
    const float bcx = bx * GBX + 0.5f * (GBX - 1) - g.nxd2;
    const float bcy = by * GBY + 0.5f * (GBY - 1) - g.nyd2;
    const float bcz = bz * GBZ + 0.5f * (GBZ - 1) - g.nzd2;
    const float ex = 0.5f * (GBX - 1) + 1.f, ey = 0.5f * (GBY - 1) + 1.f, ez = 0.5f * (GBZ - 1) + 1.f;
    const float dxg = (float)(vx - g.nxd2), dyg = (float)(vy - g.nyd2), dzg = (float)(vz - g.nzd2);
    const bool inside = vx < g.nx && vy < g.ny && vz < g.nz;

    float sr = 0.f, cr = 0.f, si = 0.f, ci = 0.f, ss = 0.f, cs = 0.f;
    const int lane = tid & 63, wave = tid >> 6;

    for (int base = 0; base < nrows; base += LIST_CAP) {
        const int nchunk = min(LIST_CAP, nrows - base);
        if (tid == 0) s_n = 0;
        __syncthreads();
        for (int r0 = 0; r0 < nchunk; r0 += 256) {
            const int r = r0 + tid;
            bool pass = false;
            if (r < nchunk)
                pass = brick_touch(g, rows[base + r], bcx, bcy, bcz, ex, ey, ez, fx2, fy2, wl2f);
            const unsigned long long mask = __ballot(pass);
            const int pre = __popcll(mask & ((1ull << lane) - 1ull));
            if (lane == 0) s_cnt[wave] = __popcll(mask);
            __syncthreads();
            int off = s_n;
            for (int w = 0; w < wave; w++) off += s_cnt[w];
            if (pass) s_list[off + pre] = base + r;
            __syncthreads();
            if (tid == 0) s_n += s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
            __syncthreads();
        }
        const int n = s_n;
        for (int q = 0; q < n; q++) {
            const int idx = __builtin_amdgcn_readfirstlane(s_list[q]);
            const RowParams& R = rows[idx];
            if (!inside) continue;
            const float* mi = R.minv;
            const float s = mi[0] * dxg + mi[1] * dyg + mi[2] * dzg;
            const float t = mi[3] * dxg + mi[4] * dyg + mi[5] * dzg;
            const float h = mi[6] * dxg + mi[7] * dyg + mi[8] * dzg;
            float zzs = 0.f, G = 0.f;
            if (g.escc) {
                zzs = -wl2f * (fx2 * s * s + fy2 * t * t);
                G = 2.f * wl2f * (fx2 * (fabsf(s) + R.wa) * R.wa + fy2 * (fabsf(t) + R.wb) * R.wb);
            }
            if (fabsf(zzs - h) > R.wz * 1.001f + G * 1.01f + 1e-3f * R.wz) continue;
            const float eps = 1e-3f;
            int alo = (int)ceilf(s - R.wa - eps), ahi = (int)floorf(s + R.wa + eps);
            alo = max(alo, -g.nxd2);
            ahi = min(ahi, g.nx - 1 - g.nxd2);
            int blo, bhi;
            if (R.y_exact) {
                blo = bhi = vy - g.nyd2;
            } else {
                blo = (int)ceilf(t - R.wb - eps);
                bhi = (int)floorf(t + R.wb + eps);
                blo = max(blo, -g.nyd2);
                bhi = min(bhi, g.ny - 1 - g.nyd2);
            }
            const cplx<float>* K = spec + (size_t)idx * g.ny * g.nx;
            for (int b = blo; b <= bhi; b++) {
                for (int a = alo; a <= ahi; a++) {
                    const GridPos P = grid_pos(g, R.m, a, b);
                    const double fnn = floor(P.dz1), fii = floor(P.dx1), fjj = floor(P.dy1);
                    if (fnn < 0.0 || fnn > (double)(g.nz - 2)) continue;
                    if (fii < 0.0 || fii > (double)(g.nx - 2)) continue;
                    if (fjj < 0.0 || fjj > (double)(g.ny - 2)) continue;
                    const int ox = vx - (int)fii, oy = vy - (int)fjj, oz = vz - (int)fnn;
                    if ((unsigned)ox > 1u || (unsigned)oy > 1u || (unsigned)oz > 1u) continue;
                    const double fx = P.dx1 - fii, fy = P.dy1 - fjj, fz = P.dz1 - fnn;
                    const double wx = ox ? fx : 1.0 - fx;
                    const double wy = oy ? fy : 1.0 - fy;
                    const double wz = oz ? fz : 1.0 - fz;
                    const float w = (float)(wx * wy * wz);
                    cplx<float> kv = K[(size_t)(b + g.nyd2) * g.nx + (a + g.nxd2)];
                    if (g.rot_centre_shift) {
                        const double ph = -(R.ph[0] * P.ox + R.ph[1] * P.oy + R.ph[2] * P.oz);
                        double sn, c;
                        sincos(ph, &sn, &c);
                        const float cf = (float)c, sf = (float)sn;
                        kv = {kv.x * cf - kv.y * sf, kv.x * sf + kv.y * cf};
                    }
                    kahan_add(sr, cr, w * kv.x);
                    kahan_add(si, ci, w * kv.y);
                    kahan_add(ss, cs, w);
                }
            }
        }
        __syncthreads();
    }
    if (!inside) return;
    const size_t vidx = ((size_t)vz * g.ny + vy) * g.nx + vx;
    if (Sout) Sout[vidx] = ss;
    Vout[vidx] = voxel_epilogue<float>(g, vx, vy, vz, cplx<float>{sr, si}, ss);
}

// --------------------- FP32 scatter into bricks, exact fixed-point sums
//
// One workgroup owns a brick of SBX x SBY x SBZ voxels.  It lists the rows
// whose Ewald sheet can reach the brick (in row order) and, for each listed
// row, the detector lines b whose points can splat into it.  Lines are
// processed 256 at a time, one per thread: the thread bounds the a-interval
// of its line that can land in the brick, evaluates each point's grid
// position in FP64 (the same expressions as the CPU restatement), and adds
// w*K (FP32 products) for the point's corners inside the brick.
//
// The sums are kept in LDS as 64-bit fixed-point integers, added with integer
// LDS atomics (on gfx950 these cost no more than plain stores, while float
// LDS atomics are several times slower).  Integer addition is associative, so
// the sums are exact, independent of the order in which points arrive, and
// the result is deterministic; the only rounding is that of each w*K product
// and of its conversion to the fixed-point grid.  The scales come from the
// largest spectrum component and the row count so that no sum can overflow:
// V3 is resolved to about 1e-11 (in units of K), Samp3 to about 1e-15.
// Weights below 2^-20 (rare: cells crossed within a millionth of a grid
// plane, mostly at exact grid alignments) go instead to float accumulators,
// so a voxel reached only by such weights keeps the relative precision that
// V3/Samp3 needs there and counts as sampled exactly as in the reference.
// Corners of weight exactly zero are skipped (adding them changes nothing).
//
// The fixed-point accumulators use odd strides (17 along y, 153 along z) so
// that lanes walking along x, y or z, as the points of a detector line do for
// any orientation, address distinct LDS banks.

constexpr int SBX = 16, SBY = 8, SBZ = 8;
constexpr int SVOX = SBX * SBY * SBZ;
constexpr int SOWN = SVOX / 256;  // voxels finalised per thread
static_assert(SVOX % 256 == 0, "brick must be a multiple of the workgroup");
constexpr int SPY = SBX + 1;           // LDS stride along y
constexpr int SPZ = SPY * (SBY + 1);   // LDS stride along z
constexpr int SPAD = SPZ * SBZ;        // LDS entries per component
constexpr int S_LIST = 256;            // rows tested per chunk
constexpr double S_TINY_D = 9.5367431640625e-07;  // 2^-20
constexpr double S_TINY_1D = 9.765625e-04;         // 2^-10

__device__ __forceinline__ int lds_idx(int lx, int ly, int lz) { return lx + SPY * ly + SPZ * lz; }

// round(d) as a 64-bit integer for |d| < 2^51 (1.5 * 2^52 aligns the units
// digit with the last mantissa bit; round to nearest even)
__device__ __forceinline__ long long dround_i64(double d) {
    const double y = d + 6755399441055744.0;
    return (long long)(__double_as_longlong(y) - 0x4338000000000000LL);
}

// round(w * k) as a 64-bit integer for |w * k| < 2^51: one fma rounds the
// exact product once, as dround_i64 of the exact product would
__device__ __forceinline__ long long fx_term(double w, double k) {
    const double y = __fma_rn(w, k, 6755399441055744.0);
    return (long long)(__double_as_longlong(y) - 0x4338000000000000LL);
}

__global__ void __launch_bounds__(256, 3) k_grid_fx(const cplx<float>* __restrict__ spec,
                                                 const RowParams* __restrict__ rows, int nrows,
                                                 const unsigned* __restrict__ rowmax, GeomDev g,
                                                 cplx<float>* __restrict__ Vout,
                                                 float* __restrict__ Sout) {
    __shared__ unsigned long long acc[3 * SPAD];
    __shared__ float tiny[3 * SVOX];
    __shared__ float s_red[4];
    __shared__ int s_list[S_LIST];     // rows reaching the brick, in row order
    __shared__ int s_lo[S_LIST];       // first detector line of each listed row
    __shared__ int s_pre[S_LIST + 1];  // line counts, then their exclusive prefix
    __shared__ int s_cnt[4], s_wsum[4];
    __shared__ int s_n;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nbx = (g.nx + SBX - 1) / SBX, nby = (g.ny + SBY - 1) / SBY;
    const int bid = blockIdx.x;
    const int bx = bid % nbx, by = (bid / nbx) % nby, bz = bid / (nbx * nby);
    const int x0 = bx * SBX, y0 = by * SBY, z0 = bz * SBZ;
    for (int t = tid; t < 3 * SPAD; t += 256) acc[t] = 0ull;
    for (int t = tid; t < 3 * SVOX; t += 256) tiny[t] = 0.f;
    // fixed-point scales: 2^sv for V3, 2^sw for Samp3.  Per row a voxel
    // receives a total weight below 8, so |sum| < 8 nrows max|K| 2^sv < 2^62;
    // each term stays below 2^51 for fx_term.
    float mk = 0.f;
    for (int r = tid; r < nrows; r += 256) mk = fmaxf(mk, __uint_as_float(rowmax[r]));
    for (int off = 32; off > 0; off >>= 1) mk = fmaxf(mk, __shfl_xor(mk, off));
    if (lane == 0) s_red[wave] = mk;
    __syncthreads();
    mk = fmaxf(fmaxf(s_red[0], s_red[1]), fmaxf(s_red[2], s_red[3]));
    int ek = 0, ea = 0;
    frexpf(mk > 0.f ? mk : 1.f, &ek);       // max|K| < 2^ek
    frexpf((float)(8 * nrows), &ea);         // 8 nrows < 2^ea
    const int sv = min(61 - ek - ea, 50 - ek);
    const int sw = min(61 - ea, 50);
    const double vscale = ldexp(1.0, sv), wscale = ldexp(1.0, sw);

    const float fx2 = (float)(g.fxst * g.fxst), fy2 = (float)(g.fyst * g.fyst);
    const float wl2f = (float)g.wl2;
    const float bcx = x0 + 0.5f * (SBX - 1) - g.nxd2;
    const float bcy = y0 + 0.5f * (SBY - 1) - g.nyd2;
    const float bcz = z0 + 0.5f * (SBZ - 1) - g.nzd2;
    const float ex = 0.5f * (SBX - 1) + 1.f, ey = 0.5f * (SBY - 1) + 1.f, ez = 0.5f * (SBZ - 1) + 1.f;
    const float bmin[3] = {(float)x0, (float)y0, (float)z0};
    const float bmax[3] = {(float)(x0 + SBX - 1), (float)(y0 + SBY - 1), (float)(z0 + SBZ - 1)};
    const float cgrid[3] = {(float)g.nxd2, (float)g.nyd2, (float)g.nzd2};

    for (int base = 0; base < nrows; base += S_LIST) {
        const int nchunk = min(S_LIST, nrows - base);
        if (tid == 0) s_n = 0;
        __syncthreads();
        // rows whose sheet can reach the brick, with their detector lines
        for (int r0 = 0; r0 < nchunk; r0 += 256) {
            const int r = r0 + tid;
            bool pass = false;
            int blo = 0, nl = 0;
            if (r < nchunk) {
                const RowParams& R = rows[base + r];
                pass = brick_touch(g, R, bcx, bcy, bcz, ex, ey, ez, fx2, fy2, wl2f);
                if (pass) {
                    int bhi;
                    if (R.y_exact) {
                        // y grid coordinate is ny/2 + b exactly: lines outside the
                        // brick's y-range only add weight 0 to it
                        blo = max(y0 - g.nyd2, -g.nyd2);
                        bhi = min(y0 + SBY - 1 - g.nyd2, g.ny - 1 - g.nyd2);
                    } else {
                        const float* mi = R.minv;
                        const float t = mi[3] * bcx + mi[4] * bcy + mi[5] * bcz;
                        const float Wb = fabsf(mi[3]) * ex + fabsf(mi[4]) * ey + fabsf(mi[5]) * ez + 0.01f;
                        blo = max((int)ceilf(t - Wb), -g.nyd2);
                        bhi = min((int)floorf(t + Wb), g.ny - 1 - g.nyd2);
                    }
                    nl = bhi - blo + 1;
                    pass = nl > 0;
                }
            }
            const unsigned long long mask = __ballot(pass);
            const int pre = __popcll(mask & ((1ull << lane) - 1ull));
            if (lane == 0) s_cnt[wave] = __popcll(mask);
            __syncthreads();
            int off = s_n;
            for (int w = 0; w < wave; w++) off += s_cnt[w];
            if (pass) {
                s_list[off + pre] = base + r;
                s_lo[off + pre] = blo;
                s_pre[off + pre] = nl;
            }
            __syncthreads();
            if (tid == 0) s_n += s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
            __syncthreads();
        }
        const int n = s_n;
        // exclusive prefix of the line counts (two entries per thread)
        {
            const int c0 = (2 * tid < n) ? s_pre[2 * tid] : 0;
            const int c1 = (2 * tid + 1 < n) ? s_pre[2 * tid + 1] : 0;
            const int pair = c0 + c1;
            int incl = pair;
            for (int o = 1; o < 64; o <<= 1) {
                const int v = __shfl_up(incl, o);
                if (lane >= o) incl += v;
            }
            if (lane == 63) s_wsum[wave] = incl;
            __syncthreads();
            int woff = 0;
            for (int w = 0; w < wave; w++) woff += s_wsum[w];
            const int excl = woff + incl - pair;
            if (2 * tid < n) s_pre[2 * tid] = excl;
            if (2 * tid + 1 < n) s_pre[2 * tid + 1] = excl + c0;
            if (tid == 0) s_pre[n] = s_wsum[0] + s_wsum[1] + s_wsum[2] + s_wsum[3];
            __syncthreads();
        }
        const int total = s_pre[n];
        for (int G = 0; G < total; G += 256) {
            const int gl = G + tid;
            if (gl < total) {
                // listed row q holding global line gl: last q with s_pre[q] <= gl
                int lo = 0, hi = n - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (s_pre[mid] <= gl) lo = mid;
                    else hi = mid - 1;
                }
                const int r = s_list[lo];
                const int b = s_lo[lo] + (gl - s_pre[lo]);
                const RowParams& R = rows[r];
                const float* mi = R.minv;
                // a-interval of this line whose points can land in the brick
                const float s = mi[0] * bcx + mi[1] * bcy + mi[2] * bcz;
                const float Wa = fabsf(mi[0]) * ex + fabsf(mi[1]) * ey + fabsf(mi[2]) * ez + 0.01f;
                float alo = fmaxf(s - Wa, (float)(-g.nxd2));
                float ahi = fminf(s + Wa, (float)(g.nx - 1 - g.nxd2));
                double m[9];
#pragma unroll
                for (int t = 0; t < 9; t++) m[t] = R.m[t];
                if (alo <= ahi) {
                    const float ac = 0.5f * (alo + ahi), W = 0.5f * (ahi - alo);
                    const float fb = (float)b;
                    const float zz0 = g.escc ? -wl2f * (fx2 * ac * ac + fy2 * fb * fb) : 0.f;
                    const float dzz = g.escc ? -2.f * wl2f * fx2 * ac : 0.f;
                    const float curv = g.escc ? wl2f * fx2 * W * W : 0.f;
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        const float m0 = (float)m[3 * d], m1 = (float)m[3 * d + 1], m2 = (float)m[3 * d + 2];
                        const float Pc = cgrid[d] + m0 * ac + m1 * fb + m2 * zz0;
                        const float Pd = m0 + m2 * dzz;
                        const float mg = fabsf(m2) * curv * 1.01f + 2e-4f * (1.f + fabsf(Pc)) + 0.01f;
                        const float lo_d = bmin[d] - 1.f - mg - Pc, hi_d = bmax[d] + 1.f + mg - Pc;
                        if (fabsf(Pd) < 1e-6f) {
                            if (lo_d > 0.f || hi_d < 0.f) alo = ahi + 1.f;
                        } else {
                            // the margins above exceed the 1-ulp error of the reciprocal
                            const float rp = __builtin_amdgcn_rcpf(Pd);
                            float t0 = lo_d * rp, t1 = hi_d * rp;
                            if (Pd < 0.f) { const float tt = t0; t0 = t1; t1 = tt; }
                            alo = fmaxf(alo, ac + t0);
                            ahi = fminf(ahi, ac + t1);
                        }
                    }
                }
                int a0 = 0, a1 = -1;
                if (alo <= ahi) {
                    a0 = max((int)ceilf(alo - 1e-3f), -g.nxd2);
                    a1 = min((int)floorf(ahi + 1e-3f), g.nx - 1 - g.nxd2);
                }
                const cplx<float>* Kline = spec + ((size_t)r * g.ny + (b + g.nyd2)) * g.nx + g.nxd2;
                double ph[3] = {0.0, 0.0, 0.0};
                if (g.rot_centre_shift) {
                    ph[0] = R.ph[0];
                    ph[1] = R.ph[1];
                    ph[2] = R.ph[2];
                }
                if (R.y_exact) {
                    // No rotation about Z: the point's y grid coordinate is ny/2 + b
                    // exactly (weight 1 on its own plane, 0 on the next), so only the
                    // four (x, z) corners can receive weight.  Same expressions as
                    // grid_pos, so the result is bit-identical to the general path.
                    const int jjg = b + g.nyd2;
                    const int ly = jjg - y0;
                    if (jjg <= g.ny - 2 && ly >= 0 && ly < SBY) {
                        const double bd = (double)b;
                        const double qy = __dmul_rn(g.fyst, bd);
                        const double qy2 = __dmul_rn(qy, qy);
                        // one detector point: position, corner terms, LDS adds
                        auto splat = [&](const int a, const cplx<float> kraw) {
                            const double ad = (double)a;
                            const double qx = __dmul_rn(g.fxst, ad);
                            const double zz = g.escc ? __dmul_rn(-g.wl2, __fma_rn(qx, qx, qy2)) : 0.0;
                            const double ox = __fma_rn(ad, m[0], __fma_rn(bd, m[1], __dmul_rn(zz, m[2])));
                            const double oz = __fma_rn(ad, m[6], __dmul_rn(zz, m[8]));
                            const double dx1 = __dadd_rn((double)g.nxd2, ox);
                            const double dz1 = __dadd_rn((double)g.nzd2, oz);
                            const double fnn = floor(dz1), fii = floor(dx1);
                            if (fnn < 0.0 || fnn > (double)(g.nz - 2)) return;
                            if (fii < 0.0 || fii > (double)(g.nx - 2)) return;
                            const int ii = (int)fii - x0, nn = (int)fnn - z0;
                            if (ii < -1 || ii >= SBX || nn < -1 || nn >= SBZ) return;
                            const double fx = dx1 - fii, fz = dz1 - fnn;
                            const double wxs[2] = {1.0 - fx, fx}, wzs[2] = {1.0 - fz, fz};
                            cplx<float> kv = kraw;
                            if (g.rot_centre_shift) {
                                const double phs = -(ph[0] * ox + ph[1] * bd + ph[2] * oz);
                                double sn, c;
                                sincos(phs, &sn, &c);
                                const float cf = (float)c, sf = (float)sn;
                                kv = {kraw.x * cf - kraw.y * sf, kraw.x * sf + kraw.y * cf};
                            }
                            // the spectrum value on the V3 fixed-point grid (exact: a power of two)
                            const double kxs = (double)kv.x * vscale, kys = (double)kv.y * vscale;
                            long long tr[4], ti[4], tw[4];
                            float wc[4];
                            bool big[4], tny[4];
                            bool any_tiny = false;
#pragma unroll
                            for (int c = 0; c < 4; c++) {
                                const int cx = c & 1, cz = c >> 1;
                                const bool inb = (unsigned)(ii + cx) < (unsigned)SBX &&
                                                 (unsigned)(nn + cz) < (unsigned)SBZ;
                                const double wd = wxs[cx] * 1.0 * wzs[cz];
                                big[c] = inb && wd >= S_TINY_D;
                                tny[c] = inb && wd != 0.0 && wd < S_TINY_D;
                                any_tiny |= tny[c];
                                wc[c] = (float)wd;
                                tr[c] = fx_term(wd, kxs);
                                ti[c] = fx_term(wd, kys);
                                tw[c] = fx_term(wd, wscale);
                            }
                            const int li0 = lds_idx(ii, ly, nn);
#pragma unroll
                            for (int c = 0; c < 4; c++) {
                                if (!big[c]) continue;
                                const int li = li0 + (c & 1) + SPZ * (c >> 1);
                                atomicAdd(&acc[li], (unsigned long long)tr[c]);
                                atomicAdd(&acc[SPAD + li], (unsigned long long)ti[c]);
                                atomicAdd(&acc[2 * SPAD + li], (unsigned long long)tw[c]);
                            }
                            if (any_tiny) {
#pragma unroll
                                for (int c = 0; c < 4; c++) {
                                    const int lx = ii + (c & 1), lz = nn + (c >> 1);
                                    const float w = wc[c];
                                    if (!tny[c]) continue;
                                    const int v = lx + SBX * (ly + SBY * lz);
                                    atomicAdd(&tiny[v], w * kv.x);
                                    atomicAdd(&tiny[SVOX + v], w * kv.y);
                                    atomicAdd(&tiny[2 * SVOX + v], w);
                                }
                            }
                        };
                        // two points in flight without register moves between the
                        // loads and their use (clamped indices keep loads in the line)
                        cplx<float> kA = {0.f, 0.f}, kB = {0.f, 0.f};
                        if (a0 <= a1) {
                            kA = Kline[a0];
                            kB = Kline[min(a0 + 1, a1)];
                        }
                        for (int a = a0; a <= a1; a += 2) {
                            splat(a, kA);
                            kA = Kline[min(a + 2, a1)];
                            if (a + 1 <= a1) splat(a + 1, kB);
                            kB = Kline[min(a + 3, a1)];
                        }
                    }
                } else {
                // the next point's value is loaded while this one is splatted
                cplx<float> knext = {0.f, 0.f};
                if (a0 <= a1) knext = Kline[a0];
                for (int a = a0; a <= a1; a++) {
                    const cplx<float> kraw = knext;
                    if (a < a1) knext = Kline[a + 1];
                    const GridPos P = grid_pos(g, m, a, b);
                    const double fnn = floor(P.dz1), fii = floor(P.dx1), fjj = floor(P.dy1);
                    if (fnn < 0.0 || fnn > (double)(g.nz - 2)) continue;
                    if (fii < 0.0 || fii > (double)(g.nx - 2)) continue;
                    if (fjj < 0.0 || fjj > (double)(g.ny - 2)) continue;
                    const int ii = (int)fii - x0, jj = (int)fjj - y0, nn = (int)fnn - z0;
                    if (ii < -1 || ii >= SBX || jj < -1 || jj >= SBY || nn < -1 || nn >= SBZ) continue;
                    const double fx = P.dx1 - fii, fy = P.dy1 - fjj, fz = P.dz1 - fnn;
                    const double wxs[2] = {1.0 - fx, fx}, wys[2] = {1.0 - fy, fy}, wzs[2] = {1.0 - fz, fz};
                    cplx<float> kv = kraw;
                    if (g.rot_centre_shift) {
                        const double phs = -(ph[0] * P.ox + ph[1] * P.oy + ph[2] * P.oz);
                        double sn, c;
                        sincos(phs, &sn, &c);
                        const float cf = (float)c, sf = (float)sn;
                        kv = {kraw.x * cf - kraw.y * sf, kraw.x * sf + kraw.y * cf};
                    }
                    // the eight corners' terms are formed first, without branches, so
                    // that their independent chains overlap; then each corner inside
                    // the brick adds its terms
                    // the spectrum value on the V3 fixed-point grid (exact: a power of two)
                    const double kxs = (double)kv.x * vscale, kys = (double)kv.y * vscale;
                    long long tr[8], ti[8], tw[8];
                    float wc[8];
                    bool big[8], tny[8];
                    bool any_tiny = false;
#pragma unroll
                    for (int c = 0; c < 8; c++) {
                        const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
                        const bool inb = (unsigned)(ii + cx) < (unsigned)SBX &&
                                         (unsigned)(jj + cy) < (unsigned)SBY &&
                                         (unsigned)(nn + cz) < (unsigned)SBZ;
                        const double wd = wxs[cx] * wys[cy] * wzs[cz];
                        big[c] = inb && wd >= S_TINY_D;
                        tny[c] = inb && wd != 0.0 && wd < S_TINY_D;
                        any_tiny |= tny[c];
                        wc[c] = (float)wd;
                        tr[c] = fx_term(wd, kxs);
                        ti[c] = fx_term(wd, kys);
                        tw[c] = fx_term(wd, wscale);
                    }
                    const int li0 = lds_idx(ii, jj, nn);
#pragma unroll
                    for (int c = 0; c < 8; c++) {
                        if (!big[c]) continue;
                        const int li = li0 + (c & 1) + SPY * ((c >> 1) & 1) + SPZ * (c >> 2);
                        atomicAdd(&acc[li], (unsigned long long)tr[c]);
                        atomicAdd(&acc[SPAD + li], (unsigned long long)ti[c]);
                        atomicAdd(&acc[2 * SPAD + li], (unsigned long long)tw[c]);
                    }
                    if (any_tiny) {
                        // weights below 2^-20: float accumulators (rare)
#pragma unroll
                        for (int c = 0; c < 8; c++) {
                            const int lx = ii + (c & 1), ly = jj + ((c >> 1) & 1), lz = nn + (c >> 2);
                            const float w = wc[c];
                            if (!tny[c]) continue;
                            const int v = lx + SBX * (ly + SBY * lz);
                            atomicAdd(&tiny[v], w * kv.x);
                            atomicAdd(&tiny[SVOX + v], w * kv.y);
                            atomicAdd(&tiny[2 * SVOX + v], w);
                        }
                    }
                }
                }
            }
        }
        __syncthreads();
    }
    const double vinv = ldexp(1.0, -sv), winv = ldexp(1.0, -sw);
#pragma unroll
    for (int k = 0; k < SOWN; k++) {
        const int v = tid + 256 * k;
        const int lx = v % SBX, ly = (v / SBX) % SBY, lz = v / (SBX * SBY);
        const int vx = x0 + lx, vy = y0 + ly, vz = z0 + lz;
        if (vx >= g.nx || vy >= g.ny || vz >= g.nz) continue;
        const int li = lds_idx(lx, ly, lz);
        const float vr = (float)((double)(long long)acc[li] * vinv + (double)tiny[v]);
        const float vi = (float)((double)(long long)acc[SPAD + li] * vinv + (double)tiny[SVOX + v]);
        const float S = (float)((double)(long long)acc[2 * SPAD + li] * winv + (double)tiny[2 * SVOX + v]);
        const size_t vidx = ((size_t)vz * g.ny + vy) * g.nx + vx;
        if (Sout) Sout[vidx] = S;
        Vout[vidx] = voxel_epilogue<float>(g, vx, vy, vz, cplx<float>{vr, vi}, S);
    }
}

// ------------- FP32 scatter into y-plane tiles (rows without Z rotation)
//
// When no row rotates about Z, a detector line b reaches exactly one y plane
// of the volume (j = ny/2 + b, weight 1) and its points move in that plane's
// (x, z) coordinates.  Each workgroup owns a PX x PZ tile of a group of
// consecutive y planes and takes, from every row, the one line that lands in
// each plane.  The lines of neighbouring planes differ only through the
// Ewald term of b, so one clip of every row's line to the tile, conservative
// over the group's b range, serves all planes of the group: the clipped
// lines are listed in row order, each wave takes a run of the listed points
// and walks it 64 consecutive points at a time (lane l on the chunk's point
// l, so the lanes read consecutive values of one or a few lines), and the
// planes are then gridded one after another with the same list and the same
// runs.  Point positions, corner weights and the fixed-point sums are those
// of k_grid_fx (the same expressions in the same order), so both kernels
// produce the same V3 and Samp3 sums.  Groups are numbered contiguously per
// XCD, so the planes of a group share one XCD's L2.

constexpr int PX = 32, PZ = 32;
constexpr int PVOX = PX * PZ;
constexpr int POWN = PVOX / 256;
static_assert(PVOX % 256 == 0, "tile must be a multiple of the workgroup");
// LDS sums of a tile with a one-cell halo: cell (x, z), x and z in
// [-1, PX] x [-1, PZ], sits at x + PSX z + P_OFF.  The stride is odd (lanes
// walking along z spread over the banks) and one short of the halo's width,
// so the halo cells x = -1 and x = PX of neighbouring rows share a slot; no
// tile cell shares one.  Every corner of a point whose cell touches the tile
// then has a slot, and corners outside the tile add to the halo, which is
// never read: the corners need no bounds tests.
constexpr int PSX = PX + 1;          // LDS stride along z
constexpr int P_OFF = PSX + 1;       // slot of cell (0, 0)
constexpr int PPAD = PX + PSX * PZ + P_OFF + 1;  // LDS entries per component
// rows clipped per chunk: the list capacity PL of k_grid_yp (384 for up to
// 384 rows, two per thread; 768 beyond, three per thread), so jobs of up to
// 768 rows list each group's lines once
constexpr int P_LIST_S = 384, P_LIST_L = 768;
constexpr int P_TQ = 256;            // points with tiny corner weights recorded per plane (one per thread)

// Fixed-point scales 2^sv (V3) and 2^sw (Samp3) from the largest spectrum
// component and the row count.  Per row a voxel receives a total weight
// below 8, so |sum| < 8 nrows max|K| 2^sv < 2^62, and each term stays below
// 2^51 for fx_term.  Ends with a barrier.
__device__ __forceinline__ void fx_scales(const unsigned* __restrict__ rowmax, int nrows,
                                          float* s_red, int& sv, int& sw) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    float mk = 0.f;
    for (int r = tid; r < nrows; r += blockDim.x) mk = fmaxf(mk, __uint_as_float(rowmax[r]));
    for (int off = 32; off > 0; off >>= 1) mk = fmaxf(mk, __shfl_xor(mk, off));
    if (lane == 0) s_red[wave] = mk;
    __syncthreads();
    mk = fmaxf(fmaxf(s_red[0], s_red[1]), fmaxf(s_red[2], s_red[3]));
    int ek = 0, ea = 0;
    frexpf(mk > 0.f ? mk : 1.f, &ek);    // max|K| < 2^ek
    frexpf((float)(8 * nrows), &ea);      // 8 nrows < 2^ea
    sv = min(61 - ek - ea, 50 - ek);
    sw = min(61 - ea, 50);
}

// Clip one row's detector line b in [fb_lo, fb_hi] to the x range
// [bmin0, bmax0] and z range [bmin1, bmax1] of a tile: the a interval
// [a0, a1] outside which no point of any of these lines can put a corner
// in the tile.  x(a) and z(a) are linear in a up to the Ewald term
// m2*zz(a), whose deviation from its tangent at the interval centre is at
// most wl2*fx2*W^2; two passes, the second about the first's interval.  The
// spread of the lines over b enters through m1*b and through fy2*b^2 in zz.
// Returns the number of points (0: the line misses the tile).
__device__ __forceinline__ int clip_line(const GeomDev& g, const RowParams& R, float fbm, float fbh,
                                         float fb2m, float fb2h, const float bmin[2],
                                         const float bmax[2], float fx2, float fy2, float wl2f,
                                         int& a0) {
    const float cgrid[2] = {(float)g.nxd2, (float)g.nzd2};
    float alo = (float)(-g.nxd2), ahi = (float)(g.nx - 1 - g.nxd2);
    for (int it = 0; it < 2 && alo <= ahi; it++) {
        const float ac = 0.5f * (alo + ahi), W = 0.5f * (ahi - alo);
        const float zz0 = g.escc ? -wl2f * (fx2 * ac * ac + fy2 * fb2m) : 0.f;
        const float dzz = g.escc ? -2.f * wl2f * fx2 * ac : 0.f;
        const float curv = g.escc ? wl2f * fx2 * W * W : 0.f;
        const float bspr = g.escc ? wl2f * fy2 * fb2h : 0.f;
#pragma unroll
        for (int d = 0; d < 2; d++) {
            const int rd = d ? 6 : 0;
            const float m0 = (float)R.m[rd], m1 = (float)R.m[rd + 1], m2 = (float)R.m[rd + 2];
            const float Pc = cgrid[d] + m0 * ac + (d ? 0.f : m1 * fbm) + m2 * zz0;
            const float Pd = m0 + m2 * dzz;
            const float mg = fabsf(m2) * (curv * 1.01f + bspr * 1.01f) + (d ? 0.f : fabsf(m1) * fbh) +
                             2e-4f * (1.f + fabsf(Pc)) + 0.01f;
            const float lo_d = bmin[d] - 1.f - mg - Pc, hi_d = bmax[d] + 1.f + mg - Pc;
            if (fabsf(Pd) < 1e-6f) {
                if (lo_d > 0.f || hi_d < 0.f) alo = ahi + 1.f;
            } else {
                // the margins above exceed the 1-ulp error of the reciprocal
                const float rp = __builtin_amdgcn_rcpf(Pd);
                float t0 = lo_d * rp, t1 = hi_d * rp;
                if (Pd < 0.f) { const float tt = t0; t0 = t1; t1 = tt; }
                alo = fmaxf(alo, ac + t0);
                ahi = fminf(ahi, ac + t1);
            }
        }
    }
    if (!(alo <= ahi)) return 0;
    a0 = max((int)ceilf(alo - 1e-3f), -g.nxd2);
    const int a1 = min((int)floorf(ahi + 1e-3f), g.nx - 1 - g.nxd2);
    return a0 <= a1 ? a1 - a0 + 1 : 0;
}

// Launch size of k_grid_yp: each XCD takes ceil(ngroups / 8) consecutive
// plane groups, all tiles of them.
__host__ __device__ constexpr unsigned yp_grid(int ntiles, int ngroups) {
    return 8u * (unsigned)((ngroups + 7) / 8) * (unsigned)ntiles;
}

// ROT: rows carry the centre-shift phase (compiled separately: its FP64
// sincos would otherwise set the register budget of the common case).
// Workgroup (tile, group) grids planes pg*group .. pg*group + pg - 1; ntiles
// tiles per plane, ngroups groups.  Workgroups are dealt to the XCDs round
// robin (workgroup L on XCD L % 8) and start in launch order, so each XCD's
// workgroups are ordered by tile rank first: tile_order lists the tiles by
// decreasing expected point count (the lines of all rows cross near the
// plane's centre), and the heaviest tiles start first.
//
// rowc: the five map coefficients the points use, per row (m0 m1 m2 m6 m8,
// padded to 6 doubles): a compact table that stays in the L1 cache.
// GATE: the voxel epilogue is the noise gate alone (no 3D filter, mode
// other than 1, nonzero NSR, not the sCTF): a straight-line instantiation
// of voxel_epilogue's noise-gate branch (the same operations).
//
// Corner weights below 2^-20 (a point within 2^-20 of a grid plane) keep
// their relative precision outside the fixed-point sums: the plane's pass
// adds the fixed-point terms of all other corners and only flags such a
// corner.  A flagged plane then converts its sums to FP64 in place and adds
// just those corners' FP32 terms in FP64: those of the points it recorded
// (up to P_TQ), or all of them in a second pass when more were met.  The
// LDS then holds only the fixed-point sums (with their halo), the list and
// the record (31 KB), and the noise-gate instantiation fits 96 VGPRs: 5
// workgroups per CU.
template <bool B>
struct PassKind {
    static constexpr bool value = B;
};

template <bool ROT, bool GATE, int PL>
__global__ void __launch_bounds__(256, ROT ? 1 : (GATE && PL <= P_LIST_S ? 5 : 4)) k_grid_yp(const cplx<float>* __restrict__ spec,
                                                 const RowParams* __restrict__ rows,
                                                 const double* __restrict__ rowc, int nrows,
                                                 const unsigned* __restrict__ rowmax, GeomDev g,
                                                 cplx<float>* __restrict__ Vout,
                                                 float* __restrict__ Sout, int ntiles, int ngroups,
                                                 int pg, const int* __restrict__ tile_order, int gblk) {
    __shared__ __align__(16) unsigned long long acc[3 * PPAD];  // fixed-point sums (FP64 after a tiny-weight pass)
    constexpr int RPT = (PL + 255) / 256;         // rows clipped per thread
    __shared__ int s_line[PL];                    // listed lines: row << 16 | first a of the clipped line
    __shared__ int s_pre[PL + 1];                 // exclusive prefix of the point counts
    __shared__ int s_wsum[4];
    __shared__ int s_tq[P_TQ];                    // points with a corner weight below 2^-20 (row << 16 | a)
    __shared__ int s_tn;                          // their count (may exceed P_TQ)
    __shared__ float s_red[4];
    double* const accd = reinterpret_cast<double*>(acc);

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // each XCD's groups are taken gblk at a time (fewer planes in flight per
    // L2), and within such a block by tile rank first
    const int ngx = (ngroups + 7) >> 3;  // groups per XCD
    const int lidx = (int)(blockIdx.x >> 3);
    const int blk = lidx / (gblk * ntiles), in = lidx - blk * gblk * ntiles;
    const int nb = min(gblk, ngx - blk * gblk);  // groups in this block
    const int rank = in / max(nb, 1);
    const int grp = (int)(blockIdx.x & 7u) * ngx + blk * gblk + (in - rank * nb);
    if (nb <= 0 || rank >= ntiles || grp >= ngroups || blk * gblk + (in - rank * nb) >= ngx) return;
    const int tile = tile_order[rank];
    const int ntx = (g.nx + PX - 1) / PX;
    const int x0 = (tile % ntx) * PX, z0 = (tile / ntx) * PZ;
    const int jfirst = grp * pg, jend = min(jfirst + pg, g.ny);
    // planes that receive weight: j <= ny - 2 (the last plane's cells would
    // need plane ny)
    const int jlast_w = min(jend, g.ny - 1) - 1;
    int sv, sw;  // fixed-point scales of the job (the same in every workgroup)
    fx_scales(rowmax, nrows, s_red, sv, sw);
    const double vscale = ldexp(1.0, sv), wscale = ldexp(1.0, sw);
    const double vinv = ldexp(1.0, -sv), winv = ldexp(1.0, -sw);
    const double xmaxd = (double)(g.nx - 2), zmaxd = (double)(g.nz - 2);

    const float fx2 = (float)(g.fxst * g.fxst), fy2 = (float)(g.fyst * g.fyst);
    const float wl2f = (float)g.wl2;
    const float bmin[2] = {(float)x0, (float)z0};
    const float bmax[2] = {(float)(x0 + PX - 1), (float)(z0 + PZ - 1)};
    // the group's detector lines b in [blo, bhi]
    const int blo = jfirst - g.nyd2, bhi = jlast_w - g.nyd2;
    const float fbm = 0.5f * (float)(blo + bhi), fbh = 0.5f * (float)(bhi - blo);
    const float b2lo = (blo <= 0 && bhi >= 0) ? 0.f : fminf((float)blo * blo, (float)bhi * bhi);
    const float b2hi = fmaxf((float)blo * blo, (float)bhi * bhi);
    const float fb2m = 0.5f * (b2lo + b2hi), fb2h = 0.5f * (b2hi - b2lo);
    const unsigned rstride = (unsigned)(g.ny * g.nx) & 0xffffffu;  // ny nx <= 2^18

    const int nchunks = (nrows + PL - 1) / PL;
    int n = 0, total = 0, q0 = 0, p0 = 0, p1 = 0;
    for (int j = jfirst; j < jend; j++) {
        for (int t = tid; t < 3 * PPAD / 2; t += 256)
            reinterpret_cast<ulonglong2*>(acc)[t] = ulonglong2{0ull, 0ull};
        if (tid == 0) {
            if (3 * PPAD % 2) acc[3 * PPAD - 1] = 0ull;
            s_tn = 0;
        }
        __syncthreads();  // LDS zeroed (and the previous plane's epilogue done)
        const int b = j - g.nyd2;
        const double bd = (double)b;
        const double qy = __dmul_rn(g.fyst, bd);
        const double qy2 = __dmul_rn(qy, qy);
        const cplx<float>* spec_j = spec + (size_t)j * g.nx + g.nxd2;
        // One point of row r: a, its spectrum value, the row's coefficients.
        // FIX false: the fixed-point terms of every in-tile corner of weight
        // >= 2^-20, and the point recorded when a smaller nonzero one is in
        // the tile.  FIX true: only those small corners, their FP32 terms
        // added in FP64 to the converted sums.
        auto splat = [&](auto kind, const int r, const int a, const cplx<float> kraw, const double m0,
                         const double m1, const double m2, const double m6, const double m8,
                         const double* ph) {
            constexpr bool FIX = decltype(kind)::value;
            const double ad = (double)a;
            const double qx = __dmul_rn(g.fxst, ad);
            const double zz = g.escc ? __dmul_rn(-g.wl2, __fma_rn(qx, qx, qy2)) : 0.0;
            const double ox = __fma_rn(ad, m0, __fma_rn(bd, m1, __dmul_rn(zz, m2)));
            const double oz = __fma_rn(ad, m6, __dmul_rn(zz, m8));
            const double dx1 = __dadd_rn(g.d_nxd2, ox);
            const double dz1 = __dadd_rn(g.d_nzd2, oz);
            const double fnn = floor(dz1), fii = floor(dx1);
            if (fnn < 0.0 || fnn > zmaxd) return;
            if (fii < 0.0 || fii > xmaxd) return;
            const int ii = (int)fii - x0, nn = (int)fnn - z0;
            if (ii < -1 || ii >= PX || nn < -1 || nn >= PZ) return;
            const double fx = dx1 - fii, fz = dz1 - fnn;
            // with both 1D weights of each axis at least 2^-10 every
            // corner weight is at least 2^-20: the per-corner tests
            // only run when a fraction lies within 2^-10 of a grid plane
            const bool near = !(fx >= S_TINY_1D && fx <= 1.0 - S_TINY_1D &&
                                fz >= S_TINY_1D && fz <= 1.0 - S_TINY_1D);
            if (FIX && !near) return;
            const double wxs[2] = {1.0 - fx, fx}, wzs[2] = {1.0 - fz, fz};
            cplx<float> kv = kraw;
            if (ROT) {
                const double phs = -(ph[0] * ox + ph[1] * bd + ph[2] * oz);
                double sn, c;
                sincos(phs, &sn, &c);
                const float cf = (float)c, sf = (float)sn;
                kv = {kraw.x * cf - kraw.y * sf, kraw.x * sf + kraw.y * cf};
            }
            const int li0 = ii + PSX * nn + P_OFF;
            if (FIX) {
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    const int cx = c & 1, cz = c >> 1;
                    const double wd = wxs[cx] * 1.0 * wzs[cz];
                    const bool inb = (unsigned)(ii + cx) < (unsigned)PX &&
                                     (unsigned)(nn + cz) < (unsigned)PZ;
                    if (!(inb && wd != 0.0 && wd < S_TINY_D)) continue;
                    const int li = li0 + cx + PSX * cz;
                    const float w = (float)wd;
                    atomicAdd(&accd[li], (double)(w * kv.x));
                    atomicAdd(&accd[PPAD + li], (double)(w * kv.y));
                    atomicAdd(&accd[2 * PPAD + li], (double)w);
                }
                return;
            }
            double wds[4];
#pragma unroll
            for (int c = 0; c < 4; c++) wds[c] = wxs[c & 1] * 1.0 * wzs[c >> 1];
            if (near) {
                // weights below 2^-20 leave the fixed-point sums (their terms
                // become 0); the point is recorded when such a corner is in
                // the tile
                bool any_tiny = false;
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    const bool tny = wds[c] != 0.0 && wds[c] < S_TINY_D;
                    any_tiny |= tny && (unsigned)(ii + (c & 1)) < (unsigned)PX &&
                                (unsigned)(nn + (c >> 1)) < (unsigned)PZ;
                    if (tny) wds[c] = 0.0;
                }
                if (any_tiny) {
                    const int slot = atomicAdd(&s_tn, 1);
                    if (slot < P_TQ) s_tq[slot] = (int)(((unsigned)r << 16) | ((unsigned)a & 0xffffu));
                }
            }
            // the spectrum value on the V3 fixed-point grid (exact: a power of
            // two); all 4 corners add (corners outside the tile to the halo,
            // zero weights a zero term)
            const double kxs = (double)kv.x * vscale, kys = (double)kv.y * vscale;
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const int li = li0 + (c & 1) + PSX * (c >> 1);
                atomicAdd(&acc[li], (unsigned long long)fx_term(wds[c], kxs));
                atomicAdd(&acc[PPAD + li], (unsigned long long)fx_term(wds[c], kys));
                atomicAdd(&acc[2 * PPAD + li], (unsigned long long)fx_term(wds[c], wscale));
            }
        };
        // One pass of splat over all the plane's listed points.
        auto pass = [&](auto kind) {
            constexpr bool FIX = decltype(kind)::value;
            for (int base = 0; base < nrows && j <= jlast_w; base += PL) {
                if (nchunks > 1 || (j == jfirst && !FIX)) {
                    // list the chunk's rows whose lines reach the tile: RPT rows
                    // per thread (RPT tid + h), one block scan of the packed
                    // (line count << 20 | point count)
                    const int nchunk = min(PL, nrows - base);
                    int npt[RPT] = {}, a0s[RPT] = {};
#pragma unroll
                    for (int h = 0; h < RPT; h++) {
                        const int ri = RPT * tid + h;
                        if (ri < nchunk)
                            npt[h] = clip_line(g, rows[base + ri], fbm, fbh, fb2m, fb2h, bmin, bmax, fx2,
                                               fy2, wl2f, a0s[h]);
                    }
                    int nl = 0, np = 0;
#pragma unroll
                    for (int h = 0; h < RPT; h++) {
                        nl += (int)(npt[h] > 0);
                        np += npt[h];
                    }
                    const int v = nl << 20 | np;
                    int incl = v;
                    for (int o = 1; o < 64; o <<= 1) {
                        const int u = __shfl_up(incl, o);
                        if (lane >= o) incl += u;
                    }
                    if (lane == 63) s_wsum[wave] = incl;
                    __syncthreads();
                    int woff = 0;
                    for (int w = 0; w < wave; w++) woff += s_wsum[w];
                    const int tot = s_wsum[0] + s_wsum[1] + s_wsum[2] + s_wsum[3];
                    const int ex = woff + incl - v;
                    int off = ex >> 20, pre = ex & 0xfffff;
#pragma unroll
                    for (int h = 0; h < RPT; h++) {
                        if (npt[h] > 0) {
                            s_line[off] = (int)(((unsigned)(base + RPT * tid + h) << 16) |
                                                ((unsigned)a0s[h] & 0xffffu));
                            s_pre[off] = pre;
                            off++;
                            pre += npt[h];
                        }
                    }
                    n = tot >> 20;
                    total = tot & 0xfffff;
                    if (tid == 0) s_pre[n] = total;
                    __syncthreads();
                    // each wave takes a run of whole 64-point chunks of the listed
                    // points (the runs differ by at most one chunk, and only the
                    // last chunk of the list is partial), in list order; q0: the
                    // listed line holding its first point
                    const int nch = (total + 63) >> 6;
                    p0 = min(total, 64 * ((wave * nch) >> 2));
                    p1 = min(total, 64 * (((wave + 1) * nch) >> 2));
                    if (p0 < p1) {
                        int lo = 0, hi = n - 1;
                        while (lo < hi) {
                            const int mid = (lo + hi + 1) >> 1;
                            if (s_pre[mid] <= p0) lo = mid;
                            else hi = mid - 1;
                        }
                        q0 = lo;
                    }
                }
                // the wave's run, 64 consecutive points at a time (lane l takes
                // the chunk's point l): the lanes read consecutive values of
                // one or a few lines and add to neighbouring cells
                int qc = __builtin_amdgcn_readfirstlane(q0);
                const int pw1 = __builtin_amdgcn_readfirstlane(p1);
                for (int pb = __builtin_amdgcn_readfirstlane(p0); pb < pw1; pb += 64) {
                    const int p = pb + lane;
                    int q = qc, pre = s_pre[qc], nxt = s_pre[qc + 1];
                    if (p < pw1) {
                        while (nxt <= p) {
                            q++;
                            pre = nxt;
                            nxt = s_pre[q + 1];
                        }
                        const int e = s_line[q];
                        const int r = (int)((unsigned)e >> 16);
                        const int a = (int)(short)(e & 0xffff) + (p - pre);
                        const double* c6 = rowc + 6 * r;
                        double ph[3] = {0.0, 0.0, 0.0};
                        if (ROT) {
                            ph[0] = rows[r].ph[0];
                            ph[1] = rows[r].ph[1];
                            ph[2] = rows[r].ph[2];
                        }
                        const cplx<float> kraw = spec_j[(size_t)r * rstride + a];
                        splat(kind, r, a, kraw, c6[0], c6[1], c6[2], c6[3], c6[4], ph);
                    }
                    qc = __builtin_amdgcn_readlane(q, 63);  // the line of the chunk's last point
                }
                __syncthreads();  // the sums of the chunk are complete; its list may be replaced
            }
        };
        pass(PassKind<false>{});
        // an opaque copy of the thread index: the voxel terms below are then
        // evaluated per plane instead of being kept live (and spilled)
        // across the plane loop
        int vt = tid;
        asm volatile("" : "+v"(vt));
        const int tn = __builtin_amdgcn_readfirstlane(s_tn);  // same in all lanes
        const bool fixd = tn != 0;
        if (fixd) {
            // each thread converts its own voxels' sums to FP64 (the values
            // the epilogue would form), then the small corners are added:
            // those of the recorded points, or a second pass over all points
            // when more were met than the record holds
#pragma unroll
            for (int k = 0; k < POWN; k++) {
                const int v = vt + 256 * k;
                const int li = (v % PX) + PSX * (v / PX) + P_OFF;
#pragma unroll
                for (int m = 0; m < 3; m++) {
                    const double d = (double)(long long)acc[m * PPAD + li] * (m < 2 ? vinv : winv);
                    accd[m * PPAD + li] = d;
                }
            }
            __syncthreads();
            if (tn <= P_TQ) {
                if (tid < tn) {
                    const int e = s_tq[tid];
                    const int r = (int)((unsigned)e >> 16), a = (int)(short)(e & 0xffff);
                    const double* c6 = rowc + 6 * r;
                    double ph[3] = {0.0, 0.0, 0.0};
                    if (ROT) {
                        ph[0] = rows[r].ph[0];
                        ph[1] = rows[r].ph[1];
                        ph[2] = rows[r].ph[2];
                    }
                    const cplx<float> kraw = spec_j[(size_t)r * rstride + a];
                    splat(PassKind<true>{}, r, a, kraw, c6[0], c6[1], c6[2], c6[3], c6[4], ph);
                }
                __syncthreads();
            } else {
                pass(PassKind<true>{});
            }
        }
        // voxel (x0 + vt % PX, j, z0 + vt / PX + k 256 / PX) for k < POWN
        const size_t kstride = (size_t)(256 / PX) * g.ny * g.nx;
        const size_t vidx0 = ((size_t)(z0 + vt / PX) * g.ny + j) * g.nx + x0 + vt % PX;
#pragma unroll
        for (int k = 0; k < POWN; k++) {
            const int v = vt + 256 * k;
            const int lx = v % PX, lz = v / PX;
            const int vx = x0 + lx, vz = z0 + lz;
            if (vx >= g.nx || vz >= g.nz) continue;
            const int li = lx + PSX * lz + P_OFF;
            float vr, vi, S;
            if (fixd) {
                vr = (float)accd[li];
                vi = (float)accd[PPAD + li];
                S = (float)accd[2 * PPAD + li];
            } else {
                vr = (float)((double)(long long)acc[li] * vinv);
                vi = (float)((double)(long long)acc[PPAD + li] * vinv);
                S = (float)((double)(long long)acc[2 * PPAD + li] * winv);
            }
            const size_t vidx = vidx0 + k * kstride;
            if (Sout) Sout[vidx] = S;
            if (GATE) {
                // voxel_epilogue's noise-gate branch (UTR.cpp:1050-1128)
                cplx<float> o = {0.f, 0.f};
                if (S != 0.f) {
                    const float fnsr = (float)(g.nsr2 * (double)vsqrt(S));
                    const float y = vsqrt(vr * vr + vi * vi) * recip_nr(fnsr);
                    const float f = recip_nr(S) * (0.5f + atanf(100.f * (y - (float)kTargetSNR)) *
                                                           0.318309886183790671538f);
                    o = {vr * f, vi * f};
                }
                Vout[vidx] = o;
            } else {
                Vout[vidx] = voxel_epilogue<float>(g, vx, j, vz, cplx<float>{vr, vi}, S);
            }
        }
        __syncthreads();  // the epilogue has read the sums before the next plane zeroes them
    }
}

// Zero fill of a small counter array (kept out of hipMemsetAsync so that
// launch sequences containing it can be captured into a graph).
__global__ void k_fill_u32(unsigned* __restrict__ p, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0u;
}

// ------------------------------------------- plain FP64 form: scatter

// One thread per spectrum point of one row; atomics into V (interleaved
// re/im) and S, as UTR.h:230-305 does from its OpenMP threads.
__global__ void __launch_bounds__(256) k_scatter(const cplx<double>* __restrict__ K,
                                                 const RowParams* __restrict__ rows, int row,
                                                 GeomDev g, double* __restrict__ V,
                                                 double* __restrict__ S) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.nx * g.ny) return;
    const int j = p / g.nx, i = p - j * g.nx;
    const RowParams& R = rows[row];
    const GridPos P = grid_pos(g, R.m, i - g.nxd2, j - g.nyd2);
    const double fnn = floor(P.dz1), fii = floor(P.dx1), fjj = floor(P.dy1);
    if (fnn < 0.0 || fnn > (double)(g.nz - 2)) return;
    if (fii < 0.0 || fii > (double)(g.nx - 2)) return;
    if (fjj < 0.0 || fjj > (double)(g.ny - 2)) return;
    const long ii = (long)fii, jj = (long)fjj, nn = (long)fnn;
    const double fx = P.dx1 - fii, fy = P.dy1 - fjj, fz = P.dz1 - fnn;
    const double wxs[2] = {1.0 - fx, fx}, wys[2] = {1.0 - fy, fy}, wzs[2] = {1.0 - fz, fz};
    cplx<double> kv = K[p];
    if (g.rot_centre_shift) {
        const double ph = -(R.ph[0] * P.ox + R.ph[1] * P.oy + R.ph[2] * P.oz);
        double sn, c;
        sincos(ph, &sn, &c);
        kv = {kv.x * c - kv.y * sn, kv.x * sn + kv.y * c};
    }
    const long nxny = (long)g.nx * g.ny;
    for (int cz = 0; cz < 2; cz++)
        for (int cy = 0; cy < 2; cy++)
            for (int cx = 0; cx < 2; cx++) {
                const double w = wxs[cx] * wys[cy] * wzs[cz];
                const long idx = (nn + cz) * nxny + (jj + cy) * g.nx + (ii + cx);
                atomicAdd(&S[idx], w);
                atomicAdd(&V[2 * idx], w * kv.x);
                atomicAdd(&V[2 * idx + 1], w * kv.y);
            }
}

template <typename T>
__global__ void __launch_bounds__(256) k_normalize(cplx<T>* __restrict__ V,
                                                   const T* __restrict__ S, GeomDev g) {
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = (int)(t % g.nx), j = (int)((t / g.nx) % g.ny), k = (int)(t / ((size_t)g.nx * g.ny));
    V[t] = voxel_epilogue<T>(g, i, j, k, V[t], S[t]);
}

// ------------------------------------------------------ 3D inverse FFT

constexpr int CW_MAX = 16;

// Row stride of the half spectrum H: nx/2 + 1 entries rounded up to 8, so
// that every row starts on a 64-byte boundary and the column segments the 3D
// passes move stay whole cache lines.
__host__ __device__ constexpr int h_stride(int nx) { return (nx / 2 + 1 + 7) & ~7; }

// Hermitian combination of V and its mirror (the symmetrisation of
// UTR.cpp:1203-1217 followed by the half the c2r plan consumes), checkerboard
// (-1)^(kz+ky+w) for the centred z/y indices, and inverse FFT along z.
// Nyquist x-plane (centred i = 0): the serial loop order of the reference
// gives H = (3a + 5 conj b)/8 when (k,j) precedes its mirror, (5a + 3 conj b)/8
// when it follows it, and Re(a) when it is its own mirror.
template <typename T, int NZ>
__global__ void __launch_bounds__(256) k_pack_z(const cplx<T>* __restrict__ V,
                                                cplx<T>* __restrict__ H, GeomDev g,
                                                const cplx<T>* __restrict__ tw, int gx, int gy) {
    constexpr int CW = lds_batch(NZ, sizeof(cplx<T>), LDS_PACKZ);
    __shared__ cplx<T> sm[2 * lds_len(CW * NZ)];
    cplx<T>* a = sm;
    cplx<T>* b = sm + lds_len(CW * NZ);
    const int nx = g.nx, ny = g.ny, W = nx / 2 + 1, WP = h_stride(nx);
    int tx, ty;
    if (!xcd_tile(gx, gy, tx, ty)) return;
    const int j = ty;
    const int w0 = tx * CW;
    const int nb = min(CW, W - w0);
    const int ja = (j == 0) ? 0 : ny - j;
    for (int t = threadIdx.x; t < nb * NZ; t += blockDim.x) {
        const int k = t / nb, c = t - k * nb;
        const int w = w0 + c;
        const int i = (w == nx / 2) ? 0 : w + g.nxd2;
        const int ia = (i == 0) ? 0 : nx - i;
        const int ka = (k == 0) ? 0 : NZ - k;
        const cplx<T> va = V[((size_t)k * ny + j) * nx + i];
        const cplx<T> vb = V[((size_t)ka * ny + ja) * nx + ia];
        cplx<T> h;
        if (i != 0) {
            h = {T(0.5) * (va.x + vb.x), T(0.5) * (va.y - vb.y)};
        } else if (k == ka && j == ja) {
            h = {va.x, T(0)};
        } else if (k < ka || (k == ka && j < ja)) {
            h = {(T(3) * va.x + T(5) * vb.x) * T(0.125), (T(3) * va.y - T(5) * vb.y) * T(0.125)};
        } else {
            h = {(T(5) * va.x + T(3) * vb.x) * T(0.125), (T(5) * va.y - T(3) * vb.y) * T(0.125)};
        }
        if ((k + j + w) & 1) {
            h.x = -h.x;
            h.y = -h.y;
        }
        a[lx(c * NZ + k)] = h;
    }
    __syncthreads();
    fft_lds<T, NZ>(a, b, nb, tw, +1, threadIdx.x, blockDim.x);
    for (int t = threadIdx.x; t < nb * NZ; t += blockDim.x) {
        const int k = t / nb, c = t - k * nb;
        H[((size_t)k * ny + j) * WP + w0 + c] = a[lx(c * NZ + k)];
    }
}

// Inverse FFT along y of the half spectrum, in place.
template <typename T, int NY>
__global__ void __launch_bounds__(256) k_fft_y(cplx<T>* __restrict__ H, GeomDev g,
                                               const cplx<T>* __restrict__ tw, int gx, int gy) {
    constexpr int CW = lds_batch(NY, sizeof(cplx<T>), LDS_FFTY);
    __shared__ cplx<T> sm[2 * lds_len(CW * NY)];
    cplx<T>* a = sm;
    cplx<T>* b = sm + lds_len(CW * NY);
    const int W = g.nx / 2 + 1, WP = h_stride(g.nx);
    int tx, ty;
    if (!xcd_tile(gx, gy, tx, ty)) return;
    const int k = ty;
    const int w0 = tx * CW;
    const int nb = min(CW, W - w0);
    cplx<T>* base = H + (size_t)k * NY * WP + w0;
    for (int t = threadIdx.x; t < nb * NY; t += blockDim.x) {
        const int y = t / nb, c = t - y * nb;
        a[lx(c * NY + y)] = base[(size_t)y * WP + c];
    }
    __syncthreads();
    fft_lds<T, NY>(a, b, nb, tw, +1, threadIdx.x, blockDim.x);
    for (int t = threadIdx.x; t < nb * NY; t += blockDim.x) {
        const int y = t / nb, c = t - y * nb;
        base[(size_t)y * WP + c] = a[lx(c * NY + y)];
    }
}

// Complex-to-real inverse FFT along x of two rows at once (z = x1 + i x2),
// then (-1)^(k+j+(nz+ny)/2)/N, sphere mask, cupping compensation and final
// scale (UTR.cpp:1235-1279).
template <typename T, int NX, typename TO>
__global__ void __launch_bounds__(256) k_c2r_x(const cplx<T>* __restrict__ H,
                                               TO* __restrict__ vol, GeomDev g,
                                               const cplx<T>* __restrict__ tw, int gx, int gy) {
    constexpr int NB = lds_batch(NX, sizeof(cplx<T>), LDS_C2R);
    __shared__ cplx<T> sm[2 * lds_len(NB * NX)];
    cplx<T>* a = sm;
    cplx<T>* b = sm + lds_len(NB * NX);
    const int ny = g.ny, WP = h_stride(NX);
    int tx, ty;
    if (!xcd_tile(gx, gy, tx, ty)) return;
    const int k = ty;
    const int p0 = tx * NB;  // first row pair
    const int np = min(NB, ny / 2 - p0);
    for (int t = threadIdx.x; t < np * NX; t += blockDim.x) {
        const int pp = t / NX, w = t - pp * NX;
        const int j = 2 * (p0 + pp);
        const cplx<T>* r1 = H + ((size_t)k * ny + j) * WP;
        const cplx<T>* r2 = r1 + WP;
        cplx<T> x1, x2;
        if (w <= NX / 2) {
            x1 = r1[w];
            x2 = r2[w];
        } else {
            x1 = r1[NX - w];
            x2 = r2[NX - w];
            x1.y = -x1.y;
            x2.y = -x2.y;
        }
        a[lx(t)] = {x1.x - x2.y, x1.y + x2.x};
    }
    __syncthreads();
    fft_lds<T, NX>(a, b, np, tw, +1, threadIdx.x, blockDim.x);
    const double invN = 1.0 / ((double)NX * ny * g.nz);
    const int gs = par_sign((g.nz + ny) / 2);
    // The cupping factor exp(dtemp*(di^2 + dj^2 + dk^2)) is the product of one
    // factor per axis: x factors go to the free FFT buffer, the block's 2*NB
    // y factors to a small table, and the z factor is one value per block.
    __shared__ double s_ey[2 * NB];
    double* s_ex = reinterpret_cast<double*>(b);
    static_assert(sizeof(double) <= NB * sizeof(cplx<T>), "x factors must fit the FFT scratch buffer");
    if (!g.no_mask) {
        for (int m = threadIdx.x; m < NX; m += blockDim.x)
            s_ex[m] = exp(g.mask_dtemp * (double)((long)(m - g.ic) * (m - g.ic)));
        if (threadIdx.x < 2 * np) {
            const int jj = 2 * p0 + threadIdx.x;
            s_ey[threadIdx.x] = g.mask_j ? exp(g.mask_dtemp * (double)((long)(jj - g.jc) * (jj - g.jc))) : 1.0;
        }
        __syncthreads();
    }
    const double ez = (g.mask_k && !g.no_mask) ? exp(g.mask_dtemp * (double)((long)(k - g.kc) * (k - g.kc))) : 1.0;
    const double base_f = g.fnorm * invN * (double)gs;
    for (int t = threadIdx.x; t < np * NX; t += blockDim.x) {
        const int pp = t / NX, m = t - pp * NX;
        const int j = 2 * (p0 + pp);
        const cplx<T> z = a[lx(t)];
        for (int h = 0; h < 2; h++) {
            const int jj = j + h;
            const T val = h ? z.y : z.x;
            long ii2 = (long)(m - g.ic) * (m - g.ic);
            if (g.mask_j) ii2 += (long)(jj - g.jc) * (jj - g.jc);
            if (g.mask_k) ii2 += (long)(k - g.kc) * (k - g.kc);
            const double sgn = (double)par_sign(k + jj);
            TO o;
            if (g.no_mask) {
                o = TO((double)val * (base_f * sgn));
            } else if (ii2 > g.nR2) {
                o = TO(0);
            } else {
                const double f = s_ex[m] * s_ey[2 * pp + h] * ez * base_f * sgn;
                o = TO((double)val * f);
            }
            vol[((size_t)k * ny + jj) * NX + m] = o;
        }
    }
}

}  // namespace utr
