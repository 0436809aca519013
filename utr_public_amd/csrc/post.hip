// post.hip -- post-processing of the reconstructed volume on gfx950 (the
// post-processing block of MultisliceCppTest/PhaseRetrieval/UTR.cpp, behind
// the utr_highpass3d*, utr_rotate3d* and utr_symmetrize3d* C ABI of
// include/utr_hip.h).  Included by utr_api.hip
// (one translation unit).
//
// 3D rotation (Rotate3, parameters 28 and 20 = 2): k_rotate3, restated
// below with the symmetry average of UTR.cpp:1341-1355.
//
// High-pass filter (UTR.cpp:1357-1378, parameter 23): K3out -= G * K3out with
// G the Gaussian of FWHM w (sigma = w / 2.355), applied in Fourier space as
// the reference's Fftwd3frc::GaussFilter does:
//   k_post_load      real volume -> complex scratch
//   k_post_lines     batched 1D FFT of contiguous lines (x)
//   k_post_cols      batched 1D FFT of strided columns (y, then z)
//   k_post_gauss     multiply by exp(-2 pi^2 sigma^2 |q|^2) / N
//   k_post_subtract  K3out -= Re(scratch)
// The Gaussian's Fourier form is restated (GaussFilter lives in the absent
// XArrayLibrary): a normalised real-space Gaussian exp(-r^2 / 2 sigma^2) on
// the periodic grid, whose transform is exp(-2 pi^2 sigma^2 q^2) with q in
// cycles per length unit.

namespace utr {

constexpr int LDS_POST = 8192;

template <typename T, int N>
__global__ void __launch_bounds__(256) k_post_lines(cplx<T>* __restrict__ data, int nlines, int sign,
                                                    const cplx<T>* __restrict__ tw) {
    constexpr int NB = lds_batch(N, sizeof(cplx<T>), LDS_POST);
    __shared__ cplx<T> sm[2 * lds_len(NB * N)];
    cplx<T>* a = sm;
    cplx<T>* b = sm + lds_len(NB * N);
    const int l0 = blockIdx.x * NB;
    const int nb = min(NB, nlines - l0);
    cplx<T>* base = data + (size_t)l0 * N;
    for (int t = threadIdx.x; t < nb * N; t += blockDim.x) a[lx(t)] = base[t];
    __syncthreads();
    fft_lds<T, N>(a, b, nb, tw, sign, threadIdx.x, blockDim.x);
    for (int t = threadIdx.x; t < nb * N; t += blockDim.x) base[t] = a[lx(t)];
}

// data viewed as [outer][N][inner]; tile (tx, ty) = NB adjacent columns of
// slab ty
template <typename T, int N>
__global__ void __launch_bounds__(256) k_post_cols(cplx<T>* __restrict__ data, int inner, int gx, int gy,
                                                   int sign, const cplx<T>* __restrict__ tw) {
    constexpr int NB = lds_batch(N, sizeof(cplx<T>), LDS_POST);
    __shared__ cplx<T> sm[2 * lds_len(NB * N)];
    cplx<T>* a = sm;
    cplx<T>* b = sm + lds_len(NB * N);
    int tx, ty;
    if (!xcd_tile(gx, gy, tx, ty)) return;
    const int c0 = tx * NB;
    const int nb = min(NB, inner - c0);
    cplx<T>* base = data + (size_t)ty * N * inner + c0;
    for (int t = threadIdx.x; t < nb * N; t += blockDim.x) {
        const int y = t / nb, c = t - y * nb;
        a[lx(c * N + y)] = base[(size_t)y * inner + c];
    }
    __syncthreads();
    fft_lds<T, N>(a, b, nb, tw, sign, threadIdx.x, blockDim.x);
    for (int t = threadIdx.x; t < nb * N; t += blockDim.x) {
        const int y = t / nb, c = t - y * nb;
        base[(size_t)y * inner + c] = a[lx(c * N + y)];
    }
}

__global__ void k_post_load(const float* __restrict__ vol, cplx<float>* __restrict__ out, size_t n) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = {vol[t], 0.f};
}

// frequency index of DFT bin k of length n: k for k < n/2, k - n otherwise
__device__ __forceinline__ int post_freq(int k, int n) { return k < n / 2 ? k : k - n; }

__global__ void k_post_gauss(cplx<float>* __restrict__ data, int nx, int ny, int nz, double cx, double cy,
                             double cz, double invN) {
    const size_t n = (size_t)nx * ny * nz;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / ((size_t)nx * ny));
    const double fi = post_freq(i, nx), fj = post_freq(j, ny), fk = post_freq(k, nz);
    const float g = (float)(exp(-(cx * fi * fi + cy * fj * fj + cz * fk * fk)) * invN);
    data[t].x *= g;
    data[t].y *= g;
}

// regularised inverse (-Laplacian): norm / (alpha + 4 pi^2 |q|^2) / N, with
// dk = 4 pi^2 / aperture^2 per axis (UTR.h:135-185 in the wrapped layout)
__global__ void k_post_invlap(cplx<float>* __restrict__ data, int nx, int ny, int nz, double dksi2,
                              double deta2, double dzeta2, double alpha, double normN) {
    const size_t n = (size_t)nx * ny * nz;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / ((size_t)nx * ny));
    const double i1 = post_freq(i, nx), j1 = post_freq(j, ny), k1 = post_freq(k, nz);
    const float f = (float)(normN / (dksi2 * i1 * i1 + deta2 * j1 * j1 + dzeta2 * k1 * k1 + alpha));
    data[t].x *= f;
    data[t].y *= f;
}

__global__ void k_post_real(float* __restrict__ vol, const cplx<float>* __restrict__ v, size_t n) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) vol[t] = v[t].x;
}

__global__ void k_post_subtract(float* __restrict__ vol, const cplx<float>* __restrict__ low, size_t n) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) vol[t] -= low[t].x;
}

// 3D rotation of a volume (XArray3DSpln::Rotate3, UTR.cpp:1349 and 1393;
// the library is absent, restated): the content turns by
// R = Rz(a2) Ry(b) Rz(a) about the array centre (nx/2, ny/2, nz/2), the
// rotation pdb.exe applies to atoms (pdb_run.cpp:419-432), so
// out(p) = in(R^T (p - c) + c); the value at a source position is the
// tricubic convolution of Keys (1981, a = -1/2) of its 4 x 4 x 4
// neighbours, 0 outside the volume.  minv: R^T row-major.  FP64, one
// rounding to float.
struct Rot3 {
    double m[9];
};

__device__ __forceinline__ void keys_w64(double t, double w[4]) {
    const double t2 = t * t, t3 = t2 * t;
    w[0] = -0.5 * t3 + t2 - 0.5 * t;
    w[1] = 1.5 * t3 - 2.5 * t2 + 1.0;
    w[2] = -1.5 * t3 + 2.0 * t2 + 0.5 * t;
    w[3] = 0.5 * t3 - 0.5 * t2;
}

__global__ void __launch_bounds__(256) k_rotate3(const float* __restrict__ in, float* __restrict__ out,
                                                 int nx, int ny, int nz, Rot3 M) {
    const size_t n = (size_t)nx * ny * nz;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / ((size_t)nx * ny));
    const double cx = 0.5 * nx, cy = 0.5 * ny, cz = 0.5 * nz;
    const double dx = i - cx, dy = j - cy, dz = k - cz;
    const double q[3] = {cx + M.m[0] * dx + M.m[1] * dy + M.m[2] * dz,
                         cy + M.m[3] * dx + M.m[4] * dy + M.m[5] * dz,
                         cz + M.m[6] * dx + M.m[7] * dy + M.m[8] * dz};
    // far outside: every neighbour is 0
    if (!(q[0] > -2.0 && q[0] < nx + 1.0 && q[1] > -2.0 && q[1] < ny + 1.0 && q[2] > -2.0 && q[2] < nz + 1.0)) {
        out[t] = 0.f;
        return;
    }
    const double f[3] = {floor(q[0]), floor(q[1]), floor(q[2])};
    double wx[4], wy[4], wz[4];
    keys_w64(q[0] - f[0], wx);
    keys_w64(q[1] - f[1], wy);
    keys_w64(q[2] - f[2], wz);
    const int ix = (int)f[0] - 1, iy = (int)f[1] - 1, iz = (int)f[2] - 1;
    double acc = 0.0;
    for (int c = 0; c < 4; c++) {
        const int zz = iz + c;
        if ((unsigned)zz >= (unsigned)nz) continue;
        for (int b = 0; b < 4; b++) {
            const int yy = iy + b;
            if ((unsigned)yy >= (unsigned)ny) continue;
            const float* rowp = in + ((size_t)zz * ny + yy) * nx;
            const double wyz = wy[b] * wz[c];
#pragma unroll
            for (int a = 0; a < 4; a++) {
                const int xx = ix + a;
                if ((unsigned)xx < (unsigned)nx) acc += wx[a] * wyz * (double)rowp[xx];
            }
        }
    }
    out[t] = (float)acc;
}

// out += w * in (the symmetry sum), elementwise
__global__ void k_post_axpy(float* __restrict__ out, const float* __restrict__ in, float w, size_t n) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] += w * in[t];
}

__global__ void k_post_scale(float* __restrict__ v, float w, size_t n) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) v[t] *= w;
}

// ThresholdLow(background, threshold) of UTR.cpp:1383: v - background, then
// values below threshold set to it (the comparison leaves NaN as NaN).
__global__ void k_post_threshold(float* __restrict__ v, size_t n, float a, float b) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float x = v[t] - a;
    if (x < b) x = b;
    v[t] = x;
}

// Reslice with axis flips (UTR.cpp:1449-1453): output slices mode 0 XY, 1 XZ,
// 2 YZ, 3 YX, 4 ZX, 5 ZY (first letter along a slice's rows, second down the
// slice, slices stacked along the remaining axis); a flip reverses an input
// axis first.  One thread per output voxel; reads gather, writes coalesce.
__global__ void k_post_reslice(const float* __restrict__ in, float* __restrict__ out, int nx, int ny, int nz,
                               int a0, int a1, int a2, int fx, int fy, int fz) {
    const int n[3] = {nx, ny, nz};
    const int ox = n[a0], oy = n[a1];
    const size_t total = (size_t)nx * ny * nz;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int i = (int)(t % ox), j = (int)((t / ox) % oy), k = (int)(t / ((size_t)ox * oy));
    int src[3];
    src[a0] = i;
    src[a1] = j;
    src[a2] = k;
    if (fx) src[0] = nx - 1 - src[0];
    if (fy) src[1] = ny - 1 - src[1];
    if (fz) src[2] = nz - 1 - src[2];
    out[t] = in[((size_t)src[2] * ny + src[1]) * nx + src[0]];
}

// Thick slices (UTR.cpp:1456-1490): slice k1 of the nz / step output is the
// MIP (mode 1) or the average (mode 2, float sum in slice order, then the
// division) of input slices k1 step .. k1 step + step - 1.
__global__ void k_post_thick(const float* __restrict__ in, float* __restrict__ out, int nx, int ny, int nzt,
                             int mode, int step) {
    const size_t plane = (size_t)nx * ny;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= plane * nzt) return;
    const size_t p = t % plane;
    const size_t k1 = t / plane;
    const float* s = in + k1 * step * plane + p;
    float v = s[0];
    if (mode == 1) {
        for (int m = 1; m < step; m++) {
            const float w = s[(size_t)m * plane];
            if (w > v) v = w;
        }
    } else {
        for (int m = 1; m < step; m++) v += s[(size_t)m * plane];
        v /= (float)step;
    }
    out[t] = v;
}

// 16-bit output (UTR.cpp:1494-1498, the absent Convert restated): values
// clamped to [minin, maxin], mapped linearly onto [0, maxout], rounded half
// away from zero (FP64).
__global__ void k_post_u16(const float* __restrict__ in, unsigned short* __restrict__ out, size_t n, double minin,
                           double maxin, double scale) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double x = fmin(fmax((double)in[t], minin), maxin);
    out[t] = (unsigned short)round((x - minin) * scale);
}

}  // namespace utr

namespace {

// 1D FFTs of every line of a complex [nz][ny][nx] array along one axis
// (0 = x, 1 = y, 2 = z), on stream s.
int post_fft_axis(utr::cplx<float>* d, int nx, int ny, int nz, int axis, int sign,
                  const utr::cplx<float>* tw, hipStream_t s) {
    const int n = axis == 0 ? nx : (axis == 1 ? ny : nz);
    bool ok = false;
    if (axis == 0) {
#define UTR_PL(N)                                                                               \
    {                                                                                           \
        constexpr int NB = utr::lds_batch(N, sizeof(utr::cplx<float>), utr::LDS_POST);          \
        const int nlines = ny * nz;                                                             \
        utr::k_post_lines<float, N><<<(nlines + NB - 1) / NB, 256, 0, s>>>(d, nlines, sign, tw); \
        ok = true;                                                                              \
    }
        UTR_DISPATCH_N(n, UTR_PL)
#undef UTR_PL
    } else {
        const int inner = axis == 1 ? nx : nx * ny;
        const int outer = axis == 1 ? nz : 1;
#define UTR_PC(N)                                                                               \
    {                                                                                           \
        constexpr int NB = utr::lds_batch(N, sizeof(utr::cplx<float>), utr::LDS_POST);          \
        const int gx = (inner + NB - 1) / NB;                                                   \
        utr::k_post_cols<float, N><<<utr::xcd_grid(gx * outer), 256, 0, s>>>(d, inner, gx, outer, \
                                                                              sign, tw);        \
        ok = true;                                                                              \
    }
        UTR_DISPATCH_N(n, UTR_PC)
#undef UTR_PC
    }
    if (!ok) return set_err(UTR_E_UNSUPPORTED, "volume edge not supported by the FFT kernels");
    UTR_HIP_TRY(hipGetLastError());
    return UTR_OK;
}

// R^T of R = Rz(a2) Ry(b) Rz(a) (angles in degrees), the rotation of
// pdb_run.cpp:419-432: Rz(a) (x, y) -> (x cos a - y sin a, x sin a + y cos a),
// Ry(b) (x, z) -> (x cos b + z sin b, -x sin b + z cos b).
utr::Rot3 rot3_inverse(double aZ_deg, double aY_deg, double aZ2_deg) {
    const double d = 3.14159265358979323846 / 180.0;
    const double ca = std::cos(aZ_deg * d), sa = std::sin(aZ_deg * d);
    const double cb = std::cos(aY_deg * d), sb = std::sin(aY_deg * d);
    const double cg = std::cos(aZ2_deg * d), sg = std::sin(aZ2_deg * d);
    const double R[9] = {cg * cb * ca - sg * sa, -cg * cb * sa - sg * ca, cg * sb,
                         sg * cb * ca + cg * sa, -sg * cb * sa + cg * ca, sg * sb,
                         -sb * ca, sb * sa, cb};
    utr::Rot3 M;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) M.m[3 * r + c] = R[3 * c + r];
    return M;
}

int check_vol(const void* p, int nx, int ny, int nz) {
    if (!p) return set_err(UTR_E_INVALID, "null volume pointer");
    if (nx <= 0 || ny <= 0 || nz <= 0) return set_err(UTR_E_INVALID, "volume sizes must be positive");
    return UTR_OK;
}

}  // namespace

extern "C" {

int utr_rotate3d_device(const float* d_in, float* d_out, int nx, int ny, int nz, double angZ_deg,
                        double angY_deg, double angZ2_deg) {
    int rc = check_vol(d_in, nx, ny, nz);
    if (rc) return rc;
    if (!d_out || d_out == d_in) return set_err(UTR_E_INVALID, "utr_rotate3d_device: output must be a distinct buffer");
    if (!std::isfinite(angZ_deg) || !std::isfinite(angY_deg) || !std::isfinite(angZ2_deg))
        return set_err(UTR_E_INVALID, "rotation angles must be finite");
    const size_t n = (size_t)nx * ny * nz;
    utr::k_rotate3<<<(unsigned)((n + 255) / 256), 256>>>(d_in, d_out, nx, ny, nz,
                                                         rot3_inverse(angZ_deg, angY_deg, angZ2_deg));
    UTR_HIP_TRY(hipGetLastError());
    UTR_HIP_TRY(hipDeviceSynchronize());
    return UTR_OK;
}

int utr_rotate3d(float* vol, int nx, int ny, int nz, double angZ_deg, double angY_deg, double angZ2_deg) {
    int rc = check_vol(vol, nx, ny, nz);
    if (rc) return rc;
    const size_t n = (size_t)nx * ny * nz;
    LocalBuf<float> a, b;
    UTR_HIP_TRY(a.alloc(n));
    UTR_HIP_TRY(b.alloc(n));
    UTR_HIP_TRY(hipMemcpy(a.p, vol, sizeof(float) * n, hipMemcpyHostToDevice));
    rc = utr_rotate3d_device(a.p, b.p, nx, ny, nz, angZ_deg, angY_deg, angZ2_deg);
    if (rc) return rc;
    UTR_HIP_TRY(hipMemcpy(vol, b.p, sizeof(float) * n, hipMemcpyDeviceToHost));
    return UTR_OK;
}

int utr_symmetrize3d_device(float* d_vol, int nx, int ny, int nz, int nsym, const double* angZ_deg,
                            const double* angY_deg, const double* angZ2_deg) {
    int rc = check_vol(d_vol, nx, ny, nz);
    if (rc) return rc;
    if (nsym < 1 || !angZ_deg || !angY_deg || !angZ2_deg)
        return set_err(UTR_E_INVALID, "utr_symmetrize3d: need nsym >= 1 rows of angles");
    for (int s = 0; s < nsym; s++)
        if (!std::isfinite(angZ_deg[s]) || !std::isfinite(angY_deg[s]) || !std::isfinite(angZ2_deg[s]))
            return set_err(UTR_E_INVALID, "rotation angles must be finite");
    if (nsym == 1) return UTR_OK;
    const size_t n = (size_t)nx * ny * nz;
    LocalBuf<float> sum, rot;
    UTR_HIP_TRY(sum.alloc(n));
    UTR_HIP_TRY(rot.alloc(n));
    UTR_HIP_TRY(hipMemcpy(sum.p, d_vol, sizeof(float) * n, hipMemcpyDeviceToDevice));
    const unsigned blocks = (unsigned)((n + 255) / 256);
    // UTR.cpp:1341-1355: rows 1 .. nsym-1 turned back (Rotate3 with -Z'', -Y', -Z)
    for (int s = 1; s < nsym; s++) {
        utr::k_rotate3<<<blocks, 256>>>(d_vol, rot.p, nx, ny, nz,
                                        rot3_inverse(-angZ2_deg[s], -angY_deg[s], -angZ_deg[s]));
        UTR_HIP_TRY(hipGetLastError());
        utr::k_post_axpy<<<blocks, 256>>>(sum.p, rot.p, 1.f, n);
        UTR_HIP_TRY(hipGetLastError());
    }
    utr::k_post_scale<<<blocks, 256>>>(sum.p, 1.f / (float)nsym, n);
    UTR_HIP_TRY(hipGetLastError());
    UTR_HIP_TRY(hipMemcpy(d_vol, sum.p, sizeof(float) * n, hipMemcpyDeviceToDevice));
    UTR_HIP_TRY(hipDeviceSynchronize());
    return UTR_OK;
}

int utr_symmetrize3d(float* vol, int nx, int ny, int nz, int nsym, const double* angZ_deg,
                     const double* angY_deg, const double* angZ2_deg) {
    int rc = check_vol(vol, nx, ny, nz);
    if (rc) return rc;
    const size_t n = (size_t)nx * ny * nz;
    LocalBuf<float> a;
    UTR_HIP_TRY(a.alloc(n));
    UTR_HIP_TRY(hipMemcpy(a.p, vol, sizeof(float) * n, hipMemcpyHostToDevice));
    rc = utr_symmetrize3d_device(a.p, nx, ny, nz, nsym, angZ_deg, angY_deg, angZ2_deg);
    if (rc) return rc;
    UTR_HIP_TRY(hipMemcpy(vol, a.p, sizeof(float) * n, hipMemcpyDeviceToHost));
    return UTR_OK;
}

}  // extern "C"

namespace {

// d_vol (real, nz*ny*nx) -> 3D FFT in work -> filter(work, blocks, stream) ->
// inverse FFT -> d_vol -= Re (subtract) or d_vol = Re, launched on stream s
// (not synchronised).  tw[0..2]: twiddle tables of nx, ny, nz on the device.
template <typename F>
int post_fourier_launch(float* d_vol, int nx, int ny, int nz, utr::cplx<float>* work,
                        utr::cplx<float>* const tw[3], hipStream_t s, F&& filter, bool subtract) {
    const size_t n = (size_t)nx * ny * nz;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    utr::k_post_load<<<blocks, 256, 0, s>>>(d_vol, work, n);
    UTR_HIP_TRY(hipGetLastError());
    for (int ax = 0; ax < 3; ax++) {
        const int rc = post_fft_axis(work, nx, ny, nz, ax, -1, tw[ax], s);
        if (rc) return rc;
    }
    filter(work, blocks, s);
    UTR_HIP_TRY(hipGetLastError());
    for (int ax = 2; ax >= 0; ax--) {
        const int rc = post_fft_axis(work, nx, ny, nz, ax, +1, tw[ax], s);
        if (rc) return rc;
    }
    if (subtract) utr::k_post_subtract<<<blocks, 256, 0, s>>>(d_vol, work, n);
    else utr::k_post_real<<<blocks, 256, 0, s>>>(d_vol, work, n);
    UTR_HIP_TRY(hipGetLastError());
    return UTR_OK;
}

int post_check_edges(int nx, int ny, int nz) {
    if (!fft_size_supported(nx) || !fft_size_supported(ny) || !fft_size_supported(nz))
        return set_err(UTR_E_UNSUPPORTED,
                       "volume edge not supported by the FFT kernels (supported: 16 32 48 64 96 128 "
                       "160 192 256 320 384 512)");
    return UTR_OK;
}

// The same with a complex work volume and twiddle tables allocated here, on a
// stream of its own, synchronised before returning.
template <typename F>
int post_fourier(float* d_vol, int nx, int ny, int nz, F&& filter, bool subtract) {
    if (const int rc = post_check_edges(nx, ny, nz)) return rc;
    const size_t n = (size_t)nx * ny * nz;
    LocalBuf<utr::cplx<float>> work, tx, ty, tz;
    UTR_HIP_TRY(work.alloc(n));
    const int ns[3] = {nx, ny, nz};
    DevBuf<utr::cplx<float>>* tws[3] = {&tx, &ty, &tz};
    for (int d = 0; d < 3; d++) {
        auto tw = make_twiddles<float>(ns[d]);
        UTR_HIP_TRY(tws[d]->alloc(tw.size()));
        UTR_HIP_TRY(hipMemcpy(tws[d]->p, tw.data(), sizeof(utr::cplx<float>) * tw.size(),
                              hipMemcpyHostToDevice));
    }
    hipStream_t s = nullptr;
    UTR_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    utr::cplx<float>* const twp[3] = {tx.p, ty.p, tz.p};
    int rc = post_fourier_launch(d_vol, nx, ny, nz, work.p, twp, s, filter, subtract);
    if (!rc) {
        const hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = set_err(UTR_E_DEVICE, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
    }
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
    return rc;
}

// Gaussian high-pass filter of parameter 23 (UTR.cpp:1357-1378) as a filter
// callback for post_fourier / post_fourier_launch.
struct HighpassFilter {
    int nx, ny, nz;
    double cx, cy, cz, invN;
    HighpassFilter(int nx_, int ny_, int nz_, double xst, double yst, double zst, double width)
        : nx(nx_), ny(ny_), nz(nz_) {
        const double pi = 3.14159265358979323846;
        const double sigma = width / 2.355;
        const double k2 = 2.0 * pi * pi * sigma * sigma;
        cx = k2 / ((nx * xst) * (nx * xst));
        cy = k2 / ((ny * yst) * (ny * yst));
        cz = k2 / ((nz * zst) * (nz * zst));
        invN = 1.0 / ((double)nx * ny * nz);
    }
    void operator()(utr::cplx<float>* d, unsigned blocks, hipStream_t s) const {
        utr::k_post_gauss<<<blocks, 256, 0, s>>>(d, nx, ny, nz, cx, cy, cz, invN);
    }
};

int highpass_check(const float* d_vol, int nx, int ny, int nz, double xst, double yst, double zst,
                   double width) {
    if (!d_vol) return set_err(UTR_E_INVALID, "null volume pointer");
    if (nx <= 0 || ny <= 0 || nz <= 0 || !(xst > 0) || !(yst > 0) || !(zst > 0))
        return set_err(UTR_E_INVALID, "utr_highpass3d: sizes and steps must be positive");
    if (!(width > 0)) return set_err(UTR_E_INVALID, "utr_highpass3d: the filter width must be positive");
    return UTR_OK;
}

// ------------------------------------------- allocation-free high-pass filter
// Work memory of utr_highpass3d_device_noalloc: three twiddle tables of
// kHpTwiddles entries, then the complex work volume of at most kHpMaxVoxels
// voxels (16 MiB; 128^3 or any other nx*ny*nz up to 2^21).  Every GPU the
// library runs on holds its own copy for the life of the process.  Host code
// never uses the array's name as an address: hp_scratch() fetches each
// device's address once with hipGetSymbolAddress and that pointer is what
// the kernels receive.
constexpr size_t kHpTwiddles = 1024;           // >= FacList::ntw for every supported edge
constexpr size_t kHpMaxVoxels = (size_t)1 << 21;
static __device__ __align__(256) utr::cplx<float> scratch[3 * kHpTwiddles + kHpMaxVoxels];

constexpr int kHpMaxDevices = 64;
std::mutex g_hp_mu;
utr::cplx<float>* g_hp_ptr[kHpMaxDevices] = {};

// Address of this device's copy of scratch, looked up on the first call on
// that device (with that device current) and kept.
int hp_scratch(utr::cplx<float>** out) {
    int dev = -1;
    UTR_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kHpMaxDevices) return set_err(UTR_E_DEVICE, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_hp_mu);
    if (!g_hp_ptr[dev]) {
        void* p = nullptr;
        UTR_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return set_err(UTR_E_DEVICE, "hipGetSymbolAddress returned a null address");
        g_hp_ptr[dev] = static_cast<utr::cplx<float>*>(p);
    }
    *out = g_hp_ptr[dev];
    return UTR_OK;
}

}  // namespace

namespace utr {

// Twiddle table of length n in the FacList layout (the table make_twiddles
// builds on the host), evaluated on the device in FP64.
__global__ void k_post_twiddles(cplx<float>* __restrict__ tw, int n) {
    const FacList F = factorize(n);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= F.ntw) return;
    for (int s = 0; s < F.nf; s++) {
        const int p = F.p[s], R = F.f[s];
        const int loc = t - F.off[s];
        if (loc < 0 || loc >= p * (R - 1)) continue;
        const int k = loc / (R - 1), r = loc - k * (R - 1) + 1;
        const double ang = -2.0 * kPi * (double)r * k / ((double)p * R);
        tw[t] = {(float)cos(ang), (float)sin(ang)};
    }
}

}  // namespace utr

extern "C" {

int utr_highpass3d_device(float* d_vol, int nx, int ny, int nz, double xst, double yst, double zst,
                          double width) {
    if (const int rc = highpass_check(d_vol, nx, ny, nz, xst, yst, zst, width)) return rc;
    return post_fourier(d_vol, nx, ny, nz, HighpassFilter(nx, ny, nz, xst, yst, zst, width), true);
}

int utr_threshold_device(float* d_vol, long long n, float background, float threshold) {
    if (!d_vol || n < 0) return set_err(UTR_E_INVALID, "utr_threshold_device: null volume or negative size");
    if (n == 0) return UTR_OK;
    utr::k_post_threshold<<<(unsigned)(((size_t)n + 255) / 256), 256>>>(d_vol, (size_t)n, background, threshold);
    UTR_HIP_TRY(hipGetLastError());
    UTR_HIP_TRY(hipDeviceSynchronize());
    return UTR_OK;
}

int utr_reslice_device(const float* d_in, float* d_out, int nx, int ny, int nz, int mode, int flipx, int flipy,
                       int flipz) {
    if (int rc = check_vol(d_in, nx, ny, nz)) return rc;
    if (!d_out || d_out == d_in) return set_err(UTR_E_INVALID, "utr_reslice_device: distinct output needed");
    if (mode < 0 || mode > 5) return set_err(UTR_E_INVALID, "utr_reslice_device: mode must be 0..5");
    static const int axes[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 2, 0}, {1, 0, 2}, {2, 0, 1}, {2, 1, 0}};
    const int* a = axes[mode];
    const size_t n = (size_t)nx * ny * nz;
    utr::k_post_reslice<<<(unsigned)((n + 255) / 256), 256>>>(d_in, d_out, nx, ny, nz, a[0], a[1], a[2], flipx != 0,
                                                             flipy != 0, flipz != 0);
    UTR_HIP_TRY(hipGetLastError());
    UTR_HIP_TRY(hipDeviceSynchronize());
    return UTR_OK;
}

int utr_thick_slices_device(const float* d_in, float* d_out, int nx, int ny, int nz, int mode, int step) {
    if (int rc = check_vol(d_in, nx, ny, nz)) return rc;
    if (!d_out || d_out == d_in) return set_err(UTR_E_INVALID, "utr_thick_slices_device: distinct output needed");
    if (mode != 1 && mode != 2) return set_err(UTR_E_INVALID, "utr_thick_slices_device: mode must be 1 (MIP) or 2 (average)");
    if (step < 1) return set_err(UTR_E_INVALID, "utr_thick_slices_device: the slice thickness must be at least one voxel");
    if (step > nz) return set_err(UTR_E_INVALID, "Slice thickness is larger than the sample thickness.");  // UTR.cpp:1460
    const int nzt = nz / step;
    const size_t n = (size_t)nx * ny * nzt;
    utr::k_post_thick<<<(unsigned)((n + 255) / 256), 256>>>(d_in, d_out, nx, ny, nzt, mode, step);
    UTR_HIP_TRY(hipGetLastError());
    UTR_HIP_TRY(hipDeviceSynchronize());
    return UTR_OK;
}

int utr_convert16_device(const float* d_in, unsigned short* d_out, long long n, float minin, float maxin,
                         long long maxout) {
    if (!d_in || !d_out || n < 0) return set_err(UTR_E_INVALID, "utr_convert16_device: null buffers or negative size");
    if (!(maxin > minin) || maxout < 1 || maxout > 65535)
        return set_err(UTR_E_INVALID, "utr_convert16_device: need min-in < max-in and 1 <= max-out <= 65535");
    if (n == 0) return UTR_OK;
    const double scale = (double)maxout / ((double)maxin - (double)minin);
    utr::k_post_u16<<<(unsigned)(((size_t)n + 255) / 256), 256>>>(d_in, d_out, (size_t)n, (double)minin,
                                                                  (double)maxin, scale);
    UTR_HIP_TRY(hipGetLastError());
    UTR_HIP_TRY(hipDeviceSynchronize());
    return UTR_OK;
}

// Host-memory forms: one upload, the kernel, one download.
int utr_threshold(float* vol, long long n, float background, float threshold) {
    if (!vol || n < 0) return set_err(UTR_E_INVALID, "utr_threshold: null volume or negative size");
    if (n == 0) return UTR_OK;
    LocalBuf<float> d;
    UTR_HIP_TRY(d.alloc((size_t)n));
    UTR_HIP_TRY(hipMemcpy(d.p, vol, sizeof(float) * n, hipMemcpyHostToDevice));
    if (int rc = utr_threshold_device(d.p, n, background, threshold)) return rc;
    UTR_HIP_TRY(hipMemcpy(vol, d.p, sizeof(float) * n, hipMemcpyDeviceToHost));
    return UTR_OK;
}

int utr_reslice(const float* in, float* out, int nx, int ny, int nz, int mode, int flipx, int flipy, int flipz) {
    if (int rc = check_vol(in, nx, ny, nz)) return rc;
    if (!out) return set_err(UTR_E_INVALID, "utr_reslice: null output");
    const size_t n = (size_t)nx * ny * nz;
    LocalBuf<float> a, b;
    UTR_HIP_TRY(a.alloc(n));
    UTR_HIP_TRY(b.alloc(n));
    UTR_HIP_TRY(hipMemcpy(a.p, in, sizeof(float) * n, hipMemcpyHostToDevice));
    if (int rc = utr_reslice_device(a.p, b.p, nx, ny, nz, mode, flipx, flipy, flipz)) return rc;
    UTR_HIP_TRY(hipMemcpy(out, b.p, sizeof(float) * n, hipMemcpyDeviceToHost));
    return UTR_OK;
}

int utr_thick_slices(const float* in, float* out, int nx, int ny, int nz, int mode, int step) {
    if (int rc = check_vol(in, nx, ny, nz)) return rc;
    if (!out) return set_err(UTR_E_INVALID, "utr_thick_slices: null output");
    if (mode != 1 && mode != 2) return set_err(UTR_E_INVALID, "utr_thick_slices: mode must be 1 (MIP) or 2 (average)");
    if (step < 1) return set_err(UTR_E_INVALID, "utr_thick_slices: the slice thickness must be at least one voxel");
    if (step > nz) return set_err(UTR_E_INVALID, "Slice thickness is larger than the sample thickness.");  // UTR.cpp:1460
    const size_t n = (size_t)nx * ny * nz, no = (size_t)nx * ny * (nz / step);
    LocalBuf<float> a, b;
    UTR_HIP_TRY(a.alloc(n));
    UTR_HIP_TRY(b.alloc(no));
    UTR_HIP_TRY(hipMemcpy(a.p, in, sizeof(float) * n, hipMemcpyHostToDevice));
    if (int rc = utr_thick_slices_device(a.p, b.p, nx, ny, nz, mode, step)) return rc;
    UTR_HIP_TRY(hipMemcpy(out, b.p, sizeof(float) * no, hipMemcpyDeviceToHost));
    return UTR_OK;
}

int utr_convert16(const float* in, unsigned short* out, long long n, float minin, float maxin, long long maxout) {
    if (!in || !out || n < 0) return set_err(UTR_E_INVALID, "utr_convert16: null buffers or negative size");
    if (n == 0) return UTR_OK;
    LocalBuf<float> a;
    LocalBuf<unsigned short> b;
    UTR_HIP_TRY(a.alloc((size_t)n));
    UTR_HIP_TRY(b.alloc((size_t)n));
    UTR_HIP_TRY(hipMemcpy(a.p, in, sizeof(float) * n, hipMemcpyHostToDevice));
    if (int rc = utr_convert16_device(a.p, b.p, n, minin, maxin, maxout)) return rc;
    UTR_HIP_TRY(hipMemcpy(out, b.p, sizeof(unsigned short) * n, hipMemcpyDeviceToHost));
    return UTR_OK;
}

int utr_highpass3d_device_noalloc(float* d_vol, int nx, int ny, int nz, double xst, double yst,
                                  double zst, double width) {
    if (const int rc = highpass_check(d_vol, nx, ny, nz, xst, yst, zst, width)) return rc;
    if (const int rc = post_check_edges(nx, ny, nz)) return rc;
    const size_t n = (size_t)nx * ny * nz;
    if (n > kHpMaxVoxels)
        return set_err(UTR_E_INVALID, "utr_highpass3d_device_noalloc: nx*ny*nz = " + std::to_string(n) +
                                          " exceeds the fixed work area of " +
                                          std::to_string(kHpMaxVoxels) + " voxels");
    utr::cplx<float>* base = nullptr;
    if (const int rc = hp_scratch(&base)) return rc;
    // Nothing needs clearing: the twiddle kernels write every table entry the
    // FFTs read and k_post_load writes every work voxel before it is read.
    utr::cplx<float>* const tw[3] = {base, base + kHpTwiddles, base + 2 * kHpTwiddles};
    const int ns[3] = {nx, ny, nz};
    hipStream_t s = nullptr;  // the null stream: no stream is created
    for (int d = 0; d < 3; d++) {
        utr::k_post_twiddles<<<(unsigned)(kHpTwiddles / 256), 256, 0, s>>>(tw[d], ns[d]);
        UTR_HIP_TRY(hipGetLastError());
    }
    const int rc = post_fourier_launch(d_vol, nx, ny, nz, base + 3 * kHpTwiddles, tw, s,
                                       HighpassFilter(nx, ny, nz, xst, yst, zst, width), true);
    if (rc) return rc;
    UTR_HIP_TRY(hipStreamSynchronize(s));
    return UTR_OK;
}

int utr_invlaplace3d_device(float* d_vol, int nx, int ny, int nz, double xst, double yst, double zst,
                            double alpha, double norm) {
    if (!d_vol) return set_err(UTR_E_INVALID, "null volume pointer");
    if (nx <= 0 || ny <= 0 || nz <= 0 || !(xst > 0) || !(yst > 0) || !(zst > 0))
        return set_err(UTR_E_INVALID, "utr_invlaplace3d: sizes and steps must be positive");
    // UTR.h:148-149
    if (!(alpha > 0)) return set_err(UTR_E_INVALID, "InverseMLaplacian3D() cannot be called with alpha <= 0");
    const size_t n = (size_t)nx * ny * nz;
    const double fact = 4.0 * 3.14159265358979323846 * 3.14159265358979323846;
    const double xa = nx * xst, ya = ny * yst, za = nz * zst;
    return post_fourier(d_vol, nx, ny, nz, [&](utr::cplx<float>* d, unsigned blocks, hipStream_t s) {
        utr::k_post_invlap<<<blocks, 256, 0, s>>>(d, nx, ny, nz, fact / (xa * xa), fact / (ya * ya),
                                                  fact / (za * za), alpha, norm / (double)n);
    }, false);
}

int utr_invlaplace3d(float* vol, int nx, int ny, int nz, double xst, double yst, double zst, double alpha,
                     double norm) {
    if (!vol) return set_err(UTR_E_INVALID, "null volume pointer");
    if (nx <= 0 || ny <= 0 || nz <= 0)
        return set_err(UTR_E_INVALID, "utr_invlaplace3d: sizes and steps must be positive");
    const size_t n = (size_t)nx * ny * nz;
    LocalBuf<float> d;
    UTR_HIP_TRY(d.alloc(n));
    UTR_HIP_TRY(hipMemcpy(d.p, vol, sizeof(float) * n, hipMemcpyHostToDevice));
    const int rc = utr_invlaplace3d_device(d.p, nx, ny, nz, xst, yst, zst, alpha, norm);
    if (rc) return rc;
    UTR_HIP_TRY(hipMemcpy(vol, d.p, sizeof(float) * n, hipMemcpyDeviceToHost));
    return UTR_OK;
}

int utr_highpass3d(float* vol, int nx, int ny, int nz, double xst, double yst, double zst, double width) {
    if (!vol) return set_err(UTR_E_INVALID, "null volume pointer");
    if (nx <= 0 || ny <= 0 || nz <= 0)
        return set_err(UTR_E_INVALID, "utr_highpass3d: sizes and steps must be positive");
    const size_t n = (size_t)nx * ny * nz;
    LocalBuf<float> d;
    UTR_HIP_TRY(d.alloc(n));
    UTR_HIP_TRY(hipMemcpy(d.p, vol, sizeof(float) * n, hipMemcpyHostToDevice));
    const int rc = utr_highpass3d_device(d.p, nx, ny, nz, xst, yst, zst, width);
    if (rc) return rc;
    UTR_HIP_TRY(hipMemcpy(vol, d.p, sizeof(float) * n, hipMemcpyDeviceToHost));
    return UTR_OK;
}

}  // extern "C"
