// post.hip -- post-processing of the reconstructed volume on gfx950 (the
// post-processing block of MultisliceCppTest/PhaseRetrieval/UTR.cpp, behind
// the utr_highpass3d* C ABI of include/utr_hip.h).  Included by utr_api.hip
// (one translation unit).
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

__global__ void k_post_subtract(float* __restrict__ vol, const cplx<float>* __restrict__ low, size_t n) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) vol[t] -= low[t].x;
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

}  // namespace

extern "C" {

int utr_highpass3d_device(float* d_vol, int nx, int ny, int nz, double xst, double yst, double zst,
                          double width) {
    if (!d_vol) return set_err(UTR_E_INVALID, "null volume pointer");
    if (nx <= 0 || ny <= 0 || nz <= 0 || !(xst > 0) || !(yst > 0) || !(zst > 0))
        return set_err(UTR_E_INVALID, "utr_highpass3d: sizes and steps must be positive");
    if (!(width > 0)) return set_err(UTR_E_INVALID, "utr_highpass3d: the filter width must be positive");
    if (!fft_size_supported(nx) || !fft_size_supported(ny) || !fft_size_supported(nz))
        return set_err(UTR_E_UNSUPPORTED,
                       "volume edge not supported by the FFT kernels (supported: 16 32 48 64 96 128 "
                       "160 192 256 320 384 512)");
    const size_t n = (size_t)nx * ny * nz;
    DevBuf<utr::cplx<float>> scratch, tx, ty, tz;
    UTR_HIP_TRY(scratch.alloc(n));
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
    auto run = [&]() -> int {
        const unsigned blocks = (unsigned)((n + 255) / 256);
        utr::k_post_load<<<blocks, 256, 0, s>>>(d_vol, scratch.p, n);
        UTR_HIP_TRY(hipGetLastError());
        for (int ax = 0; ax < 3; ax++) {
            const int rc = post_fft_axis(scratch.p, nx, ny, nz, ax, -1, tws[ax]->p, s);
            if (rc) return rc;
        }
        const double pi = 3.14159265358979323846;
        const double sigma = width / 2.355;
        const double k2 = 2.0 * pi * pi * sigma * sigma;
        utr::k_post_gauss<<<blocks, 256, 0, s>>>(scratch.p, nx, ny, nz, k2 / ((nx * xst) * (nx * xst)),
                                                 k2 / ((ny * yst) * (ny * yst)),
                                                 k2 / ((nz * zst) * (nz * zst)), 1.0 / (double)n);
        UTR_HIP_TRY(hipGetLastError());
        for (int ax = 2; ax >= 0; ax--) {
            const int rc = post_fft_axis(scratch.p, nx, ny, nz, ax, +1, tws[ax]->p, s);
            if (rc) return rc;
        }
        utr::k_post_subtract<<<blocks, 256, 0, s>>>(d_vol, scratch.p, n);
        UTR_HIP_TRY(hipGetLastError());
        UTR_HIP_TRY(hipStreamSynchronize(s));
        return UTR_OK;
    };
    const int rc = run();
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
    return rc;
}

int utr_highpass3d(float* vol, int nx, int ny, int nz, double xst, double yst, double zst, double width) {
    if (!vol) return set_err(UTR_E_INVALID, "null volume pointer");
    if (nx <= 0 || ny <= 0 || nz <= 0)
        return set_err(UTR_E_INVALID, "utr_highpass3d: sizes and steps must be positive");
    const size_t n = (size_t)nx * ny * nz;
    DevBuf<float> d;
    UTR_HIP_TRY(d.alloc(n));
    UTR_HIP_TRY(hipMemcpy(d.p, vol, sizeof(float) * n, hipMemcpyHostToDevice));
    const int rc = utr_highpass3d_device(d.p, nx, ny, nz, xst, yst, zst, width);
    if (rc) return rc;
    UTR_HIP_TRY(hipMemcpy(vol, d.p, sizeof(float) * n, hipMemcpyDeviceToHost));
    return UTR_OK;
}

}  // extern "C"
