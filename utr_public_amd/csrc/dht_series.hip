// dht_series.hip -- DHT methods 1 (IWFR) and 2 (CTFL2) on gfx950: the complex
// amplitude of each orientation at its zout is retrieved from the
// orientation's defocus series (DHT.cpp:889-917), then dht.hip's method-3
// kernels back-propagate it to every output plane and splat it
// (DHT.cpp:957-1164).  Included by utr_api.hip after dht.hip.
//
//   k_ser_rows<N, 0>  IWFR start: u = sqrt(I) - 1 of each image, FFT along x
//   k_ser_rows<N, 1>  IWFR iteration, per image: inverse FFT along x, the
//                     modulus replaced by sqrt(I) (phase kept), FFT along x
//   k_ser_cols<N>     IWFR, per orientation and column block: FFT along y of
//                     each image's field, propagated to zout and averaged
//                     (the new estimate and its change for the stopping
//                     test), then propagated back to every image plane and
//                     inverse-transformed along y
//   k_ser_check       the stopping test of each orientation
//   k_ctf_cols<N>     CTFL2, per orientation and column block: the least-
//                     squares phase spectrum at the first plane, inverse
//                     FFT along y
//   k_ctf_rows<N>     CTFL2: inverse FFT along x, u = sqrt(I_0) e^(i phi) - 1,
//                     FFT along x (k_dht_cols then finishes the spectrum)
//   k_ser_wave<N>     the retrieved wave 1 + u back in real space
// Fields are carried as u = psi - 1: every propagator passes the constant 1
// unchanged, so FP32 keeps the relative precision of the scattered wave.
// The propagation between two planes of a series is free-space Fresnel
// propagation inside the objective aperture: the objective's Cs phases are
// the same at every defocus and cancel between two planes.
//
// Published algorithms (the reference's XA_IWFR is not in its tree):
//   IWFR  Allen, McBride, O'Leary, Oxley, Ultramicroscopy 100 (2004) 91-104:
//         psi_0 = mean over n of P(zout - z_n) sqrt(I_n); iteration k:
//         psi_n = P(z_n - zout) psi_{k-1}, psi_n <- sqrt(I_n) psi_n / |psi_n|,
//         psi_k = mean of P(zout - z_n) psi_n; stop after itermax iterations
//         or when ||psi_k - psi_{k-1}||^2 < epsilon ||psi_{k-1}||^2.
//   CTFL2 the weak phase-amplitude object model over several distances
//         (Paganin, Barty, McMahon, Nugent, J. Microsc. 214 (2004) 51-61):
//         (I_n - 1)^ = 2 s_n phi^ - 2 c_n mu^ with s_n, c_n = sin, cos of
//         chi_n = pi wl (z_n - z_0) q^2, solved by Tikhonov-regularised least
//         squares, phi^ = (Sum c^2 Sum s I^ - Sum sc Sum c I^) /
//         (2 (Sum s^2 Sum c^2 - (Sum sc)^2) + epsilon), zero outside the
//         aperture; the wave at z_0 is sqrt(I_0) e^(i phi).

namespace utr {

// sqrt(I) - 1 of a raw input value, renormalised and clamped (DHT.cpp:850-858)
__device__ __forceinline__ float ser_amp_m1(float raw, const DhtGeom& g) {
    double x = (double)raw;
    if (g.use_f1) x = x / (double)g.f1;
    if (g.use_f2) x = x + (double)g.f2;
    if (x < 0.0) x = 0.0;
    return (float)((x - 1.0) / (sqrt(x) + 1.0));
}

// psi = 1 + u -> sqrt(I) psi / |psi| = (1 + am1)(1 + u) / (1 + b) with
// b = |1 + u| - 1, written as u + gq (1 + u), gq = (am1 - b) / (1 + b), so
// that only small quantities are subtracted; sqrt(I) (phase 0) where psi = 0.
__device__ __forceinline__ cplx<float> ser_replace(cplx<float> u, float am1) {
    const float m = u.x * (2.f + u.x) + u.y * u.y;  // |1 + u|^2 - 1
    if (!(1.f + m > 0.f)) return {am1, 0.f};
    const float b = m / (sqrtf(1.f + m) + 1.f);
    const float gq = (am1 - b) / (1.f + b);
    return {u.x + gq + gq * u.x, u.y + gq * u.y};
}

__device__ __forceinline__ bool ser_in_aperture(int kx, int ky, const DhtGeom& g) {
    if (!(g.k2maxo > 0)) return true;
    const double fx = dht_freq(kx, g.nx, g.xst), fy = dht_freq(ky, g.ny, g.xst);
    return fx * fx + fy * fy < g.k2maxo;
}

// Rows of images p0 + blockIdx.y (Bp, img: that image's planes at index
// blockIdx.y); grid (ceil(ny / NB), images).
// A1: sqrt(I) - 1 of the same images, written by the start (MODE 0) and
// read by the iterations (MODE 1).
template <int N, int MODE>
__global__ void __launch_bounds__(256) k_ser_rows(cplx<float>* __restrict__ Bp, const float* __restrict__ img,
                                                  float* __restrict__ A1, int p0, DhtGeom g,
                                                  const SerImg* __restrict__ im, const int* __restrict__ done,
                                                  const cplx<float>* __restrict__ tw) {
    constexpr int NB = lds_batch(N, sizeof(cplx<float>), LDS_DHT_ROWS);
    __shared__ cplx<float> sm[2 * lds_len(NB * N)];
    cplx<float>* a = sm;
    cplx<float>* b = sm + lds_len(NB * N);
    const int pr = blockIdx.y, tid = threadIdx.x, ny = g.ny;
    if (MODE == 1 && done[im[p0 + pr].angle]) return;  // uniform over the workgroup
    const int j0 = blockIdx.x * NB, nb = min(NB, ny - j0);
    const size_t off = ((size_t)pr * ny + j0) * N;
    cplx<float>* row = Bp + off;
    if (MODE == 0) {
        for (int t = tid; t < nb * N; t += 256) {
            const float am1 = ser_amp_m1(img[off + t], g);
            A1[off + t] = am1;
            a[lx(t)] = {am1, 0.f};
        }
    } else {
        for (int t = tid; t < nb * N; t += 256) a[lx(t)] = row[t];
        __syncthreads();
        fft_lds<float, N>(a, b, nb, tw, +1, tid, 256);
        const float inv = (float)(1.0 / ((double)N * ny));
        for (int t = tid; t < nb * N; t += 256) {
            const cplx<float> v = a[lx(t)];
            a[lx(t)] = ser_replace({v.x * inv, v.y * inv}, A1[off + t]);
        }
    }
    __syncthreads();
    fft_lds<float, N>(a, b, nb, tw, -1, tid, 256);
    for (int t = tid; t < nb * N; t += 256) row[t] = a[lx(t)];
}

// One IWFR half-step per orientation a0 + blockIdx.y and column block
// blockIdx.x (NB columns of N = ny rows): gather (mode 0: the start, mode 1:
// an iteration with its change into part) into U, then spread to the image
// planes for the next iteration.  E: per image nx + ny factors
// exp(-i pi wl d fx^2), exp(-i pi wl d fy^2) of its distance d from zout.
template <int N>
__global__ void __launch_bounds__(256) k_ser_cols(cplx<float>* __restrict__ Bp, int p0, cplx<float>* __restrict__ U,
                                                  DhtGeom g, const SerAng* __restrict__ an, int a0,
                                                  const cplx<float>* __restrict__ E, double* __restrict__ part,
                                                  const int* __restrict__ done, int mode,
                                                  const cplx<float>* __restrict__ tw) {
    constexpr int NB = lds_batch(N, sizeof(cplx<float>), LDS_DHT);
    constexpr int PER = (NB * N + 255) / 256;
    __shared__ cplx<float> sm[2 * lds_len(NB * N)];
    __shared__ double s_red[2][4];
    const int a = a0 + blockIdx.y, tid = threadIdx.x, nx = g.nx;
    if (mode == 1 && done[a]) return;  // uniform over the workgroup
    const int x0 = blockIdx.x * NB, nb = min(NB, nx - x0);
    const SerAng A = an[a];
    const size_t n2 = (size_t)nx * N;
    const int ne = nx + N;
    bool in[PER];
    cplx<float> acc[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int t = tid + 256 * q;
        const int y = t / nb, c = t - y * nb;
        in[q] = t < nb * N && ser_in_aperture(x0 + c, y, g);
        acc[q] = {0.f, 0.f};
    }
    for (int k = 0; k < A.count; k++) {
        const int p = A.first + k;
        cplx<float>* sa = sm;
        cplx<float>* sb = sm + lds_len(NB * N);
        const cplx<float>* base = Bp + (size_t)(p - p0) * n2 + x0;
        for (int t = tid; t < nb * N; t += 256) {
            const int y = t / nb, c = t - y * nb;
            sa[lx(c * N + y)] = base[(size_t)y * nx + c];
        }
        __syncthreads();
        fft_lds<float, N>(sa, sb, nb, tw, -1, tid, 256);
        const cplx<float>* ex = E + (size_t)p * ne;
        const cplx<float>* ey = ex + nx;
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int t = tid + 256 * q;
            if (in[q]) {
                const int y = t / nb, c = t - y * nb;
                const cplx<float> P = cmul(ex[x0 + c], ey[y]);  // back to zout: conj(P)
                const cplx<float> v = sa[lx(c * N + y)];
                acc[q].x += v.x * P.x + v.y * P.y;
                acc[q].y += v.y * P.x - v.x * P.y;
            }
        }
        __syncthreads();  // the next image overwrites both buffers
    }
    const float invc = 1.f / (float)A.count;
    double num = 0.0, den = 0.0;
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int t = tid + 256 * q;
        if (t < nb * N) {
            const int y = t / nb, c = t - y * nb;
            acc[q].x *= invc;
            acc[q].y *= invc;
            const size_t idx = (size_t)a * n2 + (size_t)y * nx + x0 + c;
            if (mode == 1) {
                const cplx<float> o = U[idx];
                const double dr = (double)acc[q].x - o.x, di = (double)acc[q].y - o.y;
                const double orr = (double)o.x + (y == 0 && x0 + c == 0 ? (double)n2 : 0.0);  // psi = 1 + u
                num += dr * dr + di * di;
                den += orr * orr + (double)o.y * o.y;
            }
            U[idx] = acc[q];
        }
    }
    if (mode == 1) {
        for (int off = 32; off > 0; off >>= 1) {
            num += __shfl_xor(num, off);
            den += __shfl_xor(den, off);
        }
        if ((tid & 63) == 0) {
            s_red[0][tid >> 6] = num;
            s_red[1][tid >> 6] = den;
        }
        __syncthreads();
        if (tid == 0) {
            double* o = part + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 2;
            o[0] = (s_red[0][0] + s_red[0][1]) + (s_red[0][2] + s_red[0][3]);
            o[1] = (s_red[1][0] + s_red[1][1]) + (s_red[1][2] + s_red[1][3]);
        }
    }
    for (int k = 0; k < A.count; k++) {
        const int p = A.first + k;
        cplx<float>* sa = sm;
        cplx<float>* sb = sm + lds_len(NB * N);
        const cplx<float>* ex = E + (size_t)p * ne;
        const cplx<float>* ey = ex + nx;
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int t = tid + 256 * q;
            if (t < nb * N) {
                const int y = t / nb, c = t - y * nb;
                sa[lx(c * N + y)] = in[q] ? cmul(acc[q], cmul(ex[x0 + c], ey[y])) : cplx<float>{0.f, 0.f};
            }
        }
        __syncthreads();
        fft_lds<float, N>(sa, sb, nb, tw, +1, tid, 256);
        cplx<float>* base = Bp + (size_t)(p - p0) * n2 + x0;
        for (int t = tid; t < nb * N; t += 256) {
            const int y = t / nb, c = t - y * nb;
            base[(size_t)y * nx + c] = sa[lx(c * N + y)];
        }
        __syncthreads();
    }
}

// Stopping test of orientations a0 .. a0+na-1 (one workgroup): an active
// orientation counts the iteration and stops when num / den < eps; active:
// the orientations still iterating.
__global__ void __launch_bounds__(256) k_ser_check(const double* __restrict__ part, int gx, int a0, int na, double eps,
                                                   int* __restrict__ done, int* __restrict__ iters,
                                                   int* __restrict__ active) {
    __shared__ int s_cnt[4];
    int cnt = 0;
    for (int i = threadIdx.x; i < na; i += 256) {
        const int a = a0 + i;
        if (done[a]) continue;
        double num = 0.0, den = 0.0;
        for (int k = 0; k < gx; k++) {
            num += part[((size_t)i * gx + k) * 2];
            den += part[((size_t)i * gx + k) * 2 + 1];
        }
        iters[a] += 1;
        if (num / den < eps) done[a] = 1;
        else cnt++;
    }
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
    if ((threadIdx.x & 63) == 0) s_cnt[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) active[0] = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
}

__global__ void k_ser_reset(int* __restrict__ done, int* __restrict__ iters, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        done[i] = 0;
        iters[i] = 0;
    }
}

// CTFL2 per orientation a0 + blockIdx.y and column block: Bp holds the 2D
// spectra of I - 1 of the orientation's images; the phase spectrum at the
// first plane is formed in FP64, inverse-transformed along y and written
// over the first image's columns.
template <int N>
__global__ void __launch_bounds__(256) k_ctf_cols(cplx<float>* __restrict__ Bp, int p0, DhtGeom g,
                                                  const SerImg* __restrict__ im, const SerAng* __restrict__ an,
                                                  int a0, const cplx<float>* __restrict__ tw) {
    constexpr int NB = lds_batch(N, sizeof(cplx<float>), LDS_DHT);
    __shared__ cplx<float> sm[2 * lds_len(NB * N)];
    cplx<float>* sa = sm;
    cplx<float>* sb = sm + lds_len(NB * N);
    const int a = a0 + blockIdx.y, tid = threadIdx.x, nx = g.nx;
    const int x0 = blockIdx.x * NB, nb = min(NB, nx - x0);
    const SerAng A = an[a];
    const size_t n2 = (size_t)nx * N;
    const double pi = 3.14159265358979323846;
    for (int t = tid; t < nb * N; t += 256) {
        const int y = t / nb, c = t - y * nb;
        double ss = 0, cc = 0, sc = 0, sr = 0, si = 0, cr = 0, ci = 0;
        if (ser_in_aperture(x0 + c, y, g)) {
            const double fx = dht_freq(x0 + c, nx, g.xst), fy = dht_freq(y, N, g.xst), q2 = fx * fx + fy * fy;
            for (int k = 0; k < A.count; k++) {
                const int p = A.first + k;
                double s, cs;
                sincos(pi * g.wl * im[p].d * q2, &s, &cs);
                const cplx<float> v = Bp[(size_t)(p - p0) * n2 + (size_t)y * nx + x0 + c];
                ss += s * s;
                cc += cs * cs;
                sc += s * cs;
                sr += s * v.x;
                si += s * v.y;
                cr += cs * v.x;
                ci += cs * v.y;
            }
        }
        const double den = 2.0 * (ss * cc - sc * sc) + g.epsilon;
        sa[lx(c * N + y)] = {(float)((cc * sr - sc * cr) / den), (float)((cc * si - sc * ci) / den)};
    }
    __syncthreads();
    fft_lds<float, N>(sa, sb, nb, tw, +1, tid, 256);
    cplx<float>* base = Bp + (size_t)(A.first - p0) * n2 + x0;
    for (int t = tid; t < nb * N; t += 256) {
        const int y = t / nb, c = t - y * nb;
        base[(size_t)y * nx + c] = sa[lx(c * N + y)];
    }
}

// CTFL2 rows of orientation a0 + blockIdx.y: phi from the first image's
// plane of Bp, u = sqrt(I_0) e^(i phi) - 1 = (am1 cos phi - 2 sin^2(phi/2),
// (1 + am1) sin phi), FFT along x into the orientation's spectrum rows.
template <int N>
__global__ void __launch_bounds__(256) k_ctf_rows(const cplx<float>* __restrict__ Bp, int p0,
                                                  const float* __restrict__ simg, cplx<float>* __restrict__ spec,
                                                  DhtGeom g, const SerAng* __restrict__ an, int a0,
                                                  const cplx<float>* __restrict__ tw) {
    constexpr int NB = lds_batch(N, sizeof(cplx<float>), LDS_DHT_ROWS);
    __shared__ cplx<float> sm[2 * lds_len(NB * N)];
    cplx<float>* a = sm;
    cplx<float>* b = sm + lds_len(NB * N);
    const int an_i = a0 + blockIdx.y, tid = threadIdx.x, ny = g.ny;
    const SerAng A = an[an_i];
    const int j0 = blockIdx.x * NB, nb = min(NB, ny - j0);
    const size_t n2 = (size_t)N * ny;
    const cplx<float>* W = Bp + (size_t)(A.first - p0) * n2 + (size_t)j0 * N;
    const float* src = simg + (size_t)A.first * n2 + (size_t)j0 * N;
    for (int t = tid; t < nb * N; t += 256) a[lx(t)] = W[t];
    __syncthreads();
    fft_lds<float, N>(a, b, nb, tw, +1, tid, 256);
    const float inv = (float)(1.0 / (double)n2);
    for (int t = tid; t < nb * N; t += 256) {
        const float phi = a[lx(t)].x * inv, am1 = ser_amp_m1(src[t], g);
        float s, cs;
        sincosf(phi, &s, &cs);
        const float h = sinf(0.5f * phi);
        a[lx(t)] = {am1 * cs - 2.f * h * h, (1.f + am1) * s};
    }
    __syncthreads();
    fft_lds<float, N>(a, b, nb, tw, -1, tid, 256);
    cplx<float>* dst = spec + (size_t)an_i * n2 + (size_t)j0 * N;
    for (int t = tid; t < nb * N; t += 256) dst[t] = a[lx(t)];
}

// Rows of the inverse-transformed-along-y spectrum of u: inverse FFT along
// x, psi = 1 + u / (nx ny).
template <int N>
__global__ void __launch_bounds__(256) k_ser_wave(const cplx<float>* __restrict__ T, cplx<float>* __restrict__ out,
                                                  int ny, const cplx<float>* __restrict__ tw) {
    constexpr int NB = lds_batch(N, sizeof(cplx<float>), LDS_DHT_ROWS);
    __shared__ cplx<float> sm[2 * lds_len(NB * N)];
    cplx<float>* a = sm;
    cplx<float>* b = sm + lds_len(NB * N);
    const int tid = threadIdx.x, j0 = blockIdx.x * NB, nb = min(NB, ny - j0);
    for (int t = tid; t < nb * N; t += 256) a[lx(t)] = T[(size_t)j0 * N + t];
    __syncthreads();
    fft_lds<float, N>(a, b, nb, tw, +1, tid, 256);
    const float inv = (float)(1.0 / ((double)N * ny));
    for (int t = tid; t < nb * N; t += 256) {
        const cplx<float> v = a[lx(t)];
        out[(size_t)j0 * N + t] = {1.f + v.x * inv, v.y * inv};
    }
}

}  // namespace utr

namespace {

size_t dht_series_bytes() {
    const char* e = std::getenv("UTR_DHT_SERIES_BYTES");
    if (e && *e) {
        const long long v = std::atoll(e);
        if (v > 0) return (size_t)v;
    }
    return size_t(160) << 20;
}

int dht_cols_nb(int ny) { return utr::lds_batch(ny, (int)sizeof(cplx<float>), utr::LDS_DHT); }
int dht_rows_nb(int nx) { return utr::lds_batch(nx, (int)sizeof(cplx<float>), utr::LDS_DHT_ROWS); }

// Images [first, first+count) of a series context, on the device and
// already through the input image correction, into their slots.
int dht_series_add(dht_ctx* c, int first, int count, const float* d_img) {
    const size_t n2 = (size_t)c->g.nx * c->g.ny;
    UTR_HIP_TRY(hipMemcpyAsync(c->simg.p + (size_t)first * n2, d_img, sizeof(float) * n2 * count,
                               hipMemcpyDeviceToDevice, c->stream));
    UTR_HIP_TRY(hipStreamSynchronize(c->stream));
    for (int r = 0; r < count; r++) {
        if (!c->added[first + r]) c->n_added++;
        c->added[first + r] = 1;
    }
    return UTR_OK;
}

// Orientations [a0, a1) of one retrieval pass and their images [p0, p1).
void ser_pass(const dht_ctx* c, int a0, int& a1, int& p0, int& p1) {
    const size_t n2 = (size_t)c->g.nx * c->g.ny;
    const size_t cap = std::max<size_t>(1, c->A1.n / n2);
    p0 = c->first[a0];
    a1 = a0 + 1;
    while (a1 < c->nang && a1 - a0 < c->ser_batch &&
           (size_t)(c->first[a1] + c->count[a1] - p0) <= cap)
        a1++;
    p1 = c->first[a1 - 1] + c->count[a1 - 1];
}

int dht_series_retrieve(dht_ctx* c) {
    const auto& g = c->g;
    const size_t n2 = (size_t)g.nx * g.ny;
    const int gx = (g.nx + dht_cols_nb(g.ny) - 1) / dht_cols_nb(g.ny);
    const int gy = (g.ny + dht_rows_nb(g.nx) - 1) / dht_rows_nb(g.nx);
    hipEvent_t e0 = c->ev[0], e1 = c->ev[1];
    UTR_HIP_TRY(hipEventRecord(e0, c->stream));
    std::fill(c->iters.begin(), c->iters.end(), 0);
    for (int a0 = 0; a0 < c->nang;) {
        int a1, p0, p1;
        ser_pass(c, a0, a1, p0, p1);
        const int na = a1 - a0, np = p1 - p0;
        bool ok = false;
        if (c->method == 1) {
#define SER_ROWS(N, M)                                                                                     \
    {                                                                                                      \
        utr::k_ser_rows<N, M><<<dim3(gy, np), 256, 0, c->stream>>>(c->B.p, c->simg.p + (size_t)p0 * n2,        \
                                                                   c->A1.p, p0, g, c->ser_img.p,            \
                                                                   c->ser_done.p, c->twx.p);                \
        ok = true;                                                                                         \
    }
#define SER_ROWS0(N) SER_ROWS(N, 0)
#define SER_ROWS1(N) SER_ROWS(N, 1)
#define SER_COLS(N)                                                                                        \
    {                                                                                                      \
        utr::k_ser_cols<N><<<dim3(gx, na), 256, 0, c->stream>>>(c->B.p, p0, c->spec.p, g, c->ser_ang.p, a0, \
                                                                c->E.p, c->ser_part.p, c->ser_done.p, mode, \
                                                                c->twy.p);                                 \
        ok = true;                                                                                         \
    }
            UTR_DISPATCH_N(g.nx, SER_ROWS0)
            if (!ok) return set_err(UTR_E_UNSUPPORTED, "unsupported image width for the FFT");
            int mode = 0;
            ok = false;
            UTR_DISPATCH_N(g.ny, SER_COLS)
            if (!ok) return set_err(UTR_E_UNSUPPORTED, "unsupported image height for the FFT");
            utr::k_ser_reset<<<(na + 255) / 256, 256, 0, c->stream>>>(c->ser_done.p + a0, c->ser_iters.p + a0, na);
            UTR_HIP_TRY(hipGetLastError());
            mode = 1;
            for (int k = 1; k <= c->itermax; k++) {
                UTR_DISPATCH_N(g.nx, SER_ROWS1)
                UTR_DISPATCH_N(g.ny, SER_COLS)
                utr::k_ser_check<<<1, 256, 0, c->stream>>>(c->ser_part.p, gx, a0, na, g.epsilon, c->ser_done.p,
                                                           c->ser_iters.p, c->ser_active.p);
                UTR_HIP_TRY(hipGetLastError());
                if (k % 8 == 0 && k < c->itermax) {
                    int active = 0;
                    UTR_HIP_TRY(hipMemcpyAsync(&active, c->ser_active.p, sizeof(int), hipMemcpyDeviceToHost,
                                               c->stream));
                    UTR_HIP_TRY(hipStreamSynchronize(c->stream));
                    if (active == 0) break;
                }
            }
#undef SER_ROWS
#undef SER_ROWS0
#undef SER_ROWS1
#undef SER_COLS
            UTR_HIP_TRY(hipMemcpyAsync(c->iters.data() + a0, c->ser_iters.p + a0, sizeof(int) * na,
                                       hipMemcpyDeviceToHost, c->stream));
        } else {
#define CTF_IN(N)                                                                                          \
    {                                                                                                      \
        constexpr int NB = utr::lds_batch(N, sizeof(cplx<float>), utr::LDS_DHT);                           \
        utr::k_dht_rows<N><<<dim3((g.ny + NB - 1) / NB, np), 256, 0, c->stream>>>(                          \
            c->simg.p + (size_t)p0 * n2, c->B.p, g.ny, g, c->twx.p);                                       \
        ok = true;                                                                                         \
    }
#define CTF_INC(N)                                                                                         \
    {                                                                                                      \
        utr::k_dht_cols<N><<<dim3(gx, np), 256, 0, c->stream>>>(c->B.p, g.nx, c->twy.p, -1);               \
        utr::k_ctf_cols<N><<<dim3(gx, na), 256, 0, c->stream>>>(c->B.p, p0, g, c->ser_img.p, c->ser_ang.p,  \
                                                                a0, c->twy.p);                             \
        ok = true;                                                                                         \
    }
#define CTF_OUT(N)                                                                                         \
    {                                                                                                      \
        utr::k_ctf_rows<N><<<dim3(gy, na), 256, 0, c->stream>>>(c->B.p, p0, c->simg.p, c->spec.p, g,       \
                                                                c->ser_ang.p, a0, c->twx.p);               \
        ok = true;                                                                                         \
    }
#define CTF_OUTC(N)                                                                                        \
    {                                                                                                      \
        utr::k_dht_cols<N><<<dim3(gx, na), 256, 0, c->stream>>>(c->spec.p + (size_t)a0 * n2, g.nx,         \
                                                                c->twy.p, -1);                             \
        ok = true;                                                                                         \
    }
            UTR_DISPATCH_N(g.nx, CTF_IN)
            if (ok) {
                ok = false;
                UTR_DISPATCH_N(g.ny, CTF_INC)
            }
            if (ok) {
                ok = false;
                UTR_DISPATCH_N(g.nx, CTF_OUT)
            }
            if (ok) {
                ok = false;
                UTR_DISPATCH_N(g.ny, CTF_OUTC)
            }
#undef CTF_IN
#undef CTF_INC
#undef CTF_OUT
#undef CTF_OUTC
            if (!ok) return set_err(UTR_E_UNSUPPORTED, "unsupported image size for the FFT");
        }
        UTR_HIP_TRY(hipGetLastError());
        a0 = a1;
    }
    UTR_HIP_TRY(hipEventRecord(e1, c->stream));
    UTR_HIP_TRY(hipEventSynchronize(e1));
    UTR_HIP_TRY(hipEventElapsedTime(&c->ret_ms, e0, e1));
    c->waves_ready = 1;
    return UTR_OK;
}

}  // namespace

extern "C" {

int dht_create_series(const dht_params* p, int nangles, const double* angZ_deg, const double* angY_deg,
                      const int* counts, const double* defocus, const float* weights, dht_ctx** out) {
    if (!out) return set_err(UTR_E_INVALID, "null output pointer");
    *out = nullptr;
    if (!p) return set_err(UTR_E_INVALID, "null parameters");
    if (nangles < 1 || !angZ_deg || !angY_deg || !defocus || !counts)
        return set_err(UTR_E_INVALID, "Error: the orientation file contains no rows.");
    std::string err = dht_validate(p);
    if (!err.empty()) return set_err(UTR_E_INVALID, err);
    long long total = 0;
    int cmin = counts[0], cmax = counts[0];
    for (int a = 0; a < nangles; a++) {
        if (counts[a] < 1) return set_err(UTR_E_INVALID, "every orientation needs at least one image");
        cmin = std::min(cmin, counts[a]);
        cmax = std::max(cmax, counts[a]);
        total += counts[a];
    }
    if (total > 65535LL * 256) return set_err(UTR_E_UNSUPPORTED, "too many input images");
    // DHT.cpp:272-283
    if ((p->method == 1 || p->method == 2) && cmin < 2)
        return set_err(UTR_E_INVALID,
                       "Error: selected phase retrieval method is incompatibe with a single defocus distance "
                       "appearing in some lines of the defocus/orientations parameter file.");
    if (p->method >= 3) {
        if (cmax != 1)
            return set_err(UTR_E_INVALID,
                           "Error: selected phase retrieval method is incompatibe with multiple defocus distances "
                           "in some lines of the defocus/orientations parameter file.");
        return dht_create(p, nangles, angZ_deg, angY_deg, defocus, weights, out);
    }
    for (long long i = 0; i < total; i++)
        if (!std::isfinite(defocus[i])) return set_err(UTR_E_INVALID, "defocus distances must be finite");
    dht_ctx* c = new dht_ctx();
    auto fail = [&](int rc) {
        dht_destroy(c);
        return rc;
    };
    c->method = p->method;
    c->series = 1;
    c->itermax = p->itermax;
    c->nrows = (int)total;
    c->added.assign((size_t)total, 0);
    c->first.resize(nangles);
    c->count.assign(counts, counts + nangles);
    c->iters.assign(nangles, 0);
    std::vector<double> zout(nangles);
    for (int a = 0, f = 0; a < nangles; a++) {
        c->first[a] = f;
        if (p->method == 1) {
            double z = 0.0;
            for (int k = 0; k < counts[a]; k++) z += defocus[f + k];
            zout[a] = z / (double)counts[a];  // DHT.cpp:895-897
        } else {
            zout[a] = defocus[f];  // DHT.cpp:908
        }
        f += counts[a];
    }
    int rc = dht_setup(c, p, nangles, angZ_deg, angY_deg, zout.data(), weights);
    if (rc) return fail(rc);
    const auto& g = c->g;
    const size_t n2 = (size_t)g.nx * g.ny;
    std::vector<utr::SerImg> im((size_t)total);
    std::vector<utr::SerAng> an(nangles);
    for (int a = 0; a < nangles; a++) {
        an[a] = {c->first[a], counts[a]};
        for (int k = 0; k < counts[a]; k++) {
            utr::SerImg& s = im[(size_t)c->first[a] + k];
            std::memset(&s, 0, sizeof s);
            s.d = defocus[c->first[a] + k] - zout[a];
            s.angle = a;
        }
    }
    // A retrieval pass iterates on the series of a group of orientations
    // whose planes (12 B per image pixel: field and amplitude) fit a budget
    // small enough to stay in the memory-side cache between the passes of
    // an iteration (UTR_DHT_SERIES_BYTES, default 160 MB; at most half the
    // DHT batch budget), and at least one orientation's series.
    size_t budget = std::min<size_t>(dht_series_bytes(), dht_batch_bytes() / 2);
    size_t cap = std::max<size_t>((size_t)cmax, budget / (n2 * (sizeof(cplx<float>) + sizeof(float))));
    cap = std::min<size_t>(cap, (size_t)total);
    cap = std::min<size_t>(cap, 65535);
    if ((size_t)cmax > cap) return fail(set_err(UTR_E_UNSUPPORTED, "too many images in one orientation"));
    c->ser_batch = 65535;
    const int gx = (g.nx + dht_cols_nb(g.ny) - 1) / dht_cols_nb(g.ny);
    if (c->simg.alloc((size_t)total * n2) || c->B.alloc(cap * n2) || c->A1.alloc(cap * n2) ||
        c->ser_img.alloc((size_t)total) ||
        c->ser_ang.alloc(nangles) || c->ser_part.alloc((size_t)std::min<size_t>(cap, 65535) * gx * 2) ||
        c->ser_done.alloc(nangles) || c->ser_iters.alloc(nangles) || c->ser_active.alloc(1))
        return fail(set_err(UTR_E_NOMEM, "device allocation of the series buffers failed"));
    if (hipMemcpy(c->ser_img.p, im.data(), sizeof(utr::SerImg) * im.size(), hipMemcpyHostToDevice) ||
        hipMemcpy(c->ser_ang.p, an.data(), sizeof(utr::SerAng) * an.size(), hipMemcpyHostToDevice))
        return fail(set_err(UTR_E_DEVICE, "upload of the series tables failed"));
    if (p->method == 1) {
        // per image: exp(-i pi wl d f^2) along x, then along y, phases in FP64 reduced mod 2 pi
        const double pi = 3.14159265358979323846, twopi = 2.0 * pi;
        const int ne = g.nx + g.ny;
        std::vector<cplx<float>> E((size_t)total * ne);
        auto freq = [&](int k, int n) { return (double)(k <= n / 2 ? k : k - n) / ((double)n * g.xst); };
        for (long long i = 0; i < total; i++)
            for (int k = 0; k < ne; k++) {
                const double f = k < g.nx ? freq(k, g.nx) : freq(k - g.nx, g.ny);
                double ph = pi * g.wl * im[i].d * (f * f);
                ph -= twopi * std::rint(ph / twopi);
                E[(size_t)i * ne + k] = {(float)std::cos(ph), (float)-std::sin(ph)};
            }
        if (c->E.alloc(E.size()) ||
            hipMemcpy(c->E.p, E.data(), sizeof(cplx<float>) * E.size(), hipMemcpyHostToDevice))
            return fail(set_err(UTR_E_DEVICE, "upload of the propagator factors failed"));
    }
    *out = c;
    return UTR_OK;
}

int dht_get_iterations(const dht_ctx* c, int* iters) {
    if (!c || !iters) return set_err(UTR_E_INVALID, "null argument");
    if (!c->series) return set_err(UTR_E_STATE, "dht_get_iterations: not a defocus-series context");
    std::copy(c->iters.begin(), c->iters.end(), iters);
    return UTR_OK;
}

int dht_get_waves(dht_ctx* c, int first, int count, float* psi) {
    if (!c || !psi) return set_err(UTR_E_INVALID, "null argument");
    if (!c->series) return set_err(UTR_E_STATE, "dht_get_waves: not a defocus-series context");
    if (!c->waves_ready) return set_err(UTR_E_STATE, "dht_get_waves before dht_finalize");
    if (first < 0 || count < 0 || first + count > c->nang)
        return set_err(UTR_E_INVALID, "orientation range outside the table");
    const auto& g = c->g;
    const size_t n2 = (size_t)g.nx * g.ny;
    const int gx = (g.nx + dht_cols_nb(g.ny) - 1) / dht_cols_nb(g.ny);
    const int gy = (g.ny + dht_rows_nb(g.nx) - 1) / dht_rows_nb(g.nx);
    LocalBuf<cplx<float>> w;
    UTR_HIP_TRY(w.alloc(n2));
    for (int a = first; a < first + count; a++) {
        UTR_HIP_TRY(hipMemcpyAsync(c->T.p, c->spec.p + (size_t)a * n2, sizeof(cplx<float>) * n2,
                                   hipMemcpyDeviceToDevice, c->stream));
        bool ok = false;
#define SER_WC(N)                                                                                   \
    {                                                                                               \
        utr::k_dht_cols<N><<<dim3(gx, 1), 256, 0, c->stream>>>(c->T.p, g.nx, c->twy.p, +1);         \
        ok = true;                                                                                  \
    }
#define SER_WR(N)                                                                                   \
    {                                                                                               \
        utr::k_ser_wave<N><<<dim3(gy, 1), 256, 0, c->stream>>>(c->T.p, w.p, g.ny, c->twx.p);        \
        ok = true;                                                                                  \
    }
        UTR_DISPATCH_N(g.ny, SER_WC)
        if (ok) {
            ok = false;
            UTR_DISPATCH_N(g.nx, SER_WR)
        }
#undef SER_WC
#undef SER_WR
        if (!ok) return set_err(UTR_E_UNSUPPORTED, "unsupported image size for the FFT");
        UTR_HIP_TRY(hipGetLastError());
        UTR_HIP_TRY(hipMemcpyAsync(psi + (size_t)(a - first) * 2 * n2, w.p, sizeof(cplx<float>) * n2,
                                   hipMemcpyDeviceToHost, c->stream));
        UTR_HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return UTR_OK;
}

}  // extern "C"
