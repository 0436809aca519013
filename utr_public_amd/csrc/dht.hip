// dht.hip -- the reconstructions of the DHT program on gfx950
// (MultisliceCppTest/PhaseRetrieval/DHT.cpp), behind the dht_* C ABI of
// include/utr_hip.h.  Included by utr_api.hip (one translation unit).
// Methods 1 and 2 retrieve each orientation's wave from a defocus series in
// dht_series.hip and then take method 3's back-propagation path below.
//
// Methods 3 and 4 (real-space path):
//   k_dht_rows / k_dht_cols    phase-retrieval input and forward 2D FFT of each
//                              orientation (DHT.cpp:850-947)
//   k_dht_prop_rows / _cols    per orientation and output plane: Fresnel
//                              propagator (method 3) or weak-object CTF
//                              inversion (method 4), inverse 2D FFT,
//                              1-|psi|^2 or Re (DHT.cpp:957-1055); phase
//                              arguments in FP64
//   k_dht_splat                real-space inverse rotation and trilinear splat
//                              with (int) truncation (DHT.cpp:1081-1164): one
//                              workgroup per 16x8x8 brick, exact 64-bit
//                              fixed-point sums in LDS, coordinates evaluated
//                              in FP64 in the reference's operation order
//   k_dht_splat_col            the same sums in columns through y when no
//                              orientation rotates about Z
//   k_dht_finish               scale of DHT.cpp:1325-1340
// Orientations are processed in batches whose plane stacks fit a memory
// budget; each batch adds its exact sums to an FP64 volume.
//
// Method 5 (sCTF) runs the UTR Ewald-sphere kernels with the DHT voxel
// filter, no sphere mask and the DHT scale (DHT.cpp:1056-1067, 1195-1416).

namespace utr {

struct DhtGeom {
    int nx, ny, nz, nplanes, method, escc, use_f1, use_f2;
    float f1, f2;
    double xst, wl, k2maxo, c3, c5, asigma, epsilon, zlodz;
    double xlo, ylo, xc, yc, zc;
    int nx2, ny2, nz2;
    int use_ap;  // aperture or Cs terms present: filters read the table ap
};

// One orientation of the DHT path.
struct alignas(16) DhtRow {
    double sinY, cosY, sinZ, cosZ;
    double zout;
    float weight;
    float c0[3];   // grid position (dx1, dy1, dz1) of source sample (i, j, n) = (0, 0, 0)
    float ui[3];   // its derivative along i, j, n
    float uj[3];
    float un[3];
    float inv[9];  // inverse of the 3x3 map [ui uj un]
    int pad;
};

// Methods 1 and 2: one input image of an orientation's defocus series
// (d = its distance minus the orientation's zout) and one orientation's
// images first .. first+count-1.
struct alignas(16) SerImg {
    double d;
    int angle;
    int pad;
};
struct SerAng {
    int first, count;
};

constexpr int DHT_BX = 16, DHT_BY = 8, DHT_BZ = 8;
constexpr int DHT_PY = DHT_BX + 1, DHT_PZ = DHT_PY * (DHT_BY + 1), DHT_PAD = DHT_PZ * DHT_BZ;
constexpr int DHT_VOX = DHT_BX * DHT_BY * DHT_BZ;

__device__ __forceinline__ double dht_freq(int k, int n, double xst) {
    return (double)(k <= n / 2 ? k : k - n) / ((double)n * xst);
}

// Image rows: renormalise and clamp (DHT.cpp:850-858), the field of the
// method (3: sqrt(I) exp(i Fact (-0.5)(I - 1)) - 1; 4: I - 1) and forward FFT
// along x in natural order.
template <int N>
__global__ void __launch_bounds__(256) k_dht_rows(const float* __restrict__ img,
                                                  cplx<float>* __restrict__ out, int ny, DhtGeom g,
                                                  const cplx<float>* __restrict__ tw) {
    constexpr int NB = lds_batch(N, sizeof(cplx<float>), LDS_DHT);
    __shared__ cplx<float> sm[2 * lds_len(NB * N)];
    cplx<float>* a = sm;
    cplx<float>* b = sm + lds_len(NB * N);
    const int row = blockIdx.y;
    const int j0 = blockIdx.x * NB;
    const int nb = min(NB, ny - j0);
    const float* src = img + ((size_t)row * ny + j0) * N;
    for (int t = threadIdx.x; t < nb * N; t += blockDim.x) {
        double x = (double)src[t];
        if (g.use_f1) x = x / (double)g.f1;
        if (g.use_f2) x = x + (double)g.f2;
        if (x < 0.0) x = 0.0;
        if (g.method == 3) {
            // the scattered part u = psi - 1 is carried instead of psi (the
            // propagators leave the constant 1 unchanged), so that 1 - |psi|^2
            // = -2 Re u - |u|^2 keeps its relative precision in FP32
            const double amp = sqrt(x), phi = g.epsilon * (-0.5) * (x - 1.0);
            double s, c;
            sincos(phi, &s, &c);
            a[lx(t)] = {(float)(amp * c - 1.0), (float)(amp * s)};
        } else {
            a[lx(t)] = {(float)(x - 1.0), 0.f};
        }
    }
    __syncthreads();
    fft_lds<float, N>(a, b, nb, tw, -1, threadIdx.x, blockDim.x);
    cplx<float>* dst = out + ((size_t)row * ny + j0) * N;
    for (int t = threadIdx.x; t < nb * N; t += blockDim.x) dst[t] = a[lx(t)];
}

// Forward (sign -1) or inverse (+1) FFT along y of `count` ny x nx arrays, in place.
template <int N>
__global__ void __launch_bounds__(256) k_dht_cols(cplx<float>* __restrict__ data, int nx,
                                                  const cplx<float>* __restrict__ tw, int sign) {
    constexpr int NB = lds_batch(N, sizeof(cplx<float>), LDS_DHT);
    __shared__ cplx<float> sm[2 * lds_len(NB * N)];
    cplx<float>* a = sm;
    cplx<float>* b = sm + lds_len(NB * N);
    const int x0 = blockIdx.x * NB;
    const int nb = min(NB, nx - x0);
    cplx<float>* base = data + (size_t)blockIdx.y * N * nx + x0;
    for (int t = threadIdx.x; t < nb * N; t += blockDim.x) {
        const int y = t / nb, c = t - y * nb;
        a[lx(c * N + y)] = base[(size_t)y * nx + c];
    }
    __syncthreads();
    fft_lds<float, N>(a, b, nb, tw, sign, threadIdx.x, blockDim.x);
    for (int t = threadIdx.x; t < nb * N; t += blockDim.x) {
        const int y = t / nb, c = t - y * nb;
        base[(size_t)y * nx + c] = a[lx(c * N + y)];
    }
}

// Per slot s = blockIdx.y of orientation a0+blockIdx.z: spectrum row times
// the filter of the slot's plane(s), inverse FFT along x.  grid (rows/NB,
// slots, na).  Method 3: slot s is plane s.  Method 4: the filtered spectrum
// of a plane is Hermitian (the spectrum of the real I - 1 times a real, even
// filter), so its inverse transform is real, and slot s carries plane 2s in
// the real part and plane 2s+1 in the imaginary part (dht_slots).
// The phase chi = f2 q^2 + f4 q^4 + f6 q^6 (DHT.cpp:962-1000) is split as
// exp(-i chi) = ex(fx) ey(fy) ap(fx, fy): ex = exp(-i f2 fx^2) and
// ey = exp(-i f2 fy^2) depend on the plane's defocus and are formed per
// workgroup with FP64 phases reduced to [-pi, pi] (FP32 sin and cos of the
// reduced phase); ap = exp(-i (f4 q^4 + f6 q^6)), zero outside the
// objective aperture, is one table per job (read when g.use_ap).
__host__ __device__ inline int dht_slots(const DhtGeom& g) {
    return g.method == 4 ? (g.nplanes + 1) / 2 : g.nplanes;
}

template <int N>
__global__ void __launch_bounds__(256) k_dht_prop_rows(const cplx<float>* __restrict__ S,
                                                       cplx<float>* __restrict__ T, DhtGeom g,
                                                       const DhtRow* __restrict__ rows, int a0,
                                                       const cplx<float>* __restrict__ tw,
                                                       const cplx<float>* __restrict__ ap) {
    constexpr int NB = lds_batch(N, sizeof(cplx<float>), LDS_DHT_ROWS);
    constexpr int PER = (NB * N + 255) / 256;  // elements per thread
    __shared__ cplx<float> sm[2 * lds_len(NB * N)];
    __shared__ cplx<float> s_ex[2][N], s_ey[2][NB];
    cplx<float>* a = sm;
    cplx<float>* b = sm + lds_len(NB * N);
    const int ny = g.ny;
    const int slot = blockIdx.y, ab = blockIdx.z;
    const bool pack = g.method == 4;
    const int p0 = pack ? 2 * slot : slot;
    const int np = pack && p0 + 1 < g.nplanes ? 2 : 1;  // planes in this slot
    const int k0 = blockIdx.x * NB;
    const int nb = min(NB, ny - k0);
    const int tid = threadIdx.x;
    const DhtRow& R = rows[a0 + ab];
    const double pi = 3.14159265358979323846, twopi = 2.0 * pi;
    const cplx<float>* src = S + ((size_t)(a0 + ab) * ny + k0) * N;
    const cplx<float>* aps = ap + (size_t)k0 * N;
    cplx<float> v[PER], h[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int t = tid + 256 * q;
        if (t < nb * N) {
            v[q] = src[t];
            h[q] = g.use_ap ? aps[t] : cplx<float>{1.f, 0.f};
        }
    }
    for (int k = tid; k < np * (N + nb); k += 256) {
        const int m = k >= N + nb ? 1 : 0, kk = k - m * (N + nb);
        const int n = g.escc ? p0 + m : g.nz / 2;
        const double zn = g.zlodz + g.xst * n;  // voutdefocus[n], DHT.cpp:388
        const double d = g.method != 4 ? zn - R.zout : R.zout - zn;  // 1-3: back-propagation
        const double f2 = pi * g.wl * d;
        const double f = kk < N ? dht_freq(kk, N, g.xst) : dht_freq(k0 + kk - N, ny, g.xst);
        // the phase and its reduction to [-pi, pi] in FP64; sin and cos of
        // the reduced phase in FP32
        double chi = f2 * (f * f);
        chi -= twopi * rint(chi * (1.0 / twopi));
        float sn, cs;
        sincosf((float)chi, &sn, &cs);
        const cplx<float> e = {cs, -sn};
        if (kk < N) s_ex[m][kk] = e;
        else s_ey[m][kk - N] = e;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int t = tid + 256 * q;
        if (t < nb * N) {
            const int r = t / N, kx = t - r * N;
            const bool in = h[q].x != 0.f || h[q].y != 0.f;  // |ap| = 1 inside the aperture, 0 outside
            if (!pack) {
                // cos chi - i sin chi (the aperture table only when present)
                cplx<float> H = cmul(s_ex[0][kx], s_ey[0][r]);
                if (g.use_ap) H = cmul(H, h[q]);
                const float fr = H.x, fi = H.y;
                a[lx(t)] = {v[q].x * fr - v[q].y * fi, v[q].x * fi + v[q].y * fr};
            } else {
                float f[2] = {0.f, 0.f};
#pragma unroll
                for (int m = 0; m < 2; m++) {
                    if (m < np && in) {
                        const cplx<float> H = cmul(cmul(s_ex[m][kx], s_ey[m][r]), h[q]);
                        const float tt = -H.y - (float)g.asigma * H.x;  // sin chi - asigma cos chi
                        f[m] = tt / (2.f * (tt * tt + (float)g.epsilon));
                    }
                }
                // plane p0 times f[0] plus i times plane p0+1 times f[1]
                a[lx(t)] = {v[q].x * f[0] - v[q].y * f[1], v[q].y * f[0] + v[q].x * f[1]};
            }
        }
    }
    __syncthreads();
    fft_lds<float, N>(a, b, nb, tw, +1, tid, blockDim.x);
    cplx<float>* dst = T + (((size_t)ab * dht_slots(g) + slot) * ny + k0) * N;
    for (int t = tid; t < nb * N; t += blockDim.x) dst[t] = a[lx(t)];
}

// Inverse FFT along y of each slot, then 1 - |psi|^2 = -2 Re u - |u|^2
// (method 3) or the real and imaginary parts as planes 2s and 2s+1
// (method 4), times the orientation's weight.
// Tile w = NB adjacent columns of one slot: column group w % gx, slot
// w / gx of the batch (orientation-major).  Persistent: in round k this
// workgroup takes tile k G + (L & 7) G/8 + L/8, so the tiles of one XCD in a
// round are contiguous and neighbouring column groups share its L2; the
// next tile's columns are loaded into registers while the current one is
// transformed.  gridDim.x is a multiple of 8.
// pmax: max |value| of the batch as float bits.
//
// COL: the planes are stored column-major ([plane][x][y], y fastest) for
// the column splat; otherwise row-major ([plane][y][x]).
template <int N, bool COL>
__global__ void __launch_bounds__(256) k_dht_prop_cols(const cplx<float>* __restrict__ T,
                                                       float* __restrict__ P, DhtGeom g,
                                                       const DhtRow* __restrict__ rows, int a0,
                                                       const cplx<float>* __restrict__ tw,
                                                       unsigned* __restrict__ pmax, int gx, int total) {
    constexpr int NB = lds_batch(N, sizeof(cplx<float>), LDS_DHT);
    constexpr int PER = (NB * N + 255) / 256;  // elements per thread
    __shared__ cplx<float> sm[2 * lds_len(NB * N)];
    __shared__ float s_wmax[4];
    const int nx = g.nx;
    const int tid = threadIdx.x;
    const unsigned L = blockIdx.x, G = gridDim.x;
    auto tile_of = [&](int k) { return (int)(k * G + (L & 7u) * (G >> 3) + (L >> 3)); };
    // element t of a tile is (row t / cols, column t % cols); a full tile
    // (the common case) divides by the constant NB
    cplx<float> vn[PER];
    auto fetch_cols = [&](const cplx<float>* base, const int cols) {
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int t = tid + 256 * q;
            if (t < cols * N) {
                const int y = t / cols, c = t - y * cols;
                vn[q] = base[(size_t)y * nx + c];
            }
        }
    };
    auto fetch = [&](int w) {
        const int pa = w / gx, x0 = (w - pa * gx) * NB, nb = min(NB, nx - x0);
        const cplx<float>* base = T + (size_t)pa * N * nx + x0;
        if (nb == NB) fetch_cols(base, NB);
        else fetch_cols(base, nb);
    };
    float vmax = 0.f;
    const int rounds = (total + (int)G - 1) / (int)G;
    if (tile_of(0) < total) fetch(tile_of(0));
    for (int rk = 0; rk < rounds; rk++) {
        const int wt = tile_of(rk);
        cplx<float> vc[PER];
#pragma unroll
        for (int q = 0; q < PER; q++) vc[q] = vn[q];
        if (rk + 1 < rounds && tile_of(rk + 1) < total) fetch(tile_of(rk + 1));
        if (wt >= total) continue;  // uniform over the workgroup
        cplx<float>* a = sm;
        cplx<float>* b = sm + lds_len(NB * N);
        const int pa = wt / gx, x0 = (wt - pa * gx) * NB, nb = min(NB, nx - x0);
        const int nslots = dht_slots(g), ab = pa / nslots, slot = pa - ab * nslots;
        // method 4: the slot's real part is plane 2 slot, its imaginary part plane 2 slot + 1
        const bool pack = g.method == 4;
        const int pl0 = ab * g.nplanes + (pack ? 2 * slot : slot);
        const bool second = pack && (pack ? 2 * slot : slot) + 1 < g.nplanes;
        const size_t plane = (size_t)pl0 * N * nx;
        auto put = [&](const int cols) {
#pragma unroll
            for (int q = 0; q < PER; q++) {
                const int t = tid + 256 * q;
                if (t < cols * N) {
                    const int y = t / cols, c = t - y * cols;
                    a[lx(c * N + y)] = vc[q];
                }
            }
        };
        if (nb == NB) put(NB);
        else put(nb);
        __syncthreads();
        fft_lds<float, N>(a, b, nb, tw, +1, tid, blockDim.x);
        const float inv = (float)(1.0 / ((double)nx * N));
        const float w = rows[a0 + ab].weight;
        auto store = [&](const int cols) {
#pragma unroll
            for (int q = 0; q < PER; q++) {
                const int t = tid + 256 * q;
                if (t < cols * N) {
                    // column-major: consecutive threads take consecutive y
                    const int y = COL ? t % N : t / cols, c = COL ? t / N : t - y * cols;
                    const cplx<float> z = a[lx(c * N + y)];
                    const float re = z.x * inv, im = z.y * inv;
                    const float v = (pack ? re : -2.f * re - (re * re + im * im)) * w;
                    const size_t o = COL ? (size_t)(x0 + c) * N + y : (size_t)y * nx + x0 + c;
                    P[plane + o] = v;
                    vmax = fmaxf(vmax, fabsf(v));
                    if (second) {
                        const float v2 = im * w;
                        P[plane + (size_t)N * nx + o] = v2;
                        vmax = fmaxf(vmax, fabsf(v2));
                    }
                }
            }
        };
        if (nb == NB) store(NB);
        else store(nb);
        __syncthreads();  // the next tile overwrites both buffers
    }
    for (int off = 32; off > 0; off >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, off));
    if ((tid & 63) == 0) s_wmax[tid >> 6] = vmax;
    __syncthreads();
    if (tid == 0)
        atomicMax(pmax, __float_as_uint(fmaxf(fmaxf(s_wmax[0], s_wmax[1]), fmaxf(s_wmax[2], s_wmax[3]))));
}

__device__ __forceinline__ int dht_lds(int lx, int ly, int lz) { return lx + DHT_PY * ly + DHT_PZ * lz; }

// Splat of the batch's plane stacks (orientations a0 .. a0+na-1) into one
// brick, added to the FP64 volume acc3 (the brick's own voxels only).
// xs/ys/zs: per orientation (x_i - xc) sinY, (x_i - xc) cosY; (y_j - yc) sinZ,
// (y_j - yc) cosZ; (z_n - zc) sinY, (z_n - zc) cosY, as DHT.cpp:1083-1107.
__global__ void __launch_bounds__(256) k_dht_splat(const float* __restrict__ P, DhtGeom g,
                                                   const DhtRow* __restrict__ rows,
                                                   const double* __restrict__ xs,
                                                   const double* __restrict__ ys,
                                                   const double* __restrict__ zs, int a0, int na,
                                                   const unsigned* __restrict__ pmax,
                                                   double* __restrict__ acc3) {
    __shared__ unsigned long long acc[DHT_PAD];
    const int tid = threadIdx.x;
    const int nbx = (g.nx + DHT_BX - 1) / DHT_BX, nby = (g.ny + DHT_BY - 1) / DHT_BY;
    const int bid = blockIdx.x;
    const int bx = bid % nbx, by = (bid / nbx) % nby, bz = bid / (nbx * nby);
    const int x0 = bx * DHT_BX, y0 = by * DHT_BY, z0 = bz * DHT_BZ;
    for (int t = tid; t < DHT_PAD; t += 256) acc[t] = 0ull;
    // fixed-point scale: per orientation a voxel takes at most 27 samples with
    // |weight| <= 8 (truncation allows dz1 in (-1, 0)), so |sum| < 2^8 na max|P|
    int ek = 0, ea = 0;
    const float mk = __uint_as_float(*pmax);
    frexpf(mk > 0.f ? mk : 1.f, &ek);
    frexpf((float)na, &ea);
    const int s = min(53 - ek - ea, 47 - ek);
    const double scale = ldexp(1.0, s);
    __syncthreads();

    const int np_stride = g.nplanes;
    const size_t n2 = (size_t)g.nx * g.ny;
    for (int ab = 0; ab < na; ab++) {
        const DhtRow& R = rows[a0 + ab];
        // source-index box of the brick's preimage: samples whose truncated
        // cell can touch the brick lie in the open box (x0-1, x0+BX) etc.
        float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
        const float bl[3] = {(float)x0 - 1.f, (float)y0 - 1.f, (float)z0 - 1.f};
        const float bh[3] = {(float)(x0 + DHT_BX), (float)(y0 + DHT_BY), (float)(z0 + DHT_BZ)};
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const float q[3] = {((c & 1) ? bh[0] : bl[0]) - R.c0[0], ((c & 2) ? bh[1] : bl[1]) - R.c0[1],
                                ((c & 4) ? bh[2] : bl[2]) - R.c0[2]};
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const float v = R.inv[3 * d] * q[0] + R.inv[3 * d + 1] * q[1] + R.inv[3 * d + 2] * q[2];
                lo[d] = fminf(lo[d], v);
                hi[d] = fmaxf(hi[d], v);
            }
        }
        // lo/hi over (i, j, n); one sample of margin
        const int jlo = max((int)floorf(lo[1]) - 1, 0), jhi = min((int)ceilf(hi[1]) + 1, g.ny - 1);
        const int nlo = max((int)floorf(lo[2]) - 1, 0), nhi = min((int)ceilf(hi[2]) + 1, g.nz - 1);
        const int ilo0 = max((int)floorf(lo[0]) - 1, 0), ihi0 = min((int)ceilf(hi[0]) + 1, g.nx - 1);
        if (jlo > jhi || nlo > nhi || ilo0 > ihi0) continue;
        const int nj = jhi - jlo + 1;
        const int nlines = nj * (nhi - nlo + 1);
        const double* X = xs + (size_t)(a0 + ab) * 2 * g.nx;
        const double* Y = ys + (size_t)(a0 + ab) * 2 * g.ny;
        const double* Z = zs + (size_t)(a0 + ab) * 2 * g.nz;
        for (int L = tid; L < nlines; L += 256) {
            const int n = nlo + L / nj, j = jlo + L % nj;
            // i-interval of this line inside the box (slab test, one sample of margin)
            float ilo = (float)ilo0, ihi = (float)ihi0;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const float base = R.c0[d] + R.uj[d] * j + R.un[d] * n;
                const float u = R.ui[d];
                const float l = bl[d] - 0.05f - base, h = bh[d] + 0.05f - base;
                if (fabsf(u) < 1e-6f) {
                    if (l > 0.f || h < 0.f) ihi = ilo - 1.f;
                } else {
                    float t0 = l / u, t1 = h / u;
                    if (u < 0.f) { const float tt = t0; t0 = t1; t1 = tt; }
                    ilo = fmaxf(ilo, floorf(t0) - 1.f);
                    ihi = fminf(ihi, ceilf(t1) + 1.f);
                }
            }
            ilo = fminf(ilo, (float)ihi0 + 1.f);  // keep the conversions in range
            ihi = fmaxf(ihi, (float)ilo0 - 1.f);
            const int i0 = (int)ilo, i1 = (int)ihi;
            const float* Kp = P + ((size_t)ab * np_stride + (g.escc ? n : 0)) * n2 + (size_t)j * g.nx;
            const double zsin = Z[2 * n], zcos = Z[2 * n + 1];
            const double ysin = Y[2 * j], ycos = Y[2 * j + 1];
            for (int i = i0; i <= i1; i++) {
                // DHT.cpp:1119-1150, operation by operation
                const double zzz = __dadd_rn(__dadd_rn(g.zc, zcos), X[2 * i]);
                double dz1 = __ddiv_rn(__dsub_rn(zzz, g.zlodz), g.xst);
                const int nn = (int)dz1;
                if (nn < 0 || nn > g.nz2) continue;
                dz1 = __dsub_rn(dz1, (double)nn);
                const double dz0 = __dsub_rn(1.0, dz1);
                const double xx = __dadd_rn(-zsin, X[2 * i + 1]);
                const double xx_sinZ = __dmul_rn(xx, R.sinZ), xx_cosZ = __dmul_rn(xx, R.cosZ);
                const double xxx = __dadd_rn(__dadd_rn(g.xc, ysin), xx_cosZ);
                double dx1 = __ddiv_rn(__dsub_rn(xxx, g.xlo), g.xst);
                const int ii = (int)dx1;
                if (ii < 0 || ii > g.nx2) continue;
                dx1 = __dsub_rn(dx1, (double)ii);
                const double dx0 = __dsub_rn(1.0, dx1);
                const double yyy = __dsub_rn(__dadd_rn(g.yc, ycos), xx_sinZ);
                double dy1 = __ddiv_rn(__dsub_rn(yyy, g.ylo), g.xst);
                const int jj = (int)dy1;
                if (jj < 0 || jj > g.ny2) continue;
                dy1 = __dsub_rn(dy1, (double)jj);
                const double dy0 = __dsub_rn(1.0, dy1);
                const int lx = ii - x0, ly = jj - y0, lz = nn - z0;
                if (lx < -1 || lx >= DHT_BX || ly < -1 || ly >= DHT_BY || lz < -1 || lz >= DHT_BZ) continue;
                const double k = (double)Kp[i];
                const double dz0K = dz0 * k, dz1K = dz1 * k;
                const double wx[2] = {dx0, dx1}, wy[2] = {dy0, dy1}, wz[2] = {dz0K, dz1K};
#pragma unroll
                for (int c = 0; c < 8; c++) {
                    const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
                    const int vx = lx + cx, vy = ly + cy, vz = lz + cz;
                    if ((unsigned)vx >= (unsigned)DHT_BX || (unsigned)vy >= (unsigned)DHT_BY ||
                        (unsigned)vz >= (unsigned)DHT_BZ)
                        continue;
                    const double w = wx[cx] * wy[cy] * wz[cz];
                    atomicAdd(&acc[dht_lds(vx, vy, vz)], (unsigned long long)dround_i64(w * scale));
                }
            }
        }
    }
    __syncthreads();
    const double inv = ldexp(1.0, -s);
    for (int v = tid; v < DHT_VOX; v += 256) {
        const int lx = v % DHT_BX, ly = (v / DHT_BX) % DHT_BY, lz = v / (DHT_BX * DHT_BY);
        const int vx = x0 + lx, vy = y0 + ly, vz = z0 + lz;
        if (vx >= g.nx || vy >= g.ny || vz >= g.nz) continue;
        const size_t idx = ((size_t)vz * g.ny + vy) * g.nx + vx;
        acc3[idx] += (double)(long long)acc[dht_lds(lx, ly, lz)] * inv;
    }
}

__global__ void k_dht_zero_u32(unsigned* __restrict__ v) {
    if (threadIdx.x == 0) v[0] = 0u;
}

__global__ void k_dht_fill(double* __restrict__ v, size_t n) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) v[t] = 0.0;
}

__global__ void k_dht_finish(const double* __restrict__ acc3, float* __restrict__ vol, double dnorm,
                             size_t n) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) vol[t] = (float)(acc3[t] * dnorm);
}

}  // namespace utr

namespace utr {

// Splat of the batch's column-major plane stacks into columns of the volume
// when no orientation rotates about Z (every angZ = 0, a tilt series about
// y).  Then sinZ = 0 and cosZ = 1: a sample's x and z grid coordinates
// depend on its (i, n) only, and its y coordinate on its line j only (line j
// lands in cell jj(j), weights 1 - dy1(j) and dy1(j)).  Each workgroup owns
// a DC_X x DC_Z (x, z) column of the volume through all y, two threads per
// source line j (one takes the even listed samples, the other the odd):
//   per orientation, the threads test the source samples (i, n) around the
//   column (FP64, the reference's expressions of DHT.cpp:1119-1150) and list
//   those whose cell (ii, nn) puts a corner in the column, with their x and
//   z weights; then every thread splats its line j of each listed sample:
//   the values P[n][i][0..ny) of one sample are contiguous, and the threads
//   add to consecutive y entries of the column's sums.
// The fixed-point terms are those of k_dht_splat expression by expression,
// so both kernels form the same sums.
constexpr int DC_X = 4, DC_Z = 4;
// candidate half-width (source samples): the cells a column's corners reach
// form a 5 x 5 square about its centre, whose corners lie 2.5 sqrt(2) = 3.54
// samples away under any rotation, so a box of floor(centre) +- 4 holds
// every sample that can reach the column; 5 leaves a sample of margin for
// the FP32 centre
constexpr int DC_R = 5;
constexpr int DC_NC = (2 * DC_R + 1) * (2 * DC_R + 1);
constexpr int DC_NYMAX = 384;                 // tallest image (column sums of 48 KB in LDS)
constexpr int DC_T = 512;                     // threads per workgroup
static_assert(DC_NC <= DC_T, "one candidate per thread");

__global__ void __launch_bounds__(DC_T) k_dht_splat_col(const float* __restrict__ P, DhtGeom g,
                                                       const DhtRow* __restrict__ rows,
                                                       const double* __restrict__ xs,
                                                       const double* __restrict__ ys,
                                                       const double* __restrict__ zs, int a0, int na,
                                                       const unsigned* __restrict__ pmax,
                                                       double* __restrict__ acc3, int ntx, int ntz) {
    extern __shared__ unsigned long long acc[];  // [DC_Z][DC_X][ny]
    __shared__ short s_i[DC_NC], s_n[DC_NC];
    __shared__ signed char s_lx[DC_NC], s_lz[DC_NC];
    __shared__ double s_dx1[DC_NC], s_dz1[DC_NC];
    __shared__ int s_cnt[DC_T / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tile = (int)blockIdx.x;
    if (tile >= ntx * ntz) return;
    const int x0 = (tile % ntx) * DC_X, z0 = (tile / ntx) * DC_Z;
    const int ny = g.ny;
    for (int t = tid; t < DC_X * DC_Z * ny; t += DC_T) acc[t] = 0ull;
    // the scale of k_dht_splat
    int ek = 0, ea = 0;
    const float mk = __uint_as_float(*pmax);
    frexpf(mk > 0.f ? mk : 1.f, &ek);
    frexpf((float)na, &ea);
    const int sc = min(53 - ek - ea, 47 - ek);
    const double scale = ldexp(1.0, sc);
    const size_t n2 = (size_t)g.nx * ny;
    const int nsrc = g.nz;
    // grid centre offsets: dx1 = cxs + (i - cxs) cosY - (n - czs) sinY and
    // dz1 = czs + (n - czs) cosY + (i - cxs) sinY (candidate box only)
    const float cxs = (float)((g.xc - g.xlo) / g.xst), czs = (float)((g.zc - g.zlodz) / g.xst);
    const float X = (float)x0 + 0.5f * DC_X - 0.5f, Z = (float)z0 + 0.5f * DC_Z - 0.5f;
    for (int ab = 0; ab < na; ab++) {
        const DhtRow& R = rows[a0 + ab];
        const double* Xt = xs + (size_t)(a0 + ab) * 2 * g.nx;
        const double* Yt = ys + (size_t)(a0 + ab) * 2 * ny;
        const double* Zt = zs + (size_t)(a0 + ab) * 2 * g.nz;
        __syncthreads();  // the previous orientation's list is consumed (and acc zeroed)
        // candidates (i, n) in a box about the column's centre
        bool pass = false;
        int ci = 0, cn = 0, lxv = 0, lzv = 0;
        double dx1 = 0.0, dz1 = 0.0;
        if (tid < DC_NC) {
            const float sY = (float)R.sinY, cY = (float)R.cosY;
            const float ua = X - cxs, vb = Z - czs;
            const float ic = cxs + ua * cY + vb * sY, nc = czs - ua * sY + vb * cY;
            ci = (int)floorf(ic) - DC_R + tid % (2 * DC_R + 1);
            cn = (int)floorf(nc) - DC_R + tid / (2 * DC_R + 1);
            if (ci >= 0 && ci < g.nx && cn >= 0 && cn < nsrc) {
                // DHT.cpp:1119-1150, operation by operation
                const double zcos = Zt[2 * cn + 1], zsin = Zt[2 * cn];
                const double zzz = __dadd_rn(__dadd_rn(g.zc, zcos), Xt[2 * ci]);
                dz1 = __ddiv_rn(__dsub_rn(zzz, g.zlodz), g.xst);
                const int nn = (int)dz1;
                const double ysin = Yt[0];  // y_j sinZ: a zero of either sign, same sum
                const double xx = __dadd_rn(-zsin, Xt[2 * ci + 1]);
                const double xx_cosZ = __dmul_rn(xx, R.cosZ);
                const double xxx = __dadd_rn(__dadd_rn(g.xc, ysin), xx_cosZ);
                dx1 = __ddiv_rn(__dsub_rn(xxx, g.xlo), g.xst);
                const int ii = (int)dx1;
                lxv = ii - x0;
                lzv = nn - z0;
                pass = nn >= 0 && nn <= g.nz2 && ii >= 0 && ii <= g.nx2 && lxv >= -1 && lxv < DC_X &&
                       lzv >= -1 && lzv < DC_Z;
                if (pass) {
                    dz1 = __dsub_rn(dz1, (double)nn);
                    dx1 = __dsub_rn(dx1, (double)ii);
                }
            }
        }
        const unsigned long long mask = __ballot(pass);
        if (lane == 0) s_cnt[wave] = __popcll(mask);
        __syncthreads();
        int off = __popcll(mask & ((1ull << lane) - 1ull)), nl = 0;
        for (int w = 0; w < DC_T / 64; w++) {
            if (w < wave) off += s_cnt[w];
            nl += s_cnt[w];
        }
        if (pass) {
            s_i[off] = (short)ci;
            s_n[off] = (short)cn;
            s_lx[off] = (signed char)lxv;
            s_lz[off] = (signed char)lzv;
            s_dx1[off] = dx1;
            s_dz1[off] = dz1;
        }
        __syncthreads();  // the list is written
        if (nl == 0) continue;  // uniform
        const float* Pab = P + (size_t)ab * g.nplanes * n2;
        // (line j, parity) pairs: thread q takes line q % ny and the listed
        // samples of parity q / ny
        for (int q = tid; q < 2 * ny; q += DC_T) {
            const int j = q % ny, par = q / ny;
            if (par >= nl) continue;
            // y of source line j as DHT.cpp:1143-1148 forms it (xx sinZ = 0)
            const double yyy = __dsub_rn(__dadd_rn(g.yc, Yt[2 * j + 1]), __dmul_rn(0.0, R.sinZ));
            double dy1 = __ddiv_rn(__dsub_rn(yyy, g.ylo), g.xst);
            const int jj = (int)dy1;
            if (jj < 0 || jj > g.ny2) continue;
            dy1 = __dsub_rn(dy1, (double)jj);
            const double wy[2] = {__dsub_rn(1.0, dy1), dy1};
            auto value = [&](int e) {
                return Pab[((size_t)(g.escc ? s_n[e] : 0) * g.nx + s_i[e]) * ny + j];
            };
            // the values of the next two samples loaded ahead
            float kv = value(par);
            float kv2 = par + 2 < nl ? value(par + 2) : 0.f;
            for (int e = par; e < nl; e += 2) {
                const float kraw = kv;
                kv = kv2;
                if (e + 4 < nl) kv2 = value(e + 4);
                const int lx = s_lx[e], lz = s_lz[e];
                const double ex1 = s_dx1[e], ez1 = s_dz1[e];
                const double wx[2] = {__dsub_rn(1.0, ex1), ex1};
                const double k = (double)kraw;
                const double wz[2] = {__dsub_rn(1.0, ez1) * k, ez1 * k};
#pragma unroll
                for (int cy = 0; cy < 2; cy++) {
                    unsigned long long* ap = acc + jj + cy;
#pragma unroll
                    for (int c = 0; c < 4; c++) {
                        const int cx = c & 1, cz = c >> 1;
                        const int vx = lx + cx, vz = lz + cz;
                        if ((unsigned)vx >= (unsigned)DC_X || (unsigned)vz >= (unsigned)DC_Z) continue;
                        const double w = wx[cx] * wy[cy] * wz[cz];
                        atomicAdd(&ap[(size_t)(vz * DC_X + vx) * ny], (unsigned long long)dround_i64(w * scale));
                    }
                }
            }
        }
    }
    __syncthreads();
    const double inv = ldexp(1.0, -sc);
    // add the column's sums to the FP64 volume: lane -> (x, y) with x fastest
    for (int t = tid; t < DC_X * DC_Z * ny; t += DC_T) {
        const int vx = t % DC_X, r = t / DC_X, y = r % ny, vz = r / ny;
        const int gx = x0 + vx, gz = z0 + vz;
        if (gx >= g.nx || gz >= g.nz) continue;
        const size_t idx = ((size_t)gz * ny + y) * g.nx + gx;
        acc3[idx] += (double)(long long)acc[(size_t)(vz * DC_X + vx) * ny + y] * inv;
    }
}

}  // namespace utr

struct dht_ctx {
    int method = 3;
    utr_ctx* sctf = nullptr;  // method 5
    // methods 3 and 4
    utr::DhtGeom g{};
    int nrows = 0;  // input images (one per orientation, or a defocus series each)
    int nang = 0;   // orientations
    int batch = 1;
    double dnorm = 0;
    int device = 0;
    int ncu = 256;  // compute units (persistent grids)
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {};
    std::vector<char> added;
    int n_added = 0;
    float add_ms = 0;
    utr_timings tim = {};
    DevBuf<utr::DhtRow> rows;
    DevBuf<double> xs, ys, zs, acc3;
    DevBuf<cplx<float>> spec, T, twx, twy, ap;
    DevBuf<float> P, stage, vol;
    DevBuf<unsigned> pmax;
    int all_z0 = 0;  // no orientation rotates about Z (y-plane splat applies)
    // input image correction (Z'' rotation, RELION shift; DHT.cpp:862-880)
    std::vector<utr::PrepRow> prep;
    DevBuf<utr::PrepRow> prep_d;
    DevBuf<float> img_prep, prep_aver;
    // methods 1 and 2: the defocus series of every orientation
    int series = 0;
    int itermax = 0;
    int ser_batch = 1;  // orientations per retrieval pass
    std::vector<int> first, count;  // images of orientation a: first[a] .. first[a]+count[a]-1
    std::vector<int> iters;         // IWFR iterations of each orientation in the last finalize
    DevBuf<float> simg;             // the renormalised-on-read input images
    DevBuf<float> A1;               // sqrt(I) - 1 of one pass's images
    DevBuf<cplx<float>> B, E;       // retrieval planes; per-image propagator factors
    DevBuf<utr::SerImg> ser_img;
    DevBuf<utr::SerAng> ser_ang;
    DevBuf<double> ser_part;
    DevBuf<int> ser_done, ser_iters, ser_active;
    int waves_ready = 0;  // a finalize has retrieved the waves
    float ret_ms = 0;     // retrieval time of the last finalize
};

namespace {
int dht_series_add(dht_ctx* c, int first, int count, const float* d_img);
int dht_series_retrieve(dht_ctx* c);
}  // namespace

namespace {

// plane-stack budget per batch of orientations (UTR_DHT_BATCH_BYTES overrides)
size_t dht_batch_bytes() {
    const char* e = std::getenv("UTR_DHT_BATCH_BYTES");
    if (e && *e) {
        const long long v = std::atoll(e);
        if (v > 0) return (size_t)v;
    }
    return size_t(6) << 30;
}

double tem_energy(double wl) {
    // UTR.cpp:352-366 (the same relation DHT.cpp uses)
    constexpr double hp = 6.62607004e-34, cc = 299792458, ee = 1.602176634e-19, m0 = 9.1093837015e-31;
    constexpr long double mc2 = m0 * cc * cc;
    long double chl2 = (long double)(cc * cc * hp * hp) / (long double)(wl * wl * 1.e-20);
    long double abra = std::sqrt(mc2 * mc2 + chl2);
    return (double)((-mc2 + abra) / (long double)(ee));
}

bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

// Validation of DHT.cpp:36-720 for the numeric path; "" if valid.
std::string dht_validate(const dht_params* p) {
    if (p->method < 1 || p->method > 5) return "Error: unknown phase retrieval method in the input parameter file.";
    if (p->asigma != 0 && !(p->method == 4 || p->method == 5))
        return "Error: selected phase retrieval method is incompatibe non-zero absorption coefficient in the input parameter file.";
    if (p->method != 5) {
        if (!is_pow2(p->nx) || !is_pow2(p->ny))
            return "Error: Width and Height parameters in input parameter file must be integer powers of 2.";
    } else if (p->nx <= 0 || p->ny <= 0 || p->nx % 2 || p->ny % 2) {
        return "Error: Width and Height parameters in input parameter file must be even.";
    }
    if (!(p->xst > 0)) return "Error: pixel size must be positive in input parameter file.";
    if (!(p->wl > 0) || p->wl > 1) return "Error: wavelength value appears to be wrong.";
    if (p->aobj_mrad < 0 || p->aobj_mrad > 1000) return "Error: objective aperture value appears to be wrong.";
    if (p->escc != 0 && p->escc != 1)
        return "Error: unknown value for Ewald sphere curvature correction switch in input parameter file.";
    if (p->norm_f1 == 0) return "The first normalization factor cannot be zero (leads to division by zero).";
    if (p->method == 1 && p->itermax < 1) return "Error: the maximal number of iterations must be >= 1.";
    if (p->method == 5 && p->critical_snr < 0) return "Error: the critical SNR value must be non-negative.";
    if (p->epsilon < 0 && p->method != 3)
        return "Error: minimal phase error / regulalization parameter must be non-negative, unless -0.5LogAmp phase retrieval method is selected.";
    if ((p->method == 2 || p->method == 4) && p->epsilon == 0)
        return "Error: the estimated regularization parameter (epsilon = 0) is not supported; give epsilon > 0.";
    if (p->inv_laplace && p->alpha < 0)
        return "Error: regularization parameter for 3D Laplacian filter must be non-negative.";
    if (p->dzextra == 0 && p->method <= 3)
        return "Error: extra defocus = 0, which is unexpected for the selected reconstruction (phase retrieval) method.";
    if (p->dzextra != 0 && (p->method == 4 || p->method == 5))
        return "ERROR: extra defocus != 0, which is unexpected for the phase retrieval method selected above.";
    return "";
}

// Per-orientation tables of DHT.cpp:1083-1107, evaluated as the reference
// does (no contraction into fused multiply-adds).
void dht_tables(const utr::DhtGeom& g, double angZ_deg, double angY_deg, double* xs, double* ys,
                double* zs, utr::DhtRow& R) {
#pragma clang fp contract(off)
    const double aZ = angZ_deg * utr::kPi180, aY = angY_deg * utr::kPi180;
    R.sinY = std::sin(aY);
    R.cosY = std::cos(aY);
    R.sinZ = std::sin(aZ);
    R.cosZ = std::cos(aZ);
    for (int i = 0; i < g.nx; i++) {
        const double xxx = g.xlo + g.xst * i - g.xc;
        xs[2 * i] = xxx * R.sinY;
        xs[2 * i + 1] = xxx * R.cosY;
    }
    for (int j = 0; j < g.ny; j++) {
        const double yyy = g.ylo + g.xst * j - g.yc;
        ys[2 * j] = yyy * R.sinZ;
        ys[2 * j + 1] = yyy * R.cosZ;
    }
    for (int n = 0; n < g.nz; n++) {
        const double zzz = (g.zlodz + g.xst * n) - g.zc;
        zs[2 * n] = zzz * R.sinY;
        zs[2 * n + 1] = zzz * R.cosY;
    }
    // affine map (i, j, n) -> (dx1, dy1, dz1) for the brick bounds
    auto pos = [&](double i, double j, double n, double* o) {
        const double X = g.xlo + g.xst * i - g.xc, Y = g.ylo + g.xst * j - g.yc,
                     Zz = (g.zlodz + g.xst * n) - g.zc;
        const double xx = X * R.cosY - Zz * R.sinY;
        o[0] = (g.xc + Y * R.sinZ + xx * R.cosZ - g.xlo) / g.xst;
        o[1] = (g.yc + Y * R.cosZ - xx * R.sinZ - g.ylo) / g.xst;
        o[2] = (g.zc + Zz * R.cosY + X * R.sinY - g.zlodz) / g.xst;
    };
    double c[3], ci[3], cj[3], cn[3];
    pos(0, 0, 0, c);
    pos(1, 0, 0, ci);
    pos(0, 1, 0, cj);
    pos(0, 0, 1, cn);
    double M[9];
    for (int d = 0; d < 3; d++) {
        R.c0[d] = (float)c[d];
        R.ui[d] = (float)(ci[d] - c[d]);
        R.uj[d] = (float)(cj[d] - c[d]);
        R.un[d] = (float)(cn[d] - c[d]);
        M[3 * d] = ci[d] - c[d];
        M[3 * d + 1] = cj[d] - c[d];
        M[3 * d + 2] = cn[d] - c[d];
    }
    // inverse of a rotation (in index units): its transpose, from the exact matrix
    const double det = M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) +
                       M[2] * (M[3] * M[7] - M[4] * M[6]);
    const double id = 1.0 / det;
    const double inv[9] = {(M[4] * M[8] - M[5] * M[7]) * id, (M[2] * M[7] - M[1] * M[8]) * id,
                           (M[1] * M[5] - M[2] * M[4]) * id, (M[5] * M[6] - M[3] * M[8]) * id,
                           (M[0] * M[8] - M[2] * M[6]) * id, (M[2] * M[3] - M[0] * M[5]) * id,
                           (M[3] * M[7] - M[4] * M[6]) * id, (M[1] * M[6] - M[0] * M[7]) * id,
                           (M[0] * M[4] - M[1] * M[3]) * id};
    for (int t = 0; t < 9; t++) R.inv[t] = (float)inv[t];
}

int dht_create_sctf(const dht_params* p, int nrows, const double* angZ, const double* angY,
                    const double* defocus, const float* weights, dht_ctx* c) {
    // the UTR Ewald-sphere path with the DHT filter, no mask and the DHT scale
    utr_params q{};
    q.nx = p->nx; q.ny = p->ny; q.modality = 0; q.xst = p->xst;
    q.zlo = p->zlo; q.zhi = p->zhi; q.dzextra = 0;
    q.dxc = p->dxc; q.dyc = p->dyc; q.dzc = p->dzc; q.wl = p->wl;
    q.aobj_mrad = p->aobj_mrad; q.cs3_mm = p->cs3_mm; q.cs5_mm = p->cs5_mm;
    q.ctf_mode = 3; q.escc = p->escc; q.asigma = 0; q.nsr = 0; q.epsilon = 1;
    q.inv_laplace = 0; q.norm_f1 = p->norm_f1; q.norm_f2 = p->norm_f2;
    int rc = utr_create(&q, nrows, angZ, angY, nullptr, defocus, weights, UTR_FP32, &c->sctf);
    if (rc) return rc;
    utr::GeomDev& g = c->sctf->plan.g;
    g.asigma = p->asigma;
    g.epsilon = p->epsilon;
    g.dht_sctf = 1;
    g.critical_snr = p->critical_snr;
    g.no_mask = 1;
    g.do_invlap = p->inv_laplace;
    g.invlap_eps = p->alpha;
    g.fnorm = p->wl * tem_energy(p->wl) / (2.0 * utr::kPi * std::sqrt(1.0 + p->asigma * p->asigma));
    utr::derive_epilogue(g);
    return UTR_OK;
}

int dht_add_plain(dht_ctx* c, int first_row, int count, const float* d_img) {
    if (c->series) return dht_series_add(c, first_row, count, d_img);
    const auto& g = c->g;
    UTR_HIP_TRY(hipEventRecord(c->ev[0], c->stream));
    bool ok = false;
#define DHT_ROWS(N)                                                                              \
    {                                                                                            \
        constexpr int NB = utr::lds_batch(N, sizeof(cplx<float>), utr::LDS_DHT);                               \
        utr::k_dht_rows<N><<<dim3((g.ny + NB - 1) / NB, count), 256, 0, c->stream>>>(            \
            d_img, c->spec.p + (size_t)first_row * g.ny * g.nx, g.ny, g, c->twx.p);              \
        ok = true;                                                                               \
    }
    UTR_DISPATCH_N(g.nx, DHT_ROWS)
#undef DHT_ROWS
    if (!ok) return set_err(UTR_E_UNSUPPORTED, "unsupported image width for the FFT");
    ok = false;
#define DHT_COLS(N)                                                                              \
    {                                                                                            \
        constexpr int NB = utr::lds_batch(N, sizeof(cplx<float>), utr::LDS_DHT);                               \
        utr::k_dht_cols<N><<<dim3((g.nx + NB - 1) / NB, count), 256, 0, c->stream>>>(            \
            c->spec.p + (size_t)first_row * g.ny * g.nx, g.nx, c->twy.p, -1);                    \
        ok = true;                                                                               \
    }
    UTR_DISPATCH_N(g.ny, DHT_COLS)
#undef DHT_COLS
    if (!ok) return set_err(UTR_E_UNSUPPORTED, "unsupported image height for the FFT");
    UTR_HIP_TRY(hipGetLastError());
    UTR_HIP_TRY(hipEventRecord(c->ev[1], c->stream));
    UTR_HIP_TRY(hipEventSynchronize(c->ev[1]));
    float ms = 0;
    UTR_HIP_TRY(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    c->add_ms += ms;
    for (int r = 0; r < count; r++) {
        if (!c->added[first_row + r]) c->n_added++;
        c->added[first_row + r] = 1;
    }
    return UTR_OK;
}

// The rows' images through the input image correction where a row needs it
// (the kernels of the UTR path with the DHT's renormalisation), in passes of
// kPrepRows images, then the 2D stage.
int dht_add_rows(dht_ctx* c, int first_row, int count, const float* d_img) {
    bool any = false;
    for (int r = 0; r < count && !c->prep.empty(); r++) any |= c->prep[first_row + r].on != 0;
    if (!any) return dht_add_plain(c, first_row, count, d_img);
    const auto& g = c->g;
    utr::GeomDev gp{};
    gp.nx = g.nx;
    gp.ny = g.ny;
    gp.use_f1 = g.use_f1;
    gp.use_f2 = g.use_f2;
    gp.f1 = g.f1;
    gp.f2 = g.f2;
    const size_t n2 = (size_t)g.nx * g.ny;
    // rotation centre (DHT.cpp:820) in pixels
    const double px = (g.xc - g.xlo) / g.xst, py = (g.yc - g.ylo) / g.xst;
    for (int r = 0; r < count; r += kPrepRows) {
        const int k = std::min(kPrepRows, count - r);
        if (c->img_prep.n < n2 * k || c->prep_aver.n < (size_t)k)
            return set_err(UTR_E_STATE, "image correction buffers not allocated");
        utr::k_prep_aver<<<k, 256, 0, c->stream>>>(d_img + r * n2, c->prep_d.p, first_row + r, gp,
                                                   c->prep_aver.p);
        UTR_HIP_TRY(hipGetLastError());
        const dim3 grid((unsigned)((n2 + 255) / 256), (unsigned)k);
        utr::k_prep_img<<<grid, 256, 0, c->stream>>>(d_img + r * n2, c->img_prep.p, c->prep_d.p, first_row + r,
                                                     gp, px, py, c->prep_aver.p);
        UTR_HIP_TRY(hipGetLastError());
        const int rc = dht_add_plain(c, first_row + r, k, c->img_prep.p);
        if (rc) return rc;
    }
    return UTR_OK;
}

int dht_run(dht_ctx* c, float* d_vol) {
    const auto& g = c->g;
    if (c->n_added != c->nrows)
        return set_err(UTR_E_STATE, "dht_finalize: " + std::to_string(c->nrows - c->n_added) +
                                        (c->series ? " input images are missing" : " orientation rows have no image"));
    const size_t N = (size_t)g.nx * g.ny * g.nz;
    if (c->series) {
        const int rc = dht_series_retrieve(c);
        if (rc) return rc;
    }
    UTR_HIP_TRY(hipEventRecord(c->ev[2], c->stream));
    utr::k_dht_fill<<<(unsigned)((N + 255) / 256), 256, 0, c->stream>>>(c->acc3.p, N);
    const long nbricks = (long)((g.nx + utr::DHT_BX - 1) / utr::DHT_BX) *
                         ((g.ny + utr::DHT_BY - 1) / utr::DHT_BY) * ((g.nz + utr::DHT_BZ - 1) / utr::DHT_BZ);
    // columns through y, from column-major plane stacks, when no orientation
    // rotates about Z (UTR_GRID_FORM=brick selects the bricks; both form the
    // same sums)
    const char* form = std::getenv("UTR_GRID_FORM");
    const bool colsplat = c->all_z0 && !(form && std::strcmp(form, "brick") == 0) && g.ny <= utr::DC_NYMAX;
    const int ctx_ = (g.nx + utr::DC_X - 1) / utr::DC_X, ctz = (g.nz + utr::DC_Z - 1) / utr::DC_Z;
    for (int a0 = 0; a0 < c->nang; a0 += c->batch) {
        const int na = std::min(c->batch, c->nang - a0);
        utr::k_dht_zero_u32<<<1, 64, 0, c->stream>>>(c->pmax.p);
        // per orientation: its plane stack's row pass into T, then the column
        // pass into its slot of P; T (one orientation) stays in the
        // memory-side cache between the two passes
        bool ok = false;
        for (int ab = 0; ab < na; ab++) {
            ok = false;
#define DHT_PR(N)                                                                                \
    {                                                                                            \
        constexpr int NB = utr::lds_batch(N, sizeof(cplx<float>), utr::LDS_DHT_ROWS);                 \
        utr::k_dht_prop_rows<N><<<dim3((g.ny + NB - 1) / NB, utr::dht_slots(g), 1), 256, 0,       \
                                  c->stream>>>(                                                  \
            c->spec.p, c->T.p, g, c->rows.p, a0 + ab, c->twx.p, c->ap.p);                        \
        ok = true;                                                                               \
    }
            UTR_DISPATCH_N(g.nx, DHT_PR)
#undef DHT_PR
            if (!ok) return set_err(UTR_E_UNSUPPORTED, "unsupported image width for the FFT");
            ok = false;
#define DHT_PC(N)                                                                                \
    {                                                                                            \
        constexpr int NB = utr::lds_batch(N, sizeof(cplx<float>), utr::LDS_DHT);                 \
        constexpr int lds = 2 * utr::lds_len(NB * N) * (int)sizeof(cplx<float>);                \
        const int gx = (g.nx + NB - 1) / NB, total = gx * utr::dht_slots(g);                     \
        const unsigned cap = (unsigned)(c->ncu * std::max(1, std::min(8, 163840 / lds))) & ~7u;  \
        const unsigned grid = std::min(utr::xcd_grid((unsigned)total), cap ? cap : 8u);          \
        if (colsplat)                                                                            \
            utr::k_dht_prop_cols<N, true><<<grid, 256, 0, c->stream>>>(                          \
                c->T.p, c->P.p + (size_t)ab * g.nplanes * g.nx * g.ny, g, c->rows.p, a0 + ab,    \
                c->twy.p, c->pmax.p, gx, total);                                                 \
        else                                                                                     \
            utr::k_dht_prop_cols<N, false><<<grid, 256, 0, c->stream>>>(                         \
                c->T.p, c->P.p + (size_t)ab * g.nplanes * g.nx * g.ny, g, c->rows.p, a0 + ab,    \
                c->twy.p, c->pmax.p, gx, total);                                                 \
        ok = true;                                                                               \
    }
            UTR_DISPATCH_N(g.ny, DHT_PC)
#undef DHT_PC
            if (!ok) break;
        }
        if (!ok) return set_err(UTR_E_UNSUPPORTED, "unsupported image height for the FFT");
        if (colsplat) {
            const size_t lds = sizeof(unsigned long long) * utr::DC_X * utr::DC_Z * g.ny;
            utr::k_dht_splat_col<<<(unsigned)(ctx_ * ctz), utr::DC_T, lds, c->stream>>>(
                c->P.p, g, c->rows.p, c->xs.p, c->ys.p, c->zs.p, a0, na, c->pmax.p, c->acc3.p, ctx_, ctz);
        } else {
            utr::k_dht_splat<<<(unsigned)nbricks, 256, 0, c->stream>>>(c->P.p, g, c->rows.p, c->xs.p, c->ys.p,
                                                                      c->zs.p, a0, na, c->pmax.p, c->acc3.p);
        }
        UTR_HIP_TRY(hipGetLastError());
    }
    UTR_HIP_TRY(hipEventRecord(c->ev[3], c->stream));
    utr::k_dht_finish<<<(unsigned)((N + 255) / 256), 256, 0, c->stream>>>(c->acc3.p, d_vol, c->dnorm, N);
    UTR_HIP_TRY(hipGetLastError());
    UTR_HIP_TRY(hipEventSynchronize(c->ev[3]));
    UTR_HIP_TRY(hipStreamSynchronize(c->stream));
    float ms = 0;
    UTR_HIP_TRY(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
    c->tim.fft2d_ctf_ms = c->add_ms + c->ret_ms;
    c->tim.backprop_ms = ms;
    c->tim.fft3d_ms = 0;
    c->tim.total_ms = c->add_ms + c->ret_ms + ms;
    return UTR_OK;
}

}  // namespace

namespace {

// Geometry, per-orientation tables and buffers of the real-space methods
// (1 to 4) for nang orientations whose fields are defined at zout[a].  On
// failure the caller destroys c.
int dht_setup(dht_ctx* c, const dht_params* p, int nang, const double* angZ_deg, const double* angY_deg,
              const double* zout, const float* weights) {
    c->nang = nang;
    utr::DhtGeom& g = c->g;
    g.nx = p->nx; g.ny = p->ny; g.method = p->method; g.escc = p->escc;
    g.use_f1 = p->norm_f1 != 1.0; g.use_f2 = p->norm_f2 != 0.0;
    g.f1 = (float)p->norm_f1; g.f2 = (float)p->norm_f2;
    g.xst = p->xst; g.wl = p->wl;
    g.k2maxo = std::pow(p->aobj_mrad * (double)0.001f / p->wl, 2.0);  // DHT.cpp:250
    const double cs3 = p->cs3_mm * 1.e+7, cs5 = p->cs5_mm * 1.e+7;
    g.c3 = p->method != 4 ? -cs3 : cs3;  // iSign = -1 reverts the aberrations (DHT.cpp:962)
    g.c5 = p->method != 4 ? -cs5 : cs5;
    g.asigma = p->asigma; g.epsilon = p->epsilon;
    g.xlo = -0.5 * p->xst * p->nx; g.ylo = -0.5 * p->xst * p->ny;
    const double xhi = g.xlo + p->xst * p->nx, yhi = g.ylo + p->xst * p->ny;
    double zlo = p->zlo, zhi = p->zhi;
    if (zlo > zhi) std::swap(zlo, zhi);
    g.zlodz = zlo + p->dzextra;
    const double zhidz = zhi + p->dzextra;
    const long nz = (long)((zhidz - g.zlodz) / p->xst + 0.5);
    if (nz <= 0) return (set_err(UTR_E_INVALID, "Error: number of output defocus planes must be positive."));
    g.nz = (int)nz;
    g.nplanes = p->escc ? g.nz : 1;
    g.xc = (xhi + g.xlo) / 2.0 + p->dxc;
    g.yc = (yhi + g.ylo) / 2.0 + p->dyc;
    g.zc = (zhidz + g.zlodz) / 2.0 + p->dzc;
    g.nx2 = g.nx - 2; g.ny2 = g.ny - 2; g.nz2 = g.nz - 2;
    if (!fft_size_supported(g.nx) || !fft_size_supported(g.ny))
        return (set_err(UTR_E_UNSUPPORTED, "image size not supported by the FFT kernels"));
    // DHT.cpp:1325-1340
    double total = 0;
    if (weights) for (int r = 0; r < nang; r++) total += weights[r];
    else total = nang;
    const double EE = tem_energy(p->wl), dAtomWidth = 1.0, dGamma = 0.1;
    if (p->method != 4) c->dnorm = EE * dGamma / dAtomWidth * (-p->wl / utr::kPi) / total;
    else c->dnorm = p->wl * EE / utr::kPi / (dAtomWidth * 8.0) / total;
    // per-orientation rows and tables
    std::vector<utr::DhtRow> rows(nang);
    std::vector<double> xs((size_t)nang * 2 * g.nx), ys((size_t)nang * 2 * g.ny), zs((size_t)nang * 2 * g.nz);
    c->all_z0 = 1;
    for (int r = 0; r < nang; r++) {
        if (angZ_deg[r] != 0.0 || g.ny > utr::DC_NYMAX) c->all_z0 = 0;  // column sums in LDS
        std::memset(&rows[r], 0, sizeof rows[r]);
        dht_tables(g, angZ_deg[r], angY_deg[r], xs.data() + (size_t)r * 2 * g.nx,
                   ys.data() + (size_t)r * 2 * g.ny, zs.data() + (size_t)r * 2 * g.nz, rows[r]);
        rows[r].zout = zout[r];
        rows[r].weight = weights ? weights[r] : 1.0f;
    }
    const size_t n2 = (size_t)g.nx * g.ny, N = n2 * g.nz;
    const size_t per_angle = (size_t)g.nplanes * n2 * sizeof(float);  // its plane stack in P
    c->batch = (int)std::max<size_t>(1, std::min<size_t>(std::min<size_t>((size_t)nang, 65535),
                                                         dht_batch_bytes() / per_angle));
    if (g.nplanes > 65535) return (set_err(UTR_E_UNSUPPORTED, "too many output planes"));
    if (hipGetDevice(&c->device) != hipSuccess) return (set_err(UTR_E_DEVICE, "no HIP device"));
    if (hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess ||
        c->ncu <= 0)
        c->ncu = 256;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return (set_err(UTR_E_DEVICE, "hipStreamCreate failed"));
    for (auto& e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) return (set_err(UTR_E_DEVICE, "hipEventCreate failed"));
    if (c->rows.alloc(nang) || c->xs.alloc(xs.size()) || c->ys.alloc(ys.size()) || c->zs.alloc(zs.size()) ||
        c->acc3.alloc(N) || c->spec.alloc((size_t)nang * n2) || c->pmax.alloc(1) ||
        c->T.alloc((size_t)g.nplanes * n2) || c->P.alloc((size_t)c->batch * g.nplanes * n2))
        return (set_err(UTR_E_NOMEM, "device allocation of the DHT buffers failed"));
    if (hipMemcpy(c->rows.p, rows.data(), sizeof(utr::DhtRow) * nang, hipMemcpyHostToDevice) ||
        hipMemcpy(c->xs.p, xs.data(), sizeof(double) * xs.size(), hipMemcpyHostToDevice) ||
        hipMemcpy(c->ys.p, ys.data(), sizeof(double) * ys.size(), hipMemcpyHostToDevice) ||
        hipMemcpy(c->zs.p, zs.data(), sizeof(double) * zs.size(), hipMemcpyHostToDevice))
        return (set_err(UTR_E_DEVICE, "upload of the orientation tables failed"));
    {
        // aperture and Cs phase table of the propagators (DHT.cpp:962-1000):
        // ap = exp(-i (f4 q^4 + f6 q^6)) inside q^2 < k2maxo, 0 outside
        const double pi = 3.14159265358979323846, twopi = 2.0 * pi;
        const double f4 = pi * g.wl * g.wl * g.wl * g.c3 / 2.0, f6 = pi * std::pow(g.wl, 5) * g.c5 / 3.0;
        g.use_ap = (g.k2maxo > 0 || f4 != 0.0 || f6 != 0.0) ? 1 : 0;
        std::vector<cplx<float>> ap(g.use_ap ? n2 : 1, cplx<float>{1.f, 0.f});
        auto freq = [&](int k, int n) { return (double)(k <= n / 2 ? k : k - n) / ((double)n * g.xst); };
        if (g.use_ap)
            for (int ky = 0; ky < g.ny; ky++)
                for (int kx = 0; kx < g.nx; kx++) {
                    const double fx = freq(kx, g.nx), fy = freq(ky, g.ny), q2 = fx * fx + fy * fy;
                    cplx<float>& e = ap[(size_t)ky * g.nx + kx];
                    if (g.k2maxo > 0 && !(q2 < g.k2maxo)) {
                        e = {0.f, 0.f};
                        continue;
                    }
                    double chi = f4 * q2 * q2 + f6 * q2 * q2 * q2;
                    chi -= twopi * std::rint(chi / twopi);
                    e = {(float)std::cos(chi), (float)-std::sin(chi)};
                }
        if (c->ap.alloc(ap.size()) ||
            hipMemcpy(c->ap.p, ap.data(), sizeof(cplx<float>) * ap.size(), hipMemcpyHostToDevice))
            return (set_err(UTR_E_DEVICE, "upload of the aperture table failed"));
    }
    auto twx = make_twiddles<float>(g.nx), twy = make_twiddles<float>(g.ny);
    if (c->twx.alloc(twx.size()) || c->twy.alloc(twy.size()) ||
        hipMemcpy(c->twx.p, twx.data(), sizeof(cplx<float>) * twx.size(), hipMemcpyHostToDevice) ||
        hipMemcpy(c->twy.p, twy.data(), sizeof(cplx<float>) * twy.size(), hipMemcpyHostToDevice))
        return (set_err(UTR_E_DEVICE, "twiddle upload failed"));
    return UTR_OK;
}


}  // namespace

extern "C" {

int dht_create(const dht_params* p, int nrows, const double* angZ_deg, const double* angY_deg,
               const double* defocus, const float* weights, dht_ctx** out) {
    if (!out) return set_err(UTR_E_INVALID, "null output pointer");
    *out = nullptr;
    if (!p) return set_err(UTR_E_INVALID, "null parameters");
    if (nrows < 1 || !angZ_deg || !angY_deg || !defocus)
        return set_err(UTR_E_INVALID, "Error: the orientation file contains no rows.");
    std::string err = dht_validate(p);
    if (!err.empty()) return set_err(UTR_E_INVALID, err);
    if (p->method == 1 || p->method == 2)  // DHT.cpp:272-276
        return set_err(UTR_E_INVALID,
                       "Error: selected phase retrieval method is incompatibe with a single defocus distance "
                       "appearing in some lines of the defocus/orientations parameter file.");
    dht_ctx* c = new dht_ctx();
    c->method = p->method;
    if (p->method == 5) {
        int rc = dht_create_sctf(p, nrows, angZ_deg, angY_deg, defocus, weights, c);
        if (rc) {
            delete c;
            return rc;
        }
        *out = c;
        return UTR_OK;
    }
    c->nrows = nrows;
    c->added.assign(nrows, 0);
    const int rc = dht_setup(c, p, nrows, angZ_deg, angY_deg, defocus, weights);
    if (rc) {
        dht_destroy(c);
        return rc;
    }
    *out = c;
    return UTR_OK;
}

int dht_set_image_correction(dht_ctx* c, const double* angZ2_deg, const double* shift_x,
                             const double* shift_y) {
    if (!c) return set_err(UTR_E_INVALID, "null context");
    if (c->sctf) return set_image_correction(c->sctf, angZ2_deg, shift_x, shift_y);
    if (c->n_added) return set_err(UTR_E_STATE, "dht_set_image_correction after images were added");
    const auto& g = c->g;
    std::vector<utr::PrepRow> prep(c->nrows);
    bool any = false;
    for (int i = 0; i < c->nrows; i++) {
        const double z2 = angZ2_deg ? angZ2_deg[i] : 0.0, sx = shift_x ? shift_x[i] : 0.0,
                     sy = shift_y ? shift_y[i] : 0.0;
        if (!std::isfinite(z2) || !std::isfinite(sx) || !std::isfinite(sy))
            return set_err(UTR_E_INVALID, "image correction values must be finite");
        // DHT.cpp:862-880: Move(floor(-shift_y / yst + 0.5), floor(-shift_x / xst + 0.5)), Rotate(-Z'')
        const double dx = std::floor(-sx / g.xst + 0.5), dy = std::floor(-sy / g.xst + 0.5);
        if (std::fabs(dx) > 1e6 || std::fabs(dy) > 1e6) return set_err(UTR_E_INVALID, "image shift too large");
        const double a = -z2 * (3.14159265358979323846 / 180.0);
        utr::PrepRow& q = prep[i];
        q.cs = std::cos(a);
        q.sn = std::sin(a);
        q.dx = (int)dx;
        q.dy = (int)dy;
        q.on = (z2 != 0 || q.dx != 0 || q.dy != 0) ? 1 : 0;
        any |= q.on != 0;
    }
    UTR_HIP_TRY(hipStreamSynchronize(c->stream));
    UTR_HIP_TRY(c->prep_d.alloc(c->nrows));
    UTR_HIP_TRY(hipMemcpy(c->prep_d.p, prep.data(), sizeof(utr::PrepRow) * c->nrows, hipMemcpyHostToDevice));
    if (any) {
        const int k = std::min(kPrepRows, c->nrows);
        UTR_HIP_TRY(c->img_prep.alloc((size_t)k * g.nx * g.ny));
        UTR_HIP_TRY(c->prep_aver.alloc(k));
    }
    c->prep.swap(prep);
    return UTR_OK;
}

int dht_add_images_device(dht_ctx* c, int first_row, int count, const float* d_images) {
    if (!c) return set_err(UTR_E_INVALID, "null context");
    if (c->sctf) return utr_add_images_device(c->sctf, first_row, count, d_images);
    if (first_row < 0 || count < 0 || first_row + count > c->nrows)
        return set_err(UTR_E_INVALID, "row range outside the orientation table");
    if (count == 0) return UTR_OK;
    if (!d_images) return set_err(UTR_E_INVALID, "null image pointer");
    return dht_add_rows(c, first_row, count, d_images);
}

int dht_add_images(dht_ctx* c, int first_row, int count, const float* images) {
    if (!c) return set_err(UTR_E_INVALID, "null context");
    if (c->sctf) return utr_add_images(c->sctf, first_row, count, images);
    if (first_row < 0 || count < 0 || first_row + count > c->nrows)
        return set_err(UTR_E_INVALID, "row range outside the orientation table");
    if (count == 0) return UTR_OK;
    if (!images) return set_err(UTR_E_INVALID, "null image pointer");
    const size_t n2 = (size_t)c->g.nx * c->g.ny;
    UTR_HIP_TRY(c->stage.alloc(n2 * kStageRows));
    for (int r = 0; r < count; r += kStageRows) {
        const int k = std::min(kStageRows, count - r);
        UTR_HIP_TRY(hipMemcpyAsync(c->stage.p, images + (size_t)r * n2, sizeof(float) * n2 * k,
                                   hipMemcpyHostToDevice, c->stream));
        int rc = dht_add_rows(c, first_row + r, k, c->stage.p);
        if (rc) return rc;
    }
    return UTR_OK;
}

int dht_finalize_device(dht_ctx* c, float* d_vol) {
    if (!c) return set_err(UTR_E_INVALID, "null context");
    if (c->sctf) return utr_finalize_device(c->sctf, d_vol);
    if (!d_vol) return set_err(UTR_E_INVALID, "null output pointer");
    return dht_run(c, d_vol);
}

int dht_finalize(dht_ctx* c, float* vol) {
    if (!c) return set_err(UTR_E_INVALID, "null context");
    if (c->sctf) return utr_finalize(c->sctf, vol);
    if (!vol) return set_err(UTR_E_INVALID, "null output pointer");
    const size_t N = (size_t)c->g.nx * c->g.ny * c->g.nz;
    UTR_HIP_TRY(c->vol.alloc(N));
    int rc = dht_run(c, c->vol.p);
    if (rc) return rc;
    UTR_HIP_TRY(hipMemcpy(vol, c->vol.p, sizeof(float) * N, hipMemcpyDeviceToHost));
    return UTR_OK;
}

int dht_reset(dht_ctx* c) {
    if (!c) return set_err(UTR_E_INVALID, "null context");
    if (c->sctf) return utr_reset(c->sctf);
    std::fill(c->added.begin(), c->added.end(), 0);
    c->n_added = 0;
    c->add_ms = 0;
    return UTR_OK;
}

int dht_dims(const dht_ctx* c, int* nx, int* ny, int* nz) {
    if (!c) return set_err(UTR_E_INVALID, "null context");
    if (c->sctf) return utr_dims(c->sctf, nx, ny, nz);
    if (nx) *nx = c->g.nx;
    if (ny) *ny = c->g.ny;
    if (nz) *nz = c->g.nz;
    return UTR_OK;
}

int dht_get_timings(const dht_ctx* c, utr_timings* t) {
    if (!c || !t) return set_err(UTR_E_INVALID, "null argument");
    if (c->sctf) return utr_get_timings(c->sctf, t);
    *t = c->tim;
    return UTR_OK;
}

int dht_destroy(dht_ctx* c) {
    if (!c) return UTR_OK;
    if (c->sctf) utr_destroy(c->sctf);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->rows.release(); c->xs.release(); c->ys.release(); c->zs.release(); c->acc3.release();
    c->spec.release(); c->T.release(); c->twx.release(); c->twy.release(); c->ap.release();
    c->prep_d.release(); c->img_prep.release(); c->prep_aver.release();
    c->P.release(); c->stage.release(); c->vol.release(); c->pmax.release();
    c->simg.release(); c->B.release(); c->A1.release(); c->E.release(); c->ser_img.release(); c->ser_ang.release();
    c->ser_part.release(); c->ser_done.release(); c->ser_iters.release(); c->ser_active.release();
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return UTR_OK;
}

}  // extern "C"
