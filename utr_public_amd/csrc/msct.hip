// msct.hip -- TEM multislice forward model on gfx950: the coherent branch of
// TemsimLibrary/autoslic.cpp (autoslic::calculate, lpartl = 0; nwobble 0 = one
// static configuration, 1 = Debye-Waller smoothing of each slice's potential
// (trlayerDW), > 1 = averaged thermal configurations), its projected atomic
// potentials (slicelib.cpp vzatom /
// vzatomLUT), the per-orientation atom rotation and the free-space
// propagation of the exit wave to each defocus distance (autosliccmd.cpp),
// behind the msct_* C ABI of include/utr_hip.h.  Included by utr_api.hip
// (one translation unit).
//
// All orientations of a batch advance through the slices together, so
// every launch covers B images.  Per slice:
//   k_ms_potential    projected potential of the slice's atoms, summed per
//                     pixel (atom order, float) and turned into the
//                     transmission function exp(i sigma V)
//                     (autoslic.cpp:917-990)
//   k_post_lines      x FFT (exponent +1, cfpix::fft) of the transmission
//   k_ms_cols<0>      y FFT, bandwidth limit k^2 < k2max, y inverse FFT
//                     (autoslic.cpp:992-1006)
//   k_ms_mul          x inverse FFTs of transmission and wave, product, x FFT
//                     of the product (autoslic.cpp:771)
//   k_ms_cols<1>      y FFT, propagator of the slice with the same limit, y
//                     inverse FFT (slicelib.cpp:811-849)
// Between slices the wave stays in the x-transformed state (normalised so
// that an inverse x FFT returns it), so no plane makes a round trip through
// the real-space form that the next slice does not need.  After the last
// slice, per defocus distance R: k_ms_cols<2> multiplies the spectrum by
// the restated Fresnel factor (autosliccmd.cpp:1072-1087) and k_ms_out
// returns |psi|^2 (autosliccmd.cpp:1196-1200) or psi.

namespace utr {

constexpr int MS_NR = 100;  // radii of the potential look-up table (slicelib.cpp:2818-2820)
// LDS budget (bytes of one buffer) of the column and product kernels.
// 8192 (4 columns of 256 per workgroup) measured faster than 16384 (8
// columns, fewer resident workgroups): 96 against 124 us per column pass
// over 360 images of 256^2.
constexpr int MS_LDS = 8192;

struct alignas(16) MsAtom {
    float x, y, occ;
    int zi;        // row of the element in the spline tables
    int ixo, iyo;  // (int)(x / dx), (int)(y / dy) (autoslic.cpp:950-951)
    float z;
    int pad;
};

struct MsGeom {
    int nx, ny, idx, idy, nsl1;
    double sx, sy, rminsq, lnx0, inv_dln;
};

__device__ __forceinline__ int ms_pmod(int a, int n) {
    const int r = a % n;
    return r < 0 ? r + n : r;
}

// vzatomLUT (slicelib.cpp:2868-2886): interval of seval (below the first
// knot 0, at or above the next-to-last n-2, else x[i] <= r2 < x[i+1]) and the
// cubic in z = r2 - x[i].
__device__ __forceinline__ double ms_lut(const double* __restrict__ knots, const double4* __restrict__ coef,
                                         double r2, const MsGeom& g) {
    int i;
    if (r2 <= knots[0]) {
        i = 0;
    } else if (r2 >= knots[MS_NR - 2]) {
        i = MS_NR - 2;
    } else {
        i = (int)((logf((float)r2) - (float)g.lnx0) * (float)g.inv_dln);
        i = min(max(i, 0), MS_NR - 2);
        while (i < MS_NR - 2 && r2 >= knots[i + 1]) i++;
        while (i > 0 && r2 < knots[i]) i--;
    }
    const double4 c = coef[i];
    const double z = r2 - knots[i];
    return c.x + (c.y + (c.z + c.w * z) * z) * z;
}

// grid (tiles of 16 x 16 pixels, orientation); one thread per pixel.  A
// pixel takes an atom when its periodic image inside the atom's window
// [ixo - idx, ixo + idx] x [iyo - idy, iyo + idy] lies within 3 A; windows
// are narrower than the image (checked on the host), so there is at most
// one such image.  Each chunk of the slice's atoms is first culled against
// the tile and compacted in atom order (wave ballots, no atomics), so every
// pixel still sums its atoms in the order of the z-sorted list.
__global__ void __launch_bounds__(256) k_ms_potential(const MsAtom* __restrict__ atoms,
                                                      const int* __restrict__ off, int natom, int s,
                                                      MsGeom g, const double* __restrict__ knots,
                                                      const double4* __restrict__ coef, double scale,
                                                      cplx<float>* __restrict__ trans, int tiles_x) {
    const int o = blockIdx.y;
    const int a0 = off[o * g.nsl1 + s], a1 = off[o * g.nsl1 + s + 1];
    if (a0 == a1) return;  // empty slice: the wave is not transmitted
    const int tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
    const int X0 = tx * 16, Y0 = ty * 16;
    const int px = X0 + (threadIdx.x & 15), py = Y0 + (threadIdx.x >> 4);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __shared__ MsAtom sa[256];
    __shared__ int wcount[4];
    const MsAtom* A = atoms + (size_t)o * natom;
    float acc = 0.f;
    for (int c0 = a0; c0 < a1; c0 += 256) {
        const int n = min(256, a1 - c0);
        bool hit = false;
        MsAtom at{};
        if ((int)threadIdx.x < n) {
            at = A[c0 + threadIdx.x];
            // offset of the tile's first pixel from the window start, mod n:
            // the 16-pixel tile meets some image of the window when it
            // starts inside one or the next one starts within the tile
            const int dx = ms_pmod(X0 - (at.ixo - g.idx), g.nx);
            const int dy = ms_pmod(Y0 - (at.iyo - g.idy), g.ny);
            hit = (dx <= 2 * g.idx || g.nx - dx <= 15) && (dy <= 2 * g.idy || g.ny - dy <= 15);
        }
        const unsigned long long m = __ballot(hit);
        const int pre = __popcll(m & ((1ull << lane) - 1ull));
        __syncthreads();  // the previous chunk's list is no longer read
        if (lane == 0) wcount[wv] = __popcll(m);
        __syncthreads();
        int base = 0;
        for (int k = 0; k < wv; k++) base += wcount[k];
        const int total = wcount[0] + wcount[1] + wcount[2] + wcount[3];
        if (hit) sa[base + pre] = at;
        __syncthreads();
        for (int k = 0; k < total; k++) {
            const MsAtom t = sa[k];
            const int x0 = t.ixo - g.idx;
            const int ix = x0 + ms_pmod(px - x0, g.nx);
            if (ix > t.ixo + g.idx) continue;
            const int y0 = t.iyo - g.idy;
            const int iy = y0 + ms_pmod(py - y0, g.ny);
            if (iy > t.iyo + g.idy) continue;
            const double rx = (double)t.x - (double)ix * g.sx;
            const double ry = (double)t.y - (double)iy * g.sy;
            double r2 = rx * rx + ry * ry;
            if (r2 > 9.0) continue;
            if (r2 < g.rminsq) r2 = g.rminsq;
            acc += (float)((double)t.occ * ms_lut(knots, coef + (size_t)t.zi * MS_NR, r2, g));
        }
    }
    float sn, cs;
    sincosf((float)(scale * (double)acc), &sn, &cs);
    trans[((size_t)o * g.ny + py) * g.nx + px] = {cs, sn};
}

// The same potential for NB whole rows of one orientation, followed by the
// rows' x FFT (exponent +1) in LDS, so the transmission function reaches
// HBM once, already x-transformed (k_ms_potential + k_post_lines in one
// pass; the sums, the FFT and so the result are the same bit for bit).
// grid (row groups, orientation).
template <int N>
__global__ void __launch_bounds__(256) k_ms_potential_rows(const MsAtom* __restrict__ atoms,
                                                           const int* __restrict__ off, int natom, int s,
                                                           MsGeom g, const double* __restrict__ knots,
                                                           const double4* __restrict__ coef, double scale,
                                                           cplx<float>* __restrict__ trans,
                                                           const cplx<float>* __restrict__ tw, int raw) {
    constexpr int NB = lds_batch(N, sizeof(cplx<float>), LDS_POST);
    constexpr int PER = (NB * N + 255) / 256;
    __shared__ cplx<float> sm[2 * lds_len(NB * N)];
    __shared__ MsAtom sa[256];
    __shared__ int wcount[4];
    const int o = blockIdx.y;
    const int a0 = off[o * g.nsl1 + s], a1 = off[o * g.nsl1 + s + 1];
    if (a0 == a1) return;  // empty slice: the wave is not transmitted
    const int y0 = blockIdx.x * NB;
    const int nb = min(NB, g.ny - y0);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const MsAtom* A = atoms + (size_t)o * natom;
    float acc[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) acc[q] = 0.f;
    for (int c0 = a0; c0 < a1; c0 += 256) {
        const int n = min(256, a1 - c0);
        bool hit = false;
        // the record travels as two 16-byte vectors (a local MsAtom
        // aggregate was given a scratch slot by the compiler)
        float4 r0 = make_float4(0.f, 0.f, 0.f, 0.f), r1 = r0;
        if ((int)threadIdx.x < n) {
            const float4* src = reinterpret_cast<const float4*>(A + c0 + threadIdx.x);
            r0 = src[0];
            r1 = src[1];
            // rows y0 .. y0+nb-1 meet some image of the atom's y window
            const int iyo = __float_as_int(r1.y);
            const int dy = ms_pmod(y0 - (iyo - g.idy), g.ny);
            hit = dy <= 2 * g.idy || g.ny - dy <= nb - 1;
        }
        const unsigned long long m = __ballot(hit);
        const int pre = __popcll(m & ((1ull << lane) - 1ull));
        __syncthreads();
        if (lane == 0) wcount[wv] = __popcll(m);
        __syncthreads();
        int base = 0;
        for (int k = 0; k < wv; k++) base += wcount[k];
        const int total = wcount[0] + wcount[1] + wcount[2] + wcount[3];
        if (hit) {
            float4* dstv = reinterpret_cast<float4*>(sa + base + pre);
            dstv[0] = r0;
            dstv[1] = r1;
        }
        __syncthreads();
        for (int k = 0; k < total; k++) {
            const MsAtom t = sa[k];
#pragma unroll
            for (int q = 0; q < PER; q++) {
                const int p = (int)threadIdx.x + q * 256;
                if (p >= nb * N) break;
                const int r = p / N, px = p - r * N, py = y0 + r;
                const int x0 = t.ixo - g.idx;
                const int ix = x0 + ms_pmod(px - x0, g.nx);
                if (ix > t.ixo + g.idx) continue;
                const int yw = t.iyo - g.idy;
                const int iy = yw + ms_pmod(py - yw, g.ny);
                if (iy > t.iyo + g.idy) continue;
                const double rx = (double)t.x - (double)ix * g.sx;
                const double ry = (double)t.y - (double)iy * g.sy;
                double r2 = rx * rx + ry * ry;
                if (r2 > 9.0) continue;
                if (r2 < g.rminsq) r2 = g.rminsq;
                acc[q] += (float)((double)t.occ * ms_lut(knots, coef + (size_t)t.zi * MS_NR, r2, g));
            }
        }
    }
    cplx<float>* a = sm;
    cplx<float>* b = sm + lds_len(NB * N);
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int p = (int)threadIdx.x + q * 256;
        if (p < nb * N) {
            if (raw) {  // the potential itself, for the Debye-Waller filter
                a[lx(p)] = {acc[q], 0.f};
            } else {
                float sn, cs;
                sincosf((float)(scale * (double)acc[q]), &sn, &cs);
                a[lx(p)] = {cs, sn};
            }
        }
    }
    __syncthreads();
    fft_lds<float, N>(a, b, nb, tw, +1, threadIdx.x, blockDim.x);
    cplx<float>* out = trans + ((size_t)o * g.ny + y0) * N;
    for (int p = threadIdx.x; p < nb * N; p += blockDim.x) out[p] = a[lx(p)];
}

// 3D electrostatic potential of the atoms on a plane at height zp, per
// pixel (xlo + xst ix, ylo + yst iy), summed over every atom in list order
// (the noutput 3 branch of autosliccmd.cpp:1004-1044, 1146-1151): float
// squared distances, vatom() of slicelib.cpp:2699-2737 in FP64, float sum.
// grid (tiles of 256 pixels, orientation * ndef + d).
__global__ void __launch_bounds__(256) k_ms_vatom(const MsAtom* __restrict__ atoms, int natom,
                                                  const double* __restrict__ fe, float xst, float yst,
                                                  const float* __restrict__ zp, int ndef, int nx, int ny,
                                                  float* __restrict__ out, int accum, float div) {
    const int od = blockIdx.y, o = od / ndef, d = od - o * ndef;
    const int p = blockIdx.x * 256 + threadIdx.x;
    const bool on = p < nx * ny;
    const int ix = on ? p % nx : 0, iy = on ? p / nx : 0;
    const float xx = 0.f + xst * (float)ix, yy = 0.f + yst * (float)iy, zz = zp[d];
    __shared__ MsAtom sa[256];
    const MsAtom* A = atoms + (size_t)o * natom;
    const double al = 150.4121417, ag = 266.5985798, pi = 3.141592654;
    float acc = 0.f;
    for (int c0 = 0; c0 < natom; c0 += 256) {
        const int n = min(256, natom - c0);
        __syncthreads();
        if ((int)threadIdx.x < n) sa[threadIdx.x] = A[c0 + threadIdx.x];
        __syncthreads();
        if (!on) continue;
        for (int k = 0; k < n; k++) {
            const MsAtom t = sa[k];
            const float dx = t.x - xx, dy = t.y - yy, dz = t.z - zz;
            const float x2 = dx * dx, y2 = dy * dy, z2 = dz * dz;
            double r = sqrt((double)x2 + (double)y2 + (double)z2);
            if (r < 1.0e-10) r = 1.0e-10;
            const double* f = fe + (size_t)t.zi * 12;
            const double xl = 2.0 * pi * r;
            double suml = 0.0, sumg = 0.0;
            for (int i = 0; i < 6; i += 2) suml += f[i] * exp(-xl * sqrt(f[i + 1]));
            const double xg = (pi * r) * (pi * r);
            for (int i = 6; i < 12; i += 2) {
                double tt = sqrt(f[i + 1]);
                tt = tt * tt * tt;
                sumg += f[i] * exp(-xg / f[i + 1]) / tt;
            }
            acc += (float)(al * suml / r + ag * sumg);
        }
    }
    if (on) {
        float* q = out + ((size_t)od * ny + iy) * nx + ix;
        *q = (accum ? *q + acc : acc) / div;  // thermal average (autosliccmd.cpp:1242-1252)
    }
}

// Column pass over [B][ny][nx] (tile = NB adjacent kx columns of one
// orientation): y FFT (exponent +1), multiply, y inverse FFT (exponent -1),
// store into dst.  MODE 0: bandwidth limit, factor = scale where
// kx2 + ky2 < k2max (float, as autoslic.cpp:1000-1003), else 0.  MODE 1: the
// same limit times the propagator py * px (slicelib.cpp:827-835); the two
// vectors stay in cache where a full (ky, kx) table is read once per image
// (measured 154 against 180 us per pass over 360 images of 256^2).  MODE 2:
// the Fresnel table F[ky][kx] times scale.
template <int N, int MODE>
__global__ void __launch_bounds__(256) k_ms_cols(const cplx<float>* __restrict__ src, cplx<float>* __restrict__ dst,
                                                 int nx, int gx, int gy, const cplx<float>* __restrict__ tw,
                                                 const float* __restrict__ kx2, const float* __restrict__ ky2,
                                                 float k2max, const cplx<float>* __restrict__ px,
                                                 const cplx<float>* __restrict__ py,
                                                 const cplx<float>* __restrict__ F, float scale,
                                                 const int* __restrict__ off, int nsl1, int s,
                                                 const float* __restrict__ dwm) {
    constexpr int NB = lds_batch(N, sizeof(cplx<float>), MS_LDS);
    __shared__ cplx<float> sm[2 * lds_len(NB * N)];
    cplx<float>* a = sm;
    cplx<float>* b = sm + lds_len(NB * N);
    int tx, o;
    if (!xcd_tile(gx, gy, tx, o)) return;
    if ((MODE == 0 || MODE == 3) && off[o * nsl1 + s] == off[o * nsl1 + s + 1]) return;
    const int c0 = tx * NB;
    const int nb = min(NB, nx - c0);
    const size_t base = (size_t)o * N * nx + c0;
    for (int t = threadIdx.x; t < nb * N; t += blockDim.x) {
        const int y = t / nb, c = t - y * nb;
        a[lx(c * N + y)] = src[base + (size_t)y * nx + c];
    }
    __syncthreads();
    fft_lds<float, N>(a, b, nb, tw, +1, threadIdx.x, blockDim.x);
    for (int t = threadIdx.x; t < nb * N; t += blockDim.x) {
        const int c = t / N, y = t - c * N;
        const int kx = c0 + c;
        cplx<float> v = a[lx(t)];
        if (MODE == 3) {
            // exp(-M) of trlayerDW (autoslic.cpp:1080-1088): every frequency
            const float e = (float)exp((double)dwm[o * nsl1 + s] * (double)(ky2[y] + kx2[kx]));
            v = {v.x * e * scale, v.y * e * scale};
        } else if (MODE == 2) {
            v = cmul(v, F[(size_t)y * nx + kx]);
            v = {v.x * scale, v.y * scale};
        } else if (kx2[kx] + ky2[y] < k2max) {
            if (MODE == 1) v = cmul(cmul(v, py[y]), px[kx]);
            v = {v.x * scale, v.y * scale};
        } else {
            v = {0.f, 0.f};
        }
        a[lx(t)] = v;
    }
    __syncthreads();
    fft_lds<float, N>(a, b, nb, tw, -1, threadIdx.x, blockDim.x);
    for (int t = threadIdx.x; t < nb * N; t += blockDim.x) {
        const int y = t / nb, c = t - y * nb;
        dst[base + (size_t)y * nx + c] = a[lx(c * N + y)];
    }
}

// Lines of [B][ny][nx]: NB lines per workgroup.  Multislice (PROJ false):
// the wave W and the transmission T are both x-transformed; inverse x FFT of
// both, W * T where the slice holds atoms for this orientation, x FFT of the
// product back into W.  Projection (PROJ true): W is real-space; inverse x
// FFT of T only and W *= T.
template <int N, bool PROJ>
__global__ void __launch_bounds__(256) k_ms_mul(cplx<float>* __restrict__ W, const cplx<float>* __restrict__ T,
                                                int ny, int nlines, const cplx<float>* __restrict__ tw,
                                                const int* __restrict__ off, int nsl1, int s) {
    constexpr int NB = lds_batch(N, sizeof(cplx<float>), MS_LDS) / 2;
    static_assert(NB >= 1, "two sequences per line must fit");
    __shared__ cplx<float> sm[2 * lds_len(2 * NB * N)];
    cplx<float>* a = sm;
    cplx<float>* b = sm + lds_len(2 * NB * N);
    const int l0 = blockIdx.x * NB;
    const int nb = min(NB, nlines - l0);
    const size_t base = (size_t)l0 * N;
    // lines of one workgroup may straddle two orientations: flag per line
    if (PROJ) {
        for (int t = threadIdx.x; t < nb * N; t += blockDim.x) a[lx(t)] = T[base + t];
        __syncthreads();
        fft_lds<float, N>(a, b, nb, tw, -1, threadIdx.x, blockDim.x);
        for (int t = threadIdx.x; t < nb * N; t += blockDim.x) {
            const int o = (l0 + t / N) / ny;
            if (off[o * nsl1 + s] == off[o * nsl1 + s + 1]) continue;
            W[base + t] = cmul(W[base + t], a[lx(t)]);
        }
        return;
    }
    for (int t = threadIdx.x; t < nb * N; t += blockDim.x) {
        a[lx(t)] = W[base + t];
        a[lx(nb * N + t)] = T[base + t];
    }
    __syncthreads();
    fft_lds<float, N>(a, b, 2 * nb, tw, -1, threadIdx.x, blockDim.x);
    for (int t = threadIdx.x; t < nb * N; t += blockDim.x) {
        const int o = (l0 + t / N) / ny;
        if (off[o * nsl1 + s] != off[o * nsl1 + s + 1]) a[lx(t)] = cmul(a[lx(t)], a[lx(nb * N + t)]);
    }
    __syncthreads();
    fft_lds<float, N>(a, b, nb, tw, +1, threadIdx.x, blockDim.x);
    for (int t = threadIdx.x; t < nb * N; t += blockDim.x) W[base + t] = a[lx(t)];
}

// Debye-Waller branch, after the filter: lines of T (the filtered potential,
// x-transformed and normalised): inverse x FFT, the transmission
// exp(i sigma Re V) (autoslic.cpp:1091-1097), x FFT back into T.
template <int N>
__global__ void __launch_bounds__(256) k_ms_dwexp(cplx<float>* __restrict__ T, int nlines, double scale,
                                                  const cplx<float>* __restrict__ tw) {
    constexpr int NB = lds_batch(N, sizeof(cplx<float>), LDS_POST);
    __shared__ cplx<float> sm[2 * lds_len(NB * N)];
    cplx<float>* a = sm;
    cplx<float>* b = sm + lds_len(NB * N);
    const int l0 = blockIdx.x * NB;
    const int nb = min(NB, nlines - l0);
    cplx<float>* base = T + (size_t)l0 * N;
    for (int t = threadIdx.x; t < nb * N; t += blockDim.x) a[lx(t)] = base[t];
    __syncthreads();
    fft_lds<float, N>(a, b, nb, tw, -1, threadIdx.x, blockDim.x);
    for (int t = threadIdx.x; t < nb * N; t += blockDim.x) {
        float sn, cs;
        sincosf((float)(scale * (double)a[lx(t)].x), &sn, &cs);
        a[lx(t)] = {cs, sn};
    }
    __syncthreads();
    fft_lds<float, N>(a, b, nb, tw, +1, threadIdx.x, blockDim.x);
    for (int t = threadIdx.x; t < nb * N; t += blockDim.x) base[t] = a[lx(t)];
}

// Lines of E [B][ny][nx] (x-transformed, normalised): inverse x FFT, then
// |psi|^2 (what 0), arg psi (what 1, CArg) or psi (what 2) into
// out[(o * ndef + d) * ny + y] (autosliccmd.cpp:1196-1239).
template <int N>
__global__ void __launch_bounds__(256) k_ms_out(const cplx<float>* __restrict__ E, float* __restrict__ out,
                                                int ny, int nlines, int ndef, int d, int what,
                                                const cplx<float>* __restrict__ tw, int accum, float div) {
    constexpr int NB = lds_batch(N, sizeof(cplx<float>), LDS_POST);
    __shared__ cplx<float> sm[2 * lds_len(NB * N)];
    cplx<float>* a = sm;
    cplx<float>* b = sm + lds_len(NB * N);
    const int l0 = blockIdx.x * NB;
    const int nb = min(NB, nlines - l0);
    for (int t = threadIdx.x; t < nb * N; t += blockDim.x) a[lx(t)] = E[(size_t)l0 * N + t];
    __syncthreads();
    fft_lds<float, N>(a, b, nb, tw, -1, threadIdx.x, blockDim.x);
    for (int t = threadIdx.x; t < nb * N; t += blockDim.x) {
        const int l = l0 + t / N, x = t % N;
        const int o = l / ny, y = l - o * ny;
        const size_t row = ((size_t)o * ndef + d) * ny + y;
        const cplx<float> v = a[lx(t)];
        if (what == 2) {
            out[2 * (row * N + x)] = v.x;
            out[2 * (row * N + x) + 1] = v.y;
        } else if (what == 1) {
            out[row * N + x] = atan2f(v.y, v.x);
        } else {
            // thermal configurations: the running sum, divided by their
            // number after the last one (autosliccmd.cpp:1200-1208)
            const float i2 = v.x * v.x + v.y * v.y;
            out[row * N + x] = (accum ? out[row * N + x] + i2 : i2) / div;
        }
    }
}

// wave before the first slice: plane wave 1.  x-transformed and normalised
// that is 1 at kx = 0 and 0 elsewhere; real-space (projection) it is 1.
__global__ void k_ms_init(cplx<float>* __restrict__ W, int nx, size_t n, int proj) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    W[t] = (proj || t % nx == 0) ? cplx<float>{1.f, 0.f} : cplx<float>{0.f, 0.f};
}

}  // namespace utr

namespace {

constexpr double kMsPi = 3.14159265358979323846;

// A&S 9.8.1-9.8.2 (slicelib.cpp:185-212)
double ms_bessi0(double x) {
    const double ax = std::fabs(x);
    if (ax <= 3.75) {
        static const double c[7] = {1.0, 3.5156229, 3.0899424, 1.2067492, 0.2659732, 0.0360768, 0.0045813};
        const double t = (x / 3.75) * (x / 3.75);
        double s = c[6];
        for (int i = 5; i >= 0; i--) s = s * t + c[i];
        return s;
    }
    static const double c[9] = {0.39894228,  0.01328592, 0.00225319,  -0.00157565, 0.00916281,
                                -0.02057706, 0.02635537, -0.01647633, 0.00392377};
    const double t = 3.75 / ax;
    double s = c[8];
    for (int i = 7; i >= 0; i--) s = s * t + c[i];
    return std::exp(ax) * s / std::sqrt(ax);
}

// A&S 9.8.5-9.8.6 (slicelib.cpp:217-245)
double ms_bessk0(double x) {
    const double ax = std::fabs(x);
    if (ax > 0.0 && ax <= 2.0) {
        static const double c[7] = {-0.57721566, 0.42278420, 0.23069756, 0.03488590,
                                    0.00262698,  0.00010750, 0.00000740};
        const double t = (ax / 2.0) * (ax / 2.0);
        double s = c[6];
        for (int i = 5; i >= 0; i--) s = s * t + c[i];
        return -std::log(ax / 2.0) * ms_bessi0(x) + s;
    }
    if (ax > 2.0) {
        static const double c[7] = {1.25331414,  -0.07832358, 0.02189568, -0.01062446,
                                    0.00587872, -0.00251540, 0.00053208};
        const double t = 2.0 / ax;
        double s = c[6];
        for (int i = 5; i >= 0; i--) s = s * t + c[i];
        return std::exp(-ax) * s / std::sqrt(ax);
    }
    return 1.0e20;
}

// projected potential of one element from its 12 fe parameters
// (slicelib.cpp:2759-2792): Kirkland's al = 4 pi^2 a0 e, ag = 2 pi^2 a0 e
// and the 10-digit pi of that routine.
double ms_vzatom(const double* fp, double r) {
    const double al = 300.8242834, ag = 150.4121417, pi = 3.141592654;
    r = std::fabs(r);
    if (r < 1.0e-10) r = 1.0e-10;
    double suml = 0.0, sumg = 0.0;
    double x = 2.0 * pi * r;
    for (int i = 0; i < 6; i += 2) suml += fp[i] * ms_bessk0(x * std::sqrt(fp[i + 1]));
    x = pi * r;
    x = x * x;
    for (int i = 6; i < 12; i += 2) sumg += fp[i] * std::exp(-x / fp[i + 1]) / fp[i + 1];
    return al * suml + ag * sumg;
}

// Akima quasi-Hermite spline (H. Akima, J. ACM 17 (1970) 589; the
// reference's splinh, slicelib.cpp:2537-2603): slopes t_i from the weighted
// neighbouring chords, two extrapolated chords at each end.
void ms_akima(const double* x, const double* y, double* b, double* c, double* d, int n) {
    const double SMALL = 1.0e-25;
    double m5 = (y[3] - y[2]) / (x[3] - x[2]);
    double m4 = (y[2] - y[1]) / (x[2] - x[1]);
    double m3 = (y[1] - y[0]) / (x[1] - x[0]);
    double m2 = m3 + m3 - m4;
    double m1 = m2 + m2 - m3;
    double m54 = std::fabs(m5 - m4), m43 = std::fabs(m4 - m3), m32 = std::fabs(m3 - m2),
           m21 = std::fabs(m2 - m1);
    double t1 = (m43 + m21) > SMALL ? (m43 * m2 + m21 * m3) / (m43 + m21) : 0.5 * (m2 + m3);
    for (int i = 0; i < n - 1; i++) {
        const double t2 = (m54 + m32) > SMALL ? (m54 * m3 + m32 * m4) / (m54 + m32) : 0.5 * (m3 + m4);
        const double h = x[i + 1] - x[i];
        b[i] = t1;
        c[i] = (3.0 * m3 - t1 - t1 - t2) / h;
        d[i] = (t1 + t2 - m3 - m3) / (h * h);
        m1 = m2;
        m2 = m3;
        m3 = m4;
        m4 = m5;
        m5 = i < n - 4 ? (y[i + 4] - y[i + 3]) / (x[i + 4] - x[i + 3]) : m4 + m4 - m3;
        m21 = m32;
        m32 = m43;
        m43 = m54;
        m54 = std::fabs(m5 - m4);
        t1 = t2;
    }
    b[n - 1] = c[n - 1] = d[n - 1] = 0.0;
}

// spatial frequency of FFT bin i (slicelib.cpp:457-476), float
float ms_freq(int i, int n, float a) {
    const int imid = (int)(n / 2.0 + 0.5);
    return i > imid ? (float)(i - n) / a : (float)i / a;
}

}  // namespace

struct msct_ctx {
    msct_params p{};
    int natom = 0;
    std::vector<int> zi;  // element row per atom
    std::vector<float> x, y, z, occ, wob;  // wob: rms thermal displacement at 300 K
    std::vector<double> dwp;               // prefix sums of wob^2 T/300 (file order)
    DevBuf<float> dwm;                     // -B/4 of trlayerDW per (orientation, slice)
    std::vector<float> dwm_h;
    int nsl = 0;          // slices per orientation
    float ctbf = 0.f, k2max = 0.f, wavlen = 0.f, aobj = 0.f;
    utr::MsGeom g{};
    double scale = 0.0;   // sigma(kev) / 1000
    DevBuf<double> knots, fe_el;  // fe_el: the 12 fe parameters per element row
    DevBuf<double4> coef;
    DevBuf<float> kx2, ky2, zp;  // zp: potential planes, ctblength + defocus (float)
    DevBuf<cplx<float>> px, py, twx, twy, F;
    DevBuf<utr::MsAtom> atoms;
    DevBuf<int> off;
    DevBuf<cplx<float>> W, T, E;
    int batch = 0;
};

namespace {

int ms_fail_invalid(const char* m) { return set_err(UTR_E_INVALID, m); }

// the reference's wavelength() and sigma() (slicelib.cpp:2441-2455, 2891-2900)
double ms_wavelength(double kev) { return 12.3984244 / std::sqrt(kev * (2 * 510.99906 + kev)); }
double ms_sigma(double kev) {
    const double x = (510.99906 + kev) / (2.0 * 510.99906 + kev);
    return 2.0 * kMsPi * x / (ms_wavelength(kev) * kev);
}

// number of slices: zslice = k * deltaz (k = 1, 2, ...) while zslice <
// zmax + deltaz (autoslic.cpp:727-744, 794), zmax = the cube side
int ms_slice_count(float ctb, float deltaz) {
    const double dz = deltaz, lim = (double)ctb + dz;
    int n = 0;
    for (double zs = dz; zs < lim; zs += dz) n++;
    return n;
}

// Fresnel factor table for the +1-exponent spectrum (restated FresnelA):
// bin (kx, ky) holds the frequency of bin (-kx, -ky) of the -1-exponent
// convention, exp(-i chi) with chi of the oracle, 0 outside the objective
// aperture.
__global__ void k_ms_fresnel(cplx<float>* __restrict__ F, int nx, int ny, double ctb, double wl, double R,
                             double cs3A, double cs5A, double adz, double ca, double sa, double k2maxo) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nx * ny) return;
    const int kx = t % nx, ky = t / nx;
    const int mx = (nx - kx) % nx, my = (ny - ky) % ny;
    const double qx = (mx < nx / 2 ? mx : mx - nx) / ctb;
    const double qy = (my < ny / 2 ? my : my - ny) / ctb;
    const double q2 = qx * qx + qy * qy;
    const double pi = 3.14159265358979323846;
    double chi = pi * wl * R * q2 + pi * wl * wl * wl * cs3A * q2 * q2 / 2.0 +
                 pi * wl * wl * wl * wl * wl * cs5A * q2 * q2 * q2 / 3.0;
    if (adz != 0.0) chi += pi * wl * adz * ((qx * qx - qy * qy) * ca + 2.0 * qx * qy * sa);
    if (k2maxo > 0.0 && q2 >= k2maxo) {
        F[t] = {0.f, 0.f};
        return;
    }
    double s, c;
    sincos(chi, &s, &c);
    F[t] = {(float)c, (float)-s};
}

int ms_run_cols(int mode, msct_ctx* c, const cplx<float>* src, cplx<float>* dst, int B, float scale, int s,
                const cplx<float>* F, hipStream_t st) {
    const int nx = c->g.nx, ny = c->g.ny;
    bool ok = false;
#define MS_COLS(N)                                                                                  \
    {                                                                                               \
        constexpr int NB = utr::lds_batch(N, sizeof(cplx<float>), utr::MS_LDS);                    \
        const int gx = (nx + NB - 1) / NB;                                                          \
        const unsigned grid = utr::xcd_grid((unsigned)(gx * B));                                    \
        if (mode == 0)                                                                              \
            utr::k_ms_cols<N, 0><<<grid, 256, 0, st>>>(src, dst, nx, gx, B, c->twy.p, c->kx2.p,     \
                                                       c->ky2.p, c->k2max, c->px.p, c->py.p,        \
                                                       F, scale, c->off.p, c->g.nsl1, s, c->dwm.p); \
        else if (mode == 1)                                                                         \
            utr::k_ms_cols<N, 1><<<grid, 256, 0, st>>>(src, dst, nx, gx, B, c->twy.p, c->kx2.p,     \
                                                       c->ky2.p, c->k2max, c->px.p, c->py.p,        \
                                                       F, scale, c->off.p, c->g.nsl1, s, c->dwm.p); \
        else if (mode == 2)                                                                         \
            utr::k_ms_cols<N, 2><<<grid, 256, 0, st>>>(src, dst, nx, gx, B, c->twy.p, c->kx2.p,     \
                                                       c->ky2.p, c->k2max, c->px.p, c->py.p,        \
                                                       F, scale, c->off.p, c->g.nsl1, s, c->dwm.p); \
        else                                                                                        \
            utr::k_ms_cols<N, 3><<<grid, 256, 0, st>>>(src, dst, nx, gx, B, c->twy.p, c->kx2.p,     \
                                                       c->ky2.p, c->k2max, c->px.p, c->py.p,        \
                                                       F, scale, c->off.p, c->g.nsl1, s, c->dwm.p); \
        ok = true;                                                                                  \
    }
    UTR_DISPATCH_N(ny, MS_COLS)
#undef MS_COLS
    if (!ok) return set_err(UTR_E_UNSUPPORTED, "msct: unsupported image height for the FFT");
    UTR_HIP_TRY(hipGetLastError());
    return UTR_OK;
}

int ms_run_lines(msct_ctx* c, int what, int B, int s, int ndef, int d, float* out, hipStream_t st,
                 int accum = 0, float div = 1.f) {
    const int nx = c->g.nx, ny = c->g.ny, nlines = B * ny;
    const bool proj = c->p.mode != 0;
    bool ok = false;
#define MS_LINES(N)                                                                                 \
    {                                                                                               \
        constexpr int NB = utr::lds_batch(N, sizeof(cplx<float>), utr::LDS_POST);                   \
        constexpr int NB2 = utr::lds_batch(N, sizeof(cplx<float>), utr::MS_LDS) / 2;               \
        if (what == 0) {                                                                            \
            utr::k_post_lines<float, N><<<(nlines + NB - 1) / NB, 256, 0, st>>>(c->T.p, nlines, +1, \
                                                                               c->twx.p);           \
        } else if (what == 1) {                                                                     \
            if (proj)                                                                               \
                utr::k_ms_mul<N, true><<<(nlines + NB2 - 1) / NB2, 256, 0, st>>>(                   \
                    c->W.p, c->T.p, ny, nlines, c->twx.p, c->off.p, c->g.nsl1, s);                  \
            else                                                                                    \
                utr::k_ms_mul<N, false><<<(nlines + NB2 - 1) / NB2, 256, 0, st>>>(                  \
                    c->W.p, c->T.p, ny, nlines, c->twx.p, c->off.p, c->g.nsl1, s);                  \
        } else if (what == 4) {                                                                     \
            utr::k_ms_potential_rows<N><<<dim3((ny + NB - 1) / NB, B), 256, 0, st>>>(                  \
                c->atoms.p, c->off.p, c->natom, s, c->g, c->knots.p, c->coef.p, c->scale, c->T.p,    \
                c->twx.p, c->p.nwobble == 1 ? 1 : 0);                                               \
        } else if (what == 5) {                                                                     \
            utr::k_ms_dwexp<N><<<(nlines + NB - 1) / NB, 256, 0, st>>>(c->T.p, nlines, c->scale,       \
                                                                      c->twx.p);                    \
        } else if (what == 2) {                                                                     \
            utr::k_post_lines<float, N><<<(nlines + NB - 1) / NB, 256, 0, st>>>(c->W.p, nlines, +1, \
                                                                               c->twx.p);           \
        } else {                                                                                    \
            utr::k_ms_out<N><<<(nlines + NB - 1) / NB, 256, 0, st>>>(c->E.p, out, ny, nlines, ndef, \
                                                                    d, c->p.output, c->twx.p, accum, div); \
        }                                                                                           \
        ok = true;                                                                                  \
    }
    UTR_DISPATCH_N(nx, MS_LINES)
#undef MS_LINES
    if (!ok) return set_err(UTR_E_UNSUPPORTED, "msct: unsupported image width for the FFT");
    UTR_HIP_TRY(hipGetLastError());
    return UTR_OK;
}

// Rotates, checks, z-sorts and slices the atoms of orientation a into
// dst (natom) and o (nsl + 1): autosliccmd.cpp:709-748 in float, then the
// bounds test of autosliccmd.cpp:857-860 and the depth sort and slice
// membership of autoslic.cpp:720-744.  Returns false when an atom leaves
// the cube.
// Standard normal number for (orientation, configuration, atom, axis):
// counter-based (SplitMix64 of the key, then Box-Muller), so batches and
// host threads do not change the configurations.  The reference draws them
// from Numerical Recipes' gasdev seeded with the clock (autosliccmd.cpp:531),
// so its configurations are not reproducible; these are, for a given seed.
double ms_splitmix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return (double)((x ^ (x >> 31)) >> 11) * (1.0 / 9007199254740992.0);
}
double ms_normal(uint64_t seed, int a, int w, int k, int axis) {
    const uint64_t key = ((uint64_t)a << 40) ^ ((uint64_t)w << 24) ^ ((uint64_t)k << 2) ^ (uint64_t)axis;
    const double u1 = ms_splitmix(seed + 2 * key) + 1.0 / 9007199254740992.0;  // (0, 1]
    const double u2 = ms_splitmix(seed + 2 * key + 1);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kMsPi * u2);
}

bool ms_prepare_one(const msct_ctx* c, int a, int w, double aZ, double aY, utr::MsAtom* dst, int* o, float* dwm,
                    std::vector<float>& X, std::vector<float>& Y, std::vector<float>& Z, std::vector<int>& idx) {
    const int n = c->natom;
    X.resize(n);
    Y.resize(n);
    Z.resize(n);
    idx.resize(n);
    const float ctb = c->ctbf;
    const float xc = (float)(c->p.ctblength / 2.0 + c->p.dxc), yc = (float)(c->p.ctblength / 2.0 + c->p.dyc),
                zc = (float)(c->p.ctblength / 2.0 + c->p.dzc);
    const float sinZ = (float)std::sin(aZ), cosZ = (float)std::cos(aZ), sinY = (float)std::sin(aY),
                cosY = (float)std::cos(aY);
    for (int k = 0; k < n; k++) {
        const float xx = xc + (c->x[k] - xc) * cosZ + (-c->y[k] + yc) * sinZ;
        const float yy = yc + (c->x[k] - xc) * sinZ + (c->y[k] - yc) * cosZ;
        const float x1 = xc + (xx - xc) * cosY + (c->z[k] - zc) * sinY;
        const float z1 = zc + (-xx + xc) * sinY + (c->z[k] - zc) * cosY;
        if (!(x1 >= 0.f && yy >= 0.f && z1 >= 0.f && x1 <= ctb && yy <= ctb && z1 <= ctb)) return false;
        X[k] = x1;
        Y[k] = yy;
        Z[k] = z1;
        idx[k] = k;
    }
    if (c->p.nwobble > 1) {
        // thermal configuration w: rms displacement wobble (at 300 K) scaled
        // by sqrt(T / 300) on the rotated coordinates (autoslic.cpp:700-711)
        const double sc = (double)(float)std::sqrt(c->p.temperature / 300.0);
        for (int k = 0; k < n; k++) {
            const double wb = c->wob[k];
            X[k] = X[k] + (float)(wb * ms_normal(c->p.seed, a, w, k, 0) * sc);
            Y[k] = Y[k] + (float)(wb * ms_normal(c->p.seed, a, w, k, 1) * sc);
            Z[k] = Z[k] + (float)(wb * ms_normal(c->p.seed, a, w, k, 2) * sc);
            // the multislice checks the depth range again after the
            // displacement (autoslic.cpp:733-738); the potential output
            // applies it unchecked (autosliccmd.cpp:1001-1009)
            if (c->p.output != 3 && !(Z[k] >= 0.f && Z[k] <= ctb)) return false;
        }
    }
    // the multislice takes the atoms in depth order (sortByZ, autoslic.cpp:720);
    // the potential output keeps the file order (no multislice runs there)
    if (c->p.output != 3) std::stable_sort(idx.begin(), idx.end(), [&](int i, int j) { return Z[i] < Z[j]; });
    for (int k = 0; k < n; k++) {
        const int i = idx[k];
        utr::MsAtom& m = dst[k];
        m.x = X[i];
        m.y = Y[i];
        m.occ = c->occ[i];
        m.zi = c->zi[i];
        m.ixo = (int)((double)X[i] / c->g.sx);
        m.iyo = (int)((double)Y[i] / c->g.sy);
        m.z = Z[i];
        m.pad = 0;
    }
    const double dz = (double)(float)c->p.deltaz;
    int istart = 0;
    double zslice = dz;
    for (int s = 0; s < c->nsl; s++) {
        o[s] = istart;
        while (istart < n && (double)Z[idx[istart]] < zslice) istart++;
        zslice += dz;
    }
    o[c->nsl] = istart;
    if (dwm) {
        // trlayerDW (autoslic.cpp:1047-1056, 1077): B = 8 pi^2 times the mean of
        // wobble^2 T/300 over the first na entries of the file-order list
        const double tp2 = 2.0 * (2.0 * kMsPi) * (2.0 * kMsPi);
        for (int s = 0; s < c->nsl; s++) {
            const int na = o[s + 1] - o[s];
            dwm[s] = na > 0 ? (float)(-0.25 * (c->dwp[na] / na) * tp2) : 0.f;
        }
    }
    return true;
}

// orientations [a0, a0 + B) into A (B x natom) and off (B x (nsl + 1)), on
// up to 16 host threads
int ms_prepare(const msct_ctx* c, int a0, int B, int w, const double* angZ, const double* angY,
               std::vector<utr::MsAtom>& A, std::vector<int>& off, std::vector<float>& dwm) {
    const int n = c->natom, nsl1 = c->nsl + 1;
    A.resize((size_t)B * n);
    off.assign((size_t)B * nsl1, 0);
    dwm.assign(c->p.nwobble == 1 ? (size_t)B * nsl1 : 0, 0.f);
    const int nt = std::max(1, std::min<int>({B, 16, (int)std::max(1u, std::thread::hardware_concurrency())}));
    std::vector<int> bad(nt, -1);
    auto work = [&](int t) {
        std::vector<float> X, Y, Z;
        std::vector<int> idx;
        for (int b = t; b < B; b += nt)
            if (!ms_prepare_one(c, a0 + b, w, angZ[a0 + b], angY[a0 + b], A.data() + (size_t)b * n,
                                off.data() + (size_t)b * nsl1,
                                c->p.nwobble == 1 ? dwm.data() + (size_t)b * nsl1 : nullptr, X, Y, Z, idx)) {
                bad[t] = b;
                return;
            }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nt; t++) th.emplace_back(work, t);
    work(0);
    for (auto& t : th) t.join();
    int first = -1;
    for (int t = 0; t < nt; t++)
        if (bad[t] >= 0 && (first < 0 || bad[t] < first)) first = bad[t];
    if (first >= 0)
        return set_err(UTR_E_INVALID, "msct: an atom leaves the cube [0, ctblength]^3 at orientation " +
                                          std::to_string(a0 + first) + " (the reference stops with the same error)");
    return UTR_OK;
}

// Orientations per batch within `budget` bytes of device memory.  Each
// orientation of a batch holds three complex planes (wave, transmission,
// propagated exit wave), its rotated atom records and its slice table (start
// offset and Debye-Waller factor per slice).  The host-output staging of
// the images (at most 2 floats per pixel and defocus distance) is allocated
// per call and not counted here.
int ms_batch_size(int nx, int ny, int natom, int nsl, size_t budget) {
    const size_t per = (size_t)3 * nx * ny * sizeof(cplx<float>) +
                       (size_t)std::max(natom, 1) * sizeof(utr::MsAtom) +
                       (size_t)(nsl + 1) * (sizeof(int) + sizeof(float));
    return (int)std::max<size_t>(1, std::min<size_t>(budget / per, 1 << 16));
}

int ms_images(msct_ctx* c, int nang, const double* angZ, const double* angY, int ndef, const double* defocus,
              float* out, bool device_out) {
    if (!c) return ms_fail_invalid("null msct context");
    if (nang < 1 || ndef < 1 || !angZ || !angY || !defocus || !out)
        return ms_fail_invalid("msct_images: need nang >= 1 orientations, ndef >= 1 defocus distances and an output");
    for (int a = 0; a < nang; a++)
        if (!std::isfinite(angZ[a]) || !std::isfinite(angY[a])) return ms_fail_invalid("msct: angles must be finite");
    for (int d = 0; d < ndef; d++)
        if (!std::isfinite(defocus[d])) return ms_fail_invalid("msct: defocus distances must be finite");
    const int nx = c->g.nx, ny = c->g.ny;
    const size_t plane = (size_t)nx * ny;
    const size_t per_img = plane * (c->p.output == 2 ? 2 : 1);
    const int B = std::min(nang, c->batch);
    UTR_HIP_TRY(c->W.alloc(plane * B));
    UTR_HIP_TRY(c->T.alloc(plane * B));
    UTR_HIP_TRY(c->E.alloc(plane * B));
    UTR_HIP_TRY(c->atoms.alloc((size_t)B * std::max(c->natom, 1)));
    UTR_HIP_TRY(c->off.alloc((size_t)B * (c->nsl + 1)));
    UTR_HIP_TRY(c->dwm.alloc((size_t)B * (c->nsl + 1)));
    UTR_HIP_TRY(c->F.alloc(plane * ndef));
    UTR_HIP_TRY(c->zp.alloc(ndef));
    LocalBuf<float> dout;
    if (!device_out) UTR_HIP_TRY(dout.alloc(per_img * ndef * B));
    hipStream_t st = nullptr;
    UTR_HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    auto run = [&]() -> int {
        // Fresnel tables, one per defocus distance: R = defocus - ctblength/2
        // (autosliccmd.cpp:1065), Cs in mm -> A (autosliccmd.cpp:1049-1050)
        const double wl = (double)c->wavlen;
        const double k2maxo = c->aobj > 0.f ? (double)((c->aobj / c->wavlen) * (c->aobj / c->wavlen)) : 0.0;
        const double aa = c->p.astig_angle_deg * kMsPi / 180.0;
        for (int d = 0; d < ndef; d++) {
            k_ms_fresnel<<<(unsigned)((plane + 255) / 256), 256, 0, st>>>(
                c->F.p + plane * d, nx, ny, (double)c->ctbf, wl, defocus[d] - 0.5 * c->p.ctblength,
                (double)(float)c->p.cs3_mm * 1.e7, (double)(float)c->p.cs5_mm * 1.e7, c->p.astig_dz,
                std::cos(2.0 * aa), std::sin(2.0 * aa), k2maxo);
            UTR_HIP_TRY(hipGetLastError());
        }
        std::vector<float> zp(ndef);
        for (int d = 0; d < ndef; d++) zp[d] = c->ctbf + (float)defocus[d];
        UTR_HIP_TRY(hipMemcpyAsync(c->zp.p, zp.data(), sizeof(float) * ndef, hipMemcpyHostToDevice, st));
        std::vector<utr::MsAtom> A;
        std::vector<int> off;
        const bool proj = c->p.mode != 0;
        // UTR_MSCT_FUSED=0: potential in 16 x 16 tiles, then a separate x FFT
        // (same result bit for bit, one more pass over the transmission)
        const char* fenv = std::getenv("UTR_MSCT_FUSED");
        const bool fused = !(fenv && std::atoi(fenv) == 0);
        const int tiles_x = (nx + 15) / 16, tiles = tiles_x * ((ny + 15) / 16);
        const float inv = 1.f / (float)plane;
        const int nw = std::max(1, c->p.nwobble);
        for (int bw = 0; bw < ((nang + B - 1) / B) * nw; bw++) {
            const int a0 = (bw / nw) * B, w = bw % nw;
            const int accum = w > 0 ? 1 : 0;
            const float div = w == nw - 1 ? (float)nw : 1.f;
            const int nb = std::min(B, nang - a0);
            int rc = ms_prepare(c, a0, nb, w, angZ, angY, A, off, c->dwm_h);
            if (rc) return rc;
            UTR_HIP_TRY(hipStreamSynchronize(st));  // the previous batch no longer reads atoms/off
            if (c->natom > 0)
                UTR_HIP_TRY(hipMemcpyAsync(c->atoms.p, A.data(), sizeof(utr::MsAtom) * A.size(),
                                           hipMemcpyHostToDevice, st));
            UTR_HIP_TRY(hipMemcpyAsync(c->off.p, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice, st));
            if (!c->dwm_h.empty())
                UTR_HIP_TRY(hipMemcpyAsync(c->dwm.p, c->dwm_h.data(), sizeof(float) * c->dwm_h.size(),
                                           hipMemcpyHostToDevice, st));
            if (c->p.output == 3) {
                // 3D potential on the plane z = zmax + defocus (autosliccmd.cpp:1034-1038)
                float* dst = device_out ? out + plane * ndef * (size_t)a0 : dout.p;
                if (c->natom > 0) {
                    utr::k_ms_vatom<<<dim3((unsigned)((plane + 255) / 256), nb * ndef), 256, 0, st>>>(
                        c->atoms.p, c->natom, c->fe_el.p, c->ctbf / (float)nx, c->ctbf / (float)ny, c->zp.p, ndef,
                        nx, ny, dst, accum, div);
                    UTR_HIP_TRY(hipGetLastError());
                } else {
                    UTR_HIP_TRY(hipMemsetAsync(dst, 0, sizeof(float) * plane * ndef * nb, st));
                }
                if (!device_out && w == nw - 1)
                    UTR_HIP_TRY(hipMemcpyAsync(out + plane * ndef * (size_t)a0, dout.p, sizeof(float) * plane * ndef * nb,
                                               hipMemcpyDeviceToHost, st));
                continue;
            }
            utr::k_ms_init<<<(unsigned)((plane * nb + 255) / 256), 256, 0, st>>>(c->W.p, nx, plane * nb,
                                                                                 proj ? 1 : 0);
            UTR_HIP_TRY(hipGetLastError());
            for (int s = 0; s < c->nsl; s++) {
                if (c->natom > 0) {
                    if (c->p.nwobble == 1) {
                        // Debye-Waller: raw potential, its filter, then the phase
                        if ((rc = ms_run_lines(c, 4, nb, s, 0, 0, nullptr, st))) return rc;
                        if ((rc = ms_run_cols(3, c, c->T.p, c->T.p, nb, inv, s, nullptr, st))) return rc;
                        if ((rc = ms_run_lines(c, 5, nb, s, 0, 0, nullptr, st))) return rc;
                    } else if (fused) {
                        if ((rc = ms_run_lines(c, 4, nb, s, 0, 0, nullptr, st))) return rc;
                    } else {
                        utr::k_ms_potential<<<dim3(tiles, nb), 256, 0, st>>>(c->atoms.p, c->off.p, c->natom, s,
                                                                              c->g, c->knots.p, c->coef.p, c->scale,
                                                                              c->T.p, tiles_x);
                        UTR_HIP_TRY(hipGetLastError());
                        if ((rc = ms_run_lines(c, 0, nb, s, 0, 0, nullptr, st))) return rc;
                    }
                    if ((rc = ms_run_cols(0, c, c->T.p, c->T.p, nb, inv, s, nullptr, st))) return rc;
                    if ((rc = ms_run_lines(c, 1, nb, s, 0, 0, nullptr, st))) return rc;
                } else if (!proj) {
                    if ((rc = ms_run_lines(c, 1, nb, s, 0, 0, nullptr, st))) return rc;
                }
                if (!proj && (rc = ms_run_cols(1, c, c->W.p, c->W.p, nb, inv, s, nullptr, st))) return rc;
            }
            // exit wave -> spectrum along x: already so (multislice, normalised)
            // or one x FFT of the real-space wave (projection, unnormalised)
            if (proj && (rc = ms_run_lines(c, 2, nb, 0, 0, 0, nullptr, st))) return rc;
            const float fs = proj ? inv : 1.f / (float)ny;
            float* dst = device_out ? out + per_img * ndef * (size_t)a0 : dout.p;
            for (int d = 0; d < ndef; d++) {
                if ((rc = ms_run_cols(2, c, c->W.p, c->E.p, nb, fs, 0, c->F.p + plane * d, st))) return rc;
                if ((rc = ms_run_lines(c, 3, nb, 0, ndef, d, dst, st, accum, div))) return rc;
            }
            if (!device_out && w == nw - 1)
                UTR_HIP_TRY(hipMemcpyAsync(out + per_img * ndef * (size_t)a0, dout.p,
                                           sizeof(float) * per_img * ndef * nb, hipMemcpyDeviceToHost, st));
        }
        UTR_HIP_TRY(hipStreamSynchronize(st));
        return UTR_OK;
    };
    const int rc = run();
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
    return rc;
}

}  // namespace

extern "C" {

int msct_create(const msct_params* p, const double* fe_params, int natom, const int* Z, const float* x,
                const float* y, const float* z, const float* occ, const float* wobble, msct_ctx** out) {
    if (!out) return ms_fail_invalid("null output handle");
    *out = nullptr;
    if (!p || !fe_params) return ms_fail_invalid("msct_create: null parameters or fe table");
    if (!fft_size_supported(p->nx) || !fft_size_supported(p->ny))
        return set_err(UTR_E_UNSUPPORTED, "msct: image size not supported by the FFT kernels (supported: 16 32 48 "
                                          "64 96 128 160 192 256 320 384 512)");
    const double vals[] = {p->ctblength, p->deltaz, p->kev, p->aobj_mrad, p->cs3_mm, p->cs5_mm,
                           p->astig_dz, p->astig_angle_deg, p->dxc, p->dyc, p->dzc};
    for (double v : vals)
        if (!std::isfinite(v)) return ms_fail_invalid("msct: parameters must be finite");
    if (!(p->ctblength > 0) || !(p->deltaz > 0) || !(p->kev > 0))
        return ms_fail_invalid("msct: cube side, slice thickness and energy must be positive");
    if (p->mode < 0 || p->mode > 2) return ms_fail_invalid("msct: mode must be 0 (multislice), 1 or 2 (projection)");
    if (p->output < 0 || p->output > 3)
        return ms_fail_invalid("msct: output must be 0 (intensity), 1 (phase), 2 (complex) or 3 (3D potential)");
    if (natom < 0 || (natom > 0 && (!Z || !x || !y || !z || !occ))) return ms_fail_invalid("msct: bad atom arrays");
    if (p->nwobble < 0 || !std::isfinite(p->temperature) || p->temperature < 0)
        return ms_fail_invalid("msct: nwobble must be >= 0 and the temperature finite and >= 0");
    if (p->nwobble == 1 && p->output == 3)
        return set_err(UTR_E_UNSUPPORTED, "msct: the Debye-Waller factor of the potential output is not built");
    if (p->nwobble > 1 && (p->output == 1 || p->output == 2))
        return ms_fail_invalid("msct: thermal configurations average intensities or potentials (output 0 or 3)");
    for (int k = 0; wobble && k < natom; k++)
        if (!std::isfinite(wobble[k])) return ms_fail_invalid("msct: thermal displacements must be finite");
    for (int k = 0; k < natom; k++) {
        if (Z[k] < 1 || Z[k] > 103) return ms_fail_invalid("msct: atomic numbers must be in 1..103");
        if (!std::isfinite(x[k]) || !std::isfinite(y[k]) || !std::isfinite(z[k]) || !std::isfinite(occ[k]))
            return ms_fail_invalid("msct: atom coordinates and occupancies must be finite");
    }
    for (int i = 0; i < 103 * 12; i++)
        if (!std::isfinite(fe_params[i])) return ms_fail_invalid("msct: fe table must be finite");
    msct_ctx* c = new msct_ctx;
    c->p = *p;
    auto fail = [&](int rc) {
        msct_destroy(c);
        return rc;
    };
    const int nx = p->nx, ny = p->ny;
    const float ax = (float)p->ctblength;  // param[pAX] (autosliccmd.cpp:894-895)
    c->ctbf = ax;
    const float kevf = (float)p->kev;
    c->wavlen = (float)ms_wavelength(kevf);
    c->scale = ms_sigma(kevf) / 1000.0;
    c->aobj = std::fabs((float)p->aobj_mrad * 0.001f);  // autosliccmd.cpp:312
    // trlayer geometry (autoslic.cpp:930-941): float steps, double radii
    utr::MsGeom& g = c->g;
    g.nx = nx;
    g.ny = ny;
    g.sx = (double)(ax / (float)nx);
    g.sy = (double)(ax / (float)ny);
    double rmin = ax / (double)nx;
    const double r = ax / (double)ny;
    rmin = 0.25 * std::sqrt(0.5 * (rmin * rmin + r * r));
    g.rminsq = rmin * rmin;
    g.idx = (int)(nx * 3.0 / ax) + 1;
    g.idy = (int)(ny * 3.0 / ax) + 1;
    if (2 * g.idx + 1 > nx || 2 * g.idy + 1 > ny)
        return fail(ms_fail_invalid("msct: the 3 A atom window is wider than the image (cube side too small)"));
    c->nsl = ms_slice_count(ax, (float)p->deltaz);
    g.nsl1 = c->nsl + 1;
    // bandwidth limit (autoslic.cpp:382-393) and propagator (autoslic.cpp:395-415), float
    float k2max = (float)nx / (2.0f * ax);
    const float tcty = (float)ny / (2.0f * ax);
    if (tcty < k2max) k2max = tcty;
    k2max = (2.0f / 3.0f) * k2max;
    c->k2max = k2max * k2max;
    const float pscale = 3.14159265f * (float)p->deltaz;
    std::vector<float> kx2(nx), ky2(ny);
    std::vector<cplx<float>> px(nx), py(ny);
    for (int i = 0; i < nx; i++) {
        const float k = ms_freq(i, nx, ax);
        kx2[i] = k * k;
        const float t = pscale * (kx2[i] * c->wavlen);
        px[i] = {(float)std::cos((double)t), (float)-std::sin((double)t)};
    }
    for (int i = 0; i < ny; i++) {
        const float k = ms_freq(i, ny, ax);
        ky2[i] = k * k;
        const float t = pscale * (ky2[i] * c->wavlen);
        py[i] = {(float)std::cos((double)t), (float)-std::sin((double)t)};
    }
    // element tables: vzatomLUT knots and Akima coefficients (slicelib.cpp:2831-2866)
    int zrow[104];
    for (int i = 0; i < 104; i++) zrow[i] = -1;
    std::vector<int> elems;
    c->zi.resize(natom);
    for (int k = 0; k < natom; k++) {
        if (zrow[Z[k]] < 0) {
            zrow[Z[k]] = (int)elems.size();
            elems.push_back(Z[k]);
        }
        c->zi[k] = zrow[Z[k]];
    }
    const double RMIN = 0.01, RMAX = 5.0;
    const double dlnr = std::log(RMAX / RMIN) / (utr::MS_NR - 1);
    std::vector<double> knots(utr::MS_NR);
    for (int i = 0; i < utr::MS_NR; i++) {
        const double rr = RMIN * std::exp(i * dlnr);
        knots[i] = rr * rr;
    }
    g.lnx0 = std::log(knots[0]);
    g.inv_dln = 1.0 / (2.0 * dlnr);
    std::vector<double4> coef(std::max<size_t>(1, elems.size() * utr::MS_NR));
    std::vector<double> yv(utr::MS_NR), bv(utr::MS_NR), cv(utr::MS_NR), dv(utr::MS_NR);
    for (size_t e = 0; e < elems.size(); e++) {
        const double* fp = fe_params + (size_t)(elems[e] - 1) * 12;
        for (int i = 0; i < utr::MS_NR; i++) yv[i] = ms_vzatom(fp, std::sqrt(knots[i]));
        ms_akima(knots.data(), yv.data(), bv.data(), cv.data(), dv.data(), utr::MS_NR);
        for (int i = 0; i < utr::MS_NR; i++) coef[e * utr::MS_NR + i] = make_double4(yv[i], bv[i], cv[i], dv[i]);
    }
    c->natom = natom;
    c->x.assign(x, x + natom);
    c->y.assign(y, y + natom);
    c->z.assign(z, z + natom);
    c->occ.assign(occ, occ + natom);
    if (wobble) c->wob.assign(wobble, wobble + natom);
    else c->wob.assign(natom, 0.f);
    c->dwp.assign(natom + 1, 0.0);
    for (int k = 0; k < natom; k++)
        c->dwp[k + 1] = c->dwp[k] + (double)(c->wob[k] * c->wob[k]) * (p->temperature / 300.0);
    auto up = [](auto& buf, const auto& v) -> hipError_t {
        hipError_t e = buf.alloc(v.size());
        if (e != hipSuccess) return e;
        return hipMemcpy(buf.p, v.data(), sizeof(v[0]) * v.size(), hipMemcpyHostToDevice);
    };
    hipError_t e = hipSuccess;
    std::vector<double> fe_el(std::max<size_t>(1, elems.size() * 12));
    for (size_t e2 = 0; e2 < elems.size(); e2++)
        for (int i = 0; i < 12; i++) fe_el[e2 * 12 + i] = fe_params[(size_t)(elems[e2] - 1) * 12 + i];
    if ((e = up(c->knots, knots)) != hipSuccess || (e = up(c->coef, coef)) != hipSuccess ||
        (e = up(c->fe_el, fe_el)) != hipSuccess ||
        (e = up(c->kx2, kx2)) != hipSuccess || (e = up(c->ky2, ky2)) != hipSuccess ||
        (e = up(c->px, px)) != hipSuccess || (e = up(c->py, py)) != hipSuccess ||
        (e = up(c->twx, make_twiddles<float>(nx))) != hipSuccess ||
        (e = up(c->twy, make_twiddles<float>(ny))) != hipSuccess)
        return fail(set_err(e == hipErrorOutOfMemory ? UTR_E_NOMEM : UTR_E_DEVICE,
                            std::string("msct_create: upload failed: ") + hipGetErrorString(e)));
    // orientations per batch: UTR_MSCT_BATCH_BYTES caps the device memory of
    // everything held per orientation of a batch
    size_t budget = (size_t)6 << 30;
    if (const char* s = std::getenv("UTR_MSCT_BATCH_BYTES")) budget = std::max<size_t>(1, std::strtoull(s, nullptr, 10));
    c->batch = ms_batch_size(nx, ny, natom, c->nsl, budget);
    *out = c;
    return UTR_OK;
}

int msct_images(msct_ctx* ctx, int nang, const double* angZ_rad, const double* angY_rad, int ndef,
                const double* defocus, float* out) {
    return ms_images(ctx, nang, angZ_rad, angY_rad, ndef, defocus, out, false);
}

int msct_images_device(msct_ctx* ctx, int nang, const double* angZ_rad, const double* angY_rad, int ndef,
                       const double* defocus, float* d_out) {
    return ms_images(ctx, nang, angZ_rad, angY_rad, ndef, defocus, d_out, true);
}

int msct_slices(const msct_ctx* ctx) { return ctx ? ctx->nsl : -1; }

int msct_batch(const msct_ctx* ctx) { return ctx ? ctx->batch : -1; }

int msct_destroy(msct_ctx* c) {
    if (!c) return UTR_OK;
    c->knots.release(); c->fe_el.release(); c->coef.release(); c->kx2.release(); c->ky2.release(); c->zp.release();
    c->px.release(); c->py.release(); c->twx.release(); c->twy.release(); c->F.release();
    c->atoms.release(); c->off.release(); c->dwm.release(); c->W.release(); c->T.release(); c->E.release();
    delete c;
    return UTR_OK;
}

}  // extern "C"
