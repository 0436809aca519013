// corrector.hip -- the learned 3D residual corrector on gfx950: a 3-level
// U-Net (3x3x3 convolutions, 16/32/64 channels, ReLU, average pooling,
// nearest upsampling, skip connections by channel concatenation) applied to
// UTR reconstructions in tiles, trained on the GPU by the kernels below
// (forward, backward, Adam).  Included by utr_api.hip (one translation
// unit).  No counterpart in the reference, which has no learned component;
// the C ABI is utr_corr_* in include/utr_hip.h.
//
// Numbers: activations and activation gradients are BF16 (channels last,
// [n][z][y][x][c]); weights are FP32 masters updated by Adam and repacked to
// BF16 MFMA fragments after every step; all products accumulate in FP32 on
// the matrix cores (v_mfma_f32_16x16x32_bf16 for the convolutions,
// v_mfma_f32_16x16x16_bf16 for the weight gradients); the head (1x1x1
// convolution to one channel plus alpha times the input) and the loss are
// FP32.
//
// Convolution as an implicit GEMM: M = 16 consecutive x positions of one
// output row, N = 16 output channels, K = (tap, input channel).  A work-group
// owns a 16 x 4 x 4 (x, y, z) output block; each 16-channel slice of the
// input block with its one-voxel halo (18 x 6 x 6) is staged in LDS, and two
// taps form one K = 32 step.  The backward pass with respect to the input is
// the same kernel with the taps mirrored and the channel roles swapped
// (weights repacked once per step), its epilogue multiplying by the ReLU
// mask of the layer below.  Weight gradients contract over voxels (K = 16
// voxels of an x row) with the input slice and the output-gradient block
// both staged in LDS; each work-group sums a fixed subset of blocks and a
// second kernel adds the partial sums in a fixed order, so gradients are
// deterministic.
namespace utr {

typedef short cr_s4 __attribute__((ext_vector_type(4)));
typedef float cr_f4 __attribute__((ext_vector_type(4)));
typedef __bf16 cr_b8 __attribute__((ext_vector_type(8)));

// BF16 <-> FP32 (round to nearest even; NaN stays NaN)
__host__ __device__ __forceinline__ float cr_bf2f(unsigned short h) {
    union { unsigned u; float f; } v;
    v.u = (unsigned)h << 16;
    return v.f;
}
__host__ __device__ __forceinline__ unsigned short cr_f2bf(float f) {
    union { unsigned u; float f; } v;
    v.f = f;
    unsigned u = v.u;
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

constexpr int CR_BX = 16, CR_BY = 4, CR_BZ = 4;        // output block of a work-group
constexpr int CR_HX = CR_BX + 2, CR_HY = CR_BY + 2, CR_HZ = CR_BZ + 2;
constexpr int CR_HVOX = CR_HX * CR_HY * CR_HZ;         // 648 staged voxels
constexpr int CR_VS = 24;                              // LDS halves per staged voxel (16 + pad, 48 B)
constexpr int CR_TSTEPS = 14;                          // K = 32 steps per 16-channel slice (27 taps + 1 zero)

struct CrShape {
    int n, d, h, w;         // batch and spatial extent
    int tx, ty, tz;         // blocks along x, y, z
};

// Voxel index of (b, z, y, x).
__device__ __forceinline__ size_t cr_vox(const CrShape& s, int b, int z, int y, int x) {
    return (((size_t)b * s.d + z) * s.h + y) * s.w + x;
}

// 3x3x3 convolution, stride 1, zero padding 1, NB blocks of 16 output
// channels.  in: cin channels at channel offset in_co of rows of in_cs
// halves; wf: BF16 fragments ((slice * 14 + step) * NB + nb) * 64 + lane,
// 8 halves each; bias may be null; relu: max(0, .) after the bias; mask
// (may be null): multiply by (mask > 0) at the output voxel and channel.
template <int NB>
__global__ void __launch_bounds__(256) k_cr_conv(const unsigned short* __restrict__ in, int in_cs, int in_co,
                                                 int cin, const uint4* __restrict__ wf,
                                                 const float* __restrict__ bias, int relu,
                                                 const unsigned short* __restrict__ mask, int mask_cs,
                                                 int mask_co, unsigned short* __restrict__ out, int out_cs,
                                                 int out_co, CrShape s) {
    __shared__ __align__(16) unsigned short xs[CR_HVOX * CR_VS];
    int blk = blockIdx.x;
    const int bx = blk % s.tx; blk /= s.tx;
    const int by = blk % s.ty; blk /= s.ty;
    const int bz = blk % s.tz;
    const int b = blk / s.tz;
    const int x0 = bx * CR_BX, y0 = by * CR_BY, z0 = bz * CR_BZ;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, kq = lane >> 4;
    cr_f4 acc[CR_BY][NB];
#pragma unroll
    for (int r = 0; r < CR_BY; r++)
#pragma unroll
        for (int nb = 0; nb < NB; nb++) acc[r][nb] = cr_f4{0.f, 0.f, 0.f, 0.f};
    const int nslice = cin >> 4;
    for (int c = 0; c < nslice; c++) {
        __syncthreads();
        for (int t = threadIdx.x; t < CR_HVOX * 2; t += 256) {
            const int v = t >> 1, half = t & 1;
            const int lx = v % CR_HX, ly = (v / CR_HX) % CR_HY, lz = v / (CR_HX * CR_HY);
            const int gx = x0 - 1 + lx, gy = y0 - 1 + ly, gz = z0 - 1 + lz;
            uint4 val = {0u, 0u, 0u, 0u};
            if (gx >= 0 && gx < s.w && gy >= 0 && gy < s.h && gz >= 0 && gz < s.d)
                val = *reinterpret_cast<const uint4*>(in + cr_vox(s, b, gz, gy, gx) * in_cs + in_co + c * 16 +
                                                      half * 8);
            *reinterpret_cast<uint4*>(xs + v * CR_VS + half * 8) = val;
        }
        __syncthreads();
        const uint4* wc = wf + (size_t)c * CR_TSTEPS * NB * 64;
#pragma unroll 2
        for (int st = 0; st < CR_TSTEPS; st++) {
            const int tap = 2 * st + (kq >> 1);
            const bool live = tap < 27;
            const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
            cr_b8 bfr[NB];
#pragma unroll
            for (int nb = 0; nb < NB; nb++) {
                const uint4 q = wc[(st * NB + nb) * 64 + lane];
                bfr[nb] = *reinterpret_cast<const cr_b8*>(&q);
            }
#pragma unroll
            for (int r = 0; r < CR_BY; r++) {
                uint4 a = {0u, 0u, 0u, 0u};
                if (live) {
                    const int v = ((wave + 1 + dz) * CR_HY + (r + 1 + dy)) * CR_HX + (m + 1 + dx);
                    a = *reinterpret_cast<const uint4*>(xs + v * CR_VS + (kq & 1) * 8);
                }
                const cr_b8 afr = *reinterpret_cast<const cr_b8*>(&a);
#pragma unroll
                for (int nb = 0; nb < NB; nb++)
                    acc[r][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr, bfr[nb], acc[r][nb], 0, 0, 0);
            }
        }
    }
    // D[row = 4 kq + i][col = m]: voxel x0 + 4 kq + i, channel nb * 16 + m
    const int z = z0 + wave;
    if (z >= s.d) return;
#pragma unroll
    for (int r = 0; r < CR_BY; r++) {
        const int y = y0 + r;
        if (y >= s.h) continue;
#pragma unroll
        for (int nb = 0; nb < NB; nb++) {
            const int ch = nb * 16 + m;
            const float bv = bias ? bias[ch] : 0.f;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int x = x0 + 4 * kq + i;
                if (x >= s.w) continue;
                const size_t vox = cr_vox(s, b, z, y, x);
                float val = acc[r][nb][i] + bv;
                if (relu) val = fmaxf(val, 0.f);
                if (mask) {
                    const unsigned short mv = mask[vox * mask_cs + mask_co + ch];
                    if (!(mv != 0 && (mv & 0x8000u) == 0)) val = 0.f;
                }
                out[vox * out_cs + out_co + ch] = cr_f2bf(val);
            }
        }
    }
}

// Weight gradients of one 16-channel input slice (blockIdx.y) and one
// 16-channel output block (blockIdx.z), summed over the output blocks
// g, g + G, g + 2G, ... (g = blockIdx.x): part[g][tap][ci][co] (FP32).
// Waves take taps w, w + 4, w + 8, ...; K = the 16 voxels of one x row.
__global__ void __launch_bounds__(256) k_cr_wgrad(const unsigned short* __restrict__ in, int in_cs, int in_co,
                                                  int cin, const unsigned short* __restrict__ dy, int dy_cs,
                                                  int dy_co, int cout, float* __restrict__ part, CrShape s,
                                                  int G) {
    __shared__ unsigned short xs[CR_HVOX * 17];
    __shared__ unsigned short ys[CR_BX * CR_BY * CR_BZ * 17];
    const int g = blockIdx.x, c = blockIdx.y, nb = blockIdx.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, kq = lane >> 4;
    cr_f4 acc[7];
#pragma unroll
    for (int t = 0; t < 7; t++) acc[t] = cr_f4{0.f, 0.f, 0.f, 0.f};
    const int nblocks = s.n * s.tz * s.ty * s.tx;
    for (int blk0 = g; blk0 < nblocks; blk0 += G) {
        int blk = blk0;
        const int bx = blk % s.tx; blk /= s.tx;
        const int by = blk % s.ty; blk /= s.ty;
        const int bz = blk % s.tz;
        const int b = blk / s.tz;
        const int x0 = bx * CR_BX, y0 = by * CR_BY, z0 = bz * CR_BZ;
        __syncthreads();
        for (int t = threadIdx.x; t < CR_HVOX * 16; t += 256) {
            const int v = t >> 4, ch = t & 15;
            const int lx = v % CR_HX, ly = (v / CR_HX) % CR_HY, lz = v / (CR_HX * CR_HY);
            const int gx = x0 - 1 + lx, gy = y0 - 1 + ly, gz = z0 - 1 + lz;
            unsigned short val = 0;
            if (gx >= 0 && gx < s.w && gy >= 0 && gy < s.h && gz >= 0 && gz < s.d)
                val = in[cr_vox(s, b, gz, gy, gx) * in_cs + in_co + c * 16 + ch];
            xs[v * 17 + ch] = val;
        }
        for (int t = threadIdx.x; t < CR_BX * CR_BY * CR_BZ * 16; t += 256) {
            const int v = t >> 4, ch = t & 15;
            const int lx = v % CR_BX, ly = (v / CR_BX) % CR_BY, lz = v / (CR_BX * CR_BY);
            const int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;
            unsigned short val = 0;
            if (gx < s.w && gy < s.h && gz < s.d) val = dy[cr_vox(s, b, gz, gy, gx) * dy_cs + dy_co + nb * 16 + ch];
            ys[v * 17 + ch] = val;
        }
        __syncthreads();
#pragma unroll 1
        for (int row = 0; row < CR_BY * CR_BZ; row++) {
            const int ry = row & 3, rz = row >> 2;
            cr_s4 bb;
#pragma unroll
            for (int i = 0; i < 4; i++) bb[i] = (short)ys[((rz * CR_BY + ry) * CR_BX + 4 * kq + i) * 17 + m];
#pragma unroll
            for (int ti = 0; ti < 7; ti++) {
                const int tap = wave + 4 * ti;
                if (tap < 27) {
                    const int dz = tap / 9 - 1, dyy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
                    const int vrow = ((rz + 1 + dz) * CR_HY + (ry + 1 + dyy)) * CR_HX + 1 + dx;
                    cr_s4 a;
#pragma unroll
                    for (int i = 0; i < 4; i++) a[i] = (short)xs[(vrow + 4 * kq + i) * 17 + m];
                    acc[ti] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc[ti], 0, 0, 0);
                }
            }
        }
    }
    // D[ci = 4 kq + i][co = m]
#pragma unroll
    for (int ti = 0; ti < 7; ti++) {
        const int tap = wave + 4 * ti;
        if (tap >= 27) break;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int ci = c * 16 + 4 * kq + i, co = nb * 16 + m;
            part[(((size_t)g * 27 + tap) * cin + ci) * cout + co] = acc[ti][i];
        }
    }
}

// Sum of the G partial gradients (in order g = 0 .. G-1) into grad.
__global__ void k_cr_reduce(const float* __restrict__ part, int G, size_t n, float* __restrict__ grad) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float sum = 0.f;
    for (int g = 0; g < G; g++) sum += part[(size_t)g * n + t];
    grad[t] = sum;
}

// Bias gradient partial sums: part[g][c] = sum over voxels v = g, g + G, ...
// of dy[v][c] (one work-group per g, threads over channels and voxel lanes).
__global__ void __launch_bounds__(256) k_cr_bgrad(const unsigned short* __restrict__ dy, int dy_cs, int dy_co,
                                                  int cout, size_t nvox, float* __restrict__ part, int G) {
    __shared__ float red[256];
    const int g = blockIdx.x;
    const int ch = threadIdx.x % cout, lanes = 256 / cout, li = threadIdx.x / cout;
    float sum = 0.f;
    if (li < lanes)
        for (size_t v = (size_t)g * lanes + li; v < nvox; v += (size_t)G * lanes)
            sum += cr_bf2f(dy[v * dy_cs + dy_co + ch]);
    red[threadIdx.x] = (li < lanes) ? sum : 0.f;
    __syncthreads();
    if (threadIdx.x < cout) {
        float t = 0.f;
        for (int k = 0; k < lanes; k++) t += red[k * cout + threadIdx.x];
        part[(size_t)g * cout + threadIdx.x] = t;
    }
}

// 2 x 2 x 2 average pooling: out (c channels) from in at half resolution.
__global__ void k_cr_pool(const unsigned short* __restrict__ in, int in_cs, int in_co,
                          unsigned short* __restrict__ out, int out_cs, int out_co, int c, CrShape so) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nvox = (size_t)so.n * so.d * so.h * so.w;
    if (t >= nvox * c) return;
    const int ch = (int)(t % c);
    size_t v = t / c;
    const int x = (int)(v % so.w); v /= so.w;
    const int y = (int)(v % so.h); v /= so.h;
    const int z = (int)(v % so.d);
    const int b = (int)(v / so.d);
    const int W = 2 * so.w, H = 2 * so.h, D = 2 * so.d;
    float sum = 0.f;
    for (int k = 0; k < 8; k++) {
        const int xx = 2 * x + (k & 1), yy = 2 * y + ((k >> 1) & 1), zz = 2 * z + (k >> 2);
        sum += cr_bf2f(in[((((size_t)b * D + zz) * H + yy) * W + xx) * in_cs + in_co + ch]);
    }
    out[(t / c) * out_cs + out_co + ch] = cr_f2bf(0.125f * sum);
}

// Backward of the pooling added to the fine gradient, times the ReLU mask
// of the pooled activation: g[v][c] = (g[v][c] + gc[v/2][c] / 8) * (act > 0).
__global__ void k_cr_pool_bwd(unsigned short* __restrict__ gf, int g_cs, int g_co,
                              const unsigned short* __restrict__ gc, int gc_cs, int gc_co,
                              const unsigned short* __restrict__ act, int a_cs, int a_co, int c, CrShape sf) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nvox = (size_t)sf.n * sf.d * sf.h * sf.w;
    if (t >= nvox * c) return;
    const int ch = (int)(t % c);
    size_t v = t / c;
    const size_t vf = v;
    const int x = (int)(v % sf.w); v /= sf.w;
    const int y = (int)(v % sf.h); v /= sf.h;
    const int z = (int)(v % sf.d);
    const int b = (int)(v / sf.d);
    const size_t vc = (((size_t)b * (sf.d / 2) + z / 2) * (sf.h / 2) + y / 2) * (sf.w / 2) + x / 2;
    const unsigned short a = act[vf * a_cs + a_co + ch];
    float val = cr_bf2f(gf[vf * g_cs + g_co + ch]) + 0.125f * cr_bf2f(gc[vc * gc_cs + gc_co + ch]);
    if (!(a != 0 && (a & 0x8000u) == 0)) val = 0.f;
    gf[vf * g_cs + g_co + ch] = cr_f2bf(val);
}

// Nearest 2x upsampling: out (fine, c channels) from in (coarse).
__global__ void k_cr_up(const unsigned short* __restrict__ in, int in_cs, int in_co,
                        unsigned short* __restrict__ out, int out_cs, int out_co, int c, CrShape sf) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nvox = (size_t)sf.n * sf.d * sf.h * sf.w;
    if (t >= nvox * c) return;
    const int ch = (int)(t % c);
    size_t v = t / c;
    const size_t vf = v;
    const int x = (int)(v % sf.w); v /= sf.w;
    const int y = (int)(v % sf.h); v /= sf.h;
    const int z = (int)(v % sf.d);
    const int b = (int)(v / sf.d);
    const size_t vc = (((size_t)b * (sf.d / 2) + z / 2) * (sf.h / 2) + y / 2) * (sf.w / 2) + x / 2;
    out[vf * out_cs + out_co + ch] = in[vc * in_cs + in_co + ch];
}

// Backward of the upsampling: gc[u][c] = sum of the 8 fine gradients.
__global__ void k_cr_up_bwd(const unsigned short* __restrict__ gf, int g_cs, int g_co,
                            unsigned short* __restrict__ gc, int gc_cs, int gc_co, int c, CrShape sc) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nvox = (size_t)sc.n * sc.d * sc.h * sc.w;
    if (t >= nvox * c) return;
    const int ch = (int)(t % c);
    size_t v = t / c;
    const int x = (int)(v % sc.w); v /= sc.w;
    const int y = (int)(v % sc.h); v /= sc.h;
    const int z = (int)(v % sc.d);
    const int b = (int)(v / sc.d);
    const int W = 2 * sc.w, H = 2 * sc.h, D = 2 * sc.d;
    float sum = 0.f;
    for (int k = 0; k < 8; k++) {
        const int xx = 2 * x + (k & 1), yy = 2 * y + ((k >> 1) & 1), zz = 2 * z + (k >> 2);
        sum += cr_bf2f(gf[((((size_t)b * D + zz) * H + yy) * W + xx) * g_cs + g_co + ch]);
    }
    gc[(t / c) * gc_cs + gc_co + ch] = cr_f2bf(sum);
}

// Network input: FP32 crops [n][T][T][T] -> x0 (BF16, 16 channels, the value
// times scale in channel 0) and x0f (FP32, the value times scale).
__global__ void k_cr_input(const float* __restrict__ vol, float scale, size_t nvox,
                           unsigned short* __restrict__ x0, float* __restrict__ x0f) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nvox) return;
    const float v = vol[t] * scale;
    x0f[t] = v;
    uint4 z = {0u, 0u, 0u, 0u};
    z.x = cr_f2bf(v);
    *reinterpret_cast<uint4*>(x0 + t * 16) = z;
    *reinterpret_cast<uint4*>(x0 + t * 16 + 8) = uint4{0u, 0u, 0u, 0u};
}

// Head: out = alpha * x0f + b1 + sum_c w1[c] * act[c]  (FP32); hp = {w1[16], b1, alpha}.
__global__ void k_cr_head(const unsigned short* __restrict__ act, const float* __restrict__ x0f,
                          const float* __restrict__ hp, size_t nvox, float* __restrict__ out) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nvox) return;
    float s = hp[16] + hp[17] * x0f[t];
#pragma unroll
    for (int c = 0; c < 16; c++) s += hp[c] * cr_bf2f(act[t * 16 + c]);
    out[t] = s;
}

// Loss and head backward over the central tile^3 of every T^3 crop
// (halo h): L = sum (out - tgt)^2 / count; g = 2 (out - tgt) / count there and
// 0 in the halo; gact[v][c] = g w1[c] (act > 0); per work-group partial sums
// part[blk][0..19] = {dw1[16], db1, dalpha, loss, -}.
__global__ void __launch_bounds__(256) k_cr_head_bwd(const unsigned short* __restrict__ act,
                                                     const float* __restrict__ x0f, const float* __restrict__ out,
                                                     const float* __restrict__ tgt, const float* __restrict__ hp,
                                                     int T, int halo, size_t nvox, float inv_count,
                                                     unsigned short* __restrict__ gact, float* __restrict__ part) {
    __shared__ float red[256 * 20];
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    float loc[20];
#pragma unroll
    for (int k = 0; k < 20; k++) loc[k] = 0.f;
    if (t < nvox) {
        const int x = (int)(t % T), y = (int)((t / T) % T), z = (int)((t / ((size_t)T * T)) % T);
        const bool in = x >= halo && x < T - halo && y >= halo && y < T - halo && z >= halo && z < T - halo;
        const float d = out[t] - tgt[t];
        const float g = in ? 2.f * d * inv_count : 0.f;
        uint4 w[2];
        unsigned short* wp = reinterpret_cast<unsigned short*>(w);
#pragma unroll
        for (int c = 0; c < 16; c++) {
            const unsigned short a = act[t * 16 + c];
            const float av = cr_bf2f(a);
            loc[c] = g * av;
            wp[c] = cr_f2bf((a != 0 && (a & 0x8000u) == 0) ? g * hp[c] : 0.f);
        }
        *reinterpret_cast<uint4*>(gact + t * 16) = w[0];
        *reinterpret_cast<uint4*>(gact + t * 16 + 8) = w[1];
        loc[16] = g;
        loc[17] = g * x0f[t];
        loc[18] = in ? d * d * inv_count : 0.f;
    }
#pragma unroll
    for (int k = 0; k < 20; k++) red[k * 256 + threadIdx.x] = loc[k];
    __syncthreads();
    if (threadIdx.x < 20) {
        float s = 0.f;
        for (int i = 0; i < 256; i++) s += red[threadIdx.x * 256 + i];
        part[(size_t)blockIdx.x * 20 + threadIdx.x] = s;
    }
}

// Sum of nblk partial rows of width 20 (fixed order) into res[20].
__global__ void k_cr_head_reduce(const float* __restrict__ part, int nblk, float* __restrict__ res) {
    const int k = threadIdx.x;
    if (k >= 20) return;
    double s = 0.0;
    for (int b = 0; b < nblk; b++) s += part[(size_t)b * 20 + k];
    res[k] = (float)s;
}

// Adam (Kingma & Ba 2015) on the flat parameter vector with bias-corrected
// step lr_t = lr sqrt(1 - b2^t) / (1 - b1^t).
__global__ void k_cr_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                          float* __restrict__ v, size_t n, float lr_t, float b1, float b2, float eps) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float gg = g[t];
    const float mm = b1 * m[t] + (1.f - b1) * gg;
    const float vv = b2 * v[t] + (1.f - b2) * gg * gg;
    m[t] = mm;
    v[t] = vv;
    w[t] -= lr_t * mm / (sqrtf(vv) + eps);
}

// BF16 MFMA fragments of one layer's weights W[27][cin][cout] (FP32):
// forward (fwd != 0: K = (tap, ci), N = co) or input gradient (K = (mirrored
// tap, co), N = ci).  Fragment ((slice * 14 + step) * NB + nb) * 64 + lane
// holds, for lane (n = lane % 16, kq = lane / 16), the 8 values of tap
// 2 step + kq / 2, K channels slice * 16 + (kq % 2) * 8 + j, N channel
// nb * 16 + n (zero for the 28th tap).
__global__ void k_cr_pack(const float* __restrict__ W, int cin, int cout, int fwd, uint4* __restrict__ wf) {
    const int kin = fwd ? cin : cout, nout = fwd ? cout : cin;
    const int NB = nout / 16, ns = kin / 16;
    const size_t total = (size_t)ns * CR_TSTEPS * NB * 64;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int lane = (int)(t % 64);
    size_t r = t / 64;
    const int nb = (int)(r % NB); r /= NB;
    const int st = (int)(r % CR_TSTEPS);
    const int c = (int)(r / CR_TSTEPS);
    const int n = lane & 15, kq = lane >> 4;
    const int tap = 2 * st + (kq >> 1);
    uint4 q;
    unsigned short* h = reinterpret_cast<unsigned short*>(&q);
    for (int j = 0; j < 8; j++) {
        const int k = c * 16 + (kq & 1) * 8 + j, o = nb * 16 + n;
        float val = 0.f;
        if (tap < 27) val = fwd ? W[((size_t)tap * cin + k) * cout + o] : W[((size_t)(26 - tap) * cin + o) * cout + k];
        h[j] = cr_f2bf(val);
    }
    wf[t] = q;
}

// Tiles of a volume for inference: crop b of the (tz, ty, tx) tile grid
// covers [t0 - halo, t0 + tile + halo) per axis, zeros outside the volume.
__global__ void k_cr_gather(const float* __restrict__ vol, int nx, int ny, int nz, int tile, int halo,
                            int ntx, int nty, int first, int nb, float* __restrict__ crops) {
    const int T = tile + 2 * halo;
    const size_t per = (size_t)T * T * T;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= per * nb) return;
    const int b = (int)(t / per);
    size_t r = t % per;
    const int x = (int)(r % T), y = (int)((r / T) % T), z = (int)(r / ((size_t)T * T));
    const int id = first + b;
    const int tx = id % ntx, ty = (id / ntx) % nty, tz = id / (ntx * nty);
    const int gx = tx * tile - halo + x, gy = ty * tile - halo + y, gz = tz * tile - halo + z;
    float v = 0.f;
    if (gx >= 0 && gx < nx && gy >= 0 && gy < ny && gz >= 0 && gz < nz) v = vol[((size_t)gz * ny + gy) * nx + gx];
    crops[t] = v;
}

// Central tile^3 of each output crop back into the volume (divided by scale).
__global__ void k_cr_scatter(const float* __restrict__ crops, int nx, int ny, int nz, int tile, int halo,
                             int ntx, int nty, int first, int nb, float inv_scale, float* __restrict__ vol) {
    const int T = tile + 2 * halo;
    const size_t per = (size_t)tile * tile * tile;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= per * nb) return;
    const int b = (int)(t / per);
    size_t r = t % per;
    const int x = (int)(r % tile), y = (int)((r / tile) % tile), z = (int)(r / ((size_t)tile * tile));
    const int id = first + b;
    const int tx = id % ntx, ty = (id / ntx) % nty, tz = id / (ntx * nty);
    const int gx = tx * tile + x, gy = ty * tile + y, gz = tz * tile + z;
    if (gx >= nx || gy >= ny || gz >= nz) return;
    const size_t src = (size_t)b * T * T * T + ((size_t)(z + halo) * T + (y + halo)) * T + (x + halo);
    vol[((size_t)gz * ny + gy) * nx + gx] = crops[src] * inv_scale;
}

}  // namespace utr

namespace {

// Layers of the U-Net (3x3x3 convolutions), in the order of the flat
// parameter vector: W[27][cin][cout] then b[cout] per layer, then the head
// {w1[16], b1, alpha}.
struct CrLayerDef {
    int cin, cout;
};
constexpr CrLayerDef kCrLayers[10] = {
    {16, 16},  // e1a  (input: the volume in channel 0 of 16)
    {16, 16},  // e1b  -> C1[32:48]
    {16, 32},  // e2a  (input: pooled e1b)
    {32, 32},  // e2b  -> C2[64:96]
    {32, 64},  // e3a  (input: pooled e2b)
    {64, 64},  // e3b
    {96, 32},  // d2a  (input: C2 = up(e3b) | e2b)
    {32, 32},  // d2b
    {48, 16},  // d1a  (input: C1 = up(d2b) | e1b)
    {16, 16},  // d1b
};
constexpr int kCrHead = 18;

using bf16_t = unsigned short;

struct CrLayer {
    int cin, cout;
    size_t woff, boff;
    DevBuf<uint4> wf, wb;  // forward and input-gradient fragments
};

}  // namespace

struct utr_corr {
    int tile = 64, halo = 8, T = 80, batch = 1, device = 0;
    size_t np = 0, headoff = 0;
    CrLayer L[10];
    DevBuf<float> w, g, m, v, res;
    int step = 0;
    float scale = 1.f;
    // activations (BF16, channels last) at the three resolutions
    DevBuf<bf16_t> x0, e1a, C1, p1, e2a, C2, p2, e3a, e3b, d2a, d2b, d1a, d1b;
    DevBuf<float> x0f, out;
    // activation gradients
    DevBuf<bf16_t> gd1b, gd1a, gC1, gd2b, gd2a, gC2, ge3b, ge3a, gp2, ge2a, gp1, ge1a;
    DevBuf<float> part, hpart, crops_in, crops_out;
    hipStream_t st = nullptr;
    double last_ms = 0.0;
};

namespace {

constexpr int kCrG = 256;  // partial sums per weight gradient

utr::CrShape cr_shape(int n, int e) {
    utr::CrShape s;
    s.n = n;
    s.d = s.h = s.w = e;
    s.tx = (e + utr::CR_BX - 1) / utr::CR_BX;
    s.ty = (e + utr::CR_BY - 1) / utr::CR_BY;
    s.tz = (e + utr::CR_BZ - 1) / utr::CR_BZ;
    return s;
}

int cr_conv(utr_corr* c, int layer, bool fwd, const bf16_t* in, int in_cs, int in_co, bf16_t* out, int out_cs,
            int out_co, const bf16_t* mask, int mask_cs, int mask_co, utr::CrShape s) {
    const CrLayer& Ly = c->L[layer];
    const int kin = fwd ? Ly.cin : Ly.cout, nout = fwd ? Ly.cout : Ly.cin;
    const unsigned grid = (unsigned)(s.n * s.tz * s.ty * s.tx);
    const uint4* wf = fwd ? Ly.wf.p : Ly.wb.p;
    const float* bias = fwd ? c->w.p + Ly.boff : nullptr;
    const int relu = fwd ? 1 : 0;
#define CR_CONV(NB)                                                                                           \
    utr::k_cr_conv<NB><<<grid, 256, 0, c->st>>>(in, in_cs, in_co, kin, wf, bias, relu, mask, mask_cs, mask_co, \
                                                out, out_cs, out_co, s)
    switch (nout / 16) {
        case 1: CR_CONV(1); break;
        case 2: CR_CONV(2); break;
        case 3: CR_CONV(3); break;
        case 4: CR_CONV(4); break;
        case 6: CR_CONV(6); break;
        default: return set_err(UTR_E_INVALID, "corrector: unsupported channel count");
    }
#undef CR_CONV
    UTR_HIP_TRY(hipGetLastError());
    return UTR_OK;
}

int cr_wgrad(utr_corr* c, int layer, const bf16_t* in, int in_cs, int in_co, const bf16_t* dy, int dy_cs,
             int dy_co, utr::CrShape s) {
    const CrLayer& Ly = c->L[layer];
    utr::k_cr_wgrad<<<dim3(kCrG, Ly.cin / 16, Ly.cout / 16), 256, 0, c->st>>>(in, in_cs, in_co, Ly.cin, dy, dy_cs,
                                                                            dy_co, Ly.cout, c->part.p, s, kCrG);
    UTR_HIP_TRY(hipGetLastError());
    const size_t nw = (size_t)27 * Ly.cin * Ly.cout;
    utr::k_cr_reduce<<<(unsigned)((nw + 255) / 256), 256, 0, c->st>>>(c->part.p, kCrG, nw, c->g.p + Ly.woff);
    UTR_HIP_TRY(hipGetLastError());
    const size_t nvox = (size_t)s.n * s.d * s.h * s.w;
    utr::k_cr_bgrad<<<kCrG, 256, 0, c->st>>>(dy, dy_cs, dy_co, Ly.cout, nvox, c->part.p, kCrG);
    UTR_HIP_TRY(hipGetLastError());
    utr::k_cr_reduce<<<1, 256, 0, c->st>>>(c->part.p, kCrG, (size_t)Ly.cout, c->g.p + Ly.boff);
    UTR_HIP_TRY(hipGetLastError());
    return UTR_OK;
}

unsigned cr_blocks(size_t n) { return (unsigned)((n + 255) / 256); }

// Largest 27 * cin * cout over the layers (the weight-gradient partials of
// one layer; the bias partials need less).
size_t cr_max_wsize() {
    size_t mx = 0;
    for (const auto& d : kCrLayers) mx = std::max(mx, (size_t)27 * d.cin * d.cout);
    return mx;
}

int cr_alloc(utr_corr* c) {
    const size_t n1 = (size_t)c->batch * c->T * c->T * c->T;
    const size_t n2 = n1 / 8, n3 = n1 / 64;
    const hipError_t e[] = {
        c->x0.alloc(n1 * 16), c->e1a.alloc(n1 * 16), c->C1.alloc(n1 * 48), c->p1.alloc(n2 * 16),
        c->e2a.alloc(n2 * 32), c->C2.alloc(n2 * 96), c->p2.alloc(n3 * 32), c->e3a.alloc(n3 * 64),
        c->e3b.alloc(n3 * 64), c->d2a.alloc(n2 * 32), c->d2b.alloc(n2 * 32), c->d1a.alloc(n1 * 16),
        c->d1b.alloc(n1 * 16), c->x0f.alloc(n1), c->out.alloc(n1),
        c->gd1b.alloc(n1 * 16), c->gd1a.alloc(n1 * 16), c->gC1.alloc(n1 * 48), c->gd2b.alloc(n2 * 32),
        c->gd2a.alloc(n2 * 32), c->gC2.alloc(n2 * 96), c->ge3b.alloc(n3 * 64), c->ge3a.alloc(n3 * 64),
        c->gp2.alloc(n3 * 32), c->ge2a.alloc(n2 * 32), c->gp1.alloc(n2 * 16), c->ge1a.alloc(n1 * 16),
        c->part.alloc((size_t)kCrG * cr_max_wsize()), c->hpart.alloc(((n1 + 255) / 256) * 20), c->res.alloc(32),
        c->crops_in.alloc(n1), c->crops_out.alloc(n1)};
    for (hipError_t x : e)
        if (x != hipSuccess)
            return set_err(x == hipErrorOutOfMemory ? UTR_E_NOMEM : UTR_E_DEVICE,
                           std::string("corrector: device allocation failed: ") + hipGetErrorString(x));
    return UTR_OK;
}

int cr_repack(utr_corr* c) {
    for (int l = 0; l < 10; l++) {
        CrLayer& Ly = c->L[l];
        const size_t nf = (size_t)(Ly.cin / 16) * utr::CR_TSTEPS * (Ly.cout / 16) * 64;
        utr::k_cr_pack<<<cr_blocks(nf), 256, 0, c->st>>>(c->w.p + Ly.woff, Ly.cin, Ly.cout, 1, Ly.wf.p);
        UTR_HIP_TRY(hipGetLastError());
        utr::k_cr_pack<<<cr_blocks(nf), 256, 0, c->st>>>(c->w.p + Ly.woff, Ly.cin, Ly.cout, 0, Ly.wb.p);
        UTR_HIP_TRY(hipGetLastError());
    }
    return UTR_OK;
}

// Forward pass of nb crops already in x0 / x0f; result in c->out.
int cr_forward(utr_corr* c, int nb) {
    const int T = c->T;
    const utr::CrShape s1 = cr_shape(nb, T), s2 = cr_shape(nb, T / 2), s3 = cr_shape(nb, T / 4);
    const size_t n1 = (size_t)nb * T * T * T, n2 = n1 / 8, n3 = n1 / 64;
    int rc;
    if ((rc = cr_conv(c, 0, true, c->x0.p, 16, 0, c->e1a.p, 16, 0, nullptr, 0, 0, s1))) return rc;
    if ((rc = cr_conv(c, 1, true, c->e1a.p, 16, 0, c->C1.p, 48, 32, nullptr, 0, 0, s1))) return rc;
    utr::k_cr_pool<<<cr_blocks(n2 * 16), 256, 0, c->st>>>(c->C1.p, 48, 32, c->p1.p, 16, 0, 16, s2);
    if ((rc = cr_conv(c, 2, true, c->p1.p, 16, 0, c->e2a.p, 32, 0, nullptr, 0, 0, s2))) return rc;
    if ((rc = cr_conv(c, 3, true, c->e2a.p, 32, 0, c->C2.p, 96, 64, nullptr, 0, 0, s2))) return rc;
    utr::k_cr_pool<<<cr_blocks(n3 * 32), 256, 0, c->st>>>(c->C2.p, 96, 64, c->p2.p, 32, 0, 32, s3);
    if ((rc = cr_conv(c, 4, true, c->p2.p, 32, 0, c->e3a.p, 64, 0, nullptr, 0, 0, s3))) return rc;
    if ((rc = cr_conv(c, 5, true, c->e3a.p, 64, 0, c->e3b.p, 64, 0, nullptr, 0, 0, s3))) return rc;
    utr::k_cr_up<<<cr_blocks(n2 * 64), 256, 0, c->st>>>(c->e3b.p, 64, 0, c->C2.p, 96, 0, 64, s2);
    if ((rc = cr_conv(c, 6, true, c->C2.p, 96, 0, c->d2a.p, 32, 0, nullptr, 0, 0, s2))) return rc;
    if ((rc = cr_conv(c, 7, true, c->d2a.p, 32, 0, c->d2b.p, 32, 0, nullptr, 0, 0, s2))) return rc;
    utr::k_cr_up<<<cr_blocks(n1 * 32), 256, 0, c->st>>>(c->d2b.p, 32, 0, c->C1.p, 48, 0, 32, s1);
    if ((rc = cr_conv(c, 8, true, c->C1.p, 48, 0, c->d1a.p, 16, 0, nullptr, 0, 0, s1))) return rc;
    if ((rc = cr_conv(c, 9, true, c->d1a.p, 16, 0, c->d1b.p, 16, 0, nullptr, 0, 0, s1))) return rc;
    utr::k_cr_head<<<cr_blocks(n1), 256, 0, c->st>>>(c->d1b.p, c->x0f.p, c->w.p + c->headoff, n1, c->out.p);
    UTR_HIP_TRY(hipGetLastError());
    return UTR_OK;
}

// Loss and gradients of nb crops after cr_forward, targets in tgt (FP32,
// already scaled); gradient vector in c->g, {dw1, db1, dalpha, loss} in res.
int cr_backward(utr_corr* c, int nb, const float* tgt) {
    const int T = c->T;
    const utr::CrShape s1 = cr_shape(nb, T), s2 = cr_shape(nb, T / 2), s3 = cr_shape(nb, T / 4);
    const size_t n1 = (size_t)nb * T * T * T, n2 = n1 / 8, n3 = n1 / 64;
    const double cnt = (double)nb * c->tile * c->tile * c->tile;
    const unsigned hb = cr_blocks(n1);
    int rc;
    utr::k_cr_head_bwd<<<hb, 256, 0, c->st>>>(c->d1b.p, c->x0f.p, c->out.p, tgt, c->w.p + c->headoff, T, c->halo, n1,
                                              (float)(1.0 / cnt), c->gd1b.p, c->hpart.p);
    UTR_HIP_TRY(hipGetLastError());
    utr::k_cr_head_reduce<<<1, 64, 0, c->st>>>(c->hpart.p, (int)hb, c->res.p);
    UTR_HIP_TRY(hipGetLastError());
    UTR_HIP_TRY(hipMemcpyAsync(c->g.p + c->headoff, c->res.p, sizeof(float) * kCrHead, hipMemcpyDeviceToDevice,
                               c->st));
    // d1b, d1a
    if ((rc = cr_wgrad(c, 9, c->d1a.p, 16, 0, c->gd1b.p, 16, 0, s1))) return rc;
    if ((rc = cr_conv(c, 9, false, c->gd1b.p, 16, 0, c->gd1a.p, 16, 0, c->d1a.p, 16, 0, s1))) return rc;
    if ((rc = cr_wgrad(c, 8, c->C1.p, 48, 0, c->gd1a.p, 16, 0, s1))) return rc;
    if ((rc = cr_conv(c, 8, false, c->gd1a.p, 16, 0, c->gC1.p, 48, 0, c->C1.p, 48, 0, s1))) return rc;
    // up(d2b) -> d2b, d2a
    utr::k_cr_up_bwd<<<cr_blocks(n2 * 32), 256, 0, c->st>>>(c->gC1.p, 48, 0, c->gd2b.p, 32, 0, 32, s2);
    if ((rc = cr_wgrad(c, 7, c->d2a.p, 32, 0, c->gd2b.p, 32, 0, s2))) return rc;
    if ((rc = cr_conv(c, 7, false, c->gd2b.p, 32, 0, c->gd2a.p, 32, 0, c->d2a.p, 32, 0, s2))) return rc;
    if ((rc = cr_wgrad(c, 6, c->C2.p, 96, 0, c->gd2a.p, 32, 0, s2))) return rc;
    if ((rc = cr_conv(c, 6, false, c->gd2a.p, 32, 0, c->gC2.p, 96, 0, c->C2.p, 96, 0, s2))) return rc;
    // up(e3b) -> e3b, e3a
    utr::k_cr_up_bwd<<<cr_blocks(n3 * 64), 256, 0, c->st>>>(c->gC2.p, 96, 0, c->ge3b.p, 64, 0, 64, s3);
    if ((rc = cr_wgrad(c, 5, c->e3a.p, 64, 0, c->ge3b.p, 64, 0, s3))) return rc;
    if ((rc = cr_conv(c, 5, false, c->ge3b.p, 64, 0, c->ge3a.p, 64, 0, c->e3a.p, 64, 0, s3))) return rc;
    if ((rc = cr_wgrad(c, 4, c->p2.p, 32, 0, c->ge3a.p, 64, 0, s3))) return rc;
    if ((rc = cr_conv(c, 4, false, c->ge3a.p, 64, 0, c->gp2.p, 32, 0, nullptr, 0, 0, s3))) return rc;
    // pool(e2b) -> e2b (skip and pooled paths), e2a
    utr::k_cr_pool_bwd<<<cr_blocks(n2 * 32), 256, 0, c->st>>>(c->gC2.p, 96, 64, c->gp2.p, 32, 0, c->C2.p, 96, 64, 32,
                                                               s2);
    if ((rc = cr_wgrad(c, 3, c->e2a.p, 32, 0, c->gC2.p, 96, 64, s2))) return rc;
    if ((rc = cr_conv(c, 3, false, c->gC2.p, 96, 64, c->ge2a.p, 32, 0, c->e2a.p, 32, 0, s2))) return rc;
    if ((rc = cr_wgrad(c, 2, c->p1.p, 16, 0, c->ge2a.p, 32, 0, s2))) return rc;
    if ((rc = cr_conv(c, 2, false, c->ge2a.p, 32, 0, c->gp1.p, 16, 0, nullptr, 0, 0, s2))) return rc;
    // pool(e1b) -> e1b, e1a
    utr::k_cr_pool_bwd<<<cr_blocks(n1 * 16), 256, 0, c->st>>>(c->gC1.p, 48, 32, c->gp1.p, 16, 0, c->C1.p, 48, 32, 16,
                                                               s1);
    if ((rc = cr_wgrad(c, 1, c->e1a.p, 16, 0, c->gC1.p, 48, 32, s1))) return rc;
    if ((rc = cr_conv(c, 1, false, c->gC1.p, 48, 32, c->ge1a.p, 16, 0, c->e1a.p, 16, 0, s1))) return rc;
    if ((rc = cr_wgrad(c, 0, c->x0.p, 16, 0, c->ge1a.p, 16, 0, s1))) return rc;
    UTR_HIP_TRY(hipGetLastError());
    return UTR_OK;
}

int cr_load_crops(utr_corr* c, const float* d_in, int nb) {
    const size_t n1 = (size_t)nb * c->T * c->T * c->T;
    utr::k_cr_input<<<cr_blocks(n1), 256, 0, c->st>>>(d_in, c->scale, n1, c->x0.p, c->x0f.p);
    UTR_HIP_TRY(hipGetLastError());
    return UTR_OK;
}

// Target crops scaled like the input, into crops_out (reused as scratch).
__global__ void k_cr_scale_copy(const float* __restrict__ src, float s, size_t n, float* __restrict__ dst) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = src[t] * s;
}

int cr_check(utr_corr* c) {
    if (!c) return set_err(UTR_E_INVALID, "null corrector");
    UTR_HIP_TRY(hipSetDevice(c->device));
    return UTR_OK;
}

}  // namespace

extern "C" {

int utr_corr_create(int tile, int halo, int batch, unsigned long long seed, utr_corr** out) {
    if (!out) return set_err(UTR_E_INVALID, "null output pointer");
    *out = nullptr;
    if (tile < 4 || halo < 0 || batch < 1 || (tile + 2 * halo) % 4 != 0 || tile + 2 * halo > 512)
        return set_err(UTR_E_INVALID,
                       "utr_corr_create: tile >= 4, halo >= 0, batch >= 1 and tile + 2 halo a multiple of 4 up to 512");
    utr_corr* c = new utr_corr;
    auto fail = [&](int rc) {
        utr_corr_destroy(c);
        return rc;
    };
    if (hipGetDevice(&c->device) != hipSuccess) return fail(set_err(UTR_E_DEVICE, "no HIP device"));
    c->tile = tile;
    c->halo = halo;
    c->T = tile + 2 * halo;
    c->batch = batch;
    size_t off = 0;
    for (int l = 0; l < 10; l++) {
        c->L[l].cin = kCrLayers[l].cin;
        c->L[l].cout = kCrLayers[l].cout;
        c->L[l].woff = off;
        off += (size_t)27 * c->L[l].cin * c->L[l].cout;
        c->L[l].boff = off;
        off += c->L[l].cout;
    }
    c->headoff = off;
    c->np = off + kCrHead;
    // He initialisation (seeded): N(0, 2 / fan_in), fan_in = 27 cin (27 for
    // the first layer, whose input has one live channel); biases 0; head
    // w1 ~ N(0, 1e-4), b1 = 0, alpha = 1 (the network starts as the identity).
    std::vector<float> w(c->np, 0.f);
    uint64_t state = seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    auto gauss = [&]() {
        auto u = [&]() {
            state += 0x9E3779B97F4A7C15ull;
            uint64_t z = state;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            return ((double)((z ^ (z >> 31)) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        };
        const double u1 = u(), u2 = u();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * utr::kPi * u2);
    };
    for (int l = 0; l < 10; l++) {
        const double fan = 27.0 * (l == 0 ? 1 : c->L[l].cin);
        const double sd = std::sqrt(2.0 / fan);
        for (size_t k = 0; k < (size_t)27 * c->L[l].cin * c->L[l].cout; k++) w[c->L[l].woff + k] = (float)(sd * gauss());
    }
    for (int k = 0; k < 16; k++) w[c->headoff + k] = (float)(1e-2 * gauss());
    w[c->headoff + 17] = 1.f;
    if (c->w.alloc(c->np) || c->g.alloc(c->np) || c->m.alloc(c->np) || c->v.alloc(c->np))
        return fail(set_err(UTR_E_NOMEM, "corrector: parameter allocation failed"));
    for (int l = 0; l < 10; l++) {
        const size_t nf = (size_t)(c->L[l].cin / 16) * utr::CR_TSTEPS * (c->L[l].cout / 16) * 64;
        const size_t nb = (size_t)(c->L[l].cout / 16) * utr::CR_TSTEPS * (c->L[l].cin / 16) * 64;
        if (c->L[l].wf.alloc(nf) || c->L[l].wb.alloc(nb))
            return fail(set_err(UTR_E_NOMEM, "corrector: fragment allocation failed"));
    }
    if (int rc = cr_alloc(c)) return fail(rc);
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess)
        return fail(set_err(UTR_E_DEVICE, "corrector: stream creation failed"));
    if (hipMemcpy(c->w.p, w.data(), sizeof(float) * c->np, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(c->m.p, 0, sizeof(float) * c->np) != hipSuccess ||
        hipMemset(c->v.p, 0, sizeof(float) * c->np) != hipSuccess ||
        hipMemset(c->g.p, 0, sizeof(float) * c->np) != hipSuccess)
        return fail(set_err(UTR_E_DEVICE, "corrector: parameter upload failed"));
    if (int rc = cr_repack(c)) return fail(rc);
    if (hipStreamSynchronize(c->st) != hipSuccess) return fail(set_err(UTR_E_DEVICE, "corrector: set-up failed"));
    *out = c;
    return UTR_OK;
}

int utr_corr_nparams(const utr_corr* c) { return c ? (int)c->np : -1; }

int utr_corr_tile(const utr_corr* c, int* tile, int* halo, int* batch) {
    if (!c) return set_err(UTR_E_INVALID, "null corrector");
    if (tile) *tile = c->tile;
    if (halo) *halo = c->halo;
    if (batch) *batch = c->batch;
    return UTR_OK;
}

int utr_corr_get_params(utr_corr* c, float* host, float* scale) {
    if (int rc = cr_check(c)) return rc;
    if (host) UTR_HIP_TRY(hipMemcpy(host, c->w.p, sizeof(float) * c->np, hipMemcpyDeviceToHost));
    if (scale) *scale = c->scale;
    return UTR_OK;
}

int utr_corr_set_params(utr_corr* c, const float* host, float scale) {
    if (int rc = cr_check(c)) return rc;
    if (!(scale > 0) || !std::isfinite(scale)) return set_err(UTR_E_INVALID, "utr_corr_set_params: scale must be positive");
    c->scale = scale;
    if (host) {
        UTR_HIP_TRY(hipMemcpy(c->w.p, host, sizeof(float) * c->np, hipMemcpyHostToDevice));
        UTR_HIP_TRY(hipMemsetAsync(c->m.p, 0, sizeof(float) * c->np, c->st));
        UTR_HIP_TRY(hipMemsetAsync(c->v.p, 0, sizeof(float) * c->np, c->st));
        c->step = 0;
        if (int rc = cr_repack(c)) return rc;
    }
    UTR_HIP_TRY(hipStreamSynchronize(c->st));
    return UTR_OK;
}

int utr_corr_forward(utr_corr* c, const float* d_in, int nb, float* d_out) {
    if (int rc = cr_check(c)) return rc;
    if (!d_in || !d_out || nb < 1 || nb > c->batch) return set_err(UTR_E_INVALID, "utr_corr_forward: 1 <= nb <= batch crops");
    if (int rc = cr_load_crops(c, d_in, nb)) return rc;
    if (int rc = cr_forward(c, nb)) return rc;
    const size_t n1 = (size_t)nb * c->T * c->T * c->T;
    k_cr_scale_copy<<<cr_blocks(n1), 256, 0, c->st>>>(c->out.p, 1.f / c->scale, n1, d_out);
    UTR_HIP_TRY(hipGetLastError());
    UTR_HIP_TRY(hipStreamSynchronize(c->st));
    return UTR_OK;
}

int utr_corr_gradient(utr_corr* c, const float* d_in, const float* d_tgt, int nb, float* loss, float* grad_host) {
    if (int rc = cr_check(c)) return rc;
    if (!d_in || !d_tgt || nb < 1 || nb > c->batch)
        return set_err(UTR_E_INVALID, "utr_corr_gradient: 1 <= nb <= batch crops");
    const size_t n1 = (size_t)nb * c->T * c->T * c->T;
    if (int rc = cr_load_crops(c, d_in, nb)) return rc;
    k_cr_scale_copy<<<cr_blocks(n1), 256, 0, c->st>>>(d_tgt, c->scale, n1, c->crops_out.p);
    UTR_HIP_TRY(hipGetLastError());
    if (int rc = cr_forward(c, nb)) return rc;
    if (int rc = cr_backward(c, nb, c->crops_out.p)) return rc;
    float res[20];
    UTR_HIP_TRY(hipMemcpyAsync(res, c->res.p, sizeof(res), hipMemcpyDeviceToHost, c->st));
    UTR_HIP_TRY(hipStreamSynchronize(c->st));
    if (loss) *loss = res[18];
    if (grad_host) UTR_HIP_TRY(hipMemcpy(grad_host, c->g.p, sizeof(float) * c->np, hipMemcpyDeviceToHost));
    return UTR_OK;
}

int utr_corr_train_step(utr_corr* c, const float* d_in, const float* d_tgt, int nb, float lr, float* loss) {
    if (int rc = cr_check(c)) return rc;
    if (!d_in || !d_tgt || nb < 1 || nb > c->batch)
        return set_err(UTR_E_INVALID, "utr_corr_train_step: 1 <= nb <= batch crops");
    if (!(lr > 0) || !std::isfinite(lr)) return set_err(UTR_E_INVALID, "utr_corr_train_step: lr must be positive");
    const size_t n1 = (size_t)nb * c->T * c->T * c->T;
    if (int rc = cr_load_crops(c, d_in, nb)) return rc;
    k_cr_scale_copy<<<cr_blocks(n1), 256, 0, c->st>>>(d_tgt, c->scale, n1, c->crops_out.p);
    UTR_HIP_TRY(hipGetLastError());
    if (int rc = cr_forward(c, nb)) return rc;
    if (int rc = cr_backward(c, nb, c->crops_out.p)) return rc;
    c->step++;
    const double b1 = 0.9, b2 = 0.999;
    const double lr_t = lr * std::sqrt(1.0 - std::pow(b2, c->step)) / (1.0 - std::pow(b1, c->step));
    utr::k_cr_adam<<<cr_blocks(c->np), 256, 0, c->st>>>(c->w.p, c->g.p, c->m.p, c->v.p, c->np, (float)lr_t, (float)b1,
                                                        (float)b2, 1e-8f);
    UTR_HIP_TRY(hipGetLastError());
    if (int rc = cr_repack(c)) return rc;
    // returns with the step done, so the caller may reuse the crop buffers
    float res[20];
    UTR_HIP_TRY(hipMemcpyAsync(res, c->res.p, sizeof(res), hipMemcpyDeviceToHost, c->st));
    UTR_HIP_TRY(hipStreamSynchronize(c->st));
    if (loss) *loss = res[18];
    return UTR_OK;
}

int utr_corr_apply(utr_corr* c, const float* d_vol, float* d_out, int nx, int ny, int nz) {
    if (int rc = cr_check(c)) return rc;
    if (!d_vol || !d_out || d_vol == d_out) return set_err(UTR_E_INVALID, "utr_corr_apply: distinct input and output");
    if (nx < 1 || ny < 1 || nz < 1) return set_err(UTR_E_INVALID, "utr_corr_apply: sizes must be positive");
    const int ntx = (nx + c->tile - 1) / c->tile, nty = (ny + c->tile - 1) / c->tile, ntz = (nz + c->tile - 1) / c->tile;
    const int ntiles = ntx * nty * ntz;
    const size_t per = (size_t)c->T * c->T * c->T;
    for (int first = 0; first < ntiles; first += c->batch) {
        const int nb = std::min(c->batch, ntiles - first);
        utr::k_cr_gather<<<cr_blocks(per * nb), 256, 0, c->st>>>(d_vol, nx, ny, nz, c->tile, c->halo, ntx, nty, first,
                                                                 nb, c->crops_in.p);
        UTR_HIP_TRY(hipGetLastError());
        if (int rc = cr_load_crops(c, c->crops_in.p, nb)) return rc;
        if (int rc = cr_forward(c, nb)) return rc;
        const size_t pt = (size_t)c->tile * c->tile * c->tile;
        utr::k_cr_scatter<<<cr_blocks(pt * nb), 256, 0, c->st>>>(c->out.p, nx, ny, nz, c->tile, c->halo, ntx, nty,
                                                                 first, nb, 1.f / c->scale, d_out);
        UTR_HIP_TRY(hipGetLastError());
    }
    UTR_HIP_TRY(hipStreamSynchronize(c->st));
    return UTR_OK;
}

int utr_corr_destroy(utr_corr* c) {
    if (!c) return UTR_OK;
    (void)hipSetDevice(c->device);
    if (c->st) {
        (void)hipStreamSynchronize(c->st);
        (void)hipStreamDestroy(c->st);
    }
    for (auto& Ly : c->L) {
        Ly.wf.release();
        Ly.wb.release();
    }
    DevBuf<bf16_t>* bb[] = {&c->x0, &c->e1a, &c->C1, &c->p1, &c->e2a, &c->C2, &c->p2, &c->e3a, &c->e3b,
                            &c->d2a, &c->d2b, &c->d1a, &c->d1b, &c->gd1b, &c->gd1a, &c->gC1, &c->gd2b,
                            &c->gd2a, &c->gC2, &c->ge3b, &c->ge3a, &c->gp2, &c->ge2a, &c->gp1, &c->ge1a};
    for (auto* b : bb) b->release();
    DevBuf<float>* fb[] = {&c->w, &c->g, &c->m, &c->v, &c->res, &c->x0f, &c->out, &c->part, &c->hpart,
                           &c->crops_in, &c->crops_out};
    for (auto* b : fb) b->release();
    delete c;
    return UTR_OK;
}

}  // extern "C"
