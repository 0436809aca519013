// utr_fft.hpp -- LDS-resident Stockham FFT for gfx950 (device side).
//
// A workgroup transforms `nb` sequences of length N that sit contiguously in
// LDS (sequence-major).  Each radix-R pass with p = product of the previous
// radices reads x[i + r*N/R], twiddles it by w^(r*(i mod p)) with
// w = exp(sign*2*pi*i/(p*R)), applies a length-R DFT and writes
// y[(i/p)*p*R + (i mod p) + q*p]; after the last pass the data are in natural
// order (Stockham autosort, no bit reversal).  Two LDS buffers are used in
// ping-pong.  Twiddles come from a table computed in FP64 on the host.
//
// The forward transform replaces the FFTW plans of Fftwd2fc / Fftwd3frc used
// at UTR.cpp:851,994,1223-1232 (exponent sign -1 forward, +1 inverse,
// unnormalised).
#pragma once

#include <hip/hip_runtime.h>

namespace utr {

template <typename T>
struct cplx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) { return {a.x - b.x, a.y - b.y}; }
// multiply by sign*i
template <typename T>
__device__ __forceinline__ cplx<T> mul_si(cplx<T> a, int sign) {
    return sign > 0 ? cplx<T>{-a.y, a.x} : cplx<T>{a.y, -a.x};
}

// ------------------------------------------------------------- LDS layout
// One pad entry after every 8: the strided writes of the first Stockham
// passes (stride R, then R*p for p < 64) then fall on distinct LDS banks
// instead of 4 to 16 lanes sharing one.  (The transposed column loads of
// 8 sequences of 256 still share banks; two layouts that spread them, an
// extra pad per 128 entries and an XOR swizzle, measured slower overall.)
// A buffer of n logical entries holds lds_len(n); entry i sits at lx(i).
__host__ __device__ constexpr int lds_len(int n) { return n + (n >> 3); }
__device__ __forceinline__ int lx(int i) { return i + (i >> 3); }

// ---------------------------------------------------------------- factors
struct FacList {
    int nf;
    int f[24];
    int p[24];    // product of earlier radices
    int off[24];  // twiddle-table offset of each pass
    int ntw;      // table length
};

// Radix order: 8s first, then 4, 2, 3, 5, 7 and any larger prime.
constexpr FacList factorize(int n) {
    FacList L{};
    L.nf = 0;
    int m = n;
    while (m % 8 == 0 && m > 8) { L.f[L.nf++] = 8; m /= 8; }
    if (m % 8 == 0) { L.f[L.nf++] = 8; m /= 8; }
    while (m % 4 == 0) { L.f[L.nf++] = 4; m /= 4; }
    while (m % 2 == 0) { L.f[L.nf++] = 2; m /= 2; }
    while (m % 3 == 0) { L.f[L.nf++] = 3; m /= 3; }
    while (m % 5 == 0) { L.f[L.nf++] = 5; m /= 5; }
    for (int q = 7; m > 1; q += 2)
        while (m % q == 0) { L.f[L.nf++] = q; m /= q; }
    int p = 1, off = 0;
    for (int s = 0; s < L.nf; s++) {
        L.p[s] = p;
        L.off[s] = off;
        off += p * (L.f[s] - 1);
        p *= L.f[s];
    }
    L.ntw = off;
    return L;
}

// ---------------------------------------------------------- small DFTs
// In-place length-R DFT of u[0..R-1] with exponent sign*2*pi*i*r*q/R.
template <typename T, int R>
__device__ __forceinline__ void dft_small(cplx<T>* u, int sign) {
    if constexpr (R == 2) {
        cplx<T> a = u[0], b = u[1];
        u[0] = cadd(a, b);
        u[1] = csub(a, b);
    } else if constexpr (R == 4) {
        cplx<T> a0 = cadd(u[0], u[2]), a1 = csub(u[0], u[2]);
        cplx<T> b0 = cadd(u[1], u[3]), b1 = mul_si(csub(u[1], u[3]), sign);
        u[0] = cadd(a0, b0);
        u[2] = csub(a0, b0);
        u[1] = cadd(a1, b1);
        u[3] = csub(a1, b1);
    } else if constexpr (R == 8) {
        const T h = T(0.70710678118654752440);
        // radix-2 split: even/odd 4-point DFTs
        cplx<T> e[4] = {u[0], u[2], u[4], u[6]};
        cplx<T> o[4] = {u[1], u[3], u[5], u[7]};
        dft_small<T, 4>(e, sign);
        dft_small<T, 4>(o, sign);
        // twiddle o[q] by exp(sign*2*pi*i*q/8)
        cplx<T> o1 = {h * (o[1].x - sign * o[1].y), h * (o[1].y + sign * o[1].x)};
        cplx<T> o2 = mul_si(o[2], sign);
        cplx<T> o3 = {h * (-o[3].x - sign * o[3].y), h * (-o[3].y + sign * o[3].x)};
        u[0] = cadd(e[0], o[0]);
        u[4] = csub(e[0], o[0]);
        u[1] = cadd(e[1], o1);
        u[5] = csub(e[1], o1);
        u[2] = cadd(e[2], o2);
        u[6] = csub(e[2], o2);
        u[3] = cadd(e[3], o3);
        u[7] = csub(e[3], o3);
    } else if constexpr (R == 3) {
        const T c1 = T(-0.5), s1 = T(0.86602540378443864676) * T(sign);
        cplx<T> a = u[0], b = u[1], c = u[2];
        cplx<T> t = cadd(b, c), d = csub(b, c);
        u[0] = cadd(a, t);
        cplx<T> m = {a.x + c1 * t.x, a.y + c1 * t.y};
        cplx<T> r = {-s1 * d.y, s1 * d.x};
        u[1] = cadd(m, r);
        u[2] = csub(m, r);
    } else if constexpr (R == 5) {
        const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);
        const T s1 = T(0.95105651629515357212) * T(sign), s2 = T(0.58778525229247312917) * T(sign);
        cplx<T> a = u[0];
        cplx<T> t1 = cadd(u[1], u[4]), d1 = csub(u[1], u[4]);
        cplx<T> t2 = cadd(u[2], u[3]), d2 = csub(u[2], u[3]);
        u[0] = {a.x + t1.x + t2.x, a.y + t1.y + t2.y};
        cplx<T> m1 = {a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y};
        cplx<T> m2 = {a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y};
        cplx<T> r1 = {-(s1 * d1.y + s2 * d2.y), s1 * d1.x + s2 * d2.x};
        cplx<T> r2 = {-(s2 * d1.y - s1 * d2.y), s2 * d1.x - s1 * d2.x};
        u[1] = cadd(m1, r1);
        u[4] = csub(m1, r1);
        u[2] = cadd(m2, r2);
        u[3] = csub(m2, r2);
    } else {
        // generic odd prime: direct O(R^2) DFT with sincos (rare sizes only)
        cplx<T> v[R];
        for (int q = 0; q < R; q++) {
            T sr = 0, si = 0;
            for (int r = 0; r < R; r++) {
                double ang = 6.283185307179586476925 * (double)((r * q) % R) / R * sign;
                T c = T(cos(ang)), s = T(sin(ang));
                sr += u[r].x * c - u[r].y * s;
                si += u[r].x * s + u[r].y * c;
            }
            v[q] = {sr, si};
        }
        for (int q = 0; q < R; q++) u[q] = v[q];
    }
}

// One Stockham pass for all nb sequences.
template <typename T, int N, int R>
__device__ __forceinline__ void fft_pass(const cplx<T>* __restrict__ src, cplx<T>* __restrict__ dst,
                                         int nb, int p, const cplx<T>* __restrict__ tw, int sign,
                                         int tid, int nthr) {
    constexpr int M = N / R;
    const int total = nb * M;
    for (int t = tid; t < total; t += nthr) {
        const int seq = t / M;
        const int i = t - seq * M;
        const int k = i % p;
        const int s0 = seq * N + i;
        cplx<T> u[R];
#pragma unroll
        for (int r = 0; r < R; r++) u[r] = src[lx(s0 + r * M)];
        if (p > 1) {
            const cplx<T>* w = tw + k * (R - 1);
#pragma unroll
            for (int r = 1; r < R; r++) {
                cplx<T> c = w[r - 1];
                if (sign > 0) c.y = -c.y;
                u[r] = cmul(u[r], c);
            }
        }
        dft_small<T, R>(u, sign);
        const int d0 = seq * N + (i / p) * p * R + k;
#pragma unroll
        for (int q = 0; q < R; q++) dst[lx(d0 + q * p)] = u[q];
    }
}

template <typename T, int N, int S>
__device__ __forceinline__ void fft_passes(cplx<T>*& a, cplx<T>*& b, int nb,
                                           const cplx<T>* __restrict__ tw, int sign, int tid,
                                           int nthr) {
    constexpr FacList F = factorize(N);
    if constexpr (S < F.nf) {
        constexpr int R = F.f[S];
        fft_pass<T, N, R>(a, b, nb, F.p[S], tw + F.off[S], sign, tid, nthr);
        __syncthreads();
        cplx<T>* t = a;
        a = b;
        b = t;
        fft_passes<T, N, S + 1>(a, b, nb, tw, sign, tid, nthr);
    }
}

// Transforms nb sequences of length N held in `a` (LDS, entry i of the
// concatenated sequences at lx(i)).  `b` is scratch of the same size.  On return `a` points at the buffer holding the result.
// The caller must __syncthreads() before the call (data loaded) and reads
// the result after it (this function ends with a barrier).
// tw: forward-sign twiddle table for length N (FacList layout).
template <typename T, int N>
__device__ __forceinline__ void fft_lds(cplx<T>*& a, cplx<T>*& b, int nb,
                                        const cplx<T>* __restrict__ tw, int sign, int tid,
                                        int nthr) {
    fft_passes<T, N, 0>(a, b, nb, tw, sign, tid, nthr);
}

}  // namespace utr
