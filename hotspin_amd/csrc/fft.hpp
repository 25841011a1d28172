// fft.hpp -- LDS-staged mixed-radix FFT for gfx950, written for this library.
//
// A block transforms W pencils of compile-time length L held in LDS with the
// layout lds[t*W + p] (pencil index fastest, so a wave64 touches 64
// consecutive complex words).  Each Stockham stage of radix R reads R
// elements per butterfly into registers, applies twiddles from a global
// table tw[t] = exp(-2 pi i t/L), does the radix-R DFT in registers and
// writes the outputs back in self-sorted order:
//     v_r = x[j + r L/R] * w^(r (j mod Ns)),  y[(j/Ns) Ns R + j mod Ns + r Ns] = DFT_R(v)_r
// with Ns the product of the earlier radices.  Radices 8, 4, 2, 3, 5.
#pragma once

#include <hip/hip_runtime.h>

template <typename T>
struct cpx {
    T x, y;
};

template <typename T>
__host__ __device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__host__ __device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__host__ __device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> cscale(cpx<T> a, T s) { return {a.x * s, a.y * s}; }
// (s i) * a  for s = +-1
template <typename T, int S>
__host__ __device__ __forceinline__ cpx<T> mul_si(cpx<T> a) {
    return S > 0 ? cpx<T>{-a.y, a.x} : cpx<T>{a.y, -a.x};
}

// ---- radix-R DFTs in registers; S = -1 forward, +1 backward --------------
template <typename T, int S>
__host__ __device__ __forceinline__ void dft2(cpx<T>* v) {
    cpx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}
template <typename T, int S>
__host__ __device__ __forceinline__ void dft4(cpx<T>* v) {
    cpx<T> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    cpx<T> t2 = cadd(v[1], v[3]), t3 = mul_si<T, S>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
}
template <typename T, int S>
__host__ __device__ __forceinline__ void dft8(cpx<T>* v) {
    cpx<T> e[4] = {v[0], v[2], v[4], v[6]};
    cpx<T> o[4] = {v[1], v[3], v[5], v[7]};
    dft4<T, S>(e);
    dft4<T, S>(o);
    const T h = T(0.70710678118654752440);
    // w8^q = exp(S 2 pi i q / 8)
    cpx<T> o1 = {h * (o[1].x - T(S) * o[1].y), h * (o[1].y + T(S) * o[1].x)};
    cpx<T> o2 = mul_si<T, S>(o[2]);
    cpx<T> o3 = {h * (-o[3].x - T(S) * o[3].y), h * (-o[3].y + T(S) * o[3].x)};
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
}
template <typename T, int S>
__host__ __device__ __forceinline__ void dft3(cpx<T>* v) {
    const T c = T(-0.5), sn = T(S) * T(0.86602540378443864676);
    cpx<T> t = cadd(v[1], v[2]);
    cpx<T> d = csub(v[1], v[2]);
    cpx<T> m1 = {v[0].x + c * t.x, v[0].y + c * t.y};
    cpx<T> m2 = {-sn * d.y, sn * d.x};
    v[0] = cadd(v[0], t);
    v[1] = cadd(m1, m2);
    v[2] = csub(m1, m2);
}
template <typename T, int S>
__host__ __device__ __forceinline__ void dft5(cpx<T>* v) {
    const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);
    const T s1 = T(S) * T(0.95105651629515357212), s2 = T(S) * T(0.58778525229247312917);
    cpx<T> t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    cpx<T> t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    cpx<T> a0 = v[0];
    v[0] = cadd(a0, cadd(t1, t2));
    cpx<T> b1 = {a0.x + c1 * t1.x + c2 * t2.x, a0.y + c1 * t1.y + c2 * t2.y};
    cpx<T> b2 = {a0.x + c2 * t1.x + c1 * t2.x, a0.y + c2 * t1.y + c1 * t2.y};
    cpx<T> e1 = {s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y};
    cpx<T> e2 = {s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y};
    cpx<T> d1 = {-e1.y, e1.x}, d2 = {-e2.y, e2.x};
    v[1] = cadd(b1, d1);
    v[4] = csub(b1, d1);
    v[2] = cadd(b2, d2);
    v[3] = csub(b2, d2);
}
template <typename T, int R, int S>
__host__ __device__ __forceinline__ void dftR(cpx<T>* v) {
    if constexpr (R == 2) dft2<T, S>(v);
    else if constexpr (R == 3) dft3<T, S>(v);
    else if constexpr (R == 4) dft4<T, S>(v);
    else if constexpr (R == 5) dft5<T, S>(v);
    else if constexpr (R == 8) dft8<T, S>(v);
}

// ---- compile-time radix schedule -----------------------------------------
struct Radices {
    int n;
    int r[24];
};
constexpr Radices radices_of(int L) {
    Radices f{0, {}};
    int m = L;
    while (m % 8 == 0 && m != 16) { f.r[f.n++] = 8; m /= 8; }
    while (m % 4 == 0) { f.r[f.n++] = 4; m /= 4; }
    while (m % 2 == 0) { f.r[f.n++] = 2; m /= 2; }
    while (m % 3 == 0) { f.r[f.n++] = 3; m /= 3; }
    while (m % 5 == 0) { f.r[f.n++] = 5; m /= 5; }
    if (m != 1) f.n = -1;  // unsupported length
    return f;
}
constexpr bool fft_supported(int L) { return L >= 1 && radices_of(L).n >= 0; }

// One Stockham stage over W pencils held in lds[t*W+p].
template <typename T, int L, int W, int NT, int R, int NS, int S>
__device__ __forceinline__ void fft_stage(cpx<T>* lds, const cpx<T>* __restrict__ tw, int tid) {
    constexpr int NB = L / R;        // butterflies per pencil
    constexpr int TOT = NB * W;      // butterflies in the block
    constexpr int BPT = (TOT + NT - 1) / NT;
    constexpr int TWS = L / (NS * R);
    cpx<T> v[BPT][R];
#pragma unroll
    for (int q = 0; q < BPT; q++) {
        int b = tid + q * NT;
        if (TOT % NT == 0 || b < TOT) {
            int p = b % W, j = b / W;
            int jl = j % NS;
#pragma unroll
            for (int r = 0; r < R; r++) {
                cpx<T> a = lds[(j + r * NB) * W + p];
                if (NS > 1 && r > 0) {
                    cpx<T> w = tw[r * jl * TWS];
                    if (S > 0) w.y = -w.y;
                    a = cmul(a, w);
                }
                v[q][r] = a;
            }
            dftR<T, R, S>(v[q]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < BPT; q++) {
        int b = tid + q * NT;
        if (TOT % NT == 0 || b < TOT) {
            int p = b % W, j = b / W;
            int jl = j % NS;
            int d = (j / NS) * NS * R + jl;
#pragma unroll
            for (int r = 0; r < R; r++) lds[(d + r * NS) * W + p] = v[q][r];
        }
    }
    __syncthreads();
}

template <typename T, int L, int W, int NT, int S, int ST, int NS>
__device__ __forceinline__ void fft_stages(cpx<T>* lds, const cpx<T>* __restrict__ tw, int tid) {
    constexpr Radices F = radices_of(L);
    if constexpr (ST < F.n) {
        constexpr int R = F.r[ST];
        fft_stage<T, L, W, NT, R, NS, S>(lds, tw, tid);
        fft_stages<T, L, W, NT, S, ST + 1, NS * R>(lds, tw, tid);
    }
}

// In-place, unnormalised transform of the W pencils in lds (layout [t][p]).
// Must be entered by all NT threads with the data already visible (after a
// __syncthreads); returns with the result visible to all threads.
template <typename T, int L, int W, int NT, int S>
__device__ __forceinline__ void fft_lds(cpx<T>* lds, const cpx<T>* __restrict__ tw, int tid) {
    static_assert(radices_of(L).n >= 0, "FFT length must be a product of 2, 3 and 5");
    if constexpr (L > 1) fft_stages<T, L, W, NT, S, 0, 1>(lds, tw, tid);
}

// Supported lengths (the plan refuses others with HS_ESIZE).
#define HS_FFT_SIZES(X)                                                                                    \
    X(1) X(2) X(4) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(3) X(6) X(12) X(24) \
        X(48) X(96) X(192) X(384) X(768) X(1536) X(5) X(10) X(20) X(40) X(80) X(160) X(320) X(640)   \
        X(1280)
