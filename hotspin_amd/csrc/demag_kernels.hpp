// demag_kernels.hpp -- the five FFT pass kernels of the demag convolution
// (see demag.hip for the pipeline).  Tile shapes (pencils per block W,
// threads NT) are template parameters so the plan can time variants.
#pragma once

#include <type_traits>

#include "fft.hpp"

// ------------------------------------------------------------------ tiling
template <typename T, int L>
constexpr int col_w() {  // adjacent columns per block: <= 128 KB of LDS
    int w = 16;
    while (w > 1 && (long)L * w * (long)sizeof(cpx<T>) > 131072) w /= 2;
    return w;
}
constexpr int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
template <typename T, int L>
constexpr int col_nt() {
    return clampi(L * col_w<T, L>() / 16, 64, 1024);
}
template <int M>
constexpr int row_w() {
    return clampi(4096 / (M > 0 ? M : 1), 1, 16);
}
template <int M>
constexpr int row_nt() {
    return clampi(M * row_w<M>() / 8, 64, 512);
}
// columns per block of the axis-0 convolution pass (pencil pair: 6W pencils
// of length L in LDS, at most 128 KB)
template <typename T, int L>
constexpr int z_w() {
    int w = 16;
    while (w > 1 && (long)L * 6 * w * (long)sizeof(cpx<T>) > 131072) w /= 2;
    return w;
}
template <typename T, int L>
constexpr int z_nt() {
    return clampi(L * 6 * z_w<T, L>() / 16, 64, 1024);
}

// Per-solve liveness of a batched pass: work units [q*per, (q+1)*per) (rows,
// pencils or solves) belong to solve q; flag null = every solve live.  A
// block whose units all belong to one solve with flag 0 exits at once (idle
// slots of a continuous batch cost no bandwidth).
struct Live {
    const int* flag;
    long per;
};
__device__ __forceinline__ bool all_dead(const Live& lv, long u0, long u1) {
    if (!lv.flag) return false;
    const long q0 = u0 / lv.per, q1 = u1 / lv.per;
    return q0 == q1 && !lv.flag[q0];
}

#include "xcd.hpp"

struct Geo {  // pencil a starts at (a / A2) * S1 + (a % A2) * S2; element t at + t * ts
    long S1, S2, ts;
    int A2;
};
__device__ __forceinline__ long geo_base(const Geo& g, int a) { return (long)(a / g.A2) * g.S1 + (long)(a % g.A2) * g.S2; }

// ------------------------------------------------------------------ F1: R2C rows
// x: rows of n2 reals (row r at x + r*n2), optional mask of the same layout
// (applied per cell, rows taken modulo mrows), out: rows of H2 complex at
// stride H2p.  z[n] = x[2n] + i x[2n+1]; X[k] = E + w^k O with
// E = (Z[k] + conj Z[M-k])/2, O = -i (Z[k] - conj Z[M-k])/2, w = exp(-2 pi i/P2).
template <typename T, typename TI, int M, int W = row_w<M>(), int NT = row_nt<M>()>
__global__ void __launch_bounds__(NT) k_r2c_rows(const TI* __restrict__ x, const TI* __restrict__ msk,
                                                         long mrows, cpx<T>* __restrict__ out, long nrows, int n2,
                                                         int H2p, const cpx<T>* __restrict__ twM,
                                                         const cpx<T>* __restrict__ twP, Live lv) {
    __shared__ cpx<T> lds[M * W];
    const int tid = threadIdx.x;
    long row0 = (long)blockIdx.x * W;
    if (all_dead(lv, row0, (row0 + W <= nrows ? row0 + W : nrows) - 1)) return;
    const bool pairs = (n2 & 1) == 0;  // rows of even length: each complex input is one aligned pair
    // loads in equal batches of at most 24 per thread (no short last batch), each batch in flight before its LDS stores
    constexpr int NE = M * W, IT = (NE + NT - 1) / NT, NB = (IT + 23) / 24, CH = (IT + NB - 1) / NB;
    for (int q0 = 0; q0 < IT; q0 += CH) {
        cpx<T> g[CH];
#pragma unroll
        for (int q = 0; q < CH; q++) {
            const int e = tid + (q0 + q) * NT;
            int p = e / M, n = e % M;
            long r = row0 + p;
            cpx<T> v = {T(0), T(0)};
            if (e < NE && r < nrows) {
                long base = r * n2;
                long mb = (r % mrows) * n2;
                int i0 = 2 * n, i1 = 2 * n + 1;
                if (pairs) {
                    if (i0 < n2) {
                        using V2 = typename std::conditional<sizeof(TI) == 4, float2, double2>::type;
                        const V2 a = *(const V2*)(x + base + i0);
                        v = {T(a.x), T(a.y)};
                        if (msk) {
                            const V2 w = *(const V2*)(msk + mb + i0);
                            v.x *= T(w.x);
                            v.y *= T(w.y);
                        }
                    }
                } else {
                    if (i0 < n2) v.x = T(x[base + i0]) * (msk ? T(msk[mb + i0]) : T(1));
                    if (i1 < n2) v.y = T(x[base + i1]) * (msk ? T(msk[mb + i1]) : T(1));
                }
            }
            g[q] = v;
        }
#pragma unroll
        for (int q = 0; q < CH; q++) {
            const int e = tid + (q0 + q) * NT;
            const int p = e / M, n = e % M;
            if (e < NE) lds[n * W + p] = g[q];
        }
    }
    __syncthreads();
    fft_lds<T, M, W, NT, -1>(lds, twM, tid);
    for (int e = tid; e < (M + 1) * W; e += NT) {
        int p = e / (M + 1), k = e % (M + 1);
        long r = row0 + p;
        if (r >= nrows) continue;
        cpx<T> zk = lds[(k % M) * W + p];
        cpx<T> zm = lds[((M - k) % M) * W + p];
        cpx<T> E = {T(0.5) * (zk.x + zm.x), T(0.5) * (zk.y - zm.y)};
        cpx<T> O = {T(0.5) * (zk.y + zm.y), T(-0.5) * (zk.x - zm.x)};
        cpx<T> w = twP[k];
        out[r * H2p + k] = cadd(E, cmul(w, O));
    }
}

// ------------------------------------------------------------------ I1: C2R rows
// in: rows of H2 = M+1 complex (stride H2p), out: rows of n2 reals keeping
// the first n2 of P2 = 2M.  Z[k] = (X[k] + conj X[M-k]) + i w^-k (X[k] - conj X[M-k]),
// z = IFFT_M(Z), x[2n] = Re z[n], x[2n+1] = Im z[n] (unnormalised C2R).
template <typename T, typename TO, int M, int W = row_w<M>(), int NT = row_nt<M>()>
__global__ void __launch_bounds__(NT) k_c2r_rows(const cpx<T>* __restrict__ in, TO* __restrict__ x,
                                                         long nrows, int n2, int H2p, const cpx<T>* __restrict__ twM,
                                                         const cpx<T>* __restrict__ twP, Live lv) {
    __shared__ cpx<T> lds[(M + 1) * W];
    const int tid = threadIdx.x;
    long row0 = (long)blockIdx.x * W;
    if (all_dead(lv, row0, (row0 + W <= nrows ? row0 + W : nrows) - 1)) return;
    {  // loads in equal batches of at most 24 per thread (no short last batch), each batch in flight before its LDS stores
        constexpr int NE = (M + 1) * W, IT = (NE + NT - 1) / NT, NB = (IT + 23) / 24, CH = (IT + NB - 1) / NB;
        for (int q0 = 0; q0 < IT; q0 += CH) {
            cpx<T> g[CH];
#pragma unroll
            for (int q = 0; q < CH; q++) {
                const int e = tid + (q0 + q) * NT;
                const int p = e / (M + 1), k = e % (M + 1);
                const long r = row0 + p;
                g[q] = {T(0), T(0)};
                if (e < NE && r < nrows) g[q] = in[r * H2p + k];
            }
#pragma unroll
            for (int q = 0; q < CH; q++) {
                const int e = tid + (q0 + q) * NT;
                const int p = e / (M + 1), k = e % (M + 1);
                if (e < NE) lds[k * W + p] = g[q];
            }
        }
    }
    __syncthreads();
    constexpr int PER = (M * W + NT - 1) / NT;
    cpx<T> z[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
        int e = tid + q * NT;
        if (e < M * W) {
            int k = e / W, p = e % W;
            cpx<T> a = lds[k * W + p];
            cpx<T> b = lds[(M - k) * W + p];
            cpx<T> s = {a.x + b.x, a.y - b.y};      // X[k] + conj X[M-k]
            cpx<T> d = {a.x - b.x, a.y + b.y};      // X[k] - conj X[M-k]
            cpx<T> w = twP[k];
            w.y = -w.y;                            // w^-k
            cpx<T> wd = cmul(w, d);
            z[q] = {s.x - wd.y, s.y + wd.x};       // s + i w^-k d
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; q++) {
        int e = tid + q * NT;
        if (e < M * W) lds[e] = z[q];
    }
    __syncthreads();
    fft_lds<T, M, W, NT, +1>(lds, twM, tid);
    const bool pairs = (n2 & 1) == 0;
    for (int e = tid; e < M * W; e += NT) {
        int p = e / M, n = e % M;
        long r = row0 + p;
        if (r >= nrows) continue;
        cpx<T> v = lds[n * W + p];
        int i0 = 2 * n, i1 = 2 * n + 1;
        if (pairs) {
            if (i0 < n2) {
                using V2 = typename std::conditional<sizeof(TO) == 4, float2, double2>::type;
                V2 o;
                o.x = TO(v.x);
                o.y = TO(v.y);
                *(V2*)(x + r * n2 + i0) = o;
            }
        } else {
            if (i0 < n2) x[r * n2 + i0] = TO(v.x);
            if (i1 < n2) x[r * n2 + i1] = TO(v.y);
        }
    }
}

// ------------------------------------------------------------------ F2 / I2: columns
// Pencil a (blockIdx.y), columns col0..col0+W (blockIdx.x); input rows t < nin
// (zero beyond), output rows t < nkeep.  in and out may be the same buffer.
template <typename T, int L, int S, int W = col_w<T, L>(), int NT = col_nt<T, L>()>
__global__ void __launch_bounds__(NT) k_cols(const cpx<T>* in, cpx<T>* out, Geo gi, Geo go, int nin, int nkeep,
                                             int ncol, const cpx<T>* __restrict__ tw, Live lv) {
    __shared__ cpx<T> lds[L * W];
    const int tid = threadIdx.x;
    const int lb = xcd_swizzle(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int a = lb / gridDim.x;
    if (all_dead(lv, a, a)) return;
    const int col0 = (lb % gridDim.x) * W;
    const long bi = geo_base(gi, a) + col0;
    const long bo = geo_base(go, a) + col0;
    {  // loads in equal batches of at most 24 per thread (no short last batch), each batch in flight before its LDS stores
        constexpr int NE = L * W, IT = (NE + NT - 1) / NT, NB = (IT + 23) / 24, CH = (IT + NB - 1) / NB;
        for (int q0 = 0; q0 < IT; q0 += CH) {
            cpx<T> g[CH];
#pragma unroll
            for (int q = 0; q < CH; q++) {
                const int e = tid + (q0 + q) * NT;
                const int t = e / W, p = e % W;
                g[q] = {T(0), T(0)};
                if (e < NE && t < nin && col0 + p < ncol) g[q] = in[bi + t * gi.ts + p];
            }
#pragma unroll
            for (int q = 0; q < CH; q++) {
                const int e = tid + (q0 + q) * NT;
                if (e < NE) lds[e] = g[q];
            }
        }
    }
    __syncthreads();
    fft_lds<T, L, W, NT, S>(lds, tw, tid);
    for (int e = tid; e < nkeep * W; e += NT) {
        int t = e / W, p = e % W;
        if (col0 + p < ncol) out[bo + t * go.ts + p] = lds[e];
    }
}

// ------------------------------------------------------------------ F3K: axis 0 + kernel product
// X2 layout (3, N0, P1, H2p).  A block takes a tile of W columns of pencil
// j' = blockIdx.y and of its mirror P1 - j' (both use kernel row |j'|, so the
// kernel octant is read once for the pair), all 3 components: 6W pencils of
// length L in LDS as lds[t * 6W + s * 3W + c * W + p], s = 0 (j') / 1 (mirror).
// Kernel octants: 6 planes (XX,YY,ZZ,YZ,XZ,XY) of (K0, K1, H2), rows at
// pitch kp (kp >= H2; padded so FP32 rows start on 128-byte lines).  Grid y =
// P1/2 + 1; j' = 0 and j' = P1/2 (even P1) have no distinct mirror.
template <typename T, typename TK, int L, int W = z_w<T, L>(), int NT = z_nt<T, L>()>
__global__ void __launch_bounds__(NT) k_zconv(cpx<T>* X, const TK* __restrict__ kern, int N0, int P1, int H2,
                                              int H2p, int K1, int kp, T srcMul, const cpx<T>* __restrict__ tw,
                                              Live lv) {
    constexpr int WP = 3 * W, W2 = 2 * WP;
    if (all_dead(lv, blockIdx.z, blockIdx.z)) return;
    X += (long)blockIdx.z * 3 * N0 * P1 * H2p;  // solve of a batch
    __shared__ cpx<T> lds[L * W2];
    const int tid = threadIdx.x;
    const int lb = xcd_swizzle(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int jp = lb / gridDim.x;                              // 0 .. P1/2
    const int jm = (jp == 0 || 2 * jp == P1) ? -1 : P1 - jp;    // mirror pencil, if distinct
    const int col0 = (lb % gridDim.x) * W;
    const long cstride = (long)N0 * P1 * H2p;
    const long tstride = (long)P1 * H2p;
    for (int e = tid; e < L * W2; e += NT) {
        int t = e / W2, q = e % W2;
        int sm = q / WP, r = q % WP;
        int c = r / W, p = r % W;
        int j = sm ? jm : jp;
        cpx<T> v = {T(0), T(0)};
        if (t < N0 && col0 + p < H2 && j >= 0) v = X[c * cstride + (long)j * H2p + col0 + t * tstride + p];
        lds[e] = v;
    }
    __syncthreads();
    fft_lds<T, L, W2, NT, -1>(lds, tw, tid);
    const long kplane = (long)(L / 2 + 1) * K1 * kp;
    for (int e = tid; e < L * W; e += NT) {
        int t = e / W, p = e % W;
        int col = col0 + p;
        if (col >= H2) continue;
        int ii = (t > L / 2) ? L - t : t;
        T si = (t > L / 2) ? T(-1) : T(1);
        long K = ((long)ii * K1 + jp) * kp + col;
        T kxx = T(kern[K]), kyy = T(kern[kplane + K]), kzz = T(kern[2 * kplane + K]);
        T kyz0 = T(kern[3 * kplane + K]);
        T kxz = si * T(kern[4 * kplane + K]);
        T kxy0 = si * T(kern[5 * kplane + K]);
#pragma unroll
        for (int sm = 0; sm < 2; sm++) {  // mirror pencil: YZ and XY change sign
            const T sj = sm ? T(-1) : T(1);
            const T kyz = sj * kyz0, kxy = sj * kxy0;
            cpx<T>* a = lds + t * W2 + sm * WP + p;
            cpx<T> mx = cscale(a[0], srcMul);
            cpx<T> my = cscale(a[W], srcMul);
            cpx<T> mz = cscale(a[2 * W], srcMul);
            a[0] = {kxx * mx.x + kxy * my.x + kxz * mz.x, kxx * mx.y + kxy * my.y + kxz * mz.y};
            a[W] = {kxy * mx.x + kyy * my.x + kyz * mz.x, kxy * mx.y + kyy * my.y + kyz * mz.y};
            a[2 * W] = {kxz * mx.x + kyz * my.x + kzz * mz.x, kxz * mx.y + kyz * my.y + kzz * mz.y};
        }
    }
    __syncthreads();
    fft_lds<T, L, W2, NT, +1>(lds, tw, tid);
    for (int e = tid; e < N0 * W2; e += NT) {
        int t = e / W2, q = e % W2;
        int sm = q / WP, r = q % WP;
        int c = r / W, p = r % W;
        int j = sm ? jm : jp;
        if (col0 + p < H2 && j >= 0) X[c * cstride + (long)j * H2p + col0 + t * tstride + p] = lds[e];
    }
}
