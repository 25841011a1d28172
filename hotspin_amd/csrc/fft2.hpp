// fft2.hpp -- register-blocked two-level FFT passes.
//
// A length L = A * B transform done as A-point DFTs in registers, a twiddle,
// one exchange through LDS, and B-point DFTs in registers:
//     X[k1 + A k2] = sum_n2 w_B^(n2 k2) w_L^(n2 k1) sum_n1 x[B n1 + n2] w_A^(n1 k1)
// so each element is written to and read from LDS once, instead of once per
// radix stage.  The A- and B-point DFTs are mixed-radix (2,3,4,5,8) in
// registers with compile-time twiddles.
#pragma once

#include <type_traits>

#include "demag_kernels.hpp"

// ---- compile-time roots of unity -----------------------------------------
constexpr double kPi2 = 6.283185307179586476925286766559;
constexpr double kPi = 3.141592653589793238462643383279;
constexpr double series_sin(double x) {  // |x| <= pi/4
    double x2 = x * x, term = x, sum = x;
    for (int n = 1; n < 14; n++) {
        term *= -x2 / ((2.0 * n) * (2.0 * n + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double series_cos(double x) {
    double x2 = x * x, term = 1.0, sum = 1.0;
    for (int n = 1; n < 14; n++) {
        term *= -x2 / ((2.0 * n - 1.0) * (2.0 * n));
        sum += term;
    }
    return sum;
}
// cos and sin of 2 pi m / R by octant symmetry on exact fractions of the
// circle, the series evaluated on [0, pi/4] only
struct CS {
    double c, s;
};
constexpr CS cis_quadrant(long m, long R) {  // angle 2 pi m / R in [0, pi/2]
    if (8 * m > R) {
        const double psi = kPi2 * (double)(R - 4 * m) / (4.0 * (double)R);
        return CS{series_sin(psi), series_cos(psi)};
    }
    const double t = kPi2 * (double)m / (double)R;
    return CS{series_cos(t), series_sin(t)};
}
constexpr CS cis2pi(long m, long R) {
    m %= R;
    if (m < 0) m += R;
    bool neg_s = false;
    if (2 * m > R) {  // theta > pi: reflect
        m = R - m;
        neg_s = true;
    }
    CS r{0.0, 0.0};
    if (4 * m > R) {  // theta > pi/2: pi - theta = 2 pi (R - 2m) / (2R)
        const CS q = cis_quadrant(R - 2 * m, 2 * R);
        r = CS{-q.c, q.s};
    } else {
        r = cis_quadrant(m, R);
    }
    if (neg_s) r.s = -r.s;
    return r;
}

template <int R>
struct Roots {
    double c[R > 0 ? R : 1], s[R > 0 ? R : 1];
    constexpr Roots() : c(), s() {
        for (int m = 0; m < R; m++) {
            CS v = cis2pi(m, R);
            c[m] = v.c;
            s[m] = v.s;
        }
    }
};
template <int R>
struct RootTable {
    static constexpr Roots<R> v{};
};

// w_R^m = exp(S 2 pi i m / R)
template <typename T, int R, int S>
__host__ __device__ __forceinline__ cpx<T> root(int m) {
    return {T(RootTable<R>::v.c[m]), T(S) * T(RootTable<R>::v.s[m])};
}

constexpr int first_factor(int R) {
    const int rs[5] = {8, 4, 2, 3, 5};
    for (int r : rs)
        if (R % r == 0 && R != r) return r;
    return R;
}
constexpr bool reg_dft_ok(int R) {
    int m = R;
    while (m % 2 == 0) m /= 2;
    while (m % 3 == 0) m /= 3;
    while (m % 5 == 0) m /= 5;
    return m == 1 && R >= 1 && R <= 64;
}

// In-place DFT of v[0..R) in registers, natural order (S = -1 forward).
template <typename T, int R, int S>
__host__ __device__ __forceinline__ void dft_reg(cpx<T>* v) {
    static_assert(reg_dft_ok(R), "register DFT length must be 2^a 3^b 5^c <= 64");
    if constexpr (R == 1) {
    } else if constexpr (R == 2 || R == 3 || R == 4 || R == 5 || R == 8) {
        dftR<T, R, S>(v);
    } else {
        constexpr int R1 = first_factor(R), R2 = R / R1;
        cpx<T> u[R];
#pragma unroll
        for (int n2 = 0; n2 < R2; n2++) {
            cpx<T> g[R1];
#pragma unroll
            for (int n1 = 0; n1 < R1; n1++) g[n1] = v[R2 * n1 + n2];
            dft_reg<T, R1, S>(g);
#pragma unroll
            for (int k1 = 0; k1 < R1; k1++) u[n2 * R1 + k1] = (n2 * k1) ? cmul(g[k1], root<T, R, S>(n2 * k1)) : g[k1];
        }
#pragma unroll
        for (int k1 = 0; k1 < R1; k1++) {
            cpx<T> g[R2];
#pragma unroll
            for (int n2 = 0; n2 < R2; n2++) g[n2] = u[n2 * R1 + k1];
            dft_reg<T, R2, S>(g);
#pragma unroll
            for (int k2 = 0; k2 < R2; k2++) v[k1 + R1 * k2] = g[k2];
        }
    }
}

// Split L = A * B with A, B register-DFT lengths, A <= B <= 32 preferred.
struct Split2 {
    int A, B;
};
constexpr Split2 split2(int L) {
    Split2 best{0, 0};
    for (int a = 1; a <= 64; a++) {
        if (L % a) continue;
        int b = L / a;
        if (!reg_dft_ok(a) || !reg_dft_ok(b) || a > b || b > 32) continue;
        best = Split2{a, b};  // keeps the largest a <= b
    }
    return best;
}
constexpr bool two_level_ok(int L) { return split2(L).A > 0; }

// ---- column pass (F2 / I2) ------------------------------------------------
// Phase-1 loads of the two-level column passes: thread (n2, column p) takes
// rows t = B n1 + n2, its own row plus a uniform step per n1.  For each n1 a
// uniform test decides: every lane's row is an input row (B (n1+1) <= nin),
// none is (B n1 >= nin: zero), or the lanes differ (a lane past the input
// rows loads row 0 of column col0 and zeroes the value).  A column past ncol reads column col0: its values stay in
// their own column of the tile and are never stored.  Every load is in flight
// before the first wait, and the arrays stay in registers (one unrolled loop
// of constant indices).
template <typename T, int A, int B>
__device__ __forceinline__ void cols2_load(cpx<T> (&v)[A], const cpx<T>* in, long bi, long ts, int n2, int p,
                                           bool pin, int nin) {
    const cpx<T>* lin = in + (pin ? bi + (long)n2 * ts + p : bi);
    const long bstep = (long)B * ts;
#pragma unroll
    for (int n1 = 0; n1 < A; n1++) {
        if (B * (n1 + 1) <= nin) {
            v[n1] = lin[n1 * bstep];
        } else if (B * n1 >= nin) {
            v[n1] = {T(0), T(0)};
        } else {  // a lane past the input rows reads row 0 of column col0 (always present)
            const bool ok = B * n1 + n2 < nin;
            v[n1] = *(ok ? lin + n1 * bstep : in + bi);
            if (!ok) v[n1] = {T(0), T(0)};
        }
    }
}

// Phase-2 stores: thread (q1, column p) holds rows t = q1 + A k2 (q1 < A),
// its own row plus a uniform step per k2; rows t >= nkeep are dropped (a
// uniform test when A (k2+1) <= nkeep or A k2 >= nkeep).
template <typename T, int A, int B>
__device__ __forceinline__ void cols2_store(const cpx<T> (&u)[B], cpx<T>* out, long bo, long ts, int q1, int p,
                                            int nkeep) {
    cpx<T>* lout = out + bo + (long)q1 * ts + p;
    const long astep = (long)A * ts;
#pragma unroll
    for (int k2 = 0; k2 < B; k2++) {
        if (A * (k2 + 1) <= nkeep) {
            lout[k2 * astep] = u[k2];
        } else if (A * k2 < nkeep) {
            if (q1 + A * k2 < nkeep) lout[k2 * astep] = u[k2];
        }
    }
}

// Pencil a, columns col0..col0+W; input rows t < nin (zero beyond), output
// rows t < nkeep; in and out may alias only if they are the same buffer with
// nin == nkeep == L (each block reads its tile completely before writing).
// Threads: W * B in phase 1 (one A-point DFT over n1 for each (n2, column)),
// W * A in phase 2 (one B-point DFT over n2 for each (k1, column)).
template <typename T, int L, int S, int W, int NT = W * split2(L).B>
__global__ void __launch_bounds__(NT) k_cols2(const cpx<T>* in, cpx<T>* out, Geo gi, Geo go, int nin, int nkeep,
                                              int ncol, const cpx<T>* __restrict__ tw, Live lv) {
    constexpr int A = split2(L).A, B = split2(L).B;
    static_assert(A * B == L && NT == W * B && A <= B, "two-level split");
    __shared__ cpx<T> lds[L * W];
    __shared__ cpx<T> tws[L];
    const int tid = threadIdx.x;
    const int lb = xcd_swizzle(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int a = lb / gridDim.x;
    if (all_dead(lv, a, a)) return;
    const int col0 = (lb % gridDim.x) * W;
    const long bi = geo_base(gi, a) + col0;
    const long bo = geo_base(go, a) + col0;
    for (int e = tid; e < L; e += NT) tws[e] = tw[e];
    // phase 1: thread (n2, p) gathers x[B n1 + n2], n1 < A
    const int p = tid % W, n2 = tid / W;
    const bool pin = col0 + p < ncol;
    cpx<T> v[B];  // A used in phase 1, B in phase 2
#pragma unroll
    for (int n1 = 0; n1 < A; n1++) {
        const int t = B * n1 + n2;
        cpx<T> x = {T(0), T(0)};
        if (pin && t < nin) x = in[bi + (long)t * gi.ts + p];
        v[n1] = x;
    }
    __syncthreads();  // twiddles visible
    dft_reg<T, A, S>(v);
#pragma unroll
    for (int k1 = 0; k1 < A; k1++) {
        cpx<T> w = tws[n2 * k1];
        if (S > 0) w.y = -w.y;
        lds[(k1 * B + n2) * W + p] = (n2 * k1) ? cmul(v[k1], w) : v[k1];
    }
    __syncthreads();
    // phase 2: thread (k1, p) gathers Y[k1][n2], n2 < B
    if (tid < W * A) {
        const int k1 = tid / W;
#pragma unroll
        for (int m = 0; m < B; m++) v[m] = lds[(k1 * B + m) * W + p];
        dft_reg<T, B, S>(v);
        if (pin) {
#pragma unroll
            for (int k2 = 0; k2 < B; k2++) {
                const int t = k1 + A * k2;
                if (t < nkeep) out[bo + (long)t * go.ts + p] = v[k2];
            }
        }
    }
}


// ---- column pass with a split LDS exchange --------------------------------
// The k_cols2 contract with the phase-1 -> phase-2 exchange done in two real
// halves (real parts, then imaginary parts) through an L*W array of T: half
// the LDS of k_cols2, so two blocks of W = 16 columns of length 1024 share a
// CU and overlap their load, transform and store phases.
template <typename T, int L, int S, int W, int NT = W * split2(L).B>
__global__ void __launch_bounds__(NT) k_cols2s(const cpx<T>* in, cpx<T>* out, Geo gi, Geo go, int nin, int nkeep,
                                               int ncol, const cpx<T>* __restrict__ tw, Live lv) {
    constexpr int A = split2(L).A, B = split2(L).B;
    static_assert(A * B == L && NT == W * B && A <= B, "two-level split");
    __shared__ T lds[L * W];
    __shared__ cpx<T> tws[L];
    const int tid = threadIdx.x;
    const int lb = xcd_swizzle(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int a = lb / gridDim.x;
    if (all_dead(lv, a, a)) return;
    const int col0 = (lb % gridDim.x) * W;
    const long bi = geo_base(gi, a) + col0;
    const long bo = geo_base(go, a) + col0;
    const int p = tid % W, n2 = tid / W;
    const bool pin = col0 + p < ncol;
    // every global load in flight before the first wait: the pencil data from
    // clamped (valid) addresses, zeroed afterwards outside the input rows,
    // then the twiddle table into registers
    cpx<T> v[A];
    cols2_load<T, A, B>(v, in, bi, gi.ts, n2, p, pin, nin);
    {
        constexpr int TW = (L + NT - 1) / NT;
        cpx<T> tv[TW];
#pragma unroll
        for (int q = 0; q < TW; q++) tv[q] = tw[tid + q * NT < L ? tid + q * NT : 0];
#pragma unroll
        for (int q = 0; q < TW; q++)
            if (tid + q * NT < L) tws[tid + q * NT] = tv[q];
    }
    __syncthreads();  // twiddles visible
    dft_reg<T, A, S>(v);
    // w_L^(n2 k1) for every k1 > 0 (tws[0] = 1 exactly, so n2 = 0 needs no
    // branch and the table reads go out together)
#pragma unroll
    for (int k1 = 1; k1 < A; k1++) {
        cpx<T> w = tws[n2 * k1];
        if (S > 0) w.y = -w.y;
        v[k1] = cmul(v[k1], w);
    }
#pragma unroll
    for (int k1 = 0; k1 < A; k1++) lds[(k1 * B + n2) * W + p] = v[k1].x;
    __syncthreads();
    const bool ph2 = tid < W * A;
    const int q1 = tid / W;
    cpx<T> u[B];
    if (ph2) {
#pragma unroll
        for (int m = 0; m < B; m++) u[m].x = lds[(q1 * B + m) * W + p];
    }
    __syncthreads();
#pragma unroll
    for (int k1 = 0; k1 < A; k1++) lds[(k1 * B + n2) * W + p] = v[k1].y;
    __syncthreads();
    if (ph2) {
#pragma unroll
        for (int m = 0; m < B; m++) u[m].y = lds[(q1 * B + m) * W + p];
        dft_reg<T, B, S>(u);
        if (pin) cols2_store<T, A, B>(u, out, bo, go.ts, q1, p, nkeep);
    }
}

// The split-exchange column pass with the phase-2 reads free of LDS bank
// conflicts: each k1 row of the exchange array is padded by W floats, so the
// 32 lanes of a half-wave (32 / W values of k1, W columns each) fall on 32
// distinct banks (three W = 8 blocks per CU instead of four).
template <typename T, int L, int S, int W, int MINW = 1, int NT = W * split2(L).B>
__global__ void __launch_bounds__(NT, MINW) k_cols2p(const cpx<T>* in, cpx<T>* out, Geo gi, Geo go, int nin, int nkeep,
                                               int ncol, const cpx<T>* __restrict__ tw, Live lv) {
    constexpr int A = split2(L).A, B = split2(L).B, RS = B * W + W;  // padded k1 row stride
    static_assert(A * B == L && NT == W * B && A <= B, "two-level split");
    __shared__ T lds[A * RS];
    __shared__ cpx<T> tws[L];
    const int tid = threadIdx.x;
    const int lb = xcd_swizzle(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int a = lb / gridDim.x;
    if (all_dead(lv, a, a)) return;
    const int col0 = (lb % gridDim.x) * W;
    const long bi = geo_base(gi, a) + col0;
    const long bo = geo_base(go, a) + col0;
    const int p = tid % W, n2 = tid / W;
    const bool pin = col0 + p < ncol;
    // every global load in flight before the first wait: the pencil data from
    // clamped (valid) addresses, zeroed afterwards outside the input rows,
    // then the twiddle table into registers
    cpx<T> v[A];
    cols2_load<T, A, B>(v, in, bi, gi.ts, n2, p, pin, nin);
    {
        constexpr int TW = (L + NT - 1) / NT;
        cpx<T> tv[TW];
#pragma unroll
        for (int q = 0; q < TW; q++) tv[q] = tw[tid + q * NT < L ? tid + q * NT : 0];
#pragma unroll
        for (int q = 0; q < TW; q++)
            if (tid + q * NT < L) tws[tid + q * NT] = tv[q];
    }
    __syncthreads();  // twiddles visible
    dft_reg<T, A, S>(v);
    // w_L^(n2 k1) for every k1 > 0 (tws[0] = 1 exactly, so n2 = 0 needs no
    // branch and the table reads go out together)
#pragma unroll
    for (int k1 = 1; k1 < A; k1++) {
        cpx<T> w = tws[n2 * k1];
        if (S > 0) w.y = -w.y;
        v[k1] = cmul(v[k1], w);
    }
#pragma unroll
    for (int k1 = 0; k1 < A; k1++) lds[k1 * RS + n2 * W + p] = v[k1].x;
    __syncthreads();
    const bool ph2 = tid < W * A;
    const int q1 = tid / W;
    cpx<T> u[B];
    if (ph2) {
#pragma unroll
        for (int m = 0; m < B; m++) u[m].x = lds[q1 * RS + m * W + p];
    }
    __syncthreads();
#pragma unroll
    for (int k1 = 0; k1 < A; k1++) lds[k1 * RS + n2 * W + p] = v[k1].y;
    __syncthreads();
    if (ph2) {
#pragma unroll
        for (int m = 0; m < B; m++) u[m].y = lds[q1 * RS + m * W + p];
        dft_reg<T, B, S>(u);
        if (pin) cols2_store<T, A, B>(u, out, bo, go.ts, q1, p, nkeep);
    }
}

// ---- half-length column pass (F2 / I2 of a zero-padded axis) -------------
// L = 2H with the data on one side of the padding: the forward transform of
// an input with rows t < nin <= H, or the inverse transform whose output rows
// t < nkeep <= H are kept.  The length-L transform splits into two length-H
// transforms (even and odd frequencies):
//   forward: X[2k] = DFT_H(x)[k],  X[2k+1] = DFT_H(x[n] w_L^n)[k]
//   inverse: x[n] = IDFT_H(X[2k])[n] + w_L^-n IDFT_H(X[2k+1])[n]
// each done two-level (H = A B) through an LDS tile of H x W points, so a
// block holds half the LDS of the length-L pass.  tw: w_L^j, j < L
// (w_H^j = w_L^2j).  Threads: W * B (phase 1), W * A (phase 2).
template <typename T, int L, int S, int W, int NT = W * split2(L / 2).B>
__global__ void __launch_bounds__(NT) k_cols2h(const cpx<T>* in, cpx<T>* out, Geo gi, Geo go, int nin, int nkeep,
                                               int ncol, const cpx<T>* __restrict__ tw, Live lv) {
    constexpr int H = L / 2, A = split2(H).A, B = split2(H).B;
    static_assert(A * B == H && NT == W * B && A <= B, "two-level split");
    __shared__ cpx<T> lds[H * W];
    __shared__ cpx<T> tws[L];
    const int tid = threadIdx.x;
    const int lb = xcd_swizzle(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int a = lb / gridDim.x;
    if (all_dead(lv, a, a)) return;
    const int col0 = (lb % gridDim.x) * W;
    const long bi = geo_base(gi, a) + col0;
    const long bo = geo_base(go, a) + col0;
    for (int e = tid; e < L; e += NT) tws[e] = tw[e];
    const int p = tid % W, n2 = tid / W;
    const bool pin = col0 + p < ncol;
    auto twl = [&](int j) {  // w_L^j in the transform's direction
        cpx<T> w = tws[j];
        if (S > 0) w.y = -w.y;
        return w;
    };
    // phase 1 inputs: forward x[B n1 + n2]; inverse X[2(B n1 + n2)] and X[2(B n1 + n2) + 1]
    cpx<T> ve[A], vo[A];
#pragma unroll
    for (int n1 = 0; n1 < A; n1++) {
        const int t = B * n1 + n2;
        cpx<T> x0 = {T(0), T(0)}, x1 = {T(0), T(0)};
        if (S < 0) {
            if (pin && t < nin) x0 = in[bi + (long)t * gi.ts + p];
        } else {
            if (pin && 2 * t < nin) x0 = in[bi + (long)(2 * t) * gi.ts + p];
            if (pin && 2 * t + 1 < nin) x1 = in[bi + (long)(2 * t + 1) * gi.ts + p];
        }
        ve[n1] = x0;
        vo[n1] = x1;
    }
    __syncthreads();  // twiddles visible
    if (S < 0) {
#pragma unroll
        for (int n1 = 0; n1 < A; n1++) vo[n1] = (B * n1 + n2) ? cmul(ve[n1], twl(B * n1 + n2)) : ve[n1];
    }
    cpx<T> acc[B];  // inverse: the even half of the output rows
    const int k1 = tid / W;
#pragma unroll
    for (int half = 0; half < 2; half++) {
        cpx<T>* v = half ? vo : ve;
        dft_reg<T, A, S>(v);
        if (half) __syncthreads();  // the previous half's phase 2 has read the tile
#pragma unroll
        for (int q = 0; q < A; q++) lds[(q * B + n2) * W + p] = (n2 * q) ? cmul(v[q], twl(2 * n2 * q)) : v[q];
        __syncthreads();
        if (tid < W * A) {
            cpx<T> u[B];
#pragma unroll
            for (int m = 0; m < B; m++) u[m] = lds[(k1 * B + m) * W + p];
            dft_reg<T, B, S>(u);
            if (S < 0) {  // forward: frequencies 2k + half
                if (pin) {
#pragma unroll
                    for (int k2 = 0; k2 < B; k2++) {
                        const int t = 2 * (k1 + A * k2) + half;
                        if (t < nkeep) out[bo + (long)t * go.ts + p] = u[k2];
                    }
                }
            } else if (half == 0) {
#pragma unroll
                for (int k2 = 0; k2 < B; k2++) acc[k2] = u[k2];
            } else if (pin) {  // inverse: x[n] = even[n] + w_L^-n odd[n], n < nkeep
#pragma unroll
                for (int k2 = 0; k2 < B; k2++) {
                    const int t = k1 + A * k2;
                    const cpx<T> o = t ? cmul(u[k2], twl(t)) : u[k2];
                    if (t < nkeep) out[bo + (long)t * go.ts + p] = cadd(acc[k2], o);
                }
            }
        }
    }
}

// ---- row passes (F1 R2C / I1 C2R) -----------------------------------------
// M = P2/2-point complex FFT per row as in demag_kernels.hpp's k_r2c_rows /
// k_c2r_rows, done two-level.  W rows per block; phase-1 threads take
// (n2, row) with n2 fastest so a wave reads consecutive pairs of a row as
// 8-byte words.  Exchange layout [row][k1][n2] with row pitch B + 1 (odd
// stride in 8-byte words: no bank conflicts in either phase).  Rows of even
// length n2 move as aligned pairs (one 8-byte word per complex element);
// odd rows fall back to single words.
template <int M, int W>
constexpr int rows2_lds() {
    constexpr int A = split2(M).A, B = split2(M).B;
    return (W * A * (B + 1) > W * (M + 1)) ? W * A * (B + 1) : W * (M + 1);
}

template <typename T, int M, int W, int NT = W * split2(M).B>
__global__ void __launch_bounds__(NT) k_r2c_rows2(const T* __restrict__ x, const T* __restrict__ msk, long mrows,
                                                  cpx<T>* __restrict__ out, long nrows, int n2, int H2p,
                                                  const cpx<T>* __restrict__ twM, const cpx<T>* __restrict__ twP,
                                                  Live lv) {
    constexpr int A = split2(M).A, B = split2(M).B, BP = B + 1;
    static_assert(A * B == M && NT == W * B && A <= B, "two-level split");
    __shared__ cpx<T> lds[rows2_lds<M, W>()];
    __shared__ cpx<T> tws[M + 1];  // twM for phase 1, then twP[0..M] for the split
    const int tid = threadIdx.x;
    const long row0 = (long)blockIdx.x * W;
    if (all_dead(lv, row0, (row0 + W <= nrows ? row0 + W : nrows) - 1)) return;
    // Every global load of the block is issued before the first wait: the
    // twiddles of both tables into registers, the row data from clamped
    // (always valid) addresses with the out-of-range lanes zeroed after.
    constexpr int TA = (M + NT - 1) / NT, TB = (M + 1 + NT - 1) / NT;
    cpx<T> ta[TA], tb[TB];
#pragma unroll
    for (int q = 0; q < TA; q++) ta[q] = twM[tid + q * NT < M ? tid + q * NT : 0];
#pragma unroll
    for (int q = 0; q < TB; q++) tb[q] = twP[tid + q * NT <= M ? tid + q * NT : 0];
    {  // phase 1: thread (n2, row p) gathers z[B n1 + n2] = x[2n] + i x[2n+1]
        const int nb = tid % B, p = tid / B;
        const long r = row0 + p;
        const bool rin = r < nrows;
        const long rc = rin ? r : row0;  // a valid row for the lanes past the end
        const T* xr = x + rc * n2;
        const T* mr = msk ? msk + (rc % mrows) * n2 : nullptr;
        cpx<T> v[A];
        if ((n2 & 1) == 0) {  // uniform: pairs of reals as one 8- or 16-byte word
            using V2 = typename std::conditional<sizeof(T) == 4, float2, double2>::type;
            V2 q[A];
#pragma unroll
            for (int n1 = 0; n1 < A; n1++) {
                const int n = B * n1 + nb;
                q[n1] = *(const V2*)(xr + 2 * (2 * n < n2 ? n : 0));
            }
            if (mr) {  // uniform
#pragma unroll
                for (int n1 = 0; n1 < A; n1++) {
                    const int n = B * n1 + nb;
                    const V2 w = *(const V2*)(mr + 2 * (2 * n < n2 ? n : 0));
                    q[n1].x *= w.x;
                    q[n1].y *= w.y;
                }
            }
#pragma unroll
            for (int n1 = 0; n1 < A; n1++) {
                const int n = B * n1 + nb;
                const bool ok = rin && 2 * n < n2;
                v[n1] = {ok ? T(q[n1].x) : T(0), ok ? T(q[n1].y) : T(0)};
            }
        } else {
#pragma unroll
            for (int n1 = 0; n1 < A; n1++) {
                const int n = B * n1 + nb;
                cpx<T> z = {T(0), T(0)};
                if (rin) {
                    if (2 * n < n2) z.x = xr[2 * n] * (mr ? mr[2 * n] : T(1));
                    if (2 * n + 1 < n2) z.y = xr[2 * n + 1] * (mr ? mr[2 * n + 1] : T(1));
                }
                v[n1] = z;
            }
        }
#pragma unroll
        for (int q = 0; q < TA; q++)
            if (tid + q * NT < M) tws[tid + q * NT] = ta[q];
        __syncthreads();  // twiddles visible
        dft_reg<T, A, -1>(v);
#pragma unroll
        for (int k1 = 0; k1 < A; k1++) lds[(p * A + k1) * BP + nb] = (nb * k1) ? cmul(v[k1], tws[nb * k1]) : v[k1];
    }
    __syncthreads();  // phase 1 done with twM: the table takes twP for the split
#pragma unroll
    for (int q = 0; q < TB; q++)
        if (tid + q * NT <= M) tws[tid + q * NT] = tb[q];
    cpx<T> u[B];
    const int k1 = tid % A, p2 = tid / A;
    if (tid < W * A) {  // phase 2: thread (k1, row) gathers Y[k1][n2]
#pragma unroll
        for (int m = 0; m < B; m++) u[m] = lds[(p2 * A + k1) * BP + m];
        dft_reg<T, B, -1>(u);
    }
    __syncthreads();
    if (tid < W * A) {
#pragma unroll
        for (int k2 = 0; k2 < B; k2++) lds[p2 * M + k1 + A * k2] = u[k2];
    }
    __syncthreads();
    // split: X[k] = E + w^k O, k = 0..M; row p's k < M by threads k = tid +
    // q NT (rows in order), then the M-th points of the W rows
    auto split = [&](int p, int k) {
        cpx<T> zk = lds[p * M + (k % M)];
        cpx<T> zm = lds[p * M + ((M - k) % M)];
        cpx<T> E = {T(0.5) * (zk.x + zm.x), T(0.5) * (zk.y - zm.y)};
        cpx<T> O = {T(0.5) * (zk.y + zm.y), T(-0.5) * (zk.x - zm.x)};
        out[(row0 + p) * H2p + k] = cadd(E, cmul(tws[k], O));
    };
    if (M % NT == 0) {
#pragma unroll
        for (int p = 0; p < W; p++)
            if (row0 + p < nrows) {  // uniform
#pragma unroll
                for (int q = 0; q < M / NT; q++) split(p, tid + q * NT);
            }
        if (tid < W && row0 + tid < nrows) split(tid, M);
    } else {
        for (int e = tid; e < (M + 1) * W; e += NT) {
            const int p = e / (M + 1), k = e % (M + 1);
            if (row0 + p < nrows) split(p, k);
        }
    }
}

template <typename T, int M, int W, int MINW = 1, int NT = W * split2(M).B>
__global__ void __launch_bounds__(NT, MINW) k_c2r_rows2(const cpx<T>* __restrict__ in, T* __restrict__ x, long nrows,
                                                  int n2, int H2p, const cpx<T>* __restrict__ twM,
                                                  const cpx<T>* __restrict__ twP, Live lv) {
    constexpr int A = split2(M).A, B = split2(M).B, BP = B + 1;
    static_assert(A * B == M && NT == W * B && A <= B, "two-level split");
    __shared__ cpx<T> lds[rows2_lds<M, W>()];
    __shared__ cpx<T> tws[M];
    __shared__ cpx<T> twp[M];
    const int tid = threadIdx.x;
    const long row0 = (long)blockIdx.x * W;
    if (all_dead(lv, row0, (row0 + W <= nrows ? row0 + W : nrows) - 1)) return;
    if constexpr (M % NT == 0) {
        // row by row: thread k = tid + q NT of each row (the row's start is
        // uniform), then the M-th points of the W rows; rows past nrows read
        // row row0 and are zeroed.  Every global load is in flight before the
        // first wait; the twiddles of both tables go through registers.
        constexpr int QR = M / NT;
        cpx<T> g[W][QR], gm, ta[QR], tb[QR];
#pragma unroll
        for (int p = 0; p < W; p++) {
            const cpx<T>* rp = in + (row0 + p < nrows ? row0 + p : row0) * H2p;
#pragma unroll
            for (int q = 0; q < QR; q++) g[p][q] = rp[tid + q * NT];
        }
        const bool mrow = tid < W && row0 + tid < nrows;
        gm = in[(mrow ? row0 + tid : row0) * H2p + M];
#pragma unroll
        for (int q = 0; q < QR; q++) {
            ta[q] = twM[tid + q * NT];
            tb[q] = twP[tid + q * NT];
        }
#pragma unroll
        for (int p = 0; p < W; p++) {
            const bool rin = row0 + p < nrows;  // uniform
#pragma unroll
            for (int q = 0; q < QR; q++) lds[p * (M + 1) + tid + q * NT] = rin ? g[p][q] : cpx<T>{T(0), T(0)};
        }
        if (tid < W) lds[tid * (M + 1) + M] = mrow ? gm : cpx<T>{T(0), T(0)};
#pragma unroll
        for (int q = 0; q < QR; q++) {
            tws[tid + q * NT] = ta[q];
            twp[tid + q * NT] = tb[q];
        }
    } else {  // every global load in flight before the first wait: spectrum rows in
       // equal batches of at most 24 per thread from clamped (valid) addresses,
       // the twiddles of both tables into registers
        constexpr int NE = (M + 1) * W, IT = (NE + NT - 1) / NT, NB = (IT + 23) / 24, CH = (IT + NB - 1) / NB;
        constexpr int TA = (M + NT - 1) / NT;
        cpx<T> ta[TA], tb[TA];
        for (int q0 = 0; q0 < IT; q0 += CH) {
            cpx<T> g[CH];
#pragma unroll
            for (int q = 0; q < CH; q++) {
                const int e = tid + (q0 + q) * NT;
                const int p = e / (M + 1), k = e % (M + 1);
                const long r = row0 + p;
                const bool ok = e < NE && r < nrows;
                g[q] = in[ok ? r * H2p + k : row0 * H2p];
                if (!ok) g[q] = {T(0), T(0)};
            }
            if (q0 == 0) {
#pragma unroll
                for (int q = 0; q < TA; q++) {
                    const int e = tid + q * NT;
                    ta[q] = twM[e < M ? e : 0];
                    tb[q] = twP[e < M ? e : 0];
                }
            }
#pragma unroll
            for (int q = 0; q < CH; q++) {
                const int e = tid + (q0 + q) * NT;
                if (e < NE) lds[e] = g[q];  // e = p (M + 1) + k
            }
        }
#pragma unroll
        for (int q = 0; q < TA; q++) {
            const int e = tid + q * NT;
            if (e < M) {
                tws[e] = ta[q];
                twp[e] = tb[q];
            }
        }
    }
    __syncthreads();
    const int nb = tid % B, p = tid / B;
    cpx<T> v[A];
    {  // phase 1: Z[k] = (X[k] + conj X[M-k]) + i w^-k (X[k] - conj X[M-k]), k = B n1 + n2
#pragma unroll
        for (int n1 = 0; n1 < A; n1++) {
            const int k = B * n1 + nb;
            cpx<T> a = lds[p * (M + 1) + k];
            cpx<T> b = lds[p * (M + 1) + (M - k)];
            cpx<T> s = {a.x + b.x, a.y - b.y};
            cpx<T> d = {a.x - b.x, a.y + b.y};
            cpx<T> w = twp[k];
            w.y = -w.y;
            cpx<T> wd = cmul(w, d);
            v[n1] = {s.x - wd.y, s.y + wd.x};
        }
    }
    __syncthreads();  // X no longer read: the buffer takes the exchange
    dft_reg<T, A, +1>(v);
#pragma unroll
    for (int k1 = 0; k1 < A; k1++) {
        cpx<T> w = tws[nb * k1];
        w.y = -w.y;
        lds[(p * A + k1) * BP + nb] = (nb * k1) ? cmul(v[k1], w) : v[k1];
    }
    __syncthreads();
    if (tid < W * A) {  // phase 2 and the real output pairs x[2n], x[2n+1], n = k1 + A k2
        const int k1 = tid % A, p2 = tid / A;
        const long r = row0 + p2;
        cpx<T> u[B];
#pragma unroll
        for (int m = 0; m < B; m++) u[m] = lds[(p2 * A + k1) * BP + m];
        dft_reg<T, B, +1>(u);
        if (r < nrows && n2 == M && B % 2 == 0) {
            // zero-padded x axis (P2 = 2 n2): the pairs n = k1 + A k2 < n2 / 2 are
            // exactly k2 < B / 2
            using V2 = typename std::conditional<sizeof(T) == 4, float2, double2>::type;
            T* xr = x + r * n2 + 2 * k1;
#pragma unroll
            for (int k2 = 0; k2 < B / 2; k2++) {
                V2 q;
                q.x = u[k2].x;
                q.y = u[k2].y;
                *(V2*)(xr + 2 * A * k2) = q;
            }
        } else if (r < nrows) {
            T* xr = x + r * n2;
#pragma unroll
            for (int k2 = 0; k2 < B; k2++) {
                const int n = k1 + A * k2;
                if ((n2 & 1) == 0) {
                    if (2 * n < n2) {
                        using V2 = typename std::conditional<sizeof(T) == 4, float2, double2>::type;
                        V2 q;
                        q.x = u[k2].x;
                        q.y = u[k2].y;
                        *(V2*)(xr + 2 * n) = q;
                    }
                } else {
                    if (2 * n < n2) xr[2 * n] = u[k2].x;
                    if (2 * n + 1 < n2) xr[2 * n + 1] = u[k2].y;
                }
            }
        }
    }
}

// ---- axis-0 convolution pass (F3K) with register FFTs ---------------------
// Same contract as demag_kernels.hpp's k_zconv (pencil j' and its mirror,
// W columns, 3 components), with one thread per pencil for the transforms:
// the thread loads the pencil's N0 non-zero values, does the L-point DFT in
// registers and posts the spectrum to LDS; the 3x3 kernel product runs over
// the LDS points as in k_zconv; each thread then takes its pencil back,
// transforms it in registers and stores its N0 rows.  NT = 6 W threads,
// tid = (s * 3 + c) * W + p, LDS layout [t][s][c][p] as in k_zconv.
template <typename T, typename TK, int L, int W>
__global__ void __launch_bounds__(6 * W) k_zconv2(cpx<T>* X, const TK* __restrict__ kern, int N0, int P1, int H2,
                                                  int H2p, int K1, int kp, T srcMul, const cpx<T>* __restrict__ tw,
                                                  Live lv) {
    constexpr int WP = 3 * W, W2 = 2 * WP, NT = 6 * W;
    static_assert(reg_dft_ok(L), "register DFT length");
    if (all_dead(lv, blockIdx.z, blockIdx.z)) return;
    X += (long)blockIdx.z * 3 * N0 * P1 * H2p;
    __shared__ cpx<T> lds[L * W2];
    const int tid = threadIdx.x;
    const int lb = xcd_swizzle(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int jp = lb / gridDim.x;
    const int jm = (jp == 0 || 2 * jp == P1) ? -1 : P1 - jp;
    const int col0 = (lb % gridDim.x) * W;
    {
        const int p = tid % W, c = (tid / W) % 3, sm = tid / WP;
        const int j = sm ? jm : jp;
        const bool live = j >= 0 && col0 + p < H2;
        const long cstride = (long)N0 * P1 * H2p, tstride = (long)P1 * H2p;
        cpx<T>* base = X + c * cstride + (long)(j < 0 ? 0 : j) * H2p + col0 + p;
        cpx<T> v[L];
#pragma unroll
        for (int t = 0; t < L; t++) {
            cpx<T> x = {T(0), T(0)};
            if (live && t < N0) x = base[t * tstride];
            v[t] = x;
        }
        dft_reg<T, L, -1>(v);
#pragma unroll
        for (int t = 0; t < L; t++) lds[t * W2 + tid] = v[t];
    }
    __syncthreads();
    const long kplane = (long)(L / 2 + 1) * K1 * kp;
    for (int e = tid; e < L * W; e += NT) {
        const int t = e / W, p = e % W;
        const int col = col0 + p;
        if (col >= H2) continue;
        const int ii = (t > L / 2) ? L - t : t;
        const T si = (t > L / 2) ? T(-1) : T(1);
        const long K = ((long)ii * K1 + jp) * kp + col;
        const T kxx = T(kern[K]), kyy = T(kern[kplane + K]), kzz = T(kern[2 * kplane + K]);
        const T kyz0 = T(kern[3 * kplane + K]);
        const T kxz = si * T(kern[4 * kplane + K]);
        const T kxy0 = si * T(kern[5 * kplane + K]);
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
    {
        const int p = tid % W, c = (tid / W) % 3, sm = tid / WP;
        const int j = sm ? jm : jp;
        const bool live = j >= 0 && col0 + p < H2;
        const long cstride = (long)N0 * P1 * H2p, tstride = (long)P1 * H2p;
        cpx<T>* base = X + c * cstride + (long)(j < 0 ? 0 : j) * H2p + col0 + p;
        cpx<T> v[L];
#pragma unroll
        for (int t = 0; t < L; t++) v[t] = lds[t * W2 + tid];
        dft_reg<T, L, +1>(v);
        if (live) {
#pragma unroll
            for (int t = 0; t < L; t++)
                if (t < N0) base[t * tstride] = v[t];
        }
    }
}

// The axis-0 pass with the kernel product done in two real halves: K is real,
// so the real and the imaginary parts of the spectrum go through the 3x3
// product separately, each through an L x 6W array of T (half the LDS of
// k_zconv2, more blocks per CU).
template <typename T, typename TK, int L, int W>
__device__ __forceinline__ void zconv_half(T* lds, const TK* __restrict__ kern, int jp, int col0, int H2, int K1,
                                           int kp, T srcMul) {
    constexpr int WP = 3 * W, W2 = 2 * WP, NT = 6 * W;
    const long kplane = (long)(L / 2 + 1) * K1 * kp;
    for (int e = threadIdx.x; e < L * W; e += NT) {
        const int t = e / W, p = e % W;
        const int col = col0 + p;
        if (col >= H2) continue;
        const int ii = (t > L / 2) ? L - t : t;
        const T si = (t > L / 2) ? T(-1) : T(1);
        const long K = ((long)ii * K1 + jp) * kp + col;
        const T kxx = T(kern[K]), kyy = T(kern[kplane + K]), kzz = T(kern[2 * kplane + K]);
        const T kyz0 = T(kern[3 * kplane + K]);
        const T kxz = si * T(kern[4 * kplane + K]);
        const T kxy0 = si * T(kern[5 * kplane + K]);
#pragma unroll
        for (int sm = 0; sm < 2; sm++) {  // mirror pencil: YZ and XY change sign
            const T sj = sm ? T(-1) : T(1);
            const T kyz = sj * kyz0, kxy = sj * kxy0;
            T* a = lds + t * W2 + sm * WP + p;
            const T mx = a[0] * srcMul, my = a[W] * srcMul, mz = a[2 * W] * srcMul;
            a[0] = kxx * mx + kxy * my + kxz * mz;
            a[W] = kxy * mx + kyy * my + kyz * mz;
            a[2 * W] = kxz * mx + kyz * my + kzz * mz;
        }
    }
}
template <typename T, typename TK, int L, int W, int MINW = 1>
__global__ void __launch_bounds__(6 * W, MINW) k_zconv2s(cpx<T>* X, const TK* __restrict__ kern, int N0, int P1, int H2,
                                                   int H2p, int K1, int kp, T srcMul, const cpx<T>* __restrict__ tw,
                                                   Live lv) {
    constexpr int WP = 3 * W, W2 = 2 * WP;
    static_assert(reg_dft_ok(L), "register DFT length");
    if (all_dead(lv, blockIdx.z, blockIdx.z)) return;
    X += (long)blockIdx.z * 3 * N0 * P1 * H2p;
    __shared__ T lds[L * W2];
    const int tid = threadIdx.x;
    const int lb = xcd_swizzle(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int jp = lb / gridDim.x;
    const int jm = (jp == 0 || 2 * jp == P1) ? -1 : P1 - jp;
    const int col0 = (lb % gridDim.x) * W;
    const int p = tid % W, c = (tid / W) % 3, sm = tid / WP;
    const int j = sm ? jm : jp;
    const bool live = j >= 0 && col0 + p < H2;
    const long cstride = (long)N0 * P1 * H2p, tstride = (long)P1 * H2p;
    cpx<T>* base = X + c * cstride + (long)(j < 0 ? 0 : j) * H2p + col0 + p;
    cpx<T> v[L];
#pragma unroll
    for (int t = 0; t < L; t++) {
        cpx<T> x = {T(0), T(0)};
        if (live && t < N0) x = base[t * tstride];
        v[t] = x;
    }
    dft_reg<T, L, -1>(v);
#pragma unroll
    for (int t = 0; t < L; t++) lds[t * W2 + tid] = v[t].x;
    __syncthreads();
    zconv_half<T, TK, L, W>(lds, kern, jp, col0, H2, K1, kp, srcMul);
    __syncthreads();
#pragma unroll
    for (int t = 0; t < L; t++) v[t].x = lds[t * W2 + tid];
    __syncthreads();
#pragma unroll
    for (int t = 0; t < L; t++) lds[t * W2 + tid] = v[t].y;
    __syncthreads();
    zconv_half<T, TK, L, W>(lds, kern, jp, col0, H2, K1, kp, srcMul);
    __syncthreads();
#pragma unroll
    for (int t = 0; t < L; t++) v[t].y = lds[t * W2 + tid];
    dft_reg<T, L, +1>(v);
    if (live) {
#pragma unroll
        for (int t = 0; t < L; t++)
            if (t < N0) base[t * tstride] = v[t];
    }
}

// k_zconv2 with non-temporal pencil loads and stores: the spectrum is read
// once and written once per convolution, so its lines need not stay in L2
// (the kernel octant, reused by the mirror pencil and the t, L - t pair,
// keeps the normal policy).
template <typename T, typename TK, int L, int W>
__global__ void __launch_bounds__(6 * W) k_zconv2n(cpx<T>* X, const TK* __restrict__ kern, int N0, int P1, int H2,
                                                   int H2p, int K1, int kp, T srcMul, const cpx<T>* __restrict__ tw,
                                                   Live lv) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    constexpr int WP = 3 * W, W2 = 2 * WP, NT = 6 * W;
    static_assert(reg_dft_ok(L), "register DFT length");
    if (all_dead(lv, blockIdx.z, blockIdx.z)) return;
    X += (long)blockIdx.z * 3 * N0 * P1 * H2p;
    __shared__ cpx<T> lds[L * W2];
    const int tid = threadIdx.x;
    const int lb = xcd_swizzle(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int jp = lb / gridDim.x;
    const int jm = (jp == 0 || 2 * jp == P1) ? -1 : P1 - jp;
    const int col0 = (lb % gridDim.x) * W;
    const int p = tid % W, c = (tid / W) % 3, sm = tid / WP;
    const int j = sm ? jm : jp;
    const bool live = j >= 0 && col0 + p < H2;
    const long cstride = (long)N0 * P1 * H2p, tstride = (long)P1 * H2p;
    cpx<T>* base = X + c * cstride + (long)(j < 0 ? 0 : j) * H2p + col0 + p;
    {
        cpx<T> v[L];
#pragma unroll
        for (int t = 0; t < L; t++) {
            cpx<T> x = {T(0), T(0)};
            if (live && t < N0) {
                const v2 q = __builtin_nontemporal_load((const v2*)(base + t * tstride));
                x = {q.x, q.y};
            }
            v[t] = x;
        }
        dft_reg<T, L, -1>(v);
#pragma unroll
        for (int t = 0; t < L; t++) lds[t * W2 + tid] = v[t];
    }
    __syncthreads();
    // the product entries of this thread, e = tid + q NT, unrolled with
    // unconditional kernel loads (clamped to a valid column) so that all of
    // them are in flight together instead of one iteration at a time
    const long kplane = (long)(L / 2 + 1) * K1 * kp;
    constexpr int NE = L * W, IT = (NE + NT - 1) / NT;
    T kv[IT][6];
#pragma unroll
    for (int q = 0; q < IT; q++) {
        const int e = tid + q * NT;
        const int t = e < NE ? e / W : 0, pp = e % W;
        const bool ok = e < NE && col0 + pp < H2;
        const int ii = (t > L / 2) ? L - t : t;
        const long K = ((long)ii * K1 + jp) * kp + (ok ? col0 + pp : col0);
#pragma unroll
        for (int k = 0; k < 6; k++) kv[q][k] = T(kern[k * kplane + K]);
    }
#pragma unroll
    for (int q = 0; q < IT; q++) {
        const int e = tid + q * NT;
        const int t = e < NE ? e / W : 0, pp = e % W;
        const bool ok = e < NE && col0 + pp < H2;
        const T si = (t > L / 2) ? T(-1) : T(1);
        const T kxx = kv[q][0], kyy = kv[q][1], kzz = kv[q][2], kyz0 = kv[q][3];
        const T kxz = si * kv[q][4], kxy0 = si * kv[q][5];
        if (!ok) continue;
#pragma unroll
        for (int s2 = 0; s2 < 2; s2++) {
            const T sj = s2 ? T(-1) : T(1);
            const T kyz = sj * kyz0, kxy = sj * kxy0;
            cpx<T>* a = lds + t * W2 + s2 * WP + pp;
            cpx<T> mx = cscale(a[0], srcMul);
            cpx<T> my = cscale(a[W], srcMul);
            cpx<T> mz = cscale(a[2 * W], srcMul);
            a[0] = {kxx * mx.x + kxy * my.x + kxz * mz.x, kxx * mx.y + kxy * my.y + kxz * mz.y};
            a[W] = {kxy * mx.x + kyy * my.x + kyz * mz.x, kxy * mx.y + kyy * my.y + kyz * mz.y};
            a[2 * W] = {kxz * mx.x + kyz * my.x + kzz * mz.x, kxz * mx.y + kyz * my.y + kzz * mz.y};
        }
    }
    __syncthreads();
    {
        cpx<T> v[L];
#pragma unroll
        for (int t = 0; t < L; t++) v[t] = lds[t * W2 + tid];
        dft_reg<T, L, +1>(v);
        if (live) {
#pragma unroll
            for (int t = 0; t < L; t++)
                if (t < N0) {
                    v2 q;
                    q.x = v[t].x;
                    q.y = v[t].y;
                    __builtin_nontemporal_store(q, (v2*)(base + t * tstride));
                }
        }
    }
}

// Axis-0 pass with Q = 2 or 4 lanes per pencil, L = Q H with H <= 32 (for
// L > 32 the L-point pencil in one lane would hold 2L floats in registers;
// with two lanes the block is 12 W lanes, whole waves for W = 16).  Lane r
// (r < Q) loads the rows t = Q m + r of its pencil and does the H-point DFT
// of them in registers, E_r.  The product phase forms the spectrum
//     X[k + qH] = sum_r w_Q^(rq) (w^(rk) E_r[k]),   w = w_L
// (a Q-point DFT of the twiddled E_r), applies the 3x3 kernel at the Q
// points t = k + qH, and writes back
//     G_r[k] = w^(-rk) sum_q w_Q^(-rq) Y[k + qH],
// so that lane r's inverse H-point DFT of G_r gives the rows t = Q m + r.
// NT = 6 Q W threads, tid = r * 6W + (s * 3 + c) * W + p; LDS [t][s][c][p]
// as in k_zconv2 (lane r posts E_r at t = r H + k).  tw holds w_L^t, t < L.
template <typename T, typename TK, int L, int W, int Q>
__global__ void __launch_bounds__(6 * Q * W) k_zconv2m(cpx<T>* X, const TK* __restrict__ kern, int N0, int P1, int H2,
                                                        int H2p, int K1, int kp, T srcMul,
                                                        const cpx<T>* __restrict__ tw, Live lv) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    constexpr int H = L / Q, WP = 3 * W, W2 = 2 * WP, NT = 6 * Q * W;
    static_assert((Q == 2 || Q == 4) && L % Q == 0 && reg_dft_ok(H), "lanes per pencil, sub-DFT length");
    if (all_dead(lv, blockIdx.z, blockIdx.z)) return;
    X += (long)blockIdx.z * 3 * N0 * P1 * H2p;
    __shared__ cpx<T> lds[L * W2];
    const int tid = threadIdx.x;
    const int r = tid / W2, u = tid % W2;
    const int lb = xcd_swizzle(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int jp = lb / gridDim.x;
    const int jm = (jp == 0 || 2 * jp == P1) ? -1 : P1 - jp;
    const int col0 = (lb % gridDim.x) * W;
    const int p = u % W, c = (u / W) % 3, sm = u / WP;
    const int j = sm ? jm : jp;
    const bool live = j >= 0 && col0 + p < H2;
    const long cstride = (long)N0 * P1 * H2p, tstride = (long)P1 * H2p;
    // row r of the pencil (row 0 when N0 <= r leaves no row r: never read then)
    cpx<T>* base = X + c * cstride + (long)(j < 0 ? 0 : j) * H2p + col0 + p + (r < N0 ? r : 0) * tstride;
    {
        cpx<T> v[H];
#pragma unroll
        for (int m = 0; m < H; m++) {
            cpx<T> x = {T(0), T(0)};
            if (live && Q * m + r < N0) {
                const v2 q = __builtin_nontemporal_load((const v2*)(base + Q * m * tstride));
                x = {q.x, q.y};
            }
            v[m] = x;
        }
        dft_reg<T, H, -1>(v);
#pragma unroll
        for (int k = 0; k < H; k++) lds[(r * H + k) * W2 + u] = v[k];
    }
    __syncthreads();
    // product entries e = tid + it NT over (k < H, column p); the kernel values
    // of the Q points are loaded for all of them first (clamped to a valid
    // column) so that they are in flight together
    const long kplane = (long)(L / 2 + 1) * K1 * kp;
    constexpr int NE = H * W, IT = (NE + NT - 1) / NT;
    T kv[IT][Q][6];
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int e = tid + it * NT;
        const int k = e < NE ? e / W : 0, pp = e % W;
        const bool ok = e < NE && col0 + pp < H2;
        const long colk = ok ? col0 + pp : col0;
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const int t = k + q * H;
            const int ii = t <= L / 2 ? t : L - t;  // octant index (t > L/2 is mirrored)
            const long K = ((long)ii * K1 + jp) * kp + colk;
#pragma unroll
            for (int n = 0; n < 6; n++) kv[it][q][n] = T(kern[n * kplane + K]);
        }
    }
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int e = tid + it * NT;
        const int k = e < NE ? e / W : 0, pp = e % W;
        const bool ok = e < NE && col0 + pp < H2;
        if (!ok) continue;
        cpx<T> w[Q];  // w^(rk)
        w[0] = {T(1), T(0)};
#pragma unroll
        for (int rr = 1; rr < Q; rr++) w[rr] = tw[rr * k];
#pragma unroll
        for (int s2 = 0; s2 < 2; s2++) {
            const T sj = s2 ? T(-1) : T(1);
            cpx<T> xq[3][Q];
#pragma unroll
            for (int cc = 0; cc < 3; cc++) {
#pragma unroll
                for (int rr = 0; rr < Q; rr++) {
                    const cpx<T> a = cscale(lds[(rr * H + k) * W2 + s2 * WP + cc * W + pp], srcMul);
                    xq[cc][rr] = rr ? cmul(w[rr], a) : a;
                }
                dft_reg<T, Q, -1>(xq[cc]);  // xq[cc][q] = X[k + qH]
            }
            cpx<T> yq[3][Q];
#pragma unroll
            for (int q = 0; q < Q; q++) {
                const int t = k + q * H;
                const T si = t > L / 2 ? T(-1) : T(1);  // mirrored t: XZ and XY change sign
                const T kxx = kv[it][q][0], kyy = kv[it][q][1], kzz = kv[it][q][2];
                const T kyz = sj * kv[it][q][3], kxz = si * kv[it][q][4], kxy = sj * si * kv[it][q][5];
                const cpx<T> x0 = xq[0][q], x1 = xq[1][q], x2 = xq[2][q];
                yq[0][q] = {kxx * x0.x + kxy * x1.x + kxz * x2.x, kxx * x0.y + kxy * x1.y + kxz * x2.y};
                yq[1][q] = {kxy * x0.x + kyy * x1.x + kyz * x2.x, kxy * x0.y + kyy * x1.y + kyz * x2.y};
                yq[2][q] = {kxz * x0.x + kyz * x1.x + kzz * x2.x, kxz * x0.y + kyz * x1.y + kzz * x2.y};
            }
#pragma unroll
            for (int cc = 0; cc < 3; cc++) {
                dft_reg<T, Q, +1>(yq[cc]);  // yq[cc][r] = sum_q w_Q^(-rq) Y[k + qH]
#pragma unroll
                for (int rr = 0; rr < Q; rr++) {
                    const cpx<T> wc = {w[rr].x, -w[rr].y};
                    lds[(rr * H + k) * W2 + s2 * WP + cc * W + pp] = rr ? cmul(wc, yq[cc][rr]) : yq[cc][rr];
                }
            }
        }
    }
    __syncthreads();
    {
        cpx<T> v[H];
#pragma unroll
        for (int k = 0; k < H; k++) v[k] = lds[(r * H + k) * W2 + u];
        dft_reg<T, H, +1>(v);
        if (live) {
#pragma unroll
            for (int m = 0; m < H; m++)
                if (Q * m + r < N0) {
                    v2 q;
                    q.x = v[m].x;
                    q.y = v[m].y;
                    __builtin_nontemporal_store(q, (v2*)(base + Q * m * tstride));
                }
        }
    }
}
