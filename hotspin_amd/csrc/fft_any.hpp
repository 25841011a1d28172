// fft_any.hpp -- demag FFT passes for padded lengths chosen at run time.
//
// The compiled families (fft.hpp, fft2.hpp) cover lengths 2^k, 3*2^k and
// 5*2^k with compile-time codelets.  Every other length -- 7 or 9 in a
// factor, 100 = 2^2 5^2, an odd x length of a periodic axis, a one-cell
// axis -- goes through these kernels, which take the length and its radix
// sequence as arguments, like cuFFT's plans in the reference
// (src/hotspin-core/gpu/fftplanX.go:55-60 plans D2Z/Z2D of any size).
//
// Each block holds W pencils in LDS (pencil p at lds[p * Lp + t]) and runs
// Stockham stages between two buffers: stage s of radix R, with Ns the
// product of the earlier radices, maps for j < L/R
//     v_r = x[j + r L/R] w_L^(S r (j mod Ns) L/(Ns R)),  r < R,
//     y[(j/Ns) Ns R + (j mod Ns) + q Ns] = sum_r v_r w_R^(S r q).
// Radices 2, 3, 4, 5, 7 and 8 use register codelets; any other prime factor
// takes the O(R^2) sum with roots from the length's twiddle table.  The
// axis-0 pass with the kernel product becomes three kernels here: the
// forward axis-0 transform into a spectrum of all P0 rows, the product in
// place, and the inverse transform keeping the N0 rows.
#pragma once

#include "fft2.hpp"

// Radix sequence of a length (host side).
struct GenFFT {
    int L, nst;
    int r[32];
};

inline GenFFT gen_fft_plan(int L) {
    GenFFT g{};
    g.L = L;
    int m = L;
    while (m % 8 == 0) { g.r[g.nst++] = 8; m /= 8; }
    while (m % 4 == 0) { g.r[g.nst++] = 4; m /= 4; }
    while (m % 2 == 0) { g.r[g.nst++] = 2; m /= 2; }
    for (int p = 3; m > 1; p += 2) {
        while (m % p == 0) { g.r[g.nst++] = p; m /= p; }
        if ((long)p * p > m && m > 1) { g.r[g.nst++] = m; m = 1; }
    }
    return g;
}

// Largest padded length the generic passes take (two buffers of W >= 1
// pencils of L + 1 complex values in at most 128 KB of LDS).
template <typename T>
constexpr int gen_max_len() { return (int)(131072 / (2 * sizeof(cpx<T>))) - 1; }

// w_L^(S t) from the table tw[t] = exp(-2 pi i t / L)
template <typename T, int S>
__device__ __forceinline__ cpx<T> gen_tw(const cpx<T>* __restrict__ tw, int t) {
    cpx<T> w = tw[t];
    if (S > 0) w.y = -w.y;
    return w;
}

// radix-R DFT of registers by its definition, roots at compile time (R = 7)
template <typename T, int R, int S>
__device__ __forceinline__ void dft_direct(cpx<T>* v) {
    cpx<T> y[R];
#pragma unroll
    for (int q = 0; q < R; q++) {
        cpx<T> a = v[0];
#pragma unroll
        for (int r = 1; r < R; r++) a = cadd(a, cmul(v[r], root<T, R, S>((r * q) % R)));
        y[q] = a;
    }
#pragma unroll
    for (int q = 0; q < R; q++) v[q] = y[q];
}

template <typename T, int R, int S>
__device__ __forceinline__ void gen_butterfly(const cpx<T>* s, cpx<T>* d, int j, int m, int jl, int tws, int Ns,
                                              const cpx<T>* __restrict__ tw) {
    cpx<T> v[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        v[r] = s[j + r * m];
        if (r && jl) v[r] = cmul(v[r], gen_tw<T, S>(tw, r * jl * tws));
    }
    if constexpr (R == 7) {
        dft_direct<T, 7, S>(v);
    } else {
        dftR<T, R, S>(v);
    }
#pragma unroll
    for (int q = 0; q < R; q++) d[q * Ns] = v[q];
}

// Any radix R (a prime above 7): y_q = sum_r v_r w_R^(S r q), w_R = w_L^(L/R).
template <typename T, int S>
__device__ __forceinline__ void gen_butterfly_any(const cpx<T>* s, cpx<T>* d, int R, int j, int m, int jl, int tws,
                                                  int Ns, const cpx<T>* __restrict__ tw) {
    for (int q = 0; q < R; q++) {
        cpx<T> a = {T(0), T(0)};
        for (int r = 0; r < R; r++) {
            cpx<T> v = s[j + r * m];
            if (r && jl) v = cmul(v, gen_tw<T, S>(tw, r * jl * tws));
            const int e = (r * q) % R;
            a = cadd(a, e ? cmul(v, gen_tw<T, S>(tw, e * m)) : v);
        }
        d[q * Ns] = a;
    }
}

// All stages of the W pencils of a block; returns the buffer holding the
// result (a or b).  Every thread of the block calls it.
template <typename T, int S>
__device__ cpx<T>* gen_fft(cpx<T>* a, cpx<T>* b, const GenFFT& g, int W, int Lp, const cpx<T>* __restrict__ tw) {
    const int L = g.L, tid = threadIdx.x, NT = blockDim.x;
    int Ns = 1;
    for (int s = 0; s < g.nst; s++) {
        const int R = g.r[s], m = L / R, tws = L / (Ns * R);
        for (int e = tid; e < W * m; e += NT) {
            const int p = e / m, j = e % m, jl = j % Ns;
            const cpx<T>* src = a + p * Lp;
            cpx<T>* dst = b + p * Lp + (j / Ns) * Ns * R + jl;
            switch (R) {
                case 2: gen_butterfly<T, 2, S>(src, dst, j, m, jl, tws, Ns, tw); break;
                case 3: gen_butterfly<T, 3, S>(src, dst, j, m, jl, tws, Ns, tw); break;
                case 4: gen_butterfly<T, 4, S>(src, dst, j, m, jl, tws, Ns, tw); break;
                case 5: gen_butterfly<T, 5, S>(src, dst, j, m, jl, tws, Ns, tw); break;
                case 7: gen_butterfly<T, 7, S>(src, dst, j, m, jl, tws, Ns, tw); break;
                case 8: gen_butterfly<T, 8, S>(src, dst, j, m, jl, tws, Ns, tw); break;
                default: gen_butterfly_any<T, S>(src, dst, R, j, m, jl, tws, Ns, tw); break;
            }
        }
        __syncthreads();
        cpx<T>* t = a;
        a = b;
        b = t;
        Ns *= R;
    }
    return a;
}

// ---- F1: R2C rows --------------------------------------------------------
// Rows of n2 reals (row r at x + r n2; mask rows taken modulo mrows) to rows
// of H2 = P2/2 + 1 complex at pitch H2p.  Even P2: the P2/2-point complex
// FFT of z[n] = x[2n] + i x[2n+1] and the split X[k] = E + w_P2^k O (as the
// compiled row passes).  Odd P2: the P2-point complex FFT of x.  g: plan of
// that FFT; tw its table; twP: exp(-2 pi i t / P2).
template <typename T>
__global__ void k_gen_r2c(const T* __restrict__ x, const T* __restrict__ msk, long mrows, cpx<T>* __restrict__ out,
                          long nrows, int n2, int P2, int H2p, GenFFT g, int W, const cpx<T>* __restrict__ tw,
                          const cpx<T>* __restrict__ twP, Live lv) {
    extern __shared__ __align__(16) unsigned char gen_smem[];
    const int L = g.L, Lp = L + 1, tid = threadIdx.x, NT = blockDim.x;
    cpx<T>* a = reinterpret_cast<cpx<T>*>(gen_smem);
    cpx<T>* b = a + W * Lp;
    const long row0 = (long)blockIdx.x * W;
    if (all_dead(lv, row0, (row0 + W <= nrows ? row0 + W : nrows) - 1)) return;
    const bool even = (P2 % 2) == 0;
    for (int e = tid; e < W * L; e += NT) {
        const int p = e / L, n = e % L;
        const long r = row0 + p;
        cpx<T> z = {T(0), T(0)};
        if (r < nrows) {
            const T* xr = x + r * n2;
            const T* mr = msk ? msk + (r % mrows) * n2 : nullptr;
            if (even) {
                if (2 * n < n2) z.x = xr[2 * n] * (mr ? mr[2 * n] : T(1));
                if (2 * n + 1 < n2) z.y = xr[2 * n + 1] * (mr ? mr[2 * n + 1] : T(1));
            } else if (n < n2) {
                z.x = xr[n] * (mr ? mr[n] : T(1));
            }
        }
        a[p * Lp + n] = z;
    }
    __syncthreads();
    const cpx<T>* Z = gen_fft<T, -1>(a, b, g, W, Lp, tw);
    const int H2 = P2 / 2 + 1;
    for (int e = tid; e < W * H2; e += NT) {
        const int p = e / H2, k = e % H2;
        const long r = row0 + p;
        if (r >= nrows) continue;
        cpx<T> v;
        if (even) {
            const cpx<T> zk = Z[p * Lp + (k % L)], zm = Z[p * Lp + ((L - k) % L)];
            const cpx<T> E = {T(0.5) * (zk.x + zm.x), T(0.5) * (zk.y - zm.y)};
            const cpx<T> O = {T(0.5) * (zk.y + zm.y), T(-0.5) * (zk.x - zm.x)};
            v = cadd(E, cmul(twP[k], O));
        } else {
            v = Z[p * Lp + k];
        }
        out[r * H2p + k] = v;
    }
}

// ---- I1: C2R rows (unnormalised, keeps n2 reals) --------------------------
// Even P2: Z[k] = (X[k] + conj X[M-k]) + i w_P2^-k (X[k] - conj X[M-k]),
// the inverse M-point FFT, x[2n] = Re z[n], x[2n+1] = Im z[n].  Odd P2: the
// Hermitian extension Y[k] = conj X[P2-k] for k >= H2 and the inverse
// P2-point FFT, x[n] = Re y[n].
template <typename T>
__global__ void k_gen_c2r(const cpx<T>* __restrict__ in, T* __restrict__ x, long nrows, int n2, int P2, int H2p,
                          GenFFT g, int W, const cpx<T>* __restrict__ tw, const cpx<T>* __restrict__ twP, Live lv) {
    extern __shared__ __align__(16) unsigned char gen_smem[];
    const int L = g.L, Lp = L + 1, tid = threadIdx.x, NT = blockDim.x;
    const int H2 = P2 / 2 + 1;
    cpx<T>* a = reinterpret_cast<cpx<T>*>(gen_smem);
    cpx<T>* b = a + W * Lp;
    const long row0 = (long)blockIdx.x * W;
    if (all_dead(lv, row0, (row0 + W <= nrows ? row0 + W : nrows) - 1)) return;
    const bool even = (P2 % 2) == 0;
    for (int e = tid; e < W * H2; e += NT) {  // X into b (H2 <= Lp)
        const int p = e / H2, k = e % H2;
        const long r = row0 + p;
        b[p * Lp + k] = (r < nrows) ? in[r * H2p + k] : cpx<T>{T(0), T(0)};
    }
    __syncthreads();
    for (int e = tid; e < W * L; e += NT) {
        const int p = e / L, k = e % L;
        const cpx<T>* X = b + p * Lp;
        cpx<T> v;
        if (even) {
            const cpx<T> xa = X[k], xb = X[L - k];
            const cpx<T> s = {xa.x + xb.x, xa.y - xb.y};
            const cpx<T> d = {xa.x - xb.x, xa.y + xb.y};
            cpx<T> w = twP[k];
            w.y = -w.y;
            const cpx<T> wd = cmul(w, d);
            v = {s.x - wd.y, s.y + wd.x};
        } else {
            v = (k < H2) ? X[k] : cpx<T>{X[P2 - k].x, -X[P2 - k].y};
        }
        a[p * Lp + k] = v;
    }
    __syncthreads();
    const cpx<T>* z = gen_fft<T, +1>(a, b, g, W, Lp, tw);
    for (int e = tid; e < W * n2; e += NT) {
        const int p = e / n2, n = e % n2;
        const long r = row0 + p;
        if (r >= nrows) continue;
        T v;
        if (even) {
            const cpx<T> zz = z[p * Lp + n / 2];
            v = (n & 1) ? zz.y : zz.x;
        } else {
            v = z[p * Lp + n].x;
        }
        x[r * n2 + n] = v;
    }
}

// ---- F2 / I2 and the axis-0 transforms: columns of a strided axis ---------
// Pencil a at geo_base(gi, a) (+ col), element t at + t gi.ts; W adjacent
// columns per block; nin input elements (zero above), nkeep outputs kept.
template <typename T, int S>
__global__ void k_gen_cols(const cpx<T>* __restrict__ in, cpx<T>* __restrict__ out, Geo gi, Geo go, int nin,
                           int nkeep, int ncol, GenFFT g, int W, const cpx<T>* __restrict__ tw, Live lv) {
    extern __shared__ __align__(16) unsigned char gen_smem[];
    const int L = g.L, Lp = L + 1, tid = threadIdx.x, NT = blockDim.x;
    cpx<T>* a = reinterpret_cast<cpx<T>*>(gen_smem);
    cpx<T>* b = a + W * Lp;
    const int pen = blockIdx.y;
    if (all_dead(lv, pen, pen)) return;
    const int col0 = blockIdx.x * W;
    const long bi = geo_base(gi, pen) + col0;
    const long bo = geo_base(go, pen) + col0;
    for (int e = tid; e < W * L; e += NT) {
        const int t = e / W, p = e % W;
        cpx<T> v = {T(0), T(0)};
        if (t < nin && col0 + p < ncol) v = in[bi + t * gi.ts + p];
        a[p * Lp + t] = v;
    }
    __syncthreads();
    const cpx<T>* y = gen_fft<T, S>(a, b, g, W, Lp, tw);
    for (int e = tid; e < W * nkeep; e += NT) {
        const int t = e / W, p = e % W;
        if (col0 + p < ncol) out[bo + t * go.ts + p] = y[p * Lp + t];
    }
}

// ---- the kernel product on a spectrum of all P0 rows ----------------------
// Xz layout (nsolve, 3, P0, P1, H2p); B^ = srcMul K^ M^ with the symmetric
// kernel's octant and mirror signs (tensSYMMVecMul, src/libhotspin/mul.cu:31-90):
// row i > P0/2 reads octant row P0 - i and flips XZ, XY; row j > P1/2 reads
// P1 - j and flips YZ, XY.
template <typename T, typename TK>
__global__ void k_gen_kprod(cpx<T>* __restrict__ X, const TK* __restrict__ kern, int P0, int P1, int H2, int H2p,
                            int K1, int kp, T srcMul, int nsolve, Live lv) {
    const long per = (long)P0 * P1 * H2;
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= per * nsolve) return;
    const int sol = (int)(g / per);
    if (lv.flag && !lv.flag[sol]) return;
    long r = g % per;
    const int col = (int)(r % H2);
    r /= H2;
    const int j = (int)(r % P1), i = (int)(r / P1);
    const int ii = (i > P0 / 2) ? P0 - i : i, jj = (j > P1 / 2) ? P1 - j : j;
    const T si = (i > P0 / 2) ? T(-1) : T(1), sj = (j > P1 / 2) ? T(-1) : T(1);
    const long kplane = (long)(P0 / 2 + 1) * K1 * kp;
    const long K = ((long)ii * K1 + jj) * kp + col;
    const T kxx = T(kern[K]), kyy = T(kern[kplane + K]), kzz = T(kern[2 * kplane + K]);
    const T kyz = sj * T(kern[3 * kplane + K]);
    const T kxz = si * T(kern[4 * kplane + K]);
    const T kxy = si * sj * T(kern[5 * kplane + K]);
    const long cs = (long)P0 * P1 * H2p;
    cpx<T>* a = X + (long)sol * 3 * cs + ((long)i * P1 + j) * H2p + col;
    const cpx<T> mx = cscale(a[0], srcMul), my = cscale(a[cs], srcMul), mz = cscale(a[2 * cs], srcMul);
    a[0] = {kxx * mx.x + kxy * my.x + kxz * mz.x, kxx * mx.y + kxy * my.y + kxz * mz.y};
    a[cs] = {kxy * mx.x + kyy * my.x + kyz * mz.x, kxy * mx.y + kyy * my.y + kyz * mz.y};
    a[2 * cs] = {kxz * mx.x + kyz * my.x + kzz * mz.x, kxz * mx.y + kyz * my.y + kzz * mz.y};
}

// ---- launches (host) --------------------------------------------------------
namespace gen {
template <typename T>
inline int pencils_per_block(int L) {
    const long per = 2L * (L + 1) * (long)sizeof(cpx<T>);
    int w = (int)(32768 / per);
    return w < 1 ? 1 : (w > 16 ? 16 : w);
}
inline int threads_for(long work) {
    long t = (work + 63) / 64 * 64;
    return (int)(t < 64 ? 64 : (t > 256 ? 256 : t));
}
template <typename K>
inline void allow_lds(K kern, size_t bytes) {
    if (bytes > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)bytes);
}

template <typename T>
void r2c(const T* x, const T* msk, long mrows, cpx<T>* out, long nrows, int n2, int P2, int H2p, const GenFFT& g,
         const cpx<T>* tw, const cpx<T>* twP, hipStream_t st, Live lv) {
    const int W = pencils_per_block<T>(g.L);
    const size_t lds = 2ull * W * (g.L + 1) * sizeof(cpx<T>);
    allow_lds(k_gen_r2c<T>, lds);
    hipLaunchKernelGGL(k_gen_r2c<T>, dim3((unsigned)((nrows + W - 1) / W)), dim3(threads_for((long)W * g.L / 2)),
                       lds, st, x, msk, mrows, out, nrows, n2, P2, H2p, g, W, tw, twP, lv);
}
template <typename T>
void c2r(const cpx<T>* in, T* x, long nrows, int n2, int P2, int H2p, const GenFFT& g, const cpx<T>* tw,
         const cpx<T>* twP, hipStream_t st, Live lv) {
    const int W = pencils_per_block<T>(g.L);
    const size_t lds = 2ull * W * (g.L + 1) * sizeof(cpx<T>);
    allow_lds(k_gen_c2r<T>, lds);
    hipLaunchKernelGGL(k_gen_c2r<T>, dim3((unsigned)((nrows + W - 1) / W)), dim3(threads_for((long)W * g.L / 2)),
                       lds, st, in, x, nrows, n2, P2, H2p, g, W, tw, twP, lv);
}
template <typename T, int S>
void cols(const cpx<T>* in, cpx<T>* out, Geo gi, Geo go, int npencil, int nin, int nkeep, int ncol, const GenFFT& g,
          const cpx<T>* tw, hipStream_t st, Live lv) {
    const int W = pencils_per_block<T>(g.L);
    const size_t lds = 2ull * W * (g.L + 1) * sizeof(cpx<T>);
    allow_lds(k_gen_cols<T, S>, lds);
    dim3 grid((unsigned)((ncol + W - 1) / W), (unsigned)npencil);
    hipLaunchKernelGGL((k_gen_cols<T, S>), grid, dim3(threads_for((long)W * g.L / 2)), lds, st, in, out, gi, go, nin,
                       nkeep, ncol, g, W, tw, lv);
}
// F3K of a length without a compiled axis-0 pass: X2 (nsolve, 3, N0, P1, H2p)
// -> Xz (nsolve, 3, P0, P1, H2p) forward along axis 0, the product, back into
// the N0 rows of X2.
template <typename T, typename TK>
void zconv(cpx<T>* X2, cpx<T>* Xz, const TK* kern, int N0, int P0, int P1, int H2, int H2p, int K1, int kp, T srcMul,
           const GenFFT& g, const cpx<T>* tw, hipStream_t st, int nsolve, Live lv) {
    const Geo g2 = {(long)N0 * P1 * H2p, (long)H2p, (long)P1 * H2p, P1};
    const Geo gz = {(long)P0 * P1 * H2p, (long)H2p, (long)P1 * H2p, P1};
    const Live lp{lv.flag, 3L * P1};
    cols<T, -1>(X2, Xz, g2, gz, 3 * P1 * nsolve, N0, P0, H2, g, tw, st, lp);
    const long n = (long)P0 * P1 * H2 * nsolve;
    hipLaunchKernelGGL((k_gen_kprod<T, TK>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, Xz, kern, P0, P1, H2,
                       H2p, K1, kp, srcMul, nsolve, Live{lv.flag, 1});
    cols<T, +1>(Xz, X2, gz, g2, 3 * P1 * nsolve, P0, N0, H2, g, tw, st, lp);
}
}  // namespace gen
