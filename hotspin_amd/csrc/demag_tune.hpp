// demag_tune.hpp -- registry of FP32 tile variants of the demag FFT passes.
// Each supported length registers several (pencils per block W, threads NT)
// instantiations; the demag plan times them on its own buffers at creation
// and keeps the fastest per pass.
#pragma once

#include <vector>

#include "demag_kernels.hpp"
#include "fft2.hpp"

namespace hsdm {
using r2c32_t = void (*)(const float*, const float*, long, cpx<float>*, long, int, int, const cpx<float>*,
                         const cpx<float>*, hipStream_t, Live);
using c2r32_t = void (*)(const cpx<float>*, float*, long, int, int, const cpx<float>*, const cpx<float>*,
                         hipStream_t, Live);
using cols32_t = void (*)(const cpx<float>*, cpx<float>*, Geo, Geo, int, int, int, int, const cpx<float>*,
                          hipStream_t, Live);
using z32_t = void (*)(cpx<float>*, const float*, int, int, int, int, int, int, float, const cpx<float>*,
                       hipStream_t, int, Live);
struct RowV {
    int W, NT;
    r2c32_t r2c;
    c2r32_t c2r;
};
struct ColV {
    int W, NT;
    cols32_t fwd, inv;
};
struct ZV {
    int W, NT;
    z32_t fn;
};
struct Variants {
    std::vector<RowV> rows;
    std::vector<ColV> cols;
    std::vector<ZV> z;
};
Variants& variants(int L);  // registry (defined in demag.hip)

constexpr long kLdsMax = 163840;
constexpr int round64(int v) { return (v + 63) / 64 * 64; }
constexpr int clamp_nt(long v) { return round64((int)(v < 64 ? 64 : (v > 1024 ? 1024 : v))); }

template <int M, int W, int NT>
void launch_r2c32(const float* x, const float* msk, long mrows, cpx<float>* out, long nrows, int n2, int H2p,
                  const cpx<float>* twM, const cpx<float>* twP, hipStream_t st, Live lv) {
    hipLaunchKernelGGL((k_r2c_rows<float, float, M, W, NT>), dim3((unsigned)((nrows + W - 1) / W)), dim3(NT), 0,
                       st, x, msk, mrows, out, nrows, n2, H2p, twM, twP, lv);
}
template <int M, int W, int NT>
void launch_c2r32(const cpx<float>* in, float* x, long nrows, int n2, int H2p, const cpx<float>* twM,
                  const cpx<float>* twP, hipStream_t st, Live lv) {
    hipLaunchKernelGGL((k_c2r_rows<float, float, M, W, NT>), dim3((unsigned)((nrows + W - 1) / W)), dim3(NT), 0,
                       st, in, x, nrows, n2, H2p, twM, twP, lv);
}
template <int L, int S, int W, int NT>
void launch_cols32(const cpx<float>* in, cpx<float>* out, Geo gi, Geo go, int npencil, int nin, int nkeep,
                   int ncol, const cpx<float>* tw, hipStream_t st, Live lv) {
    dim3 g((unsigned)((ncol + W - 1) / W), (unsigned)npencil);
    hipLaunchKernelGGL((k_cols<float, L, S, W, NT>), g, dim3(NT), 0, st, in, out, gi, go, nin, nkeep, ncol, tw, lv);
}
template <int L, int W, int NT>
void launch_z32(cpx<float>* X, const float* kern, int N0, int P1, int H2, int H2p, int K1, int kp, float srcMul,
                const cpx<float>* tw, hipStream_t st, int nsolve, Live lv) {
    dim3 g((unsigned)((H2 + W - 1) / W), (unsigned)(P1 / 2 + 1), (unsigned)nsolve);
    hipLaunchKernelGGL((k_zconv<float, float, L, W, NT>), g, dim3(NT), 0, st, X, kern, N0, P1, H2, H2p, K1, kp,
                       srcMul, tw, lv);
}

template <int L, int S, int W>
void launch_cols2_32(const cpx<float>* in, cpx<float>* out, Geo gi, Geo go, int npencil, int nin, int nkeep,
                     int ncol, const cpx<float>* tw, hipStream_t st, Live lv) {
    dim3 g((unsigned)((ncol + W - 1) / W), (unsigned)npencil);
    hipLaunchKernelGGL((k_cols2<float, L, S, W>), g, dim3(W * split2(L).B), 0, st, in, out, gi, go, nin, nkeep, ncol,
                       tw, lv);
}

template <int M, int W, int E>
void add_row(Variants& v) {
    constexpr int NT = clamp_nt((long)M * W / E);
    if constexpr ((long)(M + 1) * W * 8 <= kLdsMax) v.rows.push_back({W, NT, &launch_r2c32<M, W, NT>, &launch_c2r32<M, W, NT>});
}
template <int L, int W, int E>
void add_col(Variants& v) {
    constexpr int NT = clamp_nt((long)L * W / E);
    if constexpr ((long)L * W * 8 <= kLdsMax) v.cols.push_back({W, NT, &launch_cols32<L, -1, W, NT>, &launch_cols32<L, 1, W, NT>});
}
template <int L, int W>
void launch_z2_32(cpx<float>* X, const float* kern, int N0, int P1, int H2, int H2p, int K1, int kp, float srcMul,
                  const cpx<float>* tw, hipStream_t st, int nsolve, Live lv) {
    dim3 g((unsigned)((H2 + W - 1) / W), (unsigned)(P1 / 2 + 1), (unsigned)nsolve);
    hipLaunchKernelGGL((k_zconv2<float, float, L, W>), g, dim3(6 * W), 0, st, X, kern, N0, P1, H2, H2p, K1, kp,
                       srcMul, tw, lv);
}
// register-FFT axis-0 pass (fft2.hpp): one thread per pencil, L <= 64
template <int L, int W>
void add_z2(Variants& v) {
    if constexpr (reg_dft_ok(L) && L > 1 && L <= 64 && (long)L * 6 * W * 8 <= kLdsMax)
        v.z.push_back({W, 6 * W, &launch_z2_32<L, W>});
}

// the same with the kernel product in two real halves (fft2.hpp k_zconv2s)
template <int L, int W>
void launch_z2s_32(cpx<float>* X, const float* kern, int N0, int P1, int H2, int H2p, int K1, int kp, float srcMul,
                   const cpx<float>* tw, hipStream_t st, int nsolve, Live lv) {
    dim3 g((unsigned)((H2 + W - 1) / W), (unsigned)(P1 / 2 + 1), (unsigned)nsolve);
    hipLaunchKernelGGL((k_zconv2s<float, float, L, W>), g, dim3(6 * W), 0, st, X, kern, N0, P1, H2, H2p, K1, kp,
                       srcMul, tw, lv);
}
// ... and with at most 128 VGPRs (4 waves per SIMD)
template <int L, int W>
void launch_z2s4_32(cpx<float>* X, const float* kern, int N0, int P1, int H2, int H2p, int K1, int kp, float srcMul,
                    const cpx<float>* tw, hipStream_t st, int nsolve, Live lv) {
    dim3 g((unsigned)((H2 + W - 1) / W), (unsigned)(P1 / 2 + 1), (unsigned)nsolve);
    hipLaunchKernelGGL((k_zconv2s<float, float, L, W, 4>), g, dim3(6 * W), 0, st, X, kern, N0, P1, H2, H2p, K1, kp,
                       srcMul, tw, lv);
}
template <int L, int W>
void add_z2s4(Variants& v) {
    if constexpr (reg_dft_ok(L) && L > 1 && L <= 64 && (long)L * 6 * W * 4 <= kLdsMax)
        v.z.push_back({W, 6 * W, &launch_z2s4_32<L, W>});
}
template <int L, int W>
void add_z2s(Variants& v) {
    if constexpr (reg_dft_ok(L) && L > 1 && L <= 64 && (long)L * 6 * W * 4 <= kLdsMax)
        v.z.push_back({W, 6 * W, &launch_z2s_32<L, W>});
}

// the same with non-temporal pencil loads and stores (fft2.hpp k_zconv2n)
template <int L, int W>
void launch_z2n_32(cpx<float>* X, const float* kern, int N0, int P1, int H2, int H2p, int K1, int kp, float srcMul,
                   const cpx<float>* tw, hipStream_t st, int nsolve, Live lv) {
    dim3 g((unsigned)((H2 + W - 1) / W), (unsigned)(P1 / 2 + 1), (unsigned)nsolve);
    hipLaunchKernelGGL((k_zconv2n<float, float, L, W>), g, dim3(6 * W), 0, st, X, kern, N0, P1, H2, H2p, K1, kp,
                       srcMul, tw, lv);
}
template <int L, int W>
void add_z2n(Variants& v) {
    if constexpr (reg_dft_ok(L) && L > 1 && L <= 64 && (long)L * 6 * W * 8 <= kLdsMax)
        v.z.push_back({W, 6 * W, &launch_z2n_32<L, W>});
}

// the same with Q = 2 or 4 lanes per pencil, L = Q H with 4 <= H <= 32
// (fft2.hpp k_zconv2m)
template <int L, int W, int Q>
void launch_z2m_32(cpx<float>* X, const float* kern, int N0, int P1, int H2, int H2p, int K1, int kp, float srcMul,
                   const cpx<float>* tw, hipStream_t st, int nsolve, Live lv) {
    dim3 g((unsigned)((H2 + W - 1) / W), (unsigned)(P1 / 2 + 1), (unsigned)nsolve);
    hipLaunchKernelGGL((k_zconv2m<float, float, L, W, Q>), g, dim3(6 * Q * W), 0, st, X, kern, N0, P1, H2, H2p, K1,
                       kp, srcMul, tw, lv);
}
template <int L, int W, int Q>
void add_z2m(Variants& v) {
    if constexpr (L % Q == 0 && L / Q >= 4 && L / Q <= 32 && L >= 16 && reg_dft_ok(L / Q) &&
                  (long)L * 6 * W * 8 <= kLdsMax && 6 * Q * W <= 1024)
        v.z.push_back({W, 6 * Q * W, &launch_z2m_32<L, W, Q>});
}

// two-level register FFT columns (fft2.hpp): W columns, W * B threads
template <int L, int W>
void add_col2(Variants& v) {
    if constexpr (two_level_ok(L) && L > 1) {
        constexpr int NT = W * split2(L).B;
        if constexpr ((long)(L * W + L) * 8 <= kLdsMax && NT >= 64 && NT <= 1024)
            v.cols.push_back({W, NT, &launch_cols2_32<L, -1, W>, &launch_cols2_32<L, 1, W>});
    }
}

// column pass with a split real/imaginary LDS exchange (fft2.hpp k_cols2s)
template <int L, int S, int W>
void launch_cols2s_32(const cpx<float>* in, cpx<float>* out, Geo gi, Geo go, int npencil, int nin, int nkeep,
                      int ncol, const cpx<float>* tw, hipStream_t st, Live lv) {
    dim3 g((unsigned)((ncol + W - 1) / W), (unsigned)npencil);
    hipLaunchKernelGGL((k_cols2s<float, L, S, W>), g, dim3(W * split2(L).B), 0, st, in, out, gi, go, nin, nkeep,
                       ncol, tw, lv);
}
template <int L, int W>
void add_col2s(Variants& v) {
    if constexpr (two_level_ok(L) && L > 1) {
        constexpr int NT = W * split2(L).B;
        if constexpr ((long)L * W * 4 + (long)L * 8 <= kLdsMax && NT >= 64 && NT <= 1024)
            v.cols.push_back({W, NT, &launch_cols2s_32<L, -1, W>, &launch_cols2s_32<L, 1, W>});
    }
}

// the same with a bank-conflict-free padded exchange (fft2.hpp k_cols2p)
template <int L, int S, int W>
void launch_cols2p_32(const cpx<float>* in, cpx<float>* out, Geo gi, Geo go, int npencil, int nin, int nkeep,
                      int ncol, const cpx<float>* tw, hipStream_t st, Live lv) {
    dim3 g((unsigned)((ncol + W - 1) / W), (unsigned)npencil);
    hipLaunchKernelGGL((k_cols2p<float, L, S, W, 3>), g, dim3(W * split2(L).B), 0, st, in, out, gi, go, nin, nkeep,
                       ncol, tw, lv);
}
template <int L, int W>
void add_col2p(Variants& v) {
    if constexpr (two_level_ok(L) && L > 1) {
        constexpr int NT = W * split2(L).B;
        if constexpr ((long)split2(L).A * (split2(L).B * W + W) * 4 + (long)L * 8 <= kLdsMax && NT >= 64 &&
                      NT <= 1024)
            v.cols.push_back({W, NT, &launch_cols2p_32<L, -1, W>, &launch_cols2p_32<L, 1, W>});
    }
}
// half-length column pass (fft2.hpp k_cols2h) for a zero-padded axis: the
// forward pass needs nin <= L/2 and the inverse nkeep <= L/2; otherwise (a
// periodic axis, no padding) the launcher runs the length-L k_cols2
template <int L, int S, int W>
void launch_cols2h_32(const cpx<float>* in, cpx<float>* out, Geo gi, Geo go, int npencil, int nin, int nkeep,
                      int ncol, const cpx<float>* tw, hipStream_t st, Live lv) {
    dim3 g((unsigned)((ncol + W - 1) / W), (unsigned)npencil);
    const bool half = (S < 0) ? (2 * nin <= L) : (2 * nkeep <= L);
    if (half)
        hipLaunchKernelGGL((k_cols2h<float, L, S, W>), g, dim3(W * split2(L / 2).B), 0, st, in, out, gi, go, nin,
                           nkeep, ncol, tw, lv);
    else
        hipLaunchKernelGGL((k_cols2<float, L, S, W>), g, dim3(W * split2(L).B), 0, st, in, out, gi, go, nin, nkeep,
                           ncol, tw, lv);
}
template <int L, int W>
void add_col2h(Variants& v) {
    if constexpr (L % 2 == 0 && L >= 4 && two_level_ok(L) && two_level_ok(L / 2)) {
        constexpr int NTh = W * split2(L / 2).B, NT = W * split2(L).B;
        if constexpr ((long)(L * W + L) * 8 <= kLdsMax && NTh >= 64 && NTh <= 1024 && NT >= 64 && NT <= 1024)
            v.cols.push_back({W, NTh, &launch_cols2h_32<L, -1, W>, &launch_cols2h_32<L, 1, W>});
    }
}

// two-level register FFT rows (fft2.hpp): W rows, W * B threads
template <int M, int W>
void launch_r2c2_32(const float* x, const float* msk, long mrows, cpx<float>* out, long nrows, int n2, int H2p,
                    const cpx<float>* twM, const cpx<float>* twP, hipStream_t st, Live lv) {
    hipLaunchKernelGGL((k_r2c_rows2<float, M, W>), dim3((unsigned)((nrows + W - 1) / W)), dim3(W * split2(M).B), 0,
                       st, x, msk, mrows, out, nrows, n2, H2p, twM, twP, lv);
}
template <int M, int W>
void launch_c2r2_32(const cpx<float>* in, float* x, long nrows, int n2, int H2p, const cpx<float>* twM,
                    const cpx<float>* twP, hipStream_t st, Live lv) {
    hipLaunchKernelGGL((k_c2r_rows2<float, M, W>), dim3((unsigned)((nrows + W - 1) / W)), dim3(W * split2(M).B), 0,
                       st, in, x, nrows, n2, H2p, twM, twP, lv);
}
template <int M, int W>
void add_row2(Variants& v) {
    if constexpr (two_level_ok(M) && M > 1) {
        constexpr int NT = W * split2(M).B;
        if constexpr (((long)rows2_lds<M, W>() + 2L * M) * 8 <= kLdsMax && NT >= 64 && NT <= 1024)
            v.rows.push_back({W, NT, &launch_r2c2_32<M, W>, &launch_c2r2_32<M, W>});
    }
}

template <int L, int W, int E>
void add_z(Variants& v) {  // a block holds a pencil pair: 6W pencils
    constexpr int NT = clamp_nt((long)L * 6 * W / E);
    if constexpr (L <= 256 && (long)L * 6 * W * 8 <= kLdsMax) v.z.push_back({W, NT, &launch_z32<L, W, NT>});
}

template <int L>
void register_size() {
    Variants& v = variants(L);
    add_row<L, 2, 8>(v);
    add_row<L, 4, 8>(v);
    add_row<L, 8, 8>(v);
    add_row<L, 4, 4>(v);
    add_row<L, 16, 8>(v);
    add_row2<L, 2>(v);
    add_row2<L, 4>(v);
    add_row2<L, 8>(v);
    add_col<L, 8, 8>(v);
    add_col<L, 8, 16>(v);
    add_col<L, 16, 16>(v);
    add_col<L, 16, 8>(v);
    add_col<L, 4, 8>(v);
    add_col<L, 4, 4>(v);
    add_col2<L, 4>(v);
    add_col2<L, 8>(v);
    add_col2<L, 16>(v);
    add_col2h<L, 8>(v);
    add_col2h<L, 16>(v);
    add_col2s<L, 8>(v);
    add_col2s<L, 16>(v);
    add_col2s<L, 32>(v);
    add_col2p<L, 8>(v);
    add_col2p<L, 16>(v);
    add_z<L, 8, 12>(v);
    add_z<L, 16, 12>(v);
    add_z<L, 32, 12>(v);
    add_z<L, 16, 24>(v);
    add_z<L, 32, 24>(v);
    add_z<L, 8, 6>(v);
    add_z2<L, 16>(v);
    add_z2<L, 32>(v);
    add_z2s<L, 16>(v);
    add_z2s<L, 32>(v);
    add_z2n<L, 16>(v);
    add_z2n<L, 32>(v);
    add_z2s4<L, 32>(v);
    add_z2m<L, 8, 2>(v);
    add_z2m<L, 16, 2>(v);
    add_z2m<L, 32, 2>(v);
    add_z2m<L, 8, 4>(v);
    add_z2m<L, 16, 4>(v);
}
}  // namespace hsdm
