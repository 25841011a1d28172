// legacy.hip -- the FP64 libhotspin ABI (include/libhotspin.h) on gfx950.
//
// One lane per cell over SoA planes, 256-lane blocks (4 wave64s).  The
// stencils (exchange6, Qspat) map 64 lanes along the contiguous k axis and 4
// rows along j so every gather is a coalesced row segment; neighbour rows are
// served from L2/L1.  Formulas restate the cited reference kernels.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "hs_common.hpp"

// ------------------------------------------------------------------ errors
namespace {
thread_local int g_err = 0;
thread_local char g_msg[256] = {0};
}  // namespace

void hs::set_error(int code, const char* what) {
    if (g_err == 0) {
        g_err = code;
        snprintf(g_msg, sizeof(g_msg), "%s", what ? what : "");
    }
}
void hs::check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        char b[256];
        snprintf(b, sizeof(b), "%s: %s", what, hipGetErrorString(e));
        hs::set_error((int)e, b);
    }
}
extern "C" int hs_last_error(void) { return g_err; }
extern "C" const char* hs_last_error_string(void) { return g_msg; }
extern "C" void hs_clear_error(void) {
    g_err = 0;
    g_msg[0] = 0;
}
extern "C" const char* hs_version(void) { return "hotspin_amd 0.1 gfx950"; }

#define ARGCHECK(cond, name)                              \
    do {                                                  \
        if (!(cond)) {                                    \
            hs::set_error(HS_EINVAL, name); \
            return;                                       \
        }                                                 \
    } while (0)

// true when every pointer is non-null: the arrays a kernel always
// dereferences (masks, which may be NULL = uniform, are not passed here)
static inline bool nn() { return true; }
template <class P, class... R>
static inline bool nn(P p, R... r) {
    return p != nullptr && nn(r...);
}

// ------------------------------------------------------------------ stencils
// exchange6Kern, src/libhotspin/exchange6.cu:11-127 (neighbour order X, Z, Y)
__global__ void __launch_bounds__(256) k_exchange6(double* __restrict__ hx, double* __restrict__ hy,
                                                   double* __restrict__ hz, const double* __restrict__ mx,
                                                   const double* __restrict__ my, const double* __restrict__ mz,
                                                   const double* __restrict__ lexMsk, int N0, int N1, int N2,
                                                   int w0, int w1, int w2, double ms, double p0, double p1,
                                                   double p2) {
    int k = blockIdx.x * 64 + threadIdx.x;
    int j = blockIdx.y * 4 + threadIdx.y;
    int i = blockIdx.z;
    if (k >= N2 || j >= N1) return;
    long N12 = (long)N1 * N2;
    long I = i * N12 + (long)j * N2 + k;
    double lex = mask1(lexMsk, I);
    double lex02 = lex * lex;
    double m0x = mx[I], m0y = my[I], m0z = mz[I];
    double H[3];
    const int ax_order[3] = {0, 2, 1};
    const double pre[3] = {p0, p1, p2};
#pragma unroll
    for (int q = 0; q < 3; q++) {
        int ax = ax_order[q];
        int n = ax == 0 ? N0 : (ax == 1 ? N1 : N2);
        int w = ax == 0 ? w0 : (ax == 1 ? w1 : w2);
        int c = ax == 0 ? i : (ax == 1 ? j : k);
        long stride = ax == 0 ? N12 : (ax == 1 ? (long)N2 : 1L);
        int lo = c - 1;
        lo = (lo < 0 && w) ? n + lo : lo;
        lo = max(lo, 0);
        int hi = c + 1;
        hi = (hi == n && w) ? hi - n : hi;
        hi = min(hi, n - 1);
        long a1 = I + (long)(lo - c) * stride, a2 = I + (long)(hi - c) * stride;
        double l1 = mask1(lexMsk, a1), l2 = mask1(lexMsk, a2);
        double pre1 = avg_geom_zero(lex02, l1 * l1);
        double pre2 = avg_geom_zero(lex02, l2 * l2);
        double dx = pre[ax] * ((pre1 * mx[a1] + pre2 * mx[a2]) - (pre1 * m0x + pre2 * m0x));
        double dy = pre[ax] * ((pre1 * my[a1] + pre2 * my[a2]) - (pre1 * m0y + pre2 * m0y));
        double dz = pre[ax] * ((pre1 * mz[a1] + pre2 * mz[a2]) - (pre1 * m0z + pre2 * m0z));
        if (q == 0) {
            H[0] = dx; H[1] = dy; H[2] = dz;
        } else {
            H[0] += dx; H[1] += dy; H[2] += dz;
        }
    }
    hx[I] = ms * H[0];
    hy[I] = ms * H[1];
    hz[I] = ms * H[2];
}

extern "C" void exchange6Async(double* hx, double* hy, double* hz, double* mx, double* my, double* mz,
                               double* msat0T0, double* lex, int N0, int N1, int N2, int p0, int p1, int p2,
                               double msat0T0Mul, double c0, double c1, double c2, CUstream stream) {
    (void)msat0T0;  // accepted but unused, as in exchange6.cu:134-147
    ARGCHECK(N0 > 0 && N1 > 0 && N2 > 0 && nn(hx, hy, hz, mx, my, mz), "exchange6Async");
    dim3 blk(64, 4), grd(hs::div_up(N2, 64), hs::div_up(N1, 4), N0);
    hipLaunchKernelGGL(k_exchange6, grd, blk, 0, hs::S(stream), hx, hy, hz, mx, my, mz, lex, N0, N1, N2, p0, p1,
                       p2, msat0T0Mul, c0, c1, c2);
    HS_LAUNCH_CHECK("exchange6Async");
}

// QspatKern, src/libhotspin/Qspat.cu:13-103
__global__ void __launch_bounds__(256) k_qspat(double* __restrict__ Q, const double* __restrict__ T,
                                               const double* __restrict__ kMsk, double kMul, int sx, int sy,
                                               int sz, double c2x, double c2y, double c2z, int px, int py,
                                               int pz) {
    int k = blockIdx.x * 64 + threadIdx.x;
    int j = blockIdx.y * 4 + threadIdx.y;
    int i = blockIdx.z;
    if (k >= sz || j >= sy) return;
    long syz = (long)sy * sz;
    long I = i * syz + (long)j * sz + k;
    double T0 = T[I], k0 = mask1(kMsk, I), ddT = 0;
    const int ax_order[3] = {0, 2, 1};
    const double c2[3] = {c2x, c2y, c2z};
#pragma unroll
    for (int q = 0; q < 3; q++) {
        int ax = ax_order[q];
        int n = ax == 0 ? sx : (ax == 1 ? sy : sz);
        int w = ax == 0 ? px : (ax == 1 ? py : pz);
        int c = ax == 0 ? i : (ax == 1 ? j : k);
        long stride = ax == 0 ? syz : (ax == 1 ? (long)sz : 1L);
        int lo = c - 1;
        lo = (lo < 0 && w) ? n + lo : lo;
        lo = max(lo, 0);
        int hi = c + 1;
        hi = (hi == n && w) ? hi - n : hi;
        hi = min(hi, n - 1);
        long a1 = I + (long)(lo - c) * stride, a2 = I + (long)(hi - c) * stride;
        double pre1 = avg_geom_zero(k0, mask1(kMsk, a1));
        double pre2 = avg_geom_zero(k0, mask1(kMsk, a2));
        double d = c2[ax] * ((pre1 * T[a1] + pre2 * T[a2]) - (pre1 * T0 + pre2 * T0));
        ddT = (q == 0) ? d : ddT + d;
    }
    Q[I] = kMul * ddT;
}

extern "C" void Qspat_async(double* Q, double* T, double* k, const double kMul, const int sx, const int sy,
                            const int sz, const double csx, const double csy, const double csz, const int px,
                            const int py, const int pz, CUstream stream) {
    ARGCHECK(sx > 0 && sy > 0 && sz > 0 && Q && T, "Qspat_async");
    dim3 blk(64, 4), grd(hs::div_up(sz, 64), hs::div_up(sy, 4), sx);
    hipLaunchKernelGGL(k_qspat, grd, blk, 0, hs::S(stream), Q, T, k, kMul, sx, sy, sz, 1.0 / (csx * csx),
                       1.0 / (csy * csy), 1.0 / (csz * csz), px, py, pz);
    HS_LAUNCH_CHECK("Qspat_async");
}

// ------------------------------------------------------------------ LLB terms
#define LANE_LOOP(N) \
    int i = blockIdx.x * blockDim.x + threadIdx.x; \
    if (i >= (N)) return;

// llbarTorqueKern, llbar-torque.cu:11-45: t = cross(H, M) = M x H
__global__ void k_torque(double* tx, double* ty, double* tz, const double* Mx, const double* My, const double* Mz,
                         const double* hx, const double* hy, const double* hz, const double* msk, int N) {
    LANE_LOOP(N)
    double ms = msk ? msk[i] : 1.0;
    if (ms == 0.0) {
        tx[i] = 0.0; ty[i] = 0.0; tz[i] = 0.0;
        return;
    }
    double Hx = hx[i], Hy = hy[i], Hz = hz[i], mX = Mx[i], mY = My[i], mZ = Mz[i];
    tx[i] = -Hy * mZ + Hz * mY;
    ty[i] = -Hz * mX + Hx * mZ;
    tz[i] = -Hx * mY + Hy * mX;
}
extern "C" void llbar_torque_async(double* tx, double* ty, double* tz, double* Mx, double* My, double* Mz,
                                   double* hx, double* hy, double* hz, double* msat0T0, CUstream stream,
                                   int N) {
    ARGCHECK(N >= 0 && (N == 0 || nn(tx, ty, tz, Mx, My, Mz, hx, hy, hz)), "llbar_torque_async");
    if (N == 0) return;
    hipLaunchKernelGGL(k_torque, hs::grid1d(N), hs::kBlock, 0, hs::S(stream), tx, ty, tz, Mx, My, Mz, hx, hy, hz,
                       msat0T0, N);
    HS_LAUNCH_CHECK("llbar_torque_async");
}

// llbarLocal00ncKern, llbar-local00nc.cu:11-62
__global__ void k_local00nc(double* tx, double* ty, double* tz, const double* hx, const double* hy,
                            const double* hz, const double* msk, const double* lxx, const double* lyy,
                            const double* lzz, double mxx, double myy, double mzz, int N) {
    LANE_LOOP(N)
    double ms = msk ? msk[i] : 1.0;
    if (ms == 0.0) {
        tx[i] = 0.0; ty[i] = 0.0; tz[i] = 0.0;
        return;
    }
    double a = lxx ? lxx[i] * mxx : mxx;
    double b = lyy ? lyy[i] * myy : myy;
    double c = lzz ? lzz[i] * mzz : mzz;
    tx[i] = a * hx[i];
    ty[i] = b * hy[i];
    tz[i] = c * hz[i];
}
extern "C" void llbar_local00nc_async(double* tx, double* ty, double* tz, double* hx, double* hy, double* hz,
                                      double* msat0T0, double* lxx, double* lyy, double* lzz, const double mxx,
                                      const double myy, const double mzz, CUstream stream, int N) {
    ARGCHECK(N >= 0 && (N == 0 || nn(tx, ty, tz, hx, hy, hz)), "llbar_local00nc_async");
    if (N == 0) return;
    hipLaunchKernelGGL(k_local00nc, hs::grid1d(N), hs::kBlock, 0, hs::S(stream), tx, ty, tz, hx, hy, hz, msat0T0,
                       lxx, lyy, lzz, mxx, myy, mzz, N);
    HS_LAUNCH_CHECK("llbar_local00nc_async");
}

// llbarLocal02ncKern, llbar-local02nc.cu:11-73: t = diag(mu) (m.H) m
__global__ void k_local02nc(double* tx, double* ty, double* tz, const double* mx, const double* my,
                            const double* mz, const double* hx, const double* hy, const double* hz,
                            const double* msk, const double* uxx, const double* uyy, const double* uzz,
                            double mxx, double myy, double mzz, int N) {
    LANE_LOOP(N)
    double ms = msk ? msk[i] : 1.0;
    if (ms == 0.0) {
        tx[i] = 0.0; ty[i] = 0.0; tz[i] = 0.0;
        return;
    }
    double m0 = mx[i], m1 = my[i], m2 = mz[i];
    double mH = m0 * hx[i] + m1 * hy[i] + m2 * hz[i];
    double a = uxx ? uxx[i] * mxx : mxx;
    double b = uyy ? uyy[i] * myy : myy;
    double c = uzz ? uzz[i] * mzz : mzz;
    tx[i] = a * (mH * m0);
    ty[i] = b * (mH * m1);
    tz[i] = c * (mH * m2);
}
extern "C" void llbar_local02nc_async(double* tx, double* ty, double* tz, double* mx, double* my, double* mz,
                                      double* hx, double* hy, double* hz, double* msat0T0, double* uxx,
                                      double* uyy, double* uzz, const double mxx, const double myy,
                                      const double mzz, CUstream stream, int N) {
    ARGCHECK(N >= 0 && (N == 0 || nn(tx, ty, tz, mx, my, mz, hx, hy, hz)), "llbar_local02nc_async");
    if (N == 0) return;
    hipLaunchKernelGGL(k_local02nc, hs::grid1d(N), hs::kBlock, 0, hs::S(stream), tx, ty, tz, mx, my, mz, hx, hy,
                       hz, msat0T0, uxx, uyy, uzz, mxx, myy, mzz, N);
    HS_LAUNCH_CHECK("llbar_local02nc_async");
}

// llbarLocal02cKern, llbar-local02c.cu:11-69: t = cross(mu.cross(m,H), m)
__global__ void k_local02c(double* tx, double* ty, double* tz, const double* mx, const double* my,
                           const double* mz, const double* hx, const double* hy, const double* hz,
                           const double* msk, const double* uxx, const double* uyy, const double* uzz, double mxx,
                           double myy, double mzz, int N) {
    LANE_LOOP(N)
    double ms = msk ? msk[i] : 1.0;
    if (ms == 0.0) {
        tx[i] = 0.0; ty[i] = 0.0; tz[i] = 0.0;
        return;
    }
    double ax = mx[i], ay = my[i], az = mz[i], bx = hx[i], by = hy[i], bz = hz[i];
    double cx = -ay * bz + az * by, cy = -az * bx + ax * bz, cz = -ax * by + ay * bx;
    double px = (uxx ? uxx[i] * mxx : mxx) * cx;
    double py = (uyy ? uyy[i] * myy : myy) * cy;
    double pz = (uzz ? uzz[i] * mzz : mzz) * cz;
    tx[i] = -py * az + pz * ay;
    ty[i] = -pz * ax + px * az;
    tz[i] = -px * ay + py * ax;
}
extern "C" void llbar_local02c_async(double* tx, double* ty, double* tz, double* mx, double* my, double* mz,
                                     double* hx, double* hy, double* hz, double* msat0T0, double* uxx,
                                     double* uyy, double* uzz, const double mxx, const double myy,
                                     const double mzz, CUstream stream, int N) {
    ARGCHECK(N >= 0 && (N == 0 || nn(tx, ty, tz, mx, my, mz, hx, hy, hz)), "llbar_local02c_async");
    if (N == 0) return;
    hipLaunchKernelGGL(k_local02c, hs::grid1d(N), hs::kBlock, 0, hs::S(stream), tx, ty, tz, mx, my, mz, hx, hy,
                       hz, msat0T0, uxx, uyy, uzz, mxx, myy, mzz, N);
    HS_LAUNCH_CHECK("llbar_local02c_async");
}

// long_field_Kern, long_field.cu:13-72
__global__ void k_long_field(double* hx, double* hy, double* hz, const double* mx, const double* my,
                             const double* mz, const double* msMsk, const double* SMsk, const double* nMsk,
                             const double* TcMsk, const double* TsMsk, double msMul, double SMul, double nMul,
                             double TcMul, double TsMul, int N) {
    LANE_LOOP(N)
    double Ms0T0 = msMul * mask1(msMsk, i);
    double S = SMul * mask1(SMsk, i);
    double n = nMul * mask1(nMsk, i);
    double Tc = TcMul * mask1(TcMsk, i);
    double Ts = TsMul * mask1(TsMsk, i);
    double m0 = mx[i], m1 = my[i], m2 = mz[i];
    double abs_mf = sqrt(m0 * m0 + m1 * m1 + m2 * m2);
    if (Ms0T0 == 0.0 || n == 0.0 || abs_mf <= HS_ZERO) {
        hx[i] = 0.0; hy[i] = 0.0; hz[i] = 0.0;
        return;
    }
    double inv = 1.0 / abs_mf;
    double preS = (S > HS_INFSPIN) ? 1.0 : S / (S + 1.0);
    double b = 3.0 * preS * Tc / Ts;
    double W, dW;
    brillouin_w_dw(S, abs_mf * b, W, dW);
    double mult = n * HS_KB * Ts * (W - abs_mf) / (HS_MU0 * Ms0T0 * dW);
    hx[i] = mult * (m0 * inv);
    hy[i] = mult * (m1 * inv);
    hz[i] = mult * (m2 * inv);
}
extern "C" void long_field_async(double* hx, double* hy, double* hz, double* mx, double* my, double* mz,
                                 double* msat0T0, double* S, double* n, double* Tc, double* Ts, double msat0T0Mul,
                                 double SMul, double nMul, double TcMul, double TsMul, int N, CUstream stream) {
    ARGCHECK(N >= 0 && (N == 0 || nn(hx, hy, hz, mx, my, mz)), "long_field_async");
    if (N == 0) return;
    hipLaunchKernelGGL(k_long_field, hs::grid1d(N), hs::kBlock, 0, hs::S(stream), hx, hy, hz, mx, my, mz,
                       msat0T0, S, n, Tc, Ts, msat0T0Mul, SMul, nMul, TcMul, TsMul, N);
    HS_LAUNCH_CHECK("long_field_async");
}

// temperature_scaleAnisKern, temperature.cu:14-71
__global__ void k_temp_scale(double* hx, double* hy, double* hz, const double* uxx, const double* uyy,
                             const double* uzz, const double* T, const double* msk, double mxx, double myy,
                             double mzz, double K, int N) {
    LANE_LOOP(N)
    double ms = mask1(msk, i);
    if (ms == 0.0) {
        hx[i] = 0.0; hy[i] = 0.0; hz[i] = 0.0;
        return;
    }
    double a = sqrt(mxx * mask1(uxx, i)), b = sqrt(myy * mask1(uyy, i)), c = sqrt(mzz * mask1(uzz, i));
    double ux = a * hx[i], uy = b * hy[i], uz = c * hz[i];
    double pre = sqrt((mask1(T, i) * K) / ms);
    hx[i] = pre * ux;
    hy[i] = pre * uy;
    hz[i] = pre * uz;
}
extern "C" void temperature_scaleAnizNoise(double* hx, double* hy, double* hz, double* uxx, double* uyy,
                                           double* uzz, double* T, double* msk, double mxx, double myy,
                                           double mzz, double K, CUstream stream, int N) {
    ARGCHECK(N >= 0 && (N == 0 || nn(hx, hy, hz)), "temperature_scaleAnizNoise");
    if (N == 0) return;
    hipLaunchKernelGGL(k_temp_scale, hs::grid1d(N), hs::kBlock, 0, hs::S(stream), hx, hy, hz, uxx, uyy, uzz, T,
                       msk, mxx, myy, mzz, K, N);
    HS_LAUNCH_CHECK("temperature_scaleAnizNoise");
}

// Model + findroot_Ridders + brillouinKern, brillouin.cu:13-138
__device__ double d_model(double n, double J, double pre) {
    double W, dW;
    brillouin_w_dw(J, pre * n, W, dW);
    return W - n;
}
__device__ double d_ridders(double J, double mult, double xa, double xb) {
    return ridders_root([=](double n) { return d_model(n, J, mult); }, xa, xb);
}
__global__ void k_brillouin(double* msat0, const double* msMsk, const double* T, const double* TcMsk,
                            const double* SMsk, double msat0Mul, double msMul, double TcMul, double SMul, int N) {
    LANE_LOOP(N)
    double Temp = T[i];
    double ms0T0 = msMul * mask1(msMsk, i);
    if (ms0T0 == 0.0) { msat0[i] = 0.0; return; }
    if (Temp == 0.0) { msat0[i] = ms0T0 / msat0Mul; return; }
    double Tc = TcMul * mask1(TcMsk, i);
    if (Temp > Tc) { msat0[i] = 0.0; return; }
    double S = SMul * mask1(SMsk, i);
    double preS = (S > HS_INFSPIN) ? 1.0 : S / (S + 1.0);
    double pre = 3.0 * preS * (Tc / Temp);
    double dT = (Tc - Temp) / Tc;
    double lo = (dT < 0.0004) ? -0.1 : 0.01, hi = (dT < 0.0004) ? 0.5 : 1.1;
    msat0[i] = ms0T0 * fabs(d_ridders(S, pre, lo, hi)) / msat0Mul;
}
extern "C" void brillouinAsync(double* msat0, double* msat0T0, double* T, double* Tc, double* S,
                               const double msat0Mul, const double msat0T0Mul, const double TcMul,
                               const double SMul, int N, CUstream stream) {
    ARGCHECK(N >= 0 && (N == 0 || nn(msat0, T)), "brillouinAsync");
    if (N == 0) return;
    hipLaunchKernelGGL(k_brillouin, hs::grid1d(N), hs::kBlock, 0, hs::S(stream), msat0, msat0T0, T, Tc, S,
                       msat0Mul, msat0T0Mul, TcMul, SMul, N);
    HS_LAUNCH_CHECK("brillouinAsync");
}

// normalizeKern, normalize.cu:13-34
__global__ void k_normalize(double* mx, double* my, double* mz, int N) {
    LANE_LOOP(N)
    double a = mx[i], b = my[i], c = mz[i];
    double n = sqrt(a * a + b * b + c * c);
    n = (n == 0.0) ? 0.0 : 1.0 / n;
    mx[i] = a * n; my[i] = b * n; mz[i] = c * n;
}
extern "C" void normalizeAsync(double* mx, double* my, double* mz, CUstream stream, int N) {
    ARGCHECK(N >= 0 && (N == 0 || nn(mx, my, mz)), "normalizeAsync");
    if (N == 0) return;
    hipLaunchKernelGGL(k_normalize, hs::grid1d(N), hs::kBlock, 0, hs::S(stream), mx, my, mz, N);
    HS_LAUNCH_CHECK("normalizeAsync");
}

// QinterKern, Qinter.cu:13-28
__global__ void k_qinter(double* Qi, const double* Ti, const double* Tj, const double* G, double GMul, int N) {
    LANE_LOOP(N)
    double g = G ? GMul * G[i] : GMul;
    Qi[i] = g * (Tj[i] - Ti[i]);
}
extern "C" void QinterAsync(double* Qi, double* Ti, double* Tj, double* Gij, double GijMul, int N, CUstream s) {
    ARGCHECK(N >= 0 && (N == 0 || nn(Qi, Ti, Tj)), "QinterAsync");
    if (N == 0) return;
    hipLaunchKernelGGL(k_qinter, hs::grid1d(N), hs::kBlock, 0, hs::S(s), Qi, Ti, Tj, Gij, GijMul, N);
    HS_LAUNCH_CHECK("QinterAsync");
}

// ------------------------------------------------------------------ add.cu, mul.cu, div.cu, dot.cu
__global__ void k_add(double* d, const double* a, const double* b, int N) { LANE_LOOP(N) d[i] = a[i] + b[i]; }
__global__ void k_lc2(double* d, const double* a, double ma, const double* b, double mb, int N) {
    LANE_LOOP(N) d[i] = ma * a[i] + mb * b[i];
}
__global__ void k_lc3(double* d, const double* a, double ma, const double* b, double mb, const double* c, double mc,
                      int N) {
    LANE_LOOP(N) d[i] = ma * a[i] + mb * b[i] + mc * c[i];
}
__global__ void k_addmadd(double* d, const double* a, const double* b, const double* c, double m, int N) {
    LANE_LOOP(N) d[i] = a[i] + m * (b[i] + c[i]);
}
__global__ void k_madd1(double* a, const double* b, double mb, int N) {
    LANE_LOOP(N) a[i] += mb * (b ? b[i] : 1.0);
}
__global__ void k_madd2(double* a, const double* b, double mb, const double* c, double mc, int N) {
    LANE_LOOP(N)
    double B = b ? b[i] : 1.0, C = c ? c[i] : 1.0;
    a[i] += mb * B + mc * C;
}
__global__ void k_madd(double* d, const double* a, const double* b, double mb, int N) {
    LANE_LOOP(N) d[i] = b ? a[i] + mb * b[i] : a[i] + mb;
}
__global__ void k_vecmadd(double* dx, double* dy, double* dz, const double* ax, const double* ay, const double* az,
                          const double* bx, const double* by, const double* bz, double mx, double my, double mz,
                          int N) {
    LANE_LOOP(N)
    if (bx) {
        dx[i] = ax[i] + mx * bx[i];
        dy[i] = ay[i] + my * by[i];
        dz[i] = az[i] + mz * bz[i];
    } else {
        dx[i] = ax[i] + mx;
        dy[i] = ay[i] + my;
        dz[i] = az[i] + mz;
    }
}
__global__ void k_cmadd(double* dst, double a, double b, const double* kern, const double* src, int N) {
    LANE_LOOP(N)
    double Sa = src[2 * i], Sb = src[2 * i + 1], kk = kern[i];
    dst[2 * i] += kk * (a * Sa - b * Sb);
    dst[2 * i + 1] += kk * (b * Sa + a * Sb);
}
__global__ void k_mul(double* d, const double* a, const double* b, int N) { LANE_LOOP(N) d[i] = a[i] * b[i]; }
__global__ void k_div(double* d, const double* a, const double* b, int N) {
    LANE_LOOP(N)
    double bb = b ? b[i] : 1.0;
    d[i] = (bb == 0.0) ? 0.0 : a[i] / bb;
}
__global__ void k_dotmask(double* d, const double* ax, const double* ay, const double* az, const double* bx,
                          const double* by, const double* bz, double axm, double aym, double azm, double bxm,
                          double bym, double bzm, int N) {
    LANE_LOOP(N)
    double a0 = axm * mask1(ax, i), a1 = aym * mask1(ay, i), a2 = azm * mask1(az, i);
    double b0 = bxm * mask1(bx, i), b1 = bym * mask1(by, i), b2 = bzm * mask1(bz, i);
    d[i] = a0 * b0 + a1 * b1 + a2 * b2;
}
__global__ void k_dot(double* d, const double* ax, const double* ay, const double* az, const double* bx,
                      const double* by, const double* bz, int N) {
    LANE_LOOP(N) d[i] = ax[i] * bx[i] + ay[i] * by[i] + az[i] * bz[i];
}
__global__ void k_dotsign(double* d, const double* ax, const double* ay, const double* az, const double* bx,
                          const double* by, const double* bz, const double* cx, const double* cy, const double* cz,
                          int N) {
    LANE_LOOP(N)
    double p = ax[i] * bx[i] + ay[i] * by[i] + az[i] * bz[i];
    double s = -(bx[i] * cx[i] + by[i] * cy[i] + bz[i] * cz[i]);
    d[i] = copysign(p, s);
}

#define L1(req, k, ...)                                                              \
    do {                                                                             \
        if (N <= 0) return;                                                          \
        ARGCHECK(req, #k ": NULL operand");                                          \
        hipLaunchKernelGGL(k, hs::grid1d(N), hs::kBlock, 0, hs::S(stream), __VA_ARGS__); \
        HS_LAUNCH_CHECK(#k);                                                         \
    } while (0)

extern "C" void addAsync(double* dst, double* a, double* b, CUstream stream, int N) { L1(nn(dst, a, b), k_add, dst, a, b, N); }
extern "C" void linearCombination2Async(double* dst, double* a, double ma, double* b, double mb, CUstream stream,
                                        int N) {
    L1(nn(dst, a, b), k_lc2, dst, a, ma, b, mb, N);
}
extern "C" void linearCombination3Async(double* dst, double* a, double ma, double* b, double mb, double* c,
                                        double mc, CUstream stream, int N) {
    L1(nn(dst, a, b, c), k_lc3, dst, a, ma, b, mb, c, mc, N);
}
extern "C" void addMaddAsync(double* dst, double* a, double* b, double* c, double m, CUstream stream, int N) {
    L1(nn(dst, a, b, c), k_addmadd, dst, a, b, c, m, N);
}
extern "C" void madd1Async(double* a, double* b, double mb, CUstream stream, int N) { L1(nn(a), k_madd1, a, b, mb, N); }
extern "C" void madd2Async(double* a, double* b, double mb, double* c, double mc, CUstream stream, int N) {
    L1(nn(a), k_madd2, a, b, mb, c, mc, N);
}
extern "C" void maddAsync(double* dst, double* a, double* b, double mb, CUstream stream, int N) {
    L1(nn(dst, a), k_madd, dst, a, b, mb, N);
}
extern "C" void vecMaddAsync(double* dx, double* dy, double* dz, double* ax, double* ay, double* az, double* bx,
                             double* by, double* bz, double mx, double my, double mz, CUstream stream, int N) {
    L1(nn(dx, dy, dz, ax, ay, az) && (!bx || nn(by, bz)), k_vecmadd, dx, dy, dz, ax, ay, az, bx, by, bz, mx, my, mz, N);
}
extern "C" void cmaddAsync(double* dst, double a, double b, double* kern, double* src, CUstream stream, int N) {
    L1(nn(dst, kern, src), k_cmadd, dst, a, b, kern, src, N);
}
extern "C" void mulAsync(double* dst, double* a, double* b, CUstream stream, int N) { L1(nn(dst, a, b), k_mul, dst, a, b, N); }
extern "C" void divAsync(double* dst, double* a, double* b, CUstream stream, int N) { L1(nn(dst, a), k_div, dst, a, b, N); }
extern "C" void dotMaskAsync(double* dst, double* ax, double* ay, double* az, double* bx, double* by, double* bz,
                             double axm, double aym, double azm, double bxm, double bym, double bzm, CUstream stream,
                             int N) {
    L1(nn(dst), k_dotmask, dst, ax, ay, az, bx, by, bz, axm, aym, azm, bxm, bym, bzm, N);
}
extern "C" void dotAsync(double* dst, double* ax, double* ay, double* az, double* bx, double* by, double* bz,
                         CUstream stream, int N) {
    L1(nn(dst, ax, ay, az, bx, by, bz), k_dot, dst, ax, ay, az, bx, by, bz, N);
}
extern "C" void dotSignAsync(double* dst, double* ax, double* ay, double* az, double* bx, double* by, double* bz,
                             double* cx, double* cy, double* cz, CUstream stream, int N) {
    L1(nn(dst, ax, ay, az, bx, by, bz, cx, cy, cz), k_dotsign, dst, ax, ay, az, bx, by, bz, cx, cy, cz, N);
}

// tensSYMMVecMulKern, mul.cu:31-90 (in place allowed: dst may alias src)
__global__ void k_tens(double* dX, double* dY, double* dZ, const double* sX, const double* sY, const double* sZ,
                       const double* kXX, const double* kYY, const double* kZZ, const double* kYZ,
                       const double* kXZ, const double* kXY, double srcMul, int Nx, int Ny, int Nz) {
    long n = (long)Nx * Ny * Nz;
    long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= n) return;
    int k = (int)(I % Nz);
    long r = I / Nz;
    int j = (int)(r % Ny);
    int i = (int)(r / Ny);
    long e = 2 * I;
    double RMX = srcMul * sX[e], IMX = srcMul * sX[e + 1];
    double RMY = srcMul * sY[e], IMY = srcMul * sY[e + 1];
    double RMZ = srcMul * sZ[e], IMZ = srcMul * sZ[e + 1];
    double sXY = 1.0, sXZ = 1.0, sYZ = 1.0;
    if (j > Ny / 2) { sXY = -sXY; sYZ = -sYZ; }
    int jj = (j > Ny / 2) ? Ny - j : j;
    if (i > Nx / 2) { sXY = -sXY; sXZ = -sXZ; }
    int ii = (i > Nx / 2) ? Nx - i : i;
    long K = ((long)ii * (Ny / 2 + 1) + jj) * Nz + k;
    double KXX = mask0(kXX, K), KXZ = sXZ * mask0(kXZ, K), KXY = sXY * mask0(kXY, K);
    double KYZ = sYZ * mask0(kYZ, K), KYY = mask0(kYY, K), KZZ = mask0(kZZ, K);
    double x0 = KXX * RMX + KXY * RMY + KXZ * RMZ, x1 = KXX * IMX + KXY * IMY + KXZ * IMZ;
    double y0 = KXY * RMX + KYY * RMY + KYZ * RMZ, y1 = KXY * IMX + KYY * IMY + KYZ * IMZ;
    double z0 = KXZ * RMX + KYZ * RMY + KZZ * RMZ, z1 = KXZ * IMX + KYZ * IMY + KZZ * IMZ;
    dX[e] = x0; dX[e + 1] = x1;
    dY[e] = y0; dY[e + 1] = y1;
    dZ[e] = z0; dZ[e + 1] = z1;
}
extern "C" void tensSYMMVecMul(double* dstX, double* dstY, double* dstZ, double* srcX, double* srcY, double* srcZ,
                               double* kXX, double* kYY, double* kZZ, double* kYZ, double* kXZ, double* kXY,
                               const double srcMul, const int Nx, const int Ny, const int Nz, CUstream stream) {
    ARGCHECK(Nx > 0 && Ny > 0 && Nz > 0 && nn(dstX, dstY, dstZ, srcX, srcY, srcZ), "tensSYMMVecMul");
    long n = (long)Nx * Ny * Nz;
    hipLaunchKernelGGL(k_tens, hs::grid1d(n), hs::kBlock, 0, hs::S(stream), dstX, dstY, dstZ, srcX, srcY, srcZ, kXX,
                       kYY, kZZ, kYZ, kXZ, kXY, srcMul, Nx, Ny, Nz);
    HS_LAUNCH_CHECK("tensSYMMVecMul");
}

// copyPad3DKern / copyUnPad3DKern / zeroArrayKern, copypad.cu:12-97 (all components in one launch)
__global__ void k_copy3d(double* dst, int D1, int D2, long dstComp, const double* src, int S1, int S2,
                         long srcComp, int E0, int E1, int E2) {
    long n = (long)E0 * E1 * E2;
    long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= n) return;
    int c = blockIdx.y;
    int k = (int)(I % E2);
    long r = I / E2;
    int j = (int)(r % E1);
    int i = (int)(r / E1);
    dst[c * dstComp + ((long)i * D1 + j) * D2 + k] = src[c * srcComp + ((long)i * S1 + j) * S2 + k];
}
extern "C" void copyPad3DAsync(double* dst, int D0, int D1, int D2, double* src, int S0, int S1, int S2, int Ncomp,
                               CUstream stream) {
    ARGCHECK(S0 <= D0 && S1 <= D1 && S2 <= D2 && S0 > 0 && S1 > 0 && S2 > 0 && Ncomp > 0 && nn(dst, src),
             "copyPad3DAsync");
    long n = (long)S0 * S1 * S2;
    dim3 g(hs::grid1d(n), Ncomp);
    hipLaunchKernelGGL(k_copy3d, g, hs::kBlock, 0, hs::S(stream), dst, D1, D2, (long)D0 * D1 * D2, src, S1, S2, n,
                       S0, S1, S2);
    HS_LAUNCH_CHECK("copyPad3DAsync");
}
extern "C" void copyUnPad3DAsync(double* dst, int D0, int D1, int D2, double* src, int S0, int S1, int S2,
                                 int Ncomp, CUstream stream) {
    ARGCHECK(D0 <= S0 && D1 <= S1 && D2 <= S2 && D0 > 0 && D1 > 0 && D2 > 0 && Ncomp > 0 && nn(dst, src),
             "copyUnPad3DAsync");
    long n = (long)D0 * D1 * D2;
    dim3 g(hs::grid1d(n), Ncomp);
    hipLaunchKernelGGL(k_copy3d, g, hs::kBlock, 0, hs::S(stream), dst, D1, D2, n, src, S1, S2,
                       (long)S0 * S1 * S2, D0, D1, D2);
    HS_LAUNCH_CHECK("copyUnPad3DAsync");
}
__global__ void k_zero(double* a, int N) { LANE_LOOP(N) a[i] = 0.0; }
extern "C" void zeroArrayAsync(double* A, int N, CUstream stream) { L1(nn(A), k_zero, A, N); }

// ------------------------------------------------------------------ partial reductions, reduce.cu
// Each of the caller's `blocks` blocks folds a grid-stride slice (wave64
// shuffles, then LDS across the 4 waves) and writes one partial.
enum RedOp { R_SUM, R_SDOT, R_MAX, R_MIN, R_MAXABS, R_MAXDIFF, R_MAXSUM, R_NORM3, R_NORM3DIFF };

template <int OP>
__device__ __forceinline__ double red_identity() {
    if (OP == R_SUM || OP == R_SDOT) return 0.0;
    if (OP == R_MAX) return -INFINITY;
    if (OP == R_MIN) return INFINITY;
    return 0.0;  // max of non-negative quantities
}
template <int OP>
__device__ __forceinline__ double red_combine(double a, double b) {
    if (OP == R_SUM || OP == R_SDOT) return a + b;
    if (OP == R_MIN) return fmin(a, b);
    return fmax(a, b);
}
template <int OP>
__global__ void __launch_bounds__(256) k_partial(const double* a, const double* b, const double* c,
                                                 const double* d, const double* e, const double* f, double* out,
                                                 long N) {
    double acc = red_identity<OP>();
    long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
        double v;
        if (OP == R_SUM || OP == R_MAX || OP == R_MIN) v = a[i];
        else if (OP == R_SDOT) v = a[i] * b[i];
        else if (OP == R_MAXABS) v = fabs(a[i]);
        else if (OP == R_MAXDIFF) v = fabs(a[i] - b[i]);
        else if (OP == R_MAXSUM) v = fabs(a[i] + b[i]);
        else if (OP == R_NORM3) v = a[i] * a[i] + b[i] * b[i] + c[i] * c[i];
        else {
            double x = a[i] - d[i], y = b[i] - e[i], z = c[i] - f[i];
            v = x * x + y * y + z * z;
        }
        acc = red_combine<OP>(acc, v);
    }
    for (int o = 32; o > 0; o >>= 1) acc = red_combine<OP>(acc, __shfl_xor(acc, o, 64));
    __shared__ double part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = part[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) r = red_combine<OP>(r, part[w]);
        out[blockIdx.x] = r;
    }
}
template <int OP>
static void launch_partial(const char* name, const double* a, const double* b, const double* c, const double* d,
                           const double* e, const double* f, double* out, int blocks, int N, CUstream stream) {
    const int nops = (OP == R_NORM3DIFF) ? 6 : (OP == R_NORM3) ? 3 : (OP == R_SDOT || OP == R_MAXDIFF || OP == R_MAXSUM) ? 2 : 1;
    const double* ops[6] = {a, b, c, d, e, f};
    bool ok = blocks > 0 && N >= 0 && out;
    for (int i = 0; i < nops && N > 0; i++) ok = ok && ops[i];
    ARGCHECK(ok, name);
    hipLaunchKernelGGL(k_partial<OP>, blocks, 256, 0, hs::S(stream), a, b, c, d, e, f, out, (long)N);
    HS_LAUNCH_CHECK(name);
}
extern "C" void partialSumAsync(double* in, double* out, int blocks, int thr, int N, CUstream s) {
    (void)thr;
    launch_partial<R_SUM>("partialSumAsync", in, 0, 0, 0, 0, 0, out, blocks, N, s);
}
extern "C" void partialSDotAsync(double* a, double* b, double* out, int blocks, int thr, int N, CUstream s) {
    (void)thr;
    launch_partial<R_SDOT>("partialSDotAsync", a, b, 0, 0, 0, 0, out, blocks, N, s);
}
extern "C" void partialMaxAsync(double* in, double* out, int blocks, int thr, int N, CUstream s) {
    (void)thr;
    launch_partial<R_MAX>("partialMaxAsync", in, 0, 0, 0, 0, 0, out, blocks, N, s);
}
extern "C" void partialMinAsync(double* in, double* out, int blocks, int thr, int N, CUstream s) {
    (void)thr;
    launch_partial<R_MIN>("partialMinAsync", in, 0, 0, 0, 0, 0, out, blocks, N, s);
}
extern "C" void partialMaxAbsAsync(double* in, double* out, int blocks, int thr, int N, CUstream s) {
    (void)thr;
    launch_partial<R_MAXABS>("partialMaxAbsAsync", in, 0, 0, 0, 0, 0, out, blocks, N, s);
}
extern "C" void partialMaxDiffAsync(double* a, double* b, double* out, int blocks, int thr, int N, CUstream s) {
    (void)thr;
    launch_partial<R_MAXDIFF>("partialMaxDiffAsync", a, b, 0, 0, 0, 0, out, blocks, N, s);
}
extern "C" void partialMaxSumAsync(double* a, double* b, double* out, int blocks, int thr, int N, CUstream s) {
    (void)thr;
    launch_partial<R_MAXSUM>("partialMaxSumAsync", a, b, 0, 0, 0, 0, out, blocks, N, s);
}
extern "C" void partialMaxNorm3SqAsync(double* x, double* y, double* z, double* out, int blocks, int thr, int N,
                                       CUstream s) {
    (void)thr;
    launch_partial<R_NORM3>("partialMaxNorm3SqAsync", x, y, z, 0, 0, 0, out, blocks, N, s);
}
extern "C" void partialMaxNorm3SqDiffAsync(double* x1, double* y1, double* z1, double* x2, double* y2, double* z2,
                                           double* out, int blocks, int thr, int N, CUstream s) {
    (void)thr;
    launch_partial<R_NORM3DIFF>("partialMaxNorm3SqDiffAsync", x1, y1, z1, x2, y2, z2, out, blocks, N, s);
}
