// hs_common.hpp -- shared host/device helpers of libhotspin_amd (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/libhotspin.h"
#include "../../include/hotspin_amd.h"

// ---- error state (hs_last_error) -------------------------------------------
namespace hs {
void set_error(int code, const char* what);
void check_launch(const char* what);
inline hipStream_t S(CUstream s) { return reinterpret_cast<hipStream_t>(s); }
// per-kernel event timing (hs_prof_enable); categories HS_PROF_* of hotspin_amd.h
bool prof_on();
void prof_begin(int cat, hipStream_t st);
void prof_end(int cat, hipStream_t st);
inline int div_up(long a, long b) { return (int)((a + b - 1) / b); }
constexpr int kBlock = 256;
inline unsigned grid1d(long n) {
    long g = (n + kBlock - 1) / kBlock;
    return (unsigned)(g > 0 ? g : 1);
}
}  // namespace hs

#define HS_LAUNCH_CHECK(name) hs::check_launch(name)

// ---- device constants: src/libhotspin/common_func.h:24-33 --------------------
#define HS_MU0 (4.0e-7 * 3.14159265358979323846)
#define HS_KB 1.38064881313131313E-23
#define HS_ZERO 1.0e-32
#define HS_EPS 1.0e-8
#define HS_LINRANGE 2.0e-1
#define HS_INFSPIN 1.0e5

template <typename T>
__device__ __forceinline__ T mask1(const T* m, long i) { return m ? m[i] : T(1); }
template <typename T>
__device__ __forceinline__ T mask0(const T* m, long i) { return m ? m[i] : T(0); }

// avgGeomZero, common_func.h:155-160
template <typename T>
__device__ __forceinline__ T avg_geom_zero(T a, T b) {
    T s = a + b;
    T v = T(2) * a * b / s;
    return (s == T(0)) ? T(0) : v;
}

// Brillouin B_S(x) and dB_S/dx with the series branch below linRange
// (common_func.h:68-118); Langevin L(x), dL/dx (common_func.h:120-136).
template <typename T>
__device__ __forceinline__ void brillouin_w_dw(T S, T x, T& W, T& dW) {
    if (S > T(HS_INFSPIN)) {
        if (fabs(x) < T(HS_LINRANGE)) {
            T x2 = x * x;
            W = (x / T(3)) - (x2 * x / T(45)) + (x2 * x2 * x * T(2) / T(945)) - (x2 * x2 * x2 * x / T(4725));
            dW = (T(1) / T(3)) - (x2 / T(15)) + (x2 * x2 * T(2) / T(189)) - (x2 * x2 * x2 / T(675));
        } else {
            T c = T(1) / tanh(x);
            W = c - T(1) / x;
            dW = T(1) - c * c + T(1) / (x * x);
        }
        return;
    }
    T lpre = T(1) / (T(2) * S);
    T gpre = (T(2) * S + T(1)) * lpre;
    T gpre2 = gpre * gpre, lpre2 = lpre * lpre;
    T lim = fmin(T(HS_LINRANGE) / gpre, T(HS_LINRANGE) / lpre);
    if (fabs(x) < lim) {
        T lpre4 = lpre2 * lpre2, gpre4 = gpre2 * gpre2;
        T lpre6 = lpre4 * lpre2, gpre6 = gpre4 * gpre2;
        T lpre8 = lpre4 * lpre4, gpre8 = gpre4 * gpre4;
        T x2 = x * x;
        W = ((gpre2 - lpre2) * x / T(3)) - ((gpre4 - lpre4) * x2 * x / T(45)) +
            ((gpre6 - lpre6) * x2 * x2 * x * T(2) / T(945)) - ((gpre8 - lpre8) * x2 * x2 * x2 * x / T(4725));
        dW = ((gpre2 - lpre2) / T(3)) - ((gpre4 - lpre4) * x2 / T(15)) +
             ((gpre6 - lpre6) * x2 * x2 * T(2) / T(189)) - ((gpre8 - lpre8) * x2 * x2 * x2 / T(675));
    } else {
        T cg = T(1) / tanh(gpre * x), cl = T(1) / tanh(lpre * x);
        W = gpre * cg - lpre * cl;
        dW = (gpre2 - lpre2) + lpre2 * cl * cl - gpre2 * cg * cg;
    }
}

// Ridders' root bracketing as the reference's findroot_Ridders and
// findroot_Ridders_Ts (brillouin.cu:13-76, Ts.cu:13-64) run it: the same
// update sequence, |f| < HS_ZERO at an end point returns it, and the
// iteration stops when the relative change of x drops below HS_EPS or after
// 1000 iterations.  f: the model as a device callable of x.
template <typename F>
__device__ double ridders_root(F f, double xa, double xb) {
    auto sgn = [](double x) { return x == 0.0 ? 0.0 : (signbit(x) == 0 ? 1.0 : -1.0); };
    double ya = f(xa);
    if (fabs(ya) < HS_ZERO) return xa;
    double yb = f(xb);
    if (fabs(yb) < HS_ZERO) return xb;
    double y1 = ya, x1 = xa, y2 = yb, x2 = xb;
    double x = 1.0e10, tx = x, teps = x;
    int iter = 0;
    while (teps > HS_EPS && iter < 1000) {
        double x3 = 0.5 * (x2 + x1);
        double y3 = f(x3);
        double dy = (y3 * y3 - y1 * y2);
        if (dy == 0.0) {
            x = x3;
            break;
        }
        double dx = (x3 - x1) * sgn(y1 - y2) * y3 / sqrt(dy);
        x = x3 + dx;
        double y = f(x);
        bool sy = signbit(y) != 0, sy3 = signbit(y3) != 0, sy1 = signbit(y1) != 0;
        double ny2 = (sy == sy3) ? y2 : y3, nx2 = (sy == sy3) ? x2 : x3;
        y2 = ny2;
        x2 = nx2;
        ny2 = (sy == sy1 || x2 == x3) ? y2 : y1;
        nx2 = (sy == sy1 || x2 == x3) ? x2 : x1;
        y2 = ny2;
        x2 = nx2;
        y1 = y;
        x1 = x;
        teps = fabs((x - tx) / (tx + x));
        tx = x;
        iter++;
    }
    return x;
}
