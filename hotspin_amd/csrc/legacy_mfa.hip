// legacy_mfa.hip -- the mean-field (MFA) helpers, anisotropy, nonlocal
// damping and the remaining element-wise calls of the FP64 libhotspin ABI
// (include/libhotspin.h) on gfx950.
//
// One lane per cell over SoA planes, 256-lane blocks; the nonlocal stencil
// maps 64 lanes along the contiguous k axis and 4 rows along j, like the
// exchange and heat stencils.  Formulas restate the cited reference kernels.
#include <math.h>

#include "hs_common.hpp"

#define ARGCHECK(cond, name)                 \
    do {                                     \
        if (!(cond)) {                       \
            hs::set_error(HS_EINVAL, name);  \
            return;                          \
        }                                    \
    } while (0)

#define LANE(N)                                     \
    int i = blockIdx.x * blockDim.x + threadIdx.x; \
    if (i >= (N)) return;

#define LAUNCH1(k, name, ...)                                                                  \
    do {                                                                                       \
        ARGCHECK(N >= 0, name);                                                                \
        if (N == 0) return;                                                                    \
        hipLaunchKernelGGL(k, hs::grid1d(N), hs::kBlock, 0, hs::S(stream), __VA_ARGS__);      \
        HS_LAUNCH_CHECK(name);                                                                 \
    } while (0)

namespace {

constexpr double kPi = 3.14159265358979323846;

// Brillouin B_S(x) alone (no Langevin switch), as ModelTs calls it (Ts.cu:77-82).
__device__ double bj_only(double S, double x) {
    double lpre = 1.0 / (2.0 * S);
    double gpre = (2.0 * S + 1.0) * lpre;
    double gpre2 = gpre * gpre, lpre2 = lpre * lpre;
    double lim = fmin(HS_LINRANGE / gpre, HS_LINRANGE / lpre);
    if (fabs(x) < lim) {
        double lpre4 = lpre2 * lpre2, gpre4 = gpre2 * gpre2;
        double lpre6 = lpre4 * lpre2, gpre6 = gpre4 * gpre2;
        double lpre8 = lpre4 * lpre4, gpre8 = gpre4 * gpre4;
        double x2 = x * x;
        return ((gpre2 - lpre2) * x / 3.0) - ((gpre4 - lpre4) * x2 * x / 45.0) +
               ((gpre6 - lpre6) * x2 * x2 * x * 2.0 / 945.0) - ((gpre8 - lpre8) * x2 * x2 * x2 * x / 4725.0);
    }
    return gpre / tanh(gpre * x) - lpre / tanh(lpre * x);
}

// findroot_Ridders_Ts with ModelTs (Ts.cu:13-82): the temperature T in [0, Tc]
// with B_S(pre/T) = m.  Same update sequence as brillouinAsync's root finder.
__device__ double ts_model(double T, double S, double pre, double m) {
    double x = (T == 0.0) ? 1.0e38 : pre / T;
    return bj_only(S, x) - m;
}
__device__ double ts_root(double S, double pre, double m, double xa, double xb) {
    return ridders_root([=](double T) { return ts_model(T, S, pre, m); }, xa, xb);
}

// Debye integrand x^4 e^x / (e^x - 1)^2, 0 at x <= 1e-32 (common_func.h:168-173).
__device__ double debye(double x) {
    double e = exp(x);
    double em1 = e - 1.0;
    return (x <= HS_ZERO) ? 0.0 : (x * x * x * x * e) / (em1 * em1);
}
constexpr int kDebyeSteps = 61;  // INTMAXSTEPS, common_func.h:27

// kappaKern, kappa.cu:15-61.  n is its mask alone (no multiplier).
__global__ void k_kappa(double* kappa, const double* ms0Msk, const double* ms0T0Msk, const double* T,
                        const double* TcMsk, const double* SMsk, const double* nMsk, double ms0Mul,
                        double ms0T0Mul, double TcMul, double SMul, int N) {
    LANE(N)
    double u = mask1(ms0T0Msk, i);
    double ms0T0 = ms0T0Mul * u;
    double Temp = T[i];
    if (ms0T0 == 0.0 || Temp == 0.0) {
        kappa[i] = 0.0;
        return;
    }
    double S = SMul * mask1(SMsk, i);
    double Tc = TcMul * mask1(TcMsk, i);
    double ms0 = ms0Mul * mask1(ms0Msk, i);
    double preS = (S > HS_INFSPIN) ? 1.0 : S / (S + 1.0);
    double J0 = 3.0 * Tc;
    double n = mask1(nMsk, i);
    double mul = u * u / (preS * J0 * n);
    double me = ms0 / ms0T0;
    double b = preS * J0 / Temp;
    double W, dW;
    brillouin_w_dw(S, me * b, W, dW);
    double f = b * dW;
    kappa[i] = mul * (f / (1.0 - f));
}

// tsKern, Ts.cu:84-120
__global__ void k_ts(double* Ts, const double* msMsk, const double* ms0T0Msk, const double* TcMsk,
                     const double* SMsk, double msMul, double ms0T0Mul, double TcMul, double SMul, int N) {
    LANE(N)
    double ms0T0 = ms0T0Mul * mask1(ms0T0Msk, i);
    if (ms0T0 == 0.0) {
        Ts[i] = 0.0;
        return;
    }
    double ms = msMul * mask1(msMsk, i);
    if (ms == ms0T0) {
        Ts[i] = 0.0;
        return;
    }
    double Tc = TcMul * mask1(TcMsk, i);
    if (ms == 0.0) {
        Ts[i] = Tc;
        return;
    }
    double S = SMul * mask1(SMsk, i);
    double J0 = 3.0 * Tc / (S * (S + 1.0));
    double m = ms / ms0T0;
    double pre = S * S * J0 * m;
    Ts[i] = ts_root(S, pre, m, 0.0, Tc);
}

// cpKern, Cp.cu:14-83: n * (T/Td)^3 * int_0^{Td/T} x^4 e^x/(e^x-1)^2 dx by the
// 61-interval trapezoid rule; below T/Td = 0.01 the integral's T^3 limit
// (4 pi^4 / 15).  9 n kB stays in the caller's multiplier.
__global__ void k_cp(double* Cp, const double* T, const double* TdMsk, const double* nMsk, double TdMul, int N) {
    LANE(N)
    double n = mask1(nMsk, i);
    double Td = TdMul * mask1(TdMsk, i);
    double Temp = T[i];
    if (Td == 0.0) {
        Cp[i] = 0.0;
        return;
    }
    double xx = Td / Temp, ixx = Temp / Td;
    if (ixx < 0.01) {
        Cp[i] = n * (12.0 / 5.0) * (1.0 / 9.0) * (kPi * kPi * kPi * kPi) * (ixx * ixx * ixx);
        return;
    }
    double h = xx / (double)kDebyeSteps;
    double inner = 0.0;
    for (int j = 1; j < kDebyeSteps; j++) inner += debye(j * h);
    double v = (debye(0.0) + debye(xx) + 2.0 * inner) * (0.5 * h);
    Cp[i] = n * v / (xx * xx * xx);
}

// energyFlowKern, energy_flow.cu:15-48.  Tc and n are masks alone.
__global__ void k_energy_flow(double* w, const double* mx, const double* my, const double* mz, const double* Rx,
                              const double* Ry, const double* Rz, const double* TcMsk, const double* SMsk,
                              const double* nMsk, double SMul, int N) {
    LANE(N)
    double S = SMul * mask1(SMsk, i);
    double Tc = mask1(TcMsk, i);
    double n = mask1(nMsk, i);
    if (n == 0.0 || S == 0.0 || Tc == 0.0) {
        w[i] = 0.0;
        return;
    }
    double rW = mx[i] * Rx[i] + my[i] * Ry[i] + mz[i] * Rz[i];
    double mult = Tc * S * S * n / (S * (S + 1.0));
    w[i] = mult * rW;
}

// uniaxialAnisotropyKern, uniaxialanisotropy.cu:12-52: the projection m.u
// uses the normalized axis, the field direction the axis as given.
__global__ void k_uniaxial(double* hx, double* hy, double* hz, const double* mx, const double* my,
                           const double* mz, const double* KuMsk, const double* msMsk, double K2Mul,
                           const double* uxMsk, double uxMul, const double* uyMsk, double uyMul,
                           const double* uzMsk, double uzMul, int N) {
    LANE(N)
    double ms = mask1(msMsk, i);
    if (ms == 0.0) {
        hx[i] = 0.0; hy[i] = 0.0; hz[i] = 0.0;
        return;
    }
    double K = (K2Mul / ms) * mask1(KuMsk, i);
    double ux = uxMul * mask1(uxMsk, i), uy = uyMul * mask1(uyMsk, i), uz = uzMul * mask1(uzMsk, i);
    double len = sqrt(ux * ux + uy * uy + uz * uz);
    double inv = (len != 0.0) ? 1.0 / len : 0.0;
    double mu = mx[i] * (ux * inv) + my[i] * (uy * inv) + mz[i] * (uz * inv);
    hx[i] = K * mu * ux;
    hy[i] = K * mu * uy;
    hz[i] = K * mu * uz;
}

// llbarNonlocal00ncKern, llbar-nonlocal00nc.cu:11-208: t = -(1/mu0) sum_d
// (lambda_e mul_c / c_d^2) sum_+- avgGeom(l0, l+-) (mu0 H+- - mu0 H0) per
// component c; neighbour order X(i), Z(k), Y(j); clamp or wrap at the edges.
__global__ void __launch_bounds__(256) k_nonlocal(double* __restrict__ tx, double* __restrict__ ty,
                                                  double* __restrict__ tz, const double* __restrict__ hx,
                                                  const double* __restrict__ hy, const double* __restrict__ hz,
                                                  const double* __restrict__ msk, const double* __restrict__ lxx,
                                                  const double* __restrict__ lyy, const double* __restrict__ lzz,
                                                  double lmx, double lmy, double lmz, int s0, int s1, int s2,
                                                  double c20, double c21, double c22, int w0, int w1, int w2) {
    int k = blockIdx.x * 64 + threadIdx.x;
    int j = blockIdx.y * 4 + threadIdx.y;
    int i = blockIdx.z;
    if (k >= s2 || j >= s1) return;
    long s12 = (long)s1 * s2;
    long I = i * s12 + (long)j * s2 + k;
    if (mask1(msk, I) == 0.0) {
        tx[I] = 0.0; ty[I] = 0.0; tz[I] = 0.0;
        return;
    }
    const double* lam[3] = {lxx, lyy, lzz};
    const double* H[3] = {hx, hy, hz};
    const double lm[3] = {lmx, lmy, lmz};
    double l0[3], H0[3], R[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < 3; c++) {
        l0[c] = mask1(lam[c], I);
        H0[c] = HS_MU0 * H[c][I];
    }
    const int ax_order[3] = {0, 2, 1};
    const double c2[3] = {c20, c21, c22};
#pragma unroll
    for (int q = 0; q < 3; q++) {
        int ax = ax_order[q];
        int n = ax == 0 ? s0 : (ax == 1 ? s1 : s2);
        int w = ax == 0 ? w0 : (ax == 1 ? w1 : w2);
        int cc = ax == 0 ? i : (ax == 1 ? j : k);
        long stride = ax == 0 ? s12 : (ax == 1 ? (long)s2 : 1L);
        int lo = cc - 1;
        lo = (lo < 0 && w) ? n + lo : lo;
        lo = max(lo, 0);
        int hi = cc + 1;
        hi = (hi == n && w) ? hi - n : hi;
        hi = min(hi, n - 1);
        long a1 = I + (long)(lo - cc) * stride, a2 = I + (long)(hi - cc) * stride;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            double p1 = avg_geom_zero(l0[c], mask1(lam[c], a1));
            double p2 = avg_geom_zero(l0[c], mask1(lam[c], a2));
            double H1 = HS_MU0 * H[c][a1], H2 = HS_MU0 * H[c][a2];
            R[c] += (lm[c] * c2[ax]) * ((p1 * H1 + p2 * H2) - H0[c] * (p1 + p2));
        }
    }
    double pre = 1.0 / HS_MU0;
    tx[I] = -pre * R[0];
    ty[I] = -pre * R[1];
    tz[I] = -pre * R[2];
}

// divMulPowKern, divMulPow.cu:13-30 (powf on the values rounded to float)
__global__ void k_div_mul_pow(double* dst, const double* a, const double* b, const double* c, double p, int N) {
    LANE(N)
    double bb = b ? b[i] : 1.0;
    double val = (bb == 0.0) ? 0.0 : a[i] / bb;
    double cc = c ? c[i] : 1.0;
    cc = (cc == 0.0) ? 0.0 : (double)powf((float)cc, (float)p);
    dst[i] = val * cc;
}

// wavgKern + weightedAvgZero, wavg.cu:13-30, common_func.h:162-166
__global__ void k_wavg(double* dst, const double* a, const double* b, const double* w0Msk, const double* w1Msk,
                       const double* RMsk, double w0Mul, double w1Mul, double RMul, int N) {
    LANE(N)
    double w0 = w0Mul * mask1(w0Msk, i), w1 = w1Mul * mask1(w1Msk, i), R = RMul * mask1(RMsk, i);
    double x0 = a[i], x1 = b[i];
    double den = w0 + w1 + 2.0 * R * sqrt(w0 * w1);
    dst[i] = (den == 0.0) ? 0.0 : (w0 * x0 + w1 * x1 + 2.0 * R * sqrt(w0 * x0 * w1 * x1)) / den;
}

}  // namespace

extern "C" void kappaAsync(double* kappa, double* msat0, double* msat0T0, double* T, double* Tc, double* S,
                           double* n, const double msat0Mul, const double msat0T0Mul, const double TcMul,
                           const double SMul, int N, CUstream stream) {
    ARGCHECK(kappa && T, "kappaAsync");
    LAUNCH1(k_kappa, "kappaAsync", kappa, msat0, msat0T0, T, Tc, S, n, msat0Mul, msat0T0Mul, TcMul, SMul, N);
}

extern "C" void tsAsync(double* Ts, double* msat, double* msat0T0, double* Tc, double* S, const double msatMul,
                        const double msat0T0Mul, const double TcMul, const double SMul, int N, CUstream stream) {
    ARGCHECK(Ts != nullptr, "tsAsync");
    LAUNCH1(k_ts, "tsAsync", Ts, msat, msat0T0, Tc, S, msatMul, msat0T0Mul, TcMul, SMul, N);
}

extern "C" void cpAsync(double* Cp, double* T, double* Td, double* n, const double TdMul, const int N,
                        CUstream stream) {
    ARGCHECK(Cp && T, "cpAsync");
    LAUNCH1(k_cp, "cpAsync", Cp, T, Td, n, TdMul, N);
}

extern "C" void energyFlowAsync(double* w, double* mx, double* my, double* mz, double* Rx, double* Ry, double* Rz,
                                double* Tc, double* S, double* n, double SMul, int N, CUstream stream) {
    ARGCHECK(w && mx && my && mz && Rx && Ry && Rz, "energyFlowAsync");
    LAUNCH1(k_energy_flow, "energyFlowAsync", w, mx, my, mz, Rx, Ry, Rz, Tc, S, n, SMul, N);
}

extern "C" void uniaxialAnisotropyAsync(double* hx, double* hy, double* hz, double* mx, double* my, double* mz,
                                        double* Ku1_map, double* MSat_map, double Ku2_Mu0Msat_mul,
                                        double* anisU_mapx, double anisU_mulx, double* anisU_mapy,
                                        double anisU_muly, double* anisU_mapz, double anisU_mulz, CUstream stream,
                                        int N) {
    ARGCHECK(hx && hy && hz && mx && my && mz, "uniaxialAnisotropyAsync");
    LAUNCH1(k_uniaxial, "uniaxialAnisotropyAsync", hx, hy, hz, mx, my, mz, Ku1_map, MSat_map, Ku2_Mu0Msat_mul,
            anisU_mapx, anisU_mulx, anisU_mapy, anisU_muly, anisU_mapz, anisU_mulz, N);
}

extern "C" void llbar_nonlocal00nc_async(double* tx, double* ty, double* tz, double* hx, double* hy, double* hz,
                                         double* msat0T0, double* lambda_e_xx, double* lambda_e_yy,
                                         double* lambda_e_zz, const double lambda_eMul_xx,
                                         const double lambda_eMul_yy, const double lambda_eMul_zz, const int sx,
                                         const int sy, const int sz, const double csx, const double csy,
                                         const double csz, const int pbc_x, const int pbc_y, const int pbc_z,
                                         CUstream stream) {
    ARGCHECK(sx > 0 && sy > 0 && sz > 0 && tx && ty && tz && hx && hy && hz, "llbar_nonlocal00nc_async");
    dim3 blk(64, 4), grd(hs::div_up(sz, 64), hs::div_up(sy, 4), sx);
    hipLaunchKernelGGL(k_nonlocal, grd, blk, 0, hs::S(stream), tx, ty, tz, hx, hy, hz, msat0T0, lambda_e_xx,
                       lambda_e_yy, lambda_e_zz, lambda_eMul_xx, lambda_eMul_yy, lambda_eMul_zz, sx, sy, sz,
                       1.0 / (csx * csx), 1.0 / (csy * csy), 1.0 / (csz * csz), pbc_x, pbc_y, pbc_z);
    HS_LAUNCH_CHECK("llbar_nonlocal00nc_async");
}

extern "C" void divMulPowAsync(double* dst, double* a, double* b, double* c, double p, CUstream stream, int N) {
    ARGCHECK(dst && a, "divMulPowAsync");
    LAUNCH1(k_div_mul_pow, "divMulPowAsync", dst, a, b, c, p, N);
}

extern "C" void wavgAsync(double* dst, double* a, double* b, double* w0, double* w1, double* R, double w0Mul,
                          double w1Mul, double RMul, CUstream stream, int N) {
    ARGCHECK(dst && a && b, "wavgAsync");
    LAUNCH1(k_wavg, "wavgAsync", dst, a, b, w0, w1, R, w0Mul, w1Mul, RMul, N);
}
