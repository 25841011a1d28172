/*
 * hs_oracle.c -- FP64 CPU restatement of hotspin's per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see hs_oracle.h).  Each function restates the
 * reference formula it cites, in the same floating-point operation order
 * where the reference fixes one.  Build: oracle/build.py (gcc -O2
 * -ffp-contract=off -fopenmp).
 */
#include "hs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_PI 3.14159265358979323846

/* thresholds, src/libhotspin/common_func.h:24-28 */
#define OR_ZERO 1.0e-32
#define OR_EPS 1.0e-8
#define OR_LINRANGE 2.0e-1
#define OR_INFSPIN 1.0e5

void or_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
int or_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------ helpers */
/* getMaskUnity / getMaskZero / avgGeomZero: common_func.h:140-160 */
static inline double mask1(const double* m, long i) { return m ? m[i] : 1.0; }
static inline double mask0(const double* m, long i) { return m ? m[i] : 0.0; }
static inline double avg_geom_zero(double a, double b) {
    double a_b = a + b;
    double val = 2.0 * a * b / a_b;
    return (a_b == 0.0) ? 0.0 : val;
}
static inline double coth(double x) { return 1.0 / tanh(x); }

/* Brillouin function and derivative, common_func.h:68-118 */
static double Bj(double J, double x) {
    double lpre = 1.0 / (2.0 * J);
    double gpre = (2.0 * J + 1.0) * lpre;
    double gpre2 = gpre * gpre, lpre2 = lpre * lpre;
    double lpre4 = lpre2 * lpre2, gpre4 = gpre2 * gpre2;
    double lpre6 = lpre4 * lpre2, gpre6 = gpre4 * gpre2;
    double lpre8 = lpre4 * lpre4, gpre8 = gpre4 * gpre4;
    double limA = OR_LINRANGE / gpre, limB = OR_LINRANGE / lpre;
    double lim = fmin(limA, limB);
    if (fabs(x) < lim)
        return ((gpre2 - lpre2) * x / 3.0) - ((gpre4 - lpre4) * x * x * x / 45.0) +
               ((gpre6 - lpre6) * x * x * x * x * x * 2.0 / 945.0) -
               ((gpre8 - lpre8) * x * x * x * x * x * x * x / 4725.0);
    return gpre * coth(gpre * x) - lpre * coth(lpre * x);
}
static double dBjdx(double J, double x) {
    double lpre = 1.0 / (2.0 * J);
    double gpre = (2.0 * J + 1.0) * lpre;
    double gpre2 = gpre * gpre, lpre2 = lpre * lpre;
    double lpre4 = lpre2 * lpre2, gpre4 = gpre2 * gpre2;
    double lpre6 = lpre4 * lpre2, gpre6 = gpre4 * gpre2;
    double lpre8 = lpre4 * lpre4, gpre8 = gpre4 * gpre4;
    double limA = OR_LINRANGE / gpre, limB = OR_LINRANGE / lpre;
    double lim = fmin(limA, limB);
    if (fabs(x) < lim)
        return ((gpre2 - lpre2) / 3.0) - ((gpre4 - lpre4) * x * x / 15.0) +
               ((gpre6 - lpre6) * x * x * x * x * 2.0 / 189.0) -
               ((gpre8 - lpre8) * x * x * x * x * x * x / 675.0);
    return (gpre2 - lpre2) + lpre2 * coth(lpre * x) * coth(lpre * x) -
           gpre2 * coth(gpre * x) * coth(gpre * x);
}
/* Langevin function and derivative, common_func.h:120-136 */
static double Lf(double x) {
    if (fabs(x) < OR_LINRANGE)
        return (x / 3.0) - (x * x * x / 45.0) + (x * x * x * x * x * 2.0 / 945.0) -
               (x * x * x * x * x * x * x / 4725.0);
    return coth(x) - (1.0 / x);
}
static double dLdx(double x) {
    if (fabs(x) < OR_LINRANGE)
        return (1 / 3.0) - (x * x / 15.0) + (x * x * x * x * 2.0 / 189.0) -
               (x * x * x * x * x * x / 675.0);
    return 1.0 - (coth(x) * coth(x)) + (1.0 / (x * x));
}
static inline double sgn(double x) {
    double val = (signbit(x) == 0) ? 1.0 : -1.0;
    return (x == 0.0) ? 0.0 : val;
}

/* ------------------------------------------------------------------ exchange */
/* exchange6Kern, src/libhotspin/exchange6.cu:11-127.  Neighbour order X(i), Z(k), Y(j);
 * open boundary = clamped index, periodic = wrapped.  The msat0T0 mask of
 * exchange6Async is never passed to the kernel (exchange6.cu:134-147). */
void or_exchange6(double* hx, double* hy, double* hz,
                  const double* mx, const double* my, const double* mz,
                  const double* lexMsk, int N0, int N1, int N2,
                  int w0, int w1, int w2, double msat0T0Mul,
                  double p0, double p1, double p2) {
    long N12 = (long)N1 * N2;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < N0; i++)
        for (int j = 0; j < N1; j++)
            for (int k = 0; k < N2; k++) {
                long I = i * N12 + (long)j * N2 + k;
                double lex = mask1(lexMsk, I);
                double lex02 = lex * lex, lex2, pre1, pre2;
                double m0x = mx[I], m0y = my[I], m0z = mz[I];
                double Hx, Hy, Hz;
                long a;
                int idx;
                double m1x, m1y, m1z, m2x, m2y, m2z;
                /* X (axis 0) */
                idx = i - 1; idx = (idx < 0 && w0) ? N0 + idx : idx; if (idx < 0) idx = 0;
                a = idx * N12 + (long)j * N2 + k;
                m1x = mx[a]; m1y = my[a]; m1z = mz[a];
                lex = mask1(lexMsk, a); lex2 = lex * lex; pre1 = avg_geom_zero(lex02, lex2);
                idx = i + 1; idx = (idx == N0 && w0) ? idx - N0 : idx; if (idx > N0 - 1) idx = N0 - 1;
                a = idx * N12 + (long)j * N2 + k;
                m2x = mx[a]; m2y = my[a]; m2z = mz[a];
                lex = mask1(lexMsk, a); lex2 = lex * lex; pre2 = avg_geom_zero(lex02, lex2);
                Hx = p0 * ((pre1 * m1x + pre2 * m2x) - (pre1 * m0x + pre2 * m0x));
                Hy = p0 * ((pre1 * m1y + pre2 * m2y) - (pre1 * m0y + pre2 * m0y));
                Hz = p0 * ((pre1 * m1z + pre2 * m2z) - (pre1 * m0z + pre2 * m0z));
                /* Z (axis 2) */
                idx = k - 1; idx = (idx < 0 && w2) ? N2 + idx : idx; if (idx < 0) idx = 0;
                a = i * N12 + (long)j * N2 + idx;
                m1x = mx[a]; m1y = my[a]; m1z = mz[a];
                lex = mask1(lexMsk, a); lex2 = lex * lex; pre1 = avg_geom_zero(lex02, lex2);
                idx = k + 1; idx = (idx == N2 && w2) ? idx - N2 : idx; if (idx > N2 - 1) idx = N2 - 1;
                a = i * N12 + (long)j * N2 + idx;
                m2x = mx[a]; m2y = my[a]; m2z = mz[a];
                lex = mask1(lexMsk, a); lex2 = lex * lex; pre2 = avg_geom_zero(lex02, lex2);
                Hx += p2 * ((pre1 * m1x + pre2 * m2x) - (pre1 * m0x + pre2 * m0x));
                Hy += p2 * ((pre1 * m1y + pre2 * m2y) - (pre1 * m0y + pre2 * m0y));
                Hz += p2 * ((pre1 * m1z + pre2 * m2z) - (pre1 * m0z + pre2 * m0z));
                /* Y (axis 1) */
                idx = j - 1; idx = (idx < 0 && w1) ? N1 + idx : idx; if (idx < 0) idx = 0;
                a = i * N12 + (long)idx * N2 + k;
                m1x = mx[a]; m1y = my[a]; m1z = mz[a];
                lex = mask1(lexMsk, a); lex2 = lex * lex; pre1 = avg_geom_zero(lex02, lex2);
                idx = j + 1; idx = (idx == N1 && w1) ? idx - N1 : idx; if (idx > N1 - 1) idx = N1 - 1;
                a = i * N12 + (long)idx * N2 + k;
                m2x = mx[a]; m2y = my[a]; m2z = mz[a];
                lex = mask1(lexMsk, a); lex2 = lex * lex; pre2 = avg_geom_zero(lex02, lex2);
                Hx += p1 * ((pre1 * m1x + pre2 * m2x) - (pre1 * m0x + pre2 * m0x));
                Hy += p1 * ((pre1 * m1y + pre2 * m2y) - (pre1 * m0y + pre2 * m0y));
                Hz += p1 * ((pre1 * m1z + pre2 * m2z) - (pre1 * m0z + pre2 * m0z));
                hx[I] = msat0T0Mul * Hx;
                hy[I] = msat0T0Mul * Hy;
                hz[I] = msat0T0Mul * Hz;
            }
}

/* ------------------------------------------------------------------ LLB terms */
/* llbarTorqueKern, src/libhotspin/llbar-torque.cu:11-45: t = cross(H, M) with
 * common_func.h:42-45 cross(a,b) = b x a, i.e. t = M x H (internal frame). */
void or_llbar_torque(double* tx, double* ty, double* tz,
                     const double* mx, const double* my, const double* mz,
                     const double* hx, const double* hy, const double* hz,
                     const double* msk, int N) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
        double ms = msk ? msk[i] : 1.0;
        if (ms == 0.0) { tx[i] = ty[i] = tz[i] = 0.0; continue; }
        double Hx = hx[i], Hy = hy[i], Hz = hz[i];
        double Mx = mx[i], My = my[i], Mz = mz[i];
        tx[i] = -Hy * Mz + Hz * My;
        ty[i] = -Hz * Mx + Hx * Mz;
        tz[i] = -Hx * My + Hy * Mx;
    }
}

/* llbarLocal00ncKern, src/libhotspin/llbar-local00nc.cu:11-62: t = diag(lambda) H */
void or_llbar_local00nc(double* tx, double* ty, double* tz,
                        const double* hx, const double* hy, const double* hz,
                        const double* msk,
                        const double* lxx, const double* lyy, const double* lzz,
                        double lmx, double lmy, double lmz, int N) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
        double ms = msk ? msk[i] : 1.0;
        if (ms == 0.0) { tx[i] = ty[i] = tz[i] = 0.0; continue; }
        double l_xx = lxx ? lxx[i] * lmx : lmx;
        double l_yy = lyy ? lyy[i] * lmy : lmy;
        double l_zz = lzz ? lzz[i] * lmz : lmz;
        tx[i] = l_xx * hx[i];
        ty[i] = l_yy * hy[i];
        tz[i] = l_zz * hz[i];
    }
}

/* llbarLocal02ncKern, src/libhotspin/llbar-local02nc.cu:11-73: t = diag(mu) (m.H) m */
void or_llbar_local02nc(double* tx, double* ty, double* tz,
                        const double* mx, const double* my, const double* mz,
                        const double* hx, const double* hy, const double* hz,
                        const double* msk,
                        const double* uxx, const double* uyy, const double* uzz,
                        double umx, double umy, double umz, int N) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
        double ms = msk ? msk[i] : 1.0;
        if (ms == 0.0) { tx[i] = ty[i] = tz[i] = 0.0; continue; }
        double m0 = mx[i], m1 = my[i], m2 = mz[i];
        double mH = m0 * hx[i] + m1 * hy[i] + m2 * hz[i];
        double a = mH * m0, b = mH * m1, c = mH * m2;
        double u0 = uxx ? uxx[i] * umx : umx;
        double u1 = uyy ? uyy[i] * umy : umy;
        double u2 = uzz ? uzz[i] * umz : umz;
        tx[i] = u0 * a; ty[i] = u1 * b; tz[i] = u2 * c;
    }
}

/* llbarLocal02cKern, src/libhotspin/llbar-local02c.cu:11-69: t = cross(mu.cross(m,H), m) */
void or_llbar_local02c(double* tx, double* ty, double* tz,
                       const double* mx, const double* my, const double* mz,
                       const double* hx, const double* hy, const double* hz,
                       const double* msk,
                       const double* uxx, const double* uyy, const double* uzz,
                       double umx, double umy, double umz, int N) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
        double ms = msk ? msk[i] : 1.0;
        if (ms == 0.0) { tx[i] = ty[i] = tz[i] = 0.0; continue; }
        double ax = mx[i], ay = my[i], az = mz[i];
        double bx = hx[i], by = hy[i], bz = hz[i];
        /* cross(m, H) = H x m */
        double cx = -ay * bz + az * by;
        double cy = -az * bx + ax * bz;
        double cz = -ax * by + ay * bx;
        double u0 = uxx ? uxx[i] * umx : umx;
        double u1 = uyy ? uyy[i] * umy : umy;
        double u2 = uzz ? uzz[i] * umz : umz;
        double px = u0 * cx, py = u1 * cy, pz = u2 * cz;
        /* cross(p, m) = m x p */
        tx[i] = -py * az + pz * ay;
        ty[i] = -pz * ax + px * az;
        tz[i] = -px * ay + py * ax;
    }
}

/* long_field_Kern, src/libhotspin/long_field.cu:13-72 */
void or_long_field(double* hx, double* hy, double* hz,
                   const double* mx, const double* my, const double* mz,
                   const double* msat0T0Msk, const double* SMsk, const double* nMsk,
                   const double* TcMsk, const double* TsMsk,
                   double msat0T0Mul, double SMul, double nMul, double TcMul, double TsMul,
                   int N) {
#pragma omp parallel for schedule(static)
    for (int I = 0; I < N; I++) {
        double Ms0T0 = msat0T0Mul * mask1(msat0T0Msk, I);
        double S = SMul * mask1(SMsk, I);
        double n = nMul * mask1(nMsk, I);
        double Tc = TcMul * mask1(TcMsk, I);
        double Ts = TsMul * mask1(TsMsk, I);
        double m0 = mx[I], m1 = my[I], m2 = mz[I];
        double abs_mf = sqrt(m0 * m0 + m1 * m1 + m2 * m2);
        if (Ms0T0 == 0.0 || n == 0.0 || abs_mf <= OR_ZERO) {
            hx[I] = hy[I] = hz[I] = 0.0;
            continue;
        }
        /* normalize(), common_func.h:53-57 */
        double inv = (abs_mf != 0.0) ? 1.0 / abs_mf : 0.0;
        double s0 = m0 * inv, s1 = m1 * inv, s2 = m2 * inv;
        double preS = (S > OR_INFSPIN) ? 1.0 : S / (S + 1.0);
        double b = 3.0 * preS * Tc / Ts;
        double meb = abs_mf * b;
        double M = abs_mf;
        double W = (S > OR_INFSPIN) ? Lf(meb) : Bj(S, meb);
        double dWdx = (S > OR_INFSPIN) ? dLdx(meb) : dBjdx(S, meb);
        double mult = n * OR_KB * Ts * (W - M) / (OR_MU0 * Ms0T0 * dWdx);
        hx[I] = mult * s0;
        hy[I] = mult * s1;
        hz[I] = mult * s2;
    }
}

/* temperature_scaleAnisKern, src/libhotspin/temperature.cu:14-71 */
void or_temperature_scale_noise(double* hx, double* hy, double* hz,
                                const double* mu_xx, const double* mu_yy, const double* mu_zz,
                                const double* T, const double* msat0T0,
                                double mxx, double myy, double mzz, double K, int N) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
        double ms = mask1(msat0T0, i);
        if (ms == 0.0) { hx[i] = hy[i] = hz[i] = 0.0; continue; }
        double a = sqrt(mxx * mask1(mu_xx, i));
        double b = sqrt(myy * mask1(mu_yy, i));
        double c = sqrt(mzz * mask1(mu_zz, i));
        double ux = a * hx[i], uy = b * hy[i], uz = c * hz[i];
        double Ti = mask1(T, i);
        double pre = sqrt((Ti * K) / ms);
        hx[i] = pre * ux;
        hy[i] = pre * uy;
        hz[i] = pre * uz;
    }
}

/* Model + findroot_Ridders + brillouinKern, src/libhotspin/brillouin.cu:13-138 */
static double brillouin_model(double n, double J, double pre) {
    double x = pre * n;
    double val = (J > OR_INFSPIN) ? Lf(x) : Bj(J, x);
    return val - n;
}
static double ridders(double J, double mult, double xa, double xb) {
    double ya = brillouin_model(xa, J, mult);
    if (fabs(ya) < OR_ZERO) return xa;
    double yb = brillouin_model(xb, J, mult);
    if (fabs(yb) < OR_ZERO) return xb;
    double y1 = ya, x1 = xa, y2 = yb, x2 = xb;
    double x = 1.0e10, y, tx = x, teps = x;
    double x3, y3, dx, dy;
    int iter = 0;
    while (teps > OR_EPS && iter < 1000) {
        x3 = 0.5 * (x2 + x1);
        y3 = brillouin_model(x3, J, mult);
        dy = (y3 * y3 - y1 * y2);
        if (dy == 0.0) { x = x3; break; }
        dx = (x3 - x1) * sgn(y1 - y2) * y3 / (sqrt(dy));
        x = x3 + dx;
        y = brillouin_model(x, J, mult);
        int sy = signbit(y) != 0, sy3 = signbit(y3) != 0, sy1 = signbit(y1) != 0;
        double ny2 = (sy == sy3) ? y2 : y3;
        double nx2 = (sy == sy3) ? x2 : x3;
        y2 = ny2; x2 = nx2;
        ny2 = (sy == sy1 || x2 == x3) ? y2 : y1;
        nx2 = (sy == sy1 || x2 == x3) ? x2 : x1;
        y2 = ny2; x2 = nx2;
        y1 = y; x1 = x;
        teps = fabs((x - tx) / (tx + x));
        tx = x;
        iter++;
    }
    return x;
}
void or_brillouin(double* msat0, const double* msat0T0Msk, const double* T,
                  const double* TcMsk, const double* SMsk,
                  double msat0Mul, double msat0T0Mul, double TcMul, double SMul, int N) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < N; i++) {
        double Temp = T[i];
        double ms0T0 = msat0T0Mul * mask1(msat0T0Msk, i);
        if (ms0T0 == 0.0) { msat0[i] = 0.0; continue; }
        if (Temp == 0.0) { msat0[i] = ms0T0 / msat0Mul; continue; }
        double Tc = TcMul * mask1(TcMsk, i);
        if (Temp > Tc) { msat0[i] = 0.0; continue; }
        double S = SMul * mask1(SMsk, i);
        double preS = (S > OR_INFSPIN) ? 1.0 : S / (S + 1.0);
        double pre = 3.0 * preS * (Tc / Temp);
        double dT = (Tc - Temp) / Tc;
        double lo = (dT < 0.0004) ? -0.1 : 0.01;
        double hi = (dT < 0.0004) ? 0.5 : 1.1;
        double r = ridders(S, pre, lo, hi);
        msat0[i] = ms0T0 * fabs(r) / msat0Mul;
    }
}

/* normalizeKern, src/libhotspin/normalize.cu:13-34 */
void or_normalize(double* mx, double* my, double* mz, int N) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
        double Mx = mx[i], My = my[i], Mz = mz[i];
        double n = sqrt(Mx * Mx + My * My + Mz * Mz);
        n = (n == 0.0) ? 0.0 : 1.0 / n;
        mx[i] = Mx * n; my[i] = My * n; mz[i] = Mz * n;
    }
}

/* QspatKern, src/libhotspin/Qspat.cu:13-103 (neighbour order X, Z, Y) */
void or_qspat(double* Q, const double* T, const double* kMsk, double kMul,
              int sx, int sy, int sz, double csx, double csy, double csz,
              int px, int py, int pz) {
    double c2x = 1.0 / (csx * csx), c2y = 1.0 / (csy * csy), c2z = 1.0 / (csz * csz);
    long syz = (long)sy * sz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < sx; i++)
        for (int j = 0; j < sy; j++)
            for (int k = 0; k < sz; k++) {
                long I = i * syz + (long)j * sz + k, a;
                double T0 = T[I], T1, T2, k0 = mask1(kMsk, I), k1, k2, pre1, pre2, ddT;
                int idx;
                idx = i - 1; idx = (idx < 0 && px) ? sx + idx : idx; if (idx < 0) idx = 0;
                a = idx * syz + (long)j * sz + k; k1 = mask1(kMsk, a); T1 = T[a]; pre1 = avg_geom_zero(k0, k1);
                idx = i + 1; idx = (idx == sx && px) ? idx - sx : idx; if (idx > sx - 1) idx = sx - 1;
                a = idx * syz + (long)j * sz + k; k2 = mask1(kMsk, a); T2 = T[a]; pre2 = avg_geom_zero(k0, k2);
                ddT = c2x * ((pre1 * T1 + pre2 * T2) - (pre1 * T0 + pre2 * T0));
                idx = k - 1; idx = (idx < 0 && pz) ? sz + idx : idx; if (idx < 0) idx = 0;
                a = i * syz + (long)j * sz + idx; k1 = mask1(kMsk, a); T1 = T[a]; pre1 = avg_geom_zero(k0, k1);
                idx = k + 1; idx = (idx == sz && pz) ? idx - sz : idx; if (idx > sz - 1) idx = sz - 1;
                a = i * syz + (long)j * sz + idx; k2 = mask1(kMsk, a); T2 = T[a]; pre2 = avg_geom_zero(k0, k2);
                ddT += c2z * ((pre1 * T1 + pre2 * T2) - (pre1 * T0 + pre2 * T0));
                idx = j - 1; idx = (idx < 0 && py) ? sy + idx : idx; if (idx < 0) idx = 0;
                a = i * syz + (long)idx * sz + k; k1 = mask1(kMsk, a); T1 = T[a]; pre1 = avg_geom_zero(k0, k1);
                idx = j + 1; idx = (idx == sy && py) ? idx - sy : idx; if (idx > sy - 1) idx = sy - 1;
                a = i * syz + (long)idx * sz + k; k2 = mask1(kMsk, a); T2 = T[a]; pre2 = avg_geom_zero(k0, k2);
                ddT += c2y * ((pre1 * T1 + pre2 * T2) - (pre1 * T0 + pre2 * T0));
                Q[I] = kMul * ddT;
            }
}

/* QinterKern, src/libhotspin/Qinter.cu:13-28 */
void or_qinter(double* Qi, const double* Ti, const double* Tj, const double* G, double GMul, int N) {
    for (int i = 0; i < N; i++) {
        double g = G ? GMul * G[i] : GMul;
        Qi[i] = g * (Tj[i] - Ti[i]);
    }
}

/* ------------------------------------------------------------------ MFA helpers, anisotropy, nonlocal */
/* kappaKern, src/libhotspin/kappa.cu:15-61.  n is its mask alone (no multiplier). */
void or_kappa(double* kappa, const double* ms0Msk, const double* ms0T0Msk, const double* T,
              const double* TcMsk, const double* SMsk, const double* nMsk,
              double ms0Mul, double ms0T0Mul, double TcMul, double SMul, int N) {
    for (int i = 0; i < N; i++) {
        double u = mask1(ms0T0Msk, i);
        double ms0T0 = ms0T0Mul * u;
        double Temp = T[i];
        if (ms0T0 == 0.0 || Temp == 0.0) { kappa[i] = 0.0; continue; }
        double S = SMul * mask1(SMsk, i);
        double Tc = TcMul * mask1(TcMsk, i);
        double ms0 = ms0Mul * mask1(ms0Msk, i);
        double preS = (S > OR_INFSPIN) ? 1.0 : S / (S + 1.0);
        double J0 = 3.0 * Tc;
        double n = mask1(nMsk, i);
        double mul = u * u / (preS * J0 * n);
        double me = ms0 / ms0T0;
        double b = preS * J0 / Temp;
        double f = b * ((S > OR_INFSPIN) ? dLdx(me * b) : dBjdx(S, me * b));
        kappa[i] = mul * (f / (1.0 - f));
    }
}

/* ModelTs + findroot_Ridders_Ts + tsKern, src/libhotspin/Ts.cu:13-120: the spin
 * temperature T in [0, Tc] with B_S(pre/T) = m, pre = S^2 J0 m, J0 = 3Tc/(S(S+1)). */
static double ts_model(double T, double J, double pre, double m) {
    double x = (T == 0.0) ? 1.0e38 : pre / T;
    return Bj(J, x) - m;
}
static double ts_ridders(double J, double pre, double m, double xa, double xb) {
    double ya = ts_model(xa, J, pre, m);
    if (fabs(ya) < OR_ZERO) return xa;
    double yb = ts_model(xb, J, pre, m);
    if (fabs(yb) < OR_ZERO) return xb;
    double y1 = ya, x1 = xa, y2 = yb, x2 = xb;
    double x = 1.0e10, y, tx = x, teps = x;
    double x3, y3, dx, dy;
    int iter = 0;
    while (teps > OR_EPS && iter < 1000) {
        x3 = 0.5 * (x2 + x1);
        y3 = ts_model(x3, J, pre, m);
        dy = (y3 * y3 - y1 * y2);
        if (dy == 0.0) { x = x3; break; }
        dx = (x3 - x1) * sgn(y1 - y2) * y3 / (sqrt(dy));
        x = x3 + dx;
        y = ts_model(x, J, pre, m);
        int sy = signbit(y) != 0, sy3 = signbit(y3) != 0, sy1 = signbit(y1) != 0;
        double ny2 = (sy == sy3) ? y2 : y3;
        double nx2 = (sy == sy3) ? x2 : x3;
        y2 = ny2; x2 = nx2;
        ny2 = (sy == sy1 || x2 == x3) ? y2 : y1;
        nx2 = (sy == sy1 || x2 == x3) ? x2 : x1;
        y2 = ny2; x2 = nx2;
        y1 = y; x1 = x;
        teps = fabs((x - tx) / (tx + x));
        tx = x;
        iter++;
    }
    return x;
}
void or_ts(double* Ts, const double* msMsk, const double* ms0T0Msk, const double* TcMsk, const double* SMsk,
           double msMul, double ms0T0Mul, double TcMul, double SMul, int N) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < N; i++) {
        double ms0T0 = ms0T0Mul * mask1(ms0T0Msk, i);
        if (ms0T0 == 0.0) { Ts[i] = 0.0; continue; }
        double ms = msMul * mask1(msMsk, i);
        if (ms == ms0T0) { Ts[i] = 0.0; continue; }
        double Tc = TcMul * mask1(TcMsk, i);
        if (ms == 0.0) { Ts[i] = Tc; continue; }
        double S = SMul * mask1(SMsk, i);
        double J0 = 3.0 * Tc / (S * (S + 1.0));
        double m = ms / ms0T0;
        double pre = S * S * J0 * m;
        Ts[i] = ts_ridders(S, pre, m, 0.0, Tc);
    }
}

/* Debye + cpKern, src/libhotspin/common_func.h:168-173, Cp.cu:14-83:
 * n (T/Td)^3 int_0^{Td/T} x^4 e^x/(e^x-1)^2 dx by the 61-interval trapezoid rule
 * (INTMAXSTEPS, common_func.h:27); below T/Td = 0.01 the T^3 limit 4 pi^4/15. */
static double debye(double x) {
    double e = exp(x);
    double em1 = e - 1.0;
    return (x <= OR_ZERO) ? 0.0 : (x * x * x * x * e) / (em1 * em1);
}
void or_cp(double* Cp, const double* T, const double* TdMsk, const double* nMsk, double TdMul, int N) {
    const double pi = 3.14159265358979323846;
    for (int i = 0; i < N; i++) {
        double n = mask1(nMsk, i);
        double Td = TdMul * mask1(TdMsk, i);
        double Temp = T[i];
        if (Td == 0.0) { Cp[i] = 0.0; continue; }
        double xx = Td / Temp, ixx = Temp / Td;
        if (ixx < 0.01) {
            Cp[i] = n * (12.0 / 5.0) * (1.0 / 9.0) * (pi * pi * pi * pi) * (ixx * ixx * ixx);
            continue;
        }
        double h = xx / 61.0, inner = 0.0;
        for (int j = 1; j < 61; j++) inner += debye(j * h);
        double v = (debye(0.0) + debye(xx) + 2.0 * inner) * (0.5 * h);
        Cp[i] = n * v / (xx * xx * xx);
    }
}

/* energyFlowKern, src/libhotspin/energy_flow.cu:15-48.  Tc and n are masks alone. */
void or_energy_flow(double* w, const double* mx, const double* my, const double* mz,
                    const double* Rx, const double* Ry, const double* Rz,
                    const double* TcMsk, const double* SMsk, const double* nMsk, double SMul, int N) {
    for (int i = 0; i < N; i++) {
        double S = SMul * mask1(SMsk, i), Tc = mask1(TcMsk, i), n = mask1(nMsk, i);
        if (n == 0.0 || S == 0.0 || Tc == 0.0) { w[i] = 0.0; continue; }
        double rW = mx[i] * Rx[i] + my[i] * Ry[i] + mz[i] * Rz[i];
        w[i] = (Tc * S * S * n / (S * (S + 1.0))) * rW;
    }
}

/* uniaxialAnisotropyKern, src/libhotspin/uniaxialanisotropy.cu:12-52: m.u with the
 * normalized axis, the field along the axis as given. */
void or_uniaxial_anisotropy(double* hx, double* hy, double* hz,
                            const double* mx, const double* my, const double* mz,
                            const double* KuMsk, const double* msMsk, double K2Mul,
                            const double* uxMsk, double uxMul, const double* uyMsk, double uyMul,
                            const double* uzMsk, double uzMul, int N) {
    for (int i = 0; i < N; i++) {
        double ms = mask1(msMsk, i);
        if (ms == 0.0) { hx[i] = 0.0; hy[i] = 0.0; hz[i] = 0.0; continue; }
        double K = (K2Mul / ms) * mask1(KuMsk, i);
        double ux = uxMul * mask1(uxMsk, i), uy = uyMul * mask1(uyMsk, i), uz = uzMul * mask1(uzMsk, i);
        double len = sqrt(ux * ux + uy * uy + uz * uz);
        double inv = (len != 0.0) ? 1.0 / len : 0.0;
        double mu = mx[i] * (ux * inv) + my[i] * (uy * inv) + mz[i] * (uz * inv);
        hx[i] = K * mu * ux; hy[i] = K * mu * uy; hz[i] = K * mu * uz;
    }
}

/* llbarNonlocal00ncKern, src/libhotspin/llbar-nonlocal00nc.cu:11-208:
 * t_c = -(1/mu0) sum_d (lmul_c/cs_d^2) sum_+- avgGeom(l0_c, l+-_c)(mu0 H+-_c - mu0 H0_c),
 * neighbour order X(i), Z(k), Y(j).  The +1 neighbour along j is addressed
 * i*sy*sz + idx*sz + k (the reference's line 183 has i*sy*sy + idx*sy + k). */
void or_llbar_nonlocal00nc(double* t[3], double* const H[3], const double* msk, double* const lam[3],
                           const double lmul[3], int sx, int sy, int sz,
                           double csx, double csy, double csz, int px, int py, int pz) {
    double c2[3] = {1.0 / (csx * csx), 1.0 / (csy * csy), 1.0 / (csz * csz)};
    int n[3] = {sx, sy, sz}, w[3] = {px, py, pz};
    long str[3] = {(long)sy * sz, (long)sz, 1L};
    const int order[3] = {0, 2, 1};
    for (int i = 0; i < sx; i++)
        for (int j = 0; j < sy; j++)
            for (int k = 0; k < sz; k++) {
                long I = i * str[0] + j * str[1] + k;
                if (mask1(msk, I) == 0.0) { t[0][I] = 0.0; t[1][I] = 0.0; t[2][I] = 0.0; continue; }
                int pos[3] = {i, j, k};
                double R[3] = {0.0, 0.0, 0.0};
                for (int q = 0; q < 3; q++) {
                    int ax = order[q], c = pos[ax];
                    int lo = c - 1; lo = (lo < 0 && w[ax]) ? n[ax] + lo : lo; if (lo < 0) lo = 0;
                    int hi = c + 1; hi = (hi == n[ax] && w[ax]) ? hi - n[ax] : hi; if (hi > n[ax] - 1) hi = n[ax] - 1;
                    long a1 = I + (long)(lo - c) * str[ax], a2 = I + (long)(hi - c) * str[ax];
                    for (int cc = 0; cc < 3; cc++) {
                        double l0 = mask1(lam[cc], I);
                        double p1 = avg_geom_zero(l0, mask1(lam[cc], a1));
                        double p2 = avg_geom_zero(l0, mask1(lam[cc], a2));
                        double H0 = OR_MU0 * H[cc][I], H1 = OR_MU0 * H[cc][a1], H2 = OR_MU0 * H[cc][a2];
                        R[cc] += (lmul[cc] * c2[ax]) * ((p1 * H1 + p2 * H2) - H0 * (p1 + p2));
                    }
                }
                double pre = 1.0 / OR_MU0;
                for (int cc = 0; cc < 3; cc++) t[cc][I] = -pre * R[cc];
            }
}

/* divMulPowKern, src/libhotspin/divMulPow.cu:13-30: the power in single precision
 * (powf on the values rounded to float, :27). */
void or_div_mul_pow(double* dst, const double* a, const double* b, const double* c, double p, int N) {
    for (int i = 0; i < N; i++) {
        double bb = b ? b[i] : 1.0;
        double val = (bb == 0.0) ? 0.0 : a[i] / bb;
        double cc = c ? c[i] : 1.0;
        cc = (cc == 0.0) ? 0.0 : (double)powf((float)cc, (float)p);
        dst[i] = val * cc;
    }
}

/* wavgKern + weightedAvgZero, src/libhotspin/wavg.cu:13-30, common_func.h:162-166 */
void or_wavg(double* dst, const double* a, const double* b, const double* w0Msk, const double* w1Msk,
             const double* RMsk, double w0Mul, double w1Mul, double RMul, int N) {
    for (int i = 0; i < N; i++) {
        double w0 = w0Mul * mask1(w0Msk, i), w1 = w1Mul * mask1(w1Msk, i), R = RMul * mask1(RMsk, i);
        double x0 = a[i], x1 = b[i];
        double den = w0 + w1 + 2.0 * R * sqrt(w0 * w1);
        dst[i] = (den == 0.0) ? 0.0 : (w0 * x0 + w1 * x1 + 2.0 * R * sqrt(w0 * x0 * w1 * x1)) / den;
    }
}

/* ------------------------------------------------------------------ linear algebra, add.cu */
void or_madd(double* dst, const double* a, const double* b, double mulB, int N) {
    for (int i = 0; i < N; i++) dst[i] = b ? a[i] + mulB * b[i] : a[i] + mulB; /* add.cu:41-56,113-126 */
}
void or_madd2(double* a, const double* b, double mulB, const double* c, double mulC, int N) {
    for (int i = 0; i < N; i++) { /* add.cu:161-174 */
        double B = b ? b[i] : 1.0, C = c ? c[i] : 1.0;
        a[i] += mulB * B + mulC * C;
    }
}
void or_vec_madd(double* dst[3], double* const a[3], double* const b[3], const double mulB[3], int N) {
    for (int c = 0; c < 3; c++) /* add.cu:58-90,128-146 */
        for (int i = 0; i < N; i++) dst[c][i] = b[c] ? a[c][i] + mulB[c] * b[c][i] : a[c][i] + mulB[c];
}
void or_lincomb2(double* dst, const double* a, double mulA, const double* b, double mulB, int N) {
    for (int i = 0; i < N; i++) dst[i] = mulA * a[i] + mulB * b[i]; /* add.cu:205-219 */
}
void or_add_madd(double* dst, const double* a, const double* b, const double* c, double mul, int N) {
    for (int i = 0; i < N; i++) dst[i] = a[i] + mul * (b[i] + c[i]); /* add.cu:31-39 */
}
void or_mul(double* dst, const double* a, const double* b, int N) {
    for (int i = 0; i < N; i++) dst[i] = a[i] * b[i]; /* mul.cu:13-20 */
}
void or_div(double* dst, const double* a, const double* b, int N) {
    for (int i = 0; i < N; i++) { /* div.cu:13-22 */
        double bb = b ? b[i] : 1.0;
        dst[i] = (bb == 0.0) ? 0.0 : a[i] / bb;
    }
}
void or_dot_mask(double* dst, const double* ax, const double* ay, const double* az,
                 const double* bx, const double* by, const double* bz,
                 double axm, double aym, double azm, double bxm, double bym, double bzm, int N) {
    for (int i = 0; i < N; i++) { /* dot.cu:10-24 */
        double a0 = axm * mask1(ax, i), a1 = aym * mask1(ay, i), a2 = azm * mask1(az, i);
        double b0 = bxm * mask1(bx, i), b1 = bym * mask1(by, i), b2 = bzm * mask1(bz, i);
        dst[i] = a0 * b0 + a1 * b1 + a2 * b2;
    }
}
/* copyPad3DKern / copyUnPad3DKern, src/libhotspin/copypad.cu:34-97 (one component) */
void or_copy_pad3d(double* dst, int D0, int D1, int D2, const double* src, int S0, int S1, int S2) {
    (void)D0;
    for (int i = 0; i < S0; i++)
        for (int j = 0; j < S1; j++)
            for (int k = 0; k < S2; k++)
                dst[((long)i * D1 + j) * D2 + k] = src[((long)i * S1 + j) * S2 + k];
}
void or_copy_unpad3d(double* dst, int D0, int D1, int D2, const double* src, int S0, int S1, int S2) {
    (void)S0;
    for (int i = 0; i < D0; i++)
        for (int j = 0; j < D1; j++)
            for (int k = 0; k < D2; k++)
                dst[((long)i * D1 + j) * D2 + k] = src[((long)i * S1 + j) * S2 + k];
}

/* tensSYMMVecMulKern, src/libhotspin/mul.cu:31-90.  Interleaved complex over (Nx,Ny,Nz)
 * with the real kernel octant (Nx/2+1, Ny/2+1, Nz); mirror sign flips of XY/XZ/YZ. */
void or_tens_symm_vec_mul(double* dstX, double* dstY, double* dstZ,
                          const double* srcX, const double* srcY, const double* srcZ,
                          const double* kXX, const double* kYY, const double* kZZ,
                          const double* kYZ, const double* kXZ, const double* kXY,
                          double srcMul, int Nx, int Ny, int Nz) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < Nx; i++)
        for (int j = 0; j < Ny; j++)
            for (int k = 0; k < Nz; k++) {
                long I = ((long)i * Ny + j) * Nz + k;
                long e = 2 * I;
                double RMX = srcMul * srcX[e], IMX = srcMul * srcX[e + 1];
                double RMY = srcMul * srcY[e], IMY = srcMul * srcY[e + 1];
                double RMZ = srcMul * srcZ[e], IMZ = srcMul * srcZ[e + 1];
                double sXY = 1.0, sXZ = 1.0, sYZ = 1.0;
                sXY = (j > Ny / 2) ? -sXY : sXY;
                sYZ = (j > Ny / 2) ? -sYZ : sYZ;
                int jj = (j > Ny / 2) ? Ny - j : j;
                sXY = (i > Nx / 2) ? -sXY : sXY;
                sXZ = (i > Nx / 2) ? -sXZ : sXZ;
                int ii = (i > Nx / 2) ? Nx - i : i;
                long K = ((long)ii * (Ny / 2 + 1) + jj) * Nz + k;
                double KXX = mask0(kXX, K);
                double KXZ = sXZ * mask0(kXZ, K);
                double KXY = sXY * mask0(kXY, K);
                dstX[e] = KXX * RMX + KXY * RMY + KXZ * RMZ;
                dstX[e + 1] = KXX * IMX + KXY * IMY + KXZ * IMZ;
                double KYZ = sYZ * mask0(kYZ, K);
                double KYY = mask0(kYY, K);
                dstY[e] = KXY * RMX + KYY * RMY + KYZ * RMZ;
                dstY[e + 1] = KXY * IMX + KYY * IMY + KYZ * IMZ;
                double KZZ = mask0(kZZ, K);
                dstZ[e] = KXZ * RMX + KYZ * RMY + KZZ * RMZ;
                dstZ[e + 1] = KXZ * IMX + KYZ * IMY + KZZ * IMZ;
            }
}

/* ------------------------------------------------------------------ reductions */
/* Intended results of reduce.cu's partial kernels + gpu/reductor.go:69-202 host
 * finish.  The reference's 1024-thread max/min fold defect (reduce.cu:336 etc.)
 * is not reproduced: these are the true maxima. */
double or_sum(const double* a, long N) { double s = 0; for (long i = 0; i < N; i++) s += a[i]; return s; }
double or_sdot(const double* a, const double* b, long N) { double s = 0; for (long i = 0; i < N; i++) s += a[i] * b[i]; return s; }
double or_max(const double* a, long N) { double m = a[0]; for (long i = 1; i < N; i++) if (a[i] > m) m = a[i]; return m; }
double or_min(const double* a, long N) { double m = a[0]; for (long i = 1; i < N; i++) if (a[i] < m) m = a[i]; return m; }
double or_maxabs(const double* a, long N) { double m = 0; for (long i = 0; i < N; i++) if (fabs(a[i]) > m) m = fabs(a[i]); return m; }
double or_maxdiff(const double* a, const double* b, long N) {
    double m = 0; for (long i = 0; i < N; i++) { double d = fabs(a[i] - b[i]); if (d > m) m = d; } return m;
}
double or_maxsum(const double* a, const double* b, long N) {
    double m = 0; for (long i = 0; i < N; i++) { double d = fabs(a[i] + b[i]); if (d > m) m = d; } return m;
}
double or_maxnorm3sq(const double* x, const double* y, const double* z, long N) {
    double m = 0; for (long i = 0; i < N; i++) { double d = x[i] * x[i] + y[i] * y[i] + z[i] * z[i]; if (d > m) m = d; } return m;
}
double or_maxnorm3sqdiff(const double* x1, const double* y1, const double* z1,
                         const double* x2, const double* y2, const double* z2, long N) {
    double m = 0;
    for (long i = 0; i < N; i++) {
        double a = x1[i] - x2[i], b = y1[i] - y2[i], c = z1[i] - z2[i];
        double d = a * a + b * b + c * c; if (d > m) m = d;
    }
    return m;
}

/* ------------------------------------------------------------------ FFT (own, FP64) */
/* Stockham autosort, mixed radix (4,2,3,5,7, generic odd prime), unnormalised.
 * Stage: for j < n/R: v[r] = x[j + r n/R] * w^(r (j mod L)), L = product of
 * earlier radices; DFT_R(v); y[(j/L) L R + (j mod L) + r L] = v[r].
 * Roots come from one table w_n^t = exp(-2 pi i t/n), t < n. */
typedef struct {
    int n, nf, f[64];
    double* w;
} fftplan;
static void plan_init(fftplan* p, int n) {
    p->n = n;
    p->nf = 0;
    int m = n;
    while (m % 4 == 0) { p->f[p->nf++] = 4; m /= 4; }
    while (m % 2 == 0) { p->f[p->nf++] = 2; m /= 2; }
    for (int q = 3; m > 1; q += 2) {
        while (m % q == 0) { p->f[p->nf++] = q; m /= q; }
        if ((long)q * q > m && m > 1) { p->f[p->nf++] = m; m = 1; }
    }
    p->w = (double*)malloc(sizeof(double) * 2 * (n > 0 ? n : 1));
    for (int t = 0; t < n; t++) {
        double ang = -2.0 * OR_PI * (double)t / (double)n;
        p->w[2 * t] = cos(ang);
        p->w[2 * t + 1] = sin(ang);
    }
}
static void plan_free(fftplan* p) { free(p->w); }
/* x: 2n doubles in/out; work: 2n doubles scratch; sign -1 forward, +1 backward */
static void fft_exec(const fftplan* p, double* x, double* work, int sign) {
    int n = p->n;
    if (n <= 1) return;
    double* src = x;
    double* dst = work;
    int L = 1;
    double vr[128], vi[128];
    double s = (sign < 0) ? 1.0 : -1.0; /* conjugate table for backward */
    for (int st = 0; st < p->nf; st++) {
        int R = p->f[st];
        int m = n / R;
        int tw = n / (L * R), dr = n / R;
        for (int j = 0; j < m; j++) {
            int jl = j % L;
            for (int r = 0; r < R; r++) {
                double a = src[2 * (j + r * m)], b = src[2 * (j + r * m) + 1];
                long t = ((long)r * jl * tw) % n;
                double c = p->w[2 * t], sn = s * p->w[2 * t + 1];
                vr[r] = a * c - b * sn;
                vi[r] = a * sn + b * c;
            }
            int d = (j / L) * L * R + jl;
            for (int q = 0; q < R; q++) {
                double sr = 0, si = 0;
                for (int r = 0; r < R; r++) {
                    long t = ((long)((r * q) % R)) * dr;
                    double c = p->w[2 * t], sn = s * p->w[2 * t + 1];
                    sr += vr[r] * c - vi[r] * sn;
                    si += vr[r] * sn + vi[r] * c;
                }
                dst[2 * (d + q * L)] = sr;
                dst[2 * (d + q * L) + 1] = si;
            }
        }
        double* t = src; src = dst; dst = t;
        L *= R;
    }
    if (src != x) memcpy(x, src, sizeof(double) * 2 * n);
}
int or_fft1d(double* x, int n, int sign) {
    if (n <= 1) return 0;
    fftplan p;
    plan_init(&p, n);
    for (int i = 0; i < p.nf; i++) if (p.f[i] > 128) { plan_free(&p); return -1; }
    double* work = (double*)malloc(sizeof(double) * 2 * n);
    fft_exec(&p, x, work, sign);
    free(work);
    plan_free(&p);
    return 0;
}

/* 3-D real-to-complex forward transform of a real block of extent S (zero
 * outside) embedded in the logic size P; out: (P0,P1,H) complex, H=P2/2+1. */
static void r2c3d(const double* in, const int S[3], const int P[3], double* out) {
    int H = P[2] / 2 + 1;
    long rowc = (long)H * 2;
    memset(out, 0, sizeof(double) * (long)P[0] * P[1] * rowc);
    fftplan pl[3];
    for (int a = 0; a < 3; a++) plan_init(&pl[a], P[a]);
    int nmax = P[0] > P[1] ? P[0] : P[1];
    if (P[2] > nmax) nmax = P[2];
#pragma omp parallel
    {
        double* buf = (double*)malloc(sizeof(double) * 2 * nmax);
        double* work = (double*)malloc(sizeof(double) * 2 * nmax);
#pragma omp for collapse(2) schedule(static)
        for (int i = 0; i < S[0]; i++)
            for (int j = 0; j < S[1]; j++) {
                memset(buf, 0, sizeof(double) * 2 * P[2]);
                const double* r = in + ((long)i * S[1] + j) * S[2];
                for (int k = 0; k < S[2]; k++) buf[2 * k] = r[k];
                fft_exec(&pl[2], buf, work, -1);
                double* o = out + ((long)i * P[1] + j) * rowc;
                memcpy(o, buf, sizeof(double) * rowc);
            }
#pragma omp for collapse(2) schedule(static)
        for (int i = 0; i < S[0]; i++)
            for (int k = 0; k < H; k++) {
                for (int j = 0; j < P[1]; j++) {
                    const double* o = out + ((long)i * P[1] + j) * rowc + 2 * k;
                    buf[2 * j] = o[0]; buf[2 * j + 1] = o[1];
                }
                fft_exec(&pl[1], buf, work, -1);
                for (int j = 0; j < P[1]; j++) {
                    double* o = out + ((long)i * P[1] + j) * rowc + 2 * k;
                    o[0] = buf[2 * j]; o[1] = buf[2 * j + 1];
                }
            }
        if (P[0] > 1) {
#pragma omp for collapse(2) schedule(static)
            for (int j = 0; j < P[1]; j++)
                for (int k = 0; k < H; k++) {
                    for (int i = 0; i < P[0]; i++) {
                        const double* o = out + ((long)i * P[1] + j) * rowc + 2 * k;
                        buf[2 * i] = o[0]; buf[2 * i + 1] = o[1];
                    }
                    fft_exec(&pl[0], buf, work, -1);
                    for (int i = 0; i < P[0]; i++) {
                        double* o = out + ((long)i * P[1] + j) * rowc + 2 * k;
                        o[0] = buf[2 * i]; o[1] = buf[2 * i + 1];
                    }
                }
        }
        free(buf);
        free(work);
    }
    for (int a = 0; a < 3; a++) plan_free(&pl[a]);
}

/* 3-D complex-to-real backward transform (unnormalised), keeping only the
 * output block of extent S; 'spec' is overwritten.  The imaginary parts of
 * the k=0 and k=P2/2 bins do not contribute (C2R semantics). */
static void c2r3d(double* spec, const int P[3], const int S[3], double* out) {
    int H = P[2] / 2 + 1;
    long rowc = (long)H * 2;
    fftplan pl[3];
    for (int a = 0; a < 3; a++) plan_init(&pl[a], P[a]);
    int nmax = P[0] > P[1] ? P[0] : P[1];
    if (P[2] > nmax) nmax = P[2];
#pragma omp parallel
    {
        double* buf = (double*)malloc(sizeof(double) * 2 * nmax);
        double* work = (double*)malloc(sizeof(double) * 2 * nmax);
        if (P[0] > 1) {
#pragma omp for collapse(2) schedule(static)
            for (int j = 0; j < P[1]; j++)
                for (int k = 0; k < H; k++) {
                    for (int i = 0; i < P[0]; i++) {
                        const double* o = spec + ((long)i * P[1] + j) * rowc + 2 * k;
                        buf[2 * i] = o[0]; buf[2 * i + 1] = o[1];
                    }
                    fft_exec(&pl[0], buf, work, +1);
                    for (int i = 0; i < S[0]; i++) {
                        double* o = spec + ((long)i * P[1] + j) * rowc + 2 * k;
                        o[0] = buf[2 * i]; o[1] = buf[2 * i + 1];
                    }
                }
        }
#pragma omp for collapse(2) schedule(static)
        for (int i = 0; i < S[0]; i++)
            for (int k = 0; k < H; k++) {
                for (int j = 0; j < P[1]; j++) {
                    const double* o = spec + ((long)i * P[1] + j) * rowc + 2 * k;
                    buf[2 * j] = o[0]; buf[2 * j + 1] = o[1];
                }
                fft_exec(&pl[1], buf, work, +1);
                for (int j = 0; j < S[1]; j++) {
                    double* o = spec + ((long)i * P[1] + j) * rowc + 2 * k;
                    o[0] = buf[2 * j]; o[1] = buf[2 * j + 1];
                }
            }
#pragma omp for collapse(2) schedule(static)
        for (int i = 0; i < S[0]; i++)
            for (int j = 0; j < S[1]; j++) {
                const double* o = spec + ((long)i * P[1] + j) * rowc;
                for (int k = 0; k < H; k++) { buf[2 * k] = o[2 * k]; buf[2 * k + 1] = o[2 * k + 1]; }
                for (int k = H; k < P[2]; k++) {
                    int kk = P[2] - k;
                    buf[2 * k] = o[2 * kk];
                    buf[2 * k + 1] = -o[2 * kk + 1];
                }
                buf[1] = 0.0;
                if (P[2] % 2 == 0) buf[2 * (P[2] / 2) + 1] = 0.0;
                fft_exec(&pl[2], buf, work, +1);
                double* r = out + ((long)i * S[1] + j) * S[2];
                for (int k = 0; k < S[2]; k++) r[k] = buf[2 * k];
            }
        free(buf);
        free(work);
    }
    for (int a = 0; a < 3; a++) plan_free(&pl[a]);
}

/* ------------------------------------------------------------------ Kernel_Arne */
static inline int wrapi(int number, int max) {
    while (number < 0) number += max;
    while (number >= max) number -= max;
    return number;
}
static inline double delta_d(int d) {
    if (d < 0) d = -d;
    if (d > 0) d -= 1;
    return (double)d;
}
/* FullTensorIdx, src/hotspin-core/common/index.go:44-47 (XX=0,YY=1,ZZ=2,YZ=3,XZ=4,XY=5) */
static const int full_tensor_idx[3][3] = {{0, 5, 4}, {5, 1, 3}, {4, 3, 2}};

/* Kernel_Arne, src/hotspin-core/modules/kernel_magnetostatic.go:25-241, with
 * kernelRanges (:244-257), delta (:261-270) and wrap (:273-281).  The loop over
 * destination images keeps the reference's += order per destination cell. */
void or_kernel_arne(const int size[3], const double cs[3], const int pbc[3], int accuracy_,
                    double* out) {
    long P = (long)size[0] * size[1] * size[2];
    memset(out, 0, sizeof(double) * 6 * P);
    double accuracy = (double)accuracy_;
    int r1[3], r2[3];
    for (int c = 0; c < 3; c++) {
        if (pbc[c] == 0) { r1[c] = -(size[c] - 1) / 2; r2[c] = (size[c] - 1) / 2; }
        else { r1[c] = -(size[c] * pbc[c] - 1); r2[c] = (size[c] * pbc[c] - 1); }
    }
    if (size[2] == 1 && pbc[2] == 0) r2[2] = 0;
    double L = cs[0];
    if (cs[1] < L) L = cs[1];
    if (cs[2] < L) L = cs[2];
    int nzw = size[2] / 2 + 1;
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
    for (int s = 0; s < 3; s++)
        for (int zwant = 0; zwant < nzw; zwant++) {
            int u = s, v = (s + 1) % 3, w = (s + 2) % 3;
            double R[3], R2[3], pole[3] = {0, 0, 0};
            for (int z = r1[2]; z <= r2[2]; z++) {
                int zw = wrapi(z, size[2]);
                if (zw > size[2] / 2 || zw != zwant) continue;
                R[2] = (double)z * cs[2];
                for (int y = r1[1]; y <= r2[1]; y++) {
                    int yw = wrapi(y, size[1]);
                    if (yw > size[1] / 2) continue;
                    R[1] = (double)y * cs[1];
                    for (int x = r1[0]; x <= r2[0]; x++) {
                        int xw = wrapi(x, size[0]);
                        if (xw > size[0] / 2) continue;
                        R[0] = (double)x * cs[0];
                        double dx = delta_d(x) * cs[0], dy = delta_d(y) * cs[1], dz = delta_d(z) * cs[2];
                        double d = sqrt(dx * dx + dy * dy + dz * dz);
                        if (d == 0) d = L;
                        double maxSize = d / accuracy;
                        int nv = (int)(fmax(cs[v] / maxSize, 1) + 0.5);
                        int nw = (int)(fmax(cs[w] / maxSize, 1) + 0.5);
                        int nx = (int)(fmax(cs[0] / maxSize, 1) + 0.5);
                        int ny = (int)(fmax(cs[1] / maxSize, 1) + 0.5);
                        int nz = (int)(fmax(cs[2] / maxSize, 1) + 0.5);
                        nv *= 2;
                        nw *= 2;
                        double scale = 1 / (double)((long)nv * nw * nx * ny * nz);
                        double surface = cs[v] * cs[w];
                        double charge = surface * scale;
                        double pu1 = cs[u] / 2.;
                        double pu2 = -pu1;
                        double B[3] = {0, 0, 0};
                        for (int i = 0; i < nv; i++) {
                            double pv = -(cs[v] / 2.) + cs[v] / (double)(2 * nv) + (double)i * (cs[v] / (double)nv);
                            pole[v] = pv;
                            for (int j = 0; j < nw; j++) {
                                double pw = -(cs[w] / 2.) + cs[w] / (double)(2 * nw) + (double)j * (cs[w] / (double)nw);
                                pole[w] = pw;
                                for (int a = 0; a < nx; a++) {
                                    double rx = R[0] - cs[0] / 2 + cs[0] / (double)(2 * nx) + (cs[0] / (double)nx) * (double)a;
                                    for (int b = 0; b < ny; b++) {
                                        double ry = R[1] - cs[1] / 2 + cs[1] / (double)(2 * ny) + (cs[1] / (double)ny) * (double)b;
                                        for (int g = 0; g < nz; g++) {
                                            double rz = R[2] - cs[2] / 2 + cs[2] / (double)(2 * nz) + (cs[2] / (double)nz) * (double)g;
                                            pole[u] = pu1;
                                            R2[0] = rx - pole[0]; R2[1] = ry - pole[1]; R2[2] = rz - pole[2];
                                            double r = sqrt(R2[0] * R2[0] + R2[1] * R2[1] + R2[2] * R2[2]);
                                            double qr = charge / (4 * OR_PI * r * r * r);
                                            double bx = R2[0] * qr, by = R2[1] * qr, bz = R2[2] * qr;
                                            pole[u] = pu2;
                                            R2[0] = rx - pole[0]; R2[1] = ry - pole[1]; R2[2] = rz - pole[2];
                                            r = sqrt(R2[0] * R2[0] + R2[1] * R2[1] + R2[2] * R2[2]);
                                            qr = -charge / (4 * OR_PI * r * r * r);
                                            B[0] += (bx + R2[0] * qr);
                                            B[1] += (by + R2[1] * qr);
                                            B[2] += (bz + R2[2] * qr);
                                        }
                                    }
                                }
                            }
                        }
                        for (int dd = s; dd < 3; dd++) {
                            int I = full_tensor_idx[s][dd];
                            out[I * P + ((long)xw * size[1] + yw) * size[2] + zw] += B[dd];
                        }
                    }
                }
            }
        }
    /* symmetry reconstruction, kernel_magnetostatic.go:190-234 */
    double* XX = out + 0 * P; double* YY = out + 1 * P; double* ZZ = out + 2 * P;
    double* YZ = out + 3 * P; double* XZ = out + 4 * P; double* XY = out + 5 * P;
#define KI(x, y, z) (((long)(x) * size[1] + (y)) * size[2] + (z))
    for (int z = 0; z < size[2]; z++)
        for (int y = 0; y < size[1]; y++)
            for (int x = size[0] / 2 + 1; x < size[0]; x++) {
                int x2 = size[0] - x;
                XX[KI(x, y, z)] = +XX[KI(x2, y, z)];
                XY[KI(x, y, z)] = -XY[KI(x2, y, z)];
                XZ[KI(x, y, z)] = -XZ[KI(x2, y, z)];
                YY[KI(x, y, z)] = +YY[KI(x2, y, z)];
                YZ[KI(x, y, z)] = +YZ[KI(x2, y, z)];
                ZZ[KI(x, y, z)] = +ZZ[KI(x2, y, z)];
            }
    for (int z = 0; z < size[2]; z++)
        for (int y = size[1] / 2 + 1; y < size[1]; y++) {
            int y2 = size[1] - y;
            for (int x = 0; x < size[0]; x++) {
                XX[KI(x, y, z)] = +XX[KI(x, y2, z)];
                XY[KI(x, y, z)] = -XY[KI(x, y2, z)];
                XZ[KI(x, y, z)] = +XZ[KI(x, y2, z)];
                YY[KI(x, y, z)] = +YY[KI(x, y2, z)];
                YZ[KI(x, y, z)] = -YZ[KI(x, y2, z)];
                ZZ[KI(x, y, z)] = +ZZ[KI(x, y2, z)];
            }
        }
    for (int z = size[2] / 2 + 1; z < size[2]; z++) {
        int z2 = size[2] - z;
        for (int y = 0; y < size[1]; y++)
            for (int x = 0; x < size[0]; x++) {
                XX[KI(x, y, z)] = +XX[KI(x, y, z2)];
                XY[KI(x, y, z)] = +XY[KI(x, y, z2)];
                XZ[KI(x, y, z)] = -XZ[KI(x, y, z2)];
                YY[KI(x, y, z)] = +YY[KI(x, y, z2)];
                YZ[KI(x, y, z)] = -YZ[KI(x, y, z2)];
                ZZ[KI(x, y, z)] = +ZZ[KI(x, y, z2)];
            }
    }
#undef KI
}

/* Newell demag tensor (no counterpart in the reference, which integrates
 * Kernel_Arne): A. J. Newell, W. Williams, D. J. Dunlop, J. Geophys. Res. 98
 * (1993) 9551.  N_xx of two uniformly magnetised cuboids at displacement R is
 * the 27-point second difference of f over the cell sizes, N_xy that of g,
 * divided by 4 pi V; the other components permute the axes.  Beyond
 * R_far = 35 hmax^0.6 hmin^0.4 the differences lose digits in FP64 and the
 * cell-averaged dipole tensor is used instead:
 * N ~ -V/(4 pi) [T(R)/2 + sum_a (T(R + h_a e_a) + T(R - h_a e_a))/12],
 * T = (3 R R^T - R^2 I)/R^5, i.e. T averaged over both cells to O(h^4/R^4).
 * The kernel value is -N (the sign of Kernel_Arne's field kernel).  Image
 * ranges and the wrapped layout are those of or_kernel_arne. */
static double nw_f(double x, double y, double z) {
    x = fabs(x);
    y = fabs(y);
    z = fabs(z);
    const double x2 = x * x, y2 = y * y, z2 = z * z, R = sqrt(x2 + y2 + z2);
    double r = (2 * x2 - y2 - z2) * R / 6;
    if (y > 0 && x2 + z2 > 0) r += 0.5 * y * (z2 - x2) * asinh(y / sqrt(x2 + z2));
    if (z > 0 && x2 + y2 > 0) r += 0.5 * z * (y2 - x2) * asinh(z / sqrt(x2 + y2));
    if (x > 0 && y > 0 && z > 0) r -= x * y * z * atan(y * z / (x * R));
    return r;
}
static double nw_g(double x, double y, double z) {
    double sg = 1;
    if (x < 0) { sg = -sg; x = -x; }
    if (y < 0) { sg = -sg; y = -y; }
    z = fabs(z);
    const double x2 = x * x, y2 = y * y, z2 = z * z, R = sqrt(x2 + y2 + z2);
    double r = -x * y * R / 3;
    if (z > 0 && x2 + y2 > 0) r += x * y * z * asinh(z / sqrt(x2 + y2));
    if (x > 0 && y2 + z2 > 0) r += y / 6 * (3 * z2 - y2) * asinh(x / sqrt(y2 + z2));
    if (y > 0 && x2 + z2 > 0) r += x / 6 * (3 * z2 - x2) * asinh(y / sqrt(x2 + z2));
    if (z > 0) r -= z2 * z / 6 * atan(x * y / (z * R));
    if (y > 0) r -= z * y2 / 2 * atan(x * z / (y * R));
    if (x > 0) r -= z * x2 / 2 * atan(y * z / (x * R));
    return sg * r;
}
/* sum over i,j,k in {-1,0,1} of w_i w_j w_k F(X + i a, Y + j b, Z + k c), w_0 = 2, w_1 = -1 */
static double nw_stencil(double (*F)(double, double, double), double X, double Y, double Z, double a, double b,
                         double c) {
    double s = 0;
    for (int i = -1; i <= 1; i++)
        for (int j = -1; j <= 1; j++)
            for (int k = -1; k <= 1; k++) {
                const double w = (i ? -1.0 : 2.0) * (j ? -1.0 : 2.0) * (k ? -1.0 : 2.0);
                s += w * F(X + i * a, Y + j * b, Z + k * c);
            }
    return s;
}
static void nw_dipole(const double R[3], double t[6]) {
    const double r2 = R[0] * R[0] + R[1] * R[1] + R[2] * R[2], r5 = r2 * r2 * sqrt(r2);
    t[0] = (3 * R[0] * R[0] - r2) / r5;
    t[1] = (3 * R[1] * R[1] - r2) / r5;
    t[2] = (3 * R[2] * R[2] - r2) / r5;
    t[3] = 3 * R[1] * R[2] / r5;
    t[4] = 3 * R[0] * R[2] / r5;
    t[5] = 3 * R[0] * R[1] / r5;
}
/* N (XX,YY,ZZ,YZ,XZ,XY) of displacement R between cells of size h */
void or_newell_tensor(const double R[3], const double h[3], double N[6]) {
    const double V = h[0] * h[1] * h[2];
    double hmax = h[0], hmin = h[0];
    for (int a = 1; a < 3; a++) {
        if (h[a] > hmax) hmax = h[a];
        if (h[a] < hmin) hmin = h[a];
    }
    const double rfar = 35.0 * pow(hmax, 0.6) * pow(hmin, 0.4);
    const double r = sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
    if (r >= rfar) {
        double t[6], acc[6];
        nw_dipole(R, t);
        for (int c = 0; c < 6; c++) acc[c] = 0.5 * t[c];
        for (int a = 0; a < 3; a++)
            for (int sgn = -1; sgn <= 1; sgn += 2) {
                double Q[3] = {R[0], R[1], R[2]};
                Q[a] += sgn * h[a];
                nw_dipole(Q, t);
                for (int c = 0; c < 6; c++) acc[c] += t[c] / 12;
            }
        for (int c = 0; c < 6; c++) N[c] = -V / (4 * OR_PI) * acc[c];
        return;
    }
    const double k = 1.0 / (4 * OR_PI * V);
    const double X = R[0], Y = R[1], Z = R[2], a = h[0], b = h[1], c = h[2];
    N[0] = k * nw_stencil(nw_f, X, Y, Z, a, b, c);
    N[1] = k * nw_stencil(nw_f, Y, X, Z, b, a, c);
    N[2] = k * nw_stencil(nw_f, Z, Y, X, c, b, a);
    N[3] = k * nw_stencil(nw_g, Y, Z, X, b, c, a);
    N[4] = k * nw_stencil(nw_g, X, Z, Y, a, c, b);
    N[5] = k * nw_stencil(nw_g, X, Y, Z, a, b, c);
}
/* the full padded kernel, every element summed over its images directly
 * (no octant symmetry used) */
void or_kernel_newell(const int size[3], const double cs[3], const int pbc[3], double* out) {
    const long P = (long)size[0] * size[1] * size[2];
    int r1[3], r2[3];
    for (int c = 0; c < 3; c++) {
        if (pbc[c] == 0) { r1[c] = -(size[c] - 1) / 2; r2[c] = (size[c] - 1) / 2; }
        else { r1[c] = -(size[c] * pbc[c] - 1); r2[c] = (size[c] * pbc[c] - 1); }
    }
    if (size[2] == 1 && pbc[2] == 0) r2[2] = 0;
#pragma omp parallel for schedule(dynamic, 16)
    for (long g = 0; g < P; g++) {
        const int zw = (int)(g % size[2]), yw = (int)((g / size[2]) % size[1]), xw = (int)(g / ((long)size[1] * size[2]));
        double acc[6] = {0, 0, 0, 0, 0, 0};
        for (int z = r1[2]; z <= r2[2]; z++) {
            if (wrapi(z, size[2]) != zw) continue;
            for (int y = r1[1]; y <= r2[1]; y++) {
                if (wrapi(y, size[1]) != yw) continue;
                for (int x = r1[0]; x <= r2[0]; x++) {
                    if (wrapi(x, size[0]) != xw) continue;
                    const double R[3] = {x * cs[0], y * cs[1], z * cs[2]};
                    double N[6];
                    or_newell_tensor(R, cs, N);
                    for (int c = 0; c < 6; c++) acc[c] -= N[c];
                }
            }
        }
        for (int c = 0; c < 6; c++) out[c * P + g] = acc[c];
    }
}

/* ------------------------------------------------------------------ demag plan */
struct or_demag {
    int N[3], P[3], K[3];
    long Ncell, Pcell, Kcell;
    double* kern;    /* 6 octants */
    double* spec[3]; /* (P0,P1,H) complex */
    double* Mbuf;    /* 3N */
    double max_imag_ratio;
};

/* MaxwellPlan.init/loadDipoleKernel/LoadKernel/extract,
 * src/hotspin-core/modules/maxwell_convolution.go:42-100,135-223,350-388;
 * PadSize, src/hotspin-core/engine/zeropad.go:12-22. */
or_demag* or_demag_new(const int N[3], const double cell[3], const int pbc[3], int accuracy) {
    return or_demag_new_kernel(N, cell, pbc, accuracy, 0);
}
/* kernel: 0 = Kernel_Arne (accuracy), 1 = Newell */
or_demag* or_demag_new_kernel(const int N[3], const double cell[3], const int pbc[3], int accuracy, int kernel) {
    or_demag* d = (or_demag*)calloc(1, sizeof(or_demag));
    for (int c = 0; c < 3; c++) {
        d->N[c] = N[c];
        d->P[c] = (N[c] > 1 && pbc[c] == 0) ? 2 * N[c] : N[c];
    }
    d->K[0] = d->P[0] / 2 + 1;
    d->K[1] = d->P[1] / 2 + 1;
    d->K[2] = (d->P[2] + 2) / 2;
    d->Ncell = (long)N[0] * N[1] * N[2];
    d->Pcell = (long)d->P[0] * d->P[1] * d->P[2];
    d->Kcell = (long)d->K[0] * d->K[1] * d->K[2];
    long H = d->P[2] / 2 + 1;
    long specn = (long)d->P[0] * d->P[1] * H * 2;
    for (int c = 0; c < 3; c++) d->spec[c] = (double*)malloc(sizeof(double) * specn);
    d->Mbuf = (double*)malloc(sizeof(double) * 3 * d->Ncell);
    d->kern = (double*)calloc(6 * d->Kcell, sizeof(double));

    double* real = (double*)malloc(sizeof(double) * 6 * d->Pcell);
    if (kernel == 1) or_kernel_newell(d->P, cell, pbc, real);
    else or_kernel_arne(d->P, cell, pbc, accuracy, real);
    /* IsZero (maxwell_convolution.go:227-232) uses the max over all components */
    double gmax = 0;
    for (long i = 0; i < 6 * d->Pcell; i++) if (fabs(real[i]) > gmax) gmax = fabs(real[i]);
    double maxImg = 0, maxReal = 0;
    double norm = 1.0 / (double)d->Pcell;
    for (int c = 0; c < 6; c++) {
        double* comp = real + c * d->Pcell;
        double cmax = 0;
        for (long i = 0; i < d->Pcell; i++) if (fabs(comp[i]) > cmax) cmax = fabs(comp[i]);
        if (cmax / gmax < 1e-12) continue; /* stays zero, as a nil kernel array */
        r2c3d(comp, d->P, d->P, d->spec[0]);
        double* dst = d->kern + c * d->Kcell;
        for (int i = 0; i < d->K[0]; i++)
            for (int j = 0; j < d->K[1]; j++)
                for (int k = 0; k < d->K[2]; k++) {
                    const double* s = d->spec[0] + (((long)i * d->P[1] + j) * H + k) * 2;
                    dst[((long)i * d->K[1] + j) * d->K[2] + k] = s[0] * norm;
                    if (fabs(s[1]) > maxImg) maxImg = fabs(s[1]);
                    if (fabs(s[0]) > maxReal) maxReal = fabs(s[0]);
                }
    }
    d->max_imag_ratio = maxReal > 0 ? maxImg / maxReal : 0;
    free(real);
    return d;
}
void or_demag_free(or_demag* d) {
    if (!d) return;
    for (int c = 0; c < 3; c++) free(d->spec[c]);
    free(d->Mbuf); free(d->kern); free(d);
}
void or_demag_sizes(const or_demag* d, int P[3], int K[3]) {
    for (int c = 0; c < 3; c++) { P[c] = d->P[c]; K[c] = d->K[c]; }
}
const double* or_demag_kernel(const or_demag* d) { return d->kern; }
double or_demag_max_imag_ratio(const or_demag* d) { return d->max_imag_ratio; }

/* MaxwellPlan.UpdateB, maxwell_convolution.go:103-131 */
void or_demag_B(or_demag* d, const double* m, const double* msat0T0, double msat0T0Mul, double* B) {
    long N = d->Ncell;
    for (int c = 0; c < 3; c++) {
        if (msat0T0) or_mul(d->Mbuf + c * N, m + c * N, msat0T0, (int)N);
        else memcpy(d->Mbuf + c * N, m + c * N, sizeof(double) * N);
    }
    double mMul = msat0T0Mul * OR_MU0;
    for (int c = 0; c < 3; c++) r2c3d(d->Mbuf + c * N, d->N, d->P, d->spec[c]);
    const double* k = d->kern;
    long Kc = d->Kcell;
    int H = d->P[2] / 2 + 1;
    (void)H;
    or_tens_symm_vec_mul(d->spec[0], d->spec[1], d->spec[2], d->spec[0], d->spec[1], d->spec[2],
                         k + 0 * Kc, k + 1 * Kc, k + 2 * Kc, k + 3 * Kc, k + 4 * Kc, k + 5 * Kc,
                         mMul, d->P[0], d->P[1], d->K[2]);
    for (int c = 0; c < 3; c++) c2r3d(d->spec[c], d->P, d->N, B + c * N);
}

/* ------------------------------------------------------------------ Philox */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): multipliers 0xD2511F53,
 * 0xCD9E8D57, Weyl key increments 0x9E3779B9, 0xBB67AE85, ten rounds. */
void or_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 24-bit uniform in (0,1): exact in FP32 and FP64, never 0 or 1. */
static inline double u24(uint32_t x) { return ((double)(x >> 8) + 0.5) * (1.0 / 16777216.0); }

/* Box-Muller on two uniforms: r = sqrt(-2 ln u0), angle 2 pi u1. */
void or_thermal_normals(uint64_t seed, uint32_t stream, uint32_t step, uint32_t attempt,
                        long N, double* xi) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
#pragma omp parallel for schedule(static)
    for (long I = 0; I < N; I++) {
        uint32_t ctr[4] = {(uint32_t)I, step, attempt, stream}, o[4];
        or_philox4x32_10(ctr, key, o);
        double r0 = sqrt(-2.0 * log(u24(o[0])));
        double a0 = 2.0 * OR_PI * u24(o[1]);
        double r1 = sqrt(-2.0 * log(u24(o[2])));
        double a1 = 2.0 * OR_PI * u24(o[3]);
        xi[I] = r0 * cos(a0);
        xi[N + I] = r0 * sin(a0);
        xi[2 * N + I] = r1 * cos(a1);
    }
}

/* ------------------------------------------------------------------ LLB-r problem */
/* Test hook: a factor on the Brown prefactor 2 kB T / (mu0 V gamma dt Ms0)
 * of or_llb_rhs, 1 except in tests that plant a wrong amplitude on purpose
 * to show that the fluctuation-dissipation checks catch it. */
static double g_thermal_prefactor_scale = 1.0;
void or_debug_set_thermal_prefactor_scale(double f) { g_thermal_prefactor_scale = f; }

struct or_llb {
    or_llb_params p;
    long N;
    or_demag* demag;
    double *H, *Hex, *B, *Hlf, *Hth, *tau, *loc;
    double *k1, *k2, *k3, *k4, *y0, *ys;
};

or_llb* or_llb_new(const or_llb_params* p) {
    or_llb* s = (or_llb*)calloc(1, sizeof(or_llb));
    s->p = *p;
    s->N = (long)p->N[0] * p->N[1] * p->N[2];
    long n3 = 3 * s->N;
    if (p->demag)
        s->demag = or_demag_new_kernel(p->N, p->cell, p->pbc, p->demag_accuracy > 0 ? p->demag_accuracy : 6,
                                       p->demag_kernel);
    double** bufs[] = {&s->H, &s->Hex, &s->B, &s->Hlf, &s->Hth, &s->tau, &s->loc,
                       &s->k1, &s->k2, &s->k3, &s->k4, &s->y0, &s->ys};
    for (unsigned i = 0; i < sizeof(bufs) / sizeof(bufs[0]); i++) *bufs[i] = (double*)calloc(n3, sizeof(double));
    return s;
}
void or_llb_free(or_llb* s) {
    if (!s) return;
    or_demag_free(s->demag);
    double* bufs[] = {s->H, s->Hex, s->B, s->Hlf, s->Hth, s->tau, s->loc, s->k1, s->k2, s->k3, s->k4, s->y0, s->ys};
    for (unsigned i = 0; i < sizeof(bufs) / sizeof(bufs[0]); i++) free(bufs[i]);
    free(s);
}
const double* or_llb_heff(const or_llb* s) { return s->H; }
const double* or_llb_bdemag(const or_llb* s) { return s->B; }

/* llbar_RHS.Update (engine/quant.go:417-445 dependency walk):
 *   m normalizer (mod_normalize.go:53-55) -> H_eff SumUpdater (update_sum.go:29-75)
 *   parents H_ex, B/mu0, B_ext/mu0, H_lf, H_therm -> llbar_RHS = gamma*torque + gamma*local00nc. */
void or_llb_rhs(or_llb* s, double* m, double* k, uint32_t step, uint32_t attempt, double dt) {
    const or_llb_params* p = &s->p;
    long N = s->N;
    int n = (int)N;
    double* mx = m; double* my = m + N; double* mz = m + 2 * N;
    if (p->normalize) or_normalize(mx, my, mz, n);
    double* H = s->H;
    memset(H, 0, sizeof(double) * 3 * N);
    double* Hc[3] = {H, H + N, H + 2 * N};
    if (p->exchange) {
        double lexMul2 = p->lexMul * p->lexMul;
        double* e = s->Hex;
        or_exchange6(e, e + N, e + 2 * N, mx, my, mz, p->lex, p->N[0], p->N[1], p->N[2],
                     p->pbc[0], p->pbc[1], p->pbc[2], p->msat0T0Mul,
                     lexMul2 / (p->cell[0] * p->cell[0]), lexMul2 / (p->cell[1] * p->cell[1]),
                     lexMul2 / (p->cell[2] * p->cell[2]));
        double* b[3] = {e, e + N, e + 2 * N};
        double w[3] = {1, 1, 1};
        or_vec_madd(Hc, Hc, b, w, n);
    }
    if (p->demag) {
        or_demag_B(s->demag, m, p->msat0T0, p->msat0T0Mul, s->B);
        double* b[3] = {s->B, s->B + N, s->B + 2 * N};
        double w[3] = {1 / OR_MU0, 1 / OR_MU0, 1 / OR_MU0};
        or_vec_madd(Hc, Hc, b, w, n);
    }
    if (p->zeeman) {
        double* b[3] = {NULL, NULL, NULL};
        double w[3] = {1 / OR_MU0 * p->Bext[0], 1 / OR_MU0 * p->Bext[1], 1 / OR_MU0 * p->Bext[2]};
        or_vec_madd(Hc, Hc, b, w, n);
    }
    if (p->longfield) {
        double* e = s->Hlf;
        or_long_field(e, e + N, e + 2 * N, mx, my, mz, p->msat0T0, p->J, p->n, p->Tc, p->T,
                      p->msat0T0Mul, p->JMul, p->nMul, p->TcMul, p->TMul, n);
        double* b[3] = {e, e + N, e + 2 * N};
        double w[3] = {1, 1, 1};
        or_vec_madd(Hc, Hc, b, w, n);
    }
    if (p->thermal) {
        /* AnizBrownUpdater.Update, modules/mod_brown.go:103-149 */
        double* e = s->Hth;
        or_thermal_normals(p->seed, p->stream, step, attempt, N, e);
        double V = p->cell[0] * p->cell[1] * p->cell[2];
        double KB2TMul = OR_KB * 2.0 * p->TMul;
        double deltat = fmax(p->cutoff_dt, dt);
        double denom = OR_MU0 * V * p->gammaLL * deltat * p->msat0T0Mul;
        double K = KB2TMul / denom * g_thermal_prefactor_scale;
        or_temperature_scale_noise(e, e + N, e + 2 * N, p->alpha[0], p->alpha[1], p->alpha[2],
                                   p->T, p->msat0T0, p->alphaMul[0], p->alphaMul[1], p->alphaMul[2], K, n);
        double* b[3] = {e, e + N, e + 2 * N};
        double w[3] = {1, 1, 1};
        or_vec_madd(Hc, Hc, b, w, n);
    }
    memset(k, 0, sizeof(double) * 3 * N);
    double* kc[3] = {k, k + N, k + 2 * N};
    if (p->torque) {
        double* t = s->tau;
        or_llbar_torque(t, t + N, t + 2 * N, mx, my, mz, Hc[0], Hc[1], Hc[2], p->msat0T0, n);
        double* b[3] = {t, t + N, t + 2 * N};
        double w[3] = {p->gammaLL, p->gammaLL, p->gammaLL};
        or_vec_madd(kc, kc, b, w, n);
    }
    if (p->local00nc) {
        double* t = s->loc;
        or_llbar_local00nc(t, t + N, t + 2 * N, Hc[0], Hc[1], Hc[2], p->msat0T0,
                           p->alpha[0], p->alpha[1], p->alpha[2],
                           p->alphaMul[0], p->alphaMul[1], p->alphaMul[2], n);
        double* b[3] = {t, t + N, t + 2 * N};
        double w[3] = {p->gammaLL, p->gammaLL, p->gammaLL};
        or_vec_madd(kc, kc, b, w, n);
    }
}

static void axpy_to(double* dst, const double* y0, double h, const double* k, long n) {
    for (long i = 0; i < n; i++) dst[i] = y0[i] + h * k[i];
}

/* EulerSolver.Step, engine/solve_euler.go:22-50: y += dt*k */
int or_llb_step_euler(or_llb* s, double* m, double dt, uint32_t step) {
    long n3 = 3 * s->N;
    or_llb_rhs(s, m, s->k1, step, 0, dt);
    or_madd(m, m, s->k1, dt, (int)n3);
    return 1;
}

/* HeunSolver.Step, engine/solve_heun.go:27-79 */
int or_llb_step_heun(or_llb* s, double* m, double dt, uint32_t step) {
    long n3 = 3 * s->N;
    or_llb_rhs(s, m, s->k1, step, 0, dt);
    or_madd(m, m, s->k1, dt, (int)n3);
    or_llb_rhs(s, m, s->k2, step, 0, dt);
    or_madd2(m, s->k2, 0.5 * dt, s->k1, -0.5 * dt, (int)n3);
    return 1;
}

/* Classical RK4 (not in the reference; defined here in its lazy-normaliser
 * style: every stage state is normalised in place when read). */
int or_llb_step_rk4(or_llb* s, double* m, double dt, uint32_t step) {
    long n3 = 3 * s->N;
    or_llb_rhs(s, m, s->k1, step, 0, dt);          /* normalises m = y0 */
    axpy_to(s->ys, m, 0.5 * dt, s->k1, n3);
    or_llb_rhs(s, s->ys, s->k2, step, 0, dt);
    axpy_to(s->ys, m, 0.5 * dt, s->k2, n3);
    or_llb_rhs(s, s->ys, s->k3, step, 0, dt);
    axpy_to(s->ys, m, dt, s->k3, n3);
    or_llb_rhs(s, s->ys, s->k4, step, 0, dt);
    double h6 = dt / 6.0;
    for (long i = 0; i < n3; i++)
        m[i] = m[i] + h6 * (s->k1[i] + 2.0 * s->k2[i] + 2.0 * s->k3[i] + s->k4[i]);
    return 1;
}

/* |e| for the adaptive error estimates, NaN mapped to +inf so that a
 * diverged attempt can never look converged.  A non-finite estimate is never
 * accepted (the reference would take it when dt reaches mindt): the attempt
 * is retried at the smallest factor while retries remain and dt can still
 * shrink, otherwise the step fails. */
static inline double err_abs(double e) { double a = fabs(e); return a == a ? a : INFINITY; }

/* RK12Solver.Step, engine/solve_rk12.go:91-214 (single equation). */
int or_llb_step_rk12(or_llb* s, double* m, double* dtio, double maxErr,
                     double mindt, double maxdt, uint32_t step, double* stats) {
    long n3 = 3 * s->N;
    or_llb_rhs(s, m, s->k1, step, 0, *dtio);
    memcpy(s->y0, m, sizeof(double) * n3);
    const int maxTry = 10;
    const double headRoom = 0.8;
    int tr = 0;
    double err = 0;
    for (;;) {
        double dt = *dtio;
        if (tr > 0) memcpy(m, s->y0, sizeof(double) * n3);
        or_madd(m, m, s->k1, dt, (int)n3);
        or_llb_rhs(s, m, s->k2, step, (uint32_t)tr, dt);
        double h = dt;
        or_madd2(m, s->k2, 0.5 * h, s->k1, -0.5 * h, (int)n3);
        err = 0;
        for (long i = 0; i < n3; i++) {
            double d = err_abs(s->k2[i] - s->k1[i]);
            if (d > err) err = d;
        }
        err *= h;
        int bad = err > maxErr;
        double factor = bad ? pow(maxErr / err, 1. / 3.) : sqrt(maxErr / err);
        factor *= headRoom;
        if (factor > 1.5) factor = 1.5;
        if (factor < 0.1) factor = 0.1;
        double newDt = dt * factor;
        if (newDt < mindt) newDt = mindt;
        if (newDt > maxdt) newDt = maxdt;
        *dtio = newDt;
        if (!isfinite(err)) {
            if (tr > maxTry || !(newDt < dt)) return -1;
            tr++;
            continue;
        }
        if (!bad || newDt == mindt) break;
        if (tr > maxTry) return -1;
        tr++;
    }
    if (stats) { stats[0] = err; stats[1] = tr + 1; }
    return 1;
}

/* Bogacki-Shampine RK23 (not in the reference), with the reference's step
 * controller (headroom 0.8, factor clamped to [0.1,1.5], mindt/maxdt, at most
 * 10 retries) and exponent 1/3.  err = h*max|(-5/72)k1+(1/12)k2+(1/9)k3-(1/8)k4|
 * over all 3N scalars.  Attempt a of step n uses thermal noise (n, a) in every
 * stage (the reference redraws noise on a retry, mod_brown.go:71). */
int or_llb_step_rk23(or_llb* s, double* m, double* dtio, double maxErr,
                     double mindt, double maxdt, uint32_t step, double* stats) {
    long n3 = 3 * s->N;
    const int maxTry = 10;
    const double headRoom = 0.8;
    int tr = 0;
    double err = 0;
    memcpy(s->y0, m, sizeof(double) * n3);
    for (;;) {
        double dt = *dtio, h = dt;
        memcpy(m, s->y0, sizeof(double) * n3);
        or_llb_rhs(s, m, s->k1, step, (uint32_t)tr, dt); /* normalises y0 */
        if (tr == 0) memcpy(s->y0, m, sizeof(double) * n3);
        axpy_to(s->ys, m, 0.5 * h, s->k1, n3);
        or_llb_rhs(s, s->ys, s->k2, step, (uint32_t)tr, dt);
        axpy_to(s->ys, m, 0.75 * h, s->k2, n3);
        or_llb_rhs(s, s->ys, s->k3, step, (uint32_t)tr, dt);
        for (long i = 0; i < n3; i++)
            m[i] = m[i] + h * ((2.0 / 9.0) * s->k1[i] + (1.0 / 3.0) * s->k2[i] + (4.0 / 9.0) * s->k3[i]);
        or_llb_rhs(s, m, s->k4, step, (uint32_t)tr, dt); /* normalises y1 */
        err = 0;
        for (long i = 0; i < n3; i++) {
            double e = h * ((-5.0 / 72.0) * s->k1[i] + (1.0 / 12.0) * s->k2[i] +
                            (1.0 / 9.0) * s->k3[i] - (1.0 / 8.0) * s->k4[i]);
            if (err_abs(e) > err) err = err_abs(e);
        }
        int bad = err > maxErr;
        double factor = headRoom * pow(maxErr / err, 1. / 3.);
        if (factor > 1.5) factor = 1.5;
        if (factor < 0.1) factor = 0.1;
        double newDt = dt * factor;
        if (newDt < mindt) newDt = mindt;
        if (newDt > maxdt) newDt = maxdt;
        *dtio = newDt;
        if (!isfinite(err)) {
            if (tr > maxTry || !(newDt < dt)) return -1;
            tr++;
            continue;
        }
        if (!bad || newDt == mindt) break;
        if (tr > maxTry) return -1;
        tr++;
    }
    if (stats) { stats[0] = err; stats[1] = tr + 1; }
    return 1;
}

/* ------------------------------------------------------------------ implicit drivers */
/* Error of one fixed-point update (solve_am01.go:122-132): with the maxima
 * over cells per component p,
 *   tErr = (1/sqrt(3)) sqrt(sum_p (max|y_p - yb_p| / (absErr + max|yb_p| relErr))^2). */
static double am_norm(const double* y, const double* yb, long N, double absErr, double relErr) {
    double t = 0.0;
    for (int p = 0; p < 3; p++) {
        double d = or_maxdiff(y + p * N, yb + p * N, N);
        double mx = or_maxabs(yb + p * N, N);
        t += pow(d / (absErr + mx * relErr), 2.0);
    }
    return (1.0 / sqrt(3.0)) * sqrt(t);
}

/* BDFAM12.Step, engine/solve_am01.go:45-300, for the single equation m:
 * explicit-Euler start, implicit-Euler predictor and trapezoidal corrector,
 * each iterated to a fixed point (tErr <= 0.1, at most 6 updates, stop when
 * the contraction alpha = tErr/tErr_prev reaches 1); the step error is the
 * corrector-predictor difference; dt *= (0.8/err)^(1/2), restricted to
 * (0.6/alpha)^(1/2) after an iterator failure; at most 10 retries.  Noise
 * attempt a increments whenever dt changes within the step (the reference
 * redraws H_therm when dt changes, mod_brown.go:71).
 * stats: [0] step error, [1] attempts (rejected steps + 1), [2] fixed-point
 * iterations, [3] RHS evaluations, [4] dt of the accepted step. */
int or_llb_step_am12(or_llb* s, double* m, double* dtio, double absErr, double relErr,
                     double mindt, double maxdt, uint32_t step, double* stats) {
    const long N = s->N, n3 = 3 * N;
    const int maxTry = 10, maxIter = 5;
    const double headRoom = 0.8, maxIterErr = 0.1, alpha_ref = 0.6;
    double *dy0 = s->k1, *dy = s->k2, *dyE = s->k3, *y1 = s->k4, *yb = s->ys;
    uint32_t a = 0;
    long evals = 0, iterations = 0;
    or_llb_rhs(s, m, dy0, step, a, *dtio); /* normalises m = y0 */
    evals++;
    memcpy(s->y0, m, sizeof(double) * n3);
    int tr = 0, restrict_step = 0;
    double saved_alpha = 0.0, tErr = 0.0, used_dt = 0.0;
    for (;;) {
        double dt = *dtio, h = dt;
        int badIt = 0;
        double er, alpha = 0.0;
        for (long i = 0; i < n3; i++) m[i] = s->y0[i] + h * dy0[i];
        iterations++;
        or_llb_rhs(s, m, dy, step, a, dt);
        evals++;
        memcpy(dyE, dy, sizeof(double) * n3);
        er = maxIterErr;
        for (int iter = 0;;) {
            for (long i = 0; i < n3; i++) yb[i] = s->y0[i] + h * dy[i];
            double t = am_norm(m, yb, N, absErr, relErr);
            alpha = t / er;
            saved_alpha = alpha;
            er = t;
            memcpy(m, yb, sizeof(double) * n3);
            iter++;
            iterations++;
            if (er <= maxIterErr) break;
            if (alpha >= 1.0 || iter > maxIter) { badIt = 1; break; }
            or_llb_rhs(s, m, dy, step, a, dt);
            evals++;
        }
        if (badIt && tr == maxTry - 1) return -1;
        if (badIt) {
            double ha = 0.5 * dt;
            if (alpha > alpha_ref) ha = dt * pow(alpha_ref / alpha, 0.5);
            *dtio = ha;
            restrict_step = 1;
            a++;
            continue;
        }
        memcpy(y1, m, sizeof(double) * n3);
        memcpy(dy, dyE, sizeof(double) * n3);
        er = maxIterErr;
        for (int iter = 0;;) {
            for (long i = 0; i < n3; i++) yb[i] = s->y0[i] + (0.5 * h) * (dy[i] + dy0[i]);
            double t = am_norm(m, yb, N, absErr, relErr);
            alpha = t / er;
            er = t;
            memcpy(m, yb, sizeof(double) * n3);
            iter++;
            iterations++;
            if (er <= maxIterErr) break;
            if (alpha >= 1.0 || iter > maxIter) { badIt = 1; break; }
            or_llb_rhs(s, m, dy, step, a, dt);
            evals++;
        }
        if (badIt && tr == maxTry - 1) return -1;
        if (badIt) {
            double ha = 0.5 * dt;
            if (alpha > alpha_ref) ha = dt * pow(alpha_ref / alpha, 0.5);
            *dtio = ha;
            a++;
            continue;
        }
        tErr = am_norm(m, y1, N, absErr, relErr);
        used_dt = dt;
        int bad = tErr > 1.0;
        double new_dt = dt * pow(headRoom / tErr, 0.5);
        if (restrict_step) {
            double ha = dt * pow(alpha_ref / saved_alpha, 0.5);
            new_dt = fmin(new_dt, ha);
            restrict_step = 0;
        }
        if (new_dt < mindt) new_dt = mindt;
        if (new_dt > maxdt) new_dt = maxdt;
        *dtio = new_dt;
        if (!bad || new_dt == mindt) break;
        if (tr > maxTry) return -1;
        tr++;
        a++;
    }
    if (stats) {
        stats[0] = tErr; stats[1] = tr + 1; stats[2] = (double)iterations; stats[3] = (double)evals;
        stats[4] = used_dt;
    }
    return 1;
}

/* BDFEuler.Step, engine/solve_euler_backward.go:29-86, for the single equation
 * m: explicit-Euler start, then y <- y0 + dt f(y) until
 * max|y - yb| / max|y + yb| <= maxIterErr (over all 3N values); more than
 * maxIter updates fail.  The RHS of the last iterate is not needed and not
 * evaluated.  stats: [0] last iterator error, [1] iterations, [2] RHS evaluations. */
int or_llb_step_bdf_euler(or_llb* s, double* m, double dt, uint32_t step, double maxIterErr, int maxIter,
                          double* stats) {
    const long n3 = 3 * s->N;
    double *dy = s->k1, *yb = s->ys;
    long evals = 0, iterations = 0;
    or_llb_rhs(s, m, dy, step, 0, dt); /* normalises m = y0 */
    evals++;
    memcpy(s->y0, m, sizeof(double) * n3);
    for (long i = 0; i < n3; i++) m[i] = s->y0[i] + dt * dy[i];
    iterations++;
    double err = 1.0e38;
    int iter = 0;
    while (err > maxIterErr) {
        or_llb_rhs(s, m, dy, step, 0, dt);
        evals++;
        for (long i = 0; i < n3; i++) yb[i] = s->y0[i] + dt * dy[i];
        err = or_maxdiff(m, yb, n3);
        double sum = or_maxsum(m, yb, n3);
        if (sum > 0.0) err = err / sum;
        iter++;
        iterations++;
        memcpy(m, yb, sizeof(double) * n3);
        if (iter > maxIter) return -1;
    }
    if (stats) { stats[0] = err; stats[1] = (double)iterations; stats[2] = (double)evals; }
    return 1;
}

/* ------------------------------------------------------------------ two-temperature model */
/* Qe summed in the order Qspat, Qe-p, Qlaser; Qspat = or_qspat with a
 * uniform conductance (Qspat.cu:13-134), Qe-p = or_qinter (Qinter.cu:13-45). */
/* Debye Cp of one cell with n = 1 (cpKern, Cp.cu:14-83, as or_cp) */
static double cp_debye(double T, double Td) {
    double out;
    or_cp(&out, &T, NULL, NULL, Td, 1);
    return out;
}

/* Spin contribution to Qe at the state of the last or_llb_rhs call on m
 * (H = s->H, R = llbar_local00nc = s->loc): the molecular energy-density
 * rate ws (mod_molecular-energy-density-rate.go:77-101 with energyFlowKern,
 * energy_flow.cu:15-48) minus the dissipation function qs
 * (mod_energy-density-dissipation-rate.go:71-92), laser-demag.py:84-86:
 *   ws = 3 kB TcMul nMul gammaLL [Tc S^2 n / (S (S+1))] (m . R)
 *   qs = -mu0 msat0T0Mul gammaLL (H . R) msat0T0 */
static void or_spin_heat(const or_llb* s, const double* m, double* Qs);

void or_tm_rhs(const or_tm_params* q, const int N[3], const double cell[3], const int pbc[3], const double* Te,
               const double* Tp, double t, const double* Qs, double* dTe, double* dTp, double* work) {
    long n = (long)N[0] * N[1] * N[2];
    or_qspat(work, Te, NULL, q->kappa, N[0], N[1], N[2], cell[0], cell[1], cell[2], pbc[0], pbc[1], pbc[2]);
    const double ir2 = q->sigma_r > 0 ? 1.0 / (2.0 * q->sigma_r * q->sigma_r) : 0.0;
    const double it2 = q->sigma_t > 0 ? 1.0 / (2.0 * q->sigma_t * q->sigma_t) : 0.0;
    const double d0 = t - q->t0;
#pragma omp parallel for schedule(static)
    for (long I = 0; I < n; I++) {
        const int j = (int)((I / N[2]) % N[1]), k = (int)(I % N[2]);
        const double x = (k + 0.5) * cell[2] - q->xc, y = (j + 0.5) * cell[1] - q->yc;
        const double ql = q->Q0 * exp(-(x * x + y * y) * ir2) * exp(-(d0 * d0) * it2);
        double qe = work[I] + q->G * (Tp[I] - Te[I]) + ql;
        if (Qs) qe += Qs[I];
        const double cl = q->Td > 0.0 ? (9.0 * q->nD * OR_KB) * cp_debye(Tp[I], q->Td) : q->Cl;
        dTe[I] = qe / (q->gammaE * Te[I]);
        dTp[I] = q->G * (Te[I] - Tp[I]) / cl;
    }
}

static void or_spin_heat(const or_llb* s, const double* m, double* Qs) {
    const or_llb_params* p = &s->p;
    const long N = s->N;
    const double cw = 3.0 * OR_KB * p->TcMul * p->nMul * p->gammaLL;
    const double cq = OR_MU0 * p->msat0T0Mul * p->gammaLL;
    for (long i = 0; i < N; i++) {
        const double* R = s->loc;
        const double* H = s->H;
        double mR = m[i] * R[i] + m[N + i] * R[N + i] + m[2 * N + i] * R[2 * N + i];
        double HR = H[i] * R[i] + H[N + i] * R[N + i] + H[2 * N + i] * R[2 * N + i];
        double S = p->JMul * mask1(p->J, i), Tc = mask1(p->Tc, i), n = mask1(p->n, i);
        double ws = (n == 0.0 || S == 0.0 || Tc == 0.0) ? 0.0 : cw * (Tc * S * S * n / (S * (S + 1.0))) * mR;
        double qs = -cq * HR * mask1(p->msat0T0, i);
        Qs[i] = ws - qs;
    }
}

void or_llb_spin_heat(const or_llb* s, const double* m, double* Qs) { or_spin_heat(s, m, Qs); }

/* the spin term of Qe at the state of the or_llb_rhs call just made on X
 * (or NULL without spin heating); needs a buffer Qsb of N doubles in scope */
#define SPIN_QS(X) (q->spin_heat ? (or_spin_heat(s, (X), Qsb), (const double*)Qsb) : (const double*)NULL)

static double maxabs_err(const double* k1, const double* k2, const double* k3, const double* k4, double h, long n) {
    double e = 0;
    for (long i = 0; i < n; i++) {
        double v = h * ((-5.0 / 72.0) * k1[i] + (1.0 / 12.0) * k2[i] + (1.0 / 9.0) * k3[i] - (1.0 / 8.0) * k4[i]);
        if (err_abs(v) > e) e = err_abs(v);
    }
    return e;
}

int or_llbtm_step_rk23(or_llb* s, const or_tm_params* q, double* m, double* Te, double* Tp, double* t,
                       double* dtio, double maxErr, double mindt, double maxdt, uint32_t step, double* stats) {
    const long N = s->N, n3 = 3 * N;
    const int maxTry = 10;
    const double headRoom = 0.8;
    double* buf = (double*)malloc(sizeof(double) * 15 * N);
    double *Te0 = buf, *Tp0 = buf + N, *Tes = buf + 2 * N, *Tps = buf + 3 * N, *work = buf + 4 * N;
    double* Qsb = buf + 14 * N;
    double* ke[4] = {buf + 5 * N, buf + 6 * N, buf + 7 * N, buf + 8 * N};
    double* kp[4] = {buf + 9 * N, buf + 10 * N, buf + 11 * N, buf + 12 * N};
    memcpy(Te0, Te, sizeof(double) * N);
    memcpy(Tp0, Tp, sizeof(double) * N);
    memcpy(s->y0, m, sizeof(double) * n3);
    const double* Tsave = s->p.T;
    const double TMsave = s->p.TMul;
    s->p.TMul = 1.0;
    int tr = 0, rc = 1;
    double err[3] = {0, 0, 0};
    const double cs[4] = {0.0, 0.5, 0.75, 1.0};
    for (;;) {
        const double dt = *dtio, h = dt;
        memcpy(m, s->y0, sizeof(double) * n3);
        /* stage 1 at (y0, t) */
        s->p.T = Te0;
        or_llb_rhs(s, m, s->k1, step, (uint32_t)tr, dt); /* normalises y0 */
        if (tr == 0) memcpy(s->y0, m, sizeof(double) * n3);
        or_tm_rhs(q, s->p.N, s->p.cell, s->p.pbc, Te0, Tp0, *t + cs[0] * h, SPIN_QS(m), ke[0], kp[0], work);
        /* stage 2 at y0 + h/2 k1 */
        axpy_to(s->ys, m, 0.5 * h, s->k1, n3);
        axpy_to(Tes, Te0, 0.5 * h, ke[0], N);
        axpy_to(Tps, Tp0, 0.5 * h, kp[0], N);
        s->p.T = Tes;
        or_llb_rhs(s, s->ys, s->k2, step, (uint32_t)tr, dt);
        or_tm_rhs(q, s->p.N, s->p.cell, s->p.pbc, Tes, Tps, *t + cs[1] * h, SPIN_QS(s->ys), ke[1], kp[1], work);
        /* stage 3 at y0 + 3h/4 k2 */
        axpy_to(s->ys, m, 0.75 * h, s->k2, n3);
        axpy_to(Tes, Te0, 0.75 * h, ke[1], N);
        axpy_to(Tps, Tp0, 0.75 * h, kp[1], N);
        s->p.T = Tes;
        or_llb_rhs(s, s->ys, s->k3, step, (uint32_t)tr, dt);
        or_tm_rhs(q, s->p.N, s->p.cell, s->p.pbc, Tes, Tps, *t + cs[2] * h, SPIN_QS(s->ys), ke[2], kp[2], work);
        /* y1 and stage 4 at (y1, t + h) */
        for (long i = 0; i < n3; i++)
            m[i] = m[i] + h * ((2.0 / 9.0) * s->k1[i] + (1.0 / 3.0) * s->k2[i] + (4.0 / 9.0) * s->k3[i]);
        for (long i = 0; i < N; i++) {
            Tes[i] = Te0[i] + h * ((2.0 / 9.0) * ke[0][i] + (1.0 / 3.0) * ke[1][i] + (4.0 / 9.0) * ke[2][i]);
            Tps[i] = Tp0[i] + h * ((2.0 / 9.0) * kp[0][i] + (1.0 / 3.0) * kp[1][i] + (4.0 / 9.0) * kp[2][i]);
        }
        s->p.T = Tes;
        or_llb_rhs(s, m, s->k4, step, (uint32_t)tr, dt); /* normalises y1 */
        or_tm_rhs(q, s->p.N, s->p.cell, s->p.pbc, Tes, Tps, *t + cs[3] * h, SPIN_QS(m), ke[3], kp[3], work);
        err[0] = maxabs_err(s->k1, s->k2, s->k3, s->k4, h, n3);
        err[1] = maxabs_err(ke[0], ke[1], ke[2], ke[3], h, N);
        err[2] = maxabs_err(kp[0], kp[1], kp[2], kp[3], h, N);
        /* solve_rk12.go:150-190: bad if any equation exceeds its bound, smallest factor */
        int bad = 0;
        double minF = 2.0;
        for (int e = 0; e < 3; e++) {
            const double me = e ? q->maxErrT : maxErr;
            if (err[e] > me) bad = 1;
            double f = headRoom * pow(me / err[e], 1. / 3.);
            if (f > 1.5) f = 1.5;
            if (f < 0.1) f = 0.1;
            if (f < minF) minF = f;
        }
        double newDt = dt * minF;
        if (newDt < mindt) newDt = mindt;
        if (newDt > maxdt) newDt = maxdt;
        *dtio = newDt;
        const int finite = isfinite(err[0]) && isfinite(err[1]) && isfinite(err[2]);
        if (!finite && (tr > maxTry || !(newDt < dt))) {
            rc = -1;
            break;
        }
        if (finite && (!bad || newDt == mindt)) {
            memcpy(Te, Tes, sizeof(double) * N);
            memcpy(Tp, Tps, sizeof(double) * N);
            *t += h;
            break;
        }
        if (tr > maxTry) {
            rc = -1;
            break;
        }
        tr++;
    }
    s->p.T = Tsave;
    s->p.TMul = TMsave;
    free(buf);
    if (stats) {
        stats[0] = err[0];
        stats[1] = err[1];
        stats[2] = err[2];
        stats[3] = tr + 1;
    }
    return rc;
}

/* Error of a one-component equation (solve_am01.go:122-132 with COMP = 1). */
static double t_norm(const double* y, const double* yb, long N, double absErr, double relErr) {
    double d = or_maxdiff(y, yb, N), mx = or_maxabs(yb, N);
    double t = pow(d / (absErr + mx * relErr), 2.0);
    return (1.0 / sqrt(1.0)) * sqrt(t);
}
static void sort3(double* v) {
    for (int i = 0; i < 3; i++)
        for (int j = i + 1; j < 3; j++)
            if (v[j] < v[i]) { double x = v[i]; v[i] = v[j]; v[j] = x; }
}

/* BDFAM12.Step, engine/solve_am01.go:45-300, for the equations of the
 * laser recipe in their registration order (examples/laser-demag.py:35-60:
 * temperature/ETM, temperature/PTM, then llbar): e = 0 Te, 1 Tp, 2 m, with
 * the LLB temperature T = Te.  Every fixed-point update takes the RHS of
 * all equations at the previous iterate (one UpdateEqRHS per iteration).
 * As in the reference, the per-equation previous errors er[] and
 * contractions alp[] are sorted in place after each iteration, so er[i]
 * pairs with equation i by rank, not by name; the controller's
 * alpha-restriction applies to the first equation only (the flag is cleared
 * inside the loop, solve_am01.go:253-258).  absErr, relErr: per equation
 * (Te, Tp, m).  stats: [0..2] step errors (Te, Tp, m), [3] attempts, [4]
 * fixed-point iterations, [5] dt of the accepted step. */
int or_llbtm_step_am12(or_llb* s, const or_tm_params* q, double* m, double* Te, double* Tp, double* t,
                       double* dtio, const double absErr[3], const double relErr[3], double mindt, double maxdt,
                       uint32_t step, double* stats) {
    const long N = s->N, n3 = 3 * N;
    const int maxTry = 10, maxIter = 5;
    const double headRoom = 0.8, maxIterErr = 0.1, alpha_ref = 0.6;
    double* buf = (double*)malloc(sizeof(double) * 16 * N);
    double* Qsb = buf + 15 * N;
    double *Te0 = buf, *Tp0 = buf + N, *d0e = buf + 2 * N, *d0p = buf + 3 * N, *dEe = buf + 4 * N;
    double *dEp = buf + 5 * N, *ke = buf + 6 * N, *kp = buf + 7 * N, *y1e = buf + 8 * N, *y1p = buf + 9 * N;
    double *ybe = buf + 10 * N, *ybp = buf + 11 * N, *work = buf + 12 * N;
    double *dy0 = s->k1, *dy = s->k2, *dyE = s->k3, *y1 = s->k4, *yb = s->ys;
    const double* Tsave = s->p.T;
    const double TMsave = s->p.TMul;
    s->p.TMul = 1.0;
    const double t0 = *t;
    uint32_t a = 0;
    long iterations = 0;
    s->p.T = Te;
    or_llb_rhs(s, m, dy0, step, a, *dtio); /* normalises m = y0 */
    or_tm_rhs(q, s->p.N, s->p.cell, s->p.pbc, Te, Tp, t0, SPIN_QS(m), d0e, d0p, work);
    memcpy(s->y0, m, sizeof(double) * n3);
    memcpy(Te0, Te, sizeof(double) * N);
    memcpy(Tp0, Tp, sizeof(double) * N);
    int tr = 0, restrict_step = 0, rc = 1;
    double salpha[3] = {0, 0, 0}, tErr[3] = {0, 0, 0}, used = 0.0;
    for (;;) {
        const double dt = *dtio, h = dt;
        int badIt = 0;
        double er[3], alp[3], alpha = 0.0;
        for (long i = 0; i < N; i++) { Te[i] = Te0[i] + h * d0e[i]; Tp[i] = Tp0[i] + h * d0p[i]; }
        for (long i = 0; i < n3; i++) m[i] = s->y0[i] + h * dy0[i];
        iterations++;
        s->p.T = Te;
        or_llb_rhs(s, m, dy, step, a, dt);
        or_tm_rhs(q, s->p.N, s->p.cell, s->p.pbc, Te, Tp, t0 + dt, SPIN_QS(m), ke, kp, work);
        memcpy(dyE, dy, sizeof(double) * n3);
        memcpy(dEe, ke, sizeof(double) * N);
        memcpy(dEp, kp, sizeof(double) * N);
        for (int e = 0; e < 3; e++) er[e] = maxIterErr;
        for (int iter = 0;;) {
            double te;
            for (long i = 0; i < N; i++) ybe[i] = Te0[i] + h * ke[i];
            te = t_norm(Te, ybe, N, absErr[0], relErr[0]);
            alp[0] = te / er[0]; salpha[0] = alp[0]; er[0] = te;
            memcpy(Te, ybe, sizeof(double) * N);
            for (long i = 0; i < N; i++) ybp[i] = Tp0[i] + h * kp[i];
            te = t_norm(Tp, ybp, N, absErr[1], relErr[1]);
            alp[1] = te / er[1]; salpha[1] = alp[1]; er[1] = te;
            memcpy(Tp, ybp, sizeof(double) * N);
            for (long i = 0; i < n3; i++) yb[i] = s->y0[i] + h * dy[i];
            te = am_norm(m, yb, N, absErr[2], relErr[2]);
            alp[2] = te / er[2]; salpha[2] = alp[2]; er[2] = te;
            memcpy(m, yb, sizeof(double) * n3);
            sort3(er);
            sort3(alp);
            alpha = alp[2];
            iter++;
            iterations++;
            if (er[2] <= maxIterErr) break;
            if (alpha >= 1.0 || iter > maxIter) { badIt = 1; break; }
            s->p.T = Te;
            or_llb_rhs(s, m, dy, step, a, dt);
            or_tm_rhs(q, s->p.N, s->p.cell, s->p.pbc, Te, Tp, t0 + dt, SPIN_QS(m), ke, kp, work);
        }
        if (badIt && tr == maxTry - 1) { rc = -1; break; }
        if (badIt) {
            double ha = 0.5 * dt;
            if (alpha > alpha_ref) ha = dt * pow(alpha_ref / alpha, 0.5);
            *dtio = ha;
            restrict_step = 1;
            a++;
            continue;
        }
        memcpy(y1, m, sizeof(double) * n3);
        memcpy(y1e, Te, sizeof(double) * N);
        memcpy(y1p, Tp, sizeof(double) * N);
        memcpy(dy, dyE, sizeof(double) * n3);
        memcpy(ke, dEe, sizeof(double) * N);
        memcpy(kp, dEp, sizeof(double) * N);
        for (int e = 0; e < 3; e++) er[e] = maxIterErr;
        for (int iter = 0;;) {
            double te;
            for (long i = 0; i < N; i++) ybe[i] = Te0[i] + (0.5 * h) * (ke[i] + d0e[i]);
            te = t_norm(Te, ybe, N, absErr[0], relErr[0]);
            alp[0] = te / er[0]; er[0] = te;
            memcpy(Te, ybe, sizeof(double) * N);
            for (long i = 0; i < N; i++) ybp[i] = Tp0[i] + (0.5 * h) * (kp[i] + d0p[i]);
            te = t_norm(Tp, ybp, N, absErr[1], relErr[1]);
            alp[1] = te / er[1]; er[1] = te;
            memcpy(Tp, ybp, sizeof(double) * N);
            for (long i = 0; i < n3; i++) yb[i] = s->y0[i] + (0.5 * h) * (dy[i] + dy0[i]);
            te = am_norm(m, yb, N, absErr[2], relErr[2]);
            alp[2] = te / er[2]; er[2] = te;
            memcpy(m, yb, sizeof(double) * n3);
            sort3(er);
            sort3(alp);
            alpha = alp[2];
            iter++;
            iterations++;
            if (er[2] <= maxIterErr) break;
            if (alpha >= 1.0 || iter > maxIter) { badIt = 1; break; }
            s->p.T = Te;
            or_llb_rhs(s, m, dy, step, a, dt);
            or_tm_rhs(q, s->p.N, s->p.cell, s->p.pbc, Te, Tp, t0 + dt, SPIN_QS(m), ke, kp, work);
        }
        if (badIt && tr == maxTry - 1) { rc = -1; break; }
        if (badIt) {
            double ha = 0.5 * dt;
            if (alpha > alpha_ref) ha = dt * pow(alpha_ref / alpha, 0.5);
            *dtio = ha;
            a++;
            continue;
        }
        tErr[0] = t_norm(Te, y1e, N, absErr[0], relErr[0]);
        tErr[1] = t_norm(Tp, y1p, N, absErr[1], relErr[1]);
        tErr[2] = am_norm(m, y1, N, absErr[2], relErr[2]);
        int bad = 0;
        double nd[3];
        for (int e = 0; e < 3; e++) {
            if (tErr[e] > 1.0) bad = 1;
            double x = dt * pow(headRoom / tErr[e], 0.5);
            if (restrict_step) {
                double ha = dt * pow(alpha_ref / salpha[e], 0.5);
                x = fmin(x, ha);
                restrict_step = 0;
            }
            if (x < mindt) x = mindt;
            if (x > maxdt) x = maxdt;
            nd[e] = x;
        }
        sort3(nd);
        *dtio = nd[0];
        used = dt;
        if (!bad || nd[0] == mindt) { *t = t0 + dt; break; }
        if (tr > maxTry) { rc = -1; break; }
        tr++;
        a++;
    }
    s->p.T = Tsave;
    s->p.TMul = TMsave;
    free(buf);
    if (stats) {
        stats[0] = tErr[0]; stats[1] = tErr[1]; stats[2] = tErr[2];
        stats[3] = tr + 1; stats[4] = (double)iterations; stats[5] = used;
    }
    return rc;
}

/* LoadEnergy, modules/mod_energy.go:25-61: the terms of the fields this
 * solver loads, each weight * Ms0T0Mul * sum (m . F) msat0T0
 * (EnergyUpdater.Update, :98-116): E[0] Zeeman (uniform B_ext, weight -V),
 * E[1] exchange (H_ex, -V mu0 / 2), E[2] demag (B, -V / 2), E[3]
 * longitudinal field (H_lf, -V mu0), E[4] their sum.  The fields are those
 * of the last or_llb_rhs call; m is the (normalised) m that call used. */
void or_llb_energy(const or_llb* s, const double* m, double E[5]) {
    const or_llb_params* p = &s->p;
    const long N = s->N;
    const double V = p->cell[0] * p->cell[1] * p->cell[2];
    const double one[3] = {1.0, 1.0, 1.0};
    for (int t = 0; t < 5; t++) E[t] = 0.0;
    if (p->zeeman) {
        double sum = 0.0;
        for (long i = 0; i < N; i++) {
            double w = m[i] * p->Bext[0] + m[N + i] * p->Bext[1] + m[2 * N + i] * p->Bext[2];
            if (p->msat0T0) w = w * p->msat0T0[i];
            sum += w;
        }
        E[0] = sum * p->msat0T0Mul * (-V);
    }
    if (p->exchange) E[1] = or_energy_term(m, s->Hex, one, one, p->msat0T0, p->msat0T0Mul, -0.5 * V * OR_MU0, N);
    if (p->demag) E[2] = or_energy_term(m, s->B, one, one, p->msat0T0, p->msat0T0Mul, -0.5 * V, N);
    if (p->longfield) E[3] = or_energy_term(m, s->Hlf, one, one, p->msat0T0, p->msat0T0Mul, -V * OR_MU0, N);
    E[4] = E[0] + E[1] + E[2] + E[3];
}

/* EnergyUpdater.Update, modules/mod_energy.go:98-116 */
double or_energy_term(const double* m, const double* H, const double mMul[3], const double HMul[3],
                      const double* msat, double msatMul, double weight, long N) {
    double s = 0;
    for (long i = 0; i < N; i++) {
        double w = mMul[0] * m[i] * (HMul[0] * H[i]) + mMul[1] * m[N + i] * (HMul[1] * H[N + i]) +
                   mMul[2] * m[2 * N + i] * (HMul[2] * H[2 * N + i]);
        if (msat) w = w * msat[i];
        s += w;
    }
    return s * msatMul * weight;
}
