// solver.hip -- the resident LLB-r solver: FP32 fields, FP64 state.
//
// One RK stage = [demag passes on the FP32 stage input m_s -> B] + one fused
// stage kernel.  The fused kernel marches a 64(k) x 4(j) tile through all
// planes i, holding planes i-1, i, i+1 of m_s in registers and the plane's
// (j,k) halo in LDS (6-neighbour exchange of exchange6.cu:11-127), adds
// B/mu0, B_ext/mu0, the longitudinal field (long_field.cu:13-72) and the
// Philox thermal field (temperature.cu:14-71 scaling, mod_brown.go:125-145
// prefactor), forms k = gamma (m x H + alpha .* H) (llbar-torque.cu:39,
// llbar-local00nc.cu:37-56, summed as update_sum.go:41-62 does) and applies
// the RK stage update to the FP64 state in the same pass: accumulator,
// next stage input N(y0 + a h k) in FP32, error norm max|e| in FP64.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <limits>
#include <vector>

#include "hs_common.hpp"
#include "philox.hpp"
#include "xcd.hpp"

int hs_demag_field_f32_internal(hs_demag* d, const float* m, const float* msk, double msat0T0Mul, float* B,
                                hipStream_t st, int nsolve, const int* live);
extern "C" int hs_demag_reserve(hs_demag* d, int nsolve);

namespace {

constexpr int TX = 64, TY = 4;
constexpr double kPiD = 3.14159265358979323846;

// Per-solve parameters of a batch (solves share the mesh and materials).
struct SolveP {
    float TMul;     // temperature multiplier
    float bext[3];  // B_ext / mu0
    double thC;     // 2 kB TMul / (mu0 V gamma Ms0T0Mul); thermal K = thC / max(cutoff_dt, dt)
    double lf_b;    // uniform longitudinal field: x = |m| * lf_b = 3 preS Tc / T |m|
    double lf_c1;   // and H_lf = lf_c1 (B(x) - |m|) / (msk dB/dx |m|) m, lf_c1 = n kB T / (mu0 Ms0T0)
    double t_origin;  // two-temperature solves: laser time at the start of the current run
    uint64_t seed;
    uint32_t stream;
    float sthC;  // sqrt(thC): H_therm amplitude = sqrt(T / msk) sqrt(thC) / sqrt(max(cutoff_dt, dt))
    double bextT[3];  // B_ext in T (FP64, for the Zeeman energy)
};

// Material constants of the longitudinal field when J, n, Tc and T are
// uniform (no masks): B_S(x) and dB_S/dx of common_func.h:68-136 with the
// per-spin constants folded in, coth by one exp, and dB/dx written as
// l^2/sinh^2(l x) - g^2/sinh^2(g x) (no O(1) terms cancelling).
struct LfConst {
    int uniform;   // fast path usable
    int langevin;  // S > 1e5
    int gpow;      // 2S+1 when it is an integer <= 16: exp(-2gx) = exp(-2lx)^(2S+1); else 0
    double l, g, lim;
    double a1, a3, a5, a7;  // series of B: x (a1 - x^2 (a3 - x^2 (a5 - x^2 a7)))
    double d0, d2, d4, d6;  // series of dB/dx
};

// Per-solve clock and slot record (hs_slot): read by the stage kernels when
// the step, attempt and dt live on the device (graph replay, slots).
using Clock = hs_slot;
constexpr int kMaxTry = 10;  // solve_rk12.go:44 (maxTry), retries of one step
// The error max of a solve is gathered in kErrLanes partial maxima, one per
// 128-byte line (stride kErrStride), so the per-block atomics of a launch
// spread over many L2 lines instead of serialising on one address.
constexpr int kErrLanes = 32, kErrStride = 16;
constexpr int kErrPerEq = kErrLanes * kErrStride;
// equations with their own error max: m, and for two-temperature solves Te, Tp
constexpr int kNeq = 3;
constexpr int kErrPerSolve = kNeq * kErrPerEq;

struct Phys {
    int N0, N1, N2, w0, w1, w2;
    int chunk;  // planes marched by one block (grid.z = nsolve * ceil(N0 / chunk))
    int nsolve;
    const SolveP* sp;
    float cutoff_dt;
    int exchange, demag, zeeman, longfield, thermal, torque, local, normalize;
    float ex0, ex1, ex2;  // lexMul^2 / c_d^2
    float ms0T0Mul;
    const float* lex;
    const float* msk;  // msat0T0 mask
    float invMu0;
    float JMul, nMul, TcMul;
    const float *J, *n, *Tc, *T;
    float gamma;
    float aMul[3];
    float sqa[3];  // sqrt(aMul[c]): the thermal amplitude of a component without an alpha mask
    const float* alpha[3];
    // The uniform longitudinal field's constants sit in device memory and are
    // loaded inside the branch that uses them: as kernel arguments their 11
    // doubles stay live in scalar registers through the whole stage kernel,
    // which then spills scalar registers into vector lanes.
    int lf_uniform;
    const LfConst* lfc;
    double sh_w, sh_q;  // spin heating: 3 kB TcMul nMul gammaLL and mu0 Ms0T0Mul gammaLL
};

// Roles of the fused stage kernel.
enum Role {
    R_RHS = 0,      // k and H_eff out (testing)
    R_RK4_1, R_RK4_2, R_RK4_3, R_RK4_4,
    R_BS_1, R_BS_2, R_BS_3, R_BS_4,  // Bogacki-Shampine RK23
    R_RK12_1, R_RK12_2,               // Heun-Euler (reference RK12 / Heun)
    R_EULER
};

struct StageIO {
    const float* m;   // stage input (3 planes)
    const float* B;   // demag field (3 planes) or null
    double* y0;       // FP64 state at step start
    double* acc;      // FP64 accumulator
    float* e;         // FP32 error partial (RK23) / stored k1 (RK12)
    float *k1, *k2, *k3;  // stored stage slopes (RK4, RK23)
    float* mout;      // next stage input
    float* kout;      // R_RHS: k
    float* hout;      // R_RHS: H_eff
    unsigned long long* errmax;  // max |e| as double bits: kErrLanes lanes per solve
    const float* T;   // two-temperature solves: Te of the stage input (FP32, [solve][N]); overrides the T mask
    float* qs;        // two-temperature solves with spin heating: the spin term of Qe per cell (out)
    double h;
    uint32_t step, attempt;
    const Clock* clk;  // when set, step/attempt/h of each solve are read from here
};

// Offset the buffers of `io` to solve `sol` (layout [solve][3][N]).
// |e| for the error maxima, with NaN mapped to +inf: fmax drops a NaN
// operand, so without this a diverged state would report a finite error.
__device__ __forceinline__ double err_abs(double e) {
    const double a = fabs(e);
    return a == a ? a : __builtin_huge_val();
}

__device__ __forceinline__ StageIO solve_io(const StageIO& io, int sol, long N) {
    StageIO r = io;
    const long b = (long)sol * 3 * N;
    r.m += b;
    if (r.B) r.B += b;
    if (r.y0) r.y0 += b;
    if (r.acc) r.acc += b;
    if (r.e) r.e += b;
    if (r.k1) r.k1 += b;
    if (r.k2) r.k2 += b;
    if (r.k3) r.k3 += b;
    if (r.qs) r.qs += (long)sol * N;
    if (r.mout) r.mout += b;
    if (r.kout) r.kout += b;
    if (r.hout) r.hout += b;
    if (r.T) r.T += (long)sol * N;
    r.errmax += (long)sol * kErrPerSolve;
    if (io.clk) {
        r.step = io.clk[sol].step;
        r.attempt = io.clk[sol].attempt;
        r.h = io.clk[sol].h;
    }
    return r;
}

// Blocks of a solve that is not running (idle / done / failed slot) exit.
__device__ __forceinline__ bool solve_idle(const StageIO& io, int sol) {
    return io.clk && io.clk[sol].state != HS_SLOT_RUNNING;
}

// Lean form of the stage kernels (LEAN = true): no per-cell masks, no
// two-temperature coupling, and the longitudinal field off or uniform.  The
// kernel overrides those fields of its arguments with compile-time nulls, so
// the mask loads and tests and the general longitudinal-field path (FP64
// Brillouin with per-cell J, n, Tc, T) drop out of the code: 74 -> 46 VGPRs
// and no scalar-register spills.  The arithmetic of the remaining terms is
// the same, so both forms give the same result.
template <bool LEAN>
__device__ __forceinline__ Phys lean_phys(const Phys& P) {
    Phys Q = P;
    if constexpr (LEAN) {
        Q.lex = Q.msk = Q.J = Q.n = Q.Tc = Q.T = nullptr;
        Q.alpha[0] = Q.alpha[1] = Q.alpha[2] = nullptr;
        Q.lf_uniform = 1;
    }
    return Q;
}
template <bool LEAN>
__device__ __forceinline__ StageIO lean_io(const StageIO& io) {
    StageIO r = io;
    if constexpr (LEAN) {
        r.T = nullptr;
        r.qs = nullptr;
    }
    return r;
}


__device__ __forceinline__ int wrapc(int c, int n, int w) {
    if (c < 0) c = w ? c + n : 0;
    if (c >= n) c = w ? c - n : n - 1;
    return c;
}

// Load the (j,k) tile (bx, by) of plane i with its one-cell halo into LDS
// (the halo rows/columns are wrapped or clamped as the stencil needs).
__device__ __forceinline__ void load_tile(const Phys& P, const float* m, long N, int i, int bx, int by, int jc, int kc,
                                          int tx, int ty, const float own[3], float lown,
                                          float (*tile)[TY + 2][TX + 2], float (*ltile)[TX + 2]) {
    const int N1 = P.N1, N2 = P.N2;
    const long N12 = (long)N1 * N2;
    for (int c = 0; c < 3; c++) tile[c][ty + 1][tx + 1] = own[c];
    if (P.lex) ltile[ty + 1][tx + 1] = lown;
    if (!P.exchange) return;
    const long base = i * N12;
    if (ty == 0 || ty == TY - 1) {
        int jn = (ty == 0) ? wrapc(by * TY - 1, N1, P.w1) : wrapc(min(by * TY + TY, N1), N1, P.w1);
        int row = (ty == 0) ? 0 : TY + 1;
        long a2 = base + (long)jn * N2 + kc;
        for (int c = 0; c < 3; c++) tile[c][row][tx + 1] = m[c * N + a2];
        if (P.lex) ltile[row][tx + 1] = P.lex[a2];
    }
    if (tx == 0 || tx == TX - 1) {
        int kn = (tx == 0) ? wrapc(bx * TX - 1, N2, P.w2) : wrapc(min(bx * TX + TX, N2), N2, P.w2);
        int col = (tx == 0) ? 0 : TX + 1;
        long a2 = base + (long)jc * N2 + kn;
        for (int c = 0; c < 3; c++) tile[c][ty + 1][col] = m[c * N + a2];
        if (P.lex) ltile[ty + 1][col] = P.lex[a2];
    }
}

// The same tile in two halves, so that a kernel can issue the halo loads
// together with its other loads and write the LDS tile after one wait:
// load_halo reads the halo values of an edge lane into registers, store_tile
// writes the tile.
struct Halo {
    float y[3], x[3], ly, lx;
};
__device__ __forceinline__ void load_halo(const Phys& P, const float* m, long N, int i, int bx, int by, int jc, int kc,
                                          int tx, int ty, Halo& h) {
    const int N1 = P.N1, N2 = P.N2;
    const long N12 = (long)N1 * N2;
    const long base = i * N12;
    if (!P.exchange) return;
    if (ty == 0 || ty == TY - 1) {
        int jn = (ty == 0) ? wrapc(by * TY - 1, N1, P.w1) : wrapc(min(by * TY + TY, N1), N1, P.w1);
        long a2 = base + (long)jn * N2 + kc;
        for (int c = 0; c < 3; c++) h.y[c] = m[c * N + a2];
        if (P.lex) h.ly = P.lex[a2];
    }
    if (tx == 0 || tx == TX - 1) {
        int kn = (tx == 0) ? wrapc(bx * TX - 1, N2, P.w2) : wrapc(min(bx * TX + TX, N2), N2, P.w2);
        long a2 = base + (long)jc * N2 + kn;
        for (int c = 0; c < 3; c++) h.x[c] = m[c * N + a2];
        if (P.lex) h.lx = P.lex[a2];
    }
}
__device__ __forceinline__ void store_tile(const Phys& P, int tx, int ty, const float own[3], float lown,
                                           const Halo& h, float (*tile)[TY + 2][TX + 2], float (*ltile)[TX + 2]) {
    for (int c = 0; c < 3; c++) tile[c][ty + 1][tx + 1] = own[c];
    if (P.lex) ltile[ty + 1][tx + 1] = lown;
    if (!P.exchange) return;
    if (ty == 0 || ty == TY - 1) {
        int row = (ty == 0) ? 0 : TY + 1;
        for (int c = 0; c < 3; c++) tile[c][row][tx + 1] = h.y[c];
        if (P.lex) ltile[row][tx + 1] = h.ly;
    }
    if (tx == 0 || tx == TX - 1) {
        int col = (tx == 0) ? 0 : TX + 1;
        for (int c = 0; c < 3; c++) tile[c][ty + 1][col] = h.x[c];
        if (P.lex) ltile[ty + 1][col] = h.lx;
    }
}

// 6-neighbour exchange field of cell (i,j,k), exchange6.cu:11-127 in FP32
// with the differences formed first: pre_d (pre1 (m1-m0) + pre2 (m2-m0)).
// lo/hi: planes i-1, i+1 (registers); the j,k neighbours come from the tile
// (row ty+1, column cc holds the cell itself; halo rows/columns hold the
// neighbours of the tile's edge cells).  At a periodic domain edge the
// wrapped neighbour is read from memory; at an open edge it is the cell itself.
template <int TW>
__device__ __forceinline__ void exchange_cell(const Phys& P, const float* m, long N, int i, int j, int k, int jc,
                                              int cc, int ty, const float m0[3], const float lo[3],
                                              const float hi[3], float l0, float llo, float lhi,
                                              float (*tile)[TY + 2][TW + 2], float (*ltile)[TW + 2], float H[3]) {
    const int N0 = P.N0, N1 = P.N1, N2 = P.N2;
    const long N12 = (long)N1 * N2;
    const float l02 = l0 * l0;
    float e0[3], e1[3], e2[3];
    // harmonic-mean weights of the two neighbours; exactly 1 without a lex mask
    // (a uniform branch: no divisions on the common path)
    auto weights = [&](float l1, float l2, float& pre1, float& pre2) {
        if (P.lex) {
            pre1 = avg_geom_zero(l02, l1 * l1);
            pre2 = avg_geom_zero(l02, l2 * l2);
        } else {
            pre1 = 1.f;
            pre2 = 1.f;
        }
    };
    {  // axis 0 (X): planes i-1, i+1; clamped at open ends = self
        bool lo_self = (i == 0 && !P.w0), hi_self = (i == N0 - 1 && !P.w0);
        float l1 = lo_self ? l0 : llo, l2 = hi_self ? l0 : lhi;
        float pre1, pre2;
        weights(l1, l2, pre1, pre2);
        for (int c = 0; c < 3; c++) {
            float v1 = lo_self ? m0[c] : lo[c], v2 = hi_self ? m0[c] : hi[c];
            e0[c] = P.ex0 * (pre1 * (v1 - m0[c]) + pre2 * (v2 - m0[c]));
        }
    }
    {  // axis 2 (Z, contiguous k)
        float v1[3], v2[3], l1, l2;
        if (k - 1 < 0 && P.w2) {
            long aa = i * N12 + (long)jc * N2 + (N2 - 1);
            for (int c = 0; c < 3; c++) v1[c] = m[c * N + aa];
            l1 = P.lex ? P.lex[aa] : 1.f;
        } else {
            int col = (k - 1 < 0) ? cc : cc - 1;
            for (int c = 0; c < 3; c++) v1[c] = tile[c][ty + 1][col];
            l1 = P.lex ? ltile[ty + 1][col] : 1.f;
        }
        if (k + 1 >= N2 && P.w2) {
            long aa = i * N12 + (long)jc * N2;
            for (int c = 0; c < 3; c++) v2[c] = m[c * N + aa];
            l2 = P.lex ? P.lex[aa] : 1.f;
        } else {
            int col = (k + 1 >= N2) ? cc : cc + 1;
            for (int c = 0; c < 3; c++) v2[c] = tile[c][ty + 1][col];
            l2 = P.lex ? ltile[ty + 1][col] : 1.f;
        }
        float pre1, pre2;
        weights(l1, l2, pre1, pre2);
        for (int c = 0; c < 3; c++) e2[c] = P.ex2 * (pre1 * (v1[c] - m0[c]) + pre2 * (v2[c] - m0[c]));
    }
    {  // axis 1 (Y)
        float v1[3], v2[3], l1, l2;
        if (j - 1 < 0 && P.w1) {
            long aa = i * N12 + (long)(N1 - 1) * N2 + k;
            for (int c = 0; c < 3; c++) v1[c] = m[c * N + aa];
            l1 = P.lex ? P.lex[aa] : 1.f;
        } else {
            int row = (j - 1 < 0) ? ty + 1 : ty;
            for (int c = 0; c < 3; c++) v1[c] = tile[c][row][cc];
            l1 = P.lex ? ltile[row][cc] : 1.f;
        }
        if (j + 1 >= N1 && P.w1) {
            long aa = i * N12 + k;
            for (int c = 0; c < 3; c++) v2[c] = m[c * N + aa];
            l2 = P.lex ? P.lex[aa] : 1.f;
        } else {
            int row = (j + 1 >= N1) ? ty + 1 : ty + 2;
            for (int c = 0; c < 3; c++) v2[c] = tile[c][row][cc];
            l2 = P.lex ? ltile[row][cc] : 1.f;
        }
        float pre1, pre2;
        weights(l1, l2, pre1, pre2);
        for (int c = 0; c < 3; c++) e1[c] = P.ex1 * (pre1 * (v1[c] - m0[c]) + pre2 * (v2[c] - m0[c]));
    }
    for (int c = 0; c < 3; c++) H[c] = P.ms0T0Mul * (e0[c] + e2[c] + e1[c]);
}

// Operands of one cell's stage update, loaded at the top of the kernel so a
// wave has all its memory requests in flight at once.
struct Pre {
    float ms, B[3], J, n, Tc, T, al[3];
    double y[3];    // y0 (or acc for RK12_2 / BS_4 with the normaliser)
    float ka[3];    // k1 / e
    float kb[3];    // k2
    float kc[3];    // k3
};

template <int ROLE>
__device__ __forceinline__ void preload(const Phys& P, const StageIO& io, long I, long N, Pre& q) {
    q.ms = P.msk ? P.msk[I] : 1.f;
    if (P.demag)
        for (int c = 0; c < 3; c++) q.B[c] = io.B[c * N + I];
    if (P.longfield || io.qs) {
        q.J = P.J ? P.J[I] : 1.f;
        q.n = P.n ? P.n[I] : 1.f;
        q.Tc = P.Tc ? P.Tc[I] : 1.f;
    }
    q.T = io.T ? io.T[I] : (P.T ? P.T[I] : 1.f);
    for (int c = 0; c < 3; c++) q.al[c] = P.alpha[c] ? P.alpha[c][I] : 1.f;
    if constexpr (ROLE == R_RK4_1 || ROLE == R_BS_1 || ROLE == R_RK12_1 || ROLE == R_EULER || ROLE == R_RK4_2 ||
                  ROLE == R_RK4_3 || ROLE == R_BS_2 || ROLE == R_RK4_4 || ROLE == R_BS_3)
        for (int c = 0; c < 3; c++) q.y[c] = io.y0[c * N + I];
    if constexpr (ROLE == R_RK12_2 || ROLE == R_BS_4) {
        if (ROLE == R_RK12_2 || P.normalize)
            for (int c = 0; c < 3; c++) q.y[c] = io.acc[c * N + I];
    }
    if constexpr (ROLE == R_RK4_4 || ROLE == R_BS_3)
        for (int c = 0; c < 3; c++) {
            q.ka[c] = io.k1[c * N + I];
            q.kb[c] = io.k2[c * N + I];
        }
    if constexpr (ROLE == R_RK4_4)
        for (int c = 0; c < 3; c++) q.kc[c] = io.k3[c * N + I];
    if constexpr (ROLE == R_BS_4 || ROLE == R_RK12_2)
        for (int c = 0; c < 3; c++) q.ka[c] = io.e[c * N + I];
}

// Everything after the exchange term for one cell: B/mu0, B_ext/mu0, H_lf,
// H_therm and k = gamma (m x H + alpha H).
__device__ __forceinline__ void field_torque(const Phys& P, const SolveP* sp, const StageIO& io, long I,
                                             const float m0[3], float H[3], const Pre& q, float kk[3]) {
    const float ms = q.ms;
    if (P.demag) {
        H[0] += P.invMu0 * q.B[0];
        H[1] += P.invMu0 * q.B[1];
        H[2] += P.invMu0 * q.B[2];
    }
    if (P.zeeman) {
        H[0] += sp->bext[0];
        H[1] += sp->bext[1];
        H[2] += sp->bext[2];
    }
    if (P.longfield && P.lf_uniform) {
        const LfConst C = *P.lfc;
        double mx = m0[0], my = m0[1], mz = m0[2];
        double am2 = mx * mx + my * my + mz * mz;
        if (!(ms == 0.f || am2 <= HS_ZERO * HS_ZERO)) {
            // |m| and 1/|m|: hardware rsqrt refined by one FP64 Newton step (rel. error ~1e-14)
            double ia, am;
            if (am2 > 1e-30) {
                ia = (double)__builtin_amdgcn_rsqf((float)am2);
                ia = ia * (1.5 - 0.5 * am2 * ia * ia);
                am = am2 * ia;
            } else {
                am = sqrt(am2);
                ia = 1.0 / am;
            }
            double x = am * sp->lf_b;
            // H_lf = c1 (W - |m|) / (ms dW) m/|m| as num / den: one FP64 division
            double num, den;
            if (x < C.lim) {
                double x2 = x * x;
                double W = x * (C.a1 - x2 * (C.a3 - x2 * (C.a5 - x2 * C.a7)));
                num = W - am;
                den = C.d0 - x2 * (C.d2 - x2 * (C.d4 - x2 * C.d6));
            } else if (!C.langevin) {
                // coth(gx), coth(lx) through e^{-2gx}, e^{-2lx}; g = (2S+1) l, so for
                // integer 2S+1 the first is a power of the second (one exp).  With
                // a = 1 - e^{-2gx}, b = 1 - e^{-2lx}:  W = Wn / (ab),  dW = dWn / (ab)^2,
                // Wn = g (1 + e^{-2gx}) b - l (1 + e^{-2lx}) a,
                // dWn = 4 l^2 e^{-2lx} a^2 - 4 g^2 e^{-2gx} b^2,
                // so (W - |m|) / dW = (Wn - |m| ab) ab / dWn.
                double el = exp(-2.0 * C.l * x), eg;
                if (C.gpow) {
                    eg = el;
                    for (int q2 = 1; q2 < C.gpow; q2++) eg *= el;
                } else {
                    eg = exp(-2.0 * C.g * x);
                }
                const double a = 1.0 - eg, b = 1.0 - el, D = a * b;
                num = (C.g * (1.0 + eg) * b - C.l * (1.0 + el) * a - am * D) * D;
                den = 4.0 * (C.l * C.l) * el * (a * a) - 4.0 * (C.g * C.g) * eg * (b * b);
            } else {
                // L(x) = coth x - 1/x, L' = 1/x^2 - 1/sinh^2 x; with a = 1 - e^{-2x}:
                // (L - |m|) / L' = ((1 + e^{-2x}) x - a - |m| a x) a x / (a^2 - 4 e^{-2x} x^2)
                const double e = exp(-2.0 * x), a = 1.0 - e;
                num = ((1.0 + e) * x - a - am * a * x) * (a * x);
                den = a * a - 4.0 * e * (x * x);
            }
            double mult = sp->lf_c1 * num * ia / ((double)ms * den);
            H[0] += (float)(mult * mx);
            H[1] += (float)(mult * my);
            H[2] += (float)(mult * mz);
        }
    } else if (P.longfield) {
        // FP64: dB_S/dx is a difference of O(1) terms (common_func.h:95-98)
        double Ms0T0 = (double)P.ms0T0Mul * ms;
        double S = (double)P.JMul * q.J;
        double n = (double)P.nMul * q.n;
        double Tc = (double)P.TcMul * q.Tc;
        double Ts = (double)sp->TMul * q.T;
        double mx = m0[0], my = m0[1], mz = m0[2];
        double am = sqrt(mx * mx + my * my + mz * mz);
        if (!(Ms0T0 == 0.0 || n == 0.0 || am <= HS_ZERO)) {
            double preS = (S > HS_INFSPIN) ? 1.0 : S / (S + 1.0);
            double b = 3.0 * preS * Tc / Ts;
            double W, dW;
            brillouin_w_dw<double>(S, am * b, W, dW);
            double mult = n * (HS_KB / HS_MU0) * Ts * (W - am) / (Ms0T0 * dW) / am;
            H[0] += (float)(mult * mx);
            H[1] += (float)(mult * my);
            H[2] += (float)(mult * mz);
        }
    }
    if (P.thermal && ms != 0.f) {
        float x0, x1, x2;
        thermal_normals_hw((uint32_t)I, io.step, io.attempt, sp->stream, sp->seed, x0, x1, x2);
        const float rdt = __builtin_amdgcn_rsqf((float)fmax((double)P.cutoff_dt, io.h));
        float pre = __builtin_amdgcn_sqrtf(q.T * __builtin_amdgcn_rcpf(ms)) * (sp->sthC * rdt);
        float s0 = P.alpha[0] ? __builtin_amdgcn_sqrtf(P.aMul[0] * q.al[0]) : P.sqa[0];
        float s1 = P.alpha[1] ? __builtin_amdgcn_sqrtf(P.aMul[1] * q.al[1]) : P.sqa[1];
        float s2 = P.alpha[2] ? __builtin_amdgcn_sqrtf(P.aMul[2] * q.al[2]) : P.sqa[2];
        H[0] += pre * (s0 * x0);
        H[1] += pre * (s1 * x1);
        H[2] += pre * (s2 * x2);
    }
    if (io.qs) {  // spin heating of a two-temperature solve: ws - qs with R = diag(alpha) H_eff
        double qv = 0.0;
        if (ms != 0.f && P.local) {  // R is the local damping term: zero without it
            float R[3];
            for (int c = 0; c < 3; c++) R[c] = (P.aMul[c] * q.al[c]) * H[c];
            const double mR = (double)m0[0] * R[0] + (double)m0[1] * R[1] + (double)m0[2] * R[2];
            const double HR = (double)H[0] * R[0] + (double)H[1] * R[1] + (double)H[2] * R[2];
            const double S = (double)P.JMul * q.J, Tc = q.Tc, n = q.n;
            const double ws =
                (n == 0.0 || S == 0.0 || Tc == 0.0) ? 0.0 : P.sh_w * (Tc * S * S * n / (S * (S + 1.0))) * mR;
            qv = ws + P.sh_q * HR * (double)ms;  // ws - qs, qs = -mu0 Ms0T0 gamma (H . R) msk
        }
        io.qs[I] = (float)qv;
    }
    kk[0] = kk[1] = kk[2] = 0.f;
    if (ms != 0.f) {
        if (P.torque) {  // cross(H, m) = m x H
            kk[0] = P.gamma * (-H[1] * m0[2] + H[2] * m0[1]);
            kk[1] = P.gamma * (-H[2] * m0[0] + H[0] * m0[2]);
            kk[2] = P.gamma * (-H[0] * m0[1] + H[1] * m0[0]);
        }
        if (P.local) {
            for (int c = 0; c < 3; c++) {
                float al = P.aMul[c] * q.al[c];
                kk[c] += P.gamma * (al * H[c]);
            }
        }
    }
}

// What a stage writes for one cell (which fields are used depends on ROLE):
// k -> kout / k1 / k2 / k3 / e (RK12 k1); m -> next stage input; h -> hout
// (RHS) / e (BS_3); y -> y0 or acc.
struct Out {
    float k[3], m[3], h[3];
    double y[3];
};

// The RK stage update of one cell (FP64 state) into `o`.
template <int ROLE>
__device__ __forceinline__ void stage_update(const Phys& P, const StageIO& io, const Pre& q, const float kk[3],
                                             const float H[3], Out& o, double& errloc) {
    const double h = io.h;
    auto nrm = [&](double* v) {
        if (P.normalize) {
            double l = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            double sc = (l == 0.0) ? 0.0 : 1.0 / l;
            v[0] *= sc;
            v[1] *= sc;
            v[2] *= sc;
        }
    };
    auto set_m = [&](const double* v) {
        double w[3] = {v[0], v[1], v[2]};
        nrm(w);
        for (int c = 0; c < 3; c++) o.m[c] = (float)w[c];
    };
    double y[3], acc[3];
    if constexpr (ROLE == R_RHS) {
        for (int c = 0; c < 3; c++) {
            o.k[c] = kk[c];
            o.h[c] = H[c];
        }
        return;
    }
    if constexpr (ROLE == R_RK4_1 || ROLE == R_BS_1 || ROLE == R_RK12_1 || ROLE == R_EULER) {
        for (int c = 0; c < 3; c++) y[c] = q.y[c];
        nrm(y);
        for (int c = 0; c < 3; c++) o.y[c] = y[c];  // y0 written back only with the normaliser
    }
    if constexpr (ROLE == R_RK4_1 || ROLE == R_BS_1) {
        for (int c = 0; c < 3; c++) {
            o.k[c] = kk[c];
            acc[c] = y[c] + (0.5 * h) * (double)kk[c];
        }
        set_m(acc);
    } else if constexpr (ROLE == R_RK4_2 || ROLE == R_RK4_3 || ROLE == R_BS_2) {
        const double cm = (ROLE == R_RK4_3) ? h : ((ROLE == R_BS_2) ? 0.75 * h : 0.5 * h);
        for (int c = 0; c < 3; c++) {
            o.k[c] = kk[c];
            acc[c] = q.y[c] + cm * (double)kk[c];
        }
        set_m(acc);
    } else if constexpr (ROLE == R_RK4_4) {
        const double h6 = h / 6.0;
        for (int c = 0; c < 3; c++) {
            double k1 = q.ka[c], k2 = q.kb[c], k3 = q.kc[c];
            acc[c] = q.y[c] + h6 * (k1 + 2.0 * k2 + 2.0 * k3 + (double)kk[c]);
            o.y[c] = acc[c];
        }
        set_m(acc);
    } else if constexpr (ROLE == R_BS_3) {
        for (int c = 0; c < 3; c++) {
            double k1 = q.ka[c], k2 = q.kb[c], k3 = kk[c];
            acc[c] = q.y[c] + h * ((2.0 / 9.0) * k1 + (1.0 / 3.0) * k2 + (4.0 / 9.0) * k3);
            o.y[c] = acc[c];
            o.h[c] = (float)(h * ((-5.0 / 72.0) * k1 + (1.0 / 12.0) * k2 + (1.0 / 9.0) * k3));
        }
        set_m(acc);
    } else if constexpr (ROLE == R_BS_4) {
        for (int c = 0; c < 3; c++) {
            double ev = (double)q.ka[c] - (h / 8.0) * (double)kk[c];
            errloc = fmax(errloc, err_abs(ev));
        }
        for (int c = 0; c < 3; c++) acc[c] = q.y[c];
        nrm(acc);
        for (int c = 0; c < 3; c++) o.y[c] = acc[c];  // written only with the normaliser
    } else if constexpr (ROLE == R_RK12_1) {
        for (int c = 0; c < 3; c++) {
            o.k[c] = kk[c];
            acc[c] = y[c] + h * (double)kk[c];
            o.y[c] = acc[c];
        }
        set_m(acc);
    } else if constexpr (ROLE == R_RK12_2) {
        for (int c = 0; c < 3; c++) acc[c] = q.y[c];
        nrm(acc);
        for (int c = 0; c < 3; c++) {
            float k1 = q.ka[c];
            acc[c] += (0.5 * h) * (double)kk[c] + (-0.5 * h) * (double)k1;
            o.y[c] = acc[c];
            errloc = fmax(errloc, err_abs((double)kk[c] - (double)k1));
        }
        set_m(acc);
    } else if constexpr (ROLE == R_EULER) {
        for (int c = 0; c < 3; c++) {
            acc[c] = y[c] + h * (double)kk[c];
            o.y[c] = acc[c];
        }
        set_m(acc);
    }
}

// Destinations of a role's outputs.
template <int ROLE>
__device__ __forceinline__ float* k_dst(const StageIO& io) {
    if constexpr (ROLE == R_RHS) return io.kout;
    else if constexpr (ROLE == R_RK4_1 || ROLE == R_BS_1) return io.k1;
    else if constexpr (ROLE == R_RK4_2 || ROLE == R_BS_2) return io.k2;
    else if constexpr (ROLE == R_RK4_3) return io.k3;
    else if constexpr (ROLE == R_RK12_1) return io.e;
    else return nullptr;
}
template <int ROLE>
__device__ __forceinline__ float* h_dst(const StageIO& io) {
    if constexpr (ROLE == R_RHS) return io.hout;
    else if constexpr (ROLE == R_BS_3) return io.e;
    else return nullptr;
}
template <int ROLE>
__device__ __forceinline__ double* y_dst(const Phys& P, const StageIO& io) {
    if constexpr (ROLE == R_RK4_4 || ROLE == R_EULER) return io.y0;
    else if constexpr (ROLE == R_RK4_1 || ROLE == R_BS_1) return P.normalize ? io.y0 : nullptr;
    else if constexpr (ROLE == R_RK12_1) return io.acc;  // the normalised y0 is written by store_out
    else if constexpr (ROLE == R_BS_3 || ROLE == R_RK12_2) return io.acc;
    else if constexpr (ROLE == R_BS_4) return P.normalize ? io.acc : nullptr;
    else return nullptr;
}
template <int ROLE>
__device__ __forceinline__ bool has_m(const StageIO& io) {
    return !(ROLE == R_RHS || ROLE == R_BS_4) && io.mout;
}

// Stored slopes and error partials of an RK stage are read back only one or
// two demag evaluations later: non-temporal stores keep them from displacing
// the lines the passes in between reuse (R_RHS outputs go to the host).
template <int ROLE>
__device__ __forceinline__ void st_slope(float* a, float v) {
    if constexpr (ROLE == R_RHS) *a = v;
    else __builtin_nontemporal_store(v, a);
}
template <int ROLE>
__device__ __forceinline__ void st_slope2(float* a, float v0, float v1) {
    typedef float f2 __attribute__((ext_vector_type(2)));
    if constexpr (ROLE == R_RHS) {
        *(float2*)a = make_float2(v0, v1);
    } else {
        f2 q;
        q.x = v0;
        q.y = v1;
        __builtin_nontemporal_store(q, (f2*)a);
    }
}

// Scalar stores of one cell's outputs.
template <int ROLE>
__device__ __forceinline__ void store_out(const Phys& P, const StageIO& io, long I, long N, const Out& o,
                                          const Pre& q) {
    if (float* d = k_dst<ROLE>(io))
        for (int c = 0; c < 3; c++) st_slope<ROLE>(d + c * N + I, o.k[c]);
    if (float* d = h_dst<ROLE>(io))
        for (int c = 0; c < 3; c++) st_slope<ROLE>(d + c * N + I, o.h[c]);
    if (has_m<ROLE>(io))
        for (int c = 0; c < 3; c++) io.mout[c * N + I] = o.m[c];
    if (double* d = y_dst<ROLE>(P, io)) {
        if constexpr (ROLE == R_BS_3) {  // y_{n+1}: read back at the next step, after another demag
            for (int c = 0; c < 3; c++) __builtin_nontemporal_store(o.y[c], d + c * N + I);
        } else {
            for (int c = 0; c < 3; c++) d[c * N + I] = o.y[c];
        }
    }
    if constexpr (ROLE == R_RK12_1 || ROLE == R_EULER) {  // lazily normalised y0 written back
        if (P.normalize && ROLE == R_RK12_1) {
            double y[3] = {q.y[0], q.y[1], q.y[2]};
            double l = sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
            double sc = (l == 0.0) ? 0.0 : 1.0 / l;
            for (int c = 0; c < 3; c++) io.y0[c * N + I] = y[c] * sc;
        }
    }
}

// Both cells (I, I+1) of a thread at once: 8-byte (FP32) and 16-byte (FP64)
// stores.  I is even and the planes start 8-byte aligned.
template <int ROLE>
__device__ __forceinline__ void store_out2(const Phys& P, const StageIO& io, long I, long N, const Out& a,
                                           const Out& b, const Pre& qa, const Pre& qb) {
    if (float* d = k_dst<ROLE>(io))
        for (int c = 0; c < 3; c++) st_slope2<ROLE>(d + c * N + I, a.k[c], b.k[c]);
    if (float* d = h_dst<ROLE>(io))
        for (int c = 0; c < 3; c++) st_slope2<ROLE>(d + c * N + I, a.h[c], b.h[c]);
    if (has_m<ROLE>(io))
        for (int c = 0; c < 3; c++) *(float2*)(io.mout + c * N + I) = make_float2(a.m[c], b.m[c]);
    if (double* d = y_dst<ROLE>(P, io)) {
        if constexpr (ROLE == R_BS_3) {  // y_{n+1}: read back at the next step, after another demag
            typedef double d2 __attribute__((ext_vector_type(2)));
            for (int c = 0; c < 3; c++) {
                d2 q;
                q.x = a.y[c];
                q.y = b.y[c];
                __builtin_nontemporal_store(q, (d2*)(d + c * N + I));
            }
        } else {
            for (int c = 0; c < 3; c++) *(double2*)(d + c * N + I) = make_double2(a.y[c], b.y[c]);
        }
    }
    if constexpr (ROLE == R_RK12_1) {
        if (P.normalize) {
            double ya[3] = {qa.y[0], qa.y[1], qa.y[2]}, yb[3] = {qb.y[0], qb.y[1], qb.y[2]};
            double la = sqrt(ya[0] * ya[0] + ya[1] * ya[1] + ya[2] * ya[2]);
            double lb = sqrt(yb[0] * yb[0] + yb[1] * yb[1] + yb[2] * yb[2]);
            double sa = (la == 0.0) ? 0.0 : 1.0 / la, sb = (lb == 0.0) ? 0.0 : 1.0 / lb;
            for (int c = 0; c < 3; c++) *(double2*)(io.y0 + c * N + I) = make_double2(ya[c] * sa, yb[c] * sb);
        }
    }
}

template <int ROLE>
__device__ __forceinline__ void finish_cell(const Phys& P, const SolveP* sp, const StageIO& io, long I, long N,
                                            const float m0[3], float H[3], const Pre& q, double& errloc) {
    float kk[3];
    field_torque(P, sp, io, I, m0, H, q, kk);
    Out o;
    stage_update<ROLE>(P, io, q, kk, H, o, errloc);
    store_out<ROLE>(P, io, I, N, o, q);
}

// max of a non-negative double with the value of another lane, picked by a
// DPP control (both 32-bit halves moved, then one FP64 max)
template <int CTRL>
__device__ __forceinline__ double dpp_max(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
    return fmax(v, __longlong_as_double(((long long)hi << 32) | (unsigned)lo));
}
// Max over the 64 lanes of a wave (every lane gets it): within each row of
// 16 lanes by DPP (lane swaps in pairs and quads, then the mirrored half row
// and row), then the four row maxima read into scalar registers.
__device__ __forceinline__ double wave_max(double v) {
    v = dpp_max<0xB1>(v);   // quad_perm [1,0,3,2]
    v = dpp_max<0x4E>(v);   // quad_perm [2,3,0,1]
    v = dpp_max<0x141>(v);  // row_half_mirror: lane i <-> 7-i within 8
    v = dpp_max<0x140>(v);  // row_mirror: lane i <-> 15-i within 16
    const long long b = __double_as_longlong(v);
    double r = 0.0;
#pragma unroll
    for (int row = 0; row < 4; row++) {
        const int lo = __builtin_amdgcn_readlane((int)b, 16 * row);
        const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 16 * row);
        r = fmax(r, __longlong_as_double(((long long)hi << 32) | (unsigned)lo));
    }
    return r;
}

// Error max of the block: a wave max (wave_max), then the block's waves
// through LDS, then one atomic max (ordered double bits) into lane
// blockIdx % kErrLanes.  The error values are >= 0 (err_abs).
template <int ROLE>
__device__ __forceinline__ void flush_err(double errloc, const StageIO& io) {
    if constexpr (ROLE == R_BS_4 || ROLE == R_RK12_2) {
        __shared__ double wmax[TX * TY / 64];
        errloc = wave_max(errloc);
        const int t = threadIdx.y * TX + threadIdx.x;
        if ((t & 63) == 0) wmax[t >> 6] = errloc;
        __syncthreads();
        if (t == 0) {
            double v = wmax[0];
            for (int w = 1; w < TX * TY / 64; w++) v = fmax(v, wmax[w]);
            const unsigned lane = (blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)) % kErrLanes;
            atomicMax(io.errmax + lane * kErrStride, (unsigned long long)__double_as_longlong(v));
        }
    }
}

// Marching variant: a 64x4 tile walks planes [i0, i0+chunk) keeping planes
// i-1, i, i+1 of m in registers (each m value is read from HBM once).
template <int ROLE, bool LEAN>
__global__ void __launch_bounds__(TX* TY) k_stage(Phys P_, StageIO io_) {
    const Phys P = lean_phys<LEAN>(P_);
    const StageIO io0 = lean_io<LEAN>(io_);
    __shared__ float tile[3][TY + 2][TX + 2];
    __shared__ float ltile[TY + 2][TX + 2];
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int k = blockIdx.x * TX + tx, j = blockIdx.y * TY + ty;
    const int N0 = P.N0, N1 = P.N1, N2 = P.N2;
    const long N12 = (long)N1 * N2;
    const long N = (long)N0 * N12;
    const int nchunk = (N0 + P.chunk - 1) / P.chunk;
    const int sol = blockIdx.z / nchunk;
    if (solve_idle(io0, sol)) return;
    const StageIO io = solve_io(io0, sol, N);
    // per-solve constants read where they are used (a copy here would hold
    // them all in scalar registers through the kernel)
    const SolveP* sp = P.sp + sol;
    const bool in = (k < N2) && (j < N1);
    const float* m = io.m;
    const int kc = min(k, N2 - 1), jc = min(j, N1 - 1);
    double errloc = 0.0;
    const int i0 = (blockIdx.z % nchunk) * P.chunk, i1 = min(N0, i0 + P.chunk);
    float a[3][3];
    float la[3] = {1.f, 1.f, 1.f};
    {
        long I0 = (long)jc * N2 + kc;
        int ip = wrapc(i0 - 1, N0, P.w0);
        int in1 = wrapc(i0 + 1, N0, P.w0);
        for (int c = 0; c < 3; c++) {
            a[1][c] = m[c * N + i0 * N12 + I0];
            a[0][c] = m[c * N + ip * N12 + I0];
            a[2][c] = m[c * N + in1 * N12 + I0];
        }
        if (P.lex) {
            la[1] = P.lex[i0 * N12 + I0];
            la[0] = P.lex[ip * N12 + I0];
            la[2] = P.lex[in1 * N12 + I0];
        }
    }
    for (int i = i0; i < i1; i++) {
        const long I = i * N12 + (long)jc * N2 + kc;
        float nxt[3] = {0.f, 0.f, 0.f};
        float lnx = 1.f;
        if (i + 1 < i1) {  // prefetch plane i+2 (wrapped, or clamped at an open end)
            int i2 = i + 2;
            if (i2 >= N0) i2 = P.w0 ? i2 - N0 : N0 - 1;
            long a2 = i2 * N12 + (long)jc * N2 + kc;
            for (int c = 0; c < 3; c++) nxt[c] = m[c * N + a2];
            if (P.lex) lnx = P.lex[a2];
        }
        Pre q;
        preload<ROLE>(P, io, I, N, q);
        load_tile(P, m, N, i, blockIdx.x, blockIdx.y, jc, kc, tx, ty, a[1], la[1], tile, ltile);
        __syncthreads();
        float m0[3] = {a[1][0], a[1][1], a[1][2]};
        float H[3] = {0.f, 0.f, 0.f};
        if (P.exchange && in)
            exchange_cell<TX>(P, m, N, i, j, k, jc, tx + 1, ty, m0, a[0], a[2], la[1], la[0], la[2], tile, ltile, H);
        __syncthreads();
        if (in) finish_cell<ROLE>(P, sp, io, I, N, m0, H, q, errloc);
        for (int c = 0; c < 3; c++) {
            a[0][c] = a[1][c];
            a[1][c] = a[2][c];
            a[2][c] = nxt[c];
        }
        la[0] = la[1];
        la[1] = la[2];
        la[2] = lnx;
    }
    flush_err<ROLE>(errloc, io);
}

// Flat variant: one plane per block, planes i-1 and i+1 read directly (they
// are the interior of other blocks and come from L2 / Infinity Cache).
// ZORD: a 1-D grid whose blocks, renumbered so that consecutive numbers run
// on one XCD (xcd_swizzle), take the planes of a 64x4 tile one after the
// other; the m values of planes i-1 and i+1 that a block reads were loaded
// into that XCD's L2 by the blocks just before and after it.
template <int ROLE, bool ZORD, bool LEAN>
__global__ void __launch_bounds__(TX* TY, 6) k_stage_flat(Phys P_, StageIO io_) {
    const Phys P = lean_phys<LEAN>(P_);
    const StageIO io0 = lean_io<LEAN>(io_);
    __shared__ float tile[3][TY + 2][TX + 2];
    __shared__ float ltile[TY + 2][TX + 2];
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int N0 = P.N0, N1 = P.N1, N2 = P.N2;
    int bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
    if constexpr (ZORD) {
        const int gx = (N2 + TX - 1) / TX, gy = (N1 + TY - 1) / TY;
        const int q = xcd_swizzle(blockIdx.x, gridDim.x);
        const int r = q / N0;
        bx = r % gx;
        by = (r / gx) % gy;
        bz = (r / (gx * gy)) * N0 + q % N0;
    }
    const int k = bx * TX + tx, j = by * TY + ty, i = bz % N0;
    const int sol = bz / N0;
    if (solve_idle(io0, sol)) return;
    const long N12 = (long)N1 * N2;
    const long N = (long)N0 * N12;
    const StageIO io = solve_io(io0, sol, N);
    // per-solve constants read where they are used (a copy here would hold
    // them all in scalar registers through the kernel)
    const SolveP* sp = P.sp + sol;
    const bool in = (k < N2) && (j < N1);
    const float* m = io.m;
    const int kc = min(k, N2 - 1), jc = min(j, N1 - 1);
    const long I = i * N12 + (long)jc * N2 + kc;
    double errloc = 0.0;
    float m0[3], lo[3] = {0.f, 0.f, 0.f}, hi[3] = {0.f, 0.f, 0.f};
    float l0 = 1.f, llo = 1.f, lhi = 1.f;
    for (int c = 0; c < 3; c++) m0[c] = m[c * N + I];
    Pre q;
    preload<ROLE>(P, io, I, N, q);
    if (P.lex) l0 = P.lex[I];
    if (P.exchange) {
        long Ilo = wrapc(i - 1, N0, P.w0) * N12 + (long)jc * N2 + kc;
        long Ihi = wrapc(i + 1, N0, P.w0) * N12 + (long)jc * N2 + kc;
        for (int c = 0; c < 3; c++) {
            lo[c] = m[c * N + Ilo];
            hi[c] = m[c * N + Ihi];
        }
        if (P.lex) {
            llo = P.lex[Ilo];
            lhi = P.lex[Ihi];
        }
    }
    Halo hl{};
    load_halo(P, m, N, i, bx, by, jc, kc, tx, ty, hl);
    store_tile(P, tx, ty, m0, l0, hl, tile, ltile);
    __syncthreads();
    float H[3] = {0.f, 0.f, 0.f};
    if (P.exchange && in) exchange_cell<TX>(P, m, N, i, j, k, jc, tx + 1, ty, m0, lo, hi, l0, llo, lhi, tile, ltile, H);
    if (in) finish_cell<ROLE>(P, sp, io, I, N, m0, H, q, errloc);
    flush_err<ROLE>(errloc, io);
}

// ---- two cells per thread ------------------------------------------------
// A 64 x 4 thread block covers 128 (k) x 4 (j) cells of one plane; each
// thread owns cells (k, k+1), so every FP32 plane moves as 8-byte and the
// FP64 state as 16-byte accesses (4-byte lanes cap a streaming kernel well
// below the HBM rate).  Needs N2 even.
constexpr int TX2 = 2 * TX;

__device__ __forceinline__ float2 ld2(const float* p, long i) { return *(const float2*)(p + i); }
__device__ __forceinline__ float2 ld2m(const float* p, long i) { return p ? ld2(p, i) : make_float2(1.f, 1.f); }

template <int ROLE>
__device__ __forceinline__ void preload2(const Phys& P, const StageIO& io, long I, long N, Pre q[2]) {
    float2 v;
    v = ld2m(P.msk, I);
    q[0].ms = v.x, q[1].ms = v.y;
    if (P.demag)
        for (int c = 0; c < 3; c++) {
            v = ld2(io.B, c * N + I);
            q[0].B[c] = v.x, q[1].B[c] = v.y;
        }
    if (P.longfield || io.qs) {
        v = ld2m(P.J, I);
        q[0].J = v.x, q[1].J = v.y;
        v = ld2m(P.n, I);
        q[0].n = v.x, q[1].n = v.y;
        v = ld2m(P.Tc, I);
        q[0].Tc = v.x, q[1].Tc = v.y;
    }
    v = io.T ? ld2(io.T, I) : ld2m(P.T, I);
    q[0].T = v.x, q[1].T = v.y;
    for (int c = 0; c < 3; c++) {
        v = ld2m(P.alpha[c], I);
        q[0].al[c] = v.x, q[1].al[c] = v.y;
    }
    const double* ys = nullptr;
    if constexpr (ROLE == R_RK4_1 || ROLE == R_BS_1 || ROLE == R_RK12_1 || ROLE == R_EULER || ROLE == R_RK4_2 ||
                  ROLE == R_RK4_3 || ROLE == R_BS_2 || ROLE == R_RK4_4 || ROLE == R_BS_3)
        ys = io.y0;
    if constexpr (ROLE == R_RK12_2) ys = io.acc;
    if constexpr (ROLE == R_BS_4) ys = P.normalize ? io.acc : nullptr;
    if (ys)
        for (int c = 0; c < 3; c++) {
            double2 w = *(const double2*)(ys + c * N + I);
            q[0].y[c] = w.x, q[1].y[c] = w.y;
        }
    auto ldk = [&](const float* src, int which) {
        for (int c = 0; c < 3; c++) {
            float2 u = ld2(src, c * N + I);
            float* d0 = which == 0 ? q[0].ka : (which == 1 ? q[0].kb : q[0].kc);
            float* d1 = which == 0 ? q[1].ka : (which == 1 ? q[1].kb : q[1].kc);
            d0[c] = u.x;
            d1[c] = u.y;
        }
    };
    if constexpr (ROLE == R_RK4_4 || ROLE == R_BS_3) {
        ldk(io.k1, 0);
        ldk(io.k2, 1);
    }
    if constexpr (ROLE == R_RK4_4) ldk(io.k3, 2);
    if constexpr (ROLE == R_BS_4 || ROLE == R_RK12_2) ldk(io.e, 0);
}

template <int ROLE, bool LEAN>
__global__ void __launch_bounds__(TX* TY) k_stage_pair(Phys P_, StageIO io_) {
    const Phys P = lean_phys<LEAN>(P_);
    const StageIO io0 = lean_io<LEAN>(io_);
    __shared__ float tile[3][TY + 2][TX2 + 2];
    __shared__ float ltile[TY + 2][TX2 + 2];
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int N0 = P.N0, N1 = P.N1, N2 = P.N2;
    const int kb = blockIdx.x * TX2;  // first cell of the tile row
    const int k = kb + 2 * tx, j = blockIdx.y * TY + ty, i = blockIdx.z % N0;
    const int sol = blockIdx.z / N0;
    if (solve_idle(io0, sol)) return;
    const long N12 = (long)N1 * N2;
    const long N = (long)N0 * N12;
    const StageIO io = solve_io(io0, sol, N);
    // per-solve constants read where they are used (a copy here would hold
    // them all in scalar registers through the kernel)
    const SolveP* sp = P.sp + sol;
    const bool in = (k < N2) && (j < N1);  // N2 even: k < N2 implies k + 1 < N2
    const float* m = io.m;
    const int kc = min(k, N2 - 2), jc = min(j, N1 - 1);
    const long I = i * N12 + (long)jc * N2 + kc;
    float2 m0[3], lo[3], hi[3];
    float2 l0 = make_float2(1.f, 1.f), llo = l0, lhi = l0;
    for (int c = 0; c < 3; c++) m0[c] = ld2(m, c * N + I);
    if (P.lex) l0 = ld2(P.lex, I);
    Pre q[2];
    preload2<ROLE>(P, io, I, N, q);
    if (P.exchange) {
        const long Ilo = wrapc(i - 1, N0, P.w0) * N12 + (long)jc * N2 + kc;
        const long Ihi = wrapc(i + 1, N0, P.w0) * N12 + (long)jc * N2 + kc;
        for (int c = 0; c < 3; c++) {
            lo[c] = ld2(m, c * N + Ilo);
            hi[c] = ld2(m, c * N + Ihi);
        }
        if (P.lex) {
            llo = ld2(P.lex, Ilo);
            lhi = ld2(P.lex, Ihi);
        }
    }
    // halo values into registers with the other loads: every load of the
    // block is in flight before the first wait (the LDS stores below)
    float2 hy[3] = {}, hly = make_float2(1.f, 1.f);
    float hx[3] = {}, hlx = 1.f;
    const bool yedge = ty == 0 || ty == TY - 1, xedge = tx == 0 || tx == TX - 1;
    if (P.exchange) {
        if (yedge) {  // halo rows, two cells per thread
            const int jn = (ty == 0) ? wrapc(blockIdx.y * TY - 1, N1, P.w1)
                                     : wrapc(min(blockIdx.y * TY + TY, N1), N1, P.w1);
            const long a2 = i * N12 + (long)jn * N2 + kc;
            for (int c = 0; c < 3; c++) hy[c] = ld2(m, c * N + a2);
            if (P.lex) hly = ld2(P.lex, a2);
        }
        if (xedge) {  // halo columns kb-1 and kb+128
            const int kn = (tx == 0) ? wrapc(kb - 1, N2, P.w2) : wrapc(min(kb + TX2, N2), N2, P.w2);
            const long a2 = i * N12 + (long)jc * N2 + kn;
            for (int c = 0; c < 3; c++) hx[c] = m[c * N + a2];
            if (P.lex) hlx = P.lex[a2];
        }
    }
    const int cc = 1 + 2 * tx;  // tile column of cell k
    for (int c = 0; c < 3; c++) {
        tile[c][ty + 1][cc] = m0[c].x;
        tile[c][ty + 1][cc + 1] = m0[c].y;
    }
    if (P.lex) {
        ltile[ty + 1][cc] = l0.x;
        ltile[ty + 1][cc + 1] = l0.y;
    }
    if (P.exchange) {
        if (yedge) {
            const int row = (ty == 0) ? 0 : TY + 1;
            for (int c = 0; c < 3; c++) {
                tile[c][row][cc] = hy[c].x;
                tile[c][row][cc + 1] = hy[c].y;
            }
            if (P.lex) {
                ltile[row][cc] = hly.x;
                ltile[row][cc + 1] = hly.y;
            }
        }
        if (xedge) {
            const int col = (tx == 0) ? 0 : TX2 + 1;
            for (int c = 0; c < 3; c++) tile[c][ty + 1][col] = hx[c];
            if (P.lex) ltile[ty + 1][col] = hlx;
        }
    }
    __syncthreads();
    double errloc = 0.0;
    Out o[2];
#pragma unroll
    for (int e = 0; e < 2; e++) {
        const float mm[3] = {e ? m0[0].y : m0[0].x, e ? m0[1].y : m0[1].x, e ? m0[2].y : m0[2].x};
        const float ml[3] = {e ? lo[0].y : lo[0].x, e ? lo[1].y : lo[1].x, e ? lo[2].y : lo[2].x};
        const float mh[3] = {e ? hi[0].y : hi[0].x, e ? hi[1].y : hi[1].x, e ? hi[2].y : hi[2].x};
        float H[3] = {0.f, 0.f, 0.f};
        if (P.exchange && in)
            exchange_cell<TX2>(P, m, N, i, j, k + e, jc, cc + e, ty, mm, ml, mh, e ? l0.y : l0.x,
                               e ? llo.y : llo.x, e ? lhi.y : lhi.x, tile, ltile, H);
        if (in) {
            float kk[3];
            field_torque(P, sp, io, I + e, mm, H, q[e], kk);
            stage_update<ROLE>(P, io, q[e], kk, H, o[e], errloc);
        }
    }
    if (in) store_out2<ROLE>(P, io, I, N, o[0], o[1], q[0], q[1]);
    flush_err<ROLE>(errloc, io);
}

// RK12 retry: predictor from the stored k1 without a new field evaluation
// (solve_rk12.go:127-134 restores y0 and dy, then Madd).
__global__ void k_rk12_predict(const double* y0, const float* k1, double* acc, float* mout, double h, int normalize,
                               long N) {
    long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= N) return;
    const long b = (long)blockIdx.y * 3 * N;
    y0 += b;
    k1 += b;
    acc += b;
    mout += b;
    double v[3];
    for (int c = 0; c < 3; c++) {
        v[c] = y0[c * N + I] + h * (double)k1[c * N + I];
        acc[c * N + I] = v[c];
    }
    if (normalize) {
        double l = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        double s = l == 0.0 ? 0.0 : 1.0 / l;
        for (int c = 0; c < 3; c++) v[c] *= s;
    }
    for (int c = 0; c < 3; c++) mout[c * N + I] = (float)v[c];
}

// Fixed-point update of the implicit drivers (solve_am01.go:118-140 and
// 188-206, solve_euler_backward.go:66-78): from the RHS k of the current
// iterate y (acc; normalised first when norm_in is set: the reference's RHS
// update normalises m in place, so every iterate whose RHS was evaluated is
// compared normalised),
//   PRED (dy0 == null): y_new = y0 + c k          (Madd)
//   TRAP:               y_new = y0 + c (k + dy0)  (AddMadd)
// acc <- y_new, mout <- the FP32 stage input of y_new.  Maxima over every
// cell of every solve: red[p] = |y_p - y_new,p|, red[3+p] = |y_new,p|,
// red[6+p] = |y_new,p - y1,p| (y1 given), red[9] = |y + y_new| over all
// components, red[10] = 1 when a non-finite value appeared.  Each maximum
// has kImpLanes lanes (128-byte apart) that blocks spread their atomics over.
constexpr int kImpVals = 11, kImpLanes = 16, kImpStride = 16;
__global__ void __launch_bounds__(256) k_imp(const double* __restrict__ y0, double* __restrict__ acc,
                                             const float* __restrict__ k, const float* __restrict__ dy0,
                                             const double* __restrict__ y1, float* __restrict__ mout, double c,
                                             int norm_in, int norm_out, unsigned long long* __restrict__ red,
                                             long N) {
    const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long b = (long)blockIdx.y * 3 * N;
    double r[kImpVals];
    for (int v = 0; v < kImpVals; v++) r[v] = 0.0;
    if (I < N) {
        double yc[3], yn[3];
        for (int p = 0; p < 3; p++) yc[p] = acc[b + p * N + I];
        if (norm_in) {
            double l = sqrt(yc[0] * yc[0] + yc[1] * yc[1] + yc[2] * yc[2]);
            double s = (l == 0.0) ? 0.0 : 1.0 / l;
            for (int p = 0; p < 3; p++) yc[p] *= s;
        }
        for (int p = 0; p < 3; p++) {
            const long a = b + p * N + I;
            const double kv = (double)k[a];
            yn[p] = dy0 ? y0[a] + c * (kv + (double)dy0[a]) : y0[a] + c * kv;
            r[p] = fabs(yc[p] - yn[p]);
            r[3 + p] = fabs(yn[p]);
            if (y1) r[6 + p] = fabs(yn[p] - y1[a]);
            r[9] = fmax(r[9], fabs(yc[p] + yn[p]));
            if (!isfinite(yn[p]) || !isfinite(yc[p])) r[10] = 1.0;
            acc[a] = yn[p];
        }
        double w[3] = {yn[0], yn[1], yn[2]};
        if (norm_out) {
            double l = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
            double s = (l == 0.0) ? 0.0 : 1.0 / l;
            for (int p = 0; p < 3; p++) w[p] *= s;
        }
        for (int p = 0; p < 3; p++) mout[b + p * N + I] = (float)w[p];
    }
    __shared__ double wm[4][kImpVals];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int v = 0; v < kImpVals; v++) {
        double x = r[v];
        for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, 64));
        if (lane == 0) wm[wave][v] = x;
    }
    __syncthreads();
    if (threadIdx.x < kImpVals) {
        const int v = threadIdx.x;
        const double x = fmax(fmax(wm[0][v], wm[1][v]), fmax(wm[2][v], wm[3][v]));
        const unsigned ln = (blockIdx.x + gridDim.x * blockIdx.y) % kImpLanes;
        atomicMax(red + ((long)v * kImpLanes + ln) * kImpStride, (unsigned long long)__double_as_longlong(x));
    }
}

// The same fixed-point update for a one-component FP64 equation (Te, Tp):
// red[0] = |y - y_new|, red[1] = |y_new|, red[2] = |y_new - y1| (y1 given),
// red[3] = 1 on a non-finite value; fout (optional) takes the FP32 copy of
// y_new (the LLB stage's temperature).
constexpr int kImp1Vals = 4;
constexpr int kImpAll = kImpVals + 2 * kImp1Vals;  // m, Te, Tp maxima of one fixed-point update

// Maxima q < nq, each over its kImpLanes lanes (value q, lane l at
// src[(q kImpLanes + l) kImpStride]), into dst[q] for one small readback;
// the same max as the host's std::max over the lanes.  Each lane is left at
// zero for the next update's atomics.
__global__ void k_imp_fold(unsigned long long* __restrict__ src, int nq, double* __restrict__ dst) {
    const int q = threadIdx.x;
    if (q >= nq) return;
    double v = 0.0;
    for (int l = 0; l < kImpLanes; l++) {
        unsigned long long* a = src + ((long)q * kImpLanes + l) * kImpStride;
        const double x = __longlong_as_double((long long)*a);
        *a = 0ull;
        v = (v < x) ? x : v;
    }
    dst[q] = v;
}
__global__ void __launch_bounds__(256) k_imp1(const double* __restrict__ y0, double* __restrict__ acc,
                                              const double* __restrict__ k, const double* __restrict__ dy0,
                                              const double* __restrict__ y1, float* __restrict__ fout, double c,
                                              unsigned long long* __restrict__ red, long N) {
    const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double r[kImp1Vals] = {0.0, 0.0, 0.0, 0.0};
    if (I < N) {
        const double yc = acc[I];
        const double yn = dy0 ? y0[I] + c * (k[I] + dy0[I]) : y0[I] + c * k[I];
        r[0] = fabs(yc - yn);
        r[1] = fabs(yn);
        if (y1) r[2] = fabs(yn - y1[I]);
        if (!isfinite(yn) || !isfinite(yc)) r[3] = 1.0;
        acc[I] = yn;
        if (fout) fout[I] = (float)yn;
    }
    __shared__ double wm[4][kImp1Vals];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int v = 0; v < kImp1Vals; v++) {
        double x = r[v];
        for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, 64));
        if (lane == 0) wm[wave][v] = x;
    }
    __syncthreads();
    if (threadIdx.x < kImp1Vals) {
        const int v = threadIdx.x;
        const double x = fmax(fmax(wm[0][v], wm[1][v]), fmax(wm[2][v], wm[3][v]));
        const unsigned ln = blockIdx.x % kImpLanes;
        atomicMax(red + ((long)v * kImpLanes + ln) * kImpStride, (unsigned long long)__double_as_longlong(x));
    }
}

// y = y0 + c k (one component), with its FP32 copy
__global__ void k_axpy1(const double* __restrict__ y0, const double* __restrict__ k, double* __restrict__ y,
                        float* __restrict__ fout, double c, long N) {
    const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= N) return;
    const double v = y0[I] + c * k[I];
    y[I] = v;
    if (fout) fout[I] = (float)v;
}

__global__ void k_set_state(const double* y, float* mf, int normalize, long N) {
    long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= N) return;
    const long b = (long)blockIdx.y * 3 * N;
    y += b;
    mf += b;
    double v[3] = {y[I], y[N + I], y[2 * N + I]};
    if (normalize) {
        double l = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        double s = l == 0.0 ? 0.0 : 1.0 / l;
        for (int c = 0; c < 3; c++) v[c] *= s;
    }
    for (int c = 0; c < 3; c++) mf[c * N + I] = (float)v[c];
}

__global__ void k_d2f(const double* a, float* b, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

__global__ void k_clock_step(Clock* clk, int nsolve) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nsolve) clk[q].step += 1;
}

// the reference's step-size controller, solve_rk12.go:164-198
__host__ __device__ inline double controller(double dt, double err, double maxErr, bool bad, double expo_good,
                                             double expo_bad, double mindt, double maxdt) {
    double factor = bad ? pow(maxErr / err, expo_bad) : pow(maxErr / err, expo_good);
    factor *= 0.8;
    if (factor > 1.5) factor = 1.5;
    if (factor < 0.1) factor = 0.1;
    double nd = dt * factor;
    if (nd < mindt) nd = mindt;
    if (nd > maxdt) nd = maxdt;
    return nd;
}

// Per-slot accept/reject after an RK23 attempt (one thread per slot).  The
// state of iteration `par` is in buffer set par; an accepted attempt leaves
// the new state in set par^1, a rejected one is copied there by k_fixup.
// With neq = 3 (two-temperature solves) the step follows the reference's
// multi-equation rule (solve_rk12.go:150-190): bad if any equation exceeds
// its bound, factor = the smallest of the equations' factors.
__global__ void k_control(Clock* clk, unsigned long long* errmax, int* live, int nsolve, uint32_t par, int neq) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nsolve) return;
    Clock c = clk[q];
    if (c.state != HS_SLOT_RUNNING) {
        live[q] = 0;
        return;
    }
    double errs[kNeq] = {0.0, 0.0, 0.0};
    bool finite = true;
    for (int e = 0; e < neq; e++) {
        unsigned long long* em = errmax + (long)q * kErrPerSolve + e * kErrPerEq;
        for (int l = 0; l < kErrLanes; l++) {
            errs[e] = fmax(errs[e], __longlong_as_double((long long)em[l * kErrStride]));
            em[l * kErrStride] = 0ull;
        }
        finite = finite && isfinite(errs[e]);
    }
    const double err = errs[0];
    c.err = err;
    if (!finite) {
        // a non-finite estimate (the stage values overflowed or went NaN) is
        // never accepted: retry at the smallest factor while retries and dt
        // remain, otherwise the slot has diverged
        const double nd = controller(c.h, errs[0] + errs[1] + errs[2], 1.0, true, 0.5, 1.0 / 3.0, c.mindt, c.maxdt);
        if (c.attempt >= (uint32_t)kMaxTry + 1 || !(nd < c.h)) {
            c.state = HS_SLOT_FAILED;
            c.home = par;
            c.retry = 0;
        } else {
            c.h = nd;
            c.attempt += 1;
            c.rejected += 1;
            c.retry = 1;
        }
        clk[q] = c;
        live[q] = c.state == HS_SLOT_RUNNING ? 1 : 0;
        return;
    }
    bool bad = false;
    double minf = 2.0;
    for (int e = 0; e < neq; e++) {
        const double me = e ? c.maxErrT : c.maxErr;
        if (errs[e] > me) bad = true;
        double f = pow(me / errs[e], 1.0 / 3.0) * 0.8;
        if (f > 1.5) f = 1.5;
        if (f < 0.1) f = 0.1;
        minf = fmin(minf, f);
    }
    double nd = c.h * minf;
    if (nd < c.mindt) nd = c.mindt;
    if (nd > c.maxdt) nd = c.maxdt;
    if (!bad || nd == c.mindt) {
        c.t += c.h;
        c.h = nd;
        c.step += 1;
        c.attempt = 0;
        c.accepted += 1;
        c.retry = 0;
        if (c.step >= c.end) {
            c.state = HS_SLOT_DONE;
            c.home = par ^ 1u;
        }
    } else {
        c.h = nd;
        if (c.attempt >= (uint32_t)kMaxTry + 1) {
            c.state = HS_SLOT_FAILED;
            c.home = par;
            c.retry = 0;
        } else {
            c.attempt += 1;
            c.rejected += 1;
            c.retry = 1;
        }
    }
    clk[q] = c;
    live[q] = c.state == HS_SLOT_RUNNING ? 1 : 0;
}

// A rejected slot restarts its step from the state of set `cur`: copy it
// (FP64 state, FP32 stage input and its field) into set `nxt`.
__global__ void k_fixup(const Clock* clk, const double* ycur, const float* mcur, const float* Bcur, double* ynxt,
                        float* mnxt, float* Bnxt, long N, int demag) {
    const int q = blockIdx.y;
    if (clk[q].state != HS_SLOT_RUNNING || !clk[q].retry) return;
    long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= N) return;
    const long b = (long)q * 3 * N;
    for (int c = 0; c < 3; c++) {
        ynxt[b + c * N + I] = ycur[b + c * N + I];
        mnxt[b + c * N + I] = mcur[b + c * N + I];
        if (demag) Bnxt[b + c * N + I] = Bcur[b + c * N + I];
    }
}

// ---- two-temperature model ------------------------------------------------
struct TmP {
    double gammaE, Cl, G, kappa, Q0, ir2, it2, t0, xc, yc;  // ir2 = 1/(2 sigma_r^2), it2 = 1/(2 sigma_t^2)
    double Td, c9nkB;  // Debye lattice capacity (Td > 0): Cp = c9nkB * debye_cp(Tp, Td), c9nkB = 9 nD kB
    double c2x, c2y, c2z;  // 1/cell^2 of internal axes 0, 1, 2 (Qspat.cu's x, y, z)
    double hx, hy;         // cell sizes of the user x (internal 2) and y (internal 1) axes
};
struct TmIO {
    const double *Te, *Tp;        // stage input
    const double *Te0, *Tp0;      // state at step start (BS_3)
    double *Teo, *Tpo;            // next stage input (BS_3: the new state)
    float* Tfo;                   // FP32 copy of the next stage's Te for the LLB fields
    double *k1e, *k1p, *k2e, *k2p;
    double *ee, *ep;              // error partials (BS_3 -> BS_4)
    unsigned long long* errmax;   // solve base; Te lanes at +kErrPerEq, Tp at +2 kErrPerEq
    const Clock* clk;
    const SolveP* sp;
    double t;                     // laser time of stage-1 (no clock: testing)
    const float* qs;              // spin term of Qe per cell at the stage input (or null)
};

// Debye capacity of one cell with n = 1, as cpKern (Cp.cu:14-83): (T/Td)^3 times
// the 61-interval trapezoid of x^4 e^x/(e^x-1)^2 over [0, Td/T]; below
// T/Td = 0.01 the T^3 limit.  e^{x_j} by repeated products of e^h (one exp).
__device__ double debye_cp(double T, double Td) {
    const double xx = Td / T, ixx = T / Td;
    if (ixx < 0.01) return (12.0 / 5.0) * (1.0 / 9.0) * (kPiD * kPiD * kPiD * kPiD) * (ixx * ixx * ixx);
    const double h = xx / 61.0, eh = exp(h);
    double e = 1.0, inner = 0.0;
    for (int j = 1; j < 61; j++) {
        e *= eh;
        const double x = j * h, em1 = e - 1.0;
        inner += (x * x * x * x * e) / (em1 * em1);
    }
    const double ex = exp(xx), exm1 = ex - 1.0;
    const double end = (xx * xx * xx * xx * ex) / (exm1 * exm1);
    return ((0.0 + end + 2.0 * inner) * (0.5 * h)) / (xx * xx * xx);
}

// Q_e = kappa lap(Te) (Qspat.cu:13-134 with uniform conductance, neighbour
// order X, Z, Y as there) + G (Tp - Te) (Qinter.cu:13-45) + Q_laser;
// dTe/dt = Q_e / (gammaE Te) (TM.go:40-59 with mod_ce.go:73-94), dTp/dt = G (Te - Tp) / Cl.
// ROLE: R_RHS (slopes to k1e/k1p at time t) or the four Bogacki-Shampine stages.
template <int ROLE>
__global__ void __launch_bounds__(256) k_tm(Phys P, TmP Q, TmIO io) {
    __shared__ double wm[2][4];
    const int sol = blockIdx.y;
    if (io.clk && io.clk[sol].state != HS_SLOT_RUNNING) return;
    const int N0 = P.N0, N1 = P.N1, N2 = P.N2;
    const long N12 = (long)N1 * N2, N = (long)N0 * N12;
    const long b = (long)sol * N;
    const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double el = 0.0, pl = 0.0;
    if (I < N) {
        const int i = (int)(I / N12), j = (int)((I / N2) % N1), k = (int)(I % N2);
        const double* Te = io.Te + b;
        const double T0 = Te[I], Tp = io.Tp[b + I];
        auto at = [&](int ii, int jj, int kk) { return Te[(long)ii * N12 + (long)jj * N2 + kk]; };
        int a, c;
        double ddT;
        a = i - 1; a = (a < 0 && P.w0) ? N0 + a : a; a = a < 0 ? 0 : a;
        c = i + 1; c = (c == N0 && P.w0) ? c - N0 : c; c = c > N0 - 1 ? N0 - 1 : c;
        ddT = Q.c2x * ((at(a, j, k) + at(c, j, k)) - (T0 + T0));
        a = k - 1; a = (a < 0 && P.w2) ? N2 + a : a; a = a < 0 ? 0 : a;
        c = k + 1; c = (c == N2 && P.w2) ? c - N2 : c; c = c > N2 - 1 ? N2 - 1 : c;
        ddT += Q.c2z * ((at(i, j, a) + at(i, j, c)) - (T0 + T0));
        a = j - 1; a = (a < 0 && P.w1) ? N1 + a : a; a = a < 0 ? 0 : a;
        c = j + 1; c = (c == N1 && P.w1) ? c - N1 : c; c = c > N1 - 1 ? N1 - 1 : c;
        ddT += Q.c2y * ((at(i, a, k) + at(i, c, k)) - (T0 + T0));
        double h = 0.0, ts = io.t;
        if (io.clk) {
            const Clock& ck = io.clk[sol];
            h = ck.h;
            const double cs = (ROLE == R_BS_2) ? 0.5 : (ROLE == R_BS_3) ? 0.75 : (ROLE == R_BS_4) ? 1.0 : 0.0;
            ts = io.sp[sol].t_origin + ck.t + cs * h;
        }
        const double x = (k + 0.5) * Q.hx - Q.xc, y = (j + 0.5) * Q.hy - Q.yc, dt0 = ts - Q.t0;
        const double ql = Q.Q0 * exp(-(x * x + y * y) * Q.ir2) * exp(-(dt0 * dt0) * Q.it2);
        double qe = Q.kappa * ddT + Q.G * (Tp - T0) + ql;
        if (io.qs) qe += (double)io.qs[b + I];
        const double cl = Q.Td > 0.0 ? Q.c9nkB * debye_cp(Tp, Q.Td) : Q.Cl;
        const double ke = qe / (Q.gammaE * T0), kp = Q.G * (T0 - Tp) / cl;
        const long J = b + I;
        if constexpr (ROLE == R_RHS) {
            io.k1e[J] = ke;
            io.k1p[J] = kp;
        } else if constexpr (ROLE == R_BS_1 || ROLE == R_BS_2) {
            const double cm = (ROLE == R_BS_1) ? 0.5 * h : 0.75 * h;
            double* ke_ = (ROLE == R_BS_1) ? io.k1e : io.k2e;
            double* kp_ = (ROLE == R_BS_1) ? io.k1p : io.k2p;
            ke_[J] = ke;
            kp_[J] = kp;
            const double te = io.Te0[J] + cm * ke, tp = io.Tp0[J] + cm * kp;
            io.Teo[J] = te;
            io.Tpo[J] = tp;
            io.Tfo[J] = (float)te;
        } else if constexpr (ROLE == R_BS_3) {
            const double k1e = io.k1e[J], k2e = io.k2e[J], k1p = io.k1p[J], k2p = io.k2p[J];
            const double te = io.Te0[J] + h * ((2.0 / 9.0) * k1e + (1.0 / 3.0) * k2e + (4.0 / 9.0) * ke);
            const double tp = io.Tp0[J] + h * ((2.0 / 9.0) * k1p + (1.0 / 3.0) * k2p + (4.0 / 9.0) * kp);
            io.Teo[J] = te;
            io.Tpo[J] = tp;
            io.Tfo[J] = (float)te;
            io.ee[J] = h * ((-5.0 / 72.0) * k1e + (1.0 / 12.0) * k2e + (1.0 / 9.0) * ke);
            io.ep[J] = h * ((-5.0 / 72.0) * k1p + (1.0 / 12.0) * k2p + (1.0 / 9.0) * kp);
        } else if constexpr (ROLE == R_BS_4) {
            el = err_abs(io.ee[J] - (h / 8.0) * ke);
            pl = err_abs(io.ep[J] - (h / 8.0) * kp);
        }
    }
    if constexpr (ROLE == R_BS_4) {
        for (int o = 32; o > 0; o >>= 1) {
            el = fmax(el, __shfl_xor(el, o, 64));
            pl = fmax(pl, __shfl_xor(pl, o, 64));
        }
        if ((threadIdx.x & 63) == 0) {
            wm[0][threadIdx.x >> 6] = el;
            wm[1][threadIdx.x >> 6] = pl;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
                el = fmax(el, wm[0][w]);
                pl = fmax(pl, wm[1][w]);
            }
            unsigned long long* em = io.errmax + (long)sol * kErrPerSolve;
            const unsigned lane = blockIdx.x % kErrLanes;
            atomicMax(em + kErrPerEq + lane * kErrStride, (unsigned long long)__double_as_longlong(el));
            atomicMax(em + 2 * kErrPerEq + lane * kErrStride, (unsigned long long)__double_as_longlong(pl));
        }
    }
}

// k_fixup / k_gather for the temperature buffers of a buffer set
__global__ void k_fixup_tm(const Clock* clk, const double* tec, const double* tpc, const float* tfc, double* ten,
                           double* tpn, float* tfn, long N) {
    const int q = blockIdx.y;
    if (clk[q].state != HS_SLOT_RUNNING || !clk[q].retry) return;
    long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= N) return;
    const long J = (long)q * N + I;
    ten[J] = tec[J];
    tpn[J] = tpc[J];
    tfn[J] = tfc[J];
}
__global__ void k_gather_tm(const Clock* clk, const double* te1, const double* tp1, const float* tf1, double* te0,
                            double* tp0, float* tf0, long N) {
    const int q = blockIdx.y;
    if (clk[q].home != 1u) return;
    long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= N) return;
    const long J = (long)q * N + I;
    te0[J] = te1[J];
    tp0[J] = tp1[J];
    tf0[J] = tf1[J];
}

}  // namespace

struct hs_solver {
    hs_llb_params p;
    hipStream_t st;
    bool own_stream = false;
    long N;       // cells per solve
    int nsolve = 1;
    SolveP* sp = nullptr;      // device, nsolve entries
    LfConst* lfc = nullptr;    // device, the uniform longitudinal field's constants
    Clock* clk = nullptr;      // device, nsolve entries (graph replay, slots)
    hipGraphExec_t graph = nullptr;
    // slots: one captured RK23 iteration per buffer-set parity
    bool slot_mode = false;
    long slot_iter = 0;                 // iterations enqueued (parity = slot_iter & 1)
    hipGraphExec_t slot_graph[2] = {nullptr, nullptr};
    hs_slot* slot_host = nullptr;       // pinned copy of the slot records
    int* live = nullptr;                // device, 1 per running slot (demag passes skip the others)
    hipEvent_t slot_ev = nullptr;
    std::vector<SolveP> sp_host;
    int graph_method = -1;
    double graph_dt = 0;
    hs_demag* demag = nullptr;
    Phys phys;
    // device buffers
    double *y0 = nullptr, *acc = nullptr;
    float *mA = nullptr, *mB = nullptr, *mC = nullptr;  // FP32 stage inputs
    float *B0 = nullptr, *Bs = nullptr;
    float* e = nullptr;
    float *k1 = nullptr, *k2 = nullptr, *k3 = nullptr;
    float *kout = nullptr, *hout = nullptr;
    unsigned long long* errmax = nullptr;
    std::vector<float*> masks;
    bool b0_valid = false;  // B0 = demag(mA) is current
    bool lean_ok = true;     // lean stage kernels allowed (HS_STAGE_LEAN=0 forces the general form)
    int stage_kind = 1;      // 0 marching planes, 1 one plane per block, 2 two cells per thread, 3 = 1 in XCD plane order
    // two-temperature coupling (tm): Te, Tp in two buffer sets (state / candidate), two stage inputs
    bool tm = false;
    hs_tm_params tmp{};
    TmP tmc{};
    double tm_time = 0.0;  // laser clock
    double *Te[2] = {nullptr, nullptr}, *Tp[2] = {nullptr, nullptr};
    float* Tf[2] = {nullptr, nullptr};
    double *Tea = nullptr, *Tpa = nullptr, *Teb = nullptr, *Tpb = nullptr;
    float *Tfa = nullptr, *Tfb = nullptr;
    double *k1e = nullptr, *k1p = nullptr, *k2e = nullptr, *k2p = nullptr, *ee = nullptr, *ep = nullptr;
    // implicit drivers (AM12, backward Euler): fixed-point maxima, the predictor
    // result y1, the per-equation tolerances (reference defaults)
    unsigned long long* imp = nullptr;
    double* y1 = nullptr;
    double imp_abs = 1e-4, imp_rel = 1e-3, imp_iter_err = 1e-5;
    int imp_max_iter = 3;
    double imp_absT = 1e-4, imp_relT = 1e-3;  // Te and Tp (two-temperature solves)
    float* qs = nullptr;  // spin heating: the spin term of Qe per cell, written by the LLB stages
    double* epart = nullptr;  // energy: block partials, then the per-solve terms
    double* imp_host = nullptr;  // pinned: the folded maxima of a fixed-point update
    bool imp_zeroed = false;     // imp lanes cleared once; k_imp_fold clears them after each read
    double* err_dev = nullptr;   // lockstep error estimate, folded on the device
    double* err_host = nullptr;  // pinned copy of it
};

namespace {
float* upload_mask(hs_solver* s, const double* h) {
    if (!h) return nullptr;
    double* tmp = nullptr;
    float* d = nullptr;
    hipMalloc(&tmp, sizeof(double) * s->N);
    hipMalloc(&d, sizeof(float) * s->N);
    hipMemcpy(tmp, h, sizeof(double) * s->N, hipMemcpyHostToDevice);
    hipLaunchKernelGGL(k_d2f, hs::grid1d(s->N), hs::kBlock, 0, s->st, tmp, d, s->N);
    hipStreamSynchronize(s->st);
    hipFree(tmp);
    s->masks.push_back(d);
    return d;
}

template <int ROLE, bool LEAN>
void launch_stage_form(hs_solver* s, const StageIO& io) {
    const Phys& P = s->phys;
    if (s->stage_kind == 2) {
        dim3 blk(TX, TY), grd(hs::div_up(P.N2, TX2), hs::div_up(P.N1, TY), P.N0 * s->nsolve);
        hipLaunchKernelGGL((k_stage_pair<ROLE, LEAN>), grd, blk, 0, s->st, P, io);
    } else if (s->stage_kind == 1) {
        dim3 blk(TX, TY), grd(hs::div_up(P.N2, TX), hs::div_up(P.N1, TY), P.N0 * s->nsolve);
        hipLaunchKernelGGL((k_stage_flat<ROLE, false, LEAN>), grd, blk, 0, s->st, P, io);
    } else if (s->stage_kind == 3) {
        dim3 blk(TX, TY), grd(hs::div_up(P.N2, TX) * hs::div_up(P.N1, TY) * P.N0 * s->nsolve);
        hipLaunchKernelGGL((k_stage_flat<ROLE, true, LEAN>), grd, blk, 0, s->st, P, io);
    } else {
        dim3 blk(TX, TY), grd(hs::div_up(P.N2, TX), hs::div_up(P.N1, TY), hs::div_up(P.N0, P.chunk) * s->nsolve);
        hipLaunchKernelGGL((k_stage<ROLE, LEAN>), grd, blk, 0, s->st, P, io);
    }
}

// Whether a launch can take the lean form (see lean_phys).
bool stage_lean(const hs_solver* s, const StageIO& io) {
    const Phys& P = s->phys;
    return s->lean_ok && !P.lex && !P.msk && !P.J && !P.n && !P.Tc && !P.T && !P.alpha[0] && !P.alpha[1] &&
           !P.alpha[2] && !io.T && !io.qs && (!P.longfield || P.lf_uniform);
}

template <int ROLE>
void launch_stage(hs_solver* s, const StageIO& io) {
    hs::prof_begin(HS_PROF_STAGE, s->st);
    if (stage_lean(s, io)) launch_stage_form<ROLE, true>(s, io);
    else launch_stage_form<ROLE, false>(s, io);
    hs::prof_end(HS_PROF_STAGE, s->st);
}

StageIO base_io(hs_solver* s, double h, uint32_t step, uint32_t attempt) {
    StageIO io{};
    io.y0 = s->y0;
    io.acc = s->acc;
    io.e = s->e;
    io.k1 = s->k1;
    io.k2 = s->k2;
    io.k3 = s->k3;
    io.errmax = s->errmax;
    io.h = h;
    io.step = step;
    io.attempt = attempt;
    io.clk = nullptr;
    return io;
}

// Pick the fastest stage-kernel variant (marching / one plane per block /
// two cells per thread) by timing one representative role on the solver's
// own buffers.  HS_STAGE_KERNEL=march|flat|pair forces one.  The choice is
// cached per mesh shape for the life of the process (the variants may round
// differently, and a solve must give the same result in any batch).
std::mutex g_stage_mu;
std::map<std::array<int, 3>, std::array<int, 2>>& stage_cache() {  // shape -> (variant, march chunk)
    static std::map<std::array<int, 3>, std::array<int, 2>> c;
    return c;
}

void tune_stage(hs_solver* s) {
    const bool pair_ok = s->phys.N2 % 2 == 0;
    // (variant 3, the flat kernel in XCD plane order, wins this isolated timing
    // at 512x512x16 yet runs 10 % slower between the demag passes of a step,
    // so it is taken only when forced)
    const char* env = getenv("HS_STAGE_KERNEL");
    if (env && !strcmp(env, "flat")) { s->stage_kind = 1; return; }
    if (env && !strcmp(env, "flatz")) { s->stage_kind = 3; return; }
    if (env && !strcmp(env, "march")) { s->stage_kind = 0; return; }
    if (env && !strcmp(env, "pair") && pair_ok) { s->stage_kind = 2; return; }
    const std::array<int, 3> key{s->phys.N0, s->phys.N1, s->phys.N2};
    {
        std::lock_guard<std::mutex> lk(g_stage_mu);
        auto it = stage_cache().find(key);
        if (it != stage_cache().end()) {
            s->stage_kind = it->second[0];
            if (s->stage_kind == 0) s->phys.chunk = std::min(it->second[1], s->phys.N0);
            return;
        }
    }
    StageIO io = base_io(s, 1e-15, 0, 0);
    io.m = s->mA;
    io.B = s->B0;
    io.mout = s->mB;
    hipMemsetAsync(s->mA, 0, sizeof(float) * 3 * s->N * s->nsolve, s->st);
    hipMemsetAsync(s->y0, 0, sizeof(double) * 3 * s->N * s->nsolve, s->st);
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    float best = -1.f;
    int pick = 1, pick_chunk = s->phys.chunk;
    const int chunk0 = s->phys.chunk;
    // the marching variant at its default chunk and at 2, 4, 8 planes per block
    std::vector<std::array<int, 2>> cand;
    if (s->phys.N0 >= 2) {
        cand.push_back({0, chunk0});
        for (int c : {2, 4, 8})
            if (c != chunk0 && c < s->phys.N0) cand.push_back({0, c});
    }
    cand.push_back({1, chunk0});
    if (pair_ok) cand.push_back({2, chunk0});
    for (const auto& cv : cand) {
        s->stage_kind = cv[0];
        s->phys.chunk = cv[1];
        float t = 1e30f;
        launch_stage<R_RK4_2>(s, io);
        for (int round = 0; round < 2; round++) {  // best of two rounds of 3 launches
            hipEventRecord(e0, s->st);
            for (int r = 0; r < 3; r++) launch_stage<R_RK4_2>(s, io);
            hipEventRecord(e1, s->st);
            hipEventSynchronize(e1);
            float tr;
            hipEventElapsedTime(&tr, e0, e1);
            t = std::min(t, tr);
        }
        if (best < 0.f || t < best) {
            best = t;
            pick = cv[0];
            pick_chunk = cv[1];
        }
    }
    s->stage_kind = pick;
    s->phys.chunk = pick == 0 ? pick_chunk : chunk0;
    {
        std::lock_guard<std::mutex> lk(g_stage_mu);
        stage_cache()[key] = {pick, pick_chunk};
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
}

// demag field of the whole batch; with `live` (device-controlled stepping)
// the blocks of solves that are not running skip their rows
void demag_into(hs_solver* s, const float* m, float* B, const int* live = nullptr) {
    if (s->phys.demag)
        hs_demag_field_f32_internal(s->demag, m, s->phys.msk, s->p.msat0T0Mul, B, s->st, s->nsolve, live);
}

// Max of the error lanes em[q], q = 0, kErrStride, ... < n, in one block of
// 256 threads into *out: each thread keeps v = (v < x) ? x : v over its lanes
// (a NaN lane leaves v unchanged, as a sequential std::max from 0 does), then
// a tree of the same comparison over finite values.
__global__ void __launch_bounds__(256) k_err_fold(const unsigned long long* __restrict__ em, long n,
                                                  double* __restrict__ out) {
    __shared__ double red[256];
    const int t = threadIdx.x;
    double v = 0.0;
    for (long q = (long)t * kErrStride; q < n; q += 256L * kErrStride) {
        const double x = __longlong_as_double((long long)em[q]);
        v = (v < x) ? x : v;
    }
    red[t] = v;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) red[t] = (red[t] < red[t + h]) ? red[t + h] : red[t];
        __syncthreads();
    }
    if (t == 0) *out = red[0];
}

// max over the batch of the per-solve error maxima (lockstep stepping): folded
// on the device, one double back into pinned memory
double read_err(hs_solver* s) {
    const long n = (long)s->nsolve * kErrPerSolve;
    if ((!s->err_dev && hipMalloc(&s->err_dev, sizeof(double)) != hipSuccess) ||
        (!s->err_host && hipHostMalloc((void**)&s->err_host, sizeof(double)) != hipSuccess)) {
        hs::set_error(HS_ENOMEM, "error estimate: out of memory");
        return std::numeric_limits<double>::quiet_NaN();
    }
    hipLaunchKernelGGL(k_err_fold, 1, 256, 0, s->st, s->errmax, n, s->err_dev);
    hipMemcpyAsync(s->err_host, s->err_dev, sizeof(double), hipMemcpyDeviceToHost, s->st);
    hipStreamSynchronize(s->st);
    return *s->err_host;
}
void zero_err(hs_solver* s) {
    hipMemsetAsync(s->errmax, 0, sizeof(unsigned long long) * s->nsolve * kErrPerSolve, s->st);
}

void ensure_b0(hs_solver* s) {
    if (!s->b0_valid) {
        demag_into(s, s->mA, s->B0);
        s->b0_valid = true;
    }
}

// ---- energy terms ------------------------------------------------------------
// LoadEnergy (modules/mod_energy.go:25-61) and EnergyUpdater.Update (:98-116):
// E_term = weight * Ms0T0Mul * sum_cells (m . F) msat0T0 with F = B_ext
// (weight -V), H_ex (-V mu0 / 2), B (-V / 2) and H_lf (-V mu0).  One FP64
// pass over the state: m from the FP64 state, H_ex (exchange6.cu:11-127,
// neighbour order X, Z, Y) and H_lf (long_field.cu:13-72) in FP64, B from
// the FP32 demag of the state.  Block partials, then a fixed-order sum per
// solve (deterministic).
constexpr int kEnergyTerms = 4;  // zeeman, exchange, demag, longitudinal
struct EnergyC {
    double ms0T0, ex[3], JMul, nMul, TcMul;
};

__device__ double block_sum256(double v, double* red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    __syncthreads();
    if (l == 0) red[w] = v;
    __syncthreads();
    return (threadIdx.x == 0) ? ((red[0] + red[1]) + (red[2] + red[3])) : 0.0;
}

__global__ void __launch_bounds__(256) k_energy(Phys P, EnergyC C, const double* __restrict__ y,
                                                const float* __restrict__ B, const float* __restrict__ Tf,
                                                double* __restrict__ part, long N) {
    __shared__ double red[4];
    const int sol = blockIdx.y;
    const long I = (long)blockIdx.x * 256 + threadIdx.x;
    const SolveP& sp = P.sp[sol];
    double acc[kEnergyTerms] = {0.0, 0.0, 0.0, 0.0};
    if (I < N) {
        const double* ys = y + (long)sol * 3 * N;
        // m of a cell, unit length in the LLG limit (the state is normalised lazily)
        auto load_m = [&](long a, double v[3]) {
            for (int d = 0; d < 3; d++) v[d] = ys[d * N + a];
            if (P.normalize) {
                const double l = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                const double sc = (l == 0.0) ? 0.0 : 1.0 / l;
                for (int d = 0; d < 3; d++) v[d] *= sc;
            }
        };
        double m[3];
        load_m(I, m);
        const double ms = P.msk ? (double)P.msk[I] : 1.0;
        if (ms != 0.0) {
            if (P.zeeman) acc[0] = (m[0] * sp.bextT[0] + m[1] * sp.bextT[1] + m[2] * sp.bextT[2]) * ms;
            if (P.exchange) {
                const long N12 = (long)P.N1 * P.N2;
                const int ci[3] = {(int)(I / N12), (int)((I / P.N2) % P.N1), (int)(I % P.N2)};
                const int nn[3] = {P.N0, P.N1, P.N2}, ww[3] = {P.w0, P.w1, P.w2};
                const long st[3] = {N12, (long)P.N2, 1L};
                const double l0 = P.lex ? (double)P.lex[I] : 1.0, l02 = l0 * l0;
                double H[3] = {0.0, 0.0, 0.0};
                const int order[3] = {0, 2, 1};
                for (int q = 0; q < 3; q++) {
                    const int ax = order[q], n = nn[ax], c = ci[ax];
                    int lo = c - 1;
                    lo = (lo < 0 && ww[ax]) ? n + lo : lo;
                    lo = max(lo, 0);
                    int hi = c + 1;
                    hi = (hi == n && ww[ax]) ? hi - n : hi;
                    hi = min(hi, n - 1);
                    const long a1 = I + (long)(lo - c) * st[ax], a2 = I + (long)(hi - c) * st[ax];
                    const double l1 = P.lex ? (double)P.lex[a1] : 1.0, l2 = P.lex ? (double)P.lex[a2] : 1.0;
                    const double p1 = avg_geom_zero(l02, l1 * l1), p2 = avg_geom_zero(l02, l2 * l2);
                    double v1[3], v2[3];
                    load_m(a1, v1);
                    load_m(a2, v2);
                    for (int d = 0; d < 3; d++) H[d] += C.ex[ax] * ((p1 * v1[d] + p2 * v2[d]) - (p1 * m[d] + p2 * m[d]));
                }
                acc[1] = C.ms0T0 * (m[0] * H[0] + m[1] * H[1] + m[2] * H[2]) * ms;
            }
            if (P.demag) {
                const float* Bs = B + (long)sol * 3 * N;
                acc[2] = (m[0] * (double)Bs[I] + m[1] * (double)Bs[N + I] + m[2] * (double)Bs[2 * N + I]) * ms;
            }
            if (P.longfield) {
                const double Ms0T0 = C.ms0T0 * ms;
                const double S = C.JMul * (P.J ? (double)P.J[I] : 1.0);
                const double n = C.nMul * (P.n ? (double)P.n[I] : 1.0);
                const double Tc = C.TcMul * (P.Tc ? (double)P.Tc[I] : 1.0);
                const double Ts = (double)sp.TMul * (Tf ? (double)Tf[(long)sol * N + I] : (P.T ? (double)P.T[I] : 1.0));
                const double am = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
                if (!(Ms0T0 == 0.0 || n == 0.0 || am <= HS_ZERO)) {
                    const double preS = (S > HS_INFSPIN) ? 1.0 : S / (S + 1.0);
                    const double b = 3.0 * preS * Tc / Ts;
                    double W, dW;
                    brillouin_w_dw<double>(S, am * b, W, dW);
                    const double mult = n * (HS_KB / HS_MU0) * Ts * (W - am) / (Ms0T0 * dW) / am;
                    acc[3] = mult * (m[0] * m[0] + m[1] * m[1] + m[2] * m[2]) * ms;
                }
            }
        }
    }
    for (int t = 0; t < kEnergyTerms; t++) {
        const double v = block_sum256(acc[t], red);
        if (threadIdx.x == 0) part[((long)sol * gridDim.x + blockIdx.x) * kEnergyTerms + t] = v;
    }
}

// Per solve: the block partials summed in a fixed order, times Ms0T0Mul and
// the term weights; out[sol * 5 + t], t = 4 the sum of the four.
__global__ void __launch_bounds__(256) k_energy_sum(const double* __restrict__ part, int nb, double ms0T0, double V,
                                                    double* __restrict__ out) {
    __shared__ double red[4];
    const int sol = blockIdx.x;
    const double w[kEnergyTerms] = {-V, -0.5 * V * HS_MU0, -0.5 * V, -V * HS_MU0};
    double tot = 0.0;
    for (int t = 0; t < kEnergyTerms; t++) {
        double v = 0.0;
        for (int b = threadIdx.x; b < nb; b += 256) v += part[((long)sol * nb + b) * kEnergyTerms + t];
        v = block_sum256(v, red);
        if (threadIdx.x == 0) {
            const double e = v * ms0T0 * w[t];
            out[(long)sol * 5 + t] = e;
            tot += e;
        }
    }
    if (threadIdx.x == 0) out[(long)sol * 5 + 4] = tot;
}

// One fixed-size RK4 step of every solve (4 demag evaluations, 4 stage kernels).
void rk4_step(hs_solver* s, StageIO io) {
    demag_into(s, s->mA, s->B0);
    io.m = s->mA; io.B = s->B0; io.mout = s->mB;
    launch_stage<R_RK4_1>(s, io);
    demag_into(s, s->mB, s->Bs);
    io.m = s->mB; io.B = s->Bs; io.mout = s->mC;
    launch_stage<R_RK4_2>(s, io);
    demag_into(s, s->mC, s->Bs);
    io.m = s->mC; io.B = s->Bs; io.mout = s->mB;
    launch_stage<R_RK4_3>(s, io);
    demag_into(s, s->mB, s->Bs);
    io.m = s->mB; io.B = s->Bs; io.mout = s->mA;
    launch_stage<R_RK4_4>(s, io);
}

// The launches of one RK4 step of the whole batch with step/dt read from the
// device clocks, followed by the clock advance (graph body).
void rk4_graph_step(hs_solver* s) {
    StageIO io = base_io(s, 0.0, 0, 0);
    io.clk = s->clk;
    rk4_step(s, io);
    hipLaunchKernelGGL(k_clock_step, hs::div_up(s->nsolve, 256), 256, 0, s->st, s->clk, s->nsolve);
}

// Every solve running at step counter step0 with dt (fixed-step replay).
void set_clocks(hs_solver* s, long step0, double dt) {
    std::vector<Clock> hc(s->nsolve);
    for (auto& c : hc) {
        memset(&c, 0, sizeof(c));
        c.step = (uint32_t)step0;
        c.h = dt;
        c.state = HS_SLOT_RUNNING;
    }
    hipMemcpyAsync(s->clk, hc.data(), sizeof(Clock) * s->nsolve, hipMemcpyHostToDevice, s->st);
    hipStreamSynchronize(s->st);
}

// Buffer set `par` of slot mode: (FP64 state, FP32 stage input, its field).
struct Set {
    double* y;
    float* m;
    float* B;
};
Set slot_set(hs_solver* s, int par) {
    return par ? Set{s->acc, s->mB, s->Bs} : Set{s->y0, s->mA, s->B0};
}

// One RK23 attempt of every running slot with the state in set par: the
// stages leave the candidate in set par^1, k_control decides per slot and
// k_fixup restores the rejected slots' state into set par^1.
template <int ROLE>
void launch_tm(hs_solver* s, const double* Te, const double* Tp, int par, double* Teo, double* Tpo, float* Tfo) {
    TmIO t{};
    t.Te = Te;
    t.Tp = Tp;
    t.Te0 = s->Te[par];
    t.Tp0 = s->Tp[par];
    t.Teo = Teo;
    t.Tpo = Tpo;
    t.Tfo = Tfo;
    t.k1e = s->k1e, t.k1p = s->k1p, t.k2e = s->k2e, t.k2p = s->k2p, t.ee = s->ee, t.ep = s->ep;
    t.errmax = s->errmax;
    t.clk = s->clk;
    t.sp = s->sp;
    t.qs = s->qs;
    hipLaunchKernelGGL(k_tm<ROLE>, dim3(hs::grid1d(s->N), s->nsolve), 256, 0, s->st, s->phys, s->tmc, t);
}

void slot_iteration(hs_solver* s, int par) {
    const Set cur = slot_set(s, par), nxt = slot_set(s, par ^ 1);
    const int q = par ^ 1;
    StageIO io = base_io(s, 0.0, 0, 0);
    io.clk = s->clk;
    io.y0 = cur.y;
    io.acc = nxt.y;
    io.m = cur.m; io.B = cur.B; io.mout = nxt.m;
    io.qs = s->qs;  // spin heating: each LLB stage leaves the spin term for the k_tm of its stage
    if (s->tm) io.T = s->Tf[par];
    launch_stage<R_BS_1>(s, io);
    if (s->tm) launch_tm<R_BS_1>(s, s->Te[par], s->Tp[par], par, s->Tea, s->Tpa, s->Tfa);
    demag_into(s, nxt.m, nxt.B, s->live);
    io.m = nxt.m; io.B = nxt.B; io.mout = s->mC;
    if (s->tm) io.T = s->Tfa;
    launch_stage<R_BS_2>(s, io);
    if (s->tm) launch_tm<R_BS_2>(s, s->Tea, s->Tpa, par, s->Teb, s->Tpb, s->Tfb);
    demag_into(s, s->mC, nxt.B, s->live);
    io.m = s->mC; io.B = nxt.B; io.mout = nxt.m;
    if (s->tm) io.T = s->Tfb;
    launch_stage<R_BS_3>(s, io);
    if (s->tm) launch_tm<R_BS_3>(s, s->Teb, s->Tpb, par, s->Te[q], s->Tp[q], s->Tf[q]);
    demag_into(s, nxt.m, nxt.B, s->live);
    io.m = nxt.m; io.B = nxt.B; io.mout = nullptr;
    if (s->tm) io.T = s->Tf[q];
    launch_stage<R_BS_4>(s, io);
    if (s->tm) launch_tm<R_BS_4>(s, s->Te[q], s->Tp[q], par, nullptr, nullptr, nullptr);
    hipLaunchKernelGGL(k_control, hs::div_up(s->nsolve, 256), 256, 0, s->st, s->clk, s->errmax, s->live,
                       s->nsolve, (uint32_t)par, s->tm ? 3 : 1);
    hipLaunchKernelGGL(k_fixup, dim3(hs::grid1d(s->N), s->nsolve), hs::kBlock, 0, s->st, s->clk, cur.y, cur.m,
                       cur.B, nxt.y, nxt.m, nxt.B, s->N, s->phys.demag);
    if (s->tm)
        hipLaunchKernelGGL(k_fixup_tm, dim3(hs::grid1d(s->N), s->nsolve), hs::kBlock, 0, s->st, s->clk, s->Te[par],
                           s->Tp[par], s->Tf[par], s->Te[q], s->Tp[q], s->Tf[q], s->N);
}

bool capture(hs_solver* s, hipGraphExec_t* out, const char* who, void (*body)(hs_solver*, int), int arg) {
    bool prof = hs::prof_on();
    if (prof) hs_prof_enable(0);
    hipGraph_t g;
    bool ok = hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (ok) {
        body(s, arg);
        ok = hipStreamEndCapture(s->st, &g) == hipSuccess;
        if (ok) {
            ok = hipGraphInstantiate(out, g, nullptr, nullptr, 0) == hipSuccess;
            hipGraphDestroy(g);
        }
    }
    if (prof) hs_prof_enable(1);
    if (!ok) {
        *out = nullptr;
        hs::set_error(HS_EINVAL, who);
    }
    return ok;
}

// Captured graphs bake in buffer addresses: drop them when the host swaps
// the roles of the state buffers (Euler, RK12 and Heun steps).
void drop_graphs(hs_solver* s) {
    if (s->graph) hipGraphExecDestroy(s->graph);
    s->graph = nullptr;
    for (int v = 0; v < 2; v++) {
        if (s->slot_graph[v]) hipGraphExecDestroy(s->slot_graph[v]);
        s->slot_graph[v] = nullptr;
    }
}

// Buffers of device-controlled stepping (slots, and RK23 of hs_solver_run):
// pinned slot records, live flags, completion event.
bool ensure_devctl(hs_solver* s) {
    if (s->slot_host) return true;
    if (hipHostMalloc((void**)&s->slot_host, sizeof(hs_slot) * s->nsolve) != hipSuccess ||
        hipMalloc(&s->live, sizeof(int) * s->nsolve) != hipSuccess ||
        hipEventCreateWithFlags(&s->slot_ev, hipEventDisableTiming) != hipSuccess) {
        hs::set_error(HS_ENOMEM, "device-controlled stepping: out of memory");
        return false;
    }
    memset(s->slot_host, 0, sizeof(hs_slot) * s->nsolve);
    hipMemsetAsync(s->live, 0, sizeof(int) * s->nsolve, s->st);
    return true;
}

bool ensure_slot_graphs(hs_solver* s) {
    for (int v = 0; v < 2; v++)
        if (!s->slot_graph[v] &&
            !capture(s, &s->slot_graph[v], "device-controlled stepping: graph capture failed", slot_iteration, v))
            return false;
    return true;
}

// After device-controlled stepping: solves whose state ended in buffer set 1
// copy it (state, stage input, its field) to set 0.
__global__ void k_gather(const Clock* clk, const double* y1, const float* m1, const float* B1, double* y0,
                         float* m0, float* B0, long N, int demag) {
    const int q = blockIdx.y;
    if (clk[q].home != 1u) return;
    long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= N) return;
    const long b = (long)q * 3 * N;
    for (int c = 0; c < 3; c++) {
        y0[b + c * N + I] = y1[b + c * N + I];
        m0[b + c * N + I] = m1[b + c * N + I];
        if (demag) B0[b + c * N + I] = B1[b + c * N + I];
    }
}

// hs_solver_run's RK23: every solve of the batch takes nsteps accepted steps
// with its own dt, accept/reject decided on the device (k_control), the
// iterations launched as captured graphs; the host only checks completion.
int run_rk23_device(hs_solver* s, long nsteps, long step0, double* dtio, double maxErr, double mindt,
                    double maxdt, hs_run_stats* stats) {
    hs_run_stats st{};
    if (nsteps == 0) {
        if (stats) *stats = st;
        return hs_last_error();
    }
    if (!ensure_devctl(s)) return HS_ENOMEM;
    ensure_b0(s);
    std::vector<Clock> hc(s->nsolve);
    for (auto& c : hc) {
        memset(&c, 0, sizeof(c));
        c.step = (uint32_t)step0;
        c.h = *dtio;
        c.maxErr = maxErr;
        c.mindt = mindt;
        c.maxdt = maxdt;
        c.end = (uint32_t)(step0 + nsteps);
        c.state = HS_SLOT_RUNNING;
        c.maxErrT = s->tm ? s->tmp.maxErrT : 0.0;
    }
    if (s->tm) {  // laser clock of this run
        for (auto& q : s->sp_host) q.t_origin = s->tm_time;
        hipMemcpyAsync(s->sp, s->sp_host.data(), sizeof(SolveP) * s->nsolve, hipMemcpyHostToDevice, s->st);
    }
    std::vector<int> lv(s->nsolve, 1);
    hipMemcpyAsync(s->clk, hc.data(), sizeof(Clock) * s->nsolve, hipMemcpyHostToDevice, s->st);
    hipMemcpyAsync(s->live, lv.data(), sizeof(int) * s->nsolve, hipMemcpyHostToDevice, s->st);
    zero_err(s);
    hipStreamSynchronize(s->st);
    if (!ensure_slot_graphs(s)) return HS_EINVAL;
    s->slot_iter = 0;
    long todo = nsteps, iters = 0;
    const long cap = nsteps * (kMaxTry + 2) + 2;  // every solve has ended by then
    for (;;) {
        // with the event profiler on, launch the same kernels eagerly so each
        // launch is bracketed by events (graph replay carries none)
        const bool eager = hs::prof_on();
        for (long n = 0; n < todo; n++) {
            if (eager)
                slot_iteration(s, (int)(s->slot_iter & 1));
            else
                hipGraphLaunch(s->slot_graph[s->slot_iter & 1], s->st);
            s->slot_iter++;
        }
        iters += todo;
        hipMemcpyAsync(s->slot_host, s->clk, sizeof(Clock) * s->nsolve, hipMemcpyDeviceToHost, s->st);
        hipStreamSynchronize(s->st);
        long left = 0;
        for (int q = 0; q < s->nsolve; q++)
            if (s->slot_host[q].state == HS_SLOT_RUNNING)
                left = std::max(left, (long)(s->slot_host[q].end - s->slot_host[q].step));
        if (left == 0 || iters >= cap) break;
        todo = std::max(left, 2L);
    }
    hipLaunchKernelGGL(k_gather, dim3(hs::grid1d(s->N), s->nsolve), hs::kBlock, 0, s->st, s->clk, s->acc, s->mB,
                       s->Bs, s->y0, s->mA, s->B0, s->N, s->phys.demag);
    if (s->tm)
        hipLaunchKernelGGL(k_gather_tm, dim3(hs::grid1d(s->N), s->nsolve), hs::kBlock, 0, s->st, s->clk, s->Te[1],
                           s->Tp[1], s->Tf[1], s->Te[0], s->Tp[0], s->Tf[0], s->N);
    hipStreamSynchronize(s->st);
    s->slot_iter = 0;
    s->b0_valid = s->phys.demag != 0;
    bool failed = false;
    double dtmin = 0.0;
    for (int q = 0; q < s->nsolve; q++) {
        const Clock& c = s->slot_host[q];
        failed |= c.state != HS_SLOT_DONE;
        st.steps = std::max(st.steps, (long)c.accepted);
        st.rejected += c.rejected;
        st.rhs_evals += 3L * (c.accepted + c.rejected);
        if (q == 0 || c.h < dtmin) dtmin = c.h;
    }
    const Clock& c0 = s->slot_host[0];
    st.t = c0.t;
    if (s->tm) s->tm_time += c0.t;
    st.last_err = c0.err;
    st.dt = dtmin;
    *dtio = dtmin;
    if (stats) *stats = st;
    HS_LAUNCH_CHECK("hs_solver_run");
    if (failed) {
        hs::set_error(HS_ECONV, "hs_solver_run: RK23 did not converge (or diverged)");
        return HS_ECONV;
    }
    return hs_last_error();
}

// Material constants of the uniform longitudinal field (see LfConst).
LfConst lf_constants(const hs_llb_params* p) {
    LfConst C{};
    C.uniform = p->longfield && !p->J && !p->n && !p->Tc && !p->T;
    const double S = p->JMul;
    if (S > HS_INFSPIN) {
        C.langevin = 1;
        C.lim = HS_LINRANGE;
        C.a1 = 1.0 / 3.0, C.a3 = 1.0 / 45.0, C.a5 = 2.0 / 945.0, C.a7 = 1.0 / 4725.0;
        C.d0 = 1.0 / 3.0, C.d2 = 1.0 / 15.0, C.d4 = 2.0 / 189.0, C.d6 = 1.0 / 675.0;
        return C;
    }
    const double l = 1.0 / (2.0 * S), g = (2.0 * S + 1.0) * l;
    const double tp = 2.0 * S + 1.0;
    C.gpow = (tp == floor(tp) && tp >= 1.0 && tp <= 16.0) ? (int)tp : 0;
    const double l2 = l * l, g2 = g * g, l4 = l2 * l2, g4 = g2 * g2, l6 = l4 * l2, g6 = g4 * g2;
    const double l8 = l4 * l4, g8 = g4 * g4;
    C.l = l, C.g = g;
    C.lim = fmin(HS_LINRANGE / g, HS_LINRANGE / l);
    C.a1 = (g2 - l2) / 3.0, C.a3 = (g4 - l4) / 45.0, C.a5 = (g6 - l6) * 2.0 / 945.0, C.a7 = (g8 - l8) / 4725.0;
    C.d0 = (g2 - l2) / 3.0, C.d2 = (g4 - l4) / 15.0, C.d4 = (g6 - l6) * 2.0 / 189.0, C.d6 = (g8 - l8) / 675.0;
    return C;
}

// Per-solve constants of a solve at temperature multiplier T.
void solve_constants(const hs_llb_params* p, double T, SolveP& q) {
    const double V = p->cell[0] * p->cell[1] * p->cell[2];
    const double S = p->JMul;
    const double preS = (S > HS_INFSPIN) ? 1.0 : S / (S + 1.0);
    q.TMul = (float)T;
    q.thC = (HS_KB * 2.0 * T) / (HS_MU0 * V * p->gammaLL * p->msat0T0Mul);
    q.sthC = (float)sqrt(q.thC);
    q.lf_b = 3.0 * preS * p->TcMul / T;
    q.lf_c1 = p->nMul * (HS_KB / HS_MU0) * T / p->msat0T0Mul;
}
}  // namespace

// A batch of `nsolve` solves on one mesh: per-solve TMul, seed, B_ext.
extern "C" hs_solver* hs_batch_create(const hs_llb_params* p, int nsolve, const double* TMul, const uint64_t* seed,
                                      const double* Bext, CUstream stream) {
    if (!p || p->N[0] < 1 || p->N[1] < 1 || p->N[2] < 1 || nsolve < 1) {
        hs::set_error(HS_EINVAL, "hs_batch_create: bad mesh or batch size");
        return nullptr;
    }
    hs_solver* s = new hs_solver();
    s->p = *p;
    s->nsolve = nsolve;
    s->st = hs::S(stream);
    if (!s->st) {  // own a real stream so a step can be captured into a graph; non-blocking, so
                   // that no copy on the legacy null stream waits for it (or it for them)
        hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking);
        s->own_stream = true;
    }
    s->N = (long)p->N[0] * p->N[1] * p->N[2];
    const long NN = s->N * nsolve;
    hipMalloc(&s->y0, sizeof(double) * 3 * NN);
    hipMalloc(&s->acc, sizeof(double) * 3 * NN);
    hipMalloc(&s->mA, sizeof(float) * 3 * NN);
    hipMalloc(&s->mB, sizeof(float) * 3 * NN);
    hipMalloc(&s->mC, sizeof(float) * 3 * NN);
    hipMalloc(&s->B0, sizeof(float) * 3 * NN);
    hipMalloc(&s->Bs, sizeof(float) * 3 * NN);
    hipMalloc(&s->e, sizeof(float) * 3 * NN);
    hipMalloc(&s->k1, sizeof(float) * 3 * NN);
    hipMalloc(&s->k2, sizeof(float) * 3 * NN);
    hipMalloc(&s->k3, sizeof(float) * 3 * NN);
    hipMalloc(&s->errmax, sizeof(unsigned long long) * nsolve * kErrPerSolve);
    hipMalloc(&s->sp, sizeof(SolveP) * nsolve);
    hipMalloc(&s->clk, sizeof(Clock) * nsolve);
    if (!s->y0 || !s->acc || !s->mA || !s->mB || !s->mC || !s->B0 || !s->Bs || !s->e || !s->k1 || !s->k2 ||
        !s->k3 || !s->errmax || !s->sp || !s->clk) {
        hs::set_error(HS_ENOMEM, "hs_batch_create: out of device memory");
        hs_solver_destroy(s);
        return nullptr;
    }
    hipMemsetAsync(s->B0, 0, sizeof(float) * 3 * NN, s->st);
    hipMemsetAsync(s->Bs, 0, sizeof(float) * 3 * NN, s->st);
    hipMemsetAsync(s->clk, 0, sizeof(Clock) * nsolve, s->st);
    hipMemsetAsync(s->errmax, 0, sizeof(unsigned long long) * nsolve * kErrPerSolve, s->st);
    // per-solve parameters; thermal K = thC / max(cutoff_dt, dt), mod_brown.go:125-145
    std::vector<SolveP> hsp(nsolve);
    for (int q = 0; q < nsolve; q++) {
        double T = TMul ? TMul[q] : p->TMul;
        solve_constants(p, T, hsp[q]);
        for (int c = 0; c < 3; c++) {
            hsp[q].bextT[c] = Bext ? Bext[3 * q + c] : p->Bext[c];
            hsp[q].bext[c] = (float)(1.0 / HS_MU0 * hsp[q].bextT[c]);
        }
        hsp[q].seed = seed ? seed[q] : p->seed;
        hsp[q].stream = p->stream + (uint32_t)(seed ? 0 : q);
    }
    hipMemcpy(s->sp, hsp.data(), sizeof(SolveP) * nsolve, hipMemcpyHostToDevice);
    s->sp_host = hsp;
    Phys& P = s->phys;
    P.N0 = p->N[0];
    P.N1 = p->N[1];
    P.N2 = p->N[2];
    P.w0 = p->pbc[0] != 0;
    P.w1 = p->pbc[1] != 0;
    P.w2 = p->pbc[2] != 0;
    P.nsolve = nsolve;
    P.sp = s->sp;
    P.cutoff_dt = (float)p->cutoff_dt;
    {
        // enough blocks to fill 256 CUs several times over: split the planes into chunks
        long blocks2d = (long)hs::div_up(P.N2, TX) * hs::div_up(P.N1, TY) * nsolve;
        long chunks = (4096 + blocks2d - 1) / blocks2d;
        if (chunks < 1) chunks = 1;
        if (chunks > P.N0) chunks = P.N0;
        P.chunk = (int)((P.N0 + chunks - 1) / chunks);
    }
    P.exchange = p->exchange;
    P.demag = p->demag;
    P.zeeman = p->zeeman;
    P.longfield = p->longfield;
    P.thermal = p->thermal;
    P.torque = p->torque;
    P.local = p->local00nc;
    P.normalize = p->normalize;
    double lex2 = p->lexMul * p->lexMul;
    P.ex0 = (float)(lex2 / (p->cell[0] * p->cell[0]));
    P.ex1 = (float)(lex2 / (p->cell[1] * p->cell[1]));
    P.ex2 = (float)(lex2 / (p->cell[2] * p->cell[2]));
    P.ms0T0Mul = (float)p->msat0T0Mul;
    P.lex = upload_mask(s, p->lex);
    P.msk = upload_mask(s, p->msat0T0);
    P.invMu0 = (float)(1.0 / HS_MU0);
    P.JMul = (float)p->JMul;
    P.nMul = (float)p->nMul;
    P.TcMul = (float)p->TcMul;
    P.J = upload_mask(s, p->J);
    P.n = upload_mask(s, p->n);
    P.Tc = upload_mask(s, p->Tc);
    P.T = upload_mask(s, p->T);
    P.gamma = (float)p->gammaLL;
    {
        const LfConst C = lf_constants(p);
        P.lf_uniform = C.uniform;
        hipMalloc(&s->lfc, sizeof(LfConst));
        hipMemcpy(s->lfc, &C, sizeof(LfConst), hipMemcpyHostToDevice);
        P.lfc = s->lfc;
    }
    for (int c = 0; c < 3; c++) {
        P.aMul[c] = (float)p->alphaMul[c];
        P.sqa[c] = sqrtf(P.aMul[c]);
        P.alpha[c] = upload_mask(s, p->alpha[c]);
    }
    if (p->demag) {
        s->demag = p->demag_kernel == 1
                       ? hs_demag_create_newell(p->N, p->cell, p->pbc, (CUstream)s->st)
                       : hs_demag_create(p->N, p->cell, p->pbc, p->demag_accuracy > 0 ? p->demag_accuracy : 6,
                                         (CUstream)s->st);
        if (!s->demag || hs_demag_reserve(s->demag, nsolve) != 0) {
            hs_solver_destroy(s);
            return nullptr;
        }
    }
    {
        const char* env = getenv("HS_STAGE_LEAN");
        s->lean_ok = !(env && !strcmp(env, "0"));
    }
    tune_stage(s);
    hipStreamSynchronize(s->st);
    HS_LAUNCH_CHECK("hs_batch_create");
    if (hs_last_error()) {
        hs_solver_destroy(s);
        return nullptr;
    }
    return s;
}

extern "C" hs_solver* hs_solver_create(const hs_llb_params* p, CUstream stream) {
    return hs_batch_create(p, 1, nullptr, nullptr, nullptr, stream);
}

extern "C" void hs_solver_destroy(hs_solver* s) {
    if (!s) return;
    if (s->graph) hipGraphExecDestroy(s->graph);
    for (int v = 0; v < 2; v++)
        if (s->slot_graph[v]) hipGraphExecDestroy(s->slot_graph[v]);
    if (s->slot_ev) hipEventDestroy(s->slot_ev);
    if (s->slot_host) hipHostFree(s->slot_host);
    if (s->imp_host) hipHostFree(s->imp_host);
    if (s->err_host) hipHostFree(s->err_host);
    if (s->demag) hs_demag_destroy(s->demag);
    void* ps[] = {s->y0, s->acc, s->mA, s->mB, s->mC, s->B0, s->Bs, s->e, s->k1, s->k2, s->k3, s->kout, s->hout,
                  s->errmax, s->sp, s->lfc, s->clk, s->live, s->imp, s->y1, s->qs, s->epart, s->err_dev};
    for (void* q : ps)
        if (q) hipFree(q);
    for (float* q : s->masks) hipFree(q);
    void* tmb[] = {s->Te[0], s->Te[1], s->Tp[0], s->Tp[1], s->Tf[0], s->Tf[1], s->Tea, s->Tpa, s->Teb, s->Tpb,
                   s->Tfa, s->Tfb, s->k1e, s->k1p, s->k2e, s->k2p, s->ee, s->ep};
    for (void* q : tmb)
        if (q) hipFree(q);
    if (s->own_stream) hipStreamDestroy(s->st);
    delete s;
}

extern "C" CUstream hs_solver_stream(hs_solver* s) { return s ? (CUstream)s->st : nullptr; }
extern "C" hs_demag* hs_solver_demag(hs_solver* s) { return s ? s->demag : nullptr; }
extern "C" int hs_solver_stage_variant(hs_solver* s) { return s ? s->stage_kind : -1; }
extern "C" int hs_solver_stage_lean(hs_solver* s) { return s ? (int)(!s->tm && stage_lean(s, StageIO{})) : -1; }
extern "C" int hs_solver_nsolve(hs_solver* s) { return s ? s->nsolve : -1; }

static bool whole_batch_ok(hs_solver* s, const char* who) {
    if (s->slot_mode) hs::set_error(HS_EINVAL, who);
    return !s->slot_mode;
}

extern "C" int hs_solver_set_m(hs_solver* s, const double* m) {
    if (!s || !m) return HS_EINVAL;
    if (!whole_batch_ok(s, "hs_solver_set_m: handle is in slot mode")) return HS_EINVAL;
    hipMemcpyAsync(s->y0, m, sizeof(double) * 3 * s->N * s->nsolve, hipMemcpyHostToDevice, s->st);
    hipLaunchKernelGGL(k_set_state, dim3(hs::grid1d(s->N), s->nsolve), hs::kBlock, 0, s->st, s->y0, s->mA,
                       s->p.normalize, s->N);
    s->b0_valid = false;
    hipStreamSynchronize(s->st);
    HS_LAUNCH_CHECK("hs_solver_set_m");
    return hs_last_error();
}

extern "C" int hs_solver_get_m(hs_solver* s, double* m) {
    if (!s || !m) return HS_EINVAL;
    if (!whole_batch_ok(s, "hs_solver_get_m: handle is in slot mode")) return HS_EINVAL;
    hipMemcpyAsync(m, s->y0, sizeof(double) * 3 * s->N * s->nsolve, hipMemcpyDeviceToHost, s->st);
    hipStreamSynchronize(s->st);
    HS_LAUNCH_CHECK("hs_solver_get_m");
    return hs_last_error();
}

extern "C" int hs_solver_energy(hs_solver* s, double* E) {
    if (!s || !E) {
        hs::set_error(HS_EINVAL, "hs_solver_energy: null argument");
        return HS_EINVAL;
    }
    if (!whole_batch_ok(s, "hs_solver_energy: handle is in slot mode")) return HS_EINVAL;
    const int nb = (int)hs::div_up(s->N, 256);
    if (!s->epart &&
        hipMalloc(&s->epart, sizeof(double) * ((size_t)nb * s->nsolve * kEnergyTerms + 5 * s->nsolve)) != hipSuccess) {
        s->epart = nullptr;
        hs::set_error(HS_ENOMEM, "hs_solver_energy: out of device memory");
        return HS_ENOMEM;
    }
    if (s->phys.demag) ensure_b0(s);
    const hs_llb_params& p = s->p;
    EnergyC c;
    c.ms0T0 = p.msat0T0Mul;
    for (int d = 0; d < 3; d++) c.ex[d] = p.lexMul * p.lexMul / (p.cell[d] * p.cell[d]);
    c.JMul = p.JMul;
    c.nMul = p.nMul;
    c.TcMul = p.TcMul;
    double* out = s->epart + (size_t)nb * s->nsolve * kEnergyTerms;
    hipLaunchKernelGGL(k_energy, dim3(nb, s->nsolve), 256, 0, s->st, s->phys, c, s->y0, s->B0,
                       s->tm ? s->Tf[0] : nullptr, s->epart, s->N);
    hipLaunchKernelGGL(k_energy_sum, dim3(s->nsolve), 256, 0, s->st, s->epart, nb, p.msat0T0Mul,
                       p.cell[0] * p.cell[1] * p.cell[2], out);
    hipMemcpyAsync(E, out, sizeof(double) * 5 * s->nsolve, hipMemcpyDeviceToHost, s->st);
    hipStreamSynchronize(s->st);
    HS_LAUNCH_CHECK("hs_solver_energy");
    return hs_last_error();
}

extern "C" int hs_solver_rhs(hs_solver* s, double* k_host, double* h_host, uint32_t step, uint32_t attempt,
                             double dt) {
    if (!s) return HS_EINVAL;
    const long NN = s->N * s->nsolve;
    if (!s->kout && (hipMalloc(&s->kout, sizeof(float) * 3 * NN) != hipSuccess ||
                     hipMalloc(&s->hout, sizeof(float) * 3 * NN) != hipSuccess)) {
        hs::set_error(HS_ENOMEM, "hs_solver_rhs: out of device memory");
        return HS_ENOMEM;
    }
    ensure_b0(s);
    StageIO io = base_io(s, dt, step, attempt);
    io.m = s->mA;
    io.B = s->B0;
    io.kout = s->kout;
    io.hout = s->hout;
    if (s->tm) io.T = s->Tf[0];
    launch_stage<R_RHS>(s, io);
    std::vector<float> tmp(3 * NN);
    if (k_host) {
        hipMemcpyAsync(tmp.data(), s->kout, sizeof(float) * 3 * NN, hipMemcpyDeviceToHost, s->st);
        hipStreamSynchronize(s->st);
        for (long i = 0; i < 3 * NN; i++) k_host[i] = tmp[i];
    }
    if (h_host) {
        hipMemcpyAsync(tmp.data(), s->hout, sizeof(float) * 3 * NN, hipMemcpyDeviceToHost, s->st);
        hipStreamSynchronize(s->st);
        for (long i = 0; i < 3 * NN; i++) h_host[i] = tmp[i];
    }
    HS_LAUNCH_CHECK("hs_solver_rhs");
    return hs_last_error();
}

// Fixed-step RK4 replayed from a captured hipGraph: one step (4 demag
// evaluations, 4 stage kernels, clock advance) is captured once and
// launched nsteps times; the thermal-noise step counter lives on the device.
extern "C" int hs_solver_run_graph(hs_solver* s, long nsteps, long step0, double dt) {
    if (!s || nsteps < 0) return HS_EINVAL;
    if (s->slot_mode) {
        hs::set_error(HS_EINVAL, "hs_solver_run_graph: handle is in slot mode");
        return HS_EINVAL;
    }
    if (s->graph && (s->graph_dt != dt)) {
        hipGraphExecDestroy(s->graph);
        s->graph = nullptr;
    }
    set_clocks(s, step0, dt);
    if (!s->graph) {
        if (!capture(s, &s->graph, "hs_solver_run_graph: graph capture failed",
                     [](hs_solver* q, int) { rk4_graph_step(q); }, 0))
            return HS_EINVAL;
        s->graph_dt = dt;
    }
    for (long n = 0; n < nsteps; n++) hipGraphLaunch(s->graph, s->st);
    s->b0_valid = false;
    hipStreamSynchronize(s->st);
    HS_LAUNCH_CHECK("hs_solver_run_graph");
    return hs_last_error();
}

// ---------------------------------------------------------------- implicit drivers
namespace {
struct ImpRed {
    double d[3], a[3], d1[3], sum;
    bool bad;
};

bool ensure_imp(hs_solver* s) {
    // region 0: the m maxima (kImpVals), regions 1, 2: Te, Tp (kImp1Vals each)
    // followed by kImpAll folded doubles
    if (!s->imp && hipMalloc(&s->imp, sizeof(unsigned long long) * (kImpAll * kImpLanes * kImpStride + kImpAll)) !=
                       hipSuccess)
        s->imp = nullptr;
    if (!s->imp_host && hipHostMalloc((void**)&s->imp_host, sizeof(double) * kImpAll) != hipSuccess)
        s->imp_host = nullptr;
    if (s->imp && !s->imp_zeroed) {  // the lanes start at zero; each fold leaves them so
        hipMemsetAsync(s->imp, 0, sizeof(unsigned long long) * kImpAll * kImpLanes * kImpStride, s->st);
        s->imp_zeroed = true;
    }
    if (!s->y1 && hipMalloc(&s->y1, sizeof(double) * 3 * s->N * s->nsolve) != hipSuccess) s->y1 = nullptr;
    if (!s->imp || !s->y1 || !s->imp_host) hs::set_error(HS_ENOMEM, "implicit driver: out of memory");
    return s->imp && s->y1 && s->imp_host;
}

// LLB RHS of the iterate (FP32 stage input mB, demag first) into k_out; T:
// the FP32 temperature of a two-temperature solve's iterate
void imp_rhs(hs_solver* s, float* k_out, uint32_t step, uint32_t attempt, double dt, const float* T,
             float* qs = nullptr) {
    demag_into(s, s->mB, s->Bs);
    StageIO io = base_io(s, dt, step, attempt);
    io.m = s->mB;
    io.B = s->Bs;
    io.kout = k_out;
    io.hout = nullptr;
    io.mout = nullptr;
    io.T = T;
    io.qs = qs;
    launch_stage<R_RHS>(s, io);
}

// One fixed-point update of every solve: the RHS of the iterate (FP32 stage
// input mB, demag first) into k_out, unless k_in already holds the slope,
// then k_imp; the maxima come back to the host (one sync, as the reference's
// reductor does per update).
ImpRed imp_update(hs_solver* s, const float* k_in, float* k_out, const float* dy0, const double* y1, double c,
                  bool norm_in, uint32_t step, uint32_t attempt, double dt) {
    if (!k_in) {
        imp_rhs(s, k_out, step, attempt, dt, nullptr);
        k_in = k_out;
    }
    hipLaunchKernelGGL(k_imp, dim3(hs::grid1d(s->N), s->nsolve), hs::kBlock, 0, s->st, s->y0, s->acc, k_in, dy0,
                       y1, s->mB, c, norm_in ? s->p.normalize : 0, s->p.normalize, s->imp, s->N);
    double* fold = (double*)(s->imp + (long)kImpAll * kImpLanes * kImpStride);
    hipLaunchKernelGGL(k_imp_fold, 1, 64, 0, s->st, s->imp, kImpVals, fold);
    hipMemcpyAsync(s->imp_host, fold, sizeof(double) * kImpVals, hipMemcpyDeviceToHost, s->st);
    hipStreamSynchronize(s->st);
    const double* v = s->imp_host;
    ImpRed r;
    for (int p = 0; p < 3; p++) {
        r.d[p] = v[p];
        r.a[p] = v[3 + p];
        r.d1[p] = v[6 + p];
    }
    r.sum = v[9];
    r.bad = v[10] != 0.0;
    return r;
}

// solve_am01.go:122-132: (1/sqrt(3)) sqrt(sum_p (maxdiff_p / (absErr + maxabs_p relErr))^2)
double am_norm(const double d[3], const double a[3], double absErr, double relErr) {
    double t = 0.0;
    for (int p = 0; p < 3; p++) t += pow(d[p] / (absErr + a[p] * relErr), 2.0);
    return (1.0 / sqrt(3.0)) * sqrt(t);
}

int imp_fail(hs_solver* s, const char* what, double dt, double* dtio, hs_run_stats* stats, const hs_run_stats& st) {
    hs::set_error(HS_ECONV, what);
    *dtio = dt;
    if (stats) *stats = st;
    return HS_ECONV;
}

// Explicit-Euler start of every step: dy0 = f(y0) into e (y0 normalised in
// place with the normaliser), acc = y0 + dt dy0, mB its stage input.
void implicit_start(hs_solver* s, double dt, uint32_t step, const float* T = nullptr, float* qs = nullptr) {
    ensure_b0(s);
    StageIO io = base_io(s, dt, step, 0);
    io.m = s->mA;
    io.B = s->B0;
    io.mout = s->mB;
    io.T = T;
    io.qs = qs;
    launch_stage<R_RK12_1>(s, io);
}

void implicit_accept(hs_solver* s) {
    std::swap(s->y0, s->acc);
    std::swap(s->mA, s->mB);
    drop_graphs(s);
    s->b0_valid = false;
}

// BDFAM12.Step, engine/solve_am01.go:45-300, one equation (m) for the whole
// batch: the norms take the maxima over every solve.  Noise attempt a counts
// the dt changes within a step.
int run_am12(hs_solver* s, long nsteps, long step0, double* dtio, double mindt, double maxdt, hs_run_stats* stats) {
    const int maxTry = 10, maxIter = 5;
    const double headRoom = 0.8, maxIterErr = 0.1, alpha_ref = 0.6;
    const double ab = s->imp_abs, rl = s->imp_rel;
    hs_run_stats st{};
    double dt = *dtio;
    for (long n = 0; n < nsteps; n++) {
        const uint32_t step = (uint32_t)(step0 + n);
        uint32_t a = 0;
        implicit_start(s, dt, step);
        st.rhs_evals++;
        int tr = 0;
        bool restrict_step = false, first = true;
        double saved_alpha = 0.0, tErr = 0.0;
        for (;;) {
            const double h = dt;
            if (!first)
                hipLaunchKernelGGL(k_rk12_predict, dim3(hs::grid1d(s->N), s->nsolve), hs::kBlock, 0, s->st, s->y0,
                                   s->e, s->acc, s->mB, h, s->p.normalize, s->N);
            first = false;
            st.iterations++;
            // predictor: implicit Euler, y <- y0 + h f(y); the first slope f(y_EE) kept in k1
            double er = maxIterErr, alpha = 0.0;
            bool badIt = false;
            ImpRed r{};
            for (int iter = 0;;) {
                r = imp_update(s, nullptr, iter == 0 ? s->k1 : s->k2, nullptr, nullptr, h, true, step, a, dt);
                st.rhs_evals++;
                if (r.bad) return imp_fail(s, "hs_solver_run: AM12 iterate is not finite", dt, dtio, stats, st);
                const double t = am_norm(r.d, r.a, ab, rl);
                alpha = t / er;
                saved_alpha = alpha;
                er = t;
                iter++;
                st.iterations++;
                if (er <= maxIterErr) break;
                if (alpha >= 1.0 || iter > maxIter) {
                    badIt = true;
                    break;
                }
            }
            if (badIt && tr == maxTry - 1)
                return imp_fail(s, "hs_solver_run: AM12 predictor iteration did not converge", dt, dtio, stats, st);
            if (badIt) {
                dt = (alpha > alpha_ref) ? dt * pow(alpha_ref / alpha, 0.5) : 0.5 * dt;
                restrict_step = true;
                a++;
                st.rejected++;
                continue;
            }
            double a1[3] = {r.a[0], r.a[1], r.a[2]};  // max|y1_p|
            hipMemcpyAsync(s->y1, s->acc, sizeof(double) * 3 * s->N * s->nsolve, hipMemcpyDeviceToDevice, s->st);
            // corrector: trapezoidal, y <- y0 + h/2 (f(y) + dy0), starting from the slope f(y_EE)
            er = maxIterErr;
            for (int iter = 0;;) {
                r = imp_update(s, iter == 0 ? s->k1 : nullptr, s->k2, s->e, s->y1, 0.5 * h, iter > 0, step, a, dt);
                if (iter > 0) st.rhs_evals++;
                if (r.bad) return imp_fail(s, "hs_solver_run: AM12 iterate is not finite", dt, dtio, stats, st);
                const double t = am_norm(r.d, r.a, ab, rl);
                alpha = t / er;
                er = t;
                iter++;
                st.iterations++;
                if (er <= maxIterErr) break;
                if (alpha >= 1.0 || iter > maxIter) {
                    badIt = true;
                    break;
                }
            }
            if (badIt && tr == maxTry - 1)
                return imp_fail(s, "hs_solver_run: AM12 corrector iteration did not converge", dt, dtio, stats, st);
            if (badIt) {
                dt = (alpha > alpha_ref) ? dt * pow(alpha_ref / alpha, 0.5) : 0.5 * dt;
                a++;
                st.rejected++;
                continue;
            }
            // step error: corrector - predictor, scaled by the predictor
            tErr = am_norm(r.d1, a1, ab, rl);
            const bool bad = tErr > 1.0;
            double nd = dt * pow(headRoom / tErr, 0.5);
            if (restrict_step) {
                nd = std::min(nd, dt * pow(alpha_ref / saved_alpha, 0.5));
                restrict_step = false;
            }
            nd = std::max(nd, mindt);
            nd = std::min(nd, maxdt);
            const double used = dt;
            dt = nd;
            if (!bad || nd == mindt) {
                st.t += used;
                break;
            }
            if (tr > maxTry) return imp_fail(s, "hs_solver_run: AM12 did not converge", dt, dtio, stats, st);
            tr++;
            a++;
            st.rejected++;
        }
        implicit_accept(s);
        st.steps++;
        st.last_err = tErr;
    }
    *dtio = dt;
    st.dt = dt;
    if (stats) *stats = st;
    return hs_last_error();
}

// BDFEuler.Step, engine/solve_euler_backward.go:29-86, one equation (m) for
// the whole batch: y <- y0 + dt f(y) until max|y - yb| / max|y + yb| <=
// iterErr over every value; more than maxIter updates fail.
int run_bdf_euler(hs_solver* s, long nsteps, long step0, double* dtio, hs_run_stats* stats) {
    hs_run_stats st{};
    const double dt = *dtio;
    for (long n = 0; n < nsteps; n++) {
        const uint32_t step = (uint32_t)(step0 + n);
        implicit_start(s, dt, step);
        st.rhs_evals++;
        st.iterations++;
        double err = 1.0e38;
        int iter = 0;
        while (err > s->imp_iter_err) {
            ImpRed r = imp_update(s, nullptr, s->k2, nullptr, nullptr, dt, true, step, 0, dt);
            st.rhs_evals++;
            if (r.bad) return imp_fail(s, "hs_solver_run: backward-Euler iterate is not finite", dt, dtio, stats, st);
            err = std::max(r.d[0], std::max(r.d[1], r.d[2]));
            if (r.sum > 0.0) err = err / r.sum;
            iter++;
            st.iterations++;
            if (iter > s->imp_max_iter)
                return imp_fail(s, "hs_solver_run: backward-Euler iteration did not converge", dt, dtio, stats, st);
        }
        implicit_accept(s);
        st.steps++;
        st.t += dt;
        st.last_err = err;
    }
    st.dt = dt;
    if (stats) *stats = st;
    return hs_last_error();
}
// ---- AM12 of a two-temperature solve: equations Te, Tp, m -------------------
// BDFAM12.Step (solve_am01.go:45-300) for the laser recipe's equations in
// their registration order (examples/laser-demag.py:35-60: temperature/ETM,
// temperature/PTM, then llbar), the LLB temperature T = Te of the iterate.
// Buffers: state Te[0], Tp[0], Tf[0]; iterate Te[1], Tp[1], Tf[1]; dy0 k1e,
// k1p; dyE k2e, k2p; the current RHS ee, ep; the predictor result Tea, Tpa.
struct TmRed {
    ImpRed m;
    double d[2], a[2], d1[2];
    bool bad;
};

// RHS of all three equations at the iterate (time t): m's into km, Te's and Tp's into ke, kp
void tm_rhs_all(hs_solver* s, float* km, double* ke, double* kp, uint32_t step, uint32_t attempt, double dt,
                double t) {
    imp_rhs(s, km, step, attempt, dt, s->Tf[1], s->qs);
    TmIO io{};
    io.Te = s->Te[1];
    io.Tp = s->Tp[1];
    io.k1e = ke;
    io.k1p = kp;
    io.t = t;
    io.qs = s->qs;
    hipLaunchKernelGGL(k_tm<R_RHS>, dim3(hs::grid1d(s->N), 1), 256, 0, s->st, s->phys, s->tmc, io);
}

// One fixed-point update of the three equations from the slopes km, ke, kp
// (trapezoidal with dy0 when trap); the maxima of all three in one readback.
TmRed tm_updates(hs_solver* s, const float* km, const double* ke, const double* kp, bool trap, bool norm_in,
                 bool with_y1, double h) {
    const long region = (long)kImpLanes * kImpStride;
    unsigned long long* rT = s->imp + kImpVals * region;
    unsigned long long* rP = rT + kImp1Vals * region;
    const double c = trap ? 0.5 * h : h;
    const unsigned g = hs::grid1d(s->N);
    hipLaunchKernelGGL(k_imp, dim3(g, 1), hs::kBlock, 0, s->st, s->y0, s->acc, km, trap ? s->e : nullptr,
                       with_y1 ? s->y1 : nullptr, s->mB, c, norm_in ? s->p.normalize : 0, s->p.normalize, s->imp,
                       s->N);
    hipLaunchKernelGGL(k_imp1, g, hs::kBlock, 0, s->st, s->Te[0], s->Te[1], ke, trap ? s->k1e : nullptr,
                       with_y1 ? s->Tea : nullptr, s->Tf[1], c, rT, s->N);
    hipLaunchKernelGGL(k_imp1, g, hs::kBlock, 0, s->st, s->Tp[0], s->Tp[1], kp, trap ? s->k1p : nullptr,
                       with_y1 ? s->Tpa : nullptr, (float*)nullptr, c, rP, s->N);
    double* fold = (double*)(s->imp + (long)kImpAll * region);
    hipLaunchKernelGGL(k_imp_fold, 1, 64, 0, s->st, s->imp, kImpAll, fold);
    hipMemcpyAsync(s->imp_host, fold, sizeof(double) * kImpAll, hipMemcpyDeviceToHost, s->st);
    hipStreamSynchronize(s->st);
    auto mx = [&](long q) { return s->imp_host[q]; };
    TmRed r{};
    for (int p = 0; p < 3; p++) {
        r.m.d[p] = mx(p);
        r.m.a[p] = mx(3 + p);
        r.m.d1[p] = mx(6 + p);
    }
    r.m.sum = mx(9);
    r.m.bad = mx(10) != 0.0;
    for (int e = 0; e < 2; e++) {
        const long o = kImpVals + e * kImp1Vals;
        r.d[e] = mx(o);
        r.a[e] = mx(o + 1);
        r.d1[e] = mx(o + 2);
        r.bad = r.bad || mx(o + 3) != 0.0;
    }
    r.bad = r.bad || r.m.bad;
    return r;
}

// one-component error (solve_am01.go:122-132 with COMP = 1)
double t1_norm(double d, double a, double absErr, double relErr) {
    return (1.0 / sqrt(1.0)) * sqrt(pow(d / (absErr + a * relErr), 2.0));
}

int run_am12_tm(hs_solver* s, long nsteps, long step0, double* dtio, double mindt, double maxdt,
                hs_run_stats* stats) {
    const int maxTry = 10, maxIter = 5;
    const double headRoom = 0.8, maxIterErr = 0.1, alpha_ref = 0.6;
    const double ab[3] = {s->imp_absT, s->imp_absT, s->imp_abs}, rl[3] = {s->imp_relT, s->imp_relT, s->imp_rel};
    const unsigned g = hs::grid1d(s->N);
    hs_run_stats st{};
    double dt = *dtio;
    // the errors of one update, equation order Te, Tp, m
    auto errs = [&](const TmRed& r, double out[3]) {
        out[0] = t1_norm(r.d[0], r.a[0], ab[0], rl[0]);
        out[1] = t1_norm(r.d[1], r.a[1], ab[1], rl[1]);
        out[2] = am_norm(r.m.d, r.m.a, ab[2], rl[2]);
    };
    auto sort3 = [](double* v) { std::sort(v, v + 3); };
    for (long n = 0; n < nsteps; n++) {
        const uint32_t step = (uint32_t)(step0 + n);
        const double t0 = s->tm_time;
        uint32_t a = 0;
        // RHS of the state (time t0): m's into e (and the Euler start of m), Te's and Tp's into k1e, k1p
        implicit_start(s, dt, step, s->Tf[0], s->qs);
        {
            TmIO io{};
            io.Te = s->Te[0];
            io.Tp = s->Tp[0];
            io.k1e = s->k1e;
            io.k1p = s->k1p;
            io.t = t0;
            io.qs = s->qs;
            hipLaunchKernelGGL(k_tm<R_RHS>, dim3(g, 1), 256, 0, s->st, s->phys, s->tmc, io);
        }
        st.rhs_evals++;
        int tr = 0;
        bool restrict_step = false, first = true;
        double salpha[3] = {0, 0, 0}, tErr[3] = {0, 0, 0}, used = 0.0;
        for (;;) {
            const double h = dt;
            if (!first)
                hipLaunchKernelGGL(k_rk12_predict, dim3(g, 1), hs::kBlock, 0, s->st, s->y0, s->e, s->acc, s->mB, h,
                                   s->p.normalize, s->N);
            first = false;
            hipLaunchKernelGGL(k_axpy1, g, hs::kBlock, 0, s->st, s->Te[0], s->k1e, s->Te[1], s->Tf[1], h, s->N);
            hipLaunchKernelGGL(k_axpy1, g, hs::kBlock, 0, s->st, s->Tp[0], s->k1p, s->Tp[1], (float*)nullptr, h,
                               s->N);
            st.iterations++;
            double er[3] = {maxIterErr, maxIterErr, maxIterErr}, alp[3], alpha = 0.0, te[3];
            bool badIt = false;
            TmRed r{};
            // predictor: implicit Euler; the first slopes (at the Euler point) kept as dyE
            for (int iter = 0;;) {
                float* km = iter == 0 ? s->k1 : s->k2;
                double* ke = iter == 0 ? s->k2e : s->ee;
                double* kp = iter == 0 ? s->k2p : s->ep;
                tm_rhs_all(s, km, ke, kp, step, a, dt, t0 + dt);
                st.rhs_evals++;
                r = tm_updates(s, km, ke, kp, false, true, false, h);
                if (r.bad) return imp_fail(s, "hs_solver_run: AM12 iterate is not finite", dt, dtio, stats, st);
                errs(r, te);
                for (int e = 0; e < 3; e++) {
                    alp[e] = te[e] / er[e];
                    salpha[e] = alp[e];
                    er[e] = te[e];
                }
                sort3(er);
                sort3(alp);
                alpha = alp[2];
                iter++;
                st.iterations++;
                if (er[2] <= maxIterErr) break;
                if (alpha >= 1.0 || iter > maxIter) {
                    badIt = true;
                    break;
                }
            }
            if (badIt && tr == maxTry - 1)
                return imp_fail(s, "hs_solver_run: AM12 predictor iteration did not converge", dt, dtio, stats, st);
            if (badIt) {
                dt = (alpha > alpha_ref) ? dt * pow(alpha_ref / alpha, 0.5) : 0.5 * dt;
                restrict_step = true;
                a++;
                st.rejected++;
                continue;
            }
            const double a1m[3] = {r.m.a[0], r.m.a[1], r.m.a[2]}, a1T[2] = {r.a[0], r.a[1]};
            hipMemcpyAsync(s->y1, s->acc, sizeof(double) * 3 * s->N, hipMemcpyDeviceToDevice, s->st);
            hipMemcpyAsync(s->Tea, s->Te[1], sizeof(double) * s->N, hipMemcpyDeviceToDevice, s->st);
            hipMemcpyAsync(s->Tpa, s->Tp[1], sizeof(double) * s->N, hipMemcpyDeviceToDevice, s->st);
            for (int e = 0; e < 3; e++) er[e] = maxIterErr;
            // corrector: trapezoidal, starting from the slopes at the Euler point
            for (int iter = 0;;) {
                const float* km = s->k1;
                const double *ke = s->k2e, *kp = s->k2p;
                if (iter > 0) {
                    tm_rhs_all(s, s->k2, s->ee, s->ep, step, a, dt, t0 + dt);
                    st.rhs_evals++;
                    km = s->k2;
                    ke = s->ee;
                    kp = s->ep;
                }
                r = tm_updates(s, km, ke, kp, true, iter > 0, true, h);
                if (r.bad) return imp_fail(s, "hs_solver_run: AM12 iterate is not finite", dt, dtio, stats, st);
                errs(r, te);
                for (int e = 0; e < 3; e++) {
                    alp[e] = te[e] / er[e];
                    er[e] = te[e];
                }
                sort3(er);
                sort3(alp);
                alpha = alp[2];
                iter++;
                st.iterations++;
                if (er[2] <= maxIterErr) break;
                if (alpha >= 1.0 || iter > maxIter) {
                    badIt = true;
                    break;
                }
            }
            if (badIt && tr == maxTry - 1)
                return imp_fail(s, "hs_solver_run: AM12 corrector iteration did not converge", dt, dtio, stats, st);
            if (badIt) {
                dt = (alpha > alpha_ref) ? dt * pow(alpha_ref / alpha, 0.5) : 0.5 * dt;
                a++;
                st.rejected++;
                continue;
            }
            // step errors (corrector - predictor) and the controller; the alpha
            // restriction applies to the first equation only (solve_am01.go:253-258)
            tErr[0] = t1_norm(r.d1[0], a1T[0], ab[0], rl[0]);
            tErr[1] = t1_norm(r.d1[1], a1T[1], ab[1], rl[1]);
            tErr[2] = am_norm(r.m.d1, a1m, ab[2], rl[2]);
            bool bad = false;
            double nd[3];
            for (int e = 0; e < 3; e++) {
                if (tErr[e] > 1.0) bad = true;
                double x = dt * pow(headRoom / tErr[e], 0.5);
                if (restrict_step) {
                    x = std::min(x, dt * pow(alpha_ref / salpha[e], 0.5));
                    restrict_step = false;
                }
                x = std::max(x, mindt);
                x = std::min(x, maxdt);
                nd[e] = x;
            }
            sort3(nd);
            used = dt;
            dt = nd[0];
            if (!bad || nd[0] == mindt) {
                st.t += used;
                break;
            }
            if (tr > maxTry) return imp_fail(s, "hs_solver_run: AM12 did not converge", dt, dtio, stats, st);
            tr++;
            a++;
            st.rejected++;
        }
        implicit_accept(s);
        hipMemcpyAsync(s->Te[0], s->Te[1], sizeof(double) * s->N, hipMemcpyDeviceToDevice, s->st);
        hipMemcpyAsync(s->Tp[0], s->Tp[1], sizeof(double) * s->N, hipMemcpyDeviceToDevice, s->st);
        hipMemcpyAsync(s->Tf[0], s->Tf[1], sizeof(float) * s->N, hipMemcpyDeviceToDevice, s->st);
        s->tm_time = t0 + used;
        st.steps++;
        st.last_err = std::max(tErr[0], std::max(tErr[1], tErr[2]));
    }
    *dtio = dt;
    st.dt = dt;
    if (stats) *stats = st;
    return hs_last_error();
}
}  // namespace

extern "C" int hs_solver_set_implicit_T(hs_solver* s, double maxAbsErr, double maxRelErr) {
    if (!s || !(maxAbsErr > 0.0) || !(maxRelErr > 0.0)) {
        hs::set_error(HS_EINVAL, "hs_solver_set_implicit_T: tolerances must be positive");
        return HS_EINVAL;
    }
    s->imp_absT = maxAbsErr;
    s->imp_relT = maxRelErr;
    return 0;
}

extern "C" int hs_solver_set_implicit(hs_solver* s, double maxAbsErr, double maxRelErr, double maxIterErr,
                                      int maxIter) {
    if (!s || !(maxAbsErr > 0.0) || !(maxRelErr > 0.0) || !(maxIterErr > 0.0) || maxIter < 1) {
        hs::set_error(HS_EINVAL, "hs_solver_set_implicit: tolerances must be positive");
        return HS_EINVAL;
    }
    s->imp_abs = maxAbsErr;
    s->imp_rel = maxRelErr;
    s->imp_iter_err = maxIterErr;
    s->imp_max_iter = maxIter;
    return 0;
}

extern "C" int hs_solver_run(hs_solver* s, int method, long nsteps, long step0, double* dtio, double maxErr,
                             double mindt, double maxdt, hs_run_stats* stats) {
    if (!s || !dtio || nsteps < 0) return HS_EINVAL;
    if (!whole_batch_ok(s, "hs_solver_run: handle is in slot mode")) return HS_EINVAL;
    if (method == HS_RK23) return run_rk23_device(s, nsteps, step0, dtio, maxErr, mindt, maxdt, stats);
    if (s->tm && method == HS_AM12) {
        if (!ensure_imp(s)) return HS_ENOMEM;
        int rc = run_am12_tm(s, nsteps, step0, dtio, mindt, maxdt, stats);
        hipStreamSynchronize(s->st);
        HS_LAUNCH_CHECK("hs_solver_run");
        return rc ? rc : hs_last_error();
    }
    if (s->tm) {
        hs::set_error(HS_EINVAL, "hs_solver_run: two-temperature solves take HS_RK23 or HS_AM12");
        return HS_EINVAL;
    }
    if (method == HS_AM12 || method == HS_BDF_EULER) {
        if (!ensure_imp(s)) return HS_ENOMEM;
        int rc = method == HS_AM12 ? run_am12(s, nsteps, step0, dtio, mindt, maxdt, stats)
                                   : run_bdf_euler(s, nsteps, step0, dtio, stats);
        hipStreamSynchronize(s->st);
        HS_LAUNCH_CHECK("hs_solver_run");
        return rc ? rc : hs_last_error();
    }
    hs_run_stats st{};
    double dt = *dtio;
    const int maxTry = 10;
    for (long n = 0; n < nsteps; n++) {
        uint32_t step = (uint32_t)(step0 + n);
        if (method == HS_RK4) {
            rk4_step(s, base_io(s, dt, step, 0));
            s->b0_valid = false;
            st.rhs_evals += 4;
            st.t += dt;
        } else if (method == HS_EULER) {
            ensure_b0(s);
            StageIO io = base_io(s, dt, step, 0);
            io.m = s->mA; io.B = s->B0; io.mout = s->mB;
            launch_stage<R_EULER>(s, io);
            std::swap(s->mA, s->mB);
            drop_graphs(s);
            s->b0_valid = false;
            st.rhs_evals += 1;
            st.t += dt;
        } else if (method == HS_RK12 || method == HS_HEUN) {
            int tr = 0;
            double err = 0;
            ensure_b0(s);
            StageIO io = base_io(s, dt, step, 0);
            io.m = s->mA; io.B = s->B0; io.mout = s->mB;
            launch_stage<R_RK12_1>(s, io);
            for (;;) {
                if (tr > 0) {
                    hipLaunchKernelGGL(k_rk12_predict, dim3(hs::grid1d(s->N), s->nsolve), hs::kBlock, 0, s->st, s->y0,
                                       s->e, s->acc, s->mB, dt, s->p.normalize, s->N);
                }
                demag_into(s, s->mB, s->Bs);
                zero_err(s);
                io = base_io(s, dt, step, (uint32_t)tr);
                io.m = s->mB; io.B = s->Bs; io.mout = s->mC;
                launch_stage<R_RK12_2>(s, io);
                st.rhs_evals += 1;
                if (method == HS_HEUN) {
                    st.t += dt;
                    break;
                }
                err = read_err(s) * dt;
                if (!std::isfinite(err)) {
                    // never accepted: retry at the smallest factor while retries
                    // remain and dt can still shrink
                    const double nd = controller(dt, err, maxErr, true, 0.5, 1.0 / 3.0, mindt, maxdt);
                    if (tr > maxTry || !(nd < dt)) {
                        hs::set_error(HS_ECONV, "hs_solver_run: RK12 error estimate is not finite");
                        *dtio = dt;
                        if (stats) *stats = st;
                        return HS_ECONV;
                    }
                    dt = nd;
                    tr++;
                    st.rejected++;
                    continue;
                }
                bool bad = err > maxErr;
                double nd = controller(dt, err, maxErr, bad, 0.5, 1.0 / 3.0, mindt, maxdt);
                double used = dt;
                dt = nd;
                if (!bad || nd == mindt) {
                    st.t += used;
                    break;
                }
                if (tr > maxTry) {
                    hs::set_error(HS_ECONV, "hs_solver_run: RK12 did not converge");
                    *dtio = dt;
                    if (stats) *stats = st;
                    return HS_ECONV;
                }
                tr++;
                st.rejected++;
            }
            std::swap(s->y0, s->acc);
            std::swap(s->mA, s->mC);
            drop_graphs(s);
            s->b0_valid = false;
            st.last_err = err;
        } else {
            hs::set_error(HS_EINVAL, "hs_solver_run: unknown method");
            return HS_EINVAL;
        }
        st.steps++;
    }
    *dtio = dt;
    st.dt = dt;
    hipStreamSynchronize(s->st);
    HS_LAUNCH_CHECK("hs_solver_run");
    if (stats) *stats = st;
    return hs_last_error();
}

// ---------------------------------------------------------------- slots
extern "C" int hs_slot_admit(hs_solver* s, int slot, const double* m, double TMul, uint64_t seed, uint32_t stream,
                             const double* Bext, long step0, long nsteps, double dt, double maxErr, double mindt,
                             double maxdt) {
    if (!s || !m || slot < 0 || slot >= s->nsolve || nsteps < 0 || !(dt > 0.0) || !(maxErr > 0.0) || s->tm) {
        hs::set_error(HS_EINVAL, "hs_slot_admit: bad argument (or a two-temperature solver)");
        return HS_EINVAL;
    }
    if (!s->slot_mode) {
        // first admission: every slot idle, pinned record buffer, graphs captured lazily
        if (!ensure_devctl(s)) return HS_ENOMEM;
        hipMemsetAsync(s->clk, 0, sizeof(Clock) * s->nsolve, s->st);
        hipMemsetAsync(s->errmax, 0, sizeof(unsigned long long) * s->nsolve * kErrPerSolve, s->st);
        hipMemsetAsync(s->live, 0, sizeof(int) * s->nsolve, s->st);
        memset(s->slot_host, 0, sizeof(hs_slot) * s->nsolve);
        s->slot_mode = true;
        s->slot_iter = 0;
    }
    // every copy below is ordered on the solver's own stream; waiting for it
    // alone also makes slot_host (copied back at the end of each
    // hs_slot_advance) the device's current slot records
    hipStreamSynchronize(s->st);
    Clock c = s->slot_host[slot];
    if (c.state == HS_SLOT_RUNNING) {
        hs::set_error(HS_EINVAL, "hs_slot_admit: slot is running");
        return HS_EINVAL;
    }
    // per-solve parameters (thermal prefactor as hs_batch_create)
    const hs_llb_params* p = &s->p;
    SolveP q = s->sp_host[slot];
    solve_constants(p, TMul, q);
    for (int k = 0; k < 3; k++) {
        q.bextT[k] = Bext ? Bext[k] : p->Bext[k];
        q.bext[k] = (float)(1.0 / HS_MU0 * q.bextT[k]);
    }
    q.seed = seed;
    q.stream = stream;
    s->sp_host[slot] = q;
    hipMemcpyAsync(s->sp + slot, &s->sp_host[slot], sizeof(SolveP), hipMemcpyHostToDevice, s->st);
    // state into the set that the next iteration reads, its FP32 copy and field
    const Set cur = slot_set(s, (int)(s->slot_iter & 1));
    const long off = (long)slot * 3 * s->N;
    hipMemcpyAsync(cur.y + off, m, sizeof(double) * 3 * s->N, hipMemcpyHostToDevice, s->st);
    hipLaunchKernelGGL(k_set_state, dim3(hs::grid1d(s->N), 1), hs::kBlock, 0, s->st, cur.y + off, cur.m + off,
                       s->p.normalize, s->N);
    if (s->phys.demag)
        hs_demag_field_f32_internal(s->demag, cur.m + off, s->phys.msk, s->p.msat0T0Mul, cur.B + off, s->st, 1,
                                    nullptr);
    memset(&c, 0, sizeof(c));
    c.step = (uint32_t)step0;
    c.h = dt;
    c.maxErr = maxErr;
    c.mindt = mindt;
    c.maxdt = maxdt;
    c.end = (uint32_t)(step0 + nsteps);
    c.state = nsteps > 0 ? HS_SLOT_RUNNING : HS_SLOT_DONE;
    c.home = (uint32_t)(s->slot_iter & 1);
    hipMemcpyAsync(s->clk + slot, &c, sizeof(Clock), hipMemcpyHostToDevice, s->st);
    const int lv = c.state == HS_SLOT_RUNNING ? 1 : 0;
    hipMemcpyAsync(s->live + slot, &lv, sizeof(int), hipMemcpyHostToDevice, s->st);
    hipStreamSynchronize(s->st);
    s->slot_host[slot] = c;
    HS_LAUNCH_CHECK("hs_slot_admit");
    return hs_last_error();
}

extern "C" int hs_slot_advance(hs_solver* s, int iters) {
    if (!s || iters < 0 || !s->slot_mode) {
        hs::set_error(HS_EINVAL, "hs_slot_advance: no slot admitted");
        return HS_EINVAL;
    }
    if (!ensure_slot_graphs(s)) return HS_EINVAL;
    for (int n = 0; n < iters; n++) {
        hipGraphLaunch(s->slot_graph[s->slot_iter & 1], s->st);
        s->slot_iter++;
    }
    hipMemcpyAsync(s->slot_host, s->clk, sizeof(Clock) * s->nsolve, hipMemcpyDeviceToHost, s->st);
    hipEventRecord(s->slot_ev, s->st);
    HS_LAUNCH_CHECK("hs_slot_advance");
    return hs_last_error();
}

extern "C" int hs_slot_ready(hs_solver* s) {
    if (!s || !s->slot_ev) return 1;
    return hipEventQuery(s->slot_ev) == hipSuccess ? 1 : 0;
}

extern "C" int hs_slot_status(hs_solver* s, hs_slot* out) {
    if (!s || !out || !s->slot_mode) {
        hs::set_error(HS_EINVAL, "hs_slot_status: no slot admitted");
        return HS_EINVAL;
    }
    hipEventSynchronize(s->slot_ev);
    memcpy(out, s->slot_host, sizeof(hs_slot) * s->nsolve);
    HS_LAUNCH_CHECK("hs_slot_status");
    return hs_last_error();
}

extern "C" int hs_slot_get_m(hs_solver* s, int slot, double* m) {
    if (!s || !m || !s->slot_mode || slot < 0 || slot >= s->nsolve) {
        hs::set_error(HS_EINVAL, "hs_slot_get_m: bad argument");
        return HS_EINVAL;
    }
    hipStreamSynchronize(s->st);  // slot_host is then the device's current record
    const Clock c = s->slot_host[slot];
    const int set = c.state == HS_SLOT_RUNNING ? (int)(s->slot_iter & 1) : (int)c.home;
    hipMemcpyAsync(m, slot_set(s, set).y + (long)slot * 3 * s->N, sizeof(double) * 3 * s->N, hipMemcpyDeviceToHost,
                   s->st);
    hipStreamSynchronize(s->st);
    HS_LAUNCH_CHECK("hs_slot_get_m");
    return hs_last_error();
}

extern "C" int hs_slot_release(hs_solver* s, int slot) {
    if (!s || !s->slot_mode || slot < 0 || slot >= s->nsolve) {
        hs::set_error(HS_EINVAL, "hs_slot_release: bad argument");
        return HS_EINVAL;
    }
    // ordered after the iterations already enqueued on the solver's stream
    hipMemsetAsync(s->clk + slot, 0, sizeof(Clock), s->st);
    hipMemsetAsync(s->live + slot, 0, sizeof(int), s->st);
    hipStreamSynchronize(s->st);
    memset(&s->slot_host[slot], 0, sizeof(Clock));
    HS_LAUNCH_CHECK("hs_slot_release");
    return hs_last_error();
}

long hs_demag_footprint(const hs_demag* d);

extern "C" long hs_solver_footprint(hs_solver* s) {
    if (!s) return 0;
    const long NN = s->N * s->nsolve;
    long b = NN * 3 * (2 * sizeof(double) + 9 * sizeof(float));  // y0, acc; mA..mC, B0, Bs, e, k1..k3
    b += (long)s->nsolve * (sizeof(SolveP) + sizeof(Clock) + sizeof(unsigned long long) * kErrPerSolve);
    if (s->kout) b += NN * 3 * 2 * sizeof(float);
    b += (long)s->masks.size() * s->N * sizeof(float);
    if (s->demag) b += hs_demag_footprint(s->demag);
    return b;
}

// ---------------------------------------------------------------- ensembles
struct hs_ensemble {
    std::vector<hs_solver*> b;
    hipStream_t origin = nullptr;
    hipGraphExec_t graph = nullptr;
    double dt = 0.0;
    std::vector<hipEvent_t> ev;  // fork + one join per batch
};

extern "C" hs_ensemble* hs_ensemble_create(hs_solver* const* batches, int nbatch) {
    if (!batches || nbatch < 1) {
        hs::set_error(HS_EINVAL, "hs_ensemble_create: no batches");
        return nullptr;
    }
    hs_ensemble* e = new hs_ensemble();
    for (int i = 0; i < nbatch; i++) {
        if (!batches[i] || batches[i]->slot_mode) {
            hs::set_error(HS_EINVAL, "hs_ensemble_create: bad batch");
            delete e;
            return nullptr;
        }
        for (int j = 0; j < i; j++)
            if (batches[j]->st == batches[i]->st) {
                hs::set_error(HS_EINVAL, "hs_ensemble_create: batches must have distinct streams");
                delete e;
                return nullptr;
            }
        e->b.push_back(batches[i]);
    }
    hipStreamCreateWithFlags(&e->origin, hipStreamNonBlocking);
    e->ev.resize(nbatch + 1);
    for (auto& v : e->ev) hipEventCreateWithFlags(&v, hipEventDisableTiming);
    return e;
}

extern "C" int hs_ensemble_run_graph(hs_ensemble* e, long nsteps, long step0, double dt) {
    if (!e || nsteps < 0) return HS_EINVAL;
    if (e->graph && e->dt != dt) {
        hipGraphExecDestroy(e->graph);
        e->graph = nullptr;
    }
    for (hs_solver* s : e->b) set_clocks(s, step0, dt);
    if (!e->graph) {
        bool prof = hs::prof_on();
        if (prof) hs_prof_enable(0);
        hipGraph_t g;
        bool ok = hipStreamBeginCapture(e->origin, hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (ok) {
            hipEventRecord(e->ev[0], e->origin);
            for (size_t i = 0; i < e->b.size(); i++) {
                hs_solver* s = e->b[i];
                hipStreamWaitEvent(s->st, e->ev[0], 0);
                rk4_graph_step(s);
                hipEventRecord(e->ev[i + 1], s->st);
                hipStreamWaitEvent(e->origin, e->ev[i + 1], 0);
            }
            ok = hipStreamEndCapture(e->origin, &g) == hipSuccess;
            if (ok) {
                ok = hipGraphInstantiate(&e->graph, g, nullptr, nullptr, 0) == hipSuccess;
                hipGraphDestroy(g);
            }
        }
        if (prof) hs_prof_enable(1);
        if (!ok) {
            e->graph = nullptr;
            hs::set_error(HS_EINVAL, "hs_ensemble_run_graph: graph capture failed");
            return HS_EINVAL;
        }
        e->dt = dt;
    }
    for (long n = 0; n < nsteps; n++) hipGraphLaunch(e->graph, e->origin);
    hipStreamSynchronize(e->origin);
    for (hs_solver* s : e->b) s->b0_valid = false;
    HS_LAUNCH_CHECK("hs_ensemble_run_graph");
    return hs_last_error();
}

extern "C" void hs_ensemble_destroy(hs_ensemble* e) {
    if (!e) return;
    if (e->graph) hipGraphExecDestroy(e->graph);
    for (auto& v : e->ev) hipEventDestroy(v);
    if (e->origin) hipStreamDestroy(e->origin);
    delete e;
}

// ---------------------------------------------------------------- two-temperature solves
extern "C" hs_solver* hs_solver_create_tm(const hs_llb_params* p, const hs_tm_params* tm, double T0,
                                          CUstream stream) {
    if (!p || !tm || !(tm->gammaE > 0.0) || !(tm->Cl > 0.0 || (tm->Td > 0.0 && tm->nD > 0.0)) || !(T0 > 0.0) ||
        !(tm->maxErrT > 0.0)) {
        hs::set_error(HS_EINVAL, "hs_solver_create_tm: bad parameters");
        return nullptr;
    }
    hs_llb_params q = *p;
    q.TMul = 1.0;  // T = Te per cell
    q.T = nullptr;
    hs_solver* s = hs_batch_create(&q, 1, nullptr, nullptr, nullptr, stream);
    if (!s) return nullptr;
    s->tm = true;
    s->tmp = *tm;
    s->phys.lf_uniform = 0;  // the longitudinal field sees a per-cell temperature
    TmP& c = s->tmc;
    c.gammaE = tm->gammaE, c.Cl = tm->Cl, c.G = tm->G, c.kappa = tm->kappa, c.Q0 = tm->Q0;
    c.ir2 = tm->sigma_r > 0 ? 1.0 / (2.0 * tm->sigma_r * tm->sigma_r) : 0.0;
    c.it2 = tm->sigma_t > 0 ? 1.0 / (2.0 * tm->sigma_t * tm->sigma_t) : 0.0;
    c.t0 = tm->t0, c.xc = tm->xc, c.yc = tm->yc;
    c.c2x = 1.0 / (p->cell[0] * p->cell[0]);
    c.c2y = 1.0 / (p->cell[1] * p->cell[1]);
    c.c2z = 1.0 / (p->cell[2] * p->cell[2]);
    c.hx = p->cell[2], c.hy = p->cell[1];
    c.Td = tm->Td;
    c.c9nkB = 9.0 * tm->nD * HS_KB;
    s->phys.sh_w = 3.0 * HS_KB * p->TcMul * p->nMul * p->gammaLL;
    s->phys.sh_q = HS_MU0 * p->msat0T0Mul * p->gammaLL;
    const long N = s->N;
    double** d64[] = {&s->Te[0], &s->Te[1], &s->Tp[0], &s->Tp[1], &s->Tea, &s->Tpa, &s->Teb, &s->Tpb,
                      &s->k1e, &s->k1p, &s->k2e, &s->k2p, &s->ee, &s->ep};
    float** d32[] = {&s->Tf[0], &s->Tf[1], &s->Tfa, &s->Tfb};
    bool ok = true;
    for (auto q2 : d64) ok = ok && hipMalloc(q2, sizeof(double) * N) == hipSuccess;
    for (auto q2 : d32) ok = ok && hipMalloc(q2, sizeof(float) * N) == hipSuccess;
    if (tm->spin_heat) ok = ok && hipMalloc(&s->qs, sizeof(float) * N) == hipSuccess;
    if (!ok) {
        hs::set_error(HS_ENOMEM, "hs_solver_create_tm: out of device memory");
        hs_solver_destroy(s);
        return nullptr;
    }
    std::vector<double> t(N, T0);
    if (hs_solver_set_T(s, t.data(), t.data()) != 0) {
        hs_solver_destroy(s);
        return nullptr;
    }
    return s;
}

extern "C" int hs_solver_set_T(hs_solver* s, const double* Te, const double* Tp) {
    if (!s || !s->tm) {
        hs::set_error(HS_EINVAL, "hs_solver_set_T: not a two-temperature solver");
        return HS_EINVAL;
    }
    if (Te) {
        hipMemcpyAsync(s->Te[0], Te, sizeof(double) * s->N, hipMemcpyHostToDevice, s->st);
        hipLaunchKernelGGL(k_d2f, hs::grid1d(s->N), hs::kBlock, 0, s->st, s->Te[0], s->Tf[0], s->N);
    }
    if (Tp) hipMemcpyAsync(s->Tp[0], Tp, sizeof(double) * s->N, hipMemcpyHostToDevice, s->st);
    hipStreamSynchronize(s->st);
    HS_LAUNCH_CHECK("hs_solver_set_T");
    return hs_last_error();
}

extern "C" int hs_solver_get_T(hs_solver* s, double* Te, double* Tp) {
    if (!s || !s->tm) {
        hs::set_error(HS_EINVAL, "hs_solver_get_T: not a two-temperature solver");
        return HS_EINVAL;
    }
    if (Te) hipMemcpyAsync(Te, s->Te[0], sizeof(double) * s->N, hipMemcpyDeviceToHost, s->st);
    if (Tp) hipMemcpyAsync(Tp, s->Tp[0], sizeof(double) * s->N, hipMemcpyDeviceToHost, s->st);
    hipStreamSynchronize(s->st);
    HS_LAUNCH_CHECK("hs_solver_get_T");
    return hs_last_error();
}

extern "C" int hs_solver_set_time(hs_solver* s, double t) {
    if (!s) return HS_EINVAL;
    s->tm_time = t;
    return 0;
}
extern "C" double hs_solver_time(hs_solver* s) { return s ? s->tm_time : 0.0; }

extern "C" int hs_solver_tm_rhs(hs_solver* s, double t, double* dTe, double* dTp) {
    if (!s || !s->tm) {
        hs::set_error(HS_EINVAL, "hs_solver_tm_rhs: not a two-temperature solver");
        return HS_EINVAL;
    }
    if (s->qs) {  // spin heating: the LLB RHS of the state (step 0, dt 1e-15) leaves the spin term of Qe
        ensure_b0(s);
        StageIO sio = base_io(s, 1e-15, 0, 0);
        sio.m = s->mA;
        sio.B = s->B0;
        sio.kout = sio.hout = nullptr;
        sio.T = s->Tf[0];
        sio.qs = s->qs;
        launch_stage<R_RHS>(s, sio);
    }
    TmIO io{};
    io.Te = s->Te[0];
    io.Tp = s->Tp[0];
    io.k1e = s->k1e;
    io.k1p = s->k1p;
    io.t = t;
    io.qs = s->qs;
    hipLaunchKernelGGL(k_tm<R_RHS>, dim3(hs::grid1d(s->N), 1), 256, 0, s->st, s->phys, s->tmc, io);
    if (dTe) hipMemcpyAsync(dTe, s->k1e, sizeof(double) * s->N, hipMemcpyDeviceToHost, s->st);
    if (dTp) hipMemcpyAsync(dTp, s->k1p, sizeof(double) * s->N, hipMemcpyDeviceToHost, s->st);
    hipStreamSynchronize(s->st);
    HS_LAUNCH_CHECK("hs_solver_tm_rhs");
    return hs_last_error();
}
