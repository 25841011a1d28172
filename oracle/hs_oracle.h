/*
 * hs_oracle.h -- FP64 CPU restatement of hotspin's per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity oracle: tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it as the
 * checker.  The product (libhotspin_amd.so) never links, loads or calls it.
 *
 * Conventions (same as the reference, src/hotspin-core/gpu/array.go:26-37):
 *   - internal axes (N0,N1,N2) = user (Nz,Ny,Nx); index I = i*N1*N2 + j*N2 + k
 *   - vector quantities are three SoA planes, internal component 0 = user z
 *   - a NULL mask means "uniform" (1.0, or 0.0 for the kernel octants)
 *
 * Every function cites the reference file:line whose formula it restates.
 * Compiled with -O2 -ffp-contract=off (no FMA contraction, no fast-math).
 */
#ifndef HS_ORACLE_H
#define HS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- constants: src/libhotspin/common_func.h:24-33, src/hotspin-core/common/constants.go:16-24 */
#define OR_MU0  (4.0e-7 * 3.14159265358979323846)
#define OR_KB   1.38064881313131313E-23
#define OR_MUB  9.27400968202020202E-24

/* ---- element-wise kernels (legacy ABI argument meaning, host pointers, no stream) ---- */
void or_exchange6(double* hx, double* hy, double* hz,
                  const double* mx, const double* my, const double* mz,
                  const double* lex, int N0, int N1, int N2,
                  int p0, int p1, int p2, double msat0T0Mul,
                  double pre0, double pre1, double pre2);
void or_llbar_torque(double* tx, double* ty, double* tz,
                     const double* mx, const double* my, const double* mz,
                     const double* hx, const double* hy, const double* hz,
                     const double* msat0T0, int N);
void or_llbar_local00nc(double* tx, double* ty, double* tz,
                        const double* hx, const double* hy, const double* hz,
                        const double* msat0T0,
                        const double* lxx, const double* lyy, const double* lzz,
                        double lmx, double lmy, double lmz, int N);
void or_llbar_local02nc(double* tx, double* ty, double* tz,
                        const double* mx, const double* my, const double* mz,
                        const double* hx, const double* hy, const double* hz,
                        const double* msat0T0,
                        const double* uxx, const double* uyy, const double* uzz,
                        double umx, double umy, double umz, int N);
void or_llbar_local02c(double* tx, double* ty, double* tz,
                       const double* mx, const double* my, const double* mz,
                       const double* hx, const double* hy, const double* hz,
                       const double* msat0T0,
                       const double* uxx, const double* uyy, const double* uzz,
                       double umx, double umy, double umz, int N);
void or_long_field(double* hx, double* hy, double* hz,
                   const double* mx, const double* my, const double* mz,
                   const double* msat0T0, const double* S, const double* n,
                   const double* Tc, const double* Ts,
                   double msat0T0Mul, double SMul, double nMul, double TcMul, double TsMul,
                   int N);
void or_temperature_scale_noise(double* hx, double* hy, double* hz,
                                const double* mu_xx, const double* mu_yy, const double* mu_zz,
                                const double* T, const double* msat0T0,
                                double muMul_xx, double muMul_yy, double muMul_zz,
                                double KB2TMul_mu0VgammaDtMsatMul, int N);
void or_brillouin(double* msat0, const double* msat0T0, const double* T,
                  const double* Tc, const double* S,
                  double msat0Mul, double msat0T0Mul, double TcMul, double SMul, int N);
void or_normalize(double* mx, double* my, double* mz, int N);
void or_qspat(double* Q, const double* T, const double* k, double kMul,
              int sx, int sy, int sz, double csx, double csy, double csz,
              int px, int py, int pz);
void or_qinter(double* Qi, const double* Ti, const double* Tj, const double* G,
               double GMul, int N);
void or_kappa(double* kappa, const double* ms0Msk, const double* ms0T0Msk, const double* T,
              const double* TcMsk, const double* SMsk, const double* nMsk,
              double ms0Mul, double ms0T0Mul, double TcMul, double SMul, int N);
void or_ts(double* Ts, const double* msMsk, const double* ms0T0Msk, const double* TcMsk, const double* SMsk,
           double msMul, double ms0T0Mul, double TcMul, double SMul, int N);
void or_cp(double* Cp, const double* T, const double* TdMsk, const double* nMsk, double TdMul, int N);
void or_energy_flow(double* w, const double* mx, const double* my, const double* mz,
                    const double* Rx, const double* Ry, const double* Rz,
                    const double* TcMsk, const double* SMsk, const double* nMsk, double SMul, int N);
void or_uniaxial_anisotropy(double* hx, double* hy, double* hz,
                            const double* mx, const double* my, const double* mz,
                            const double* KuMsk, const double* msMsk, double K2Mul,
                            const double* uxMsk, double uxMul, const double* uyMsk, double uyMul,
                            const double* uzMsk, double uzMul, int N);
void or_llbar_nonlocal00nc(double* t[3], double* const H[3], const double* msk, double* const lam[3],
                           const double lmul[3], int sx, int sy, int sz,
                           double csx, double csy, double csz, int px, int py, int pz);
void or_div_mul_pow(double* dst, const double* a, const double* b, const double* c, double p, int N);
void or_wavg(double* dst, const double* a, const double* b, const double* w0Msk, const double* w1Msk,
             const double* RMsk, double w0Mul, double w1Mul, double RMul, int N);
void or_madd(double* dst, const double* a, const double* b, double mulB, int N);
void or_madd2(double* a, const double* b, double mulB, const double* c, double mulC, int N);
void or_vec_madd(double* dst[3], double* const a[3], double* const b[3],
                 const double mulB[3], int N);
void or_lincomb2(double* dst, const double* a, double mulA, const double* b, double mulB, int N);
void or_add_madd(double* dst, const double* a, const double* b, const double* c, double mul, int N);
void or_mul(double* dst, const double* a, const double* b, int N);
void or_div(double* dst, const double* a, const double* b, int N);
void or_dot_mask(double* dst, const double* ax, const double* ay, const double* az,
                 const double* bx, const double* by, const double* bz,
                 double axm, double aym, double azm, double bxm, double bym, double bzm, int N);
void or_copy_pad3d(double* dst, int D0, int D1, int D2, const double* src, int S0, int S1, int S2);
void or_copy_unpad3d(double* dst, int D0, int D1, int D2, const double* src, int S0, int S1, int S2);
void or_tens_symm_vec_mul(double* dstX, double* dstY, double* dstZ,
                          const double* srcX, const double* srcY, const double* srcZ,
                          const double* kXX, const double* kYY, const double* kZZ,
                          const double* kYZ, const double* kXZ, const double* kXY,
                          double srcMul, int Nx, int Ny, int Nz);

/* ---- full reductions (reference: partial kernels + host finish) ---- */
double or_sum(const double* a, long N);
double or_max(const double* a, long N);
double or_min(const double* a, long N);
double or_maxabs(const double* a, long N);
double or_maxdiff(const double* a, const double* b, long N);
double or_maxsum(const double* a, const double* b, long N);
double or_maxnorm3sq(const double* x, const double* y, const double* z, long N);
double or_maxnorm3sqdiff(const double* x1, const double* y1, const double* z1,
                         const double* x2, const double* y2, const double* z2, long N);
double or_sdot(const double* a, const double* b, long N);

/* ---- FP64 FFT (own mixed radix), used by the demag restatement ---- */
/* in-place complex FFT of n points, stride 1, sign -1 forward / +1 backward, unnormalised */
int or_fft1d(double* re_im, int n, int sign);

/* ---- demag: Kernel_Arne + zero-padded FFT convolution ---- */
/* Kernel_Arne, src/hotspin-core/modules/kernel_magnetostatic.go:25-241.
 * size = padded (logic) size; out = 6 planes XX,YY,ZZ,YZ,XZ,XY of prod(size). */
void or_kernel_arne(const int size[3], const double cell[3], const int pbc[3],
                    int accuracy, double* out);
typedef struct or_demag or_demag;
/* Newell tensor (Newell, Williams & Dunlop 1993; no counterpart in the
 * reference): N of displacement R between cells h, and the full padded
 * kernel -N with or_kernel_arne's image ranges and layout. */
void or_newell_tensor(const double R[3], const double h[3], double N[6]);
void or_kernel_newell(const int size[3], const double cell[3], const int pbc[3], double* out);
or_demag* or_demag_new(const int N[3], const double cell[3], const int pbc[3], int accuracy);
/* kernel 0 = Kernel_Arne, 1 = Newell */
or_demag* or_demag_new_kernel(const int N[3], const double cell[3], const int pbc[3], int accuracy, int kernel);
void or_demag_free(or_demag* d);
void or_demag_sizes(const or_demag* d, int P[3], int K[3]);
/* 6 octant planes (XX,YY,ZZ,YZ,XZ,XY), each K0*K1*K2, scaled 1/(P0P1P2) */
const double* or_demag_kernel(const or_demag* d);
double or_demag_max_imag_ratio(const or_demag* d);
void or_demag_B(or_demag* d, const double* m /*3N*/, const double* msat0T0,
                double msat0T0Mul, double* B /*3N*/);

/* ---- Philox4x32-10 + Box-Muller (thermal noise stream) ---- */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* 3 standard normals for cell I: ctr={I, step, attempt, stream}, key={seed lo, seed hi} */
void or_thermal_normals(uint64_t seed, uint32_t stream, uint32_t step, uint32_t attempt,
                        long N, double* xi /*3N*/);

/* ---- LLB-r problem (module composition of examples/p2f.py and the SP4 recipe) ---- */
typedef struct {
    int N[3];
    double cell[3];
    int pbc[3];
    /* magnetisation */
    double msat0T0Mul;
    const double* msat0T0;            /* mask or NULL */
    /* exchange6 */
    int exchange;
    double lexMul;
    const double* lex;
    /* demag */
    int demag;
    int demag_accuracy;
    /* zeeman: B_ext in tesla, internal component order */
    int zeeman;
    double Bext[3];
    /* mfa/longfield */
    int longfield;
    double JMul, nMul, TcMul, TMul;
    const double *J, *n, *Tc, *T;
    /* temperature/brown (mu = alpha as in examples/p2f.py:95-99) */
    int thermal;
    uint64_t seed;
    uint32_t stream;
    double cutoff_dt;
    /* llbar */
    double gammaLL;
    int torque;
    int local00nc;
    double alphaMul[3];
    const double* alpha[3];
    /* normalizer (LLG limit) */
    int normalize;
    /* demag tensor: 0 = Kernel_Arne, 1 = Newell */
    int demag_kernel;
} or_llb_params;

typedef struct or_llb or_llb;
or_llb* or_llb_new(const or_llb_params* p);
void or_llb_free(or_llb* s);
/* one RHS evaluation: normalises m in place when the normalizer is loaded,
 * writes k (3N).  H_eff is kept and can be read with or_llb_heff. */
void or_llb_rhs(or_llb* s, double* m, double* k, uint32_t step, uint32_t attempt, double dt);
/* test hook: scales the Brown prefactor of or_llb_rhs (1 = the reference's formula) */
void or_debug_set_thermal_prefactor_scale(double f);
const double* or_llb_heff(const or_llb* s);
const double* or_llb_bdemag(const or_llb* s);
/* steppers: return accepted-step count (1) or <0 on failure */
int or_llb_step_euler(or_llb* s, double* m, double dt, uint32_t step);
int or_llb_step_heun(or_llb* s, double* m, double dt, uint32_t step);
int or_llb_step_rk4(or_llb* s, double* m, double dt, uint32_t step);
/* adaptive drivers: dt in/out; stats[0]=err of accepted step, stats[1]=tries */
int or_llb_step_rk12(or_llb* s, double* m, double* dt, double maxErr,
                     double mindt, double maxdt, uint32_t step, double* stats);
int or_llb_step_rk23(or_llb* s, double* m, double* dt, double maxErr,
                     double mindt, double maxdt, uint32_t step, double* stats);
/* implicit drivers (solve_am01.go, solve_euler_backward.go); -1 where the
 * reference panics.  AM12 stats[5]: err, attempts, iterations, RHS evals, dt used;
 * BDF-Euler stats[3]: iterator err, iterations, RHS evals. */
int or_llb_step_am12(or_llb* s, double* m, double* dt, double absErr, double relErr,
                     double mindt, double maxdt, uint32_t step, double* stats);
int or_llb_step_bdf_euler(or_llb* s, double* m, double dt, uint32_t step, double maxIterErr, int maxIter,
                          double* stats);
/* Two-temperature model with laser source (TM.go dT/dt = Q/C; Qspat.go,
 * Qinter.go; mod_ce.go Ce = gammaE Te):
 *   Qe = kappa lap(Te) + G (Tp - Te) + Q0 exp(-r^2/(2 sr^2)) exp(-(t-t0)^2/(2 st^2))
 *   dTe/dt = Qe / (gammaE Te),  dTp/dt = G (Te - Tp) / Cl
 * r = in-plane distance of the cell centre (user x = internal axis 2, y = 1)
 * from (xc, yc).  Written for the coupled LLB-r run of configs with a laser. */
typedef struct {
    double gammaE, Cl, G, kappa, Q0, sigma_r, sigma_t, t0, xc, yc, maxErrT;
    double Td, nD;  /* Td > 0: Cp(Tp) = 9 nD kB Debye(Tp/Td) (mod_cp.go:75-85, Cp.cu) in place of Cl */
    int spin_heat;  /* Qe += ws - qs, the spin system's energy flow and dissipation (laser-demag.py:70-88) */
} or_tm_params;
/* the spin term of Qe per cell, ws - qs, at the state of the last or_llb_rhs(s, m, ...) call */
void or_llb_spin_heat(const or_llb* s, const double* m, double* Qs);
/* Qs: the spin term of Qe per cell (or NULL) */
void or_tm_rhs(const or_tm_params* q, const int N[3], const double cell[3], const int pbc[3], const double* Te,
               const double* Tp, double t, const double* Qs, double* dTe, double* dTp, double* work);
/* One adaptive RK23 step of (m, Te, Tp) with the LLB temperature T = Te per
 * cell (the LLB's T mask and TMul are replaced), laser time *t advanced on
 * acceptance; the reference's multi-equation controller (solve_rk12.go:150-190)
 * with maxErr for m and q->maxErrT for Te and Tp.  stats[0..2] = errors of
 * the accepted attempt (m, Te, Tp), stats[3] = attempts. */
int or_llbtm_step_rk23(or_llb* s, const or_tm_params* q, double* m, double* Te, double* Tp, double* t,
                       double* dtio, double maxErr, double mindt, double maxdt, uint32_t step, double* stats);
/* One AM12 step of (Te, Tp, m) (solve_am01.go:45-300) with T = Te; absErr, relErr per
 * equation in the order Te, Tp, m; stats[6]: errors (Te, Tp, m), attempts, iterations, dt used. */
int or_llbtm_step_am12(or_llb* s, const or_tm_params* q, double* m, double* Te, double* Tp, double* t,
                       double* dtio, const double absErr[3], const double relErr[3], double mindt, double maxdt,
                       uint32_t step, double* stats);

/* energy term: weight*Ms0T0Mul*sum((mMul.m).(HMul.H)*msat) -- mod_energy.go:98-116 */
double or_energy_term(const double* m, const double* H, const double mMul[3], const double HMul[3],
                      const double* msat, double msatMul, double weight, long N);
void or_llb_energy(const or_llb* s, const double* m, double E[5]);

void or_set_threads(int n);
int or_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
