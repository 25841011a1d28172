/*
 * hotspin_amd.h -- entry points of libhotspin_amd.so beyond the legacy
 * libhotspin ABI (libhotspin.h).  Plain C: pointers, sizes, scalars.
 *
 * They cover what the reference does outside libhotspin today and what the
 * MI355X path adds:
 *   - error reporting (the reference aborts or drops launch errors);
 *   - the demag plan (replaces FFTPlanX + cuFFT and MaxwellPlan.LoadKernel,
 *     src/hotspin-core/gpu/fftplanX.go:36-104,
 *     src/hotspin-core/modules/maxwell_convolution.go:42-223);
 *   - counter-based thermal noise (replaces the cuRAND Philox host generator,
 *     src/hotspin-core/modules/mod_brown.go:98-122);
 *   - the resident LLB-r solver: FP32 fields, FP64 state and FP64
 *     reductions, hand-written HIP stages (replaces the engine's per-step
 *     updater walk, src/hotspin-core/engine/solve_rk12.go:91-214 and
 *     src/hotspin-core/engine/quant.go:417-445).
 */
#ifndef HOTSPIN_AMD_H
#define HOTSPIN_AMD_H

#include <stdint.h>

#include "libhotspin.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- errors */
/* 0 = no error since the last hs_clear_error(); otherwise a hipError_t code
 * or one of the HS_E* codes below.  Sticky: the first error is kept. */
#define HS_EINVAL 10001 /* invalid argument (sizes, NULL where not allowed) */
#define HS_ESIZE 10002  /* size not supported (a padded length, or half the padded x length, above 4095) */
#define HS_ECONV 10003  /* adaptive step did not converge within 10 retries */
#define HS_ENOMEM 10004
HS_API int hs_last_error(void);
HS_API const char* hs_last_error_string(void);
HS_API void hs_clear_error(void);
HS_API const char* hs_version(void);

/* ---------------------------------------------------------------- demag plan */
typedef struct hs_demag hs_demag;
/* Build the demag plan of an N0 x N1 x N2 mesh (internal order), cell sizes
 * (c0,c1,c2) in metres, periodic image counts pbc (0 = open, zero padded):
 * Kernel_Arne (kernel_magnetostatic.go:25-241) integrated on the GPU in FP64,
 * transformed in FP64 by the library's own FFT, real octant kept, scaled by
 * 1/(P0 P1 P2) and stored in FP64 and FP32.  accuracy = 6 in the reference
 * (maxwell_convolution.go:96).  Blocks until the kernel is ready.
 * Any padded lengths up to 4095 (half the padded x length when it is even):
 * 2^k, 3*2^k and 5*2^k run compiled, tuned passes; other lengths (factors 7,
 * 9, 25, larger primes, odd periodic x lengths, one-cell axes) run passes
 * that take the length at run time, as cuFFT plans do in the reference
 * (gpu/fftplanX.go:55-60).  Returns NULL on error (hs_last_error). */
HS_API hs_demag* hs_demag_create(const int N[3], const double cell[3], const int pbc[3], int accuracy,
                                 CUstream stream);
/* The same plan with the Newell demag tensor instead of Kernel_Arne: no
 * counterpart in the reference (which integrates Kernel_Arne only); for
 * callers that want the closed-form tensor of A. J. Newell, W. Williams and
 * D. J. Dunlop, J. Geophys. Res. 98 (1993) 9551.  Second differences of the
 * f and g functions in FP64 up to |R| = 35 hmax^0.6 hmin^0.4, the dipole
 * tensor averaged over both cells to O(h^4/R^4) beyond (a few 1e-6 relative
 * to the exact tensor at the switch; Kernel_Arne at accuracy 6 differs from
 * it by up to 1e-3 of the kernel's largest value).  Same image sums, layout
 * and passes as hs_demag_create. */
HS_API hs_demag* hs_demag_create_newell(const int N[3], const double cell[3], const int pbc[3], CUstream stream);
HS_API void hs_demag_destroy(hs_demag* d);
/* P = logic (padded) size, K = kernel octant size (P0/2+1, P1/2+1, P2/2+1) */
HS_API int hs_demag_sizes(const hs_demag* d, int P[3], int K[3]);
/* Copy the FP64 kernel octants (6 planes XX,YY,ZZ,YZ,XZ,XY of K0*K1*K2) to host. */
HS_API int hs_demag_kernel_f64(const hs_demag* d, double* host_out);
/* max |imag| / max |real| of the transformed kernel (the reference panics
 * above 1e-12, maxwell_convolution.go:383-385) */
HS_API double hs_demag_max_imag_ratio(const hs_demag* d);
/* FP32 tile shapes the plan's autotuner picked: (W, NT) of the row passes
 * (F1/I1), the column passes (F2/I2) and the axis-0 product pass (F3K) */
HS_API int hs_demag_tiles(const hs_demag* d, int* out6);
/* Number of registered FP32 variants of the row, column and axis-0 passes
 * for this plan's lengths (the environment variable HS_DEMAG_FORCE="r,c,z"
 * selects variants by index at plan creation instead of timing them).  The
 * variants of a pass round differently (FP32, operation order); the timed
 * choice is kept per mesh shape for the process, and HS_TUNE_FILE=<path>
 * persists it across processes (read before timing, appended after), so
 * runs that share the file give the same bits. */
HS_API int hs_demag_variant_count(const hs_demag* d, int* out3);
/* B = mu0 * msat0T0Mul * (K conv (m .* msat0T0)), maxwell_convolution.go:103-131.
 * m, B: 3 device planes of N cells (m and B may not alias); msat0T0: device
 * mask or NULL.  FP64 variant: FP64 FFT passes; FP32 variant: FP32 passes. */
HS_API int hs_demag_field_f64(hs_demag* d, const double* m, const double* msat0T0, double msat0T0Mul,
                              double* B, CUstream stream);
HS_API int hs_demag_field_f32(hs_demag* d, const float* m, const float* msat0T0, double msat0T0Mul,
                              float* B, CUstream stream);
/* Grow the plan's FP32 spectrum buffers to hold nsolve solves of the mesh
 * (batched solves share one plan and one kernel). */
HS_API int hs_demag_reserve(hs_demag* d, int nsolve);
/* hs_demag_field_f64 without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The FP64
 * spectra live in a 64 MiB static device array, one per GPU, so the mesh must
 * satisfy 48 * N0 * (N1 + P1) * H2p <= 64 MiB (H2p = P2/2 + 1 rounded up to a
 * multiple of 16; 256-byte alignment aside; plus 48 * P0 * P1 * H2p when P0
 * has no compiled axis-0 pass); larger meshes get HS_ESIZE.
 * Results equal hs_demag_field_f64's bit for bit.  Calls that overlap on one
 * GPU share that array and race; serialise them. */
HS_API int hs_demag_field_f64_noalloc(hs_demag* d, const double* m, const double* msat0T0, double msat0T0Mul,
                                      double* B, CUstream stream);

/* ---------------------------------------------------------------- thermal noise */
/* Philox4x32-10 (Salmon et al., SC'11) keyed by seed, counter
 * (cell, step, attempt, stream); Box-Muller on 24-bit uniforms
 * u = ((x>>8)+0.5)/2^24.  xi: 3 device planes of N standard normals
 * (component c from Box-Muller pair c/2).  Replaces
 * curandGenerateNormalDouble(3N) of mod_brown.go:122. */
HS_API void hs_thermal_normals_f64(double* xi, int N, uint64_t seed, uint32_t stream, uint32_t step,
                                   uint32_t attempt, CUstream s);
/* The FP32 normals the resident solver's stage kernels draw in registers
 * (same Philox words, Box-Muller from the hardware log2 / sqrt / sin / cos):
 * within ~1e-6 absolute of hs_thermal_normals_f64. */
HS_API void hs_thermal_normals_f32(float* xi, int N, uint64_t seed, uint32_t stream, uint32_t step,
                                   uint32_t attempt, CUstream s);
/* raw Philox words for n counters (device arrays, 4 words each) */
HS_API void hs_philox4x32_10(uint32_t* out, const uint32_t* ctr, int n, uint64_t key, CUstream s);

/* ---------------------------------------------------------------- LLB-r solver */
/* Module composition of the reference's LLB-r recipes (examples/p2f.py,
 * tests/stdproblem4-llbr-modular.py): H_eff = H_ex + B/mu0 + B_ext/mu0 + H_lf
 * + H_therm, dm/dt = gamma_LL (m x H_eff + alpha .* H_eff), optional
 * normalizer (LLG limit).  Host pointers; masks may be NULL (uniform). */
typedef struct {
    int N[3];
    double cell[3];
    int pbc[3];
    double msat0T0Mul;
    const double* msat0T0;
    int exchange;
    double lexMul;
    const double* lex;
    int demag;
    int demag_accuracy;
    int zeeman;
    double Bext[3];
    int longfield;
    double JMul, nMul, TcMul, TMul;
    const double *J, *n, *Tc, *T;
    int thermal;
    uint64_t seed;
    uint32_t stream;
    double cutoff_dt;
    double gammaLL;
    int torque;
    int local00nc;
    double alphaMul[3];
    const double* alpha[3];
    int normalize;
    int demag_kernel; /* 0 = Kernel_Arne (demag_accuracy), 1 = Newell (hs_demag_create_newell) */
} hs_llb_params;

#define HS_EULER 0
#define HS_HEUN 1
#define HS_RK4 2
#define HS_RK12 3
#define HS_RK23 4
#define HS_AM12 5      /* solver/am12: implicit-Euler predictor + trapezoidal corrector, adaptive */
#define HS_BDF_EULER 6 /* solver/bdf-euler: fixed-step backward Euler */

typedef struct {
    long steps;          /* accepted steps */
    long rejected;       /* retried attempts */
    long rhs_evals;      /* field evaluations (demag convolutions) */
    double t;            /* simulated time advanced */
    double last_err;     /* error estimate of the last accepted step */
    double dt;           /* dt for the next step */
    long iterations;     /* fixed-point iterations (implicit drivers) */
} hs_run_stats;

typedef struct hs_solver hs_solver;
HS_API hs_solver* hs_solver_create(const hs_llb_params* p, CUstream stream);
/* A batch of nsolve independent solves of one mesh and material, stepped by
 * the same launches: per-solve temperature multiplier TMul[q], thermal seed
 * seed[q] and applied field Bext[3q..3q+2] (tesla, internal order); NULL
 * arrays take the value of p for every solve.  State arrays of the batch
 * (set_m / get_m / rhs) are nsolve x 3 x N, solve-major.  stream NULL: the
 * solver creates its own stream. */
HS_API hs_solver* hs_batch_create(const hs_llb_params* p, int nsolve, const double* TMul, const uint64_t* seed,
                                  const double* Bext, CUstream stream);
HS_API int hs_solver_nsolve(hs_solver* s);
/* Fixed-step RK4 of all solves replayed from a captured hipGraph (one step
 * captured once, then launched nsteps times; the thermal-noise step counter
 * advances on the device). */
HS_API int hs_solver_run_graph(hs_solver* s, long nsteps, long step0, double dt);
HS_API void hs_solver_destroy(hs_solver* s);
/* state m: 3 planes of N doubles on the host (internal order) */
HS_API int hs_solver_set_m(hs_solver* s, const double* m_host);
HS_API int hs_solver_get_m(hs_solver* s, double* m_host);
/* Advance nsteps accepted steps.  Fixed-step methods use *dt; adaptive
 * methods (RK12, RK23) update *dt with the reference's controller
 * (solve_rk12.go:121-210).  step0 is the global step counter of the first
 * step (it keys the thermal noise).  stats may be NULL. */
HS_API int hs_solver_run(hs_solver* s, int method, long nsteps, long step0, double* dt, double maxErr,
                         double mindt, double maxdt, hs_run_stats* stats);
/* Tolerances of the implicit drivers, the reference's per-equation
 * quantities: AM12 m_maxAbsError and m_maxRelError (defaults 1e-4, 1e-3,
 * engine/solve_am01.go:380-384), backward Euler m_maxIterError and
 * m_maxIterations (1e-5, 3, engine/solve_euler_backward.go:128-129).
 * hs_solver_run with HS_AM12 adapts *dt like solve_am01.go:45-300 (its
 * maxErr argument is unused); HS_BDF_EULER keeps *dt.  Both treat a batch as
 * one system (maxima over every solve) and fail with HS_ECONV where the
 * reference panics.  A two-temperature solve takes HS_AM12 (not
 * HS_BDF_EULER) for its three equations in the laser recipe's order Te, Tp, m
 * (examples/laser-demag.py:35-60), with Te_/Tp_maxAbsError and
 * _maxRelError from hs_solver_set_implicit_T (defaults 1e-4, 1e-3). */
HS_API int hs_solver_set_implicit(hs_solver* s, double maxAbsErr, double maxRelErr, double maxIterErr,
                                  int maxIter);
HS_API int hs_solver_set_implicit_T(hs_solver* s, double maxAbsErr, double maxRelErr);
/* Energy terms of the current state, per solve: E[5 q + t] with t = 0 Zeeman
 * (B_ext, weight -V), 1 exchange (H_ex, -V mu0 / 2), 2 demag (B, -V / 2),
 * 3 longitudinal field (H_lf, -V mu0), 4 their sum, each
 * weight * Ms0T0Mul * sum_cells (m . F) msat0T0 (LoadEnergy and
 * EnergyUpdater.Update, modules/mod_energy.go:25-61,98-116; a term whose
 * field is not loaded is 0).  FP64 m, H_ex and H_lf; B from the FP32 demag
 * of the state; deterministic FP64 sums.  E: 5 * nsolve doubles (host).
 * Two-temperature solves use T = Te (its FP32 copy, as the stages do); a
 * handle in slot mode is refused (HS_EINVAL). */
HS_API int hs_solver_energy(hs_solver* s, double* E);
/* One RHS evaluation at the current state (no step): k and H_eff to host. */
HS_API int hs_solver_rhs(hs_solver* s, double* k_host, double* heff_host, uint32_t step, uint32_t attempt,
                         double dt);
HS_API CUstream hs_solver_stream(hs_solver* s);
/* the solver's demag plan (NULL when demag is off) */
HS_API hs_demag* hs_solver_demag(hs_solver* s);
/* stage-kernel variant picked by timing at creation: 0 = marching planes,
 * 1 = one plane per block, 2 = two cells per thread (even N2 only); 3 = one
 * plane per block with the planes of a tile consecutive on one XCD (only
 * when forced); env HS_STAGE_KERNEL=march|flat|pair|flatz overrides.  Every
 * variant does the same arithmetic per cell, so results do not depend on
 * the choice. */
HS_API int hs_solver_stage_variant(hs_solver* s);

/* 1 when the solver's LLB stage launches take the lean kernel form (no
 * per-cell masks, no two-temperature coupling, longitudinal field off or
 * uniform: the masks' loads and the general longitudinal-field path are
 * compiled out), 0 for the general form; env HS_STAGE_LEAN=0 at creation
 * forces the general form.  Both forms give the same result. */
HS_API int hs_solver_stage_lean(hs_solver* s);

/* ------------------------------------------------ two-temperature coupling
 * Laser-heated LLB-r: the LLB temperature is the electron temperature Te of
 * each cell, and Te, Tp follow the two-temperature model of the reference's
 * temperature modules (modules/TM.go dT/dt = Q/C, Qspat.go heat diffusion
 * with harmonic-mean conductances (Qspat.cu:13-134), Qinter.go electron-
 * phonon exchange G (Tp - Te) (Qinter.cu:13-45), mod_ce.go Drude heat
 * capacity Ce = gamma Te):
 *   gamma Te dTe/dt = kappa lap(Te) + G (Tp - Te) + Q_laser(x, y, t)
 *   Cl dTp/dt       = G (Te - Tp)
 * Q_laser = Q0 exp(-((x-xc)^2 + (y-yc)^2) / (2 sigma_r^2)) exp(-(t-t0)^2 / (2 sigma_t^2))
 * with x, y the user in-plane coordinates of the cell centre (internal axes
 * 2 and 1), uniform through the thickness.  Te, Tp are FP64 state, their
 * stage values reach the LLB fields as FP32.  hs_solver_run with HS_RK23
 * integrates (m, Te, Tp) together: one error per equation, the reference's
 * multi-equation controller (solve_rk12.go:150-190: a step is bad if any
 * equation is, the factor is the smallest of the equations'), maxErr for m,
 * maxErrT for Te and Tp. */
typedef struct {
    double gammaE;  /* electron heat capacity coefficient, J/(m^3 K^2) */
    double Cl;      /* lattice heat capacity, J/(m^3 K) */
    double G;       /* electron-phonon coupling, W/(m^3 K) */
    double kappa;   /* electron thermal conductivity, W/(m K) */
    double Q0;      /* laser peak power density, W/m^3 */
    double sigma_r, sigma_t, t0, xc, yc;
    double maxErrT; /* K */
    /* Td > 0: lattice heat capacity Cp(Tp) = 9 nD kB (Tp/Td)^3 int_0^{Td/Tp} x^4 e^x/(e^x-1)^2 dx
     * (temperature/Debye, mod_cp.go:75-85 with Cp.cu's 61-interval rule) in place of Cl */
    double Td, nD;
    /* 1: Qe also takes the spin system's energy flow minus its dissipation,
     * ws - qs with ws = 3 kB Tc n gammaLL S/(S+1) (m . R) and
     * qs = -mu0 Ms0T0 gammaLL (H_eff . R), R = diag(alpha) H_eff the local
     * damping term, zero when local00nc is off (laser-demag.py:70-88;
     * mod_molecular-energy-density-rate.go, mod_energy-density-dissipation-rate.go).
     * The LLB stage forms it from its FP32 field; Cl may be 0 when Td > 0. */
    int spin_heat;
} hs_tm_params;
/* Resident solver with two-temperature coupling (the LLB T and T mask of p
 * are ignored: T = Te per cell).  Te, Tp start uniform at T0. */
HS_API hs_solver* hs_solver_create_tm(const hs_llb_params* p, const hs_tm_params* tm, double T0, CUstream stream);
/* Te, Tp: N doubles each on the host (either may be NULL) */
HS_API int hs_solver_set_T(hs_solver* s, const double* Te, const double* Tp);
HS_API int hs_solver_get_T(hs_solver* s, double* Te, double* Tp);
/* simulated time of the laser clock (seconds); hs_solver_run advances it */
HS_API int hs_solver_set_time(hs_solver* s, double t);
HS_API double hs_solver_time(hs_solver* s);
/* dTe/dt and dTp/dt at the current state and laser time t (testing) */
HS_API int hs_solver_tm_rhs(hs_solver* s, double t, double* dTe_host, double* dTp_host);

/* ------------------------------------------------ slots (continuous batching)
 * No counterpart in the reference (it steps one solve per process).  The
 * nsolve solves of a batch become slots that are filled, stepped and emptied
 * independently: each running slot takes adaptive RK23 steps (Bogacki-
 * Shampine, the reference's controller of solve_rk12.go:164-198 with
 * exponent 1/3, at most 10 retries) with its own dt, accept/reject decision
 * and step count, decided on the device so that many iterations are
 * enqueued (as a captured hipGraph) without host round trips.  Idle, done
 * and failed slots are skipped by the stage kernels.  The whole-batch calls
 * above (hs_solver_run, _run_graph, _set_m, _get_m) must not be used on a
 * handle once a slot has been admitted. */
#define HS_SLOT_IDLE 0
#define HS_SLOT_RUNNING 1
#define HS_SLOT_DONE 2
#define HS_SLOT_FAILED 3 /* more than 10 retries of one step */

typedef struct {
    uint32_t step;     /* thermal-noise step counter of the step in progress */
    uint32_t attempt;  /* retry number of that step */
    double h;          /* dt of the attempt in progress (next dt when done) */
    double t;          /* simulated time advanced since admission */
    double err;        /* error estimate of the last attempt */
    double maxErr, mindt, maxdt;
    uint32_t end;      /* step counter at which the slot is done */
    int32_t state;     /* HS_SLOT_* */
    uint32_t rejected; /* rejected attempts since admission */
    uint32_t accepted; /* accepted steps since admission */
    uint32_t home;     /* internal: buffer set that holds a finished state */
    uint32_t retry;    /* internal: last attempt was rejected */
    double maxErrT;    /* two-temperature solves: error bound of Te and Tp (K) */
} hs_slot;

/* Fill slot `slot` (idle, done or failed) with a new solve: m_host is 3 x N
 * doubles (internal order); per-solve TMul, thermal seed and noise stream,
 * Bext (tesla, 3 values, NULL = the batch's); nsteps accepted steps from
 * step counter step0 starting at dt.  Synchronous on the solver's stream. */
HS_API int hs_slot_admit(hs_solver* s, int slot, const double* m_host, double TMul, uint64_t seed, uint32_t stream,
                         const double* Bext, long step0, long nsteps, double dt, double maxErr, double mindt,
                         double maxdt);
/* Enqueue `iters` RK23 iterations (one attempt of every running slot each)
 * and a copy of the slot records to host memory; returns without waiting. */
HS_API int hs_slot_advance(hs_solver* s, int iters);
/* 1 when the work enqueued by hs_slot_advance has finished, 0 if not. */
HS_API int hs_slot_ready(hs_solver* s);
/* Wait for hs_slot_advance's work and copy the nsolve slot records out. */
HS_API int hs_slot_status(hs_solver* s, hs_slot* out);
/* State of one slot (3 x N doubles); waits for the solver's stream. */
HS_API int hs_slot_get_m(hs_solver* s, int slot, double* m_host);
/* Mark a slot idle (its result has been read). */
HS_API int hs_slot_release(hs_solver* s, int slot);
/* Device bytes held by a solver handle (state, stage buffers, demag plan). */
HS_API long hs_solver_footprint(hs_solver* s);

/* ------------------------------------------------ ensembles of uneven meshes
 * No counterpart in the reference.  Several batches (one per mesh shape)
 * stepped together: one fixed-step RK4 step of every solve of every batch is
 * captured as one hipGraph whose branches run on the batches' own streams,
 * so a step of the whole ensemble is one graph launch. */
typedef struct hs_ensemble hs_ensemble;
HS_API hs_ensemble* hs_ensemble_create(hs_solver* const* batches, int nbatch);
HS_API int hs_ensemble_run_graph(hs_ensemble* e, long nsteps, long step0, double dt);
HS_API void hs_ensemble_destroy(hs_ensemble* e);

/* ---------------------------------------------------------------- timing */
/* Per-kernel durations from HIP events recorded on the launching stream
 * around each instrumented launch (off by default).  Categories: */
#define HS_PROF_F1 0    /* demag R2C rows (axis 2) */
#define HS_PROF_F2 1    /* demag forward columns (axis 1) */
#define HS_PROF_F3K 2   /* demag axis-0 FFT + kernel product + inverse */
#define HS_PROF_I2 3    /* demag inverse columns (axis 1) */
#define HS_PROF_I1 4    /* demag C2R rows (axis 2) */
#define HS_PROF_STAGE 5 /* fused LLB-r field + torque + RK stage kernel */
#define HS_PROF_NCAT 8
/* on != 0: start recording (clears earlier records); 0: stop */
HS_API void hs_prof_enable(int on);
/* synchronises the device; total milliseconds and launch counts per category */
HS_API int hs_prof_read(double* total_ms, long* count, int ncat);

#ifdef __cplusplus
}
#endif
#endif /* HOTSPIN_AMD_H */
