/*
 * libhotspin.h -- drop-in C ABI of libhotspin_amd.so for hotspin's Go core.
 *
 * Same symbol names, argument order and argument meaning as the reference's
 * libhotspin (the concatenated per-kernel headers of src/libhotspin, Makefile
 * :31-38), re-implemented as hand-written HIP kernels for gfx950 (MI355X).
 * Each declaration cites the reference header it replaces.
 *
 * Conventions kept from the reference:
 *   - device pointers to FP64 SoA planes (one plane per vector component);
 *   - a NULL mask means uniform (1.0; 0.0 for the demag kernel octants);
 *   - internal axes (N0,N1,N2) = user (z,y,x), index i*N1*N2 + j*N2 + k;
 *   - every call is asynchronous on the given stream and returns void;
 *     output aliasing of inputs is allowed where the reference allows it.
 * CUstream is an opaque handle: pass a hipStream_t (0 = the null stream).
 * Launch and argument errors are recorded and readable via hs_last_error()
 * (hotspin_amd.h) instead of being dropped as in the reference.
 */
#ifndef LIBHOTSPIN_H
#define LIBHOTSPIN_H

#ifdef __cplusplus
extern "C" {
#endif

#ifndef HS_API
#define HS_API __attribute__((visibility("default")))
#endif

typedef struct CUstream_st* CUstream;

/* ---- exchange6.h:20 */
HS_API void exchange6Async(double* hx, double* hy, double* hz,
                           double* mx, double* my, double* mz,
                           double* msat0T0,
                           double* lex,
                           int N0, int N1Part, int N2,
                           int periodic0, int periodic1, int periodic2,
                           double msat0T0Mul,
                           double lex2Mul_cellSizeX2, double lex2Mul_cellSizeY2, double lex2Mul_cellSizeZ2,
                           CUstream streams);

/* ---- llbar-torque.h:20 */
HS_API void llbar_torque_async(double* tx, double* ty, double* tz,
                               double* Mx, double* My, double* Mz,
                               double* hx, double* hy, double* hz,
                               double* msat0T0,
                               CUstream stream, int Npart);

/* ---- llbar-local00nc.h:20 */
HS_API void llbar_local00nc_async(double* tx, double* ty, double* tz,
                                  double* hx, double* hy, double* hz,
                                  double* msat0T0,
                                  double* lambda_xx, double* lambda_yy, double* lambda_zz,
                                  const double lambdaMul_xx, const double lambdaMul_yy, const double lambdaMul_zz,
                                  CUstream stream, int Npart);

/* ---- llbar-local02nc.h:19 */
HS_API void llbar_local02nc_async(double* tx, double* ty, double* tz,
                                  double* mx, double* my, double* mz,
                                  double* hx, double* hy, double* hz,
                                  double* msat0T0,
                                  double* mu_xx, double* mu_yy, double* mu_zz,
                                  const double muMul_xx, const double muMul_yy, const double muMul_zz,
                                  CUstream stream, int Npart);

/* ---- llbar-local02c.h:19 */
HS_API void llbar_local02c_async(double* tx, double* ty, double* tz,
                                 double* mx, double* my, double* mz,
                                 double* hx, double* hy, double* hz,
                                 double* msat0T0,
                                 double* mu_xx, double* mu_yy, double* mu_zz,
                                 const double muMul_xx, const double muMul_yy, const double muMul_zz,
                                 CUstream stream, int Npart);

/* ---- long_field.h:20 */
HS_API void long_field_async(double* hx, double* hy, double* hz,
                             double* mx, double* my, double* mz,
                             double* msat0T0, double* S, double* n, double* Tc, double* Ts,
                             double msat0T0Mul, double SMul, double nMul, double TcMul, double TsMul,
                             int NPart, CUstream stream);

/* ---- temperature.h:19 */
HS_API void temperature_scaleAnizNoise(double* hx, double* hy, double* hz,
                                       double* mu_xx, double* mu_yy, double* mu_zz,
                                       double* tempMask, double* msat0T0Mask,
                                       double muMul_xx, double muMul_yy, double muMul_zz,
                                       double KB2tempMul_mu0VgammaDtMSatMul,
                                       CUstream stream, int Npart);

/* ---- brillouin.h:18 */
HS_API void brillouinAsync(double* msat0, double* msat0T0, double* T, double* Tc, double* S,
                           const double msat0Mul, const double msat0T0Mul,
                           const double TcMul, const double SMul,
                           int Npart, CUstream stream);

/* ---- normalize.h:22 */
HS_API void normalizeAsync(double* mx, double* my, double* mz, CUstream stream, int Npart);

/* ---- Qinter.h:22 */
HS_API void QinterAsync(double* Qi, double* Ti, double* Tj, double* Gij, double GijMul,
                        int Npart, CUstream stream);

/* ---- Qspat.h:22 */
HS_API void Qspat_async(double* Q, double* T, double* k, const double kMul,
                        const int sx, const int sy, const int sz,
                        const double csx, const double csy, const double csz,
                        const int pbc_x, const int pbc_y, const int pbc_z,
                        CUstream stream);

/* ---- add.h:20-61 */
HS_API void addAsync(double* dst, double* a, double* b, CUstream stream, int Npart);
HS_API void linearCombination2Async(double* dst, double* a, double mulA, double* b, double mulB,
                                    CUstream stream, int NPart);
HS_API void linearCombination3Async(double* dst, double* a, double mulA, double* b, double mulB,
                                    double* c, double mulC, CUstream stream, int NPart);
HS_API void addMaddAsync(double* dst, double* a, double* b, double* c, double mul, CUstream stream, int NPart);
HS_API void madd1Async(double* a, double* b, double mulB, CUstream stream, int Npart);
HS_API void madd2Async(double* a, double* b, double mulB, double* c, double mulC, CUstream stream, int Npart);
HS_API void maddAsync(double* dst, double* a, double* b, double mulB, CUstream stream, int Npart);
HS_API void vecMaddAsync(double* dstx, double* dsty, double* dstz,
                         double* ax, double* ay, double* az,
                         double* bx, double* by, double* bz,
                         double mulBx, double mulBy, double mulBz,
                         CUstream stream, int Npart);
HS_API void cmaddAsync(double* dst, double a, double b, double* kern, double* src, CUstream stream,
                       int NComplexPart);

/* ---- mul.h:19-21 */
HS_API void mulAsync(double* dst, double* a, double* b, CUstream stream, int Npart);
HS_API void tensSYMMVecMul(double* dstX, double* dstY, double* dstZ,
                           double* srcX, double* srcY, double* srcZ,
                           double* kernXX, double* kernYY, double* kernZZ,
                           double* kernYZ, double* kernXZ, double* kernXY,
                           const double srcMul,
                           const int Nx, const int Ny, const int Nz,
                           CUstream stream);

/* ---- div.h:20 */
HS_API void divAsync(double* dst, double* a, double* b, CUstream stream, int Npart);

/* ---- divMulPow.h:21: dst = a/b * pow(c, p), the power taken in single
 * precision as the reference does (divMulPow.cu:27). */
HS_API void divMulPowAsync(double* dst, double* a, double* b, double* c, double p, CUstream stream, int Npart);

/* ---- wavg.h:20-28 */
HS_API void wavgAsync(double* dst, double* a, double* b, double* w0, double* w1, double* R,
                      double w0Mul, double w1Mul, double RMul, CUstream stream, int Npart);

/* ---- kappa.h:20-32 (MFA longitudinal susceptibility) */
HS_API void kappaAsync(double* kappa, double* msat0, double* msat0T0, double* T, double* Tc, double* S,
                       double* n, const double msat0Mul, const double msat0T0Mul, const double TcMul,
                       const double SMul, int Npart, CUstream stream);

/* ---- Ts.h:20-30 (spin temperature from |m|, inverse Brillouin) */
HS_API void tsAsync(double* Ts, double* msat, double* msat0T0, double* Tc, double* S, const double msatMul,
                    const double msat0T0Mul, const double TcMul, const double SMul, int Npart,
                    CUstream stream);

/* ---- Cp.h:22-28 (Debye lattice heat capacity) */
HS_API void cpAsync(double* Cp, double* T, double* Td, double* n, const double TdMul, const int Npart,
                    CUstream stream);

/* ---- energy_flow.h:20-30 */
HS_API void energyFlowAsync(double* w, double* mx, double* my, double* mz, double* Rx, double* Ry, double* Rz,
                            double* Tc, double* S, double* n, double SMul, int Npart, CUstream stream);

/* ---- uniaxialanisotropy.h:21-27 */
HS_API void uniaxialAnisotropyAsync(double* hx, double* hy, double* hz, double* mx, double* my, double* mz,
                                    double* Ku1_map, double* MSat_map, double Ku2_Mu0Msat_mul,
                                    double* anisU_mapx, double anisU_mulx, double* anisU_mapy,
                                    double anisU_muly, double* anisU_mapz, double anisU_mulz, CUstream stream,
                                    int Npart);

/* ---- llbar-nonlocal00nc.h:21-34 (-lambda_e Laplacian(H); the +1 neighbour
 * along axis 1 uses the index i*N1*N2 + (j+1)*N2 + k, where the reference
 * computes i*N1*N1 + (j+1)*N1 + k, llbar-nonlocal00nc.cu:183; both agree
 * when N1 == N2). */
HS_API void llbar_nonlocal00nc_async(double* tx, double* ty, double* tz, double* hx, double* hy, double* hz,
                                     double* msat0T0, double* lambda_e_xx, double* lambda_e_yy,
                                     double* lambda_e_zz, const double lambda_eMul_xx,
                                     const double lambda_eMul_yy, const double lambda_eMul_zz, const int sx,
                                     const int sy, const int sz, const double csx, const double csy,
                                     const double csz, const int pbc_x, const int pbc_y, const int pbc_z,
                                     CUstream stream);

/* ---- dot.h:21-31 */
HS_API void dotMaskAsync(double* dst,
                         double* ax, double* ay, double* az,
                         double* bx, double* by, double* bz,
                         double axMul, double ayMul, double azMul,
                         double bxMul, double byMul, double bzMul,
                         CUstream stream, int Npart);
HS_API void dotAsync(double* dst, double* ax, double* ay, double* az, double* bx, double* by, double* bz,
                     CUstream stream, int Npart);
HS_API void dotSignAsync(double* dst, double* ax, double* ay, double* az, double* bx, double* by, double* bz,
                         double* cx, double* cy, double* cz, CUstream stream, int Npart);

/* ---- copypad.h:28-33 */
HS_API void copyPad3DAsync(double* dst, int D0, int D1, int D2, double* src, int S0, int S1, int S2,
                           int Ncomp, CUstream streams);
HS_API void copyUnPad3DAsync(double* dst, int D0, int D1, int D2, double* src, int S0, int S1, int S2,
                             int Ncomp, CUstream streams);
HS_API void zeroArrayAsync(double* A, int N, CUstream streams);

/* ---- reduce.h:34-119: one partial per block into output[0..blocks); the
 * caller finishes on the host (gpu/reductor.go:69-202).  The block count is
 * honoured exactly; the thread count is a hint (the kernel uses its own). */
HS_API void partialSumAsync(double* input, double* output, int blocksPerGPU, int threadsPerBlockPerGPU,
                            int NPerGPU, CUstream streams);
HS_API void partialSDotAsync(double* input1, double* input2, double* output, int blocks, int threadsPerBlock,
                             int N, CUstream stream);
HS_API void partialMaxAsync(double* input, double* output, int blocksPerGPU, int threadsPerBlockPerGPU,
                            int NPerGPU, CUstream streams);
HS_API void partialMinAsync(double* input, double* output, int blocksPerGPU, int threadsPerBlockPerGPU,
                            int NPerGPU, CUstream streams);
HS_API void partialMaxAbsAsync(double* input, double* output, int blocksPerGPU, int threadsPerBlockPerGPU,
                               int NPerGPU, CUstream streams);
HS_API void partialMaxDiffAsync(double* a, double* b, double* output, int blocksPerGPU, int threadsPerBlockPerGPU,
                                int NPerGPU, CUstream streams);
HS_API void partialMaxSumAsync(double* a, double* b, double* output, int blocksPerGPU, int threadsPerBlockPerGPU,
                               int NPerGPU, CUstream streams);
HS_API void partialMaxNorm3SqAsync(double* x, double* y, double* z, double* output, int blocksPerGPU,
                                   int threadsPerBlockPerGPU, int NPerGPU, CUstream streams);
HS_API void partialMaxNorm3SqDiffAsync(double* x1, double* y1, double* z1, double* x2, double* y2, double* z2,
                                       double* output, int blocksPerGPU, int threadsPerBlockPerGPU, int NPerGPU,
                                       CUstream streams);

#ifdef __cplusplus
}
#endif
#endif /* LIBHOTSPIN_H */
