/*
 * padeops_hip.h -- C ABI of the MI355X (gfx950) implementation of the PadeOps
 * compact-derivative / compact-filter / periodic FFT-Poisson hot path.
 *
 * Every entry point is a plain C function over raw pointers and sizes so that
 * the Fortran solvers can bind it with ISO_C_BINDING (see
 * padeops_amd/fortran/padeops_hip_mod.F90).  Field arguments are DEVICE
 * pointers laid out exactly like the reference's explicit-shape dummies
 * (Fortran column-major, first index fastest):
 *     x-operators  f(n, na, nb)      (dd1, d2d1, filter1)
 *     y-operators  f(na, n, nb)      (dd2, d2d2, filter2)
 *     z-operators  f(na, nb, n)      (dd3, d2d3, filter3, cd06stagg)
 * Complex fields are interleaved (re, im) pairs; because every coefficient is
 * real, a complex z-operator on (na, nb, n) equals the real operator on
 * (2*na, nb, n).
 *
 * Calls are stream-ordered and asynchronous (stream = hipStream_t, NULL for
 * the default stream) and never synchronize, except pdo_sync and the
 * divergence maximum.  Objects hold only device-resident factor tables; the
 * caller owns every field buffer.
 *
 * Return codes: 0 success; the reference's own init codes (2, 3, 7: too few
 * points; 21: cd06stagg needs more than 4 points); 324 for a boundary flag
 * outside {-1,0,1} (the reference aborts with GracefulExit code 324); negative
 * values for errors of this library (pdo_last_error() gives the text).
 */
#ifndef PADEOPS_HIP_H
#define PADEOPS_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* arithmetic/storage modes of an apply call */
enum pdo_prec {
    PDO_F64 = 0,        /* FP64 storage and arithmetic                               */
    PDO_F32_REFINE = 1, /* FP32 storage; FP32 stencil and solves; one refinement
                           step with the residual b - A y formed in FP64            */
    PDO_F32_STORE = 2,  /* FP32 storage, FP64 arithmetic (no refinement needed)      */
    PDO_F32 = 3         /* FP32 storage and arithmetic, no refinement (error study)  */
};

enum pdo_status {
    PDO_OK = 0,
    PDO_ERR_ARG = -1,   /* invalid argument (null handle, size mismatch, ...)        */
    PDO_ERR_HIP = -2,   /* a HIP runtime call failed                                  */
    PDO_ERR_UNSUP = -3, /* combination the reference leaves incomplete               */
    PDO_ERR_NOGPU = -4  /* no HIP device                                              */
};

/* modified-wavenumber models for the Poisson solver (PoissonPeriodic%init
 * Get_ModK callbacks, PoissonPeriodic.F90:114-212) */
enum pdo_modk {
    PDO_MODK_FOURIER = 0, /* exact wavenumbers (no callback)                          */
    PDO_MODK_CD06 = 1,    /* collocated cd06 first derivative                         */
    PDO_MODK_CD06STAGG = 2, /* staggered cd06, getmodCD06stagg PadeDerOps.F90:1034-1053 */
    PDO_MODK_CD10 = 3     /* collocated cd10 first derivative                         */
};

typedef struct pdo_op_s *pdo_op;
typedef struct pdo_poisson_s *pdo_poisson;
typedef struct pdo_rhs_s *pdo_rhs;
typedef struct pdo_spectral_s *pdo_spectral;
typedef struct pdo_padepoisson_s *pdo_padepoisson;
typedef struct pdo_pade6stagg_s *pdo_pade6stagg;
typedef struct pdo_duidxj_s *pdo_duidxj;
typedef struct pdo_dynproc_s *pdo_dynproc;
typedef struct pdo_mlp_s *pdo_mlp;
typedef struct pdo_closure_s *pdo_closure;
typedef struct pdo_xfilter_s *pdo_xfilter;
typedef struct pdo_derivatives_s *pdo_derivatives;
typedef struct pdo_filters_s *pdo_filters;

/* ---- library ------------------------------------------------------------ */
const char *pdo_last_error(void);
int pdo_device_count(int *count);
int pdo_sync(void *stream);
int pdo_malloc(void **ptr, size_t bytes);
int pdo_free(void *ptr);
int pdo_memcpy_h2d(void *dst, const void *src, size_t bytes);
int pdo_memcpy_d2h(void *dst, const void *src, size_t bytes);
/* device-to-device copy, asynchronous on stream (nothing when dst == src):
 * the one-GPU transpose_x_to_y / y_to_z / z_to_y / y_to_x of 2DECOMP&FFT
 * (decomp_2d.f90 transpose_*), whose pencils on one process hold the whole
 * box in the same (i, j, k) order */
int pdo_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream);

/* ---- cd06: 6th-order collocated compact derivative ----------------------
 * replaces type(cd06) of src/derivatives/cd06.F90:
 *   init        cd06.F90:129-207  (ierr 3: periodic n<6; 2: non-periodic n<8)
 *   dd1/dd2/dd3 cd06.F90:775-839  (first derivative along x / y / z)
 * d2d1/d2d2/d2d3: periodic second derivative (RHS cd06.F90:745-773 over
 * dx^2, cyclic solve of the alpha = 2/11 system, cd06.F90:163-170); the
 * reference leaves this operator unwired (derivatives.F90:524-525), so the
 * non-periodic variant returns PDO_ERR_UNSUP.  Non-periodic cd06 supports
 * bc1 = bcn = 0 only (bc 1 is incomplete in cd06.F90:288-292). */
int pdo_cd06_init(pdo_op *h, int n, double dx, int periodic, int bc1, int bcn);
int pdo_cd06_dd1(pdo_op h, int prec, const void *f, void *df, int na, int nb, void *stream);
int pdo_cd06_dd2(pdo_op h, int prec, const void *f, void *df, int na, int nb, void *stream);
int pdo_cd06_dd3(pdo_op h, int prec, const void *f, void *df, int na, int nb, void *stream);
int pdo_cd06_d2d1(pdo_op h, int prec, const void *f, void *df, int na, int nb, void *stream);
int pdo_cd06_d2d2(pdo_op h, int prec, const void *f, void *df, int na, int nb, void *stream);
int pdo_cd06_d2d3(pdo_op h, int prec, const void *f, void *df, int na, int nb, void *stream);

/* ---- cd10: 10th-order collocated compact derivatives ---------------------
 * replaces type(cd10) of src/derivatives/cd10.F90:
 *   init           cd10.F90:195-315  (ierr 2: n<8)
 *   dd1/dd2/dd3    cd10.F90:2029-2237 (bc1/bcn in {0,1,-1}: one-sided closure,
 *                  symmetric, antisymmetric; 324 otherwise)
 *   d2d1/d2d2/d2d3 cd10.F90:2239-2447 */
int pdo_cd10_init(pdo_op *h, int n, double dx, int periodic, int bc1, int bcn);
int pdo_cd10_dd1(pdo_op h, int prec, const void *f, void *df, int na, int nb, int bc1, int bcn, void *stream);
int pdo_cd10_dd2(pdo_op h, int prec, const void *f, void *df, int na, int nb, int bc1, int bcn, void *stream);
int pdo_cd10_dd3(pdo_op h, int prec, const void *f, void *df, int na, int nb, int bc1, int bcn, void *stream);
int pdo_cd10_d2d1(pdo_op h, int prec, const void *f, void *df, int na, int nb, int bc1, int bcn, void *stream);
int pdo_cd10_d2d2(pdo_op h, int prec, const void *f, void *df, int na, int nb, int bc1, int bcn, void *stream);
int pdo_cd10_d2d3(pdo_op h, int prec, const void *f, void *df, int na, int nb, int bc1, int bcn, void *stream);

/* ---- cf90: 8th-order compact filter ---------------------------------------
 * replaces type(cf90) of src/filters/cf90.F90:
 *   init              cf90.F90:107-177 (ierr 7: n<10)
 *   filter1/2/3       cf90.F90:1020-1224
 *   filter3 (complex) cf90.F90:1226-1251; like the reference it always uses
 *                     the non-periodic bc (0,0) system and its RHS for the
 *                     given bc pair, so a periodic object returns PDO_ERR_UNSUP */
int pdo_cf90_init(pdo_op *h, int n, int periodic);
int pdo_cf90_filter1(pdo_op h, int prec, const void *f, void *fil, int na, int nb, int bc1, int bcn, void *stream);
int pdo_cf90_filter2(pdo_op h, int prec, const void *f, void *fil, int na, int nb, int bc1, int bcn, void *stream);
int pdo_cf90_filter3(pdo_op h, int prec, const void *f, void *fil, int na, int nb, int bc1, int bcn, void *stream);
int pdo_cf90_filter3_cmplx(pdo_op h, int prec, const void *f, void *fil, int na, int nb, int bc1, int bcn,
                           void *stream);

/* ---- cd06stagg: staggered 6th-order z operators (periodic) ---------------
 * replaces type(cd06stagg) of src/derivatives/cd06stagg.F90, periodic init
 * (:170-195).  Cell fields have n points, edge fields n+1 (edge n+1 repeats
 * edge 1).  is_complex != 0: interleaved complex (n1 complex points per row).
 *   ddz_E2C :820-845  ddz_C2E :846-873  d2dz2_C2C :1008-1026
 *   d2dz2_E2E :1027-1044  InterpZ_E2C :932-957  InterpZ_C2E :958-985 */
int pdo_cd06stagg_init_periodic(pdo_op *h, int n, double dz);
/* non-periodic init (cd06stagg.F90:197-232): even / odd reflection at each
 * wall, or the one-sided closures where is*Sided != 0 (first derivatives and
 * interpolation; D2 and ddz_E2E always use the even / odd flags).  Returns 21
 * for n <= 4 like the reference; n in 5..7 is not supported here. */
int pdo_cd06stagg_init(pdo_op *h, int n, double dz, int isTopEven, int isBotEven, int isTopSided, int isBotSided);
int pdo_cd06stagg_ddz_E2C(pdo_op h, int prec, const void *fE, void *dfC, int n1, int n2, int is_complex, void *stream);
int pdo_cd06stagg_ddz_C2E(pdo_op h, int prec, const void *fC, void *dfE, int n1, int n2, int is_complex, void *stream);
int pdo_cd06stagg_d2dz2_C2C(pdo_op h, int prec, const void *fC, void *d2fC, int n1, int n2, int is_complex,
                            void *stream);
int pdo_cd06stagg_d2dz2_E2E(pdo_op h, int prec, const void *fE, void *d2fE, int n1, int n2, int is_complex,
                            void *stream);
int pdo_cd06stagg_interpz_E2C(pdo_op h, int prec, const void *fE, void *ifC, int n1, int n2, int is_complex,
                              void *stream);
int pdo_cd06stagg_interpz_C2E(pdo_op h, int prec, const void *fC, void *ifE, int n1, int n2, int is_complex,
                              void *stream);
/* collocated first derivatives of the non-periodic object: ddz_C2C :883-903,
 * ddz_E2E :905-925 (PDO_ERR_UNSUP on a periodic object) */
int pdo_cd06stagg_ddz_C2C(pdo_op h, int prec, const void *fC, void *dfC, int n1, int n2, int is_complex, void *stream);
int pdo_cd06stagg_ddz_E2E(pdo_op h, int prec, const void *fE, void *dfE, int n1, int n2, int is_complex, void *stream);

int pdo_op_destroy(pdo_op h);

/* Allocation-free apply.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP stream capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same kernels and results
 * as the entry points above:
 *   kind   0 dd (first derivative), 1 d2d, 2 filter, 10..17 the staggered
 *          operators (numbering of pdo_op_dense); dir 1 x, 2 y, 3 z with the
 *          layouts of dd1 / dd2 / dd3 (staggered operators: dir 3, and na
 *          counts real words, 2 n1 for complex data);
 *   bc1, bcn as in the cd10 / cf90 calls (ignored by cd06 and cd06stagg).
 * It calls no hipMalloc or hipFree.  The operator variant (handle, kind, bc1,
 * bcn) must have been built by one earlier call of the ordinary entry point,
 * otherwise PDO_ERR_ARG.  Sizes outside the register path (n not a multiple
 * of 32 rows, more than 16 segments, non-periodic x pencils) use a work area
 * of n_sys * na * nb FP64 words held in a fixed device array of
 * PDO_NOALLOC_SCRATCH_BYTES (one per device, kept until the process ends);
 * larger inputs return PDO_ERR_ARG.  Every call on one device shares that
 * array: calls that overlap in time on one device (concurrent streams) race. */
#define PDO_NOALLOC_SCRATCH_BYTES (32u << 20)
int pdo_op_apply_noalloc(pdo_op h, int kind, int dir, int prec, const void *f, void *out, int na, int nb, int bc1,
                         int bcn, void *stream);

/* Host-side description of one operator variant (no device work): the dense
 * system A y = B f it solves, A n x n and B n x n_in, row-major FP64.  kind:
 * 0 first derivative, 1 second derivative, 2 filter, 10.. staggered
 * (10 ddz_E2C, 11 ddz_C2E, 12 d2dz2_C2C, 13 d2dz2_E2E, 14 InterpZ_E2C,
 * 15 InterpZ_C2E, 16 ddz_C2C, 17 ddz_E2E).  Used to check the tables against
 * the oracle. */
int pdo_op_dense(pdo_op h, int kind, int bc1, int bcn, double *A, double *B);
/* Size query for pdo_op_dense: *n_sys rows, *n_in columns of B. */
int pdo_op_dims(pdo_op h, int kind, int *n_sys, int *n_in);
/* Partition chosen for an operator variant: segment length m and count P
 * (P == 1 with m == n means the general single-pencil kernel). */
int pdo_op_partition(pdo_op h, int kind, int bc1, int bcn, int *m, int *P);
/* Copy of the partitioned-solve tables of a variant (host memory, FP64), for
 * checking them on the CPU: meta[6] = {m, P, K, nbl, nbr, xr} (the last
 * segment holds m + xr rows, the others m); fac [4][m+xr][5],
 * spk [4][m+xr][4], G [P][2K][2KP], sp [2][4][8].  general != 0 asks for the
 * single-segment tables.  Pass NULL arrays to query meta only. */
int pdo_op_tables(pdo_op h, int kind, int bc1, int bcn, int general, int *meta, double *fac, double *spk,
                  double *G, double *sp);

/* ---- periodic FFT Poisson solver -----------------------------------------
 * replaces type(PoissonPeriodic), src/utilities/PoissonPeriodic.F90:
 *   init          :114-212 (dir_id 1, x-decomposed input), modk_* choose the
 *                 modified-wavenumber callback for each direction
 *   poisson_solve :37-87 ; poisson3D_multiply :89-111
 * zero_policy 0 reproduces the reference (divide by k^2+1e-20, zero only the
 * (1,1,1) mode); 1 zeroes every mode with |k'|^2 <= 1e-14, as
 * PadePoisson.F90:104-108 does (needed for collocated modified wavenumbers,
 * which vanish at the Nyquist modes).  nx must be a power of two in [16, 1024],
 * ny and nz powers of two in [8, 1024].  rhs and f are real (nx, ny, nz); rhs == f (in place) allowed. */
int pdo_poisson_periodic_init(pdo_poisson *h, int nx, int ny, int nz, double dx, double dy, double dz, int modk_x,
                              int modk_y, int modk_z, int zero_policy);
/* PoissonPeriodic%init with the Get_ModKx/y/z callbacks already applied
 * (PoissonPeriodic.F90:188-204): kx[nx/2+1], ky[ny], kz[nz] are the
 * wavenumbers of GetWaveNums (:226-242) after the callbacks, in 1/length
 * (host arrays, copied).  The solve divides by kx^2 + ky^2 + kz^2. */
int pdo_poisson_periodic_init_k(pdo_poisson *h, int nx, int ny, int nz, const double *kx, const double *ky,
                                const double *kz, int zero_policy);
int pdo_poisson_periodic_solve(pdo_poisson h, int prec, const void *rhs, void *f, void *stream);
int pdo_poisson_destroy(pdo_poisson h);

/* ---- Pade6stagg: igrid's vertical staggered derivative object -------------
 * replaces type(Pade6stagg) of src/incompressible/PadeDerOps.F90 with scheme
 * cd06: init :57-124 (periodic: one periodic cd06stagg; otherwise the nine
 * boundary objects derOO ... derSS), and the generics ddz_C2E / ddz_E2C /
 * d2dz2_C2C / d2dz2_E2E / interpz_C2E / interpz_E2C :146-990 selected by
 * kind = 10..15 (as pdo_op_dense) and (bot, top) in {-1 odd, 0 one-sided,
 * 1 even}; pairs without an object write zeros like the reference.
 * getModifiedWavenumbers :997-1075: scheme 0 fd02, 1 cd06, 2 Fourier. */
int pdo_pade6stagg_init(pdo_pade6stagg *h, int nz, double dz, int periodic);
int pdo_pade6stagg_apply(pdo_pade6stagg h, int kind, int bot, int top, int prec, const void *in, void *out, int n1,
                         int n2, int is_complex, void *stream);
int pdo_pade6stagg_modk(int scheme, int n, double dz, const double *k, double *kp);
int pdo_pade6stagg_destroy(pdo_pade6stagg h);

/* ---- spectral: 2D transforms of the igrid solvers --------------------------
 * replaces type(spectral) of src/incompressible/spectral.F90 as igrid uses it
 * (dimTransform = 2, x input pencil, igrid.F90:489-493):
 *   init            :867-931, initializeEverything :933-1260 (TwoPeriodic),
 *                   init_periodic_inZ_procedures :755-865 (periodic_z != 0)
 *   fft / ifft      :1413-1453 -> fft_3d.F90:645-663 / :616-643: x r2c then y
 *                   c2c per z plane, unnormalized forward, inverse x 1/(nx ny);
 *                   set_oddball drops the kx = nx/2 modes before the c2r
 *   mTimes_ik1/2    :235-312 (out = i k f, in == out allowed)
 *   dealias         :314-341: without periodic z the 2/3 rule on (k1, k2)
 *                   (hard-coded 2/3, :1135-1148); with periodic z a z
 *                   transform, the dealias_fact mask on (k1, k2, k3), and the
 *                   inverse times 1/nz
 *   dealias_edgeField :343-363 (nz+1 planes, periodic z only)
 * Real fields are (nx, ny, nz), spectra (nx/2+1, ny, nz) interleaved complex,
 * both unpadded.  Wavenumbers carry the oddball with a + sign, and zero when
 * fix_oddball != 0 (igrid passes 0).  nx a power of two in [16, 1024], ny in
 * [8, 1024]; with periodic_z, nz a power of two in [8, 1024]. */
int pdo_spectral_init(pdo_spectral *h, int nx, int ny, int nz, double dx, double dy, double dz, int fix_oddball,
                      double dealias_fact, int periodic_z);
int pdo_spectral_fft(pdo_spectral h, int prec, const void *in, void *out, void *stream);
int pdo_spectral_ifft(pdo_spectral h, int prec, const void *in, void *out, int set_oddball, void *stream);
int pdo_spectral_mtimes_ik(pdo_spectral h, int prec, int dir, const void *in, void *out, void *stream);
int pdo_spectral_dealias(pdo_spectral h, int prec, void *fhat, void *stream);
int pdo_spectral_dealias_edge(pdo_spectral h, int prec, void *fhat, void *stream);
/* host copies of k1 (nx/2+1), k2 (ny), k3 (nz) and the dealias factors g1, g2,
 * g3 (0 or 1, the mask is their product); NULL skips an array */
int pdo_spectral_wavenumbers(pdo_spectral h, double *k1, double *k2, double *k3, double *g1, double *g2,
                             double *g3);
/* replace the dealias factors g1 (nx/2+1), g2 (ny), g3 (nz) by any separable
 * real transfer function (host arrays; NULL keeps one): pdo_spectral_dealias
 * then applies that filter, e.g. a Gaussian LES filter.  No reference
 * counterpart (the reference's masks are fixed at init). */
int pdo_spectral_set_factors(pdo_spectral h, const double *g1, const double *g2, const double *g3);
int pdo_spectral_destroy(pdo_spectral h);

/* ---- PadePoisson: periodic pressure projection of igrid ---------------------
 * replaces the periodic-in-z branch of type(padepoisson),
 * src/incompressible/PadePoisson.F90, with derivZ = Pade6stagg(scheme cd06,
 * periodic):
 *   init (PeriodicInZ) :130-180 + InitPeriodicPoissonSolver :76-128:
 *       |k|^2 = k1^2 + k2^2 + k3'^2, k3' the cd06stagg modified wavenumber,
 *       modes with |k|^2 <= 1e-14 get no pressure
 *   PressureProjection -> PeriodicProjection :386-432 on (uhat, vhat) of
 *       (nx/2+1, ny, nz) and what of (nx/2+1, ny, nz+1), in place
 *   DivergenceCheck :1165-1244 (steps 1-3): the real divergence field
 *       (nx, ny, nz) of the spectral velocity. */
int pdo_padepoisson_init_periodic(pdo_padepoisson *h, pdo_spectral sp, double dz);
/* the non-periodic branch (walls in z, PeriodicInZ = .false., without the
 * Stokes pressure): init :181-232, PressureProjection :434-624, the
 * reference's algorithm.  f = i (k1 u + k2 v) on the cells; the even
 * extension of f and the odd extension of w (2 nz points each) are
 * transformed in z; P = -(F + i kcm W) / |k|^2 (|k|^2 <= 1e-14: 0) and
 * W -= i kcp P, with kcm, kcp = k3' exp(-+ i k3 dz / 2) and k3' the
 * cd06stagg modified wavenumber of GetWaveNums(2 nz, dz); both go back,
 * u -= i k1 p, v -= i k2 p, and w = 0 on both walls.  The odd extension
 * assumes w = 0 on the walls on input (no penetration); w given there is
 * extended as it is, as in the reference, and the result is then not
 * divergence-free next to that wall.  sp must not be periodic in z; 2 nz a
 * power of two in [8,1024].  pdo_padepoisson_divergence then uses the
 * odd-ended E2C (the reference's derZ, :150-151). */
int pdo_padepoisson_init_nonperiodic(pdo_padepoisson *h, pdo_spectral sp, double dz);
/* computeStokesPressure = .true. with Lz (init :236-307) on an object of
 * pdo_padepoisson_init_nonperiodic.  Step 0 of the projection and of both
 * pressures (ProjectStokesPressure :320-384, GetStokesPressure :641-714)
 * then removes w on both walls line by line with the potential
 * phi = chat_bot cosh(lam (Lz - z)) + chat_top cosh(lam z), lam = |(k1, k2)|,
 * chat_bot = -w(bottom) denFact, chat_top = w(top) denFact,
 * denFact = 1 / (lam sinh(lam Lz) + 1e-13) (0 where lam Lz >= 500, below
 * 1e-16 and on the mean line); u -= i k1 phi, v -= i k2 phi,
 * w -= dphi/dz, with the reference's caps of cosh and sinh from an argument
 * of 32 on.  From Step 0 on, k1 and k2 are the plain GetWaveNums values, not
 * the spectral object's (Nyquist modes of the other sign), as in the
 * reference.  The tables are evaluated on the fly; one call per object. */
int pdo_padepoisson_set_stokes(pdo_padepoisson h, double Lz);
int pdo_padepoisson_projection(pdo_padepoisson h, int prec, void *uhat, void *vhat, void *what, void *stream);
int pdo_padepoisson_divergence(pdo_padepoisson h, int prec, const void *uhat, const void *vhat, const void *what,
                               void *divergence, void *stream);
/* getPressure (update 0; :752-898, periodic :716-750) and
 * getPressureAndUpdateRHS (update 1; :952-1163, periodic :900-949):
 * pressure (nx, ny, nz) real, of prec's storage type, the inverse 2D
 * transform of the pressure p of the projection.  update 0 leaves uhat,
 * vhat, what as they are (with walls and the Stokes pressure, its step runs
 * on copies held by the object); update 1 projects them as
 * pdo_padepoisson_projection does.  With the Stokes pressure the pressure
 * adds i phi (phat_z1 + phat_z2 of the reference, :882-890): the reference's
 * factor, although the velocity takes -grad phi.  getPressureAndUpdateRHS
 * adds the i phi of the same call; the reference adds phat_z1 + phat_z2 as
 * its last getPressure stored them, since ProjectStokesPressure stores none. */
int pdo_padepoisson_pressure(pdo_padepoisson h, int prec, void *uhat, void *vhat, void *what, void *pressure,
                             int update, void *stream);
int pdo_padepoisson_destroy(pdo_padepoisson h);

/* ---- igrid%compute_duidxj --------------------------------------------------
 * replaces igrid.F90:2553-2683 for one GPU (the y <-> z transposes of the
 * reference are identities there).  Inputs are igrid's spectral velocity:
 * uhat, vhat, whatC on cells (nx/2+1, ny, nz) and what, uEhat, vEhat on edges
 * (nx/2+1, ny, nz+1).  Outputs, in igrid's component order 1 dudx, 2 dudy,
 * 3 dudz, 4 dvdx, 5 dvdy, 6 dvdz, 7 dwdx, 8 dwdy, 9 dwdz:
 *   duidxjC (nx, ny, nz, 9) and duidxjE (nx, ny, nz+1, 9) real,
 *   duidxjChat / duidxjEhat their spectra (the reference fills them likewise),
 *   d2wdz2hatE, d2udz2hatC, d2vdz2hatC (viscous only; NULL when inviscid).
 * spC: spectral object of (nx, ny, nz); spE: of (nx, ny, nz+1); dz: the
 * Pade6stagg object.  bc[12] = {uBC_bottom, uBC_top, vBC_bottom, vBC_top,
 * wBC_bottom, wBC_top, dUdzBC_bottom, dUdzBC_top, dVdzBC_bottom, dVdzBC_top,
 * dWdzBC_bottom, dWdzBC_top} (ignored by a periodic Pade6stagg). */
typedef struct {
    const void *uhat, *vhat, *whatC, *what, *uEhat, *vEhat;
    void *duidxjC, *duidxjE, *duidxjChat, *duidxjEhat;
    void *d2wdz2hatE, *d2udz2hatC, *d2vdz2hatC;
} pdo_duidxj_fields;
int pdo_duidxj_init(pdo_duidxj *h, pdo_spectral spC, pdo_spectral spE, pdo_pade6stagg dz, const int *bc,
                    int inviscid);
int pdo_duidxj_compute(pdo_duidxj h, int prec, const pdo_duidxj_fields *fields, void *stream);
int pdo_duidxj_destroy(pdo_duidxj h);

/* ---- igrid%interp_PrimitiveVars ------------------------------------------
 * replaces igrid.F90:1423-1463 on the pdo_duidxj object (its spectral
 * objects, Pade6stagg and boundary kinds bc[0..5] = uBC_bottom, uBC_top,
 * vBC_bottom, vBC_top, wBC_bottom, wBC_top):
 *   whatC = interpz_E2C(what), wC = ifft(whatC)    (wBC)
 *   uEhat = interpz_C2E(uhat), uE = ifft(uEhat)    (uBC), v likewise
 *   with That != NULL: TEhat = interpz_C2E(That) (TBC_bottom, TBC_top); when
 *   botBC_Temp == 0 its bottom plane is zeroed and its mean mode set to
 *   Tsurf nx ny (:1449-1454); TE = ifft(TEhat).
 * Cell spectra (nx/2+1, ny, nz), edge spectra (nx/2+1, ny, nz+1); real
 * fields (nx, ny, nz) / (nx, ny, nz+1). */
typedef struct {
    const void *uhat, *vhat, *what, *That;
    void *whatC, *wC, *uEhat, *uE, *vEhat, *vE, *TEhat, *TE;
    int TBC_bottom, TBC_top, botBC_Temp;
    double Tsurf;
} pdo_interp_fields;
int pdo_igrid_interp_primitive(pdo_duidxj h, int prec, const pdo_interp_fields *fields, void *stream);

/* ---- SGS eddy-viscosity models (src/incompressible/sgs_models) -------------
 * pdo_sgs_kernel: get_Sij_from_duidxj (eddyViscosity.F90:1-21) fused with the
 * model kernel, model 0 Smagorinsky |S| = sqrt(2 S_ij S_ij)
 * (smagorinsky.F90:45-68), 1 sigma s3 (s1 - s2) (s2 - s3) / s1^2 of the
 * gradient's singular values (sigma.F90:30-100) or 2 AMD with constants
 * camd_x/y/z (AMD.F90:23-73); duidxj (npts, 9), Sij (npts, 6), nu_sgs
 * (npts); Sij or nu_sgs may be NULL.  The model constants of init_sigma
 * ((Csgs deltaLES)^2, sigma.F90:1-21) are the caller's cglobal.
 * pdo_sgs_scale: multiply_by_model_constant (eddyViscosity.F90:74-103):
 * nu *= cplane[k] on plane k of nxy points (device array, e.g. the dynamic
 * procedure's lambda), or cglobal when cplane is NULL.
 * pdo_dynproc_*: DoStandardDynamicProcedure (standardDynamicProcedure.F90:
 * 74-250) with the spectral test filter of InitTestFilter (spectral.F90:
 * 1333-1367, filtfact = deltaRat) and no vertical test filter: from u, v, w
 * (nx, ny, nz), their spectra, Sij and the unscaled Smagorinsky kernel Dsgs,
 * the planar-averaged coefficient lambda_1d (device, nz FP64 values). */
int pdo_sgs_kernel(int prec, int model, const void *duidxj, void *Sij, void *nu_sgs, long npts, double camd_x,
                   double camd_y, double camd_z, void *stream);
int pdo_sgs_scale(int prec, void *nu_sgs, long nxy, int nz, const double *cplane, double cglobal, void *stream);
int pdo_dynproc_init(pdo_dynproc *h, pdo_spectral spC, double deltaRat);
int pdo_dynproc_compute(pdo_dynproc h, int prec, const void *u, const void *v, const void *w, const void *uhat,
                        const void *vhat, const void *what, const void *Sij, const void *Dsgs, double *lambda_1d,
                        void *stream);
int pdo_dynproc_destroy(pdo_dynproc h);

/* ---- sgs_igrid%getRHS_SGS (sgsmod_igrid.F90:204-266) ------------------------
 * with getTauSGS (:156-201) for the eddy-viscosity models:
 *   S_ij and the kernel on cells (duidxjC) and on edges (duidxjE, kernel only
 *   with explicit_edge = explicitCalcEdgeEddyViscosity);
 *   multiply_by_model_constant (eddyViscosity.F90:74-103): lambda != NULL
 *   (the dynamic procedure's planar coefficient, device, nz values, e.g. from
 *   pdo_dynproc_compute) scales nuC and interpolates it to the edges;
 *   otherwise nuC (and with explicit_edge nuE) take cmodelC / cmodelE per
 *   plane (host arrays of nz / nz+1, copied at init) or cglobal when NULL,
 *   and without explicit_edge nuE = interpz_C2E(nuC) clipped at zero;
 *   tau_ij = -2 nu S_ij (11, 12, 22, 33 on cells; 13, 23 on edges);
 *   urhs -= d/dx tau11 + d/dy tau12 + ddz_E2C(tau13),
 *   vrhs -= d/dx tau12 + d/dy tau22 + ddz_E2C(tau23),
 *   wrhs -= ddz_C2E(tau33) + d/dx tau13 + d/dy tau23
 *   (x, y spectral; z Pade6stagg with (0, 0) ends).
 * urhs, vrhs: (nx/2+1, ny, nz) complex; wrhs: (nx/2+1, ny, nz+1).  prec at
 * init sizes the work buffers; compute must use the same precision.
 * pdo_sgsrhs_fields: device pointers of the last tau (cells: 4 fields of
 * nx ny nz, edges: 2 of nx ny (nz+1)), nu_sgs_C and nu_sgs_E. */
typedef struct pdo_sgsrhs_s *pdo_sgsrhs;
typedef struct {
    int model;        /* 0 Smagorinsky, 1 sigma, 2 AMD */
    int explicit_edge;
    double camd_x, camd_y, camd_z;
    double cglobal;
    const double *cmodelC, *cmodelE;
    int wall_model;   /* 0 none, 1 Moeng (WallModel 1; needs walls in z) */
    double z0;        /* roughness length of the wall model, 0 < z0 < dz/2 */
} pdo_sgs_params;
int pdo_sgsrhs_init(pdo_sgsrhs *h, pdo_spectral spC, pdo_spectral spE, pdo_pade6stagg dz, const pdo_sgs_params *p,
                    int prec);
int pdo_sgsrhs_compute(pdo_sgsrhs h, int prec, const void *duidxjC, const void *duidxjE, const double *lambda,
                       void *urhs, void *vrhs, void *wrhs, void *stream);
int pdo_sgsrhs_fields(pdo_sgsrhs h, const void **tauC, const void **tauE, const void **nuC, const void **nuE);
/* with a wall model: computeWallStress (wallmodel.F90:33-77, Moeng,
 * unstratified) from the first cell level of u, v (physical, (nx, ny, nz)):
 * Uspmn the plane mean of sqrt(u^2 + v^2), ustar = Uspmn kappa /
 * log(dz / 2 / z0) (kappa 0.4), WallMFactor = -ustar^2 / (Uspmn + 1e-13);
 * embed_WM_stress (:79-86) then sets tau13 = WallMFactor u and
 * tau23 = WallMFactor v on the bottom edge plane (after the eddy-viscosity
 * or learned stress), and the divergence follows as in pdo_sgsrhs_compute,
 * which refuses an object with a wall model.  The reference takes the
 * stresses through the spectrum of the first level and back, the same values
 * to rounding.  Its Bou-Zeid model (WallModel 2) filters with a surface
 * filter that the reference never sets up, so it is not offered.
 * pdo_sgsrhs_wall_state: {Uspmn, ustar, WallMFactor, Umn} of the last call
 * (synchronizes). */
int pdo_sgsrhs_compute_wall(pdo_sgsrhs h, int prec, const void *duidxjC, const void *duidxjE, const double *lambda,
                            const void *u, const void *v, void *urhs, void *vrhs, void *wrhs, void *stream);
int pdo_sgsrhs_wall_state(pdo_sgsrhs h, double *state);
/* replace the eddy-viscosity stress by a learned closure's (pdo_closure_stress
 * on the cell and edge gradients, filter width delta, feature dratio); with
 * fallback the eddy-viscosity stress stays where the closure's dissipation is
 * negative.  c = NULL restores the eddy-viscosity model.  No reference
 * counterpart. */
int pdo_sgsrhs_set_closure(pdo_sgsrhs h, pdo_closure c, double delta, double dratio, int fallback);
int pdo_sgsrhs_destroy(pdo_sgsrhs h);

/* ---- igrid time advancement, periodic in z or with walls ------------------
 * replaces the time loop of type(igrid) (src/incompressible/igrid.F90) for
 * domains periodic in x, y and z (the forced isotropic turbulence setup) or
 * with walls in z (PeriodicInZ = .false.: non-periodic spectral objects,
 * Pade6stagg and PadePoisson, the boundary kinds of botWall / topWall, the
 * x-y dealias of the spectra, the SGS wall model of a wall-model bottom
 * wall, the mean pressure gradient dpFdx added to the mean mode of u;
 * no HIT forcing):
 *   timeAdvance :1056-1076 with scheme 1 TVD_RK3 :1105-1170 or 2 SSP_RK45
 *     :1172-1299 (timeSteppingScheme); the stage-1 projection is taken
 *     explicitly, as with fastCalcPressure = .false.
 *   populate_rhs :1793-1912: AddNonLinearTerm_Rot :1497-1566,
 *     addViscousTerm :1914-1959 (Re > 0), getRHS_SGS (use_sgs; the
 *     pdo_sgsrhs model constants, no dynamic procedure, which the reference
 *     does not support with periodic z either), HIT_shell_forcing
 *     forcingIsotropic.F90:123-314 (use_forcing)
 *   project_and_prep :1961-1992: dealiasFields (periodic-z masks), the
 *     periodic PadePoisson projection, inverse transforms,
 *     interp_PrimitiveVars, compute_duidxj
 *   compute_deltaT :1374-1400 (CFL > 0; otherwise the fixed dt)
 * The forcing draws its random wave vectors with this library's generator
 * (splitmix64 of seed, draw counter, wave and variable), not the Fortran
 * runtime's random_number.  set_velocity: physical u, v (nx, ny, nz) and w
 * (nx, ny, nz+1, edges), then project_and_prep; time and step restart at 0.
 * pdo_igrid_field: 0 u, 1 v, 2 wC, 3 w, 4 uhat, 5 vhat, 6 what, 7 duidxjC,
 * 8 duidxjE, 9 urhs, 10 vrhs, 11 wrhs, 12 uE, 13 vE, 14 whatC (device
 * pointers of the object's storage, current after each step). */
typedef struct pdo_igrid_s *pdo_igrid;
typedef struct {
    int nx, ny, nz;
    double Lx, Ly, Lz;
    double Re;       /* <= 0: inviscid */
    int scheme;      /* 1 TVD-RK3, 2 SSP-RK45 */
    double CFL;      /* > 0: dt from the CFL number each step */
    double dt;       /* fixed time step when CFL <= 0 */
    int use_sgs;
    pdo_sgs_params sgs;
    int use_forcing;
    double kmin, kmax, eps_amplitude;
    int nwaves;
    unsigned long long seed;
    /* walls in z (0: periodic).  botWall 1 no-slip, 2 slip, 3 wall model
     * (with sgs.wall_model); topWall 1 or 2 (get_boundary_conditions_stencil
     * igrid.F90:5148-5220); dpFdx the mean pressure gradient of
     * addExtraForcingTerm (:1759-1765, 0: none); stokes_pressure the
     * PadePoisson Stokes pressure (computeStokesPressure, Lz; igrid's
     * namelist default is on, igrid.F90:330: without it the projection
     * zeroes the wall w each stage's update leaves there, and the wall cells
     * keep that divergence) */
    int walls, botWall, topWall;
    double dpFdx;
    int stokes_pressure;
} pdo_igrid_params;
int pdo_igrid_init(pdo_igrid *h, const pdo_igrid_params *p, int prec);
int pdo_igrid_set_velocity(pdo_igrid h, const void *u, const void *v, const void *w, void *stream);
int pdo_igrid_step(pdo_igrid h, int nsteps, void *stream);
int pdo_igrid_field(pdo_igrid h, int which, const void **ptr);
int pdo_igrid_time(pdo_igrid h, double *tsim, long *step, double *dt);
/* Restart files (igrid.F90:2719-2800): dump writes dir/RESTART_RunRR_u.TTTTTT,
 * _v, _w (the physical fields of the current step as raw FP64 in Fortran
 * order, global arrays as decomp_2d_write_one leaves them; w on the nz+1
 * edges) and _info.TTTTTT (tsim as one "(100g15.5)" record), TTTTTT the step
 * number, RR the run id.  read loads such a set (FP64 data converted to the
 * object's precision), prepares the stage and sets tsim and the step count. */
int pdo_igrid_dump_restart(pdo_igrid h, const char *dir, int rid, void *stream);
int pdo_igrid_read_restart(pdo_igrid h, const char *dir, int rid, long tid, void *stream);
/* SGS closure of the step (needs use_sgs with the Smagorinsky model at init):
 * kind 0 the eddy-viscosity model with the constants of pdo_igrid_params;
 * 1 dynamic Smagorinsky: the standard dynamic procedure (pdo_dynproc, test
 *   filter ratio deltaRat) gives the planar coefficient at every stage (the
 *   reference allows it with walls only; here periodic z is homogeneous);
 * 2 the learned closure c (filter width delta, feature dratio) with the
 *   dynamic Smagorinsky stress wherever its dissipation is negative;
 * 3 the learned closure alone.  No reference counterpart for kinds 2, 3. */
int pdo_igrid_set_closure(pdo_igrid h, int kind, pdo_closure c, double delta, double dratio, double deltaRat);
int pdo_igrid_destroy(pdo_igrid h);

/* ---- learned SGS closure -----------------------------------------------------
 * No counterpart in the reference: the pointwise MLP closure of this library
 * (10 inputs -> 64 -> 64 -> 64 -> 6, GELU), trained and run on the device
 * with BF16 weights / activations and FP32 accumulation on the matrix cores,
 * FP32 master parameters and Adam.  Parameters are one FP32 vector of
 * pdo_mlp_nparam() values: W1 (64 x 16, columns 10..15 zero), b1 (64),
 * W2 (64 x 64), b2, W3 (64 x 64), b3, W4 (16 x 64, rows 6..15 zero), b4 (16,
 * entries 6..15 zero), each W row-major [out][in].  X (B, 10) and the
 * targets T / outputs Y (B, 6) are row-major FP32 device arrays.  The loss is
 * the mean over B x 6 of the squared error; gradients are summed over the
 * batch in a fixed order (deterministic).  set_params resets Adam. */
int pdo_mlp_nparam(void);
int pdo_mlp_create(pdo_mlp *h, long max_batch);
int pdo_mlp_set_params(pdo_mlp h, const float *host_params);
int pdo_mlp_get_params(pdo_mlp h, float *host_params);
int pdo_mlp_get_grads(pdo_mlp h, float *host_grads);
int pdo_mlp_forward(pdo_mlp h, const float *X, float *Y, long B, void *stream);
int pdo_mlp_gradients(pdo_mlp h, const float *X, const float *T, long B, float *Y, void *stream);
int pdo_mlp_adam_step(pdo_mlp h, float lr, float b1, float b2, float eps, void *stream);
int pdo_mlp_train_step(pdo_mlp h, const float *X, const float *T, long B, float lr, float b1, float b2, float eps,
                       float *Y, void *stream);
/* training on any loss of the outputs: pdo_mlp_forward_train keeps the
 * activations of a batch (B <= max_batch) and pdo_mlp_backward takes dL/dY
 * (B, 6) of that batch, leaving the parameter gradients for
 * pdo_mlp_adam_step / pdo_mlp_get_grads */
int pdo_mlp_forward_train(pdo_mlp h, const float *X, float *Y, long B, void *stream);
int pdo_mlp_backward(pdo_mlp h, const float *dY, long B, void *stream);
int pdo_mlp_destroy(pdo_mlp h);

/* ---- learned closure: invariant features and tensor-basis stress -------------
 * No counterpart in the reference (its baseline is the eddy-viscosity family of
 * src/incompressible/sgs_models).  From a velocity-gradient field duidxj
 * (npts, 9 components npts apart, igrid's order du_i/dx_j at 3 i + j) and its
 * scale g = sqrt(mean |A|^2) the closure forms a = A / g, s = sym(a) - tr/3 I,
 * r = skew(a) and
 *   X (npts, 10): sqrt tr s^2, sqrt -tr r^2, cbrt tr s^3, cbrt tr r^2 s,
 *     (-tr r^2 s^2)^(1/4), signed 6th root of tr r^2 s r s^2, the principal
 *     strains of s (descending), dratio (= Delta / Delta_0);
 *   the stress tau = Delta^2 g^2 sum_n G_n T_n with the network outputs G
 *     (6) and Pope's basis T1 = s, T2 = s r - r s, T3 = s^2 - tr s^2/3 I,
 *     T4 = r^2 - tr r^2/3 I, T5 = r s^2 - s^2 r, T6 = r^2 s + s r^2 -
 *     2/3 tr(s r^2) I (frame-indifferent).
 * pdo_closure_features: X and, unless NULL, A9 = a (npts, 9), scale from
 *   this field.
 * pdo_closure_stress: features, network, stress into tau with which = 0
 *   (11 12 22 33, 4 components npts apart: igrid's cell layout), 1 (13 23:
 *   the edge layout) or 2 (11 12 13 22 23 33); rescale = 0 reuses the scale
 *   of the previous call; with fallback the values already in tau stay
 *   wherever the network's dissipation -tau_ij S_ij is negative, and
 *   pdo_closure_fallback_count reports how many points did (synchronises).
 * pdo_closure_loss: the training loss sum_b sum_c w_c (p_c - T_c)^2 / (9 B)
 *   of p = sum_n G_n T_n(A9) against T (B, 6), w = (1 2 2 1 2 1), and dL/dG
 *   (B, 6); loss (host) may be NULL (no synchronisation).
 * FP64 arithmetic per point; buffers X, G sized for max_points. */
int pdo_closure_create(pdo_closure *h, pdo_mlp net, long max_points);
int pdo_closure_features(pdo_closure h, int prec, const void *duidxj, long npts, double dratio, float *X, float *A9,
                         void *stream);
int pdo_closure_stress(pdo_closure h, int prec, const void *duidxj, long npts, double delta, double dratio, int which,
                       int fallback, int rescale, void *tau, void *stream);
int pdo_closure_fallback_count(pdo_closure h, long *count, void *stream);
int pdo_closure_scale(pdo_closure h, double *rms, void *stream);
int pdo_closure_loss(pdo_closure h, const float *G, const float *A9, const float *T, long B, float *dG, double *loss,
                     void *stream);
int pdo_closure_destroy(pdo_closure h);

/* ---- explicit filters: gaussianstuff / lstsqstuff ------------------------
 * gaussian%init(n, periodic) (gaussian.F90:74-93) and lstsq%init
 * (lstsq.F90:74-93): symmetric 9-point filters (Gaussian of width 4 dx;
 * least squares), with 4 one-sided closure rows per end when non-periodic.
 * n == 1 copies; otherwise n >= 8 (PDO_ERR_ARG below).
 * filter1/2/3 (gaussian.F90:104-573, lstsq.F90:104-523): same layouts and
 * arguments as the cf90 filters.  Gaussian bc1/bcn: 0 one-sided rows, +1 / -1
 * even / odd reflection about the end point; other values return 324
 * (GracefulExit 324).  The least-squares filter ignores bc1/bcn.
 * FP32 storage is filtered in FP64 arithmetic (all FP32 modes agree). */
int pdo_gaussian_init(pdo_xfilter *h, int n, int periodic);
int pdo_lstsq_init(pdo_xfilter *h, int n, int periodic);
int pdo_xfilter_filter1(pdo_xfilter h, int prec, const void *f, void *fil, int na, int nb, int bc1, int bcn,
                        void *stream);
int pdo_xfilter_filter2(pdo_xfilter h, int prec, const void *f, void *fil, int na, int nb, int bc1, int bcn,
                        void *stream);
int pdo_xfilter_filter3(pdo_xfilter h, int prec, const void *f, void *fil, int na, int nb, int bc1, int bcn,
                        void *stream);
/* complex (na, nb, n), interleaved (filter3_cmplx, gaussian.F90:574-582) */
int pdo_xfilter_filter3_cmplx(pdo_xfilter h, int prec, const void *f, void *fil, int na, int nb, int bc1, int bcn,
                              void *stream);
/* host only: the n x n matrix (row major) the filter applies with (bc1, bcn) */
int pdo_xfilter_dense(pdo_xfilter h, int bc1, int bcn, double *W);
int pdo_xfilter_destroy(pdo_xfilter h);

/* ---- derivatives / filters dispatch and the operators module -------------
 * derivatives%init (derivatives.F90:140-186,264-390) on one GPU: sizes
 * (nx, ny, nz) are the pencil sizes of every direction.  method[d]:
 * 1 "cd10", 2 "cd06", 3 "four" (Fourier collocation, ffts.F90: power-of-two
 * n, x 16..1024, y/z 8..1024, else PDO_ERR_UNSUP), 4 "cheb" (PDO_ERR_UNSUP:
 * incomplete in the reference); other values PDO_ERR_ARG.
 * ddx..d2dz2 (derivatives.F90:447-569) on a field (nx, ny, nz): cd10 takes
 * bc1/bcn, cd06 and four ignore them; cd06 second derivatives return
 * PDO_ERR_UNSUP ("CD06 is incomplete", GracefulExit 21).  A Fourier y (z)
 * derivative needs nx (nx*ny) even. */
int pdo_derivatives_init(pdo_derivatives *h, int nx, int ny, int nz, double dx, double dy, double dz,
                         const int *periodic, const int *method);
int pdo_derivatives_method(pdo_derivatives h, int dir, int *method);
int pdo_derivatives_ddx(pdo_derivatives h, int prec, const void *f, void *df, int bc1, int bcn, void *stream);
int pdo_derivatives_ddy(pdo_derivatives h, int prec, const void *f, void *df, int bc1, int bcn, void *stream);
int pdo_derivatives_ddz(pdo_derivatives h, int prec, const void *f, void *df, int bc1, int bcn, void *stream);
int pdo_derivatives_d2dx2(pdo_derivatives h, int prec, const void *f, void *df, int bc1, int bcn, void *stream);
int pdo_derivatives_d2dy2(pdo_derivatives h, int prec, const void *f, void *df, int bc1, int bcn, void *stream);
int pdo_derivatives_d2dz2(pdo_derivatives h, int prec, const void *f, void *df, int bc1, int bcn, void *stream);
int pdo_derivatives_destroy(pdo_derivatives h);
/* filters%init (filters.F90:85-175): method[d] 1 "cf90", 2 "gaussian",
 * 3 "lstsq"; filterx/y/z (filters.F90:220-269) on a field (nx, ny, nz). */
int pdo_filters_init(pdo_filters *h, int nx, int ny, int nz, const int *periodic, const int *method);
int pdo_filters_method(pdo_filters h, int dir, int *method);
int pdo_filters_filterx(pdo_filters h, int prec, const void *f, void *ff, int bc1, int bcn, void *stream);
int pdo_filters_filtery(pdo_filters h, int prec, const void *f, void *ff, int bc1, int bcn, void *stream);
int pdo_filters_filterz(pdo_filters h, int prec, const void *f, void *ff, int bc1, int bcn, void *stream);
int pdo_filters_destroy(pdo_filters h);
/* operators.F90: bc = {x_bc(1:2), y_bc(1:2), z_bc(1:2)} (NULL: zeros).
 * gradient (:18-53); curl (:55-116) into 3 consecutive fields
 * (dw/dy - dv/dz, du/dz - dw/dx, dv/dx - du/dy); divergence (:118-156) with
 * the bc signs reversed; filter3D (:158-223) filters arr in place in y, x,
 * then z, numtimes times each (numtimes < 1 means 1).  The derivatives /
 * filters object holds the work fields (allocated at the first call). */
int pdo_operators_gradient(pdo_derivatives der, int prec, const void *f, void *dfdx, void *dfdy, void *dfdz,
                           const int *bc, void *stream);
int pdo_operators_curl(pdo_derivatives der, int prec, const void *u, const void *v, const void *w, void *curl,
                       const int *bc, void *stream);
int pdo_operators_divergence(pdo_derivatives der, int prec, const void *u, const void *v, const void *w, void *div,
                             const int *bc, void *stream);
int pdo_operators_filter3d(pdo_filters fil, int prec, void *arr, int numtimes, const int *bc, void *stream);

/* ---- cgrid: compressible right-hand side, one species of ideal gas --------
 * replaces type(cgrid)'s getRHS (src/compressible/cgrid.F90:822-894, without
 * the problem's source hook) with get_primitive :763-797, getRHS_x/y/z
 * :896-1056, getLAD :1058-1257 (artificial shear and bulk viscosity and
 * conductivity, each filtered twice per direction with the Gaussian filter
 * gfil), get_tau :1360-1404, get_q :1406-1452, and filter :1259-1336.
 * Ideal gas (IdealGasEOS.F90: p = (gam-1) rho e, T = (gam-1) e / Rgas,
 * c = sqrt(gam p / rho)); power-law viscosity (PowerLawViscosity.F90:
 * mu = mu_ref (T/T_ref)^n, bulk 0, kappa = Cp mu / Pr), or none when
 * inviscid (the reference's default).  One species only: the mixture's
 * species fluxes and diffusivities are not built.
 * Sizes (nx, ny, nz) on one GPU; derivative[3] as pdo_derivatives_init
 * (1 cd10, cgrid's default; 2 cd06), filter[3] the conserved-variable filter
 * fil as pdo_filters_init (1 cf90, cgrid's default); x_bc, y_bc, z_bc the
 * reference's (bc1, bcn) per direction (ignored where periodic).
 * pdo_cgrid_rhs: W (nx, ny, nz, 5) = rho, rho u, rho v, rho w, TE; rhs the
 * same layout.  pdo_cgrid_filter: cgrid%filter of one field in place,
 * numtimes passes along y, x, z with fil (gaussian = 0) or gfil.
 * pdo_cgrid_field (device pointers of the last rhs): 0 rho, 1 u, 2 v, 3 w,
 * 4 e, 5 p, 6 T, 7 mu, 8 bulk, 9 kap (physical + artificial), 10 duidxj
 * (9 fields, du_i/dx_j at 3 i + j), 11 tau (xx xy xz yy yz zz), 12 q
 * (3 fields).  prec at init sizes the work (43 fields); rhs and filter must
 * use it.  Pointwise arithmetic is FP64 for every storage type. */
typedef struct pdo_cgrid_s *pdo_cgrid;
typedef struct {
    int nx, ny, nz;
    double dx, dy, dz;
    int periodic[3];
    int x_bc[2], y_bc[2], z_bc[2];
    int derivative[3];
    int filter[3];
    double gam, Rgas;
    double Cmu, Cbeta, Ckap;
    int inviscid;
    double mu_ref, T_ref, visc_n, Pr;
} pdo_cgrid_params;
int pdo_cgrid_init(pdo_cgrid *h, const pdo_cgrid_params *p, int prec);
int pdo_cgrid_rhs(pdo_cgrid h, int prec, const void *W, void *rhs, void *stream);
int pdo_cgrid_filter(pdo_cgrid h, int prec, void *arr, int numtimes, int gaussian, void *stream);
int pdo_cgrid_field(pdo_cgrid h, int which, const void **ptr);
int pdo_cgrid_destroy(pdo_cgrid h);

/* ---- full RHS evaluation (the benchmarked composite) ----------------------
 * On a triply periodic n^3 grid of side L, for velocity (u, v, w):
 *   grad(:,:,:,3*i+j) = d u_i / d x_j          9 cd06 first derivatives
 *   ut_i = F_z F_y F_x u_i                      9 cf90 filter passes
 *   div  = sum_j d ut_j / d x_j                 3 cd06 passes (accumulated)
 *   p    = Poisson(div), cd06 modified wavenumbers, |k'|^2 <= 1e-14 zeroed
 *   uo_i = ut_i - d p / d x_i                   3 cd06 passes
 * built from derivatives%ddx/ddy/ddz (derivatives.F90:447-510),
 * filters%filterx/y/z (filters.F90:220-269) and PoissonPeriodic.  grad holds 9
 * consecutive fields.  prec: PDO_F64 or PDO_F32_REFINE / PDO_F32_STORE /
 * PDO_F32 (all buffers then FP32).  Work buffers are allocated at init. */
int pdo_rhs_init(pdo_rhs *h, int n, double L, int prec);
int pdo_rhs_eval(pdo_rhs h, const void *u, const void *v, const void *w, void *grad, void *uo, void *vo, void *wo,
                 void *stream);
/* pressure of the last evaluation (device, real n^3) */
int pdo_rhs_pressure(pdo_rhs h, const void **p);
int pdo_rhs_destroy(pdo_rhs h);

/* max |D.u| with the cd06 divergence (DivergenceCheck, PadePoisson.F90:1165-1244),
 * synchronizes the stream; result on the host. */
int pdo_rhs_max_divergence(pdo_rhs h, const void *u, const void *v, const void *w, double *maxdiv, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* PADEOPS_HIP_H */
