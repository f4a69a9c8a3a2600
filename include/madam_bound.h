/*
 * madam_bound.h -- C ABI of the MI355X (gfx950) SDP bounding library
 * libmadam_bound.so.
 *
 * Every entry point takes plain pointers and sizes.  Matrices are FP64,
 * column-major, n x n with leading dimension n, fully symmetric.  Host
 * buffers are copied to HBM inside the call; the batched entry points keep
 * the whole ADMM loop on the device.
 *
 * Each function cites the reference interface it replaces (file:line in
 * HrgaT/MADAM).  Return value: 0 on success, a negative MADAM_E* code on
 * failure (the reference exits the process instead, e.g. admm.c:749-752).
 */
#ifndef MADAM_BOUND_H
#define MADAM_BOUND_H

#ifdef __cplusplus
extern "C" {
#endif

#define MADAM_OK 0
#define MADAM_E_ARG (-1)     /* bad size or null pointer */
#define MADAM_E_HIP (-2)     /* HIP runtime error */
#define MADAM_E_NODEV (-3)   /* no usable GPU */
#define MADAM_E_EIG (-4)     /* eigensolver did not converge */

#define MADAM_NMAX 1024      /* admm.h:131 NMAX */

/* Layout-identical to Triangle_Inequality, Pentagonal_Inequality and
 * Heptagonal_Inequality (admm.h:93-117), so arrays of the reference's
 * structs can be passed by pointer cast. */
typedef struct { int i, j, k, type; double value, y; } madam_tri;
typedef struct { int type; int perm[5]; double value, y; } madam_pent;
typedef struct { int type; int perm[7]; double value, y; } madam_hept;

/* Layout-identical to Parameters (admm.h:57-89). */
typedef struct {
    double sigma0;
    double scaleSigma;
    int ADMM_itermax;
    double tol0;
    double scaleTol;
    double minTol;
    int triag_iter;
    int pent_iter;
    int hept_iter;
    int max_outer_iter;
    int extra_iter;
    double violated_TriIneq;
    int TriIneq;
    int adjust_TriIneq;
    int PentIneq;
    int HeptaIneq;
    int Pent_Trials;
    int Hepta_Trials;
    int include_Pent;
    int include_Hepta;
    int root;
    int use_diff;
    int branchingStrategy;
} madam_params;

/* Knobs of the device ADMM that have no counterpart in the reference. */
typedef struct {
    double cg_rtol;        /* PCG relative residual for (BB^T+I)t = rhs, default 1e-12 */
    int cg_maxit;          /* default 1000 */
    int jacobi_sweeps;     /* max Jacobi sweeps per eigen-solve, default 40 */
    int check_every;       /* ADMM iterations between host checks of the stop flags, default 16 */
} madam_opts;

/* One ADMM_solver call (admm.c:49): inputs are read, outputs written.
 * Cut duals start from the cut structs' y fields (admm.c:79-95). */
typedef struct {
    int n;
    const double *L;          /* PP->L */
    double *X, *y, *Q;        /* in/out warm start, out: final iterates */
    int ntri, npent, nhept;
    const madam_tri *tri;
    const madam_pent *pent;
    const madam_hept *hept;
    double *s, *t, *u;        /* out, length ntri+npent+nhept (may be null) */
    double *sigma;            /* in/out */
    double tol;
    int itermax;              /* params.ADMM_itermax */
    double scaleSigma;        /* params.scaleSigma */
    /* outputs */
    int nbit;
    double f;                 /* safe upper bound (global f of admm.c:757) */
    double lambda_max;        /* certified lambda_max used in f */
    double *Z;                /* n x n buffer; first M columns = eigvecs*sqrt(lambda*sigma) */
    int M;
    int cg_iters;             /* total PCG iterations */
    int sweeps;               /* total Jacobi sweeps */
} madam_admm_job;

/* Initialise the library on a HIP device; idempotent. */
int madam_open(int device);
/* Release device memory. */
void madam_close(void);
/* Last error message (static storage). */
const char *madam_last_error(void);

/* Per-phase GPU timing (no reference counterpart; the reference only takes
 * MPI_Wtime stamps, admm.c:674).  Phases, in order: 0 prep (y, rhs),
 * 1 cg (BB^T+I solve), 2 w (duals + W), 3 eig (projection eigen-solve),
 * 4 post (X_+, residuals, stop test), 5 safe (safe bound), 6 ipm,
 * 7 trisep, 8 sa, 9 gw.  ms[p] = summed event time of all launches of
 * phase p, count[p] = number of timed launch groups. */
#define MADAM_NPHASE 10
int madam_profile(int on);
int madam_profile_read(int nphase, double *ms, long *count, int reset);

/* ADMM_solver (admm.c:49-771) for k independent nodes, all on the GPU in
 * one batch: per node the same iteration, stop test and safe bound as the
 * reference.  opts may be null (defaults). */
int madam_admm_batch(int k, madam_admm_job *jobs, const madam_opts *opts);

/* op_B (operators.c:62): y[c] += B(X)_c for the given cut lists. */
int madam_op_B(int n, const double *X, int ntri, const madam_tri *tri, int npent,
               const madam_pent *pent, int nhept, const madam_hept *hept, double *y);

/* op_Bt (operators.c:171): X -= B^T(t). */
int madam_op_Bt(int n, double *X, int ntri, const madam_tri *tri, int npent,
                const madam_pent *pent, int nhept, const madam_hept *hept, const double *t);

/* project_SDP (operators.c:521): X = W_+ (eigenvalues > 1e-8), Q = X - W,
 * Z = eigenvectors * sqrt(lambda) (n x M, ascending lambda), M. */
int madam_project_sdp(int n, const double *W, double *X, double *Q, double *Z, int *M);

/* Certified largest eigenvalue in FP64 (the dsyevr call of admm.c:731-747):
 * returns an upper bound lam >= lambda_max(A) within |E|_F + 1e-12 |A|_F. */
int madam_lambda_max(int n, const double *A, double *lam);

#ifdef __cplusplus
}
#endif
#endif
