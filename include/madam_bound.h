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
    int projection;        /* 0: FP64 eigen-decomposition (Jacobi); 1: mixed-precision sign
                              iteration (FP32 MFMA quintic Newton-Schulz + FP64 finish) */
    int ns_layout;         /* sign iteration: 0 automatic (the madam_load_tuning table, else
                              built-in crossovers), 1 one workgroup per node, 2 tile-parallel
                              (each product spread over many workgroups), 3 tile-parallel
                              over 2 x 2 blocks of tiles */
    double cg_adapt;       /* > 0: after an ADMM call's first iteration, PCG stops at the
                              relative residual max(cg_rtol, min(1e-3, cg_adapt * the
                              previous iteration's ADMM error)); 0: cg_rtol throughout */
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
    double *Z;                /* n x n buffer; first M columns = eigvecs*sqrt(lambda*sigma)
                                 (projection 1: a pivoted Cholesky factor, Z Z^T = X) */
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

/* Self-test of the MFMA matrix paths used by the mixed-precision projection:
 * C32 = A32 (alpha I + beta B32) with FP32 MFMA (np x np, np % 16 == 0) and
 * C64 = A64 (alpha I + beta B64) with FP64 MFMA (n x n); A, B symmetric and
 * commuting so that the product is symmetric (lower tiles are mirrored). */
int madam_mfma_selftest(int np, const float *A32, const float *B32, float *C32, int n,
                        const double *A64, const double *B64, double *C64, double alpha,
                        double beta);

/* Per-phase GPU timing (no reference counterpart; the reference only takes
 * MPI_Wtime stamps, admm.c:674).  Phases, in order: 0 prep (y, rhs),
 * 1 cg (BB^T+I solve), 2 w (duals + W), 3 eig (projection eigen-solve),
 * 4 post (X_+, residuals, stop test), 5 safe (safe bound), 6 ipm,
 * 7 trisep, 8 sa, 9 gw, 10 zfac (eigen-pairs for the rounding heuristic in
 * the sign-iteration form).  ms[p] = summed event time of all launches of
 * phase p, count[p] = number of timed launch groups. */
#define MADAM_NPHASE 11
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

/* project_SDP in the mixed-precision form: X = W_+ = (W + W sign(W)) / 2
 * with sign(W) from the FP32-MFMA quintic sign iteration and one FP64
 * correction step, Q = X - W, M = number of positive eigenvalues. */
int madam_project_sdp_ns(int n, const double *W, double *X, double *Q, int *M);

/* Autotuning of the sign-iteration projection's launch layout (no
 * reference counterpart: the reference has one dsyevr call per projection,
 * operators.c:579).  Layouts: 1 one workgroup per node, 2 one launch per
 * product over 16 x 16 tiles, 3 the same over 2 x 2 blocks of tiles; all
 * give the same bits.
 * madam_time_projection: average milliseconds of one projection of nb
 * synthetic nodes of order n in `layout` over reps calls.
 * madam_load_tuning: load the table (text lines "n_p nodes layout ...",
 * '#' comments) that picks the layout from the order rounded up to 16 and
 * the number of active nodes; null or "" clears it (the built-in
 * crossovers again).  Host only.
 * madam_tuned_layout: the loaded table's layout for order n and `active`
 * nodes (0 without a table). */
int madam_time_projection(int n, int nb, int layout, int reps, double *ms);
int madam_load_tuning(const char *path);
int madam_tuned_layout(int n, int active);

/* Testing hook, no reference counterpart: cap the LDS bytes the cut solve
 * (PCG on (BB^T + I) t = rhs) plans with, so that small problems run the
 * kernels otherwise chosen only for large cut sets (vectors partly or
 * wholly in HBM), and with one_pass = 0 keep nodes off the one-pass kernel
 * (products over precomputed expanded rows).  (0, 1) restores the
 * defaults.  Host only; applies to the calls that follow.  All kernels
 * solve the same system and the solutions agree to the PCG tolerance. */
int madam_set_cg_lds_cap(long bytes, int one_pass);

/* Testing hook, no reference counterpart: capacities of the pinned staging
 * buffers through which a batch of nodes' small host-device transfers move
 * as one copy each way (uploads, default 64 MiB; downloads, default
 * 32 MiB; 0 keeps a default), so that tests can make them wrap and split
 * on small problems.  Frees the current buffers (synchronising the
 * stream); the next transfer allocates the new size.  Host memory only:
 * results do not depend on the sizes. */
int madam_set_staging(long up_bytes, long down_bytes);

/* Launch mode of the ADMM iterations (no reference counterpart: the
 * reference runs one CPU loop, admm.c:520-704).  on = 1: the iterations
 * between two host checks of the stop flags are captured into a HIP graph
 * the second time the same launches recur (same batch, orders, cut-solve
 * plan and options) and replayed from it with one launch after that;
 * on = 0 (default): every kernel is launched directly.  Both issue the
 * same kernels with the same arguments in the same order (same bits).
 * While the phase timer is on (madam_profile) kernels are launched
 * directly.
 * madam_graph_stats: graphs captured and replays since the library was
 * loaded.  Host only. */
int madam_set_graphs(int on);
int madam_graph_stats(long *captures, long *replays);

/* Eigen-solver of the plain form (projection 0), which stands in for the
 * reference's dsyevr calls (operators.c:579, admm.c:731): orders below n use
 * the one-workgroup cyclic Jacobi kernel (A and the eigenvectors in LDS
 * where they fit), orders from n on a two-sided block Jacobi method spread
 * over many workgroups per node (16-wide blocks, 32 x 32 subproblems
 * diagonalised by one wave each, the block transformations on FP64 MFMA),
 * warm-started from the previous projection's eigenvectors.  Both stop at
 * an off-diagonal sum of squares of at most 1e-20 (Jacobi) or 1e-24 (block
 * Jacobi) of |A|_F^2.  Default 80 (measured crossover on MI355X: the
 * warm-started ADMM projection at n = 60 takes 0.33 ms by Jacobi and 0.36
 * by block Jacobi, at n = 100 0.99 and 0.67 ms, at n = 225 20.1 and
 * 1.7 ms).  Host only; applies to the calls that follow. */
int madam_set_bj_min_order(int n);
/* Diagnostics of the block Jacobi solver: out4 (may be null) receives the
 * counts since the previous call -- eigen-solves, sweeps, subproblem
 * solves, inner Jacobi sweeps of the subproblems -- which are then reset;
 * on = 1 keeps counting (one extra host synchronisation per solve). */
int madam_bj_stats(int on, long *out4);

/* Coefficients (a_k, b_k, c_k) of the sign iteration's steps
 * x -> a x + b x^3 + c x^5 used by the mixed-precision projection (no
 * reference counterpart): writes min(cap, steps) triples to coef and
 * returns the number of steps.  Runs on the host, no device needed. */
int madam_ns_schedule(int cap, float *coef);

/* Certified largest eigenvalue in FP64 (the dsyevr call of admm.c:731-747):
 * returns an upper bound lam >= lambda_max(A) within |E|_F + 1e-12 |A|_F. */
int madam_lambda_max(int n, const double *A, double *lam);

/* The same bound as the mixed-precision form computes it: an estimate
 * (Householder tridiagonalisation and Sturm multisection: one workgroup
 * with A in LDS for n up to about 195, many workgroups with A in device
 * memory beyond) certified by an FP64 Cholesky factorisation of mu I - A
 * (lam >= lambda_max(A), within ~1e-10 |A|_F). */
int madam_lambda_max_cert(int n, const double *A, double *lam);

/* getViolated_TriangleInequalities (cutting_planes.c:48-127): the at most
 * K triangle rows with B_c(X) - 1 > thresh of largest value, in the
 * reference's enumeration order (type, then i > j > k ascending); *maxval is
 * the largest row value over all 4 C(n,3) rows.  Same set as the reference
 * except where rows tie at the K-th value (then earlier rows win). */
int madam_tri_separate(int n, const double *X, int K, double thresh, madam_tri *list, int *size,
                       double *maxval);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): madam_tri_separate on device buffers, allocating
 * nothing.  X (n x n), list (K entries), size and maxval are DEVICE
 * pointers; the work is ordered on `stream` (a hipStream_t; null is HIP's
 * default stream) and the call returns without waiting for it.  Same list,
 * in the same order, as madam_tri_separate.  Working memory is one fixed
 * device array per GPU (1,013,760 bytes, held until the process ends),
 * which limits 3 <= n <= MADAM_NMAX and 0 <= K <= 50000 (the reference's
 * MaxTriIneqAdded); other sizes are rejected with MADAM_E_ARG.  Calls that
 * overlap on one GPU share that array and race: serialise them.
 * madam_open(device) must have been called for the current device. */
int madam_tri_separate_noalloc(int n, const double *X, int K, double thresh, madam_tri *list,
                               int *size, double *maxval, void *stream);

/* qap_simulated_annealing (qap_simulated_annealing.c:12-206) for nruns
 * independent runs on X: run r uses the sign pattern of types[r] (H = ee^T
 * with the first types[r]-1 entries of e negated) and its own random stream
 * (seed, stream_ids[r]); cost[r] and perms[r*k .. r*k+k-1] are the best
 * value and k-tuple, draws[r] the number of random values consumed.  Reads
 * X's lower triangle only (the reference copies it over the upper triangle
 * first, qap_simulated_annealing.c:14-17; X itself is not modified). */
int madam_qap_sa(int n, const double *X, int k, int nruns, const int *types,
                 const long long *stream_ids, unsigned long long seed, double *cost, int *perms,
                 long long *draws);

/* ipm_mc_pk (ipm_mc_pk.c:28-347): basic SDP max <L,X>, diag(X) = e, by the
 * predictor-corrector interior-point method; X, y, Z = Diag(y) - L and the
 * dual value *phi once phi - psi <= 1e-4.  MADAM_E_EIG where the reference
 * exits on a failed Cholesky factorisation. */
int madam_ipm(int n, const double *L, double *X, double *y, double *Z, double *phi, int *iters);

/* GW_heuristic + mc_1opt (heuristic.c:24-201) on a given factor Z (n x M):
 * num hyperplanes from stream (seed, stream), hyperplane h has components
 * u - 1/2 with u = (splitmix64(key + j) >> 11) * 2^-53,
 * key = splitmix64(seed ^ splitmix64((stream << 24) + h)).  cuts[h*n + i]
 * gets the 1-opt cut in +-1 form, values[h] = z'^T L z' with z' the cut
 * whose last (homogenising) entry is set to +1 (heuristic.c:76-84). */
int madam_gw_rounding(int n, const double *L, const double *Z, int M, int num,
                      unsigned long long seed, long long stream, double *values,
                      signed char *cuts);

/* Layout-identical to BabNode (admm.h:147-155). */
typedef struct {
    int xfixed[MADAM_NMAX];
    int solX[MADAM_NMAX];           /* BabSolution sol */
    double fracsol[MADAM_NMAX];
    int level;
    double upper_bound;
} madam_node;

/* The inputs of a node's latest ADMM_solver call inside SDPbound (no
 * reference counterpart; the reference keeps them in globals between the
 * calls of bounding.c:121-236): the cut lists with the duals the call
 * starts from (admm.c:79-95), the warm start X, y, Q, sigma and the
 * tolerance, so that the same call can be run again in another form
 * (madam_admm_batch) on identical inputs.  The caller sets the capacities
 * and buffers: X, Q cap_n x cap_n, y cap_n, the lists cap_tri/cap_pent/
 * cap_hept entries.  L is the node's subproblem (madam_create_subproblem). */
typedef struct {
    int cap_n, cap_tri, cap_pent, cap_hept;   /* in */
    double *X, *y, *Q;                        /* out buffers */
    madam_tri *tri;
    madam_pent *pent;
    madam_hept *hept;
    int n, ntri, npent, nhept;                /* out */
    double sigma, tol;                        /* out */
    int calls;                                /* out: ADMM calls of the node up to this one */
} madam_call_state;

/* What the reference keeps in process globals for bounding and branching:
 * the instance (SP), the parameters, the incumbent (heap.c:8-9), diff
 * (global_var.h:13), plus this library's random-stream state and counters. */
typedef struct {
    int N;                          /* SP->n: vertices; vertex N-1 is fixed to side 0 */
    const double *SPL;              /* SP->L, N x N (process_input.c:235-299) */
    madam_params params;            /* in/out: Pent/Hepta_Trials grow (bounding.c:184-187) */
    double LB;                      /* incumbent value (Bab_LBGet) */
    int best_x[MADAM_NMAX];         /* incumbent 0/1 vector (BabSol->X) */
    double diff;                    /* basic SDP bound - cut bound at the root */
    unsigned long long seed;        /* random streams of SA and GW rounding */
    long long stream_ctr;           /* in/out: next stream id */
    int projection;                 /* as madam_opts.projection */
    int batch;                      /* nodes bounded together by madam_bab */
    long long nodes;                /* out: evaluated nodes */
    long long admm_iters, admm_calls, ipm_iters;
    long long open_nodes;           /* out: nodes left when max_nodes stopped the search */
    double root_bound;
    double seconds_ipm, seconds_admm, seconds_sep, seconds_heur;  /* host wall time */
    double cg_rtol;                 /* PCG relative residual of the cut solve (<= 0: 1e-12) */
    long long cg_iters;             /* out: PCG iterations */
    long long ns_fallbacks;         /* out: sign-iteration projections redone by FP64 Jacobi */
    int verbose;                    /* > 0: one stderr line per cutting-plane round / B&B batch */
    int ns_layout;                  /* as madam_opts.ns_layout */
    int own_incumbent;              /* madam_sdpbound_batch: each node starts from LB and prunes
                                     * against its own heuristic's best only, as separate
                                     * SDPbound calls would; LB ends as the best of them.
                                     * madam_bab ignores it (one shared incumbent). */
    double cg_adapt;                /* as madam_opts.cg_adapt */
    int gw_mult;                    /* hyperplanes per rounding-heuristic call: gw_mult x N
                                     * (<= 1: N, as heuristic.c) */
    madam_call_state *last_call;    /* optional (null: off): madam_sdpbound_batch writes the
                                     * inputs of its first node's latest ADMM call here */
} madam_bab_ctx;

/* createSubproblem + getFixedValue (evaluate.c:39-139): PPL gets the n x n
 * subproblem matrix (n = free vertices + 1), *fixedvalue the constant. */
int madam_create_subproblem(int N, const double *SPL, const int *xfixed, const int *solX,
                            double *PPL, int *n_out, double *fixedvalue);

/* createSubproblem (evaluate.c:39-116) for k nodes on the device: node q's
 * fixings are xfixed[q*(N-1) ..] and solX[q*(N-1) ..]; its n_out[q] x
 * n_out[q] matrix is written to PPL + q*N*N (leading dimension n_out[q]),
 * bit-identical to madam_create_subproblem's.  The batched driver builds
 * its nodes this way. */
int madam_create_subproblem_batch(int N, const double *SPL, int k, const int *xfixed,
                                  const int *solX, double *PPL, int *n_out);

/* readData (process_input.c:182-312): the rudy file at path ("n m", then m
 * lines "i j w" with 1-based vertices) into SP->L (N x N, N = n <= capN),
 * the same construction: the Laplacian of vertices 1..n-1, its row sums in
 * the last row and column, their total in the corner.  A repeated edge
 * overwrites the earlier weight.  MADAM_E_ARG, with the reason in
 * madam_last_error(), for an unreadable file, header or edge, a vertex out
 * of range, or n outside 2..min(capN, MADAM_NMAX).  Host only: needs no
 * GPU. */
int madam_read_rudy(const char *path, int capN, int *N, double *SPL);

/* SDPbound (bounding.c:25) for k nodes at once: ranks[q] = 0 marks the
 * root (bounding.c:111, 241); bounds[q] = the node's upper bound, also
 * written to nodes[q]->upper_bound, with fracsol filled.  The incumbent in
 * ctx is updated by the heuristic as updateSolution does. */
int madam_sdpbound_batch(madam_bab_ctx *ctx, int k, madam_node *const *nodes, const int *ranks,
                         double *bounds);

/* Single-process branch and bound (bab_functions.c:39-69, 226-334,
 * heap.c) bounding ctx->batch nodes per step.  max_nodes <= 0: no limit.
 * *optimum = ctx->LB at the end (the proven optimum when open_nodes = 0). */
int madam_bab(madam_bab_ctx *ctx, long long max_nodes, double *optimum);


/* ---- Learned warm start (no reference counterpart: the reference starts
 * each node's ADMM from its interior-point solution).  A small MLP maps the
 * features of each vertex pair (i, j) of a node's cost matrix to a
 * prediction of the converged ADMM primal X*_ij; the prediction, made PSD
 * with unit diagonal, and the dual y_i = (L X)_ii with Q = (Diag(y) - L)_+
 * that complementarity gives, warm-start ADMM_solver.  Cold-start ADMM
 * stays the fallback: the safe bound does not depend on the start. */
#define MADAM_NFEAT 10
typedef struct {
    int nin;        /* features per pair (MADAM_NFEAT) */
    int nh;         /* hidden width, must be 64 */
    float *params;  /* FP32 master parameters, madam_mlp_nparams(nin) floats:
                       W1[nin][64], b1[64], W2[64][64], b2[64], W3[64], b3 */
} madam_mlp;
int madam_mlp_nparams(int nin);
/* Pair features of L (n x n): n(n-1)/2 rows of MADAM_NFEAT floats, pairs
 * (0,1), (0,2), ..., (1,2), ...; standardised by mu/sd when both given. */
int madam_pair_features(int n, const double *L, const float *mu, const float *sd, float *feat);
/* out[s] = network(feat[s]) for B samples (BF16 weights, FP32 sums). */
int madam_mlp_forward(const madam_mlp *net, int B, const float *feat, float *out);
/* Gradient of mean((out - target)^2 / 2) over B samples, and that loss. */
int madam_mlp_grad(const madam_mlp *net, int B, const float *feat, const float *target,
                   float *grad, float *loss);
/* Adam training on the GPU over N samples: epochs x (N / batch) steps of
 * shuffled minibatches (order from seed); net->params updated in place;
 * losses (may be null) receives each step's minibatch loss. */
int madam_mlp_train(madam_mlp *net, long long N, const float *feat, const float *target,
                    int epochs, int batch, float lr, unsigned long long seed, float *losses);
/* Warm start for ADMM_solver from predicted pair values pred (n(n-1)/2):
 * X = the PSD part of the prediction, rescaled to unit diagonal; y, Q as
 * above.  X, Q n x n, y n. */
int madam_warm_start(int n, const double *L, const float *pred, double *X, double *y, double *Q);

#ifdef __cplusplus
}
#endif
#endif
