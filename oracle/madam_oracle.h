/*
 * madam_oracle.h -- CPU FP64 restatement of MADAM's SDP bounding path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (madam_amd/, include/)
 * links, loads or calls this code.  tests/ and __graft_entry__.smoke()
 * compare the GPU results with it; bench.py's cpu_baseline leg times it as
 * the CPU path.
 *
 * Every function restates one reference function (file:line cited at its
 * definition) in this project's own code.  LAPACK (dsyevr, dpotrf, dpotrs,
 * dpotri, dposv, dsymm, dsymv) is called where the reference calls it, so the
 * CPU path is "LAPACK ADMM".  The one deliberate difference: the reference
 * factors BB^T+I with CHOLMOD (sparse, not in this image); the oracle factors
 * the same matrix densely with dpotrf, which gives the same solution up to
 * rounding.
 *
 * Random numbers: the reference draws from glibc rand() (srand(2020) at
 * bab_functions.c:82).  The oracle draws from an mo_rng, which is either
 * glibc rand() itself (so the oracle reproduces the reference's stream) or an
 * explicit array of values in [0, RAND_MAX] (so a GPU kernel fed the same
 * array can be compared draw for draw).
 *
 * Pinning (which check fixes which function) is listed in the header of
 * tests/test_oracle.py.  Functions with no pin say "parity unpinned" below.
 */
#ifndef MADAM_ORACLE_H
#define MADAM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define MO_NMAX 1024          /* admm.h:131 NMAX */
#define MO_BIG 1e9            /* admm.h:31 BIG_NUMBER */
#define MO_MAX_TRI 50000      /* admm.h:34-36 MaxTri/Pent/HeptaIneqAdded */

/* Same field order and types as Parameters (admm.h:57-89). */
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
} mo_params;

/* Same layouts as Triangle/Pentagonal/Heptagonal_Inequality (admm.h:93-117). */
typedef struct { int i, j, k, type; double value, y; } mo_tri;
typedef struct { int type; int perm[5]; double value, y; } mo_pent;
typedef struct { int type; int perm[7]; double value, y; } mo_hept;

/* Same layout as BabNode (admm.h:147-155). */
typedef struct {
    int xfixed[MO_NMAX];
    int solX[MO_NMAX];
    double fracsol[MO_NMAX];
    int level;
    double upper_bound;
} mo_node;

/* Random source: glibc rand() when stream == NULL, else stream[pos++]. */
typedef struct {
    const int *stream;
    long pos;
    long len;
} mo_rng;

/* All state the reference keeps in process globals (global_var.h:6-51,
 * heap.c:7-12), gathered in one struct. */
typedef struct {
    mo_params params;
    int N;                 /* SP->n */
    double *SPL;           /* SP->L, N x N */
    int n;                 /* PP->n */
    double *PPL;           /* PP->L, capacity N x N */
    double *X, *Q, *y, *Z, *W;
    double *u, *t, *s;     /* capacity 3 * MO_MAX_TRI */
    double f;
    int M;
    mo_tri *cuts, *tri_list;
    mo_pent *pcuts, *pent_list;
    mo_hept *hcuts, *hept_list;
    int ntri, npent, nhept;
    double LB;             /* incumbent value (heap.c:9) */
    int best_x[MO_NMAX];   /* incumbent (heap.c:8) */
    double diff;           /* global_var.h:13 */
    double root_bound;
    long nodes;            /* evaluated nodes */
    long admm_iters;       /* sum of nbit over all ADMM calls */
    mo_rng rng;
    double lb_override;    /* <0: none; else LB is pinned to this at start */
} mo_state;

/* ---- operators (operators.c) ---- */
void mo_laplacian_from_edges(int nv, int ne, const int *ei, const int *ej,
                             const double *w, double *L);
int mo_create_subproblem(int N, int babsize, const double *SPL,
                         const int *xfixed, const int *solX, double *PPL);
double mo_fixed_value(int N, int babsize, const double *SPL,
                      const int *xfixed, const int *solX);
void mo_op_B(int n, const double *X, int ntri, const mo_tri *tri,
             int npent, const mo_pent *pent, int nhept, const mo_hept *hept,
             double *y);
void mo_op_Bt(int n, double *X, int ntri, const mo_tri *tri,
              int npent, const mo_pent *pent, int nhept, const mo_hept *hept,
              const double *t);
int mo_project_psd(int n, const double *Wm, double *Xp, double *Q,
                   double *Z, double *evals, int *M);
double mo_lambda_max(int n, const double *A);

/* ---- ADMM (admm.c) ---- */
void mo_admm(mo_state *st, double *sigma, double tol, int *nbit);

/* ---- separation (cutting_planes.c, qap_simulated_annealing.c) ---- */
double mo_tri_separate(int n, const double *X, int K, double thresh,
                       mo_tri *list, int *size);
double mo_tri_update(mo_state *st, const double *dual, int *added, int *removed);
double mo_qap_sa(const int *H, int k, const double *X, int n, int *perm_out,
                 mo_rng *rng);
double mo_pent_separate(const double *X, int n, int trials, int K,
                        mo_pent *list, int *size, mo_rng *rng);
double mo_hept_separate(const double *X, int n, int trials, int K,
                        mo_hept *list, int *size, mo_rng *rng);
double mo_pent_update(mo_state *st, const double *dual, int offset,
                      int *added, int *removed);
double mo_hept_update(mo_state *st, const double *dual, int offset,
                      int *added, int *removed);
const int *mo_pent_H(int type);   /* 5x5 sign pattern, type 1..3 */
const int *mo_hept_H(int type);   /* 7x7 sign pattern, type 1..4 */

/* ---- heuristic (heuristic.c) ---- */
double mo_one_opt(int *x, int n, const double *L);
double mo_gw(mo_state *st, const mo_node *node, int *x, int num);
double mo_gw_explicit(int n, const double *L, const double *Z, int M,
                      const double *planes, int num, int *cut_out,
                      double *val_out);

/* ---- IPM (ipm_mc_pk.c) ---- */
int mo_ipm(const double *L, int n, double *X, double *y, double *Z,
           double *phi, int *iters);

/* ---- bounding / B&B (bounding.c, evaluate.c, bab_functions.c, heap.c) ---- */
void mo_default_params(mo_params *p);
mo_state *mo_state_new(int N, const double *SPL, const mo_params *p);
void mo_state_free(mo_state *st);
double mo_evaluate_solution(const mo_state *st, const int *sol);
int mo_update_solution(mo_state *st, const int *x);
double mo_sdpbound(mo_state *st, mo_node *node, int rank);
double mo_evaluate(mo_state *st, mo_node *node, int rank);
int mo_branch_var(const mo_state *st, const mo_node *node);
double mo_bab(mo_state *st, long max_nodes);

#ifdef __cplusplus
}
#endif
#endif
