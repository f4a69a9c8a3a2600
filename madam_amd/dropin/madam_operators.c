/*
 * madam_operators.c -- the reference's operator and annealing entry points
 * on the GPU through libmadam_bound, with the reference's prototypes
 * (admm.h:272-277, 286) and the globals they read and write:
 *
 *   diag, Diag, op_Diag       operators.c:32-54, 157-164 (O(n) host loops)
 *   op_B, op_Bt               operators.c:62-150, 171-513  -> madam_op_B, madam_op_Bt
 *   project_SDP               operators.c:521-619          -> madam_project_sdp
 *   qap_simulated_annealing   qap_simulated_annealing.c:12 -> madam_qap_sa
 *
 * Compiled inside MADAM's own tree in place of operators.c and
 * qap_simulated_annealing.c (madam_evaluate.c replaces evaluate.c).  Side
 * effects are kept where the reference's callers can see them: op_B
 * mirrors X's lower triangle to its upper (operators.c:68-71); project_SDP
 * leaves X_+ in X's lower triangle, Q = X_+ - X with a zero upper
 * triangle, the eigenvalues above 1e-8 in W, Z = eigenvectors *
 * sqrt(eigenvalue) (ascending) and their count in M (operators.c:531-617);
 * qap_simulated_annealing mirrors X's lower triangle to its upper
 * (qap_simulated_annealing.c:15-17) and writes the best k-tuple to its last
 * argument.  The annealing's random stream is
 * the library's (one run seeded by one glibc rand() draw, so the caller's
 * sequence of calls stays reproducible) rather than rand() itself.  Library
 * failures end the process with a message, as the reference does on LAPACK
 * errors.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "admm.h"
#include "madam_bound.h"

extern double *X, *Q, *Z, *W;
extern int M;
extern Triangle_Inequality *Cuts;
extern Pentagonal_Inequality *Pent_Cuts;
extern Heptagonal_Inequality *Hepta_Cuts;

static void ops_fatal(const char *where) {
    fprintf(stderr, "%s: %s\n", where, madam_last_error());
    exit(1);
}

static void ops_once(void) {
    static int opened = 0;
    if (opened) return;
    if (madam_open(0) != MADAM_OK) ops_fatal("madam_open");
    opened = 1;
}

static void *ops_alloc(size_t bytes) {
    void *p = malloc(bytes ? bytes : 1);
    if (!p) {
        fprintf(stderr, "madam_operators: out of host memory\n");
        exit(1);
    }
    return p;
}

/* lower triangle -> upper triangle, in place */
static void mirror_lower(double *A, int n) {
    for (int j = 0; j < n; ++j)
        for (int i = j + 1; i < n; ++i) A[j + (size_t)i * n] = A[i + (size_t)j * n];
}

void diag(const double *Xm, double *y, int n) {
    for (int i = 0; i < n; ++i) y[i] = Xm[i + (size_t)i * n];
}

void Diag(double *Xm, const double *y, int n) {
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) Xm[j + (size_t)i * n] = i == j ? y[i] : 0.0;
}

void op_Diag(const Problem *P, double *Xm, const double *yy) {
    for (int i = 0; i < P->n; ++i) Xm[i + (size_t)i * P->n] -= yy[i];
}

void op_B(const Problem *P, double *y, double *Xm) {
    ops_once();
    mirror_lower(Xm, P->n);
    if (madam_op_B(P->n, Xm, P->NIneq, (const madam_tri *)Cuts, P->NPentIneq,
                   (const madam_pent *)Pent_Cuts, P->NHeptaIneq, (const madam_hept *)Hepta_Cuts,
                   y) != MADAM_OK)
        ops_fatal("op_B");
}

void op_Bt(const Problem *P, double *Xm, const double *tt) {
    ops_once();
    if (madam_op_Bt(P->n, Xm, P->NIneq, (const madam_tri *)Cuts, P->NPentIneq,
                    (const madam_pent *)Pent_Cuts, P->NHeptaIneq,
                    (const madam_hept *)Hepta_Cuts, tt) != MADAM_OK)
        ops_fatal("op_Bt");
}

void project_SDP(const Problem *P) {
    ops_once();
    const int n = P->n;
    const size_t nn = (size_t)n * n;
    /* the reference reads X's lower triangle (UPLO = 'L') */
    double *Wm = ops_alloc(nn * sizeof(double));
    double *Xp = ops_alloc(nn * sizeof(double));
    double *Qp = ops_alloc(nn * sizeof(double));
    memcpy(Wm, X, nn * sizeof(double));
    mirror_lower(Wm, n);
    int m = 0;
    if (madam_project_sdp(n, Wm, Xp, Qp, Z, &m) != MADAM_OK) ops_fatal("project_SDP");
    /* X: lower = X_+, upper as it was; Q = X_out - X_in (upper triangle 0) */
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            const size_t e = i + (size_t)j * n;
            const double xin = X[e];
            if (i >= j) X[e] = Xp[e];
            Q[e] = X[e] - xin;
        }
    M = m;
    for (int c = 0; c < m; ++c) {
        double s2 = 0.0;
        for (int i = 0; i < n; ++i) s2 += Z[i + (size_t)c * n] * Z[i + (size_t)c * n];
        W[c] = s2; /* column c of Z is v * sqrt(lambda) */
    }
    free(Wm);
    free(Xp);
    free(Qp);
}

double qap_simulated_annealing(int *H, int k, double *Xm, int n, int *ineq) {
    ops_once();
    mirror_lower(Xm, n);
    /* H = e e^T with e_a = -1 for the first type-1 entries
     * (cutting_planes.c:242-244, 438-441): type from H's first row */
    int run = 1;
    while (run < k && H[k * run] == H[0]) ++run;
    const int type = run == k ? 1 : 1 + run;
    const long long stream = 0;
    const unsigned long long seed = (unsigned long long)rand();
    double cost = 0.0;
    long long draws = 0;
    int *perm = ops_alloc((size_t)k * sizeof(int));
    if (madam_qap_sa(n, Xm, k, 1, &type, &stream, seed, &cost, perm, &draws) != MADAM_OK)
        ops_fatal("qap_simulated_annealing");
    for (int a = 0; a < k; ++a) ineq[a] = perm[a];
    free(perm);
    return cost;
}
