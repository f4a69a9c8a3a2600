/*
 * madam_oracle.c -- CPU FP64 restatement of MADAM's SDP bounding path.
 * TEST INFRASTRUCTURE ONLY (see madam_oracle.h).
 *
 * Conventions.  Matrices are column-major n x n with leading dimension n.
 * The reference keeps only the lower triangle meaningful (LAPACK UPLO='L',
 * op_B mirrors lower to upper before reading, operators.c:65-69); the oracle
 * keeps every matrix fully symmetric and reads the lower triangle wherever
 * the reference does, which gives the same numbers.
 *
 * Cut rows.  Every cut is a vertex tuple v[0..k-1] (k = 3, 5, 7) with a sign
 * vector b (b_a = -1 for the first `neg` entries of a pentagonal/heptagonal
 * tuple, and for one vertex of a triangle, see cut_expand) and
 *     B(X)_c = -(1/g) * sum_{a<b} b_a b_b X[v_a, v_b],  g = (k-1)/2,
 * which is the row op_B evaluates (operators.c:85-148); B^T(t) puts
 * -(b_a b_b / g) t / 2 on both (v_a,v_b) and (v_b,v_a) (operators.c:171-513).
 */
#include "madam_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ---- LAPACK / BLAS (Fortran ABI, as declared in admm.h:167-205) ---- */
extern void dsyevr_(const char *jobz, const char *range, const char *uplo,
                    const int *n, double *a, const int *lda, const double *vl,
                    const double *vu, const int *il, const int *iu,
                    const double *abstol, int *m, double *w, double *z,
                    const int *ldz, int *isuppz, double *work,
                    const int *lwork, int *iwork, const int *liwork, int *info);
extern double dlansy_(const char *norm, const char *uplo, const int *n,
                      const double *a, const int *lda, double *work);
extern void dsyrk_(const char *uplo, const char *trans, const int *n,
                   const int *k, const double *alpha, const double *a,
                   const int *lda, const double *beta, double *c,
                   const int *ldc);
extern void dpotrf_(const char *uplo, const int *n, double *a, const int *lda,
                    int *info);
extern void dpotrs_(const char *uplo, const int *n, const int *nrhs,
                    const double *a, const int *lda, double *b,
                    const int *ldb, int *info);
extern void dpotri_(const char *uplo, const int *n, double *a, const int *lda,
                    int *info);
extern void dposv_(const char *uplo, const int *n, const int *nrhs, double *a,
                   const int *lda, double *b, const int *ldb, int *info);
extern void dsymm_(const char *side, const char *uplo, const int *m,
                   const int *n, const double *alpha, const double *a,
                   const int *lda, const double *b, const int *ldb,
                   const double *beta, double *c, const int *ldc);
extern void dsymv_(const char *uplo, const int *n, const double *alpha,
                   const double *a, const int *lda, const double *x,
                   const int *incx, const double *beta, double *y,
                   const int *incy);

static void *xcalloc(size_t n, size_t sz) {
    void *p = calloc(n ? n : 1, sz);
    if (!p) {
        fprintf(stderr, "oracle: out of memory\n");
        exit(1);
    }
    return p;
}

static inline double lowr(const double *X, int n, int a, int b) {
    return a > b ? X[a + (size_t)b * n] : X[b + (size_t)a * n];
}

static void symmetrize_from_lower(double *A, int n) {
    for (int j = 0; j < n; ++j)
        for (int i = j + 1; i < n; ++i) A[j + (size_t)i * n] = A[i + (size_t)j * n];
}

/* ------------------------------------------------------------------ */
/* random source                                                       */
/* ------------------------------------------------------------------ */
static inline int rng_next(mo_rng *r) {
    if (!r || !r->stream) return rand();
    if (r->pos >= r->len) {
        fprintf(stderr, "oracle: explicit random stream exhausted\n");
        exit(1);
    }
    return r->stream[r->pos++];
}

/* ------------------------------------------------------------------ */
/* problem construction (process_input.c:235-299, evaluate.c:39-139)   */
/* ------------------------------------------------------------------ */

/* SP->L from a rudy edge list (1-based endpoints).  Upper-left block is the
 * Laplacian of the graph restricted to vertices 0..nv-2; its row sums fill
 * the last row/column and their total the corner (process_input.c:235-299).
 * A repeated edge overwrites, as the adjacency assignment does
 * (process_input.c:221-222). */
void mo_laplacian_from_edges(int nv, int ne, const int *ei, const int *ej,
                             const double *w, double *L) {
    double *A = xcalloc((size_t)nv * nv, sizeof(double));
    for (int e = 0; e < ne; ++e) {
        int a = ei[e] - 1, b = ej[e] - 1;
        A[a + (size_t)b * nv] = w[e];
        A[b + (size_t)a * nv] = w[e];
    }
    double *deg = xcalloc(nv, sizeof(double));
    for (int a = 0; a < nv; ++a)
        for (int b = 0; b < nv; ++b) deg[a] += A[b + (size_t)a * nv];
    double total = 0.0;
    for (int a = 0; a < nv - 1; ++a) {
        double rs = 0.0;
        for (int b = 0; b < nv - 1; ++b) {
            double v = (a == b ? deg[a] : 0.0) - A[b + (size_t)a * nv];
            L[b + (size_t)a * nv] = v;
            rs += v;
        }
        L[(nv - 1) + (size_t)a * nv] = rs;
        total += rs;
    }
    for (int b = 0; b < nv - 1; ++b)
        L[b + (size_t)(nv - 1) * nv] = L[(nv - 1) + (size_t)b * nv];
    L[(nv - 1) + (size_t)(nv - 1) * nv] = total;
    free(A);
    free(deg);
}

/* createSubproblem (evaluate.c:39-116): keep free rows/cols of the Laplacian
 * block, fold rows of vertices fixed to 1 into the linear part, build the
 * last row/col and corner, scale by 1/4.  Returns PP->n. */
int mo_create_subproblem(int N, int babsize, const double *SPL,
                         const int *xfixed, const int *solX, double *PPL) {
    int nfree = 0;
    for (int i = 0; i < babsize; ++i) nfree += !xfixed[i];
    int n = nfree + 1;
    double *frow = xcalloc(nfree > 0 ? nfree : 1, sizeof(double));
    int col = 0;
    for (int i = 0; i < babsize; ++i) {
        if (xfixed[i]) {
            if (solX[i] == 1) {
                int f = 0;
                for (int j = 0; j < babsize; ++j)
                    if (!xfixed[j]) frow[f++] += SPL[j + (size_t)i * N];
            }
            continue;
        }
        int r = 0;
        double rs = 0.0;
        for (int j = 0; j < babsize; ++j) {
            if (xfixed[j]) continue;
            double v = SPL[j + (size_t)i * N];
            PPL[r + (size_t)col * n] = v;
            rs += v;
            ++r;
        }
        PPL[(n - 1) + (size_t)col * n] = rs;
        ++col;
    }
    for (int i = 0; i < n - 1; ++i)
        PPL[i + (size_t)(n - 1) * n] = PPL[(n - 1) + (size_t)i * n];
    double corner = 0.0;
    for (int i = 0; i < n - 1; ++i) {
        PPL[i + (size_t)i * n] += 2 * frow[i];
        PPL[(n - 1) + (size_t)i * n] += 2 * frow[i];
        PPL[i + (size_t)(n - 1) * n] += 2 * frow[i];
        corner += PPL[i + (size_t)(n - 1) * n];
    }
    PPL[(n - 1) + (size_t)(n - 1) * n] = corner;
    for (size_t e = 0; e < (size_t)n * n; ++e) PPL[e] *= 0.25;
    free(frow);
    return n;
}

/* getFixedValue (evaluate.c:124-139). */
double mo_fixed_value(int N, int babsize, const double *SPL,
                      const int *xfixed, const int *solX) {
    double v = 0.0;
    for (int i = 0; i < babsize; ++i)
        for (int j = 0; j < babsize; ++j)
            if (xfixed[i] && xfixed[j])
                v += SPL[j + (size_t)i * N] * solX[i] * solX[j];
    return v;
}

/* ------------------------------------------------------------------ */
/* cut maps B and B^T (operators.c:62-150, 171-513)                    */
/* ------------------------------------------------------------------ */

/* Vertex tuple and sign vector of a cut; returns k.  Triangle types 2/3/4
 * negate vertex k/j/i (operators.c:85-95); pentagonal/heptagonal type t
 * negates the first t-1 tuple entries (cutting_planes.c:239-244, 434-441). */
static int tri_expand(const mo_tri *c, int *v, int *b) {
    v[0] = c->i; v[1] = c->j; v[2] = c->k;
    b[0] = b[1] = b[2] = 1;
    if (c->type == 2) b[2] = -1;
    else if (c->type == 3) b[1] = -1;
    else if (c->type == 4) b[0] = -1;
    return 3;
}
static int kgon_expand(int k, int type, const int *perm, int *v, int *b) {
    for (int a = 0; a < k; ++a) {
        v[a] = perm[a];
        b[a] = (a < type - 1) ? -1 : 1;
    }
    return k;
}

/* sum_{a<b} (-b_a b_b) X[v_a, v_b], pairs in lexicographic order. */
static double cut_sum(const double *X, int n, int k, const int *v, const int *b) {
    double acc = 0.0;
    for (int a = 0; a < k; ++a)
        for (int c = a + 1; c < k; ++c) {
            double x = lowr(X, n, v[a], v[c]);
            acc += (b[a] * b[c] > 0) ? -x : x;
        }
    return acc;
}

void mo_op_B(int n, const double *X, int ntri, const mo_tri *tri,
             int npent, const mo_pent *pent, int nhept, const mo_hept *hept,
             double *y) {
    int v[7], b[7];
    for (int c = 0; c < ntri; ++c) {
        tri_expand(&tri[c], v, b);
        y[c] += cut_sum(X, n, 3, v, b);
    }
    for (int c = 0; c < npent; ++c) {
        kgon_expand(5, pent[c].type, pent[c].perm, v, b);
        y[ntri + c] += 0.5 * cut_sum(X, n, 5, v, b);
    }
    for (int c = 0; c < nhept; ++c) {
        kgon_expand(7, hept[c].type, hept[c].perm, v, b);
        y[ntri + npent + c] += 1.0 / 3.0 * cut_sum(X, n, 7, v, b);
    }
}

static void scatter_cut(double *X, int n, int k, const int *v, const int *b,
                        double temp) {
    for (int a = 0; a < k; ++a)
        for (int c = a + 1; c < k; ++c) {
            double d = (b[a] * b[c] > 0) ? temp : -temp;
            X[v[a] + (size_t)v[c] * n] += d;
            X[v[c] + (size_t)v[a] * n] += d;
        }
}

void mo_op_Bt(int n, double *X, int ntri, const mo_tri *tri,
              int npent, const mo_pent *pent, int nhept, const mo_hept *hept,
              const double *t) {
    int v[7], b[7];
    for (int c = 0; c < ntri; ++c) {
        tri_expand(&tri[c], v, b);
        scatter_cut(X, n, 3, v, b, 0.5 * t[c]);
    }
    for (int c = 0; c < npent; ++c) {
        kgon_expand(5, pent[c].type, pent[c].perm, v, b);
        scatter_cut(X, n, 5, v, b, 0.25 * t[ntri + c]);
    }
    for (int c = 0; c < nhept; ++c) {
        kgon_expand(7, hept[c].type, hept[c].perm, v, b);
        scatter_cut(X, n, 7, v, b, 1.0 / 6.0 * t[ntri + npent + c]);
    }
}

/* ------------------------------------------------------------------ */
/* PSD projection (operators.c:521-619) and lambda_max (admm.c:724-752)*/
/* ------------------------------------------------------------------ */

/* Xp = W_+ restricted to eigenvalues in (1e-8, min(|W|_inf, |W|_F)],
 * Q = Xp - W, Z = V diag(sqrt(lambda)) (n x M), evals[0..M-1] ascending.
 * Returns the dsyevr info (0 on success). */
int mo_project_psd(int n, const double *Wm, double *Xp, double *Q,
                   double *Z, double *evals, int *M) {
    const char U = 'L', NI = 'I', NF = 'F', JV = 'V', RV = 'V', TR = 'N';
    size_t nn = (size_t)n * n;
    double *A = xcalloc(nn, sizeof(double));
    double *work = xcalloc(26 * (size_t)n + 64, sizeof(double));
    int *iwork = xcalloc(10 * (size_t)n + 64, sizeof(int));
    int *isuppz = xcalloc(2 * (size_t)n + 2, sizeof(int));
    int lwork = 26 * n + 64, liwork = 10 * n + 64, il = 0, iu = 0, info = 0;
    memcpy(A, Wm, nn * sizeof(double));
    double ninf = dlansy_(&NI, &U, &n, A, &n, work);
    double nfro = dlansy_(&NF, &U, &n, A, &n, work);
    double vl = 1e-8, vu = ninf < nfro ? ninf : nfro, abstol = 1e-8;
    if (vu < vl) vu = 2.0 * vl;
    dsyevr_(&JV, &RV, &U, &n, A, &n, &vl, &vu, &il, &iu, &abstol, M, evals, Z,
            &n, isuppz, work, &lwork, iwork, &liwork, &info);
    if (info != 0) {
        /* retry as operators.c:583-593 does, on an intact copy */
        memcpy(A, Wm, nn * sizeof(double));
        vl = 0.0;
        abstol = 0.0;
        dsyevr_(&JV, &RV, &U, &n, A, &n, &vl, &vu, &il, &iu, &abstol, M, evals,
                Z, &n, isuppz, work, &lwork, iwork, &liwork, &info);
    }
    if (info == 0) {
        for (int j = 0; j < *M; ++j) {
            double sc = sqrt(evals[j]);
            for (int i = 0; i < n; ++i) Z[i + (size_t)j * n] *= sc;
        }
        double one = 1.0, zero = 0.0;
        int m = *M;
        if (m > 0) {
            dsyrk_(&U, &TR, &n, &m, &one, Z, &n, &zero, Xp, &n);
        } else {
            memset(Xp, 0, nn * sizeof(double));
        }
        symmetrize_from_lower(Xp, n);
        for (size_t e = 0; e < nn; ++e) Q[e] = Xp[e] - Wm[e];
    }
    free(A);
    free(work);
    free(iwork);
    free(isuppz);
    return info;
}

/* Largest eigenvalue, dsyevr RANGE='I', IL=IU=n, ABSTOL=1e-8 (admm.c:731-747). */
double mo_lambda_max(int n, const double *Ain) {
    const char U = 'L', JN = 'N', RI = 'I';
    size_t nn = (size_t)n * n;
    double *A = xcalloc(nn, sizeof(double));
    double *work = xcalloc(26 * (size_t)n + 64, sizeof(double));
    double *zdummy = xcalloc(n, sizeof(double));
    int *iwork = xcalloc(10 * (size_t)n + 64, sizeof(int));
    int *isuppz = xcalloc(2 * (size_t)n + 2, sizeof(int));
    /* dsyevr's W must hold n values whatever the range (it is also used
     * as workspace); the selected eigenvalue is wv[0] */
    double *wv = xcalloc(n, sizeof(double));
    int lwork = 26 * n + 64, liwork = 10 * n + 64, il = n, iu = n, m = 0, info;
    double vl = 0, vu = 0, abstol = 1e-8;
    memcpy(A, Ain, nn * sizeof(double));
    dsyevr_(&JN, &RI, &U, &n, A, &n, &vl, &vu, &il, &iu, &abstol, &m, wv,
            zdummy, &n, isuppz, work, &lwork, iwork, &liwork, &info);
    if (info != 0) {
        fprintf(stderr, "oracle: eigenvalue computation failed\n");
        exit(1);
    }
    const double lam = wv[0];
    free(wv);
    free(A);
    free(work);
    free(zdummy);
    free(iwork);
    free(isuppz);
    return lam;
}

/* ------------------------------------------------------------------ */
/* ADMM (admm.c:49-771)                                                */
/* ------------------------------------------------------------------ */

typedef struct {
    int k;
    int v[7];
    int b[7];
    double g; /* 1, 2, 3 */
} cutrow;

static int gather_rows(const mo_state *st, cutrow **out) {
    int m = st->ntri + st->npent + st->nhept;
    cutrow *r = xcalloc(m, sizeof(cutrow));
    for (int c = 0; c < st->ntri; ++c) {
        r[c].k = tri_expand(&st->cuts[c], r[c].v, r[c].b);
        r[c].g = 1.0;
    }
    for (int c = 0; c < st->npent; ++c) {
        cutrow *q = &r[st->ntri + c];
        q->k = kgon_expand(5, st->pcuts[c].type, st->pcuts[c].perm, q->v, q->b);
        q->g = 2.0;
    }
    for (int c = 0; c < st->nhept; ++c) {
        cutrow *q = &r[st->ntri + st->npent + c];
        q->k = kgon_expand(7, st->hcuts[c].type, st->hcuts[c].perm, q->v, q->b);
        q->g = 3.0;
    }
    *out = r;
    return m;
}

/* Dense BB^T + I (m x m): entry (c,d) = 1[c==d] + 1/2 sum over shared vertex
 * pairs of coef_c * coef_d, coef = -b_a b_b / g.  This is the matrix CHOLMOD
 * factors with beta=1 in admm.c:486-487 (B as built in admm.c:122-477 holds
 * coef/2 at both (a,b) and (b,a)). */
static double *build_bbt(int n, int m, const cutrow *r) {
    double *Mx = xcalloc((size_t)m * m, sizeof(double));
    /* incidence lists per unordered vertex pair */
    int *head = xcalloc((size_t)n * n, sizeof(int));
    for (size_t e = 0; e < (size_t)n * n; ++e) head[e] = -1;
    int total = 0;
    for (int c = 0; c < m; ++c) total += r[c].k * (r[c].k - 1) / 2;
    int *next = xcalloc(total, sizeof(int));
    int *cut = xcalloc(total, sizeof(int));
    double *coef = xcalloc(total, sizeof(double));
    int q = 0;
    for (int c = 0; c < m; ++c)
        for (int a = 0; a < r[c].k; ++a)
            for (int b = a + 1; b < r[c].k; ++b) {
                int va = r[c].v[a], vb = r[c].v[b];
                size_t key = va > vb ? (size_t)va * n + vb : (size_t)vb * n + va;
                cut[q] = c;
                coef[q] = -(double)(r[c].b[a] * r[c].b[b]) / r[c].g;
                next[q] = head[key];
                head[key] = q;
                ++q;
            }
    for (size_t key = 0; key < (size_t)n * n; ++key)
        for (int p1 = head[key]; p1 >= 0; p1 = next[p1])
            for (int p2 = head[key]; p2 >= 0; p2 = next[p2])
                Mx[cut[p1] + (size_t)cut[p2] * m] += 0.5 * coef[p1] * coef[p2];
    for (int c = 0; c < m; ++c) Mx[c + (size_t)c * m] += 1.0;
    free(head);
    free(next);
    free(cut);
    free(coef);
    return Mx;
}

void mo_admm(mo_state *st, double *sigma, double tol, int *nbit) {
    const int n = st->n;
    const size_t nn = (size_t)n * n;
    const double *L = st->PPL;
    double *X = st->X, *Q = st->Q, *y = st->y;
    double *u = st->u, *t = st->t, *s = st->s;
    const char U = 'L', NF = 'F';
    int inc1 = 1, info = 0;

    cutrow *rows;
    const int m = gather_rows(st, &rows);
    for (int c = 0; c < st->ntri; ++c) s[c] = t[c] = st->cuts[c].y;
    for (int c = 0; c < st->npent; ++c)
        s[st->ntri + c] = t[st->ntri + c] = st->pcuts[c].y;
    for (int c = 0; c < st->nhept; ++c)
        s[st->ntri + st->npent + c] = t[st->ntri + st->npent + c] = st->hcuts[c].y;
    for (int c = 0; c < m; ++c) u[c] = 0.0;

    double *Mx = NULL;
    if (m > 0) {
        Mx = build_bbt(n, m, rows);
        dpotrf_(&U, &m, Mx, &m, &info);
        if (info != 0) {
            fprintf(stderr, "oracle: BB^T+I factorization failed\n");
            exit(1);
        }
    }
    double *tmp = xcalloc(nn, sizeof(double));
    double *rhs = xcalloc(m, sizeof(double));
    double *v = xcalloc(m, sizeof(double));
    double *evals = xcalloc(n, sizeof(double));
    double *work = xcalloc(n + 8, sizeof(double));
    double normL = dlansy_(&NF, &U, &n, L, &n, work);

    *nbit = 0;
    while (*nbit <= st->params.ADMM_itermax) {
        ++(*nbit);
        const double sg = *sigma;
        /* y = diag(L + Q + X/sigma) - e/sigma  (admm.c:526-528) */
        for (int i = 0; i < n; ++i)
            y[i] = L[i + (size_t)i * n] + Q[i + (size_t)i * n] +
                   (X[i + (size_t)i * n] - 1.0) / sg;
        /* t = (BB^T+I)^{-1} (s + B(L+Q+X/sigma) + (u-e)/sigma) (admm.c:531-551) */
        for (size_t e = 0; e < nn; ++e) tmp[e] = L[e] + Q[e] + X[e] / sg;
        for (int c = 0; c < m; ++c) t[c] = s[c];
        mo_op_B(n, tmp, st->ntri, st->cuts, st->npent, st->pcuts, st->nhept,
                st->hcuts, t);
        for (int c = 0; c < m; ++c) t[c] += (u[c] - 1) / sg;
        if (m > 0) {
            dpotrs_(&U, &m, &inc1, Mx, &m, t, &m, &info);
        }
        /* s = (t - u/sigma)_+ (admm.c:557-560) */
        for (int c = 0; c < m; ++c) {
            v[c] = t[c] - u[c] / sg;
            s[c] = v[c] > 0 ? v[c] : 0;
        }
        /* W = L + X/sigma - Diag(y) - B^T(t), projected (admm.c:564-574) */
        for (size_t e = 0; e < nn; ++e) tmp[e] = L[e] + X[e] / sg;
        for (int i = 0; i < n; ++i) tmp[i + (size_t)i * n] -= y[i];
        mo_op_Bt(n, tmp, st->ntri, st->cuts, st->npent, st->pcuts, st->nhept,
                 st->hcuts, t);
        if (mo_project_psd(n, tmp, X, Q, st->Z, evals, &st->M) != 0) {
            fprintf(stderr, "oracle: eigenvalue computation failed\n");
            exit(1);
        }
        /* X = sigma X_+, Z = sqrt(sigma) Z, u = -sigma v_- (admm.c:580-591) */
        for (size_t e = 0; e < nn; ++e) X[e] *= sg;
        double sq = sqrt(sg);
        for (size_t e = 0; e < (size_t)n * st->M; ++e) st->Z[e] *= sq;
        for (int c = 0; c < m; ++c) u[c] = v[c] < 0 ? -sg * v[c] : 0;

        /* primal error (admm.c:600-620) */
        double ep = 0.0;
        for (int i = 0; i < n; ++i) {
            double d = X[i + (size_t)i * n] - 1.0;
            ep += d * d;
        }
        double err_p = sqrt(ep);
        for (int c = 0; c < m; ++c) rhs[c] = -1.0;
        mo_op_B(n, X, st->ntri, st->cuts, st->npent, st->pcuts, st->nhept,
                st->hcuts, rhs);
        double er = 0.0;
        for (int c = 0; c < m; ++c)
            if (rhs[c] > 0) er += rhs[c] * rhs[c];
        err_p += sqrt(er);
        err_p /= (1 + sqrt((double)n));
        /* dual error (admm.c:626-640) */
        double ed = 0.0;
        for (int c = 0; c < m; ++c) {
            double d = s[c] - t[c];
            ed += d * d;
        }
        double err_d = sqrt(ed);
        for (size_t e = 0; e < nn; ++e) tmp[e] = L[e] + Q[e];
        for (int i = 0; i < n; ++i) tmp[i + (size_t)i * n] -= y[i];
        mo_op_Bt(n, tmp, st->ntri, st->cuts, st->npent, st->pcuts, st->nhept,
                 st->hcuts, t);
        err_d += dlansy_(&NF, &U, &n, tmp, &n, work);
        err_d /= (1 + normL);
        /* objectives (admm.c:647-665) */
        double primal = 0.0, dual = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = i; j < n; ++j)
                primal += (i == j ? 1.0 : 2.0) * L[j + (size_t)i * n] *
                          X[j + (size_t)i * n];
        for (int i = 0; i < n; ++i) dual += y[i];
        for (int c = 0; c < m; ++c) dual += t[c];
        double gap = fabs(dual - primal) / (1 + fabs(dual) + fabs(primal));
        double err = err_p > err_d ? (err_p > gap ? err_p : gap)
                                   : (err_d > gap ? err_d : gap);
        double dp = log10(err_p), dd = log10(err_d);
        /* sigma update and stop test (admm.c:689-702) */
        if (dd - dp < -0.5)
            *sigma /= st->params.scaleSigma;
        else if (dd - dp > 0.5)
            *sigma *= st->params.scaleSigma;
        if (err < tol) break;
    }
    st->admm_iters += *nbit;

    /* safe bound: f = n lambda_max(L - Diag(y) - B^T(s)) + sum y + sum s
     * (admm.c:724-763) */
    memcpy(tmp, L, nn * sizeof(double));
    for (int i = 0; i < n; ++i) tmp[i + (size_t)i * n] -= y[i];
    mo_op_Bt(n, tmp, st->ntri, st->cuts, st->npent, st->pcuts, st->nhept,
             st->hcuts, s);
    double lam = mo_lambda_max(n, tmp);
    double f = n * lam;
    for (int i = 0; i < n; ++i) f += y[i];
    for (int c = 0; c < m; ++c) f += s[c];
    st->f = f;

    free(rows);
    free(Mx);
    free(tmp);
    free(rhs);
    free(v);
    free(evals);
    free(work);
}

/* ------------------------------------------------------------------ */
/* triangle separation (cutting_planes.c:20-214)                       */
/* ------------------------------------------------------------------ */

static inline double tri_value(const double *X, int n, int type, int i, int j, int k) {
    double xij = X[i + (size_t)j * n], xik = X[i + (size_t)k * n],
           xjk = X[j + (size_t)k * n];
    switch (type) {
    case 1: return -xij - xik - xjk - 1.0;
    case 2: return -xij + xik + xjk - 1.0;
    case 3: return xij - xik + xjk - 1.0;
    default: return xij + xik - xjk - 1.0;
    }
}

/* Enumerate type (outermost), then i > j > k ascending; keep at most K cuts
 * with value > thresh.  Once full, a candidate strictly larger than the
 * current minimum replaces the slot that holds the first-found minimum, and
 * the minimum is rescanned (cutting_planes.c:58-122).  Returns max value. */
double mo_tri_separate(int n, const double *X, int K, double thresh,
                       mo_tri *list, int *size) {
    int cnt = 0, minslot = 0;
    double minval = MO_BIG, best = -MO_BIG;
    for (int type = 1; type <= 4; ++type)
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < i; ++j)
                for (int k = 0; k < j; ++k) {
                    double val = tri_value(X, n, type, i, j, k);
                    if (val > best) best = val;
                    if (!(val > thresh)) continue;
                    if (cnt < K) {
                        mo_tri *e = &list[cnt];
                        e->type = type; e->i = i; e->j = j; e->k = k; e->value = val;
                        if (cnt == 0 || val < minval) {
                            minval = val;
                            minslot = cnt;
                        }
                        ++cnt;
                    } else if (val > minval) {
                        mo_tri *e = &list[minslot];
                        e->type = type; e->i = i; e->j = j; e->k = k; e->value = val;
                        minval = MO_BIG;
                        for (int q = 0; q < cnt; ++q)
                            if (list[q].value < minval) {
                                minval = list[q].value;
                                minslot = q;
                            }
                    }
                }
    *size = cnt;
    return best;
}

/* Purge cuts with dual < 1e-5, separate, append new cuts that are not equal
 * to a surviving one (cutting_planes.c:132-214). */
double mo_tri_update(mo_state *st, const double *dual, int *added, int *removed) {
    int keep = 0, rem = 0;
    for (int c = 0; c < st->ntri; ++c) {
        st->cuts[c].y = dual[c];
        if (st->cuts[c].y < 1e-5) {
            ++rem;
        } else {
            mo_tri tmp = st->cuts[c];
            st->cuts[keep].type = tmp.type;
            st->cuts[keep].i = tmp.i;
            st->cuts[keep].j = tmp.j;
            st->cuts[keep].k = tmp.k;
            st->cuts[keep].y = tmp.y;
            ++keep;
        }
    }
    st->ntri -= rem;
    int lsize = 0;
    double best = mo_tri_separate(st->n, st->X, st->params.TriIneq,
                                  st->params.violated_TriIneq, st->tri_list, &lsize);
    int add = 0;
    const int survivors = st->ntri;
    for (int q = 0; q < lsize; ++q) {
        if (keep == MO_MAX_TRI) break;
        const mo_tri *cand = &st->tri_list[q];
        int dup = 0;
        for (int c = 0; c < survivors; ++c)
            if (st->cuts[c].type == cand->type && st->cuts[c].i == cand->i &&
                st->cuts[c].j == cand->j && st->cuts[c].k == cand->k)
                dup = 1;
        if (dup) continue;
        st->cuts[keep] = *cand;
        st->cuts[keep].y = 0.0;
        ++keep;
        ++add;
    }
    st->ntri += add;
    *added = add;
    *removed = rem;
    return best;
}

/* ------------------------------------------------------------------ */
/* pentagonal / heptagonal separation (cutting_planes.c:224-605)       */
/* ------------------------------------------------------------------ */

/* H_t = e e^T with e_a = -1 for a < t-1 (cutting_planes.c:242-244, 438-441). */
static int PENT_H[3][25], HEPT_H[4][49];
static int h_ready = 0;
static void init_h(void) {
    if (h_ready) return;
    for (int t = 1; t <= 3; ++t)
        for (int a = 0; a < 5; ++a)
            for (int b = 0; b < 5; ++b)
                PENT_H[t - 1][b + 5 * a] = ((a < t - 1) ? -1 : 1) * ((b < t - 1) ? -1 : 1);
    for (int t = 1; t <= 4; ++t)
        for (int a = 0; a < 7; ++a)
            for (int b = 0; b < 7; ++b)
                HEPT_H[t - 1][b + 7 * a] = ((a < t - 1) ? -1 : 1) * ((b < t - 1) ? -1 : 1);
    h_ready = 1;
}
const int *mo_pent_H(int type) { init_h(); return PENT_H[type - 1]; }
const int *mo_hept_H(int type) { init_h(); return HEPT_H[type - 1]; }

/* Simulated annealing over permutations minimising <H, X[p,p]> for the
 * leading k entries of a random permutation p of 0..n-1
 * (qap_simulated_annealing.c:12-206).  perm_out gets the best k-tuple. */
double mo_qap_sa(const int *H, int k, const double *X, int n, int *perm_out,
                 mo_rng *rng) {
#define XV(a, b) lowr(X, n, (a), (b))
    double sumX = 0.0;
    for (int i = 0; i < n; ++i) {
        sumX += X[i + (size_t)i * n];
        for (int j = i + 1; j < n; ++j) sumX += 2 * fabs(X[j + (size_t)i * n]);
    }
    double temp0 = (double)(k * k) * sumX / (double)(n * (n - 1));
    double cost = MO_BIG;
    int *perm = xcalloc(n, sizeof(int));
    int *best = xcalloc(n, sizeof(int));
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int i = n - 1; i > 0; --i) {
        int r = rng_next(rng) % i;
        int sw = perm[i];
        perm[i] = perm[r];
        perm[r] = sw;
    }
    double sol = 0.0;
    for (int i = 0; i < k; ++i) {
        sol += XV(perm[i], perm[i]);
        for (int j = i + 1; j < k; ++j) sol += 2 * H[j + i * k] * XV(perm[j], perm[i]);
    }
    double T = temp0;
    int m1 = n;
    int changed = 1;
    while (changed > 0) {
        changed = 0;
        for (int it = 0; it < m1; ++it) {
            int i1 = rng_next(rng) % k;
            int i2 = rng_next(rng) % n;
            if (i2 < i1) {
                int sw = i1;
                i1 = i2;
                i2 = sw;
            }
            double delta = 0.0;
            for (int i = 0; i < k; ++i)
                delta += H[i + k * i1] * (XV(perm[i], perm[i2]) - XV(perm[i], perm[i1]));
            if (i2 < k)
                for (int i = 0; i < k; ++i)
                    delta += H[i + k * i2] * (XV(perm[i], perm[i1]) - XV(perm[i], perm[i2]));
            delta *= 2.0;
            double d1 = XV(perm[i1], perm[i1]) + XV(perm[i2], perm[i2]) -
                        2.0 * XV(perm[i2], perm[i1]);
            double d2 = 1.0;
            if (i2 < k) d2 += -2 * H[i1 + k * i2] + 1.0;
            delta += d1 * d2;
            double cand = sol + delta;
            int accept;
            if (delta > 0) {
                double r = delta / T;
                if (r > 5)
                    accept = 0;
                else
                    accept = ((double)rng_next(rng) / ((double)RAND_MAX)) < exp(-r);
            } else {
                accept = 1;
            }
            if (accept) {
                if (fabs(delta) > 0.0001) changed = 1;
                int sw = perm[i1];
                perm[i1] = perm[i2];
                perm[i2] = sw;
                sol = cand;
                if (sol < cost) {
                    cost = sol;
                    memcpy(best, perm, sizeof(int) * n);
                }
            }
        }
        T *= 0.6;
        m1 = (int)round(m1 * 1.1);
    }
    for (int i = 0; i < k; ++i) perm_out[i] = best[i];
    free(perm);
    free(best);
    return cost;
#undef XV
}

/* Trials x types SA runs; keep at most K with cost < 1.  Once full, a
 * candidate below the tracked maximum overwrites the tracked slot; the rescan
 * starts from +BIG with '>', so the tracked slot never moves
 * (cutting_planes.c:247-303, 444-501).  Returns min cost. */
#define KGON_SEPARATE(NAME, TYPE, KK, NTYPES, HFN)                             \
    double NAME(const double *X, int n, int trials, int K, TYPE *list,         \
                int *size, mo_rng *rng) {                                      \
        int cnt = 0, slot = 0;                                                 \
        double slotval = -MO_BIG, best = MO_BIG;                               \
        int tup[7];                                                            \
        for (int tr = 0; tr < trials; ++tr)                                    \
            for (int type = 1; type <= NTYPES; ++type) {                       \
                double val = mo_qap_sa(HFN(type), KK, X, n, tup, rng);         \
                if (val < best) best = val;                                    \
                if (!(val < 1.0)) continue;                                    \
                if (cnt < K) {                                                 \
                    list[cnt].type = type;                                     \
                    list[cnt].value = val;                                     \
                    for (int a = 0; a < KK; ++a) list[cnt].perm[a] = tup[a];   \
                    if (cnt == 0 || val > slotval) {                           \
                        slotval = val;                                         \
                        slot = cnt;                                            \
                    }                                                          \
                    ++cnt;                                                     \
                } else if (val < slotval) {                                    \
                    list[slot].type = type;                                    \
                    list[slot].value = val;                                    \
                    for (int a = 0; a < KK; ++a) list[slot].perm[a] = tup[a];  \
                    slotval = MO_BIG;                                          \
                    for (int q = 0; q < cnt; ++q)                              \
                        if (list[q].value > slotval) {                         \
                            slotval = list[q].value;                           \
                            slot = q;                                          \
                        }                                                      \
                }                                                              \
            }                                                                  \
        *size = cnt;                                                           \
        return best;                                                           \
    }
KGON_SEPARATE(mo_pent_separate, mo_pent, 5, 3, mo_pent_H)
KGON_SEPARATE(mo_hept_separate, mo_hept, 7, 4, mo_hept_H)

#define KGON_UPDATE(NAME, TYPE, KK, CUTS, LIST, COUNT, SEPFN, TRIALS, CAP)     \
    double NAME(mo_state *st, const double *dual, int offset, int *added,      \
                int *removed) {                                                \
        int keep = 0, rem = 0;                                                 \
        for (int c = 0; c < st->COUNT; ++c) {                                  \
            st->CUTS[c].y = dual[offset + c];                                  \
            if (st->CUTS[c].y < 1e-5) {                                        \
                ++rem;                                                         \
            } else {                                                           \
                TYPE tmp = st->CUTS[c];                                        \
                st->CUTS[keep].type = tmp.type;                                \
                st->CUTS[keep].y = tmp.y;                                      \
                for (int a = 0; a < KK; ++a) st->CUTS[keep].perm[a] = tmp.perm[a]; \
                ++keep;                                                        \
            }                                                                  \
        }                                                                      \
        st->COUNT -= rem;                                                      \
        int lsize = 0;                                                         \
        double best = SEPFN(st->X, st->n, st->params.TRIALS, st->params.CAP,   \
                            st->LIST, &lsize, &st->rng);                       \
        int add = 0;                                                           \
        const int survivors = st->COUNT;                                       \
        for (int q = 0; q < lsize; ++q) {                                      \
            if (keep == MO_MAX_TRI) break;                                     \
            const TYPE *cand = &st->LIST[q];                                   \
            int dup = 0;                                                       \
            for (int c = 0; c < survivors; ++c) {                              \
                int same = st->CUTS[c].type == cand->type;                     \
                for (int a = 0; a < KK && same; ++a)                           \
                    same = st->CUTS[c].perm[a] == cand->perm[a];               \
                if (same) dup = 1;                                             \
            }                                                                  \
            if (dup) continue;                                                 \
            st->CUTS[keep] = *cand;                                            \
            st->CUTS[keep].y = 0.0;                                            \
            ++keep;                                                            \
            ++add;                                                             \
        }                                                                      \
        st->COUNT += add;                                                      \
        *added = add;                                                          \
        *removed = rem;                                                        \
        return best;                                                           \
    }
KGON_UPDATE(mo_pent_update, mo_pent, 5, pcuts, pent_list, npent, mo_pent_separate,
            Pent_Trials, PentIneq)
KGON_UPDATE(mo_hept_update, mo_hept, 7, hcuts, hept_list, nhept, mo_hept_separate,
            Hepta_Trials, HeptaIneq)

/* ------------------------------------------------------------------ */
/* heuristic (heuristic.c:10-223, bab_functions.c:96-134)              */
/* ------------------------------------------------------------------ */

/* 1-opt local search in the +-1 model (heuristic.c:100-201).  Returns the
 * cost x^T L x after the flips. */
double mo_one_opt(int *x, int n, const double *L) {
    double *Lx = xcalloc(n, sizeof(double));
    double *d = xcalloc(n, sizeof(double));
    double *gain = xcalloc(n, sizeof(double));
    int *nz = xcalloc(n, sizeof(int));
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) Lx[i] += L[j + (size_t)i * n] * x[j];
    double cost = 0.0;
    for (int i = 0; i < n; ++i) {
        d[i] = L[i + (size_t)i * n];
        cost += x[i] * Lx[i];
        gain[i] = d[i] - x[i] * Lx[i];
    }
    double best = -MO_BIG;
    int at = 0;
    for (int i = 0; i < n; ++i)
        if (gain[i] > best) {
            best = gain[i];
            at = i;
        }
    while (best > 0.001) {
        int cnt = 0;
        for (int j = 0; j < n; ++j)
            if (fabs(L[at + (size_t)n * j]) > 0.001) nz[cnt++] = j;
        double sgn = x[at] > 0 ? -2.0 : 2.0;
        for (int q = 0; q < cnt; ++q) Lx[nz[q]] += sgn * L[at + (size_t)nz[q] * n];
        x[at] = -x[at];
        cost += 4 * best;
        for (int i = 0; i < n; ++i) gain[i] = d[i] - x[i] * Lx[i];
        best = -MO_BIG;
        at = 0;
        for (int i = 0; i < n; ++i)
            if (gain[i] > best) {
                best = gain[i];
                at = i;
            }
    }
    free(Lx);
    free(d);
    free(gain);
    free(nz);
    return cost;
}

double mo_evaluate_solution(const mo_state *st, const int *sol) {
    int bs = st->N - 1;
    double v = 0.0;
    for (int i = 0; i < bs; ++i)
        for (int j = 0; j < bs; ++j) v += st->SPL[j + (size_t)i * st->N] * sol[i] * sol[j];
    return v;
}

int mo_update_solution(mo_state *st, const int *x) {
    double v = mo_evaluate_solution(st, x);
    if (v > st->LB) {
        st->LB = v;
        memcpy(st->best_x, x, sizeof(int) * (st->N - 1));
        return 1;
    }
    return 0;
}

/* Goemans-Williamson rounding with `num` hyperplanes drawn from the random
 * source, 1-opt each, map to 0/1 ignoring the homogenising coordinate, keep
 * the best by the original objective (heuristic.c:24-91, 209-223).  Reads the
 * state's Z and M as they are (see quirk: after the IPM they are stale). */
double mo_gw(mo_state *st, const mo_node *node, int *x, int num) {
    const int n = st->n, bs = st->N - 1, M = st->M;
    int *tx = xcalloc(n, sizeof(int));
    int *sol = xcalloc(bs > 0 ? bs : 1, sizeof(int));
    double *vh = xcalloc(M > 0 ? M : 1, sizeof(double));
    double best = -MO_BIG;
    for (int cnt = 0; cnt < num; ++cnt) {
        for (int j = 0; j < M; ++j)
            vh[j] = ((double)rng_next(&st->rng) / (double)(RAND_MAX)) - 0.5;
        for (int i = 0; i < n; ++i) {
            double sca = 0.0;
            for (int j = 0; j < M; ++j) sca += vh[j] * st->Z[(size_t)j * n + i];
            tx[i] = sca < 0 ? -1 : 1;
        }
        mo_one_opt(tx, n, st->PPL);
        int idx = 0;
        for (int i = 0; i < bs; ++i) {
            if (node->xfixed[i])
                sol[i] = node->solX[i];
            else
                sol[i] = (tx[idx++] + 1) / 2;
        }
        double val = mo_evaluate_solution(st, sol);
        if (best < val) {
            memcpy(x, sol, sizeof(int) * bs);
            best = val;
        }
    }
    free(tx);
    free(sol);
    free(vh);
    return best;
}

/* Explicit-input GW on a subproblem: hyperplanes given as num x M values
 * (already centred), returns for each hyperplane the 1-opt cut (n entries of
 * +-1) and its x^T L x.  Used to compare the GPU kernel draw for draw. */
double mo_gw_explicit(int n, const double *L, const double *Z, int M,
                      const double *planes, int num, int *cut_out,
                      double *val_out) {
    double best = -MO_BIG;
    for (int cnt = 0; cnt < num; ++cnt) {
        int *tx = cut_out + (size_t)cnt * n;
        for (int i = 0; i < n; ++i) {
            double sca = 0.0;
            for (int j = 0; j < M; ++j) sca += planes[(size_t)cnt * M + j] * Z[(size_t)j * n + i];
            tx[i] = sca < 0 ? -1 : 1;
        }
        val_out[cnt] = mo_one_opt(tx, n, L);
        if (val_out[cnt] > best) best = val_out[cnt];
    }
    return best;
}

/* ------------------------------------------------------------------ */
/* basic SDP interior-point method (ipm_mc_pk.c:28-347)                */
/* ------------------------------------------------------------------ */

static void diag_scale_cols(double *out, const double *A, const double *dv,
                            int n, double sgn) {
    /* out[:, j] = sgn * dv[:] .* A[:, j], i.e. Diag(dv) * A scaled by sgn */
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i)
            out[i + (size_t)j * n] = sgn * dv[i] * A[i + (size_t)j * n];
}

static void sym_avg(double *A, int n) {
    for (int j = 0; j < n; ++j)
        for (int i = 0; i <= j; ++i) {
            double a = 0.5 * (A[i + (size_t)j * n] + A[j + (size_t)i * n]);
            A[i + (size_t)j * n] = A[j + (size_t)i * n] = a;
        }
}

/* Predictor-corrector IPM for max <L,X> s.t. diag(X)=e, X psd.  Outputs X,
 * y, Z = Diag(y) - L and phi = e^T y; stops when phi - psi <= 1e-4.
 * Returns 0, or a nonzero code where the reference would exit(). */
int mo_ipm(const double *L, int n, double *X, double *y, double *Z,
           double *phi, int *iters) {
    const char UP = 'U', SL = 'L';
    const size_t nn = (size_t)n * n;
    int inc = 1, info = 0;
    double *dX = xcalloc(nn, sizeof(double)), *dZ = xcalloc(nn, sizeof(double));
    double *dX1 = xcalloc(nn, sizeof(double)), *Zi = xcalloc(nn, sizeof(double));
    double *Mm = xcalloc(nn, sizeof(double)), *tmp = xcalloc(nn, sizeof(double));
    double *tmp2 = xcalloc(nn, sizeof(double));
    double *dy = xcalloc(n, sizeof(double)), *dy1 = xcalloc(n, sizeof(double));
    int rc = 0;

    for (int i = 0; i < n; ++i) {
        double acc = 0.0;
        for (int j = 0; j < n; ++j) acc += fabs(L[j + (size_t)i * n]);
        y[i] = acc + 1.1;
    }
    memset(X, 0, nn * sizeof(double));
    for (int i = 0; i < n; ++i) X[i + (size_t)i * n] = 1.0;
    for (size_t e = 0; e < nn; ++e) Z[e] = -L[e];
    for (int i = 0; i < n; ++i) Z[i + (size_t)i * n] += y[i];
    double ph = 0.0, psi = 0.0, mu = 0.0;
    for (int i = 0; i < n; ++i) ph += y[i];
    for (size_t e = 0; e < nn; ++e) psi += L[e] * X[e];
    for (size_t e = 0; e < nn; ++e) mu += Z[e] * X[e];
    mu /= 2.0 * n;

    int it = 0;
    while (ph - psi > 1e-4) {
        ++it;
        /* Zi = Z^{-1} */
        memcpy(Zi, Z, nn * sizeof(double));
        dpotrf_(&UP, &n, Zi, &n, &info);
        if (info) { rc = 1; break; }
        dpotri_(&UP, &n, Zi, &n, &info);
        if (info) { rc = 2; break; }
        for (int j = 0; j < n; ++j)
            for (int i = j + 1; i < n; ++i) Zi[i + (size_t)j * n] = Zi[j + (size_t)i * n];
        /* predictor: dy1 = (Zi .* X) \ (-e) */
        for (size_t e = 0; e < nn; ++e) Mm[e] = Zi[e] * X[e];
        for (int i = 0; i < n; ++i) dy1[i] = -1.0;
        dposv_(&UP, &n, &inc, Mm, &n, dy1, &n, &info);
        if (info) { rc = 3; break; }
        /* dX1 = -Zi Diag(dy1) X - X, symmetrised */
        diag_scale_cols(tmp, X, dy1, n, -1.0);
        double one = 1.0, zero = 0.0, mone = -1.0;
        dsymm_(&SL, &UP, &n, &n, &one, Zi, &n, tmp, &n, &zero, dX1, &n);
        for (size_t e = 0; e < nn; ++e) dX1[e] -= X[e];
        sym_avg(dX1, n);
        /* corrector: dy2 = (Zi .* X) \ (mu diag(Zi) - (Zi .* dX1) dy1) */
        for (int i = 0; i < n; ++i) dy[i] = Zi[i + (size_t)i * n];
        for (size_t e = 0; e < nn; ++e) tmp[e] = Zi[e] * dX1[e];
        dsymv_(&UP, &n, &mone, tmp, &n, dy1, &inc, &mu, dy, &inc);
        for (size_t e = 0; e < nn; ++e) Mm[e] = Zi[e] * X[e];
        dposv_(&UP, &n, &inc, Mm, &n, dy, &n, &info);
        if (info) { rc = 4; break; }
        /* dX2 = mu Zi - Zi (Diag(dy2) X + Diag(dy1) dX1) */
        diag_scale_cols(tmp, X, dy, n, -1.0);
        diag_scale_cols(tmp2, dX1, dy1, n, -1.0);
        for (size_t e = 0; e < nn; ++e) tmp[e] += tmp2[e];
        dsymm_(&SL, &UP, &n, &n, &one, Zi, &n, tmp, &n, &zero, dX, &n);
        for (size_t e = 0; e < nn; ++e) dX[e] += mu * Zi[e];
        for (int i = 0; i < n; ++i) dy[i] += dy1[i];
        for (size_t e = 0; e < nn; ++e) dX[e] += dX1[e];
        sym_avg(dX, n);
        memset(dZ, 0, nn * sizeof(double));
        for (int i = 0; i < n; ++i) dZ[i + (size_t)i * n] = dy[i];
        /* step lengths by Cholesky back-off (ipm_mc_pk.c:273-301) */
        double ap = 1.0;
        for (;;) {
            for (size_t e = 0; e < nn; ++e) tmp[e] = X[e] + ap * dX[e];
            dpotrf_(&UP, &n, tmp, &n, &info);
            if (!info) break;
            ap *= 0.8;
        }
        if (ap < 1.0) ap *= 0.95;
        double ad = 1.0;
        for (;;) {
            for (size_t e = 0; e < nn; ++e) tmp[e] = Z[e] + ad * dZ[e];
            dpotrf_(&UP, &n, tmp, &n, &info);
            if (!info) break;
            ad *= 0.8;
        }
        if (ad < 1.0) ad *= 0.95;
        for (size_t e = 0; e < nn; ++e) X[e] += ap * dX[e];
        for (int i = 0; i < n; ++i) y[i] += ad * dy[i];
        for (size_t e = 0; e < nn; ++e) Z[e] += ad * dZ[e];
        mu = 0.0;
        for (size_t e = 0; e < nn; ++e) mu += Z[e] * X[e];
        mu /= 2.0 * n;
        if (ap + ad > 1.6) mu *= 0.5;
        if (ap + ad > 1.9) mu *= 0.2;
        ph = 0.0;
        for (int i = 0; i < n; ++i) ph += y[i];
        psi = 0.0;
        for (size_t e = 0; e < nn; ++e) psi += L[e] * X[e];
    }
    *phi = ph;
    if (iters) *iters = it;
    free(dX); free(dZ); free(dX1); free(Zi); free(Mm); free(tmp); free(tmp2);
    free(dy); free(dy1);
    return rc;
}

/* ------------------------------------------------------------------ */
/* bounding routine and B&B (bounding.c, evaluate.c, bab_functions.c)  */
/* ------------------------------------------------------------------ */

/* Values of the reference's params file (params:1-23); TriIneq is later set
 * to 10 N as adjust_TriIneq=1 does (process_input.c:122-123). */
void mo_default_params(mo_params *p) {
    p->sigma0 = 1.0;
    p->scaleSigma = 1.001;
    p->ADMM_itermax = 2000;
    p->tol0 = 5e-2;
    p->scaleTol = 0.6;
    p->minTol = 1e-3;
    p->triag_iter = 3;
    p->pent_iter = 3;
    p->hept_iter = 3;
    p->max_outer_iter = 12;
    p->extra_iter = 5;
    p->violated_TriIneq = 1e-3;
    p->TriIneq = 5000;
    p->adjust_TriIneq = 1;
    p->PentIneq = 5000;
    p->HeptaIneq = 5000;
    p->Pent_Trials = 60;
    p->Hepta_Trials = 50;
    p->include_Pent = 1;
    p->include_Hepta = 1;
    p->root = 0;
    p->use_diff = 1;
    p->branchingStrategy = 1;
}

mo_state *mo_state_new(int N, const double *SPL, const mo_params *p) {
    mo_state *st = xcalloc(1, sizeof(mo_state));
    st->params = *p;
    if (st->params.adjust_TriIneq) st->params.TriIneq = N * 10;
    st->N = N;
    size_t NN = (size_t)N * N;
    st->SPL = xcalloc(NN, sizeof(double));
    memcpy(st->SPL, SPL, NN * sizeof(double));
    st->PPL = xcalloc(NN, sizeof(double));
    st->X = xcalloc(NN, sizeof(double));
    st->Q = xcalloc(NN, sizeof(double));
    st->Z = xcalloc(NN, sizeof(double));
    st->y = xcalloc(N, sizeof(double));
    st->W = xcalloc(N, sizeof(double));
    st->u = xcalloc(3 * MO_MAX_TRI, sizeof(double));
    st->t = xcalloc(3 * MO_MAX_TRI, sizeof(double));
    st->s = xcalloc(3 * MO_MAX_TRI, sizeof(double));
    st->cuts = xcalloc(MO_MAX_TRI, sizeof(mo_tri));
    st->pcuts = xcalloc(MO_MAX_TRI, sizeof(mo_pent));
    st->hcuts = xcalloc(MO_MAX_TRI, sizeof(mo_hept));
    st->tri_list = xcalloc(st->params.TriIneq, sizeof(mo_tri));
    st->pent_list = xcalloc(st->params.PentIneq, sizeof(mo_pent));
    st->hept_list = xcalloc(st->params.HeptaIneq, sizeof(mo_hept));
    st->n = N;
    st->LB = 0.0;
    st->lb_override = -1.0;
    st->rng.stream = NULL;
    return st;
}

void mo_state_free(mo_state *st) {
    if (!st) return;
    free(st->SPL); free(st->PPL); free(st->X); free(st->Q); free(st->Z);
    free(st->y); free(st->W); free(st->u); free(st->t); free(st->s);
    free(st->cuts); free(st->pcuts); free(st->hcuts);
    free(st->tri_list); free(st->pent_list); free(st->hept_list);
    free(st);
}

static void store_fracsol(const mo_state *st, mo_node *node) {
    int idx = 0, n = st->n;
    for (int i = 0; i < st->N - 1; ++i) {
        if (node->xfixed[i])
            node->fracsol[i] = (double)node->solX[i];
        else {
            node->fracsol[i] = 0.5 * (st->X[(n - 1) + (size_t)idx * n] + 1.0);
            ++idx;
        }
    }
}

static void heuristic_round(mo_state *st, mo_node *node) {
    int bs = st->N - 1;
    int *x = xcalloc(bs > 0 ? bs : 1, sizeof(int));
    for (int i = 0; i < bs; ++i) x[i] = node->xfixed[i] ? node->solX[i] : 0;
    mo_gw(st, node, x, st->N);
    mo_update_solution(st, x);
    free(x);
}

/* SDPbound (bounding.c:25-250). */
double mo_sdpbound(mo_state *st, mo_node *node, int rank) {
    mo_params *P = &st->params;
    double sigma = P->sigma0, tol = P->tol0;
    double fixedvalue = mo_fixed_value(st->N, st->N - 1, st->SPL, node->xfixed, node->solX);
    double viol3 = 0.0, viol5 = 0.0, oldf, bound, gap;
    int count = 0, nbit = 0, done = 0, giveup = 0, prune = 0;
    int ta = 0, tr = 0, pa = 0, pr = 0, ha = 0, hr = 0;
    st->ntri = st->npent = st->nhept = 0;

    if (mo_ipm(st->PPL, st->n, st->X, st->y, st->Q, &st->f, NULL) != 0) {
        fprintf(stderr, "oracle: IPM failed\n");
        exit(EXIT_FAILURE);
    }
    double basic_bound = st->f + fixedvalue;
    store_fracsol(st, node);
    heuristic_round(st, node);
    bound = st->f + fixedvalue;
    if (bound < st->LB + 1.0) return bound;
    if (P->use_diff && rank != 0 && bound > st->LB + st->diff + 1.0) return bound;
    viol3 = mo_tri_update(st, st->s, &ta, &tr);

    while (!done) {
        ++count;
        oldf = st->f;
        mo_admm(st, &sigma, tol, &nbit);
        tol *= P->scaleTol;
        if (tol < P->minTol) tol = P->minTol;
        bound = st->f + fixedvalue;
        prune = bound < st->LB + 1.0;
        if (!prune) {
            heuristic_round(st, node);
            prune = bound < st->LB + 1.0;
        }
        gap = bound - st->LB;
        if (count == P->triag_iter + P->pent_iter + P->hept_iter &&
            gap - 1.0 > (oldf - st->f) * (P->max_outer_iter - count))
            giveup = 1;
        if (count == P->max_outer_iter && gap - 1.0 > (oldf - st->f) * P->extra_iter)
            giveup = 1;
        if (count == P->max_outer_iter + P->extra_iter) giveup = 1;
        if (rank != 0 && giveup && !prune) {
            P->Pent_Trials += 60;
            P->Hepta_Trials += 50;
        }
        if (!prune && !giveup) {
            int triag = st->ntri, penta = st->npent;
            viol3 = mo_tri_update(st, st->s, &ta, &tr);
            if (P->include_Pent && (count >= P->triag_iter || viol3 < 0.2))
                viol5 = mo_pent_update(st, st->s, triag, &pa, &pr);
            if (P->include_Hepta &&
                ((count >= P->triag_iter + P->pent_iter) || (viol3 < 0.2 && (1 - viol5 < 0.4))))
                mo_hept_update(st, st->s, triag + penta, &ha, &hr);
        }
        done = prune || giveup || (nbit >= P->ADMM_itermax);
        store_fracsol(st, node);
    }
    bound = st->f + fixedvalue;
    if (rank == 0) st->diff = basic_bound - bound;
    return bound;
}

double mo_evaluate(mo_state *st, mo_node *node, int rank) {
    st->n = mo_create_subproblem(st->N, st->N - 1, st->SPL, node->xfixed, node->solX, st->PPL);
    st->nodes += 1;
    return mo_sdpbound(st, node, rank);
}

/* getBranchingVariable (bab_functions.c:389-423). */
int mo_branch_var(const mo_state *st, const mo_node *node) {
    int ic = -1;
    if (st->params.branchingStrategy == 0) {
        double mx = -MO_BIG;
        for (int i = 0; i < st->N - 1; ++i)
            if (!node->xfixed[i] && fabs(0.5 - node->fracsol[i]) > mx) {
                ic = i;
                mx = fabs(0.5 - node->fracsol[i]);
            }
    } else {
        double mn = MO_BIG;
        for (int i = 0; i < st->N - 1; ++i)
            if (!node->xfixed[i] && fabs(0.5 - node->fracsol[i]) < mn) {
                ic = i;
                mn = fabs(0.5 - node->fracsol[i]);
            }
    }
    return ic;
}

/* ---- max-heap on upper_bound (heap.c:27-133) ---- */
typedef struct {
    mo_node **d;
    int used, cap;
} mo_heap;

static int pri_gt(const mo_node *a, const mo_node *b) {
    return a->upper_bound > b->upper_bound;
}
static void heap_push(mo_heap *h, mo_node *x) {
    if (h->used == h->cap) {
        h->cap = h->cap ? 2 * h->cap : 64;
        h->d = realloc(h->d, sizeof(mo_node *) * h->cap);
    }
    int c = h->used++;
    h->d[c] = x;
    while (c > 0) {
        int p = (c - 1) / 2;
        if (!pri_gt(h->d[c], h->d[p])) break;
        mo_node *tmp = h->d[c];
        h->d[c] = h->d[p];
        h->d[p] = tmp;
        c = p;
    }
}
static mo_node *heap_pop(mo_heap *h) {
    mo_node *top = h->d[0];
    h->used--;
    h->d[0] = h->d[h->used];
    int c = 0;
    for (;;) {
        int l = 2 * c + 1, best = c;
        if (l < h->used && pri_gt(h->d[l], h->d[best])) best = l;
        if (l + 1 < h->used && pri_gt(h->d[l + 1], h->d[best])) best = l + 1;
        if (best == c) break;
        mo_node *tmp = h->d[c];
        h->d[c] = h->d[best];
        h->d[best] = tmp;
        c = best;
    }
    return top;
}

static mo_node *child_of(const mo_state *st, const mo_node *parent, int ic, int val) {
    mo_node *c = xcalloc(1, sizeof(mo_node));
    for (int i = 0; i < st->N - 1; ++i) {
        c->xfixed[i] = parent->xfixed[i];
        c->solX[i] = parent->xfixed[i] ? parent->solX[i] : 0;
    }
    c->xfixed[ic] = 1;
    c->solX[ic] = val;
    c->level = parent->level + 1;
    /* heap.c:143-168 leaves upper_bound unset; children are ordered by the
     * parent's bound here (best-first) */
    c->upper_bound = parent->upper_bound;
    return c;
}

/* Single-process B&B: the root is evaluated as rank 0 (bab_functions.c:39-69),
 * every other node as a worker (bab_functions.c:226-334, without MPI).
 * Returns the optimum; st->nodes counts evaluations.  max_nodes <= 0 means no
 * limit. */
double mo_bab(mo_state *st, long max_nodes) {
    if (st->lb_override >= 0) st->LB = st->lb_override;
    mo_heap h = {0};
    mo_node *root = xcalloc(1, sizeof(mo_node));
    root->upper_bound = 0.0;
    st->root_bound = mo_evaluate(st, root, 0);
    root->upper_bound = st->root_bound;
    if (!(st->LB + 1.0 < root->upper_bound) || st->params.root) {
        free(root);
        return st->LB;
    }
    int ic = mo_branch_var(st, root);
    heap_push(&h, child_of(st, root, ic, 0));
    heap_push(&h, child_of(st, root, ic, 1));
    free(root);
    while (h.used > 0) {
        if (max_nodes > 0 && st->nodes >= max_nodes) break;
        mo_node *node = heap_pop(&h);
        node->upper_bound = mo_evaluate(st, node, 1);
        if (st->LB + 1.0 < node->upper_bound) {
            ic = mo_branch_var(st, node);
            heap_push(&h, child_of(st, node, ic, 0));
            heap_push(&h, child_of(st, node, ic, 1));
        }
        free(node);
    }
    while (h.used > 0) free(heap_pop(&h));
    free(h.d);
    return st->LB;
}
