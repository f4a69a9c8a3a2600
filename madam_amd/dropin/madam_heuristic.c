/*
 * madam_heuristic.c -- the rounding heuristic's entry points of heuristic.c
 * with the reference's prototypes (admm.h:263-266) and the globals they
 * read (Z, M: the factor the last ADMM_solver call left):
 *
 *   runHeuristic   heuristic.c:10-19    GW_heuristic with P0->n hyperplanes
 *   GW_heuristic   heuristic.c:24-91    -> madam_gw_rounding: every hyperplane's
 *                                       cut and its 1-opt improvement on the GPU
 *                                       in one launch; the candidates are then
 *                                       offered to update_best in hyperplane order
 *   mc_1opt        heuristic.c:99-201   host: single-vertex flips while one gains
 *                                       more than 0.001, largest gain first
 *   update_best    heuristic.c:208-223  host, through evaluateSolution
 *                                       (bab_functions.c:96)
 *
 * Compiled inside MADAM's own tree in place of heuristic.c.  The
 * hyperplanes come from the library's random streams (one glibc rand() per
 * call seeds them) instead of num x M calls of rand().  Library failures end
 * the process with a message, as the reference does on LAPACK errors.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "admm.h"
#include "madam_bound.h"

extern int M;
extern double *Z;

static void hr_fatal(const char *where) {
    fprintf(stderr, "%s: %s\n", where, madam_last_error());
    exit(1);
}

static void hr_once(void) {
    static int opened = 0;
    if (opened) return;
    if (madam_open(0) != MADAM_OK) hr_fatal("madam_open");
    opened = 1;
}

static void *hr_alloc(size_t bytes) {
    void *p = calloc(1, bytes ? bytes : 1);
    if (!p) {
        fprintf(stderr, "madam_heuristic: out of host memory\n");
        exit(1);
    }
    return p;
}

double runHeuristic(Problem *P0, Problem *P, BabNode *node, int *x) {
    return GW_heuristic(P0, P, node, x, P0->n);
}

double GW_heuristic(Problem *P0, Problem *P, BabNode *node, int *x, int num) {
    hr_once();
    const int n = P->n, nb = P0->n - 1;
    double best = -BIG_NUMBER;
    if (num <= 0) return best;
    double *values = hr_alloc((size_t)num * sizeof(double));
    signed char *cuts = hr_alloc((size_t)num * n);
    int *sol = hr_alloc((size_t)(nb > 0 ? nb : 1) * sizeof(int));
    const unsigned long long seed = (unsigned long long)rand();
    if (madam_gw_rounding(n, P->L, Z, M, num, seed, 0, values, cuts) != MADAM_OK)
        hr_fatal("GW_heuristic");
    for (int h = 0; h < num; ++h) {
        const signed char *c = cuts + (size_t)h * n;
        int idx = 0;
        for (int i = 0; i < nb; ++i)
            sol[i] = node->xfixed[i] ? node->sol.X[i] : (c[idx++] + 1) / 2;
        update_best(x, sol, &best, P0);
    }
    free(values);
    free(cuts);
    free(sol);
    return best;
}

double mc_1opt(int *x, Problem *P) {
    const int n = P->n;
    const double *L = P->L;
    double *Lx = hr_alloc((size_t)n * sizeof(double));
    double *gain = hr_alloc((size_t)n * sizeof(double));
    /* Lx = L x and the cut's value x^T L x */
    double cost = 0.0;
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int j = 0; j < n; ++j) s += L[j + (size_t)i * n] * x[j];
        Lx[i] = s;
    }
    for (int i = 0; i < n; ++i) cost += x[i] * Lx[i];
    /* flipping x_i changes the value by 4 (L_ii - x_i (Lx)_i) */
    for (;;) {
        int arg = 0;
        double top = -BIG_NUMBER;
        for (int i = 0; i < n; ++i) {
            gain[i] = L[i + (size_t)i * n] - x[i] * Lx[i];
            if (gain[i] > top) {
                top = gain[i];
                arg = i;
            }
        }
        if (!(top > 0.001)) break;
        /* Lx -= 2 x_arg L(:, arg) over the column's entries above 0.001 */
        const double sgn = x[arg] > 0 ? -2.0 : 2.0;
        for (int j = 0; j < n; ++j) {
            const double l = L[arg + (size_t)j * n];
            if (fabs(l) > 0.001) Lx[j] += sgn * l;
        }
        x[arg] = -x[arg];
        cost += 4 * top;
    }
    free(Lx);
    free(gain);
    return cost;
}

int update_best(int *xbest, int *xnew, double *best, Problem *P0) {
    const double v = evaluateSolution(xnew);
    if (!(*best < v)) return 0;
    memcpy(xbest, xnew, sizeof(int) * (size_t)(P0->n - 1));
    *best = v;
    return 1;
}
