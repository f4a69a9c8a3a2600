/*
 * dropin_ops_check.c -- test driver for the drop-in adapters
 * madam_amd/dropin/madam_operators.c, madam_cutting_planes.c and
 * madam_heuristic.c: calls their entry points through the reference's
 * prototypes (admm.h:237-277) and globals on a case read from a file and
 * writes what the calls leave behind.  Built by oracle/Makefile (target
 * dropin_ops, with MADAM's admm.h and global_var.h) into
 * oracle/_ref/madam_ops_check; tests/test_gpu_dropin.py compares the
 * results with the oracle and the library.
 *
 * "ops in out" -- operators.  Input (native-endian): int n, ntri, npent,
 * nhept; double X[n*n]; ntri x {int i, j, k, type}; npent x {int type,
 * perm[5]}; nhept x {int type, perm[7]}; double t[m] (m = ntri + npent +
 * nhept).  Output: double y[m] = B(X), then X after op_B (mirrored), then
 * X - Bt(t), then X, Q after project_SDP, int M, double W[M], Z[n*M],
 * then diag(X), Diag(diag(X)) and Diag(diag(X)) - Diag(diag(X)) by op_Diag.
 *
 * "sep in out" -- separation and rounding (srand(seed + 1), seed + 2 and
 * seed + 3 right before the pentagonal, heptagonal and rounding calls).  Input: int
 * n, ntri, npent, nhept, TriIneq, PentIneq, HeptaIneq, Pent_Trials,
 * Hepta_Trials, Mz, num, seed; double violated_TriIneq; double X[n*n]
 * (the lower triangle is what counts); the cuts as in "ops"; double
 * y[m] (their duals); double L[n*n]; double Zf[n*Mz]; int x0[n] (+-1).
 * Output: for updateTriangle/Pentagonal/HeptagonalInequalities in turn
 * (duals from y, offsets ntri and ntri + npent): double returned value, int
 * count, added, subtracted, then the cuts ({i, j, k, type, y} resp.
 * {type, perm, value, y}); then mc_1opt(x0): double cost, int x[n]; then
 * GW_heuristic(num hyperplanes) with no fixings (evaluateSolution = the
 * value of the 0/1 cut on L's leading n-1 block): double best, int x[n-1].
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "admm.h"
#include "global_var.h"

static void rd(void *p, size_t sz, size_t cnt, FILE *f) {
    if (fread(p, sz, cnt, f) != cnt) {
        fprintf(stderr, "short input\n");
        exit(2);
    }
}

static double *g_L;
static int g_n;

/* bab_functions.c:96-107 on the test's matrix: the cut's value */
double evaluateSolution(int *sol) {
    double val = 0.0;
    for (int i = 0; i < g_n - 1; ++i)
        for (int j = 0; j < g_n - 1; ++j) val += g_L[j + (size_t)i * g_n] * sol[i] * sol[j];
    return val;
}

static void read_cuts(FILE *fi, int ntri, int npent, int nhept) {
    Cuts = calloc(MaxTriIneqAdded, sizeof(Triangle_Inequality));
    Pent_Cuts = calloc(MaxPentIneqAdded, sizeof(Pentagonal_Inequality));
    Hepta_Cuts = calloc(MaxHeptaIneqAdded, sizeof(Heptagonal_Inequality));
    for (int c = 0; c < ntri; ++c) {
        int v[4];
        rd(v, sizeof(int), 4, fi);
        Cuts[c].i = v[0];
        Cuts[c].j = v[1];
        Cuts[c].k = v[2];
        Cuts[c].type = v[3];
    }
    for (int c = 0; c < npent; ++c) {
        int v[6];
        rd(v, sizeof(int), 6, fi);
        Pent_Cuts[c].type = v[0];
        memcpy(Pent_Cuts[c].permutation, v + 1, 5 * sizeof(int));
    }
    for (int c = 0; c < nhept; ++c) {
        int v[8];
        rd(v, sizeof(int), 8, fi);
        Hepta_Cuts[c].type = v[0];
        memcpy(Hepta_Cuts[c].permutation, v + 1, 7 * sizeof(int));
    }
}

static int sep_main(FILE *fi, FILE *fo) {
    int h[12];
    rd(h, sizeof(int), 12, fi);
    const int n = h[0], ntri = h[1], npent = h[2], nhept = h[3];
    params.TriIneq = h[4];
    params.PentIneq = h[5];
    params.HeptaIneq = h[6];
    params.Pent_Trials = h[7];
    params.Hepta_Trials = h[8];
    const int Mz = h[9], num = h[10];
    /* each adapter that draws a seed gets srand(seed + its place) right
     * before its call: no other rand() user (the GPU runtime's start-up
     * included) comes in between */
    const unsigned seed = (unsigned)h[11];
    rd(&params.violated_TriIneq, sizeof(double), 1, fi);
    const int m = ntri + npent + nhept;
    const size_t nn = (size_t)n * n;
    X = malloc(nn * sizeof(double));
    rd(X, sizeof(double), nn, fi);
    read_cuts(fi, ntri, npent, nhept);
    double *yy = calloc(m + 1, sizeof(double));
    rd(yy, sizeof(double), m, fi);
    g_L = malloc(nn * sizeof(double));
    g_n = n;
    rd(g_L, sizeof(double), nn, fi);
    Z = calloc((size_t)n * (Mz > 0 ? Mz : 1), sizeof(double));
    rd(Z, sizeof(double), (size_t)n * Mz, fi);
    M = Mz;
    int *x0 = malloc(n * sizeof(int));
    rd(x0, sizeof(int), n, fi);
    List = calloc(params.TriIneq + 1, sizeof(Triangle_Inequality));
    Pent_List = calloc(params.PentIneq + 1, sizeof(Pentagonal_Inequality));
    Hepta_List = calloc(params.HeptaIneq + 1, sizeof(Heptagonal_Inequality));

    Problem P = {g_L, n, ntri, npent, nhept};
    int added, subtracted;
    double v = updateTriangleInequalities(&P, yy, &added, &subtracted);
    int hd[3] = {P.NIneq, added, subtracted};
    fwrite(&v, sizeof(double), 1, fo);
    fwrite(hd, sizeof(int), 3, fo);
    for (int c = 0; c < P.NIneq; ++c) {
        int e[4] = {Cuts[c].i, Cuts[c].j, Cuts[c].k, Cuts[c].type};
        fwrite(e, sizeof(int), 4, fo);
        fwrite(&Cuts[c].y, sizeof(double), 1, fo);
    }
    srand(seed + 1);
    v = updatePentagonalInequalities(&P, yy, &added, &subtracted, ntri);
    hd[0] = P.NPentIneq;
    hd[1] = added;
    hd[2] = subtracted;
    fwrite(&v, sizeof(double), 1, fo);
    fwrite(hd, sizeof(int), 3, fo);
    for (int c = 0; c < P.NPentIneq; ++c) {
        fwrite(&Pent_Cuts[c].type, sizeof(int), 1, fo);
        fwrite(Pent_Cuts[c].permutation, sizeof(int), 5, fo);
        fwrite(&Pent_Cuts[c].value, sizeof(double), 1, fo);
        fwrite(&Pent_Cuts[c].y, sizeof(double), 1, fo);
    }
    srand(seed + 2);
    v = updateHeptagonalInequalities(&P, yy, &added, &subtracted, ntri + npent);
    hd[0] = P.NHeptaIneq;
    hd[1] = added;
    hd[2] = subtracted;
    fwrite(&v, sizeof(double), 1, fo);
    fwrite(hd, sizeof(int), 3, fo);
    for (int c = 0; c < P.NHeptaIneq; ++c) {
        fwrite(&Hepta_Cuts[c].type, sizeof(int), 1, fo);
        fwrite(Hepta_Cuts[c].permutation, sizeof(int), 7, fo);
        fwrite(&Hepta_Cuts[c].value, sizeof(double), 1, fo);
        fwrite(&Hepta_Cuts[c].y, sizeof(double), 1, fo);
    }
    double cost = mc_1opt(x0, &P);
    fwrite(&cost, sizeof(double), 1, fo);
    fwrite(x0, sizeof(int), n, fo);
    BabNode *node = calloc(1, sizeof(BabNode));
    int *xb = calloc(n, sizeof(int));
    srand(seed + 3);
    const double best = GW_heuristic(&P, &P, node, xb, num);
    fwrite(&best, sizeof(double), 1, fo);
    fwrite(xb, sizeof(int), n - 1, fo);
    return 0;
}

int main(int argc, char **argv) {
    if (argc != 4 || (strcmp(argv[1], "ops") && strcmp(argv[1], "sep"))) {
        fprintf(stderr, "usage: %s ops|sep in out\n", argv[0]);
        return 2;
    }
    FILE *fi = fopen(argv[2], "rb");
    FILE *fo = fopen(argv[3], "wb");
    if (!fi || !fo) return 2;
    if (!strcmp(argv[1], "sep")) {
        const int rc = sep_main(fi, fo);
        fclose(fi);
        fclose(fo);
        return rc;
    }
    int hdr[4];
    rd(hdr, sizeof(int), 4, fi);
    const int n = hdr[0], ntri = hdr[1], npent = hdr[2], nhept = hdr[3];
    const int m = ntri + npent + nhept;
    const size_t nn = (size_t)n * n;
    double *X0 = malloc(nn * sizeof(double));
    rd(X0, sizeof(double), nn, fi);
    Cuts = calloc(ntri + 1, sizeof(Triangle_Inequality));
    Pent_Cuts = calloc(npent + 1, sizeof(Pentagonal_Inequality));
    Hepta_Cuts = calloc(nhept + 1, sizeof(Heptagonal_Inequality));
    for (int c = 0; c < ntri; ++c) {
        int v[4];
        rd(v, sizeof(int), 4, fi);
        Cuts[c].i = v[0];
        Cuts[c].j = v[1];
        Cuts[c].k = v[2];
        Cuts[c].type = v[3];
    }
    for (int c = 0; c < npent; ++c) {
        int v[6];
        rd(v, sizeof(int), 6, fi);
        Pent_Cuts[c].type = v[0];
        memcpy(Pent_Cuts[c].permutation, v + 1, 5 * sizeof(int));
    }
    for (int c = 0; c < nhept; ++c) {
        int v[8];
        rd(v, sizeof(int), 8, fi);
        Hepta_Cuts[c].type = v[0];
        memcpy(Hepta_Cuts[c].permutation, v + 1, 7 * sizeof(int));
    }
    double *tt = calloc(m + 1, sizeof(double));
    rd(tt, sizeof(double), m, fi);
    fclose(fi);

    Problem P = {NULL, n, ntri, npent, nhept};
    double *y = calloc(m + 1, sizeof(double));
    double *Xw = malloc(nn * sizeof(double));
    memcpy(Xw, X0, nn * sizeof(double));
    op_B(&P, y, Xw);
    fwrite(y, sizeof(double), m, fo);
    fwrite(Xw, sizeof(double), nn, fo);

    memcpy(Xw, X0, nn * sizeof(double));
    op_Bt(&P, Xw, tt);
    fwrite(Xw, sizeof(double), nn, fo);

    X = malloc(nn * sizeof(double));
    Q = malloc(nn * sizeof(double));
    Z = malloc(nn * sizeof(double));
    W = malloc(n * sizeof(double));
    memcpy(X, X0, nn * sizeof(double));
    project_SDP(&P);
    fwrite(X, sizeof(double), nn, fo);
    fwrite(Q, sizeof(double), nn, fo);
    fwrite(&M, sizeof(int), 1, fo);
    fwrite(W, sizeof(double), M, fo);
    fwrite(Z, sizeof(double), (size_t)n * M, fo);

    double *d = malloc(n * sizeof(double));
    diag(X0, d, n);
    fwrite(d, sizeof(double), n, fo);
    Diag(Xw, d, n);
    fwrite(Xw, sizeof(double), nn, fo);
    op_Diag(&P, Xw, d);
    fwrite(Xw, sizeof(double), nn, fo);
    fclose(fo);
    return 0;
}
