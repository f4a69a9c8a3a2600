/*
 * dropin_ops_check.c -- test driver for madam_amd/dropin/madam_operators.c:
 * calls op_B, op_Bt, project_SDP, diag, Diag and op_Diag through the
 * reference's prototypes (admm.h:272-277) and globals, on a case read from
 * a file, and writes what the calls leave behind.  Built by
 * oracle/Makefile (target dropin_ops, with MADAM's admm.h and global_var.h)
 * into oracle/_ref/madam_ops_check; tests/test_gpu_dropin.py compares the
 * results with the oracle.
 *
 * Input (native-endian): int n, ntri, npent, nhept; double X[n*n];
 * ntri x {int i, j, k, type}; npent x {int type, perm[5]}; nhept x
 * {int type, perm[7]}; double t[m] (m = ntri + npent + nhept).
 * Output: double y[m] = B(X), then X after op_B (mirrored), then
 * X - Bt(t), then X, Q after project_SDP, int M, double W[M], Z[n*M],
 * then diag(X), Diag(diag(X)) and Diag(diag(X)) - Diag(diag(X)) by op_Diag.
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

int main(int argc, char **argv) {
    if (argc != 3) {
        fprintf(stderr, "usage: %s in out\n", argv[0]);
        return 2;
    }
    FILE *fi = fopen(argv[1], "rb");
    FILE *fo = fopen(argv[2], "wb");
    if (!fi || !fo) return 2;
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
