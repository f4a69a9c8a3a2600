/*
 * madam_sdpbound.c -- SDPbound (bounding.c:25, prototype admm.h:234) on the
 * GPU through libmadam_bound.
 *
 * Compiled inside MADAM's own tree in place of bounding.c and admm.c (which
 * then has no caller); it includes MADAM's admm.h and this repository's
 * include/madam_bound.h.  It reads and writes the globals the unchanged
 * program uses around SDPbound: params (the grown Pent_Trials and
 * Hepta_Trials, bounding.c:184-187), diff (set at the root, bounding.c:241),
 * the incumbent (Bab_LBGet, updateSolution) and the node's fracsol and
 * upper_bound.  Failures end the process with a message, as the reference
 * does on LAPACK errors.
 *
 * Environment: MADAM_PROJECTION=0 selects the plain FP64 eigen-projection
 * (exact cut solve); the default is the mixed-precision sign iteration.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "admm.h"
#include "madam_bound.h"

extern Parameters params;
extern double diff;

_Static_assert(sizeof(madam_params) == sizeof(Parameters), "Parameters layout");
_Static_assert(sizeof(madam_node) == sizeof(BabNode), "BabNode layout");
_Static_assert(sizeof(madam_tri) == sizeof(Triangle_Inequality), "triangle cut layout");
_Static_assert(sizeof(madam_pent) == sizeof(Pentagonal_Inequality), "pentagonal cut layout");
_Static_assert(sizeof(madam_hept) == sizeof(Heptagonal_Inequality), "heptagonal cut layout");

static void madam_fatal(const char *where) {
    fprintf(stderr, "%s: %s\n", where, madam_last_error());
    exit(1);
}

double SDPbound(BabNode *node, Problem *SP, Problem *PP, int rank) {
    static int opened = 0;
    static long long stream = 0;
    static int projection = 1;
    (void)PP;  /* the library builds the subproblem from SP and the node */
    if (!opened) {
        const char *e = getenv("MADAM_PROJECTION");
        if (e && e[0] == '0') projection = 0;
        if (madam_open(0) != MADAM_OK) madam_fatal("madam_open");
        opened = 1;
    }
    madam_bab_ctx c;
    memset(&c, 0, sizeof c);
    c.N = SP->n;
    c.SPL = SP->L;
    memcpy(&c.params, &params, sizeof params);
    c.LB = Bab_LBGet();
    c.diff = diff;
    c.seed = 2020;
    c.stream_ctr = stream;
    c.projection = projection;
    c.cg_rtol = projection ? 1e-7 : 1e-12;
    c.cg_adapt = projection ? 0.01 : 0.0;   /* inexact cut solves while ADMM is far off */
    c.gw_mult = projection ? 4 : 1;          /* 4N rounding hyperplanes (heuristic.c: N) */
    c.batch = 1;
    madam_node *nodes[1] = {(madam_node *)node};
    int rk = rank;
    double bound = 0.0;
    if (madam_sdpbound_batch(&c, 1, nodes, &rk, &bound) != MADAM_OK) madam_fatal("SDPbound");
    stream = c.stream_ctr;
    params.Pent_Trials = c.params.Pent_Trials;
    params.Hepta_Trials = c.params.Hepta_Trials;
    if (rank == 0) diff = c.diff;
    if (c.LB > Bab_LBGet()) updateSolution(c.best_x);
    return bound;
}
