/*
 * madam_evaluate.c -- createSubproblem, getFixedValue and Evaluate
 * (evaluate.c:12-139, prototypes admm.h:246-248) through libmadam_bound's
 * madam_create_subproblem, which builds the same PP->L (bit for bit) and
 * constant.  Compiled inside MADAM's own tree in place of evaluate.c;
 * Evaluate calls whichever SDPbound the program links (the reference's
 * bounding.c or madam_sdpbound.c).  Library failures end the process with
 * a message.
 */
#include <stdio.h>
#include <stdlib.h>

#include "admm.h"
#include "madam_bound.h"

static void ev_fatal(const char *where) {
    fprintf(stderr, "%s: %s\n", where, madam_last_error());
    exit(1);
}

static void ev_once(void) {
    static int opened = 0;
    if (opened) return;
    if (madam_open(0) != MADAM_OK) ev_fatal("madam_open");
    opened = 1;
}

void createSubproblem(BabNode *node, Problem *SP, Problem *PP) {
    ev_once();
    double fixedvalue = 0.0;
    int n = 0;
    if (madam_create_subproblem(SP->n, SP->L, node->xfixed, node->sol.X, PP->L, &n,
                                &fixedvalue) != MADAM_OK)
        ev_fatal("createSubproblem");
    PP->n = n;
}

double getFixedValue(BabNode *node, Problem *SP) {
    ev_once();
    double fixedvalue = 0.0;
    int n = 0;
    double *L = malloc((size_t)SP->n * SP->n * sizeof(double));
    if (!L) {
        fprintf(stderr, "getFixedValue: out of host memory\n");
        exit(1);
    }
    if (madam_create_subproblem(SP->n, SP->L, node->xfixed, node->sol.X, L, &n, &fixedvalue) !=
        MADAM_OK)
        ev_fatal("getFixedValue");
    free(L);
    return fixedvalue;
}

double Evaluate(BabNode *node, Problem *SP, Problem *PP, int rank) {
    createSubproblem(node, SP, PP);
    return SDPbound(node, SP, PP, rank);
}
