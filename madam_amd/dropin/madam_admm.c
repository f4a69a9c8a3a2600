/*
 * madam_admm.c -- ADMM_solver (admm.c:49, prototype admm.h:212) and
 * ipm_mc_pk (ipm_mc_pk.c:28, prototype admm.h:269) on the GPU through
 * libmadam_bound, with the reference's prototypes and globals.
 *
 * Compiled inside MADAM's own tree in place of admm.c and ipm_mc_pk.c, so
 * that the unchanged bounding.c (its cutting-plane loop, heuristic and
 * separation on the CPU) calls the device ADMM and interior-point method.
 * It includes MADAM's admm.h and this repository's include/madam_bound.h.
 * What ADMM_solver leaves in globals for its callers is written as the
 * reference writes it: f (the safe bound, admm.c:757), s, t, u (cut
 * duals and slacks), Z and M (the factor the rounding heuristic reads,
 * admm.c:580-585).  Failures end the process with a message, as the
 * reference does on LAPACK errors.
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
extern double f;
extern int M;
extern double *Z;
extern double *u, *t, *s;
extern Triangle_Inequality *Cuts;
extern Pentagonal_Inequality *Pent_Cuts;
extern Heptagonal_Inequality *Hepta_Cuts;

_Static_assert(sizeof(madam_tri) == sizeof(Triangle_Inequality), "triangle cut layout");
_Static_assert(sizeof(madam_pent) == sizeof(Pentagonal_Inequality), "pentagonal cut layout");
_Static_assert(sizeof(madam_hept) == sizeof(Heptagonal_Inequality), "heptagonal cut layout");

static int projection = 1;

static void madam_fatal(const char *where) {
    fprintf(stderr, "%s: %s\n", where, madam_last_error());
    exit(1);
}

static void madam_once(void) {
    static int opened = 0;
    if (opened) return;
    const char *e = getenv("MADAM_PROJECTION");
    if (e && e[0] == '0') projection = 0;
    if (madam_open(0) != MADAM_OK) madam_fatal("madam_open");
    opened = 1;
}

void ADMM_solver(const Problem *PP, double *X, double *y, double *Q, double *sigma, double tol,
                 int *nbit, int print) {
    (void)print;
    madam_once();
    madam_admm_job job;
    memset(&job, 0, sizeof job);
    job.n = PP->n;
    job.L = PP->L;
    job.X = X;
    job.y = y;
    job.Q = Q;
    job.ntri = PP->NIneq;
    job.npent = PP->NPentIneq;
    job.nhept = PP->NHeptaIneq;
    job.tri = (const madam_tri *)Cuts;          /* duals start from Cuts[].y (admm.c:79-95) */
    job.pent = (const madam_pent *)Pent_Cuts;
    job.hept = (const madam_hept *)Hepta_Cuts;
    job.s = s;
    job.t = t;
    job.u = u;
    job.sigma = sigma;
    job.tol = tol;
    job.itermax = params.ADMM_itermax;
    job.scaleSigma = params.scaleSigma;
    job.Z = Z;
    madam_opts o;
    memset(&o, 0, sizeof o);
    o.cg_rtol = projection ? 1e-7 : 1e-12;
    o.cg_maxit = 1000;
    o.jacobi_sweeps = 40;
    o.check_every = 16;
    o.projection = projection;
    o.cg_adapt = projection ? 0.01 : 0.0;
    if (madam_admm_batch(1, &job, &o) != MADAM_OK) madam_fatal("ADMM_solver");
    *nbit = job.nbit;
    f = job.f;
    M = job.M;
}

void ipm_mc_pk(double *L, int n, double *X, double *y, double *Zd, double *phi, int print) {
    (void)print;
    madam_once();
    int iters = 0;
    if (madam_ipm(n, L, X, y, Zd, phi, &iters) != MADAM_OK) madam_fatal("ipm_mc_pk");
}
