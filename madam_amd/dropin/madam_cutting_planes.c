/*
 * madam_cutting_planes.c -- the separation and cut-list entry points of
 * cutting_planes.c with the reference's prototypes (admm.h:237-243) and the
 * globals they read and write (Cuts, List, Pent_Cuts, Pent_List,
 * Hepta_Cuts, Hepta_List, X, params):
 *
 *   evaluateTriangleInequality          cutting_planes.c:20-40 (host)
 *   getViolated_TriangleInequalities    cutting_planes.c:48-127  -> madam_tri_separate
 *   updateTriangleInequalities          cutting_planes.c:132-214 (purge, separate,
 *                                       drop duplicates, append; host)
 *   getViolated_PentagonalInequalities  cutting_planes.c:224-309 -> madam_qap_sa, all
 *                                       Pent_Trials x 3 runs in one launch
 *   updatePentagonalInequalities        cutting_planes.c:316-409
 *   getViolated_HeptagonalInequalities  cutting_planes.c:419-506 -> madam_qap_sa,
 *                                       Hepta_Trials x 4 runs in one launch
 *   updateHeptagonalInequalities        cutting_planes.c:512-605
 *
 * Compiled inside MADAM's own tree in place of cutting_planes.c.  The
 * triangle list holds the same rows as the reference's (ties at the K-th
 * value aside) in enumeration order rather than the reference's
 * replacement-slot order.  The annealing runs draw from the library's
 * random streams (one glibc rand() per call seeds them), not from rand()
 * itself; the pentagonal and heptagonal lists then follow the reference's
 * top-K rules over the runs in the reference's order (trial by trial, type
 * by type), including the replacement slot that stops moving after the
 * first replacement.  Duplicate checks use a hash of the cut instead of the
 * reference's scan, with the same outcome.  Library failures end the
 * process with a message, as the reference does on LAPACK errors.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "admm.h"
#include "madam_bound.h"

extern Parameters params;
extern Triangle_Inequality *Cuts, *List;
extern Pentagonal_Inequality *Pent_Cuts, *Pent_List;
extern Heptagonal_Inequality *Hepta_Cuts, *Hepta_List;
extern double *X;

static void cp_fatal(const char *where) {
    fprintf(stderr, "%s: %s\n", where, madam_last_error());
    exit(1);
}

static void cp_once(void) {
    static int opened = 0;
    if (opened) return;
    if (madam_open(0) != MADAM_OK) cp_fatal("madam_open");
    opened = 1;
}

static void *cp_alloc(size_t bytes) {
    void *p = calloc(1, bytes ? bytes : 1);
    if (!p) {
        fprintf(stderr, "madam_cutting_planes: out of host memory\n");
        exit(1);
    }
    return p;
}

/* ---------------------------------------------------------------- hashing */

/* open-addressing set of 64-bit cut keys (key 0 never occurs: types >= 1) */
typedef struct {
    uint64_t *slot;
    size_t mask;
} KeySet;

static void ks_init(KeySet *s, size_t count) {
    size_t cap = 16;
    while (cap < 2 * count + 2) cap <<= 1;
    s->slot = cp_alloc(cap * sizeof(uint64_t));
    s->mask = cap - 1;
}

static size_t ks_pos(const KeySet *s, uint64_t key) {
    uint64_t h = key * 0x9E3779B97F4A7C15ull;
    size_t p = (size_t)(h >> 17) & s->mask;
    while (s->slot[p] && s->slot[p] != key) p = (p + 1) & s->mask;
    return p;
}

static int ks_has(const KeySet *s, uint64_t key) { return s->slot[ks_pos(s, key)] == key; }
static void ks_add(KeySet *s, uint64_t key) { s->slot[ks_pos(s, key)] = key; }
static void ks_free(KeySet *s) { free(s->slot); }

static uint64_t tri_key(int type, int i, int j, int k) {
    return ((uint64_t)type << 48) | ((uint64_t)i << 32) | ((uint64_t)j << 16) | (uint64_t)k;
}

/* FNV-1a over type and vertex tuple, never 0 */
static uint64_t kgon_key(int type, const int *perm, int k) {
    uint64_t h = 1469598103934665603ull;
    h = (h ^ (uint64_t)(uint32_t)type) * 1099511628211ull;
    for (int a = 0; a < k; ++a) h = (h ^ (uint64_t)(uint32_t)perm[a]) * 1099511628211ull;
    return h ? h : 1;
}

/* the exact comparison behind a hash hit */
static int same_tuple(const int *a, const int *b, int k) {
    for (int q = 0; q < k; ++q)
        if (a[q] != b[q]) return 0;
    return 1;
}

/* lower triangle of A (n x n, column-major) copied over its upper triangle */
static void mirror_lower(double *A, int n) {
    for (int j = 0; j < n; ++j)
        for (int i = j + 1; i < n; ++i) A[j + (size_t)i * n] = A[i + (size_t)j * n];
}

/* ---------------------------------------------------------------- triangles */

double evaluateTriangleInequality(double *XX, int N, int type, int ii, int jj, int kk) {
    const double a = XX[ii + jj * N], b = XX[ii + kk * N], c = XX[jj + kk * N];
    if (type == 1) return -a - b - c - 1.0;
    if (type == 2) return -a + b + c - 1.0;
    if (type == 3) return a - b + c - 1.0;
    return a + b - c - 1.0;
}

double getViolated_TriangleInequalities(double *Xm, int N, Triangle_Inequality *Lst, int *ListSize) {
    cp_once();
    /* the reference reads the lower triangle only */
    double *Xs = cp_alloc((size_t)N * N * sizeof(double));
    memcpy(Xs, Xm, (size_t)N * N * sizeof(double));
    mirror_lower(Xs, N);
    double maxval = 0.0;
    int size = 0;
    if (madam_tri_separate(N, Xs, params.TriIneq, params.violated_TriIneq, (madam_tri *)Lst,
                           &size, &maxval) != MADAM_OK)
        cp_fatal("getViolated_TriangleInequalities");
    free(Xs);
    *ListSize = size;
    return maxval;
}

double updateTriangleInequalities(Problem *PP, double *y, int *NumAdded, int *NumSubtracted) {
    /* purge: keep the cuts whose dual is at least 1e-5, in order */
    int kept = 0, removed = 0;
    for (int c = 0; c < PP->NIneq; ++c) {
        Cuts[c].y = y[c];
        if (Cuts[c].y < 1e-5) {
            ++removed;
            continue;
        }
        Cuts[kept].type = Cuts[c].type;
        Cuts[kept].i = Cuts[c].i;
        Cuts[kept].j = Cuts[c].j;
        Cuts[kept].k = Cuts[c].k;
        Cuts[kept].y = Cuts[c].y;
        ++kept;
    }
    PP->NIneq = kept;
    int size = 0;
    const double maxval = getViolated_TriangleInequalities(X, PP->n, List, &size);
    /* append the new rows not among the survivors, up to the cap */
    KeySet have;
    ks_init(&have, (size_t)kept + size);
    for (int c = 0; c < kept; ++c) ks_add(&have, tri_key(Cuts[c].type, Cuts[c].i, Cuts[c].j, Cuts[c].k));
    int added = 0;
    for (int l = 0; l < size && kept + added < MaxTriIneqAdded; ++l) {
        const Triangle_Inequality *t = &List[l];
        if (ks_has(&have, tri_key(t->type, t->i, t->j, t->k))) continue;
        Triangle_Inequality *d = &Cuts[kept + added];
        d->type = t->type;
        d->i = t->i;
        d->j = t->j;
        d->k = t->k;
        d->value = t->value;
        d->y = 0.0;
        ++added;
    }
    ks_free(&have);
    PP->NIneq += added;
    *NumAdded = added;
    *NumSubtracted = removed;
    return maxval;
}

/* ------------------------------------------------- pentagons and heptagons */

/* One k-gon separation round: trials x types annealing runs on X (one
 * library call), then the reference's top-K list rules over the runs in
 * order into list (stride bytes per entry, type/perm/value at the given
 * offsets).  Returns the smallest run cost. */
typedef struct {
    int *type;
    int *perm;
    double *value, *y;
} KgonRef;

static KgonRef kgon_at(void *list, size_t stride, int q, int k) {
    char *e = (char *)list + (size_t)q * stride;
    KgonRef r;
    if (k == 5) {
        Pentagonal_Inequality *p = (Pentagonal_Inequality *)e;
        r.type = &p->type;
        r.perm = p->permutation;
        r.value = &p->value;
        r.y = &p->y;
    } else {
        Heptagonal_Inequality *h = (Heptagonal_Inequality *)e;
        r.type = &h->type;
        r.perm = h->permutation;
        r.value = &h->value;
        r.y = &h->y;
    }
    return r;
}

static double kgon_separate(double *Xm, int N, int k, int trials, int ntypes, int K, void *list,
                            size_t stride, int *ListSize) {
    cp_once();
    /* the reference's annealing copies X's lower triangle over its upper
     * triangle (qap_simulated_annealing.c:14-17) */
    mirror_lower(Xm, N);
    const int nruns = trials * ntypes;
    *ListSize = 0;
    if (nruns <= 0) return BIG_NUMBER;
    int *types = cp_alloc((size_t)nruns * sizeof(int));
    long long *streams = cp_alloc((size_t)nruns * sizeof(long long));
    double *cost = cp_alloc((size_t)nruns * sizeof(double));
    int *perms = cp_alloc((size_t)nruns * k * sizeof(int));
    for (int r = 0; r < nruns; ++r) {
        types[r] = 1 + r % ntypes; /* trial by trial, types 1.. in order */
        streams[r] = r;
    }
    const unsigned long long seed = (unsigned long long)rand();
    if (madam_qap_sa(N, Xm, k, nruns, types, streams, seed, cost, perms, NULL) != MADAM_OK)
        cp_fatal(k == 5 ? "getViolated_PentagonalInequalities" : "getViolated_HeptagonalInequalities");
    double best = BIG_NUMBER;
    int size = 0, slot = 0;
    double slotval = -BIG_NUMBER;
    for (int r = 0; r < nruns; ++r) {
        const double v = cost[r];
        if (v < best) best = v;
        if (!(v < 1.0)) continue;
        int at;
        if (size < K) {
            at = size++;
            if (at == 0 || v > slotval) {
                slotval = v;
                slot = at;
            }
        } else if (v < slotval) {
            at = slot;
            /* the reference looks for the largest value above BIG_NUMBER:
             * none, so the slot stays where it is */
            slotval = BIG_NUMBER;
        } else {
            continue;
        }
        KgonRef e = kgon_at(list, stride, at, k);
        *e.type = types[r];
        *e.value = v;
        memcpy(e.perm, perms + (size_t)r * k, (size_t)k * sizeof(int));
    }
    free(types);
    free(streams);
    free(cost);
    free(perms);
    *ListSize = size;
    return best;
}

/* purge (duals from y[off ..]), separate, append new cuts up to cap */
static double kgon_update(Problem *PP, double *y, int *NumAdded, int *NumSubtracted, int off, int k,
                          int *count, void *cuts, void *list, size_t stride, int trials, int ntypes,
                          int K, int cap) {
    int kept = 0, removed = 0;
    for (int c = 0; c < *count; ++c) {
        KgonRef s = kgon_at(cuts, stride, c, k);
        *s.y = y[off + c];
        if (*s.y < 1e-5) {
            ++removed;
            continue;
        }
        KgonRef d = kgon_at(cuts, stride, kept, k);
        *d.type = *s.type;
        *d.y = *s.y;
        memmove(d.perm, s.perm, (size_t)k * sizeof(int));
        ++kept;
    }
    *count = kept;
    int size = 0;
    const double best = kgon_separate(X, PP->n, k, trials, ntypes, K, list, stride, &size);
    KeySet have;
    ks_init(&have, (size_t)kept + size);
    for (int c = 0; c < kept; ++c) {
        KgonRef s = kgon_at(cuts, stride, c, k);
        ks_add(&have, kgon_key(*s.type, s.perm, k));
    }
    int added = 0;
    for (int l = 0; l < size && kept + added < cap; ++l) {
        KgonRef t = kgon_at(list, stride, l, k);
        const uint64_t key = kgon_key(*t.type, t.perm, k);
        if (ks_has(&have, key)) {
            /* a hash hit: confirm against the survivors as the reference's
             * scan would */
            int dup = 0;
            for (int c = 0; c < kept && !dup; ++c) {
                KgonRef s = kgon_at(cuts, stride, c, k);
                dup = *s.type == *t.type && same_tuple(s.perm, t.perm, k);
            }
            if (dup) continue;
        }
        KgonRef d = kgon_at(cuts, stride, kept + added, k);
        *d.type = *t.type;
        memcpy(d.perm, t.perm, (size_t)k * sizeof(int));
        *d.value = *t.value;
        *d.y = 0.0;
        ++added;
    }
    ks_free(&have);
    *count += added;
    *NumAdded = added;
    *NumSubtracted = removed;
    return best;
}

double getViolated_PentagonalInequalities(double *Xm, int N, Pentagonal_Inequality *Pent_Lst,
                                          int *ListSize) {
    return kgon_separate(Xm, N, 5, params.Pent_Trials, 3, params.PentIneq, Pent_Lst,
                         sizeof(Pentagonal_Inequality), ListSize);
}

double updatePentagonalInequalities(Problem *PP, double *y, int *NumAdded, int *NumSubtracted,
                                    int triag) {
    return kgon_update(PP, y, NumAdded, NumSubtracted, triag, 5, &PP->NPentIneq, Pent_Cuts,
                       Pent_List, sizeof(Pentagonal_Inequality), params.Pent_Trials, 3,
                       params.PentIneq, MaxPentIneqAdded);
}

double getViolated_HeptagonalInequalities(double *Xm, int N, Heptagonal_Inequality *Hepta_Lst,
                                          int *ListSize) {
    return kgon_separate(Xm, N, 7, params.Hepta_Trials, 4, params.HeptaIneq, Hepta_Lst,
                         sizeof(Heptagonal_Inequality), ListSize);
}

double updateHeptagonalInequalities(Problem *PP, double *y, int *NumAdded, int *NumSubtracted,
                                    int hept_index) {
    return kgon_update(PP, y, NumAdded, NumSubtracted, hept_index, 7, &PP->NHeptaIneq, Hepta_Cuts,
                       Hepta_List, sizeof(Heptagonal_Inequality), params.Hepta_Trials, 4,
                       params.HeptaIneq, MaxHeptaIneqAdded);
}
