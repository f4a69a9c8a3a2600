// madam_bab.hip -- node evaluation and branch-and-bound driver of
// libmadam_bound.so: createSubproblem / getFixedValue (evaluate.c:39-139),
// SDPbound's cutting-plane loop (bounding.c:25-250) for a batch of nodes in
// lockstep, and a single-process B&B (bab_functions.c:39-69, 226-334 and
// heap.c without MPI).
//
// Every numerical step runs on the GPU: the interior-point warm start
// (k_ipm), the ADMM calls (admm_loop), triangle separation, the simulated
// annealing of pentagonal/heptagonal separation and the GW rounding.  The
// host keeps what the reference keeps in process globals: the cut lists
// (purged, deduplicated and appended here exactly as cutting_planes.c does),
// the incumbent, diff and the parameters that grow on give-up.
#include <stdio.h>

#include <chrono>
#include <memory>
#include <unordered_map>
#include <unordered_set>

#include "madam_cutlists.h"
#include "madam_engine.h"

using namespace mh;
using madam::kgon_list;
using madam::purge_append;

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// createSubproblem (evaluate.c:39-116).  Returns PP->n.
int create_subproblem(int N, const double* SPL, const int* xfixed, const int* solX,
                      std::vector<double>& PPL) {
  const int bs = N - 1;
  int nfree = 0;
  for (int i = 0; i < bs; ++i) nfree += !xfixed[i];
  const int n = nfree + 1;
  PPL.assign((size_t)n * n, 0.0);
  std::vector<double> frow(std::max(nfree, 1), 0.0);
  int col = 0;
  for (int i = 0; i < bs; ++i) {
    if (xfixed[i]) {
      if (solX[i] == 1) {
        int f = 0;
        for (int j = 0; j < bs; ++j)
          if (!xfixed[j]) frow[f++] += SPL[j + (size_t)i * N];
      }
      continue;
    }
    int r = 0;
    double rs = 0.0;
    for (int j = 0; j < bs; ++j) {
      if (xfixed[j]) continue;
      const double v = SPL[j + (size_t)i * N];
      PPL[r + (size_t)col * n] = v;
      rs += v;
      ++r;
    }
    PPL[(n - 1) + (size_t)col * n] = rs;
    ++col;
  }
  for (int i = 0; i < n - 1; ++i) PPL[i + (size_t)(n - 1) * n] = PPL[(n - 1) + (size_t)i * n];
  double corner = 0.0;
  for (int i = 0; i < n - 1; ++i) {
    PPL[i + (size_t)i * n] += 2 * frow[i];
    PPL[(n - 1) + (size_t)i * n] += 2 * frow[i];
    PPL[i + (size_t)(n - 1) * n] += 2 * frow[i];
    corner += PPL[i + (size_t)(n - 1) * n];
  }
  PPL[(n - 1) + (size_t)(n - 1) * n] = corner;
  for (auto& v : PPL) v *= 0.25;
  return n;
}

// getFixedValue (evaluate.c:124-139): the same terms in the same order,
// the loops over the fixed vertices only.
double fixed_value(int N, const double* SPL, const int* xfixed, const int* solX) {
  std::vector<int> fx;
  for (int i = 0; i < N - 1; ++i)
    if (xfixed[i]) fx.push_back(i);
  double v = 0.0;
  for (int i : fx)
    for (int j : fx) v += SPL[j + (size_t)i * N] * solX[i] * solX[j];
  return v;
}

double evaluate_solution(int N, const double* SPL, const int* sol) {
  double v = 0.0;
  for (int i = 0; i < N - 1; ++i)
    if (sol[i])
      for (int j = 0; j < N - 1; ++j) v += SPL[j + (size_t)i * N] * sol[i] * sol[j];
  return v;
}

// Host state of one node during SDPbound (the reference's per-call locals
// and the cut arrays of global_var.h:33-42).
struct Run {
  madam_node* node = nullptr;
  int rank = 1, slot = 0, n = 0;
  double fixedvalue = 0, basic_bound = 0, bound = 0, oldf = 0, f = 0;
  double sigma = 1, tol = 0.05, viol3 = 0, viol5 = 0;
  double lb = 0;  // this node's incumbent when ctx->own_incumbent is set
  int count = 0, nbit = 0, done = 0, giveup = 0, prune = 0;
  bool capture = false;  // write each ADMM call's inputs to ctx->last_call
  std::vector<double> PPL;  // the subproblem when given by the caller (madam_ipm);
                            // empty: built on the device from the node's fixings
  std::vector<madam_tri> tri;
  std::vector<madam_pent> pent;
  std::vector<madam_hept> hept;
  std::vector<double> s;  // cut duals after the last ADMM call, layout [tri|pent|hept]
};

struct Ctx {
  madam_bab_ctx* c;
  Engine& E;
};

// The incumbent a node prunes against: the shared one (Bab_LBGet), or the
// node's own with ctx->own_incumbent.
inline double incumbent(const madam_bab_ctx* c, const Run* r) {
  return c->own_incumbent ? r->lb : c->LB;
}

void store_fracsol(Run& r, int N, const std::vector<double>& lastcol) {
  int idx = 0;
  for (int i = 0; i < N - 1; ++i) {
    if (r.node->xfixed[i])
      r.node->fracsol[i] = (double)r.node->solX[i];
    else
      r.node->fracsol[i] = 0.5 * (lastcol[idx++] + 1.0);
  }
}

int fetch_fracsol(std::vector<Run*>& runs, int N) {
  Engine& E = *g_eng;
  std::vector<std::vector<double>> cols(runs.size());
  DownBatch db(E.stream);
  int rc;
  for (size_t q = 0; q < runs.size(); ++q) {
    Run* r = runs[q];
    cols[q].resize(r->n);
    if ((rc = db.get(cols[q].data(), E.slots[r->slot].X.p + (size_t)(r->n - 1) * r->n,
                     (size_t)r->n)))
      return rc;
  }
  if ((rc = db.finish())) return rc;
  for (size_t q = 0; q < runs.size(); ++q) store_fracsol(*runs[q], N, cols[q]);
  return MADAM_OK;
}

// GW rounding on the runs' slots; updates the incumbent like updateSolution
// (bab_functions.c:114-134).  zero_planes: the call right after the IPM.
int heuristic(madam_bab_ctx* c, std::vector<Run*>& runs, bool zero_planes) {
  Engine& E = *g_eng;
  const hipStream_t st = E.stream;
  const int k = (int)runs.size();
  if (k == 0) return MADAM_OK;
  int rc;
  const int num = zero_planes ? 1 : c->N * (c->gw_mult > 1 ? c->gw_mult : 1);
  std::vector<madam::GwNode> g(k);
  std::vector<NodeView> hv(k);
  int nmax = 2;
  for (int q = 0; q < k; ++q) {
    Run& r = *runs[q];
    Slot& sl = E.slots[r.slot];
    if ((rc = sl.size_gw(r.n, num))) return rc;
    hv[q] = E.hviews[r.slot];
    g[q].num = num;
    g[q].zero_planes = zero_planes ? 1 : 0;
    g[q].seed = c->seed;
    g[q].stream = c->stream_ctr++;
    g[q].fixedvalue = r.fixedvalue;
    g[q].val = sl.gwval.p;
    g[q].zall = sl.gwz.p;
    g[q].best_z = sl.gwbz.p;
    g[q].best_val = sl.gwbest.p;
    nmax = std::max(nmax, r.n);
  }
  if ((rc = E.views.ensure(k)) || (rc = E.gwnodes.ensure(k))) return rc;
  if ((rc = up(E.views.p, hv.data(), k, st)) || (rc = up(E.gwnodes.p, g.data(), k, st))) return rc;
  TIMED(PH_GW, st, madam::launch_gw(E.views.p, E.gwnodes.p, k, nmax, num, st));
  HIP_TRY(hipGetLastError());
  // every run's best value and cut in one round trip
  std::vector<double> bvs(k);
  std::vector<std::vector<signed char>> zs(k);
  DownBatch db(st);
  for (int q = 0; q < k; ++q) {
    Slot& sl = E.slots[runs[q]->slot];
    zs[q].resize(runs[q]->n);
    if ((rc = db.get(&bvs[q], sl.gwbest.p, 1)) ||
        (rc = db.get(zs[q].data(), sl.gwbz.p, (size_t)runs[q]->n)))
      return rc;
  }
  if ((rc = db.finish())) return rc;
  std::vector<int> sol(c->N);
  for (int q = 0; q < k; ++q) {
    Run& r = *runs[q];
    const double bv = bvs[q];
    const std::vector<signed char>& z = zs[q];
    int idx = 0;
    for (int i = 0; i < c->N - 1; ++i)
      sol[i] = r.node->xfixed[i] ? r.node->solX[i] : (z[idx++] + 1) / 2;
    if (bv > incumbent(c, &r)) {
      // confirm on the original objective before taking it as the incumbent
      const double v = evaluate_solution(c->N, c->SPL, sol.data());
      if (v > r.lb) r.lb = v;
      if (v > c->LB) {
        c->LB = v;
        for (int i = 0; i < c->N - 1; ++i) c->best_x[i] = sol[i];
      }
    }
  }
  E.timer.flush();
  return MADAM_OK;
}

int separate_tri(madam_bab_ctx* c, std::vector<Run*>& runs) {
  const int k = (int)runs.size();
  if (k == 0) return MADAM_OK;
  std::vector<int> K(k, c->params.TriIneq), slots(k);
  for (int q = 0; q < k; ++q) slots[q] = runs[q]->slot;
  std::vector<std::vector<madam_tri>> lists;
  std::vector<double> mx;
  int rc = tri_separate_slots(slots, K.data(), c->params.violated_TriIneq, lists, mx);
  if (rc) return rc;
  parallel_for(k, [&](int q) {
    Run& r = *runs[q];
    std::vector<double> dual(r.tri.size());
    for (size_t i = 0; i < r.tri.size(); ++i) dual[i] = r.s[i];
    // dedup by key among the survivors
    std::vector<madam_tri> keep;
    for (size_t i = 0; i < r.tri.size(); ++i) {
      madam_tri x = r.tri[i];
      x.y = dual[i];
      if (x.y >= 1e-5) keep.push_back(x);
    }
    std::unordered_set<uint64_t> have;
    auto key = [](const madam_tri& t) {
      return ((uint64_t)t.type << 48) | ((uint64_t)t.i << 32) | ((uint64_t)t.j << 16) | (uint64_t)t.k;
    };
    for (auto& t : keep) have.insert(key(t));
    for (auto& cand : lists[q]) {
      if (keep.size() == 50000) break;
      if (have.count(key(cand))) continue;
      madam_tri x = cand;
      x.y = 0.0;
      keep.push_back(x);
    }
    r.tri.swap(keep);
    r.viol3 = mx[q];
  });
  return MADAM_OK;
}

// One round of pentagonal (hept = false) or heptagonal separation for a
// set of runs: the annealing runs of every run in one launch, then the
// top-K list and the cut set's purge and append on host threads.  In two
// parts so that two rounds (the pentagonal runs and the heptagonal runs
// known beforehand) can run side by side on the two streams.
struct KgonRound {
  std::vector<Run*> runs;
  std::vector<int> offsets;
  bool hept = false;
  int KK = 5, ntypes = 3, per = 0, total = 0;
  std::vector<int> types;
  std::vector<madam::SaOut> out;
  madam::SaOut* dout = nullptr;
};

// Uploads (on the main stream) and the launch on `run_on`; with run_on !=
// the main stream the caller orders it after the uploads.
// Stream ids id0, id0 + 1, ... in job order.
int kgon_launch(madam_bab_ctx* c, KgonRound& K, bool second, hipStream_t run_on, long long id0) {
  Engine& E = *g_eng;
  const hipStream_t st = E.stream;
  const int k = (int)K.runs.size();
  K.KK = K.hept ? 7 : 5;
  K.ntypes = K.hept ? 4 : 3;
  const int trials = K.hept ? c->params.Hepta_Trials : c->params.Pent_Trials;
  K.per = trials * K.ntypes;
  K.total = K.per * k;
  if (K.total == 0) return MADAM_OK;
  DBuf<double>& sumx = second ? E.sumx2 : E.sumx;
  DBuf<const double*>& xptrs = second ? E.xptrs2 : E.xptrs;
  DBuf<int>& nvals = second ? E.nvals2 : E.nvals;
  DBuf<madam::SaJob>& sajobs = second ? E.sajobs2 : E.sajobs;
  DBuf<madam::SaOut>& saout = second ? E.saout2 : E.saout;
  int rc;
  if ((rc = sumx.ensure(k)) || (rc = xptrs.ensure(k)) || (rc = nvals.ensure(k)) ||
      (rc = sajobs.ensure(K.total)) || (rc = saout.ensure(K.total)))
    return rc;
  std::vector<const double*> xp(k);
  std::vector<int> ns(k);
  int nmax = 2;
  for (int q = 0; q < k; ++q) {
    xp[q] = E.slots[K.runs[q]->slot].X.p;
    ns[q] = K.runs[q]->n;
    nmax = std::max(nmax, ns[q]);
  }
  if ((rc = up(xptrs.p, xp.data(), k, st)) || (rc = up(nvals.p, ns.data(), k, st))) return rc;
  madam::launch_sumx(xptrs.p, nvals.p, sumx.p, k, st);
  std::vector<madam::SaJob> jobs(K.total);
  K.types.assign(K.per, 0);
  for (int t = 0; t < trials; ++t)
    for (int ty = 1; ty <= K.ntypes; ++ty) K.types[t * K.ntypes + ty - 1] = ty;
  for (int q = 0; q < k; ++q)
    for (int r = 0; r < K.per; ++r) {
      madam::SaJob& J = jobs[(size_t)q * K.per + r];
      J.X = xp[q];
      J.sumX = sumx.p + q;
      J.n = ns[q];
      J.k = K.KK;
      J.neg = K.types[r] - 1;
      J.nstride = madam::sa_nstride(nmax);
      J.seed = c->seed;
      J.stream_id = id0 + (long long)q * K.per + r;
    }
  if ((rc = up(sajobs.p, jobs.data(), K.total, st))) return rc;
  if (run_on != st) {
    HIP_TRY(hipEventRecord(E.ev_fork, st));
    HIP_TRY(hipStreamWaitEvent(run_on, E.ev_fork, 0));
  }
  TIMED(PH_SA, run_on,
        madam::launch_qap_sa(sajobs.p, k, K.per, nmax, K.KK, saout.p, E.lds_limit, E.cu_count,
                             run_on));
  HIP_TRY(hipGetLastError());
  K.dout = saout.p;
  return MADAM_OK;
}

// The results (the caller has synchronised), list and cut-set upkeep.
void kgon_finish(madam_bab_ctx* c, KgonRound& K) {
  const int k = (int)K.runs.size();
  if (K.total == 0) return;
  parallel_for(k, [&](int q) {
    std::vector<double> cost(K.per);
    std::vector<int> perms((size_t)K.per * K.KK);
    Run& R = *K.runs[q];
    for (int r = 0; r < K.per; ++r) {
      const madam::SaOut& o = K.out[(size_t)q * K.per + r];
      cost[r] = o.cost;
      for (int a = 0; a < K.KK; ++a) perms[(size_t)r * K.KK + a] = o.perm[a];
    }
    const double* dual = R.s.data() + K.offsets[q];
    double best;
    if (K.hept) {
      std::vector<madam_hept> list;
      kgon_list<madam_hept, 7>(cost, perms, K.types, c->params.HeptaIneq, list, best);
      purge_append(R.hept, dual, list, [](const madam_hept& a, const madam_hept& b) {
        if (a.type != b.type) return false;
        for (int x = 0; x < 7; ++x)
          if (a.perm[x] != b.perm[x]) return false;
        return true;
      });
    } else {
      std::vector<madam_pent> list;
      kgon_list<madam_pent, 5>(cost, perms, K.types, c->params.PentIneq, list, best);
      purge_append(R.pent, dual, list, [](const madam_pent& a, const madam_pent& b) {
        if (a.type != b.type) return false;
        for (int x = 0; x < 5; ++x)
          if (a.perm[x] != b.perm[x]) return false;
        return true;
      });
      R.viol5 = best;
    }
  });
}

// Pentagonal separation of `ps` and, beside it on the second stream,
// heptagonal separation of `hs` (the runs whose heptagonal round does not
// depend on the pentagonal one's result); either may be empty.
int separate_kgon_pair(madam_bab_ctx* c, std::vector<Run*>& ps, const std::vector<int>& poff,
                       std::vector<Run*>& hs, const std::vector<int>& hoff) {
  Engine& E = *g_eng;
  const hipStream_t st = E.stream;
  KgonRound P, H;
  P.runs = ps;
  P.offsets = poff;
  H.runs = hs;
  H.offsets = hoff;
  H.hept = true;
  int rc;
  const bool both = !ps.empty() && !hs.empty();
  // stream ids: the pentagonal runs', then the heptagonal runs'
  const long long np_ids =
      (long long)ps.size() * c->params.Pent_Trials * 3;
  const long long id0 = c->stream_ctr;
  if ((!hs.empty() && (rc = kgon_launch(c, H, both, both ? E.stream2 : st, id0 + np_ids))) ||
      (!ps.empty() && (rc = kgon_launch(c, P, false, st, id0))))
    return rc;
  c->stream_ctr = id0 + P.total + H.total;
  if (both && H.total > 0) {
    HIP_TRY(hipEventRecord(E.ev_join, E.stream2));
    HIP_TRY(hipStreamWaitEvent(st, E.ev_join, 0));
  }
  DownBatch db(st);
  P.out.resize(P.total);
  H.out.resize(H.total);
  if ((rc = db.get(P.out.data(), P.dout, (size_t)P.total)) ||
      (rc = db.get(H.out.data(), H.dout, (size_t)H.total)) || (rc = db.finish()))
    return rc;
  E.timer.flush();
  kgon_finish(c, P);
  kgon_finish(c, H);
  return MADAM_OK;
}

int separate_kgon(madam_bab_ctx* c, std::vector<Run*>& runs, bool hept,
                  const std::vector<int>& offsets) {
  std::vector<Run*> none;
  const std::vector<int> noff;
  return hept ? separate_kgon_pair(c, none, noff, runs, offsets)
              : separate_kgon_pair(c, runs, offsets, none, noff);
}

// ADMM calls in flight: their runs and device views, the views kept up to
// date on the host between chunks of iterations.
struct Flight {
  std::vector<Run*> runs;
  std::vector<NodeView> hv;
  std::vector<int> m;  // cuts of each call
};

madam_opts bab_opts(const madam_bab_ctx* c) {
  madam_opts o;
  o.cg_rtol = c->cg_rtol > 0 ? c->cg_rtol : 1e-12;
  o.cg_maxit = 1000;
  o.jacobi_sweeps = 40;
  o.check_every = 16;
  o.projection = c->projection;
  o.ns_layout = c->ns_layout;
  o.cg_adapt = c->cg_adapt;
  return o;
}

// The inputs of the ADMM call run r is about to start -- X, y, Q in its
// slot (the previous call's or the IPM's iterates), the cut lists with the
// duals the call starts from, sigma and tol -- into ctx->last_call.
int capture_call(madam_bab_ctx* c, Run& r, Slot& sl, const CutSet& cs) {
  madam_call_state& S = *c->last_call;
  if (S.cap_n < r.n || S.cap_tri < (int)r.tri.size() || S.cap_pent < (int)r.pent.size() ||
      S.cap_hept < (int)r.hept.size() || !S.X || !S.y || !S.Q ||
      (!r.tri.empty() && !S.tri) || (!r.pent.empty() && !S.pent) || (!r.hept.empty() && !S.hept)) {
    set_err("last_call: buffers too small for n=%d and %zu/%zu/%zu cuts", r.n, r.tri.size(),
            r.pent.size(), r.hept.size());
    return MADAM_E_ARG;
  }
  Engine& E = *g_eng;
  const size_t nn = (size_t)r.n * r.n;
  int rc;
  if ((rc = down(S.X, sl.X.p, nn, E.stream)) || (rc = down(S.Q, sl.Q.p, nn, E.stream)) ||
      (rc = down(S.y, sl.y.p, (size_t)r.n, E.stream)))
    return rc;
  HIP_TRY(hipStreamSynchronize(E.stream));
  // the duals the call starts from are the cut structs' y (build_cutset's y0)
  const int nt = (int)r.tri.size(), np = (int)r.pent.size();
  for (int q = 0; q < nt; ++q) {
    S.tri[q] = r.tri[q];
    S.tri[q].y = cs.y0[q];
  }
  for (int q = 0; q < np; ++q) {
    S.pent[q] = r.pent[q];
    S.pent[q].y = cs.y0[nt + q];
  }
  for (size_t q = 0; q < r.hept.size(); ++q) {
    S.hept[q] = r.hept[q];
    S.hept[q].y = cs.y0[nt + np + q];
  }
  S.n = r.n;
  S.ntri = nt;
  S.npent = np;
  S.nhept = (int)r.hept.size();
  S.sigma = r.sigma;
  S.tol = r.tol;
  S.calls = r.count;
  return MADAM_OK;
}

// Start one ADMM_solver call (admm.c:49-771) per run: its cut structure,
// dual warm start and a fresh view join the calls in flight.
int begin_calls(madam_bab_ctx* c, std::vector<Run*>& runs, Flight& F) {
  Engine& E = *g_eng;
  const hipStream_t st = E.stream;
  const int k = (int)runs.size();
  if (k == 0) return MADAM_OK;
  int rc;
  for (Run* r : runs) {
    ++r->count;
    r->oldf = r->f;
  }
  std::vector<CutSet> cs(k);
  // cut structures on host threads; a node whose cuts are invalid is built
  // again on this thread so that its error message is set
  std::vector<char> bad(k, 0);
  parallel_for(k, [&](int q) {
    Run& r = *runs[q];
    if (!cuts_valid(r.n, r.tri, r.pent, r.hept)) {
      bad[q] = 1;
      return;
    }
    build_cutset(r.n, (int)r.tri.size(), r.tri.data(), (int)r.pent.size(), r.pent.data(),
                 (int)r.hept.size(), r.hept.data(), cs[q]);
  });
  // every node's cut structure, duals and clears: one copy and one launch;
  // cs outlives the batch, so the copies into the ring run on host threads
  UpBatch ub(st, true);
  for (int q = 0; q < k; ++q) {
    Run& r = *runs[q];
    if (bad[q] &&
        (rc = build_cutset(r.n, (int)r.tri.size(), r.tri.data(), (int)r.pent.size(), r.pent.data(),
                           (int)r.hept.size(), r.hept.data(), cs[q])))
      return rc;
    Slot& sl = E.slots[r.slot];
    if ((rc = sl.size_for(r.n, cs[q])) || (rc = upload_cuts(sl, cs[q], st))) return rc;
    if (c->projection == 1 && (rc = sl.size_ns(r.n))) return rc;
    if (cs[q].m > 0) {
      if ((rc = up(sl.s.p, cs[q].y0.data(), cs[q].m, st)) ||
          (rc = up(sl.t.p, cs[q].y0.data(), cs[q].m, st)))
        return rc;
      if ((rc = zero(sl.u.p, cs[q].m * sizeof(double), st))) return rc;
    }
    if (r.capture && c->last_call && (rc = capture_call(c, r, sl, cs[q]))) return rc;
    NodeView v = E.hviews[r.slot];
    NodeView fresh;
    bind_view(fresh, sl, r.n, cs[q]);
    // keep the node's scalars, take the new cut structure pointers
    fresh.normL = v.normL;
    fresh.eig_warm = 0;
    fresh.sigma = r.sigma;
    fresh.tol = r.tol;
    fresh.scaleSigma = c->params.scaleSigma;
    fresh.itermax = c->params.ADMM_itermax;
    F.runs.push_back(&r);
    F.hv.push_back(fresh);
    F.m.push_back(cs[q].m);
  }
  return ub.flush();
}

// Advance the calls in flight until at least min_done of them (all, with
// min_done <= 0) have stopped; those leave the flight with their safe bound
// (admm.c:718-763) and duals, listed in `done`.
int advance_calls(madam_bab_ctx* c, Flight& F, int min_done, std::vector<Run*>& done) {
  Engine& E = *g_eng;
  const hipStream_t st = E.stream;
  done.clear();
  const int k = (int)F.runs.size();
  if (k == 0) return MADAM_OK;
  int rc, nmax = 2, mmax = 1;
  for (int q = 0; q < k; ++q) {
    nmax = std::max(nmax, F.runs[q]->n);
    mmax = std::max(mmax, F.m[q]);
  }
  const madam_opts o = bab_opts(c);
  if ((rc = admm_advance(F.hv, o, nmax, mmax, min_done >= k ? 0 : min_done))) return rc;
  Flight rest;
  std::vector<NodeView> hd;
  std::vector<int> md;
  int nd = 2, ndm = 1;
  for (int q = 0; q < k; ++q) {
    if (F.hv[q].done) {
      done.push_back(F.runs[q]);
      hd.push_back(F.hv[q]);
      md.push_back(F.m[q]);
      nd = std::max(nd, F.runs[q]->n);
      ndm = std::max(ndm, F.m[q]);
    } else {
      rest.runs.push_back(F.runs[q]);
      rest.hv.push_back(F.hv[q]);
      rest.m.push_back(F.m[q]);
    }
  }
  if ((rc = admm_finish(hd, o, nd, ndm))) return rc;
  DownBatch db(st);
  for (size_t q = 0; q < done.size(); ++q) {
    Run& r = *done[q];
    E.hviews[r.slot] = hd[q];
    r.sigma = hd[q].sigma;
    r.nbit = hd[q].nbit;
    r.f = hd[q].f;
    c->admm_iters += hd[q].nbit;
    c->cg_iters += hd[q].cg_total;
    c->ns_fallbacks += hd[q].ns_fallbacks;
    c->admm_calls += 1;
    r.s.resize(md[q]);
    if ((rc = db.get(r.s.data(), E.slots[r.slot].s.p, (size_t)md[q]))) return rc;
  }
  if ((rc = db.finish())) return rc;
  F.runs.swap(rest.runs);
  F.hv.swap(rest.hv);
  F.m.swap(rest.m);
  return MADAM_OK;
}

}  // namespace


namespace {

int ipm_round(madam_bab_ctx* c, std::vector<Run*>& runs) {
  Engine& E = *g_eng;
  const hipStream_t st = E.stream;
  const int k = (int)runs.size();
  int rc, nmax = 2;
  std::vector<NodeView> hv(k);
  std::vector<const int*> xf, sx;
  std::vector<double*> dst;
  std::vector<int> dq;
  UpBatch ub(st);  // host-built subproblems: one copy and one launch
  for (int q = 0; q < k; ++q) {
    Run& r = *runs[q];
    Slot& sl = E.slots[r.slot];
    CutSet cs0;
    if ((rc = build_cutset(r.n, 0, nullptr, 0, nullptr, 0, nullptr, cs0)) ||
        (rc = sl.size_for(r.n, cs0)) || (rc = sl.size_ipm(r.n)))
      return rc;
    bind_view(hv[q], sl, r.n, cs0);
    if (r.PPL.empty()) {  // createSubproblem on the device, straight into the slot
      xf.push_back(r.node->xfixed);
      sx.push_back(r.node->solX);
      dst.push_back(sl.L.p);
      dq.push_back(q);
    } else {
      if ((rc = up(sl.L.p, r.PPL.data(), r.PPL.size(), st))) return rc;
      hv[q].normL = frob_norm_cols(r.PPL.data(), r.n);
    }
    nmax = std::max(nmax, r.n);
  }
  if ((rc = ub.flush())) return rc;
  if (!dq.empty()) {
    std::vector<int> nd(dq.size());
    std::vector<double> nrm(dq.size());
    if ((rc = device_subproblems(c->N, (int)dq.size(), xf.data(), sx.data(), dst.data(), nd.data(),
                                 nrm.data())))
      return rc;
    for (size_t t = 0; t < dq.size(); ++t) {
      if (nd[t] != runs[dq[t]]->n) {
        set_err("device subproblem of order %d, expected %d", nd[t], runs[dq[t]]->n);
        return MADAM_E_HIP;
      }
      hv[dq[t]].normL = nrm[t];
    }
  }
  if ((rc = E.views.ensure(k)) || (rc = up(E.views.p, hv.data(), k, st))) return rc;
  if (madam::ipm_lds_bytes(nmax) <= E.lds_limit) {
    TIMED(PH_IPM, st, madam::launch_ipm(E.views.p, k, nmax, st));
  } else {
    TIMED(PH_IPM, st, rc = ipm_tiles(E.views.p, hv.data(), k, nmax, st));
    if (rc) return rc;
  }
  HIP_TRY(hipGetLastError());
  if ((rc = down(hv.data(), E.views.p, k, st))) return rc;
  HIP_TRY(hipStreamSynchronize(st));
  E.timer.flush();
  for (int q = 0; q < k; ++q) {
    Run& r = *runs[q];
    if (hv[q].ipm_status) {
      set_err("interior-point method failed (status %d) on a node of order %d", hv[q].ipm_status,
              r.n);
      return MADAM_E_EIG;
    }
    E.hviews[r.slot] = hv[q];
    r.f = hv[q].f;
    c->ipm_iters += hv[q].ipm_iters;
  }
  return MADAM_OK;
}

// A run for node x in device slot `slot` (bounding.c:25-64: subproblem,
// fixed value, sigma0, tol0).
void init_run(madam_bab_ctx* c, Run& r, madam_node* x, int rank, int slot) {
  r = Run();
  r.node = x;
  r.rank = rank;
  r.slot = slot;
  int nfree = 0;
  for (int i = 0; i < c->N - 1; ++i) nfree += !x->xfixed[i];
  r.n = nfree + 1;  // PP->L itself is built on the device (ipm_round)
  r.fixedvalue = fixed_value(c->N, c->SPL, x->xfixed, x->solX);
  r.sigma = c->params.sigma0;
  r.tol = c->params.tol0;
  r.lb = c->LB;
}

// IPM warm start, first heuristic, the prune and diff tests and the first
// triangle separation (bounding.c:65-128) for runs just started; a run
// pruned here gets done = 2.
int start_runs(madam_bab_ctx* c, std::vector<Run*>& all) {
  madam_params& P = c->params;
  int rc;
  double t0 = now_s();
  if ((rc = ipm_round(c, all))) return rc;
  c->seconds_ipm += now_s() - t0;
  for (Run* r : all) r->basic_bound = r->f + r->fixedvalue;
  if ((rc = fetch_fracsol(all, c->N))) return rc;
  t0 = now_s();
  if ((rc = heuristic(c, all, true))) return rc;
  c->seconds_heur += now_s() - t0;
  std::vector<Run*> act;
  for (Run* r : all) {
    r->bound = r->f + r->fixedvalue;
    if (r->bound < incumbent(c, r) + 1.0) {
      r->done = 2;  // pruned before the loop: END without diff
      continue;
    }
    if (P.use_diff && r->rank != 0 && r->bound > incumbent(c, r) + c->diff + 1.0) {
      r->done = 2;
      continue;
    }
    act.push_back(r);
  }
  t0 = now_s();
  for (Run* r : act) r->s.clear();
  if ((rc = separate_tri(c, act))) return rc;
  c->seconds_sep += now_s() - t0;
  return MADAM_OK;
}

// One round of the cutting-plane loop (bounding.c:130-236) for the active
// runs: an ADMM call each, the heuristic, the give-up and prune tests and
// the separation for the runs that go on; done is set for the others.
// The rest of a cutting-plane round (bounding.c:138-236) for runs whose
// ADMM call has finished: tolerance, heuristic, prune and give-up tests,
// separation; sets done for runs that leave the loop.
int end_calls(madam_bab_ctx* c, std::vector<Run*>& act) {
  madam_params& P = c->params;
  int rc;
  double t0;
  std::vector<Run*> heur;
  for (Run* r : act) {
    r->tol *= P.scaleTol;
    if (r->tol < P.minTol) r->tol = P.minTol;
    r->bound = r->f + r->fixedvalue;
    r->prune = r->bound < incumbent(c, r) + 1.0;
    if (!r->prune) heur.push_back(r);
  }
  t0 = now_s();
  if ((rc = heuristic(c, heur, false))) return rc;
  c->seconds_heur += now_s() - t0;
  std::vector<Run*> sep;
  for (Run* r : act) {
    if (!r->prune) r->prune = r->bound < incumbent(c, r) + 1.0;
    const double gap = r->bound - incumbent(c, r);
    if (r->count == P.triag_iter + P.pent_iter + P.hept_iter &&
        gap - 1.0 > (r->oldf - r->f) * (P.max_outer_iter - r->count))
      r->giveup = 1;
    if (r->count == P.max_outer_iter && gap - 1.0 > (r->oldf - r->f) * P.extra_iter)
      r->giveup = 1;
    if (r->count == P.max_outer_iter + P.extra_iter) r->giveup = 1;
    if (r->rank != 0 && r->giveup && !r->prune) {
      P.Pent_Trials += 60;
      P.Hepta_Trials += 50;
    }
    if (!r->prune && !r->giveup) sep.push_back(r);
  }
  t0 = now_s();
  if (!sep.empty()) {
    std::vector<int> triag(sep.size()), penta(sep.size());
    for (size_t q = 0; q < sep.size(); ++q) {
      triag[q] = (int)sep[q]->tri.size();
      penta[q] = (int)sep[q]->pent.size();
    }
    if ((rc = separate_tri(c, sep))) return rc;
    // pentagonal separation; heptagonal separation beside it for the runs
    // past the pentagonal rounds, which take it whatever the pentagonal
    // result, then for the runs that take it because of that result
    std::vector<Run*> ps, hs_now, hs_after;
    std::vector<int> poff, hoff_now, hoff_after;
    for (size_t q = 0; q < sep.size(); ++q) {
      Run* r = sep[q];
      if (P.include_Pent && (r->count >= P.triag_iter || r->viol3 < 0.2)) {
        ps.push_back(r);
        poff.push_back(triag[q]);
      }
      if (P.include_Hepta && r->count >= P.triag_iter + P.pent_iter) {
        hs_now.push_back(r);
        hoff_now.push_back(triag[q] + penta[q]);
      }
    }
    if ((rc = separate_kgon_pair(c, ps, poff, hs_now, hoff_now))) return rc;
    for (size_t q = 0; q < sep.size(); ++q) {
      Run* r = sep[q];
      if (P.include_Hepta && !(r->count >= P.triag_iter + P.pent_iter) &&
          (r->viol3 < 0.2 && (1 - r->viol5 < 0.4))) {
        hs_after.push_back(r);
        hoff_after.push_back(triag[q] + penta[q]);
      }
    }
    if ((rc = separate_kgon(c, hs_after, true, hoff_after))) return rc;
  }
  c->seconds_sep += now_s() - t0;
  for (Run* r : act) r->done = r->prune || r->giveup || (r->nbit >= P.ADMM_itermax);
  return fetch_fracsol(act, c->N);
}

// Cutting-plane rounds of a set of runs.  A step starts an ADMM call for
// the runs that need one, advances the calls in flight until a fraction
// `part` of them (at least one) have stopped, and finishes the round of
// those.  A fixed batch of nodes uses an eighth: a run whose call converges
// early starts its next round without waiting for the slowest one.  The
// B&B pipeline uses 0.4: the nodes that finish are branched and replaced
// together, so their successors' IPMs run as one large batch, while the
// slowest calls no longer hold the whole pipeline (pw05_150.0 proof,
// pipeline of 1024: 217 s when every call runs to its end, 207 s at 0.4,
// 210 s at 0.5, 213 s at 0.3, 263 s at an eighth, where the IPM batches
// get small).  `pending` holds the runs between calls.
struct Rounds {
  Flight F;
  std::vector<Run*> pending;  // need a new ADMM call
};

int step_rounds(madam_bab_ctx* c, Rounds& R, double part) {
  int rc;
  double t0 = now_s();
  if ((rc = begin_calls(c, R.pending, R.F))) return rc;
  R.pending.clear();
  std::vector<Run*> fin;
  const int k = (int)R.F.runs.size();
  if ((rc = advance_calls(c, R.F, std::max(1, (int)(k * part)), fin))) return rc;
  c->seconds_admm += now_s() - t0;
  if ((rc = end_calls(c, fin))) return rc;
  for (Run* r : fin)
    if (!r->done) R.pending.push_back(r);
  return MADAM_OK;
}


// The bound a finished run reports (bounding.c:238-250).
void finish_run(madam_bab_ctx* c, Run* r) {
  if (r->done != 2) {
    r->bound = r->f + r->fixedvalue;
    if (r->rank == 0) c->diff = r->basic_bound - r->bound;
  }
  r->node->upper_bound = r->bound;
}

// SDPbound (bounding.c:25-250) for k nodes, one cutting-plane round of all
// of them per step.
int sdpbound_batch(madam_bab_ctx* c, int k, madam_node* const* nodes, const int* ranks,
                   double* bounds) {
  int rc;
  if ((rc = ensure_slots(k))) return rc;
  std::vector<Run> runs(k);
  std::vector<Run*> all;
  for (int q = 0; q < k; ++q) {
    init_run(c, runs[q], nodes[q], ranks ? ranks[q] : 1, q);
    all.push_back(&runs[q]);
  }
  runs[0].capture = c->last_call != nullptr;
  if ((rc = start_runs(c, all))) return rc;
  Rounds R;
  for (Run* r : all)
    if (!r->done) R.pending.push_back(r);
  const double t_start = now_s();
  int step = 0;
  while (!R.pending.empty() || !R.F.runs.empty()) {
    ++step;
    if ((rc = step_rounds(c, R, 0.125))) return rc;
    if (c->verbose > 0) {
      int nbmax = 0, rounds = 0, left = 0;
      size_t mmax = 0;
      for (Run* r : all) {
        nbmax = std::max(nbmax, r->nbit);
        rounds = std::max(rounds, r->count);
        mmax = std::max(mmax, r->tri.size() + r->pent.size() + r->hept.size());
        left += r->done ? 0 : 1;
      }
      fprintf(stderr, "sdpbound: step %d, %d of %d nodes active (%zu ADMM calls in flight), rounds <= %d, "
              "ADMM iterations <= %d, cuts <= %zu, %.2f s\n",
              step, left, k, R.F.runs.size(), rounds, nbmax, mmax, now_s() - t_start);
      fflush(stderr);
    }
  }
  for (Run* r : all) {
    finish_run(c, r);
    bounds[r->slot] = r->bound;
  }
  c->nodes += k;
  return MADAM_OK;
}

// getBranchingVariable (bab_functions.c:389-423).
int branch_var(const madam_bab_ctx* c, const madam_node* node) {
  int ic = -1;
  if (c->params.branchingStrategy == 0) {
    double mx = -1e9;
    for (int i = 0; i < c->N - 1; ++i)
      if (!node->xfixed[i] && fabs(0.5 - node->fracsol[i]) > mx) {
        ic = i;
        mx = fabs(0.5 - node->fracsol[i]);
      }
  } else {
    double mn = 1e9;
    for (int i = 0; i < c->N - 1; ++i)
      if (!node->xfixed[i] && fabs(0.5 - node->fracsol[i]) < mn) {
        ic = i;
        mn = fabs(0.5 - node->fracsol[i]);
      }
  }
  return ic;
}

// Max-heap on upper_bound (heap.c:27-133); children carry their parent's
// bound (heap.c:143-168 leaves it unset).
struct Heap {
  std::vector<madam_node*> d;
  static bool gt(const madam_node* a, const madam_node* b) { return a->upper_bound > b->upper_bound; }
  void push(madam_node* x) {
    d.push_back(x);
    size_t c = d.size() - 1;
    while (c > 0) {
      size_t p = (c - 1) / 2;
      if (!gt(d[c], d[p])) break;
      std::swap(d[c], d[p]);
      c = p;
    }
  }
  madam_node* pop() {
    madam_node* top = d[0];
    d[0] = d.back();
    d.pop_back();
    size_t c = 0;
    for (;;) {
      size_t l = 2 * c + 1, b = c;
      if (l < d.size() && gt(d[l], d[b])) b = l;
      if (l + 1 < d.size() && gt(d[l + 1], d[b])) b = l + 1;
      if (b == c) break;
      std::swap(d[c], d[b]);
      c = b;
    }
    return top;
  }
};

madam_node* child_of(const madam_bab_ctx* c, const madam_node* parent, int ic, int val) {
  madam_node* x = new madam_node();
  for (int i = 0; i < c->N - 1; ++i) {
    x->xfixed[i] = parent->xfixed[i];
    x->solX[i] = parent->xfixed[i] ? parent->solX[i] : 0;
  }
  x->xfixed[ic] = 1;
  x->solX[ic] = val;
  x->level = parent->level + 1;
  x->upper_bound = parent->upper_bound;
  return x;
}

}  // namespace

extern "C" {

int madam_create_subproblem(int N, const double* SPL, const int* xfixed, const int* solX,
                            double* PPL, int* n_out, double* fixedvalue) {
  if (N < 2 || N > MADAM_NMAX || !SPL || !xfixed || !solX || !PPL || !n_out) {
    set_err("madam_create_subproblem: bad arguments");
    return MADAM_E_ARG;
  }
  std::vector<double> L;
  const int n = create_subproblem(N, SPL, xfixed, solX, L);
  std::copy(L.begin(), L.end(), PPL);
  *n_out = n;
  if (fixedvalue) *fixedvalue = fixed_value(N, SPL, xfixed, solX);
  return MADAM_OK;
}

int madam_sdpbound_batch(madam_bab_ctx* c, int k, madam_node* const* nodes, const int* ranks,
                         double* bounds) {
  int rc = ensure_open();
  if (rc) return rc;
  if (!c || k <= 0 || !nodes || !bounds || c->N < 3 || c->N > MADAM_NMAX || !c->SPL) {
    set_err("madam_sdpbound_batch: bad arguments");
    return MADAM_E_ARG;
  }
  if ((rc = upload_spl(c->N, c->SPL))) return rc;
  return sdpbound_batch(c, k, nodes, ranks, bounds);
}

int madam_bab(madam_bab_ctx* c, long long max_nodes, double* optimum) {
  int rc = ensure_open();
  if (rc) return rc;
  if (!c || c->N < 3 || c->N > MADAM_NMAX || !c->SPL) {
    set_err("madam_bab: bad arguments");
    return MADAM_E_ARG;
  }
  if ((rc = upload_spl(c->N, c->SPL))) return rc;
  if (c->params.adjust_TriIneq) c->params.TriIneq = c->N * 10;
  // one shared incumbent in the tree search, whatever own_incumbent says
  struct Restore {
    int& f;
    int v;
    ~Restore() { f = v; }
  } restore{c->own_incumbent, c->own_incumbent};
  c->own_incumbent = 0;
  const int batch = std::max(1, c->batch);
  madam_node* root = new madam_node();
  int rank0 = 0;
  double b0 = 0;
  if ((rc = sdpbound_batch(c, 1, &root, &rank0, &b0))) {
    delete root;
    return rc;
  }
  c->root_bound = b0;
  root->upper_bound = b0;
  Heap h;
  if (c->LB + 1.0 < b0 && !c->params.root) {
    const int ic = branch_var(c, root);
    h.push(child_of(c, root, ic, 0));
    h.push(child_of(c, root, ic, 1));
  }
  delete root;
  // Node pipeline: up to `batch` nodes in flight, each running its own
  // cutting-plane rounds (step_rounds); a node that finishes is branched at
  // once and its slot refilled from the heap, so the batch stays full while
  // the heap has nodes (in place of bounding `batch` nodes to the last
  // straggler).
  std::vector<std::unique_ptr<Run>> pool;
  std::vector<int> free_slots;
  for (int q = batch - 1; q >= 0; --q) free_slots.push_back(q);
  if ((rc = ensure_slots(batch))) {
    for (auto* x : h.d) delete x;
    return rc;
  }
  long long started = c->nodes;
  const double t_start = now_s();
  double t_print = t_start;
  auto fail = [&](int code) {
    for (auto& r : pool) delete r->node;
    for (auto* x : h.d) delete x;
    return code;
  };
  Rounds R;
  for (;;) {
    std::vector<Run*> fresh;
    while ((int)pool.size() < batch && !h.d.empty() && (max_nodes <= 0 || started < max_nodes)) {
      madam_node* x = h.pop();
      if (!(c->LB + 1.0 < x->upper_bound)) {
        delete x;  // the parent's bound already prunes it
        continue;
      }
      pool.emplace_back(new Run());
      init_run(c, *pool.back(), x, 1, free_slots.back());
      free_slots.pop_back();
      fresh.push_back(pool.back().get());
      ++started;
    }
    if (pool.empty()) break;
    if (!fresh.empty()) {
      if ((rc = start_runs(c, fresh))) return fail(rc);
      for (Run* r : fresh)
        if (!r->done) R.pending.push_back(r);
    }
    if ((!R.pending.empty() || !R.F.runs.empty()) && (rc = step_rounds(c, R, 0.4))) return fail(rc);
    // finished nodes: bound, branch, free the slot
    for (size_t q = 0; q < pool.size();) {
      Run* r = pool[q].get();
      if (!r->done) {
        ++q;
        continue;
      }
      finish_run(c, r);
      madam_node* x = r->node;
      c->nodes += 1;
      if (c->LB + 1.0 < x->upper_bound) {
        const int ic = branch_var(c, x);
        if (ic >= 0) {
          h.push(child_of(c, x, ic, 0));
          h.push(child_of(c, x, ic, 1));
        }
      }
      delete x;
      free_slots.push_back(r->slot);
      pool[q] = std::move(pool.back());
      pool.pop_back();
    }
    if (c->verbose > 0 && (now_s() - t_print > 2.0 || pool.empty())) {
      t_print = now_s();
      fprintf(stderr, "bab: %lld nodes bounded, %zu in flight, %zu open, incumbent %.0f, %.2f s\n",
              c->nodes, pool.size(), h.d.size(), c->LB, t_print - t_start);
      fflush(stderr);
    }
  }
  c->open_nodes = (long long)h.d.size();
  for (auto* x : h.d) delete x;
  if (optimum) *optimum = c->LB;
  return MADAM_OK;
}

}  // extern "C"

extern "C" {

int madam_ipm(int n, const double* L, double* X, double* y, double* Z, double* phi, int* iters) {
  int rc = ensure_open();
  if (rc) return rc;
  if (n < 2 || n > MADAM_NMAX || !L || !X || !y || !Z || !phi) {
    set_err("madam_ipm: bad arguments");
    return MADAM_E_ARG;
  }
  if ((rc = ensure_slots(1))) return rc;
  madam_bab_ctx c{};
  Run r;
  r.slot = 0;
  r.n = n;
  r.PPL.assign(L, L + (size_t)n * n);
  std::vector<Run*> runs{&r};
  if ((rc = ipm_round(&c, runs))) return rc;
  Engine& E = *g_eng;
  Slot& sl = E.slots[0];
  const size_t nn = (size_t)n * n;
  if ((rc = down(X, sl.X.p, nn, E.stream)) || (rc = down(Z, sl.Q.p, nn, E.stream)) ||
      (rc = down(y, sl.y.p, (size_t)n, E.stream)))
    return rc;
  HIP_TRY(hipStreamSynchronize(E.stream));
  *phi = r.f;
  if (iters) *iters = (int)c.ipm_iters;
  return MADAM_OK;
}

int madam_gw_rounding(int n, const double* L, const double* Z, int M, int num,
                      unsigned long long seed, long long stream, double* values,
                      signed char* cuts) {
  int rc = ensure_open();
  if (rc) return rc;
  if (n < 2 || n > MADAM_NMAX || M < 0 || M > n || num <= 0 || !L || !values) {
    set_err("madam_gw_rounding: bad arguments");
    return MADAM_E_ARG;
  }
  if ((rc = ensure_slots(1))) return rc;
  Engine& E = *g_eng;
  const hipStream_t st = E.stream;
  Slot& sl = E.slots[0];
  CutSet cs0;
  if ((rc = build_cutset(n, 0, nullptr, 0, nullptr, 0, nullptr, cs0)) || (rc = sl.size_for(n, cs0)) ||
      (rc = sl.size_gw(n, num)))
    return rc;
  const size_t nn = (size_t)n * n;
  if ((rc = up(sl.L.p, L, nn, st))) return rc;
  if (M > 0 && (rc = up(sl.S1.p, Z, (size_t)n * M, st))) return rc;
  NodeView v;
  bind_view(v, sl, n, cs0);
  madam::GwNode g{};
  g.num = num;
  g.M = M;
  g.seed = seed;
  g.stream = stream;
  g.val = sl.gwval.p;
  g.zall = sl.gwz.p;
  g.best_z = sl.gwbz.p;
  g.best_val = sl.gwbest.p;
  if ((rc = E.views.ensure(1)) || (rc = E.gwnodes.ensure(1)) || (rc = up(E.views.p, &v, 1, st)) ||
      (rc = up(E.gwnodes.p, &g, 1, st)))
    return rc;
  madam::launch_gw_given_z(E.views.p, E.gwnodes.p, 1, n, num, st);
  HIP_TRY(hipGetLastError());
  if ((rc = down(values, sl.gwval.p, num, st))) return rc;
  if (cuts && (rc = down(cuts, sl.gwz.p, (size_t)num * n, st))) return rc;
  HIP_TRY(hipStreamSynchronize(st));
  return MADAM_OK;
}

}  // extern "C"
