// madam_host.hip -- host engine and C ABI of libmadam_bound.so.
//
// The engine owns one HIP stream and a pool of node slots in HBM.  A slot
// holds one node's matrices, cut structure and ADMM vectors (NodeView).
// madam_admm_batch() uploads k jobs into k slots and runs the ADMM loop of
// admm.c:520-704 for all of them with the same kernel launches, checking the
// per-node stop flags on the host every `check_every` iterations; nodes that
// stopped skip the remaining launches inside the kernels.
#include "madam_engine.h"
#include "madam_ns_coef.h"

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "../../include/madam_bound.h"

static_assert(sizeof(madam_tri) == 32, "Triangle_Inequality layout");
static_assert(sizeof(madam_pent) == 40, "Pentagonal_Inequality layout");
static_assert(sizeof(madam_hept) == 48, "Heptagonal_Inequality layout");


namespace mh {
char g_err[512] = "";
Engine* g_eng = nullptr;
}  // namespace mh
using namespace mh;

// Sign-iteration layout table written by the autotuner (madam_amd/tune.py,
// loaded by madam_load_tuning): cells (n_p, active nodes) -> layout (1 one
// workgroup per node, 2 tiles, 3 2 x 2 tiles).  Empty: the built-in
// crossovers of admm_advance.
namespace {
struct TuneCell {
  int np, nb, layout;
};
std::vector<TuneCell> g_tune;

// LDS bytes the cut solve may count on (0: the device's; lower caps move
// nodes to the kernels with vectors in HBM) and whether its one-pass kernel
// may serve nodes (madam_set_cg_lds_cap)
size_t g_cg_lds_cap = 0;
int g_cg_one_pass = 1;

// The cell of the smallest tabulated n_p >= np (else the largest) and of
// the largest tabulated node count <= active (else the smallest); 0
// without a table.
int tuned_layout(int np, int active) {
  if (g_tune.empty()) return 0;
  int sel = -1, top = -1;
  for (const TuneCell& c : g_tune) {
    if (c.np >= np && (sel < 0 || c.np < sel)) sel = c.np;
    top = std::max(top, c.np);
  }
  if (sel < 0) sel = top;
  int best_nb = -1, lay = 0, min_nb = 0, lay_min = 0;
  for (const TuneCell& c : g_tune) {
    if (c.np != sel) continue;
    if (c.nb <= active && c.nb > best_nb) {
      best_nb = c.nb;
      lay = c.layout;
    }
    if (lay_min == 0 || c.nb < min_nb) {
      min_nb = c.nb;
      lay_min = c.layout;
    }
  }
  return best_nb >= 0 ? lay : lay_min;
}
}  // namespace

int mh::admm_loop(std::vector<NodeView>& hv, const madam_opts& o, int nmax, int mmax) {
  int rc = admm_advance(hv, o, nmax, mmax, 0);
  if (rc) return rc;
  return admm_finish(hv, o, nmax, mmax);
}

// The iterations of one check interval as a HIP graph (off by default,
// madam_set_graphs).  Every launch of an interval is fixed by a few host
// values (batch size and orders, layout, the cut-solve plan, the options and
// the device arrays the launches take), so an interval whose values equal
// the previous interval's issues the same launches: the second such interval
// is captured into a graph and later ones replay it with one hipGraphLaunch.
// The graph holds kernel nodes only (none of these launches clears memory),
// and the phase timer runs the launches directly (its events bracket each
// phase).  Measured on MI355X, replay does not pay for the capture: a
// signature recurs for one ADMM call (~17 intervals at n = 512, fewer in
// batches), and the kernels already run back to back -- B&B step of the
// 8x8x8 torus 1.856 s with graphs against 1.834 s direct, root bound of
// G(100, 0.3) 226 against 210 ms, 64 torus nodes 3.97 against 3.87 s.
namespace {
struct ChunkGraph {
  hipGraphExec_t exec = nullptr;
  std::vector<long long> sig;   // signature of the captured interval
  std::vector<long long> seen;  // signature of the interval before
  long replays = 0, captures = 0;
};
ChunkGraph g_chunk;
int g_use_graphs = 0;

long long bits_of(double d) {
  long long b;
  memcpy(&b, &d, sizeof b);
  return b;
}
}  // namespace

void mh::release_chunk_graph() {
  if (g_chunk.exec) (void)hipGraphExecDestroy(g_chunk.exec);
  g_chunk.exec = nullptr;
  g_chunk.sig.clear();
  g_chunk.seen.clear();
}

// Run the ADMM iterations of the views in hv (fresh calls or calls left
// running by an earlier return) in chunks of check_every iterations until
// every node has stopped, or, with min_done > 0, until at least min_done
// of them have.  hv is brought up to date from the device after each chunk.
int mh::admm_advance(std::vector<NodeView>& hv, const madam_opts& o, int nmax, int mmax,
                     int min_done) {
  Engine& E = *g_eng;
  const hipStream_t st = E.stream;
  const int k = (int)hv.size();
  int rc;
  if ((rc = E.views.ensure(k))) return rc;
  if ((rc = E.postpart.ensure((size_t)k * 64 * 8))) return rc;
  if ((rc = up(E.views.p, hv.data(), k, st))) return rc;
  const int chunk = std::max(1, o.check_every);
  int itmax_all = 0, done0 = 0;
  for (int b = 0; b < k; ++b) {
    itmax_all = std::max(itmax_all, hv[b].itermax - hv[b].nbit);
    done0 += hv[b].done ? 1 : 0;
  }
  long launched = 0;
  int active = k - done0;
  if (active == 0) return MADAM_OK;
  const size_t cg_lds = g_cg_lds_cap ? std::min(g_cg_lds_cap, E.lds_limit) : E.lds_limit;
  for (;;) {
    if (launched > (long)itmax_all + 2 * chunk + 8) {
      set_err("ADMM stop flags never set after %ld iterations", launched);
      return MADAM_E_HIP;
    }
    launched += chunk;
    // tile-parallel sign iteration while the active nodes cannot fill the
    // CUs with one workgroup each.  Crossovers measured per projection:
    // order <= 112 (iterates in LDS in the one-workgroup layout, flat
    // 0.25 ms at n_p = 112, 0.125 ms at 80, 0.093 ms at 64): tiles up to
    // 96 active nodes at n_p = 112 (0.13 ms for one node), 16 at n_p = 80,
    // never at 64; larger orders: tiles below ~3/4 of the CU count (n = 150:
    // 1.0 against 1.13 ms at 192 nodes, 1.31 against 1.20 ms at 256)
    // (when no autotuned table is loaded)
    const bool small = madam::ns_lds_bytes(nmax) + 1024 <= E.lds_limit;
    const int np = (nmax + 15) & ~15;
    const int small_tiles = np >= 112 ? 96 : (np >= 96 ? 48 : (np >= 80 ? 16 : 0));
    int layout = o.ns_layout;
    if (layout == 0) layout = tuned_layout(np, active);
    if (layout == 0)
      layout = (small ? active <= small_tiles : active < (3 * E.cu_count) / 4) ? 2 : 1;
    const madam::CgPlan plan = madam::plan_cg(hv.data(), k, cg_lds, g_cg_one_pass);
    // the block Jacobi solver of the plain form reads its stop flags on the
    // host between sweeps (no graph capture then)
    const bool plain_bj = o.projection != 1 && bj_wanted(nmax);
    int nmin = nmax;
    for (int b = 0; b < k; ++b) nmin = std::min(nmin, hv[b].n);
    auto issue = [&]() -> int {
      for (int it = 0; it < chunk; ++it) {
        TIMED(PH_PREP, st, madam::launch_admm_prep(E.views.p, k, mmax, st));
        TIMED(PH_CG, st,
              madam::launch_cg_plan(E.views.p, k, plan, o.cg_rtol, o.cg_adapt, o.cg_maxit,
                                    cg_lds, g_cg_one_pass, st));
        TIMED(PH_W, st, madam::launch_build_w(E.views.p, k, nmax, mmax, 0, st));
        if (o.projection == 1) {
          TIMED(PH_EIG, st, {
            if (layout == 1)
              madam::launch_ns_project(E.views.p, k, nmax, 0, E.lds_limit, st);
            else
              madam::launch_ns_tiles(E.views.p, k, nmax, st, layout == 3);
            madam::launch_jacobi(E.views.p, k, nmax, 0, o.jacobi_sweeps, 0, E.lds_limit, st, 3);
          });
        } else if (plain_bj) {
          TIMED(PH_EIG, st, {
            rc = bj_eig(E.views.p, hv.data(), k, 0, 0, o.jacobi_sweeps, st);
            if (!rc && nmin < bj_min_order())
              madam::launch_jacobi(E.views.p, k, nmax, 0, o.jacobi_sweeps, 0, E.lds_limit, st, 1,
                                   bj_min_order());
          });
          if (rc) return rc;
        } else {
          TIMED(PH_EIG, st,
                madam::launch_jacobi(E.views.p, k, nmax, 0, o.jacobi_sweeps, 0, E.lds_limit, st,
                                     1));
        }
        TIMED(PH_POST, st, madam::launch_admm_post(E.views.p, k, nmax, mmax, E.postpart.p, st));
      }
      return MADAM_OK;
    };
    bool graphed = false;
    if (g_use_graphs && !E.timer.on && !plain_bj) {
      std::vector<long long> sig = {k, nmax, mmax, layout, chunk, o.projection,
                                    bits_of(o.cg_rtol), bits_of(o.cg_adapt), o.cg_maxit,
                                    o.jacobi_sweeps, (long long)(uintptr_t)E.views.p,
                                    (long long)(uintptr_t)E.postpart.p, (long long)cg_lds,
                                    g_cg_one_pass, (long long)E.lds_limit};
      for (int c = 0; c < 6; ++c) {
        sig.push_back(plan.any[c]);
        sig.push_back((long long)plan.need[c]);
      }
      if (g_chunk.exec && sig == g_chunk.sig) {
        HIP_TRY(hipGraphLaunch(g_chunk.exec, st));
        ++g_chunk.replays;
        graphed = true;
      } else if (sig == g_chunk.seen) {
        hipGraph_t graph = nullptr;
        bool ok = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (ok) {
          ok = issue() == MADAM_OK;
          ok = hipStreamEndCapture(st, &graph) == hipSuccess && graph != nullptr && ok;
        }
        if (ok) {
          hipGraphExecUpdateResult upd;
          hipGraphNode_t bad = nullptr;
          if (!g_chunk.exec ||
              hipGraphExecUpdate(g_chunk.exec, graph, &bad, &upd) != hipSuccess) {
            if (g_chunk.exec) (void)hipGraphExecDestroy(g_chunk.exec);
            g_chunk.exec = nullptr;
            (void)hipGetLastError();
            ok = hipGraphInstantiate(&g_chunk.exec, graph, nullptr, nullptr, 0) == hipSuccess;
          }
        }
        if (graph) (void)hipGraphDestroy(graph);
        if (ok) {
          g_chunk.sig = sig;
          ++g_chunk.captures;
          HIP_TRY(hipGraphLaunch(g_chunk.exec, st));
          ++g_chunk.replays;
          graphed = true;
        } else {
          // a runtime that cannot capture these launches: issue them directly
          // from now on
          (void)hipGetLastError();
          release_chunk_graph();
          g_use_graphs = 0;
        }
      }
      g_chunk.seen = sig;
    }
    if (!graphed && (rc = issue())) return rc;
    HIP_TRY(hipGetLastError());
    if ((rc = down(hv.data(), E.views.p, k, st))) return rc;
    HIP_TRY(hipStreamSynchronize(st));
    E.timer.flush();
    active = 0;
    for (int b = 0; b < k; ++b) active += hv[b].done ? 0 : 1;
    if (active == 0 || (min_done > 0 && k - done0 - active >= min_done)) break;
  }
  return MADAM_OK;
}

// End of the ADMM calls in hv (all stopped): the rounding heuristic's
// factor and the safe bound (admm.c:718-763), on their own view array.
int mh::admm_finish(std::vector<NodeView>& hv, const madam_opts& o, int nmax, int mmax) {
  Engine& E = *g_eng;
  const hipStream_t st = E.stream;
  const int k = (int)hv.size();
  if (k == 0) return MADAM_OK;
  int rc;
  if ((rc = E.fviews.ensure(k))) return rc;
  if ((rc = up(E.fviews.p, hv.data(), k, st))) return rc;
  if (o.projection == 1) {
    // the rounding heuristic's factor Z: pivoted Cholesky of X
    TIMED(PH_ZFAC, st, madam::launch_zfac_chol(E.fviews.p, k, st));
  }
  TIMED(PH_SAFE, st, {
    madam::launch_build_w(E.fviews.p, k, nmax, mmax, 1, st);
    // chosen form: an estimate of lambda_max proved an upper bound by a
    // Cholesky certificate -- Householder tridiagonalisation and the
    // Cholesky in LDS by one workgroup per node, or spread over many
    // workgroups for orders too large for that; plain form: all
    // eigenvalues by Jacobi, as dsyevr in the reference
    if (o.projection == 1 && madam::safe_hh_lds_bytes(nmax) + 1024 <= E.lds_limit)
      madam::launch_safe_hh(E.fviews.p, k, nmax, st);
    else if (o.projection == 1)
      rc = safe_tiles(E.fviews.p, hv.data(), k, nmax, st);
    else {
      int nmin = nmax;
      for (int b = 0; b < k; ++b) nmin = std::min(nmin, hv[b].n);
      if (bj_wanted(nmax)) rc = bj_eig(E.fviews.p, hv.data(), k, 1, 1, 4 * o.jacobi_sweeps, st);
      if (!rc && nmin < bj_min_order())
        madam::launch_jacobi(E.fviews.p, k, nmax, 1, 4 * o.jacobi_sweeps, 1, E.lds_limit, st, 0,
                             bj_min_order());
    }
    if (!rc) madam::launch_safe_bound(E.fviews.p, k, st);
  });
  if (rc) return rc;
  HIP_TRY(hipGetLastError());
  if ((rc = down(hv.data(), E.fviews.p, k, st))) return rc;
  HIP_TRY(hipStreamSynchronize(st));
  E.timer.flush();
  return MADAM_OK;
}

// =================================================================== C ABI

extern "C" {

const char* madam_last_error(void) { return g_err; }

int madam_profile(int on) {
  int rc = ensure_open();
  if (rc) return rc;
  g_eng->timer.on = on != 0;
  return MADAM_OK;
}

int madam_profile_read(int nphase, double* ms, long* count, int reset) {
  int rc = ensure_open();
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(g_eng->stream));
  g_eng->timer.flush();
  for (int p = 0; p < nphase && p < PH_COUNT; ++p) {
    if (ms) ms[p] = g_eng->timer.ms[p];
    if (count) count[p] = g_eng->timer.count[p];
  }
  if (reset)
    for (int p = 0; p < PH_COUNT; ++p) {
      g_eng->timer.ms[p] = 0;
      g_eng->timer.count[p] = 0;
    }
  return MADAM_OK;
}

int madam_open(int device) {
  if (g_eng) return MADAM_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) {
    set_err("no HIP device %d (count %d)", device, ndev);
    return MADAM_E_NODEV;
  }
  HIP_TRY(hipSetDevice(device));
  // the no-allocation entry point's device array: its address on this device
  if (int rc = noalloc_setup(device)) return rc;
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  Engine* e = new Engine();
  e->device = device;
  e->lds_limit = prop.sharedMemPerBlock;
  e->cu_count = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&e->stream2, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&e->ev_join, hipEventDisableTiming) != hipSuccess) {
    delete e;
    set_err("hipStreamCreate / hipEventCreate failed");
    return MADAM_E_HIP;
  }
  g_eng = e;
  return MADAM_OK;
}

void madam_close(void) {
  if (!g_eng) return;
  (void)hipStreamSynchronize(g_eng->stream);
  for (auto& s : g_eng->slots) s.release();
  for (auto& s : g_eng->sep) s.release();
  g_eng->sepviews.release();
  g_eng->sajobs.release();
  g_eng->sajobs2.release();
  g_eng->saout2.release();
  g_eng->sumx2.release();
  g_eng->xptrs2.release();
  g_eng->nvals2.release();
  g_eng->gwnodes.release();
  g_eng->saout.release();
  g_eng->sumx.release();
  g_eng->xptrs.release();
  g_eng->nvals.release();
  g_eng->views.release();
  g_eng->tmpd.release();
  g_eng->postpart.release();
  g_eng->fviews.release();
  release_spl();
  release_ipm_tiles();
  release_safe_tiles();
  release_chunk_graph();
  release_bj();
  (void)hipStreamSynchronize(g_eng->stream2);
  (void)hipStreamDestroy(g_eng->stream2);
  (void)hipEventDestroy(g_eng->ev_fork);
  (void)hipEventDestroy(g_eng->ev_join);
  (void)hipStreamDestroy(g_eng->stream);
  g_ring.release();
  g_dring.release();
  delete g_eng;
  g_eng = nullptr;
}

int madam_admm_batch(int k, madam_admm_job* jobs, const madam_opts* opts_in) {
  int rc = ensure_open();
  if (rc) return rc;
  if (k <= 0 || !jobs) {
    set_err("madam_admm_batch: k=%d", k);
    return MADAM_E_ARG;
  }
  const madam_opts o = opts_in ? *opts_in : default_opts();
  Engine& E = *g_eng;
  if ((rc = ensure_slots(k))) return rc;
  std::vector<CutSet> cs(k);
  int nmax = 1, mmax = 1;
  for (int b = 0; b < k; ++b) {
    madam_admm_job& J = jobs[b];
    if (J.n < 2 || J.n > MADAM_NMAX || !J.L || !J.X || !J.y || !J.Q || !J.sigma) {
      set_err("job %d: bad arguments (n=%d)", b, J.n);
      return MADAM_E_ARG;
    }
    if ((rc = build_cutset(J.n, J.ntri, J.tri, J.npent, J.pent, J.nhept, J.hept, cs[b])))
      return rc;
    nmax = std::max(nmax, J.n);
    mmax = std::max(mmax, cs[b].m);
  }
  const hipStream_t st = E.stream;
  for (int b = 0; b < k; ++b) {
    madam_admm_job& J = jobs[b];
    Slot& sl = E.slots[b];
    if ((rc = sl.size_for(J.n, cs[b]))) return rc;
    if (o.projection == 1 && (rc = sl.size_ns(J.n))) return rc;
    const size_t nn = (size_t)J.n * J.n;
    if ((rc = up(sl.L.p, J.L, nn, st)) || (rc = up(sl.X.p, J.X, nn, st)) ||
        (rc = up(sl.Q.p, J.Q, nn, st)) || (rc = up(sl.y.p, J.y, (size_t)J.n, st)) ||
        (rc = upload_cuts(sl, cs[b], st)))
      return rc;
    // s = t = cut duals, u = 0 (admm.c:79-95)
    if (cs[b].m > 0) {
      if ((rc = up(sl.s.p, cs[b].y0.data(), cs[b].m, st)) ||
          (rc = up(sl.t.p, cs[b].y0.data(), cs[b].m, st)))
        return rc;
      HIP_TRY(hipMemsetAsync(sl.u.p, 0, cs[b].m * sizeof(double), st));
    }
    NodeView& v = E.hviews[b];
    bind_view(v, sl, J.n, cs[b]);
    v.normL = frob_norm_cols(J.L, J.n);
    v.sigma = *J.sigma;
    v.tol = J.tol;
    v.scaleSigma = J.scaleSigma;
    v.itermax = J.itermax;
  }
  {
    std::vector<NodeView> hv(E.hviews.begin(), E.hviews.begin() + k);
    if ((rc = admm_loop(hv, o, nmax, mmax))) return rc;
    std::copy(hv.begin(), hv.end(), E.hviews.begin());
  }

  // ---- results
  std::vector<double> evh, Vh;
  for (int b = 0; b < k; ++b) {
    madam_admm_job& J = jobs[b];
    Slot& sl = E.slots[b];
    NodeView& v = E.hviews[b];
    const size_t nn = (size_t)J.n * J.n;
    const int m = cs[b].m;
    if ((rc = down(J.X, sl.X.p, nn, st)) || (rc = down(J.Q, sl.Q.p, nn, st)) ||
        (rc = down(J.y, sl.y.p, (size_t)J.n, st)))
      return rc;
    if (J.s && (rc = down(J.s, sl.s.p, m, st))) return rc;
    if (J.t && (rc = down(J.t, sl.t.p, m, st))) return rc;
    if (J.u && (rc = down(J.u, sl.u.p, m, st))) return rc;
    if (J.Z) {
      // Z = V_+ diag(sqrt(lambda * sigma_last)), ascending lambda, as
      // project_SDP and admm.c:583-585 leave it (the sign-iteration form:
      // the pivoted Cholesky factor of X, in pivot order).
      evh.resize(J.n);
      Vh.resize(nn);
      if ((rc = down(Vh.data(), sl.V.p, nn, st)) || (rc = down(evh.data(), sl.evp.p, J.n, st)))
        return rc;
      HIP_TRY(hipStreamSynchronize(st));
      std::vector<std::pair<double, int>> pos;
      for (int j = 0; j < J.n; ++j)
        if (evh[j] > 0.0) pos.push_back({evh[j], j});
      std::sort(pos.begin(), pos.end());
      J.M = (int)pos.size();
      for (int q = 0; q < J.M; ++q) {
        const double sc = sqrt(pos[q].first * v.sigma_last);
        const double* vj = &Vh[(size_t)pos[q].second * J.n];
        for (int r = 0; r < J.n; ++r) J.Z[r + (size_t)q * J.n] = vj[r] * sc;
      }
    } else {
      J.M = v.M;
    }
    *J.sigma = v.sigma;
    J.nbit = v.nbit;
    J.f = v.f;
    J.lambda_max = v.lam;
    J.cg_iters = v.cg_total;
    J.sweeps = v.sweeps_total;
  }
  HIP_TRY(hipStreamSynchronize(st));
  return MADAM_OK;
}

// Single-node helpers -----------------------------------------------------

static int one_node(int n, int ntri, const madam_tri* tri, int npent, const madam_pent* pent,
                    int nhept, const madam_hept* hept, CutSet& cs) {
  int rc = ensure_open();
  if (rc) return rc;
  if (n < 2 || n > MADAM_NMAX) {
    set_err("n=%d out of range", n);
    return MADAM_E_ARG;
  }
  if ((rc = build_cutset(n, ntri, tri, npent, pent, nhept, hept, cs))) return rc;
  if ((rc = ensure_slots(1))) return rc;
  Slot& sl = g_eng->slots[0];
  if ((rc = sl.size_for(n, cs))) return rc;
  if ((rc = upload_cuts(sl, cs, g_eng->stream))) return rc;
  bind_view(g_eng->hviews[0], sl, n, cs);
  return up(g_eng->views.p, g_eng->hviews.data(), 1, g_eng->stream);
}

int madam_op_B(int n, const double* X, int ntri, const madam_tri* tri, int npent,
               const madam_pent* pent, int nhept, const madam_hept* hept, double* y) {
  CutSet cs;
  int rc = one_node(n, ntri, tri, npent, pent, nhept, hept, cs);
  if (rc) return rc;
  Slot& sl = g_eng->slots[0];
  const hipStream_t st = g_eng->stream;
  if ((rc = up(sl.X.p, X, (size_t)n * n, st))) return rc;
  madam::launch_op_B(g_eng->views.p, 1, sl.X.p, sl.rhs.p, std::max(cs.m, 1), st);
  HIP_TRY(hipGetLastError());
  std::vector<double> out(cs.m);
  if ((rc = down(out.data(), sl.rhs.p, cs.m, st))) return rc;
  HIP_TRY(hipStreamSynchronize(st));
  for (int c = 0; c < cs.m; ++c) y[c] += out[c];
  return MADAM_OK;
}

int madam_op_Bt(int n, double* X, int ntri, const madam_tri* tri, int npent,
                const madam_pent* pent, int nhept, const madam_hept* hept, const double* t) {
  CutSet cs;
  int rc = one_node(n, ntri, tri, npent, pent, nhept, hept, cs);
  if (rc) return rc;
  Slot& sl = g_eng->slots[0];
  const hipStream_t st = g_eng->stream;
  if ((rc = up(sl.t.p, t, cs.m, st))) return rc;
  madam::launch_op_Bt(g_eng->views.p, 1, sl.t.p, sl.S1.p, n, st);
  HIP_TRY(hipGetLastError());
  std::vector<double> out((size_t)n * n);
  if ((rc = down(out.data(), sl.S1.p, out.size(), st))) return rc;
  HIP_TRY(hipStreamSynchronize(st));
  for (size_t e = 0; e < out.size(); ++e) X[e] += out[e];
  return MADAM_OK;
}

int madam_project_sdp(int n, const double* W, double* X, double* Q, double* Z, int* M) {
  CutSet cs;
  int rc = one_node(n, 0, nullptr, 0, nullptr, 0, nullptr, cs);
  if (rc) return rc;
  Slot& sl = g_eng->slots[0];
  const hipStream_t st = g_eng->stream;
  const size_t nn = (size_t)n * n;
  if ((rc = up(sl.W.p, W, nn, st))) return rc;
  if (bj_wanted(n)) {
    if ((rc = bj_eig(g_eng->views.p, g_eng->hviews.data(), 1, 0, 1, 60, st))) return rc;
  } else {
    madam::launch_jacobi(g_eng->views.p, 1, n, 0, 60, 1, g_eng->lds_limit, st, 1);
  }
  madam::launch_project(g_eng->views.p, 1, n, st);
  HIP_TRY(hipGetLastError());
  if ((rc = sync_views_down(1))) return rc;
  if ((rc = down(X, sl.X.p, nn, st)) || (rc = down(Q, sl.Q.p, nn, st))) return rc;
  std::vector<double> ev(n), Vh(nn);
  if ((rc = down(ev.data(), sl.ev.p, n, st)) || (rc = down(Vh.data(), sl.V.p, nn, st))) return rc;
  HIP_TRY(hipStreamSynchronize(st));
  std::vector<int> idx;
  for (int j = 0; j < n; ++j)
    if (ev[j] > 1e-8) idx.push_back(j);
  std::sort(idx.begin(), idx.end(), [&](int a, int b) { return ev[a] < ev[b]; });
  if (M) *M = (int)idx.size();
  if (Z)
    for (size_t q = 0; q < idx.size(); ++q) {
      const double sc = sqrt(ev[idx[q]]);
      for (int r = 0; r < n; ++r) Z[r + q * n] = Vh[r + (size_t)idx[q] * n] * sc;
    }
  return MADAM_OK;
}

int madam_lambda_max_cert(int n, const double* A, double* lam) {
  CutSet cs;
  int rc = one_node(n, 0, nullptr, 0, nullptr, 0, nullptr, cs);
  if (rc) return rc;
  Slot& sl = g_eng->slots[0];
  const hipStream_t st = g_eng->stream;
  if ((rc = up(sl.W.p, A, (size_t)n * n, st))) return rc;
  HIP_TRY(hipMemsetAsync(sl.y.p, 0, n * sizeof(double), st));
  if (madam::safe_hh_lds_bytes(n) + 1024 <= g_eng->lds_limit)
    madam::launch_safe_hh(g_eng->views.p, 1, n, st);
  else if ((rc = safe_tiles(g_eng->views.p, g_eng->hviews.data(), 1, n, st)))
    return rc;
  madam::launch_safe_bound(g_eng->views.p, 1, st);
  HIP_TRY(hipGetLastError());
  if ((rc = sync_views_down(1))) return rc;
  *lam = g_eng->hviews[0].lam;
  return MADAM_OK;
}

int madam_lambda_max(int n, const double* A, double* lam) {
  CutSet cs;
  int rc = one_node(n, 0, nullptr, 0, nullptr, 0, nullptr, cs);
  if (rc) return rc;
  Slot& sl = g_eng->slots[0];
  const hipStream_t st = g_eng->stream;
  if ((rc = up(sl.W.p, A, (size_t)n * n, st))) return rc;
  // safe bound with y = 0 and no cuts: f = n * lambda_max
  HIP_TRY(hipMemsetAsync(sl.y.p, 0, n * sizeof(double), st));
  if (bj_wanted(n)) {
    if ((rc = bj_eig(g_eng->views.p, g_eng->hviews.data(), 1, 1, 1, 160, st))) return rc;
  } else {
    madam::launch_jacobi(g_eng->views.p, 1, n, 1, 160, 1, g_eng->lds_limit, st, 0);
  }
  madam::launch_safe_bound(g_eng->views.p, 1, st);
  HIP_TRY(hipGetLastError());
  if ((rc = sync_views_down(1))) return rc;
  *lam = g_eng->hviews[0].lam;
  return MADAM_OK;
}


// Triangle separation on node slots [0, k): X of each slot must be current.
// Lists come back in enumeration order (type, i, j, k ascending).
}  // extern "C"
int mh::tri_separate_slots(const std::vector<int>& slots, const int* K, double thresh,
                           std::vector<std::vector<madam_tri>>& lists, std::vector<double>& maxval) {
  Engine& E = *g_eng;
  const int k = (int)slots.size();
  const hipStream_t st = E.stream;
  int rc;
  if ((int)E.sep.size() < k) E.sep.resize(k);
  if ((rc = E.sepviews.ensure(k))) return rc;
  std::vector<madam::SepView> hv(k);
  int nmax = 2;
  UpBatch ub(st);  // the clears and the views: one copy and one launch
  for (int b = 0; b < k; ++b) {
    SepSlot& sp = E.sep[b];
    if ((rc = sp.ensure(std::max(K[b], 1)))) return rc;
    if ((rc = zero(sp.ctl.p, madam::CTL_N * sizeof(uint32_t), st)) ||
        (rc = zero(sp.hist.p, madam::SEP_NBIN * sizeof(uint32_t), st)) ||
        (rc = zero(sp.maxbits.p, sizeof(unsigned long long), st)))
      return rc;
    madam::SepView& v = hv[b];
    v.n = E.hviews[slots[b]].n;
    v.K = K[b];
    v.thresh = thresh;
    v.X = E.slots[slots[b]].X.p;
    v.ctl = sp.ctl.p;
    v.maxbits = sp.maxbits.p;
    v.hist = sp.hist.p;
    v.okey = sp.okey.p;
    v.bkey = sp.bkey.p;
    v.bval = sp.bval.p;
    v.fkey = sp.fkey.p;
    v.fval = sp.fval.p;
    nmax = std::max(nmax, v.n);
  }
  if ((rc = up(E.sepviews.p, hv.data(), k, st)) || (rc = ub.flush())) return rc;
  TIMED(PH_TRISEP, st, madam::launch_tri_separate(E.sepviews.p, k, nmax, st));
  HIP_TRY(hipGetLastError());
  lists.assign(k, {});
  maxval.assign(k, 0.0);
  // two round trips for the whole batch: the counts, then the lists
  std::vector<uint32_t> ctl((size_t)k * madam::CTL_N);
  std::vector<unsigned long long> mbs(k);
  DownBatch db(st);
  for (int b = 0; b < k; ++b) {
    SepSlot& sp = E.sep[b];
    if ((rc = db.get(ctl.data() + (size_t)b * madam::CTL_N, sp.ctl.p, madam::CTL_N)) ||
        (rc = db.get(&mbs[b], sp.maxbits.p, 1)))
      return rc;
  }
  if ((rc = db.finish())) return rc;
  // a boundary bin too full to rank (many equal or crowded values): the
  // exact refinement for the flagged nodes, then their counts again
  bool over = false;
  for (int b = 0; b < k; ++b) over = over || ctl[(size_t)b * madam::CTL_N + madam::CTL_OVER] == 1;
  if (over) {
    TIMED(PH_TRISEP, st, madam::launch_tri_refine(E.sepviews.p, k, nmax, st));
    HIP_TRY(hipGetLastError());
    for (int b = 0; b < k; ++b)
      if ((rc = db.get(ctl.data() + (size_t)b * madam::CTL_N, E.sep[b].ctl.p, madam::CTL_N)))
        return rc;
    if ((rc = db.finish())) return rc;
    for (int b = 0; b < k; ++b)
      if (ctl[(size_t)b * madam::CTL_N + madam::CTL_OVER] == 1) {
        set_err("triangle separation: boundary refinement did not finish");
        return MADAM_E_HIP;
      }
  }
  std::vector<std::vector<uint32_t>> keys(k);
  std::vector<std::vector<double>> vals(k);
  for (int b = 0; b < k; ++b) {
    SepSlot& sp = E.sep[b];
    const uint32_t nf = ctl[(size_t)b * madam::CTL_N + madam::CTL_NFINAL];
    keys[b].resize(nf);
    vals[b].resize(nf);
    if ((rc = db.get(keys[b].data(), sp.fkey.p, nf)) || (rc = db.get(vals[b].data(), sp.fval.p, nf)))
      return rc;
  }
  if ((rc = db.finish())) return rc;
  for (int b = 0; b < k; ++b) {
    const unsigned long long mb = mbs[b];
    const uint32_t nf = (uint32_t)keys[b].size();
    const unsigned long long bits =
        (mb & 0x8000000000000000ull) ? (mb & 0x7fffffffffffffffull) : ~mb;
    double mx;
    memcpy(&mx, &bits, sizeof mx);
    maxval[b] = mx;
    auto& L = lists[b];
    L.resize(nf);
    for (uint32_t q = 0; q < nf; ++q) {
      const uint32_t key = keys[b][q];
      L[q].type = (int)(key >> 30) + 1;
      L[q].i = (key >> 20) & 1023;
      L[q].j = (key >> 10) & 1023;
      L[q].k = key & 1023;
      L[q].value = vals[b][q];
      L[q].y = 0.0;
    }
  }
  E.timer.flush();
  return MADAM_OK;
}

extern "C" {
int madam_tri_separate(int n, const double* X, int K, double thresh, madam_tri* list, int* size,
                       double* maxval) {
  CutSet cs;
  int rc = one_node(n, 0, nullptr, 0, nullptr, 0, nullptr, cs);
  if (rc) return rc;
  if (K < 0 || !list || !size) {
    set_err("madam_tri_separate: bad arguments");
    return MADAM_E_ARG;
  }
  if ((rc = up(g_eng->slots[0].X.p, X, (size_t)n * n, g_eng->stream))) return rc;
  std::vector<std::vector<madam_tri>> lists;
  std::vector<double> mx;
  if ((rc = tri_separate_slots(std::vector<int>{0}, &K, thresh, lists, mx))) return rc;
  *size = (int)lists[0].size();
  std::copy(lists[0].begin(), lists[0].end(), list);
  if (maxval) *maxval = mx[0];
  return MADAM_OK;
}

// Simulated-annealing runs on one X (host entry for tests and tools).
int madam_qap_sa(int n, const double* X, int k, int nruns, const int* types,
                 const long long* stream_ids, unsigned long long seed, double* cost, int* perms,
                 long long* draws) {
  CutSet cs;
  int rc = one_node(n, 0, nullptr, 0, nullptr, 0, nullptr, cs);
  if (rc) return rc;
  if ((k != 5 && k != 7) || nruns <= 0) {
    set_err("madam_qap_sa: k=%d nruns=%d", k, nruns);
    return MADAM_E_ARG;
  }
  Engine& E = *g_eng;
  const hipStream_t st = E.stream;
  Slot& sl = E.slots[0];
  if ((rc = up(sl.X.p, X, (size_t)n * n, st))) return rc;
  if ((rc = E.sumx.ensure(1)) || (rc = E.xptrs.ensure(1)) || (rc = E.nvals.ensure(1)) ||
      (rc = E.sajobs.ensure(nruns)) || (rc = E.saout.ensure(nruns)))
    return rc;
  const double* xp = sl.X.p;
  if ((rc = up(E.xptrs.p, &xp, 1, st)) || (rc = up(E.nvals.p, &n, 1, st))) return rc;
  madam::launch_sumx(E.xptrs.p, E.nvals.p, E.sumx.p, 1, st);
  std::vector<madam::SaJob> jobs(nruns);
  for (int r = 0; r < nruns; ++r) {
    jobs[r].X = sl.X.p;
    jobs[r].sumX = E.sumx.p;
    jobs[r].n = n;
    jobs[r].k = k;
    jobs[r].neg = types[r] - 1;
    jobs[r].nstride = madam::sa_nstride(n);
    jobs[r].seed = seed;
    jobs[r].stream_id = stream_ids[r];
  }
  if ((rc = up(E.sajobs.p, jobs.data(), nruns, st))) return rc;
  TIMED(PH_SA, st, madam::launch_qap_sa(E.sajobs.p, 1, nruns, n, k, E.saout.p, E.lds_limit, E.cu_count, st));
  HIP_TRY(hipGetLastError());
  std::vector<madam::SaOut> out(nruns);
  if ((rc = down(out.data(), E.saout.p, nruns, st))) return rc;
  HIP_TRY(hipStreamSynchronize(st));
  E.timer.flush();
  for (int r = 0; r < nruns; ++r) {
    cost[r] = out[r].cost;
    for (int a = 0; a < k; ++a) perms[(size_t)r * k + a] = out[r].perm[a];
    if (draws) draws[r] = out[r].draws;
  }
  return MADAM_OK;
}


int madam_mfma_selftest(int np, const float* A32, const float* B32, float* C32, int n,
                        const double* A64, const double* B64, double* C64, double alpha,
                        double beta) {
  int rc = ensure_open();
  if (rc) return rc;
  if (np <= 0 || np % 16 || n <= 0 || n > MADAM_NMAX || np > MADAM_NMAX) {
    set_err("madam_mfma_selftest: bad sizes");
    return MADAM_E_ARG;
  }
  const hipStream_t st = g_eng->stream;
  DBuf<float> a32, b32, c32;
  DBuf<double> a64, b64, c64;
  const size_t p2 = (size_t)np * np, n2 = (size_t)n * n;
  if ((rc = a32.ensure(p2)) || (rc = b32.ensure(p2)) || (rc = c32.ensure(p2)) ||
      (rc = a64.ensure(n2)) || (rc = b64.ensure(n2)) || (rc = c64.ensure(n2)))
    return rc;
  if ((rc = up(a32.p, A32, p2, st)) || (rc = up(b32.p, B32, p2, st)) || (rc = up(a64.p, A64, n2, st)) ||
      (rc = up(b64.p, B64, n2, st)))
    return rc;
  madam::launch_mfma_test(a32.p, b32.p, c32.p, np, a64.p, b64.p, c64.p, n, alpha, beta, st);
  HIP_TRY(hipGetLastError());
  if ((rc = down(C32, c32.p, p2, st)) || (rc = down(C64, c64.p, n2, st))) return rc;
  HIP_TRY(hipStreamSynchronize(st));
  a32.release(); b32.release(); c32.release(); a64.release(); b64.release(); c64.release();
  return MADAM_OK;
}


int madam_ns_schedule(int cap, float* coef) {
  static const float tab[MADAM_NS_STEPS][3] = MADAM_NS_COEF_INIT;
  if (cap < 0 || (cap > 0 && !coef)) {
    set_err("madam_ns_schedule: bad arguments");
    return -MADAM_E_ARG;
  }
  for (int k = 0; k < MADAM_NS_STEPS && k < cap; ++k)
    for (int j = 0; j < 3; ++j) coef[3 * k + j] = tab[k][j];
  return MADAM_NS_STEPS;
}

int madam_project_sdp_ns(int n, const double* W, double* X, double* Q, int* M) {
  CutSet cs;
  int rc = one_node(n, 0, nullptr, 0, nullptr, 0, nullptr, cs);
  if (rc) return rc;
  Slot& sl = g_eng->slots[0];
  if ((rc = sl.size_ns(n))) return rc;
  bind_view(g_eng->hviews[0], sl, n, cs);
  const hipStream_t st = g_eng->stream;
  const size_t nn = (size_t)n * n;
  if ((rc = up(g_eng->views.p, g_eng->hviews.data(), 1, st)) || (rc = up(sl.W.p, W, nn, st))) return rc;
  TIMED(PH_EIG, st, {
    madam::launch_ns_project(g_eng->views.p, 1, n, 1, g_eng->lds_limit, st);
    madam::launch_jacobi(g_eng->views.p, 1, n, 0, 40, 1, g_eng->lds_limit, st, 3);
  });
  madam::launch_project(g_eng->views.p, 1, n, st);
  HIP_TRY(hipGetLastError());
  if ((rc = sync_views_down(1))) return rc;
  if ((rc = down(X, sl.X.p, nn, st)) || (rc = down(Q, sl.Q.p, nn, st))) return rc;
  HIP_TRY(hipStreamSynchronize(st));
  g_eng->timer.flush();
  if (M) *M = g_eng->hviews[0].M;
  return MADAM_OK;
}

int madam_load_tuning(const char* path) {
  std::vector<TuneCell> cells;
  if (path && *path) {
    FILE* f = fopen(path, "r");
    if (!f) {
      set_err("madam_load_tuning: cannot open %s", path);
      return MADAM_E_ARG;
    }
    char line[512];
    int ln = 0;
    while (fgets(line, sizeof line, f)) {
      ++ln;
      const char* p = line;
      while (*p == ' ' || *p == '\t') ++p;
      if (*p == '#' || *p == '\n' || *p == '\r' || *p == '\0') continue;
      TuneCell c;
      if (sscanf(p, "%d %d %d", &c.np, &c.nb, &c.layout) != 3 || c.np < 16 || c.np % 16 != 0 ||
          c.np > ((MADAM_NMAX + 15) & ~15) || c.nb < 1 || c.layout < 1 || c.layout > 3) {
        fclose(f);
        set_err("madam_load_tuning: %s line %d: expected 'n_p nodes layout' (n_p a multiple of "
                "16 up to %d, nodes >= 1, layout 1..3)", path, ln, (MADAM_NMAX + 15) & ~15);
        return MADAM_E_ARG;
      }
      cells.push_back(c);
    }
    fclose(f);
    if (cells.empty()) {
      set_err("madam_load_tuning: %s holds no cells", path);
      return MADAM_E_ARG;
    }
  }
  g_tune.swap(cells);
  return MADAM_OK;
}

int madam_tuned_layout(int n, int active) { return tuned_layout((n + 15) & ~15, active); }

int madam_set_graphs(int on) {
  if (!on) release_chunk_graph();
  g_use_graphs = on ? 1 : 0;
  return MADAM_OK;
}

int madam_graph_stats(long* captures, long* replays) {
  if (captures) *captures = g_chunk.captures;
  if (replays) *replays = g_chunk.replays;
  return MADAM_OK;
}

int madam_set_cg_lds_cap(long bytes, int one_pass) {
  if (bytes < 0) {
    set_err("madam_set_cg_lds_cap: negative cap %ld", bytes);
    return MADAM_E_ARG;
  }
  g_cg_lds_cap = (size_t)bytes;
  g_cg_one_pass = one_pass ? 1 : 0;
  return MADAM_OK;
}

int madam_set_staging(long up_bytes, long down_bytes) {
  if (up_bytes < 0 || down_bytes < 0 || (up_bytes && up_bytes < 4096) ||
      (down_bytes && down_bytes < 4096)) {
    set_err("madam_set_staging: sizes must be 0 or at least 4096 bytes (%ld, %ld)", up_bytes,
            down_bytes);
    return MADAM_E_ARG;
  }
  if (g_ring.batching) {
    set_err("madam_set_staging: a batched upload is open");
    return MADAM_E_ARG;
  }
  if (g_eng) HIP_TRY(hipStreamSynchronize(g_eng->stream));
  g_ring.release();
  g_dring.release();
  g_ring.want = up_bytes ? (size_t)up_bytes : PinnedRing::kBytes;
  g_dring.want = down_bytes ? (size_t)down_bytes : DownRing::kBytes;
  return MADAM_OK;
}

int madam_time_projection(int n, int nb, int layout, int reps, double* ms) {
  int rc = ensure_open();
  if (rc) return rc;
  if (n < 2 || n > MADAM_NMAX || nb < 1 || nb > 1024 || layout < 1 || layout > 3 || reps < 1 ||
      !ms) {
    set_err("madam_time_projection: bad arguments (n=%d nb=%d layout=%d reps=%d)", n, nb, layout,
            reps);
    return MADAM_E_ARG;
  }
  if ((rc = ensure_slots(nb))) return rc;
  Engine& E = *g_eng;
  const hipStream_t st = E.stream;
  CutSet cs;
  if ((rc = build_cutset(n, 0, nullptr, 0, nullptr, 0, nullptr, cs))) return rc;
  for (int b = 0; b < nb; ++b) {
    Slot& sl = E.slots[b];
    if ((rc = sl.size_for(n, cs)) || (rc = sl.size_ns(n))) return rc;
    bind_view(E.hviews[b], sl, n, cs);
  }
  if ((rc = up(E.views.p, E.hviews.data(), nb, st))) return rc;
  madam::launch_ns_synth(E.views.p, nb, n, st);
  auto run = [&]() {
    if (layout == 1)
      madam::launch_ns_project(E.views.p, nb, n, 0, E.lds_limit, st);
    else
      madam::launch_ns_tiles(E.views.p, nb, n, st, layout == 3);
  };
  run();  // warm-up
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st));
  struct Events {  // back to the timer's pool on every return
    PhaseTimer& t;
    hipEvent_t e[2];
    ~Events() {
      t.pool.push_back(e[0]);
      t.pool.push_back(e[1]);
    }
  } evs{E.timer, {E.timer.get(), E.timer.get()}};
  hipEvent_t* ev = evs.e;
  HIP_TRY(hipEventRecord(ev[0], st));
  for (int r = 0; r < reps; ++r) run();
  HIP_TRY(hipEventRecord(ev[1], st));
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventSynchronize(ev[1]));
  float t = 0.f;
  HIP_TRY(hipEventElapsedTime(&t, ev[0], ev[1]));
  *ms = (double)t / reps;
  return MADAM_OK;
}

}  // extern "C"
