// madam_host.hip -- host engine and C ABI of libmadam_bound.so.
//
// The engine owns one HIP stream and a pool of node slots in HBM.  A slot
// holds one node's matrices, cut structure and ADMM vectors (NodeView).
// madam_admm_batch() uploads k jobs into k slots and runs the ADMM loop of
// admm.c:520-704 for all of them with the same kernel launches, checking the
// per-node stop flags on the host every `check_every` iterations; nodes that
// stopped skip the remaining launches inside the kernels.
#include "madam_internal.h"

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

namespace madam {
size_t jacobi_lds_bytes(int nmax);
}

using madam::NodeView;

namespace {

char g_err[512] = "";

void set_err(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      set_err("%s:%d %s: %s", __FILE__, __LINE__, #expr, hipGetErrorString(e_));   \
      return MADAM_E_HIP;                                                          \
    }                                                                              \
  } while (0)

// ---------------------------------------------------------------- cut rows

// Host image of the cut structure of one node (see NodeView).
struct CutSet {
  int m = 0, nact = 0;
  std::vector<int> cptr, cent, cact, aptr, aent, apair, p2a;
  std::vector<double> cscale, y0;
};

// Vertex tuple and sign vector (operators.c:85-148): triangle types 2/3/4
// negate vertex k/j/i; pentagonal/heptagonal type t negates the first t-1.
int tri_expand(const madam_tri& c, int* v, int* b) {
  v[0] = c.i;
  v[1] = c.j;
  v[2] = c.k;
  b[0] = b[1] = b[2] = 1;
  if (c.type == 2) b[2] = -1;
  else if (c.type == 3) b[1] = -1;
  else if (c.type == 4) b[0] = -1;
  return 3;
}
int kgon_expand(int k, int type, const int* perm, int* v, int* b) {
  for (int a = 0; a < k; ++a) {
    v[a] = perm[a];
    b[a] = (a < type - 1) ? -1 : 1;
  }
  return k;
}

int build_cutset(int n, int ntri, const madam_tri* tri, int npent, const madam_pent* pent,
                 int nhept, const madam_hept* hept, CutSet& cs) {
  const int m = ntri + npent + nhept;
  cs.m = m;
  cs.cptr.assign(m + 1, 0);
  cs.cent.clear();
  cs.cscale.resize(m);
  cs.y0.resize(m);
  cs.p2a.assign((size_t)n * n, -1);
  cs.apair.clear();
  std::vector<int> lo_list;
  lo_list.reserve((size_t)m * 21);
  int v[7], b[7];
  for (int c = 0; c < m; ++c) {
    int k;
    double sc, y0;
    if (c < ntri) {
      k = tri_expand(tri[c], v, b);
      sc = 1.0;
      y0 = tri[c].y;
    } else if (c < ntri + npent) {
      const madam_pent& P = pent[c - ntri];
      k = kgon_expand(5, P.type, P.perm, v, b);
      sc = 0.5;
      y0 = P.y;
    } else {
      const madam_hept& H = hept[c - ntri - npent];
      k = kgon_expand(7, H.type, H.perm, v, b);
      sc = 1.0 / 3.0;
      y0 = H.y;
    }
    for (int a = 0; a < k; ++a)
      if (v[a] < 0 || v[a] >= n) {
        set_err("cut %d has vertex %d outside 0..%d", c, v[a], n - 1);
        return MADAM_E_ARG;
      }
    cs.cscale[c] = sc;
    cs.y0[c] = y0;
    for (int a = 0; a < k; ++a)
      for (int q = a + 1; q < k; ++q) {
        const int hi = std::max(v[a], v[q]), lw = std::min(v[a], v[q]);
        if (hi == lw) {
          set_err("cut %d repeats vertex %d", c, hi);
          return MADAM_E_ARG;
        }
        const int lo = hi + n * lw;
        const int sgn = -(b[a] * b[q]);
        cs.cent.push_back(sgn > 0 ? lo : ~lo);
        lo_list.push_back(lo);
        if (cs.p2a[lo] < 0) {
          cs.p2a[lo] = (int)cs.apair.size();
          cs.apair.push_back(lo);
        }
      }
    cs.cptr[c + 1] = (int)cs.cent.size();
  }
  const int nact = (int)cs.apair.size();
  cs.nact = nact;
  cs.cact.resize(cs.cent.size());
  cs.aptr.assign(nact + 1, 0);
  for (size_t e = 0; e < cs.cent.size(); ++e) {
    const int a = cs.p2a[lo_list[e]];
    cs.cact[e] = a;
    cs.aptr[a + 1]++;
  }
  for (int a = 0; a < nact; ++a) cs.aptr[a + 1] += cs.aptr[a];
  cs.aent.assign(cs.cent.size(), 0);
  std::vector<int> fill(cs.aptr.begin(), cs.aptr.end() - 1);
  for (int c = 0; c < m; ++c)
    for (int e = cs.cptr[c]; e < cs.cptr[c + 1]; ++e) {
      const int a = cs.cact[e];
      cs.aent[fill[a]++] = cs.cent[e] >= 0 ? c : ~c;
    }
  return MADAM_OK;
}

// ---------------------------------------------------------------- engine

template <class T>
struct DBuf {
  T* p = nullptr;
  size_t cap = 0;
  int ensure(size_t n) {
    if (n <= cap) return MADAM_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    size_t want = std::max<size_t>(n, 64);
    HIP_TRY(hipMalloc(&p, want * sizeof(T)));
    cap = want;
    return MADAM_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

struct Slot {
  DBuf<double> L, X, Q, W, V, S1, S2, ev, evp, y;
  DBuf<int> cptr, cent, cact, aptr, aent, apair, p2a;
  DBuf<double> cscale, t, s, u, rhs, r, p, q, w;

  int size_for(int n, const CutSet& cs) {
    const size_t nn = (size_t)n * n;
    int rc = 0;
    for (DBuf<double>* b : {&L, &X, &Q, &W, &V, &S1, &S2})
      if ((rc = b->ensure(nn))) return rc;
    if ((rc = ev.ensure(n)) || (rc = evp.ensure(n)) || (rc = y.ensure(n))) return rc;
    const size_t m = std::max(cs.m, 1), ne = std::max<size_t>(cs.cent.size(), 1),
                 na = std::max(cs.nact, 1);
    if ((rc = cptr.ensure(m + 1)) || (rc = cent.ensure(ne)) || (rc = cact.ensure(ne)) ||
        (rc = aptr.ensure(na + 1)) || (rc = aent.ensure(ne)) || (rc = apair.ensure(na)) ||
        (rc = p2a.ensure(nn)))
      return rc;
    for (DBuf<double>* b : {&cscale, &t, &s, &u, &rhs, &r, &p, &q})
      if ((rc = b->ensure(m))) return rc;
    return w.ensure(na);
  }
  void release() {
    for (DBuf<double>* b : {&L, &X, &Q, &W, &V, &S1, &S2, &ev, &evp, &y, &cscale, &t, &s, &u, &rhs,
                            &r, &p, &q, &w})
      b->release();
    for (DBuf<int>* b : {&cptr, &cent, &cact, &aptr, &aent, &apair, &p2a}) b->release();
  }
};

// Optional per-phase GPU timing with HIP events on the engine's stream.
enum Phase { PH_PREP, PH_CG, PH_W, PH_EIG, PH_POST, PH_SAFE, PH_IPM, PH_TRISEP, PH_SA, PH_GW,
             PH_COUNT };

struct PhaseTimer {
  bool on = false;
  std::vector<hipEvent_t> pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
  double ms[PH_COUNT] = {0};
  long count[PH_COUNT] = {0};
  hipEvent_t get() {
    if (pool.empty()) {
      hipEvent_t e;
      (void)hipEventCreate(&e);
      return e;
    }
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  // resolve recorded pairs; the stream must be synchronised by the caller
  void flush() {
    for (auto& pr : pending) {
      float t = 0.f;
      (void)hipEventElapsedTime(&t, pr.second.first, pr.second.second);
      ms[pr.first] += t;
      count[pr.first] += 1;
      pool.push_back(pr.second.first);
      pool.push_back(pr.second.second);
    }
    pending.clear();
  }
};

struct Engine {
  PhaseTimer timer;
  int device = -1;
  hipStream_t stream = nullptr;
  size_t lds_limit = 0;
  std::vector<Slot> slots;
  DBuf<NodeView> views;
  DBuf<double> tmpd;
  std::vector<NodeView> hviews;
};

Engine* g_eng = nullptr;

// Run `call` (kernel launches on stream st) bracketed by events when timing.
#define TIMED(phase, st, call)                                          \
  do {                                                                  \
    if (g_eng->timer.on) {                                              \
      hipEvent_t a_ = g_eng->timer.get(), b_ = g_eng->timer.get();      \
      (void)hipEventRecord(a_, st);                                     \
      call;                                                             \
      (void)hipEventRecord(b_, st);                                     \
      g_eng->timer.pending.push_back({phase, {a_, b_}});                \
    } else {                                                            \
      call;                                                             \
    }                                                                   \
  } while (0)

int ensure_open() {
  if (g_eng) return MADAM_OK;
  return madam_open(0);
}

madam_opts default_opts() {
  madam_opts o;
  o.cg_rtol = 1e-12;
  o.cg_maxit = 1000;
  o.jacobi_sweeps = 40;
  o.check_every = 16;
  return o;
}

template <class T>
int up(T* dst, const T* src, size_t n, hipStream_t s) {
  if (n == 0) return MADAM_OK;
  HIP_TRY(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, s));
  return MADAM_OK;
}
template <class T>
int down(T* dst, const T* src, size_t n, hipStream_t s) {
  if (n == 0) return MADAM_OK;
  HIP_TRY(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, s));
  return MADAM_OK;
}

// Fill view v from slot sl (device pointers) for a node of order n.
void bind_view(NodeView& v, Slot& sl, int n, const CutSet& cs) {
  memset(&v, 0, sizeof v);
  v.n = n;
  v.m = cs.m;
  v.nact = cs.nact;
  v.L = sl.L.p;
  v.X = sl.X.p;
  v.Q = sl.Q.p;
  v.W = sl.W.p;
  v.V = sl.V.p;
  v.S1 = sl.S1.p;
  v.S2 = sl.S2.p;
  v.ev = sl.ev.p;
  v.evp = sl.evp.p;
  v.y = sl.y.p;
  v.cptr = sl.cptr.p;
  v.cent = sl.cent.p;
  v.cact = sl.cact.p;
  v.cscale = sl.cscale.p;
  v.aptr = sl.aptr.p;
  v.aent = sl.aent.p;
  v.apair = sl.apair.p;
  v.p2a = sl.p2a.p;
  v.t = sl.t.p;
  v.s = sl.s.p;
  v.u = sl.u.p;
  v.rhs = sl.rhs.p;
  v.cg_r = sl.r.p;
  v.cg_p = sl.p.p;
  v.cg_q = sl.q.p;
  v.cg_w = sl.w.p;
}

int upload_cuts(Slot& sl, const CutSet& cs, hipStream_t s) {
  int rc;
  if ((rc = up(sl.cptr.p, cs.cptr.data(), cs.cptr.size(), s)) ||
      (rc = up(sl.cent.p, cs.cent.data(), cs.cent.size(), s)) ||
      (rc = up(sl.cact.p, cs.cact.data(), cs.cact.size(), s)) ||
      (rc = up(sl.aptr.p, cs.aptr.data(), cs.aptr.size(), s)) ||
      (rc = up(sl.aent.p, cs.aent.data(), cs.aent.size(), s)) ||
      (rc = up(sl.apair.p, cs.apair.data(), cs.apair.size(), s)) ||
      (rc = up(sl.p2a.p, cs.p2a.data(), cs.p2a.size(), s)) ||
      (rc = up(sl.cscale.p, cs.cscale.data(), cs.cscale.size(), s)))
    return rc;
  return MADAM_OK;
}

int ensure_slots(int k) {
  if ((int)g_eng->slots.size() < k) g_eng->slots.resize(k);
  g_eng->hviews.resize(std::max<size_t>(g_eng->hviews.size(), k));
  return g_eng->views.ensure(k);
}

int sync_views_down(int k) {
  int rc = down(g_eng->hviews.data(), g_eng->views.p, k, g_eng->stream);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(g_eng->stream));
  return MADAM_OK;
}

}  // namespace

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
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  Engine* e = new Engine();
  e->device = device;
  e->lds_limit = prop.sharedMemPerBlock;
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
    delete e;
    set_err("hipStreamCreate failed");
    return MADAM_E_HIP;
  }
  g_eng = e;
  return MADAM_OK;
}

void madam_close(void) {
  if (!g_eng) return;
  (void)hipStreamSynchronize(g_eng->stream);
  for (auto& s : g_eng->slots) s.release();
  g_eng->views.release();
  g_eng->tmpd.release();
  (void)hipStreamDestroy(g_eng->stream);
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
    double fro = 0.0;
    for (size_t e = 0; e < nn; ++e) fro += J.L[e] * J.L[e];
    v.normL = sqrt(fro);
    v.sigma = *J.sigma;
    v.tol = J.tol;
    v.scaleSigma = J.scaleSigma;
    v.itermax = J.itermax;
  }
  if ((rc = up(E.views.p, E.hviews.data(), k, st))) return rc;

  // ---- the ADMM loop (admm.c:520-704), all nodes together
  const int chunk = std::max(1, o.check_every);
  int itmax_all = 0;
  for (int b = 0; b < k; ++b) itmax_all = std::max(itmax_all, jobs[b].itermax);
  long launched = 0;
  for (;;) {
    if (launched > (long)itmax_all + 2 * chunk + 8) {
      set_err("ADMM stop flags never set after %ld iterations", launched);
      return MADAM_E_HIP;
    }
    launched += chunk;
    for (int it = 0; it < chunk; ++it) {
      TIMED(PH_PREP, st, madam::launch_admm_prep(E.views.p, k, mmax, st));
      TIMED(PH_CG, st, madam::launch_cg(E.views.p, k, o.cg_rtol, o.cg_maxit, st));
      TIMED(PH_W, st, madam::launch_build_w(E.views.p, k, nmax, mmax, 0, st));
      TIMED(PH_EIG, st,
            madam::launch_jacobi(E.views.p, k, nmax, 0, o.jacobi_sweeps, 0, E.lds_limit, st));
      TIMED(PH_POST, st, madam::launch_admm_post(E.views.p, k, nmax, st));
    }
    HIP_TRY(hipGetLastError());
    if ((rc = sync_views_down(k))) return rc;
    E.timer.flush();
    bool all = true;
    for (int b = 0; b < k; ++b) all = all && E.hviews[b].done;
    if (all) break;
  }
  // ---- safe bound (admm.c:724-763)
  TIMED(PH_SAFE, st, {
    madam::launch_build_w(E.views.p, k, nmax, mmax, 1, st);
    madam::launch_jacobi(E.views.p, k, nmax, 1, 4 * o.jacobi_sweeps, 1, E.lds_limit, st);
    madam::launch_safe_bound(E.views.p, k, st);
  });
  HIP_TRY(hipGetLastError());
  if ((rc = sync_views_down(k))) return rc;
  E.timer.flush();

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
      // project_SDP and admm.c:583-585 leave it.
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
  madam::launch_jacobi(g_eng->views.p, 1, n, 0, 60, 1, g_eng->lds_limit, st);
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

int madam_lambda_max(int n, const double* A, double* lam) {
  CutSet cs;
  int rc = one_node(n, 0, nullptr, 0, nullptr, 0, nullptr, cs);
  if (rc) return rc;
  Slot& sl = g_eng->slots[0];
  const hipStream_t st = g_eng->stream;
  if ((rc = up(sl.W.p, A, (size_t)n * n, st))) return rc;
  // safe bound with y = 0 and no cuts: f = n * lambda_max
  HIP_TRY(hipMemsetAsync(sl.y.p, 0, n * sizeof(double), st));
  madam::launch_jacobi(g_eng->views.p, 1, n, 1, 160, 1, g_eng->lds_limit, st);
  madam::launch_safe_bound(g_eng->views.p, 1, st);
  HIP_TRY(hipGetLastError());
  if ((rc = sync_views_down(1))) return rc;
  *lam = g_eng->hviews[0].lam;
  return MADAM_OK;
}

}  // extern "C"
