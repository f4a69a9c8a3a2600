// madam_engine.h -- host engine internals of libmadam_bound.so: device
// buffers, node slots, the per-phase timer and the cut-structure build.
// Shared by the C ABI (madam_host.hip) and the bounding driver
// (madam_bab.hip).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "../../include/madam_bound.h"
#include "madam_gw.h"
#include "madam_internal.h"
#include "madam_sep.h"

namespace mh {

using madam::NodeView;


extern char g_err[512];

inline void set_err(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

// |L|_F of an n x n column-major matrix in the library's one summation
// order: each column's squares summed top to bottom, then the n column sums
// left to right.  k_subproblem sums in the same order (one thread per
// column), so a node's norm has the same bits whichever side built it.
inline double frob_norm_cols(const double* L, int n) {
  double fro = 0.0;
  for (int j = 0; j < n; ++j) {
    const double* c = L + (size_t)j * n;
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += c[i] * c[i];
    fro += s;
  }
  return sqrt(fro);
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
  std::vector<short> aent16, cact16;  // the same streams in 16 bits when m, nact < 2^15
  std::vector<double> cscale, y0;
  // sliced Gram rows of the one-pass cut solve (see build_xrows); empty
  // when not built
  std::vector<int> xoff, tcut, tmeta;
  std::vector<double> tsc, tdinv;
  std::vector<unsigned> xent;
};

// Rows of the Gram matrix BB^T for the one-pass product of the cut solve,
// sliced for coalesced loads.  (BB^T)_cc' = scale_c scale_c' G_cc' with
// G_cc' = sum over the pairs shared by c and c' of the product of their
// signs there, G_cc = k_c (the pair entries of c).  Row c lists, for each
// pair entry of c in entry order and each other cut c' of that pair in
// increasing cut order, one word: the LDS byte offset of entry 2 c' + neg
// of a table holding +-scale_c' u_c' (neg when the two signs differ), so
// that the product is q_c = v_c + scale_c / 2 (k_c ps_c + sum of the row's
// table entries) with no sign or pair logic per word.  The cuts are sorted
// by row length, longest first (ties by cut), into slots: slot i holds cut
// tcut[i] and belongs to thread i % XCUT_THREADS of the solve.  Slice s =
// slots 64 s .. 64 s + 63 (one wave's lanes at one step) stores word q of
// lane l's row at xoff[s] + 64 q + l, its rows padded to the longest
// rounded up to 8 with the offset 16 m of the table's zero entry.  tmeta[i]
// = pair entries of the cut, tsc[i] its scale and tdinv[i] the diagonal
// preconditioner 1 / (1 + k scale^2 / 2).  Not built (empty) when m >= 2^16
// or the sliced rows exceed 16 words per pair entry plus 4096.
inline void build_xrows(CutSet& cs) {
  cs.xoff.clear();
  cs.xent.clear();
  cs.tcut.clear();
  cs.tmeta.clear();
  cs.tsc.clear();
  cs.tdinv.clear();
  const int m = cs.m;
  if (m == 0 || m >= 65536) return;
  const size_t ne = cs.cent.size();
  std::vector<int> len(m, 0);
  for (int c = 0; c < m; ++c)
    for (int e = cs.cptr[c]; e < cs.cptr[c + 1]; ++e) {
      const int a = cs.cact[e] >= 0 ? cs.cact[e] : ~cs.cact[e];
      len[c] += cs.aptr[a + 1] - cs.aptr[a] - 1;
    }
  std::vector<int> order(m);
  for (int c = 0; c < m; ++c) order[c] = c;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return len[a] > len[b]; });
  const int ns = (m + 63) / 64;
  std::vector<int> off(ns + 1, 0);
  for (int sl = 0; sl < ns; ++sl) {
    off[sl + 1] = off[sl] + 64 * ((len[order[64 * sl]] + 7) & ~7);
    if ((size_t)off[sl + 1] > 16 * ne + 4096) return;
  }
  cs.xoff.swap(off);
  const unsigned zero = 16u * (unsigned)m;
  cs.xent.assign(cs.xoff[ns], zero);
  cs.tcut = order;
  cs.tmeta.resize(m);
  cs.tsc.resize(m);
  cs.tdinv.resize(m);
  for (int i = 0; i < m; ++i) {
    const int c = order[i];
    const int k = cs.cptr[c + 1] - cs.cptr[c];
    const double sc = cs.cscale[c];
    cs.tmeta[i] = k;
    cs.tsc[i] = sc;
    cs.tdinv[i] = 1.0 / (1.0 + 0.5 * k * sc * sc);
    size_t f = (size_t)cs.xoff[i / 64] + (i % 64);
    for (int e = cs.cptr[c]; e < cs.cptr[c + 1]; ++e) {
      const int ca = cs.cact[e];
      const int a = ca >= 0 ? ca : ~ca;
      for (int g = cs.aptr[a]; g < cs.aptr[a + 1]; ++g) {
        const int en = cs.aent[g];
        const int c2 = en >= 0 ? en : ~en;
        if (c2 == c) continue;
        const unsigned neg = (en < 0) != (ca < 0) ? 1u : 0u;
        cs.xent[f] = 8u * (2u * (unsigned)c2 + neg);
        f += 64;
      }
    }
  }
}

// Vertex tuple and sign vector (operators.c:85-148): triangle types 2/3/4
// negate vertex k/j/i; pentagonal/heptagonal type t negates the first t-1.
inline int tri_expand(const madam_tri& c, int* v, int* b) {
  v[0] = c.i;
  v[1] = c.j;
  v[2] = c.k;
  b[0] = b[1] = b[2] = 1;
  if (c.type == 2) b[2] = -1;
  else if (c.type == 3) b[1] = -1;
  else if (c.type == 4) b[0] = -1;
  return 3;
}
inline int kgon_expand(int k, int type, const int* perm, int* v, int* b) {
  for (int a = 0; a < k; ++a) {
    v[a] = perm[a];
    b[a] = (a < type - 1) ? -1 : 1;
  }
  return k;
}

// The checks build_cutset reports as errors, without touching shared
// state (callable from several threads): every vertex in 0..n-1, no vertex
// twice in one cut.
inline bool cuts_valid(int n, const std::vector<madam_tri>& tri, const std::vector<madam_pent>& pent,
                       const std::vector<madam_hept>& hept) {
  int v[7], b[7];
  auto ok = [&](int k) {
    for (int a = 0; a < k; ++a) {
      if (v[a] < 0 || v[a] >= n) return false;
      for (int q = 0; q < a; ++q)
        if (v[q] == v[a]) return false;
    }
    return true;
  };
  for (const auto& c : tri)
    if (!ok(tri_expand(c, v, b))) return false;
  for (const auto& c : pent)
    if (!ok(kgon_expand(5, c.type, c.perm, v, b))) return false;
  for (const auto& c : hept)
    if (!ok(kgon_expand(7, c.type, c.perm, v, b))) return false;
  return true;
}

inline int build_cutset(int n, int ntri, const madam_tri* tri, int npent, const madam_pent* pent,
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
  // the entry's sign travels with its active-pair id (a or ~a), so that the
  // cut-side pass of the BB^T product reads one index stream
  for (size_t e = 0; e < cs.cent.size(); ++e)
    if (cs.cent[e] < 0) cs.cact[e] = ~cs.cact[e];
  cs.aent16.clear();
  cs.cact16.clear();
  if (m < 32768 && nact < 32768) {
    cs.aent16.assign(cs.aent.begin(), cs.aent.end());
    cs.cact16.assign(cs.cact.begin(), cs.cact.end());
  }
  build_xrows(cs);
  return MADAM_OK;
}

// f(i) for i in [0, n) on up to 16 host threads; the
// per-node host work of a batch (cut structures, separation lists) is
// independent across nodes.  f must not touch shared state.
template <class F>
inline void parallel_for(int n, F f) {
  // up to 16 threads (one GPU's share of the host), items handed out one at
  // a time: per-node host work varies with the node's cut counts
  const unsigned hw = std::thread::hardware_concurrency();
  const int T = std::min<int>(std::min<unsigned>(hw ? hw : 1, 16u), n / 4);
  if (T <= 1) {
    for (int i = 0; i < n; ++i) f(i);
    return;
  }
  std::atomic<int> next(0);
  std::vector<std::thread> pool;
  pool.reserve(T - 1);
  auto run = [&]() {
    for (int i = next.fetch_add(1); i < n; i = next.fetch_add(1)) f(i);
  };
  for (int t = 1; t < T; ++t) pool.emplace_back(run);
  run();
  for (auto& th : pool) th.join();
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
  DBuf<int> cptr, cent, cact, aptr, aent, apair, p2a, xoff, tcut, tmeta;
  DBuf<short> aent16, cact16;
  DBuf<unsigned> xent;
  DBuf<double> cscale, t, s, u, rhs, r, p, q, w, tprev, tprev2, tsc, tdinv, xs;
  DBuf<double> ip1, ip2, ip3, ipv, ipP;  // IPM scratch
  DBuf<float> ns32;                      // sign-iteration scratch

  int size_ns(int n) {
    const size_t np = (size_t)((n + 15) & ~15);
    int rc;
    // the three FP32 iterates, then 16 doubles of the tile layout's norm sums
    if ((rc = ns32.ensure(3 * np * np + 32)) || (rc = ip1.ensure((size_t)n * n))) return rc;
    return MADAM_OK;
  }
  DBuf<double> gwval, gwbest;            // GW per-hyperplane values, best value
  DBuf<signed char> gwz, gwbz;           // GW cuts, best cut
  int n_ipm = 0;

  int size_ipm(int n) {
    const size_t nn = (size_t)n * n;
    int rc;
    if ((rc = ip1.ensure(nn)) || (rc = ip2.ensure(nn)) || (rc = ip3.ensure(nn)) ||
        (rc = ipv.ensure(4 * (size_t)n)) || (rc = ipP.ensure(nn)))
      return rc;
    return MADAM_OK;
  }
  int size_gw(int n, int num) {
    int rc;
    if ((rc = gwval.ensure(num)) || (rc = gwbest.ensure(1)) || (rc = gwz.ensure((size_t)num * n)) ||
        (rc = gwbz.ensure(n)))
      return rc;
    return MADAM_OK;
  }

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
        (rc = aent16.ensure(ne)) || (rc = cact16.ensure(ne)) ||
        (rc = p2a.ensure(nn)))
      return rc;
    if (!cs.tcut.empty() &&
        ((rc = xoff.ensure(cs.xoff.size())) || (rc = xent.ensure(cs.xent.size())) ||
         (rc = tcut.ensure(m)) || (rc = tmeta.ensure(m)) || (rc = tsc.ensure(m)) ||
         (rc = tdinv.ensure(m)) || (rc = xs.ensure(m))))
      return rc;
    for (DBuf<double>* b : {&cscale, &t, &s, &u, &rhs, &r, &p, &q, &tprev, &tprev2})
      if ((rc = b->ensure(m))) return rc;
    // pair sums of the two-pass product, or the cut-side A u of the one-pass one
    return w.ensure(std::max(na, m));
  }
  void release() {
    for (DBuf<double>* b : {&L, &X, &Q, &W, &V, &S1, &S2, &ev, &evp, &y, &cscale, &t, &s, &u, &rhs,
                            &r, &p, &q, &w, &tprev, &tprev2})
      b->release();
    for (DBuf<int>* b : {&cptr, &cent, &cact, &aptr, &aent, &apair, &p2a, &xoff, &tcut, &tmeta})
      b->release();
    xent.release();
    tsc.release();
    tdinv.release();
    xs.release();
    aent16.release();
    cact16.release();
    for (DBuf<double>* b : {&ip1, &ip2, &ip3, &ipv, &ipP, &gwval, &gwbest}) b->release();
    gwz.release();
    gwbz.release();
    ns32.release();
  }
};

// Optional per-phase GPU timing with HIP events on the engine's stream.
enum Phase { PH_PREP, PH_CG, PH_W, PH_EIG, PH_POST, PH_SAFE, PH_IPM, PH_TRISEP, PH_SA, PH_GW,
             PH_ZFAC, PH_COUNT };

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

// Per-slot separation buffers.
struct SepSlot {
  DBuf<uint32_t> ctl, hist, okey, bkey, fkey;
  DBuf<unsigned long long> maxbits;
  DBuf<double> bval, fval;
  int ensure(int K) {
    int rc;
    if ((rc = ctl.ensure(madam::CTL_N)) || (rc = hist.ensure(madam::SEP_NBIN)) ||
        (rc = okey.ensure(K)) || (rc = bkey.ensure(madam::SEP_BCAP)) || (rc = fkey.ensure(K)) ||
        (rc = maxbits.ensure(1)) || (rc = bval.ensure(madam::SEP_BCAP)) || (rc = fval.ensure(K)))
      return rc;
    return MADAM_OK;
  }
  void release() {
    ctl.release(); hist.release(); okey.release(); bkey.release(); fkey.release();
    maxbits.release(); bval.release(); fval.release();
  }
};

struct Engine {
  std::vector<SepSlot> sep;
  DBuf<madam::SepView> sepviews;
  DBuf<madam::SaJob> sajobs, sajobs2;  // (2: a second annealing launch in flight)
  DBuf<madam::GwNode> gwnodes;
  DBuf<madam::SaOut> saout, saout2;
  DBuf<double> sumx, sumx2;
  DBuf<const double*> xptrs, xptrs2;
  DBuf<int> nvals, nvals2;
  PhaseTimer timer;
  int device = -1;
  hipStream_t stream = nullptr;
  // a second stream (default priority) for work that runs beside the
  // main stream's, ordered against it by the two events
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  size_t lds_limit = 0;
  int cu_count = 256;
  std::vector<Slot> slots;
  DBuf<NodeView> views;
  DBuf<NodeView> fviews;  // nodes whose ADMM call is being finished
  DBuf<double> tmpd;
  DBuf<double> postpart;  // k_admm_post partial sums: nodes x 64 workgroups x 8
  std::vector<NodeView> hviews;
};

extern Engine* g_eng;

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

inline int ensure_open() {
  if (g_eng) return MADAM_OK;
  return madam_open(0);
}

inline madam_opts default_opts() {
  madam_opts o;
  o.cg_rtol = 1e-12;
  o.cg_maxit = 1000;
  o.jacobi_sweeps = 40;
  o.check_every = 16;
  o.projection = 0;
  o.ns_layout = 0;
  o.cg_adapt = 0.0;
  return o;
}

// Host-to-device uploads go through a pinned ring buffer: the source is
// copied into it at once (so the caller may reuse it, as with a pageable
// copy) and the DMA runs asynchronously on the stream.  When the ring is
// full the stream is synchronised before its start is reused; uploads of
// more than a quarter of the ring go directly from pageable memory.
//
// Batched transfers.  The host work of a batch of nodes moves tens of
// small arrays per node (cut structures, duals, flags, separation lists);
// one copy each costs a dispatch on the stream, and a download into
// pageable memory blocks the host until it lands.  Inside an UpBatch scope
// up() and zero() only stage their data in the ring and record the piece;
// the scope's flush moves the staged bytes in one DMA to a device mirror of
// the ring and one k_xfer launch scatters them to their buffers.  Work that
// reads those buffers must be issued after the flush.  A DownBatch gathers
// its pieces on the device into a staging buffer (one k_xfer launch),
// copies that to pinned memory in one DMA, synchronises the stream and
// hands the pieces out on the host.
struct PinnedRing {
  char* p = nullptr;
  char* dev = nullptr;  // device mirror (batched uploads)
  size_t cap = 0, head = 0;
  int batching = 0;
  size_t base = 0;                    // start of the staged, not yet copied bytes
  std::vector<madam::XDesc> pend;     // src: ring offset (or null: zero fill)
  std::vector<madam::P2aJob> jobs;    // pair maps to build after the scatter
  // copies into the ring left for the flush (UpBatch with deferred copies:
  // the sources live until then), made on host threads together
  struct HostCopy {
    char* dst;
    const void* src;
    size_t bytes;
  };
  int defer_copies = 0;
  std::vector<HostCopy> hcopies;
  size_t want = kBytes;  // capacity of the next allocation
  static constexpr size_t kBytes = 64u << 20;
  static constexpr size_t kZero = ~(size_t)0;
  void release() {
    if (p) (void)hipHostFree(p);
    if (dev) (void)hipFree(dev);
    p = dev = nullptr;
    cap = head = base = 0;
    batching = 0;
    defer_copies = 0;
    pend.clear();
    jobs.clear();
    hcopies.clear();
  }
};
inline PinnedRing g_ring;

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

inline void ring_ensure() {
  PinnedRing& R = g_ring;
  if (!R.p && hipHostMalloc((void**)&R.p, R.want, hipHostMallocDefault) == hipSuccess)
    R.cap = R.want;
}

// Ring bytes the open batch needs after its data: the descriptors of its
// pieces and its pair-map jobs, with one more of each.
inline size_t ring_tail(const PinnedRing& R) {
  return 512 + (R.pend.size() + 1) * sizeof(madam::XDesc) +
         (R.jobs.size() + 1) * sizeof(madam::P2aJob);
}

// The staged pieces of the open batch: one DMA of [base, end) to the device
// mirror (the descriptors and pair-map jobs are staged after the data), one
// scatter launch, then one launch building the pair maps.
inline int flush_up(hipStream_t s) {
  PinnedRing& R = g_ring;
  if (!R.hcopies.empty()) {
    parallel_for((int)R.hcopies.size(), [&](int i) {
      const PinnedRing::HostCopy& h = R.hcopies[i];
      memcpy(h.dst, h.src, h.bytes);
    });
    R.hcopies.clear();
  }
  if (R.pend.empty() && R.jobs.empty()) return MADAM_OK;
  if (!R.dev) HIP_TRY(hipMalloc((void**)&R.dev, R.cap));
  const size_t doff = align256(R.head);
  const size_t nd = R.pend.size(), nj = R.jobs.size();
  const size_t joff = align256(doff + nd * sizeof(madam::XDesc));
  const size_t end = joff + nj * sizeof(madam::P2aJob);
  if (end > R.cap) {  // ring_room keeps this room
    set_err("upload ring: %zu descriptors and %zu jobs do not fit after offset %zu", nd, nj, doff);
    return MADAM_E_HIP;
  }
  madam::XDesc* hd = (madam::XDesc*)(R.p + doff);
  for (size_t i = 0; i < nd; ++i) {
    const size_t o = (size_t)(uintptr_t)R.pend[i].src;
    hd[i].src = o == PinnedRing::kZero ? nullptr : R.dev + o;
    hd[i].dst = R.pend[i].dst;
    hd[i].bytes = R.pend[i].bytes;
  }
  if (nj) memcpy(R.p + joff, R.jobs.data(), nj * sizeof(madam::P2aJob));
  HIP_TRY(hipMemcpyAsync(R.dev + R.base, R.p + R.base, end - R.base, hipMemcpyHostToDevice, s));
  madam::launch_xfer((const madam::XDesc*)(R.dev + doff), (int)nd, s);
  madam::launch_p2a((const madam::P2aJob*)(R.dev + joff), (int)nj, s);
  HIP_TRY(hipGetLastError());
  R.pend.clear();
  R.jobs.clear();
  R.head = R.base = end;
  return MADAM_OK;
}

// Make room in the ring for `bytes` of data at align256(head) plus what
// the open batch needs after it; returns the offset.
inline int ring_room(size_t bytes, hipStream_t s, size_t* off) {
  PinnedRing& R = g_ring;
  size_t o = align256(R.head);
  const size_t desc = R.batching ? ring_tail(R) : 0;
  if (o + bytes + desc > R.cap) {
    int rc;
    if ((rc = flush_up(s))) return rc;
    HIP_TRY(hipStreamSynchronize(s));  // every copy out of the ring has finished
    o = 0;
    R.head = R.base = 0;
  }
  *off = o;
  return MADAM_OK;
}

inline int up_bytes(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes == 0) return MADAM_OK;
  PinnedRing& R = g_ring;
  ring_ensure();
  if (!R.p || bytes > R.cap / 4) {
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    return MADAM_OK;
  }
  size_t off;
  int rc;
  if ((rc = ring_room(bytes, s, &off))) return rc;
  if (R.batching && R.defer_copies)
    R.hcopies.push_back({R.p + off, src, bytes});
  else
    memcpy(R.p + off, src, bytes);
  R.head = off + bytes;
  if (R.batching)
    R.pend.push_back({(const char*)(uintptr_t)off, (char*)dst, (unsigned long long)bytes});
  else
    HIP_TRY(hipMemcpyAsync(dst, R.p + off, bytes, hipMemcpyHostToDevice, s));
  return MADAM_OK;
}

template <class T>
inline int up(T* dst, const T* src, size_t n, hipStream_t s) {
  return up_bytes((void*)dst, (const void*)src, n * sizeof(T), s);
}

// dst[0 .. bytes) = 0 on the stream (a piece of the open batch, if any)
inline int zero(void* dst, size_t bytes, hipStream_t s) {
  if (bytes == 0) return MADAM_OK;
  PinnedRing& R = g_ring;
  if (!R.batching) {
    HIP_TRY(hipMemsetAsync(dst, 0, bytes, s));
    return MADAM_OK;
  }
  size_t off;
  int rc;
  if ((rc = ring_room(0, s, &off))) return rc;
  R.pend.push_back({(const char*)(uintptr_t)PinnedRing::kZero, (char*)dst,
                    (unsigned long long)bytes});
  return MADAM_OK;
}

// defer: the sources of up() stay valid until the flush, which copies
// them into the ring on host threads (else each up() copies at once)
struct UpBatch {
  hipStream_t s;
  bool on = false;
  explicit UpBatch(hipStream_t st, bool defer = false) : s(st) {
    ring_ensure();
    PinnedRing& R = g_ring;
    if (R.p && !R.batching) {  // (an inner scope joins the outer batch)
      R.batching = 1;
      R.defer_copies = defer ? 1 : 0;
      R.base = R.head;
      on = true;
    }
  }
  int flush() {
    if (!on) return MADAM_OK;
    on = false;
    const int rc = flush_up(s);
    g_ring.batching = 0;
    g_ring.defer_copies = 0;
    g_ring.hcopies.clear();
    return rc;
  }
  ~UpBatch() { (void)flush(); }
  UpBatch(const UpBatch&) = delete;
  UpBatch& operator=(const UpBatch&) = delete;
};

// A cut structure's pair map (n^2 ints, -1 but at the active pairs), built
// on the device from the uploaded apair after the batch's scatter instead of
// being uploaded (host_p2a: the host's copy, uploaded only when there is no
// pinned ring).
inline int p2a_job(int* p2a, const int* apair, size_t nn, int nact, const int* host_p2a,
                   hipStream_t s) {
  PinnedRing& R = g_ring;
  ring_ensure();
  if (!R.p) return up(p2a, host_p2a, nn, s);
  if (!R.batching) {
    UpBatch ub(s);
    int rc = p2a_job(p2a, apair, nn, nact, host_p2a, s);
    return rc ? rc : ub.flush();
  }
  size_t off;
  int rc;
  if ((rc = ring_room(0, s, &off))) return rc;
  R.jobs.push_back({p2a, apair, (int)nn, nact});
  return MADAM_OK;
}

template <class T>
inline int down(T* dst, const T* src, size_t n, hipStream_t s) {
  if (n == 0) return MADAM_OK;
  HIP_TRY(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, s));
  return MADAM_OK;
}

struct DownRing {
  char* host = nullptr;          // pinned
  char* dev = nullptr;           // device staging
  madam::XDesc* hdesc = nullptr;  // pinned descriptors
  madam::XDesc* ddesc = nullptr;
  size_t cap = 0;
  size_t want = kBytes;  // capacity of the next allocation
  static constexpr size_t kBytes = 32u << 20;
  static constexpr size_t kPieces = 1u << 16;
  bool ensure() {
    if (cap) return true;
    if (hipHostMalloc((void**)&host, want, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&hdesc, kPieces * sizeof(madam::XDesc), hipHostMallocDefault) !=
            hipSuccess ||
        hipMalloc((void**)&dev, want) != hipSuccess ||
        hipMalloc((void**)&ddesc, kPieces * sizeof(madam::XDesc)) != hipSuccess) {
      (void)hipGetLastError();
      release();
      return false;
    }
    cap = want;
    return true;
  }
  void release() {
    if (host) (void)hipHostFree(host);
    if (hdesc) (void)hipHostFree(hdesc);
    if (dev) (void)hipFree(dev);
    if (ddesc) (void)hipFree(ddesc);
    host = dev = nullptr;
    hdesc = ddesc = nullptr;
    cap = 0;
  }
};
inline DownRing g_dring;

struct DownBatch {
  hipStream_t s;
  std::vector<void*> dst;
  std::vector<const char*> src;
  std::vector<size_t> bytes, off;
  size_t used = 0;
  explicit DownBatch(hipStream_t st) : s(st) {}
  int add(void* hdst, const void* dsrc, size_t nbytes) {
    if (nbytes == 0) return MADAM_OK;
    if (!g_dring.ensure() || nbytes > g_dring.cap / 4) {
      HIP_TRY(hipMemcpyAsync(hdst, dsrc, nbytes, hipMemcpyDeviceToHost, s));
      return MADAM_OK;
    }
    size_t o = align256(used);
    if (o + nbytes > g_dring.cap || dst.size() >= DownRing::kPieces) {
      int rc;
      if ((rc = finish())) return rc;
      o = 0;
    }
    dst.push_back(hdst);
    src.push_back((const char*)dsrc);
    bytes.push_back(nbytes);
    off.push_back(o);
    used = o + nbytes;
    return MADAM_OK;
  }
  template <class T>
  int get(T* hdst, const T* dsrc, size_t n) {
    return add((void*)hdst, (const void*)dsrc, n * sizeof(T));
  }
  // every piece on the host (and the stream synchronised)
  int finish() {
    const size_t nd = dst.size();
    if (nd > 0) {
      for (size_t i = 0; i < nd; ++i)
        g_dring.hdesc[i] = {src[i], g_dring.dev + off[i], (unsigned long long)bytes[i]};
      HIP_TRY(hipMemcpyAsync(g_dring.ddesc, g_dring.hdesc, nd * sizeof(madam::XDesc),
                             hipMemcpyHostToDevice, s));
      madam::launch_xfer(g_dring.ddesc, (int)nd, s);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipMemcpyAsync(g_dring.host, g_dring.dev, used, hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    for (size_t i = 0; i < nd; ++i) memcpy(dst[i], g_dring.host + off[i], bytes[i]);
    dst.clear();
    src.clear();
    bytes.clear();
    off.clear();
    used = 0;
    return MADAM_OK;
  }
};

// Fill view v from slot sl (device pointers) for a node of order n.
inline void bind_view(NodeView& v, Slot& sl, int n, const CutSet& cs) {
  memset(&v, 0, sizeof v);
  v.n = n;
  v.m = cs.m;
  v.nact = cs.nact;
  v.ne = cs.cptr.empty() ? 0 : cs.cptr[cs.m];
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
  v.aent16 = cs.aent16.empty() ? nullptr : sl.aent16.p;
  v.cact16 = cs.cact16.empty() ? nullptr : sl.cact16.p;
  const bool xr = !cs.tcut.empty();
  v.xrows = xr ? 1 : 0;
  v.xoff = xr ? sl.xoff.p : nullptr;
  v.xent = xr ? sl.xent.p : nullptr;
  v.tcut = xr ? sl.tcut.p : nullptr;
  v.tmeta = xr ? sl.tmeta.p : nullptr;
  v.tsc = xr ? sl.tsc.p : nullptr;
  v.tdinv = xr ? sl.tdinv.p : nullptr;
  v.xs = xr ? sl.xs.p : nullptr;
  v.apair = sl.apair.p;
  v.p2a = sl.p2a.p;
  v.t = sl.t.p;
  v.s = sl.s.p;
  v.u = sl.u.p;
  v.rhs = sl.rhs.p;
  v.tprev = sl.tprev.p;
  v.tprev2 = sl.tprev2.p;
  v.cg_r = sl.r.p;
  v.cg_p = sl.p.p;
  v.cg_q = sl.q.p;
  v.cg_w = sl.w.p;
  v.ip1 = sl.ip1.p;
  v.ip2 = sl.ip2.p;
  v.ip3 = sl.ip3.p;
  v.ipv = sl.ipv.p;
  v.ipP = sl.ipP.p;
  v.ns32 = sl.ns32.p;
}

inline int upload_cuts(Slot& sl, const CutSet& cs, hipStream_t s) {
  int rc;
  if ((rc = up(sl.cptr.p, cs.cptr.data(), cs.cptr.size(), s)) ||
      (rc = up(sl.cent.p, cs.cent.data(), cs.cent.size(), s)) ||
      (rc = up(sl.cact.p, cs.cact.data(), cs.cact.size(), s)) ||
      (rc = up(sl.aptr.p, cs.aptr.data(), cs.aptr.size(), s)) ||
      (rc = up(sl.aent.p, cs.aent.data(), cs.aent.size(), s)) ||
      (rc = up(sl.aent16.p, cs.aent16.data(), cs.aent16.size(), s)) ||
      (rc = up(sl.cact16.p, cs.cact16.data(), cs.cact16.size(), s)) ||
      (rc = up(sl.apair.p, cs.apair.data(), cs.apair.size(), s)) ||
      (rc = p2a_job(sl.p2a.p, sl.apair.p, cs.p2a.size(), cs.nact, cs.p2a.data(), s)) ||
      (rc = up(sl.cscale.p, cs.cscale.data(), cs.cscale.size(), s)))
    return rc;
  if (!cs.tcut.empty() &&
      ((rc = up(sl.xoff.p, cs.xoff.data(), cs.xoff.size(), s)) ||
       (rc = up(sl.xent.p, cs.xent.data(), cs.xent.size(), s)) ||
       (rc = up(sl.tcut.p, cs.tcut.data(), cs.tcut.size(), s)) ||
       (rc = up(sl.tmeta.p, cs.tmeta.data(), cs.tmeta.size(), s)) ||
       (rc = up(sl.tsc.p, cs.tsc.data(), cs.tsc.size(), s)) ||
       (rc = up(sl.tdinv.p, cs.tdinv.data(), cs.tdinv.size(), s))))
    return rc;
  return MADAM_OK;
}

inline int ensure_slots(int k) {
  if ((int)g_eng->slots.size() < k) g_eng->slots.resize(k);
  g_eng->hviews.resize(std::max<size_t>(g_eng->hviews.size(), k));
  return g_eng->views.ensure(k);
}

inline int sync_views_down(int k) {
  int rc = down(g_eng->hviews.data(), g_eng->views.p, k, g_eng->stream);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(g_eng->stream));
  return MADAM_OK;
}


// The ADMM loop of admm.c:520-704 plus the safe bound (admm.c:724-763) for
// the k nodes whose views are in hv (device buffers already filled: L, X,
// Q, y, cut structure, s = t = cut duals, u = 0).  hv is updated in place.
int admm_loop(std::vector<NodeView>& hv, const madam_opts& o, int nmax, int mmax);
int admm_advance(std::vector<NodeView>& hv, const madam_opts& o, int nmax, int mmax, int min_done);
int admm_finish(std::vector<NodeView>& hv, const madam_opts& o, int nmax, int mmax);
// Drop the captured check-interval graph of admm_advance.
void release_chunk_graph();
// FP64 eigen-solver of the plain form for orders from madam_set_bj_min_order
// on (madam_bjac.hip): two-sided block Jacobi over many workgroups per node,
// with the results the one-workgroup Jacobi kernel leaves (values_only: ev
// and off_last only; otherwise also evp, M, V and X_+ in S1).
// Nodes of order bj_min_order() or more go to bj_eig (it skips the others),
// smaller ones to k_jacobi (launch_jacobi with n_limit = bj_min_order()):
// each node's solver depends on its own order only.
bool bj_wanted(int nmax);
int bj_min_order();
int bj_eig(NodeView* d_nodes, const NodeView* hv, int nb, int values_only, int ignore_done,
           int max_sweeps, hipStream_t st);
void release_bj();

// Subproblems on the device (madam_subp.hip): upload_spl puts SP->L in HBM
// for the current call; device_subproblems builds PP->L of k nodes into
// dst[q] (device, n x n) and returns their orders and Frobenius norms.
int upload_spl(int N, const double* SPL);
int device_subproblems(int N, int k, const int* const* xfixed, const int* const* solX,
                       double* const* dst, int* n_out, double* norm_out);
void release_spl();
// The interior-point method for orders whose factor does not fit LDS,
// spread over many workgroups per node (madam_ipmt.hip); hv: host copies of
// the views in d_nodes.
int ipm_tiles(NodeView* d_nodes, const NodeView* hv, int nb, int nmax, hipStream_t st);
void release_ipm_tiles();
// The safe bound's certified lambda_max for orders beyond LDS, spread over
// many workgroups per node (madam_cert.hip): A in each node's W, the value
// in ev (what k_safe_bound reads).
int safe_tiles(NodeView* d_nodes, const NodeView* hv, int nb, int nmax, hipStream_t st);
void release_safe_tiles();

// Resolve the no-allocation path's device array on `device` (current).
int noalloc_setup(int device);

// Triangle separation on the given slots (X current on the device).
int tri_separate_slots(const std::vector<int>& slots, const int* K, double thresh,
                       std::vector<std::vector<madam_tri>>& lists, std::vector<double>& maxval);

}  // namespace mh
