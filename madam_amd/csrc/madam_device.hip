// madam_device.hip -- HIP kernels (gfx950) for the ADMM bounding loop of
// MADAM (admm.c:49-771, operators.c:62-619).
//
// Layout: every node of a batch is a NodeView (madam_internal.h); per-node
// kernels use one workgroup per node (blockIdx.x), streaming kernels use a
// 2-D grid (chunk, node).  All arithmetic is FP64.
//
// Rounding discipline: the cut maps and the fused elementwise steps are
// written with FP contraction off and in the reference's order of
// operations, so that for identical inputs they give the same bits as
// operators.c (op_B, op_Diag, op_Bt) and the elementwise lines of admm.c.
#include "madam_internal.h"
#include "madam_devutil.h"

#include <math.h>

namespace madam {

static_assert(XCUT_THREADS == WG, "the expanded rows are assigned to the cut solve's threads");

// ------------------------------------------------------------ ADMM steps

// y = diag(L+Q+X/sigma) - e/sigma ; rhs = s + B(L+Q+X/sigma) + (u-e)/sigma
// (admm.c:526-543).
__global__ __launch_bounds__(WGS) void k_admm_prep(NodeView* nodes) {
#pragma clang fp contract(off)
  NodeView& nd = nodes[blockIdx.y];
  if (nd.done) return;
  const int n = nd.n, m = nd.m;
  const double sg = nd.sigma;
  const double* __restrict__ L = nd.L;
  const double* __restrict__ Q = nd.Q;
  const double* __restrict__ X = nd.X;
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < n; i += WGS) {
      const size_t dg = (size_t)i * (n + 1);
      nd.y[i] = L[dg] + Q[dg] + (X[dg] - 1.0) / sg;
    }
  for (int c = blockIdx.x * WGS + threadIdx.x; c < m; c += gridDim.x * WGS) {
    const double acc = signed_sum<false>(nd.cent, nd.cptr[c], nd.cptr[c + 1], 0.0,
                                         [&](int lo) { return L[lo] + Q[lo] + X[lo] / sg; });
    double r = nd.s[c] + nd.cscale[c] * acc;
    nd.rhs[c] = r + (nd.u[c] - 1) / sg;
  }
}

// q = (BB^T + I) p, matrix-free: w_a = sum_c coef p_c over active pairs,
// q_c = p_c + 1/2 sum_a coef w_a.  coef = sign * scale_c.
__device__ void spmv_bbt(const NodeView& nd, const double* __restrict__ p, double* __restrict__ q,
                         double* __restrict__ w) {
  __syncthreads();
  for (int a = threadIdx.x; a < nd.nact; a += blockDim.x) {
    double acc = 0.0;
    for (int e = nd.aptr[a]; e < nd.aptr[a + 1]; ++e) {
      const int en = nd.aent[e], c = ent_idx(en);
      const double v = nd.cscale[c] * p[c];
      acc = en >= 0 ? acc + v : acc - v;
    }
    w[a] = acc;
  }
  __syncthreads();
  for (int c = threadIdx.x; c < nd.m; c += blockDim.x) {
    double acc = 0.0;
    for (int e = nd.cptr[c]; e < nd.cptr[c + 1]; ++e) {
      const int ca = nd.cact[e];
      const double v = w[ent_idx(ca)];
      acc = ca >= 0 ? acc + v : acc - v;
    }
    q[c] = p[c] + 0.5 * nd.cscale[c] * acc;
  }
  __syncthreads();
}

// Which cut-solve kernel serves a node: 0 k_cg1<2> (m <= 2 WG), 1 k_cg1<4>
// (m <= 4 WG; both with the cut and pair vectors in LDS), 2 k_cg_x (one-pass
// products over the sliced Gram rows, when built, with the +-ps table in
// LDS), 3
// k_cg_big<true> (cut and pair vectors in LDS), 4 k_cg_big<false> (only the
// cut vector fits), 5 k_cg (neither fits).  Decided per node, so that one
// large node does not move a whole batch to a slower kernel; every
// cut-solve launch skips the nodes of other classes.
__host__ __device__ inline int cg_class(int m, int nact, int xrows, size_t lds_limit) {
  const size_t base = ((size_t)m + nact + 64) * sizeof(double);
  const size_t cut_only = ((size_t)m + 64) * sizeof(double);
  if (base <= lds_limit && m <= 4 * WG) return m <= 2 * WG ? 0 : 1;
  // k_cg_x: the +-ps table (2 m + 2 doubles) and 64 reduction doubles
  if (xrows && (2 * (size_t)m + 66) * sizeof(double) <= lds_limit) return 2;
  if (base <= lds_limit) return 3;
  return cut_only <= lds_limit ? 4 : 5;
}

// Stopping tolerance of a cut solve: rtol, or with adapt > 0 after the
// first iteration of an ADMM call max(rtol, min(1e-3, adapt * the previous
// iteration's ADMM error)) -- an inexact solve while ADMM is far from
// converged, the full tolerance as it gets there.
__device__ __forceinline__ double cg_tol(const NodeView& nd, double rtol, double adapt) {
  if (adapt > 0.0 && nd.nbit > 0) return fmax(rtol, fmin(1e-3, adapt * nd.err));
  return rtol;
}

// Solve (BB^T + I) t = rhs by conjugate gradients with the diagonal as
// preconditioner, starting from the current t (the previous iterate), to a
// relative residual of rtol.  Replaces the CHOLMOD factor/solve of
// admm.c:486-551.
__global__ __launch_bounds__(WG) void k_cg(NodeView* nodes, double rtol, double adapt, int maxit,
                                           size_t lds_limit, int use_x) {
  __shared__ double red[4 * 16];
  NodeView& nd = nodes[blockIdx.x];
  if (nd.done || cg_class(nd.m, nd.nact, use_x ? nd.xrows : 0, lds_limit) != 5) return;
  const int m = nd.m;
  if (m == 0) return;
  double* __restrict__ x = nd.t;
  double* __restrict__ r = nd.cg_r;
  double* __restrict__ p = nd.cg_p;
  double* __restrict__ q = nd.cg_q;
  const double* __restrict__ b = nd.rhs;
  // starting guess extrapolated from the previous solutions (as k_cg_big)
  for (int c = threadIdx.x; c < m; c += WG) {
    const double tk = x[c], t1 = nd.tprev[c];
    if (nd.nbit > 1)
      x[c] = 3.0 * (tk - t1) + nd.tprev2[c];
    else if (nd.nbit > 0)
      x[c] = tk + (tk - t1);
    nd.tprev2[c] = t1;
    nd.tprev[c] = tk;
  }
  spmv_bbt(nd, x, q, nd.cg_w);
  double v3[3] = {0, 0, 0};  // r.z, r.r, b.b
  for (int c = threadIdx.x; c < m; c += WG) {
    const int k = nd.cptr[c + 1] - nd.cptr[c];
    const double dinv = 1.0 / (1.0 + 0.5 * k * nd.cscale[c] * nd.cscale[c]);
    const double rc = b[c] - q[c];
    r[c] = rc;
    p[c] = rc * dinv;
    v3[0] += rc * rc * dinv;
    v3[1] += rc * rc;
    v3[2] += b[c] * b[c];
  }
  block_sums<3>(v3, red);
  double rz = v3[0], rr = v3[1];
  const double bb = v3[2];
  const double rt = cg_tol(nd, rtol, adapt);
  const double thr = rt * rt * bb;
  int it = 0;
  while (it < maxit && rr > thr && bb > 0.0) {
    spmv_bbt(nd, p, q, nd.cg_w);
    double v1[1] = {0.0};
    for (int c = threadIdx.x; c < m; c += WG) v1[0] += p[c] * q[c];
    block_sums<1>(v1, red);
    const double alpha = rz / v1[0];
    double v2[2] = {0, 0};
    for (int c = threadIdx.x; c < m; c += WG) {
      const int k = nd.cptr[c + 1] - nd.cptr[c];
      const double dinv = 1.0 / (1.0 + 0.5 * k * nd.cscale[c] * nd.cscale[c]);
      x[c] += alpha * p[c];
      const double rc = r[c] - alpha * q[c];
      r[c] = rc;
      v2[0] += rc * rc * dinv;
      v2[1] += rc * rc;
    }
    block_sums<2>(v2, red);
    const double beta = v2[0] / rz;
    rz = v2[0];
    rr = v2[1];
    for (int c = threadIdx.x; c < m; c += WG) {
      const int k = nd.cptr[c + 1] - nd.cptr[c];
      const double dinv = 1.0 / (1.0 + 0.5 * k * nd.cscale[c] * nd.cscale[c]);
      p[c] = r[c] * dinv + beta * p[c];
    }
    ++it;
  }
  if (threadIdx.x == 0) {
    nd.cg_iters = it;
    nd.cg_total += it;
  }
}

// Signed sums over S index ranges at once: acc[s] (+/-) v[ent_idx(en)] over
// the entries en of ent[e[s]..e1[s]), each range in entry order (so each sum
// is the same bits as the one-at-a-time loop).  G entries of every range,
// then their values, are loaded before any is added -- S * G loads in
// flight -- and a range past its end reloads its last entry (or entry 0 of
// ent / v when empty; both must be non-empty) whose value is not added.
template <int S, int G, class Idx>
__device__ __forceinline__ void signed_sums(const Idx* __restrict__ ent, const double* v,
                                            int (&e)[S], const int (&e1)[S], double (&acc)[S]) {
  bool more = false;
#pragma unroll
  for (int s = 0; s < S; ++s) more = more || e[s] < e1[s];
  while (more) {
    int k[S][G];
    double x[S][G];
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
      for (int q = 0; q < G; ++q)
        k[s][q] = e[s] < e1[s] ? (int)ent[e[s] + q < e1[s] ? e[s] + q : e1[s] - 1] : 0;
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
      for (int q = 0; q < G; ++q) x[s][q] = v[ent_idx(k[s][q])];
    more = false;
#pragma unroll
    for (int s = 0; s < S; ++s) {
#pragma unroll
      for (int q = 0; q < G; ++q) {
        const double t = k[s][q] >= 0 ? acc[s] + x[s][q] : acc[s] - x[s][q];
        acc[s] = e[s] + q < e1[s] ? t : acc[s];
      }
      e[s] += G;
      more = more || e[s] < e1[s];
    }
  }
}

// q = (BB^T + I) v for the cut solves of one workgroup, with the cut-side
// vector ps = scale * v in LDS and the pair sums w in LDS (WLDS) or in HBM
// (nd.cg_w, when m + nact doubles do not fit).  The pair pass takes four
// pairs per thread at a time; the cut pass gives each thread cut q together
// with cut m-1-q (triangles, with 3 entries, come first and heptagonals,
// with 21, last), two rows in flight.  Every sum is in the entry order of
// the one-at-a-time loops, so the products are bit-identical to k_cg's
// spmv_bbt.  On entry ps must hold scale * v and be visible to the
// workgroup; on return q is, to the whole workgroup.
template <class Idx>
__device__ __forceinline__ void bbt_big(const NodeView& nd, const Idx* __restrict__ aent,
                                        const Idx* __restrict__ cact, const double* __restrict__ v,
                                        double* __restrict__ out, const double* ps, double* w) {
  const int m = nd.m, na = nd.nact, tid = threadIdx.x;
  const int* __restrict__ aptr = nd.aptr;
  const int* __restrict__ cptr = nd.cptr;
  for (int a = tid; a < na; a += 4 * WG) {
    int e[4], e1[4];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int as = a + s * WG;
      e[s] = as < na ? aptr[as] : 0;
      e1[s] = as < na ? aptr[as + 1] : 0;
    }
    signed_sums<4, 2>(aent, ps, e, e1, acc);
#pragma unroll
    for (int s = 0; s < 4; ++s)
      if (a + s * WG < na) w[a + s * WG] = acc[s];
  }
  __syncthreads();
  const int half = (m + 1) >> 1;
  for (int q = tid; q < half; q += WG) {
    const int c2 = m - 1 - q;
    const bool two = c2 > q;
    int e[2] = {cptr[q], two ? cptr[c2] : 0};
    const int e1[2] = {cptr[q + 1], two ? cptr[c2 + 1] : 0};
    double acc[2] = {0.0, 0.0};
    signed_sums<2, 4>(cact, w, e, e1, acc);
    out[q] = v[q] + 0.5 * nd.cscale[q] * acc[0];
    if (two) out[c2] = v[c2] + 0.5 * nd.cscale[c2] * acc[1];
  }
  __syncthreads();
}

// The PCG of k_cg for nodes whose cut-side vector fits LDS: ps (and, with
// WLDS, the pair sums) in LDS, the index streams (16-bit when m, nact <
// 2^15) and the PCG vectors in HBM; products by bbt_big.  The update of p
// is fused into the pass that forms the next product's ps.  Same
// operations in the same order as k_cg: the iterates are bit-identical.
template <class Idx, bool WLDS>
__device__ void cg_big_body(NodeView& nd, double rtol, double adapt, int maxit) {
  extern __shared__ double smem[];
  const Idx* aent = sizeof(Idx) == 2 ? (const Idx*)nd.aent16 : (const Idx*)nd.aent;
  const Idx* cact = sizeof(Idx) == 2 ? (const Idx*)nd.cact16 : (const Idx*)nd.cact;
  if (!aent) return;
  const int m = nd.m, tid = threadIdx.x;
  if (m == 0) return;
  double* ps = smem;
  double* w = WLDS ? smem + m : nd.cg_w;
  double* red = smem + m + (WLDS ? nd.nact : 0);
  double* __restrict__ x = nd.t;
  double* __restrict__ r = nd.cg_r;
  double* __restrict__ p = nd.cg_p;
  double* __restrict__ q = nd.cg_q;
  const double* __restrict__ b = nd.rhs;
  const int* __restrict__ cptr = nd.cptr;
  const double* __restrict__ cs = nd.cscale;
  // starting guess: the previous solutions extrapolated -- quadratically,
  // 3 t_k - 3 t_{k-1} + t_{k-2}, from the third iteration of an ADMM call
  // on, linearly at the second (the iterates move smoothly, so the first
  // residual is smaller)
#pragma unroll 4
  for (int c = tid; c < m; c += WG) {
    const double tk = x[c], t1 = nd.tprev[c];
    double xc = tk;
    if (nd.nbit > 1)
      xc = 3.0 * (tk - t1) + nd.tprev2[c];
    else if (nd.nbit > 0)
      xc = tk + (tk - t1);
    x[c] = xc;
    nd.tprev2[c] = t1;
    nd.tprev[c] = tk;
    ps[c] = cs[c] * xc;
  }
  __syncthreads();
  bbt_big(nd, aent, cact, x, q, ps, w);
  double v3[3] = {0, 0, 0};  // r.z, r.r, b.b
#pragma unroll 4
  for (int c = tid; c < m; c += WG) {
    const int k = cptr[c + 1] - cptr[c];
    const double dinv = 1.0 / (1.0 + 0.5 * k * cs[c] * cs[c]);
    const double rc = b[c] - q[c];
    r[c] = rc;
    const double pc = rc * dinv;
    p[c] = pc;
    ps[c] = cs[c] * pc;
    v3[0] += rc * rc * dinv;
    v3[1] += rc * rc;
    v3[2] += b[c] * b[c];
  }
  block_sums<3>(v3, red);
  double rz = v3[0], rr = v3[1];
  const double bb = v3[2];
  const double rt = cg_tol(nd, rtol, adapt);
  const double thr = rt * rt * bb;
  int it = 0;
  while (it < maxit && rr > thr && bb > 0.0) {
    bbt_big(nd, aent, cact, p, q, ps, w);
    double v1[1] = {0.0};
#pragma unroll 4
    for (int c = tid; c < m; c += WG) v1[0] += p[c] * q[c];
    block_sums<1>(v1, red);
    const double alpha = rz / v1[0];
    double v2[2] = {0, 0};
#pragma unroll 4
    for (int c = tid; c < m; c += WG) {
      const int k = cptr[c + 1] - cptr[c];
      const double dinv = 1.0 / (1.0 + 0.5 * k * cs[c] * cs[c]);
      x[c] += alpha * p[c];
      const double rc = r[c] - alpha * q[c];
      r[c] = rc;
      v2[0] += rc * rc * dinv;
      v2[1] += rc * rc;
    }
    block_sums<2>(v2, red);
    const double beta = v2[0] / rz;
    rz = v2[0];
    rr = v2[1];
    ++it;
    if (!(it < maxit && rr > thr)) break;
#pragma unroll 4
    for (int c = tid; c < m; c += WG) {
      const int k = cptr[c + 1] - cptr[c];
      const double dinv = 1.0 / (1.0 + 0.5 * k * cs[c] * cs[c]);
      const double pc = r[c] * dinv + beta * p[c];
      p[c] = pc;
      ps[c] = cs[c] * pc;
    }
    __syncthreads();
  }
  if (tid == 0) {
    nd.cg_iters = it;
    nd.cg_total += it;
  }
}

template <bool WLDS>
__global__ __launch_bounds__(WG) void k_cg_big(NodeView* nodes, double rtol, double adapt, int maxit,
                                               size_t lds_limit, int use_x) {
  NodeView& nd = nodes[blockIdx.x];
  if (nd.done || cg_class(nd.m, nd.nact, use_x ? nd.xrows : 0, lds_limit) != (WLDS ? 3 : 4)) return;
  if (nd.aent16)
    cg_big_body<short, WLDS>(nd, rtol, adapt, maxit);
  else
    cg_big_body<int, WLDS>(nd, rtol, adapt, maxit);
}

// sum of +-v[idx] over the signed entries ent[e0..e1) in entry order,
// four loads in flight at a time (the sum itself stays sequential)
template <class Idx>
__device__ __forceinline__ double signed_gather(const Idx* __restrict__ ent,
                                                const double* __restrict__ v, int e0, int e1) {
  double acc = 0.0;
  int e = e0;
  for (; e + 4 <= e1; e += 4) {
    const int k0 = ent[e], k1 = ent[e + 1], k2 = ent[e + 2], k3 = ent[e + 3];
    const double v0 = v[ent_idx(k0)], v1 = v[ent_idx(k1)], v2 = v[ent_idx(k2)],
                 v3 = v[ent_idx(k3)];
    acc = k0 >= 0 ? acc + v0 : acc - v0;
    acc = k1 >= 0 ? acc + v1 : acc - v1;
    acc = k2 >= 0 ? acc + v2 : acc - v2;
    acc = k3 >= 0 ? acc + v3 : acc - v3;
  }
  for (; e < e1; ++e) {
    const int k = ent[e];
    const double x = v[ent_idx(k)];
    acc = k >= 0 ? acc + x : acc - x;
  }
  return acc;
}

// The cut solve in the Chronopoulos-Gear form of PCG (one reduction per
// iteration instead of two: gamma = (r, u), delta = (A u, u) and |r|^2 are
// summed together, and s = A p is carried by recurrence), with every
// cut-side vector in registers: each thread owns up to CPT cuts and keeps
// x, r, p, s, the diagonal preconditioner and the cut scale of each for the
// whole solve.  ps (scale * u, gathered by the pair pass) and
// w (pair sums) are in LDS; with STAGE the index streams (aptr, aent, cact)
// are copied to LDS once per solve, so the iterations read no HBM at all.
// Per iteration: one product (two barriers) and one reduction (one
// barrier).  Same linear system, start and stopping rule as k_cg; the
// iterates agree with it to rounding.  Used when m <= CPT WG (CPT 2 or 4:
// more cuts per thread would spill registers).
//   r0 = b - A x0, u0 = D r0, w0 = A u0, alpha0 = (r0,u0) / (w0,u0)
//   p = u + beta p, s = w + beta s, x += alpha p, r -= alpha s, u = D r,
//   w = A u, gamma' = (r,u), delta = (w,u), beta = gamma' / gamma,
//   alpha = gamma' / (delta - beta gamma' / alpha)
template <int CPT, class Idx>
__device__ void cg1_body(NodeView& nd, double rtol, double adapt, int maxit, size_t lds_alloc) {
  extern __shared__ double smem[];
  const Idx* aent = sizeof(Idx) == 2 ? (const Idx*)nd.aent16 : (const Idx*)nd.aent;
  const Idx* cact = sizeof(Idx) == 2 ? (const Idx*)nd.cact16 : (const Idx*)nd.cact;
  if (!aent) return;
  const int m = nd.m, na = nd.nact, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (m == 0) return;
  double* ps = smem;
  double* w = smem + m;
  double* red = w + na;
  const int* aptr = nd.aptr;
  // the index streams are staged in LDS when this node's fit the allocation
  const bool stage = ((size_t)m + na + 64) * sizeof(double) + ((size_t)na + 1) * sizeof(int) +
                         2 * (size_t)nd.ne * sizeof(Idx) <=
                     lds_alloc;
  if (stage) {
    const int ne = nd.ne;
    int* sapt = (int*)(red + 64);
    Idx* sae = (Idx*)(sapt + na + 1);
    Idx* sca = sae + ne;
    for (int i = tid; i <= na; i += WG) sapt[i] = nd.aptr[i];
    for (int i = tid; i < ne; i += WG) {
      sae[i] = aent[i];
      sca[i] = cact[i];
    }
    aptr = sapt;
    aent = sae;
    cact = sca;  // visible after the barrier that follows the start below
  }
  // reduction with one barrier: every call below is separated from the
  // previous one by the two barriers of a product, so no thread can still
  // be reading red when another writes it
  auto reduce = [&](auto& v) {
    constexpr int NV = sizeof(v) / sizeof(double);
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = wave_sum(v[q]);
    if (lane == 0) {
#pragma unroll
      for (int q = 0; q < NV; ++q) red[q * 16 + wv] = v[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double a = 0.0;
      for (int i = 0; i < WG / 64; ++i) a += red[q * 16 + i];
      v[q] = a;
    }
  };
  // pair pass w = sum of the signed ps over the cuts of each active pair,
  // then the cut-side sums of w for the calling thread's cuts
  auto pair_pass = [&]() {
    for (int a = tid; a < na; a += WG) w[a] = signed_gather(aent, ps, aptr[a], aptr[a + 1]);
    __syncthreads();
  };
  // cuts come ordered triangles (3 pairs), pentagonal (10), heptagonal
  // (21): thread tid owns cut q and cut m-1-q for q = tid + i WG, so every
  // thread gets a like share of pair entries
  const int half = (m + 1) >> 1;
  auto cut_of = [&](int j) {
    const int q = tid + (j >> 1) * WG;
    if (q >= half) return m;  // none
    if (j & 1) return m - 1 - q > q ? m - 1 - q : m;
    return q;
  };
  // u = D r is recomputed where needed rather than kept (registers)
  double x[CPT], r[CPT], p[CPT], s[CPT], dinv[CPT], sc[CPT];
  int e0[CPT], e1[CPT];
  auto cut_acc = [&](int j) { return signed_gather(cact, w, e0[j], e1[j]); };
#pragma unroll
  for (int j = 0; j < CPT; ++j) {
    const int c = cut_of(j);
    x[j] = r[j] = p[j] = s[j] = dinv[j] = sc[j] = 0.0;
    e0[j] = e1[j] = 0;
    if (c < m) {
      // starting guess extrapolated from the previous solutions (as k_cg)
      const double tk = nd.t[c], t1 = nd.tprev[c];
      double xc = tk;
      if (nd.nbit > 1)
        xc = 3.0 * (tk - t1) + nd.tprev2[c];
      else if (nd.nbit > 0)
        xc = tk + (tk - t1);
      nd.tprev2[c] = t1;
      nd.tprev[c] = tk;
      x[j] = xc;
      e0[j] = nd.cptr[c];
      e1[j] = nd.cptr[c + 1];
      sc[j] = nd.cscale[c];
      dinv[j] = 1.0 / (1.0 + 0.5 * (e1[j] - e0[j]) * sc[j] * sc[j]);
      ps[c] = sc[j] * xc;
    }
  }
  __syncthreads();
  pair_pass();
  double v3[3] = {0, 0, 0};  // (r,u), (r,r), (b,b)
#pragma unroll
  for (int j = 0; j < CPT; ++j) {
    const int c = cut_of(j);
    if (c < m) {
      const double b = nd.rhs[c];
      const double rc = b - (x[j] + 0.5 * sc[j] * cut_acc(j));
      r[j] = rc;
      const double uc = rc * dinv[j];
      v3[0] += rc * uc;
      v3[1] += rc * rc;
      v3[2] += b * b;
    }
  }
  reduce(v3);
  double gamma = v3[0], rr = v3[1];
  const double bb = v3[2];
  const double rt = cg_tol(nd, rtol, adapt);
  const double thr = rt * rt * bb;
  double alpha = 0.0, beta = 0.0;
  int it = 0;
  bool run = it < maxit && rr > thr && bb > 0.0;
  double wl[CPT];
  if (run) {
    // w0 = A u0, alpha0 = (r0,u0) / (w0,u0)
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
      const int c = cut_of(j);
      if (c < m) ps[c] = sc[j] * (r[j] * dinv[j]);
    }
    __syncthreads();
    pair_pass();
    double v1[1] = {0.0};
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
      const int c = cut_of(j);
      wl[j] = 0.0;
      if (c < m) {
        const double uc = r[j] * dinv[j];
        wl[j] = uc + 0.5 * sc[j] * cut_acc(j);
        v1[0] += wl[j] * uc;
      }
    }
    reduce(v1);
    alpha = gamma / v1[0];
  }
  while (run) {
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
      const int c = cut_of(j);
      p[j] = r[j] * dinv[j] + beta * p[j];
      s[j] = wl[j] + beta * s[j];
      x[j] += alpha * p[j];
      r[j] -= alpha * s[j];
      if (c < m) ps[c] = sc[j] * (r[j] * dinv[j]);
    }
    __syncthreads();
    pair_pass();
    double v[3] = {0, 0, 0};  // (r,u), (w,u), (r,r)
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
      const int c = cut_of(j);
      wl[j] = 0.0;
      if (c < m) {
        const double uc = r[j] * dinv[j];
        wl[j] = uc + 0.5 * sc[j] * cut_acc(j);
        v[0] += r[j] * uc;
        v[1] += wl[j] * uc;
        v[2] += r[j] * r[j];
      }
    }
    reduce(v);
    ++it;
    rr = v[2];
    beta = v[0] / gamma;
    alpha = v[0] / (v[1] - beta * v[0] / alpha);
    gamma = v[0];
    run = it < maxit && rr > thr;
  }
#pragma unroll
  for (int j = 0; j < CPT; ++j) {
    const int c = cut_of(j);
    if (c < m) nd.t[c] = x[j];
  }
  if (tid == 0) {
    nd.cg_iters = it;
    nd.cg_total += it;
  }
}

template <int CPT>
__global__ __launch_bounds__(WG) void k_cg1(NodeView* nodes, double rtol, double adapt, int maxit,
                                            size_t lds_limit, size_t lds_alloc, int use_x) {
  NodeView& nd = nodes[blockIdx.x];
  if (nd.done || cg_class(nd.m, nd.nact, use_x ? nd.xrows : 0, lds_limit) != (CPT == 2 ? 0 : 1)) return;
  if (nd.aent16)
    cg1_body<CPT, short>(nd, rtol, adapt, maxit, lds_alloc);
  else
    cg1_body<CPT, int>(nd, rtol, adapt, maxit, lds_alloc);
}

// Sums of the sliced Gram rows (NodeView::xent) of two slots of the
// calling lane at once: row s has its words at xe[f[s] + 64 q] for q <
// L[s] (its slice's length, a multiple of 8, rows padded with the zero
// entry's offset); each word is the byte offset in the LDS table `tab` of
// the signed entry to add to acc[s], in row order.  The loop runs to L[0]
// >= L[1], row 1 only while it lasts (both the same on every lane of a
// wave); 8 words of each row in flight, each load one contiguous access of
// the wave.
__device__ __forceinline__ void x_rows2(const GMEM unsigned* __restrict__ xe, const char* tab,
                                        const int (&f)[2], const int (&L)[2], double (&acc)[2]) {
  constexpr int G = 8;
  for (int q0 = 0; q0 < L[0]; q0 += G) {
    unsigned wa[G], wb[G];
    const bool two = q0 < L[1];
#pragma unroll
    for (int q = 0; q < G; ++q) wa[q] = xe[f[0] + 64 * (q0 + q)];
    if (two) {
#pragma unroll
      for (int q = 0; q < G; ++q) wb[q] = xe[f[1] + 64 * (q0 + q)];
    }
    double va[G], vb[G];
#pragma unroll
    for (int q = 0; q < G; ++q) va[q] = *(const double*)(tab + wa[q]);
    if (two) {
#pragma unroll
      for (int q = 0; q < G; ++q) vb[q] = *(const double*)(tab + wb[q]);
    }
#pragma unroll
    for (int q = 0; q < G; ++q) acc[0] += va[q];
    if (two) {
#pragma unroll
      for (int q = 0; q < G; ++q) acc[1] += vb[q];
    }
  }
}

// The cut solve in the Chronopoulos-Gear form of k_cg1 with one-pass
// products over the sliced Gram rows of BB^T.  Slot i (cut tcut[i], rows
// longest first) belongs to thread i % WG; the PCG vectors x, r, p, s and
// w = A u live in HBM in slot order (every elementwise access of a wave is
// contiguous), with the scale and the diagonal preconditioner of each
// slot; the LDS holds the table +-ps (ps = scale * u, entries 2c and 2c+1
// for cut c, entry 2m zero) that the rows address.  The solution is
// gathered from and scattered to t (cut order) once per solve, and
// tprev/tprev2 are kept in slot order.  Each thread forms the products of
// its slots two at a time: (A u)_c = u_c + scale_c / 2 (k_c ps_c + row sum).
// Per iteration: one elementwise pass, one barrier, the product with the
// three dot products, one reduction (one barrier).  Same linear system,
// start and stopping rule as k_cg; the iterates agree with it to rounding.
__global__ __launch_bounds__(WG) void k_cg_x(NodeView* nodes, double rtol, double adapt, int maxit,
                                             size_t lds_limit) {
  NodeView& nd = nodes[blockIdx.x];
  if (nd.done || cg_class(nd.m, nd.nact, nd.xrows, lds_limit) != 2) return;
  extern __shared__ double smem[];
  const int m = nd.m, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double* tab = smem;
  double* red = smem + 2 * m + 2;
  const GMEM int* __restrict__ xoff = gmem(nd.xoff);
  const GMEM unsigned* __restrict__ xe = gmem(nd.xent);
  const GMEM int* __restrict__ tcut = gmem(nd.tcut);
  const GMEM int* __restrict__ tmeta = gmem(nd.tmeta);
  const GMEM double* __restrict__ tsc = gmem(nd.tsc);
  const GMEM double* __restrict__ tdv = gmem(nd.tdinv);
  const GMEM double* __restrict__ b = gmem(nd.rhs);
  GMEM double* __restrict__ t = gmem(nd.t);
  GMEM double* __restrict__ x = gmem(nd.xs);
  GMEM double* __restrict__ r = gmem(nd.cg_r);
  GMEM double* __restrict__ p = gmem(nd.cg_p);
  GMEM double* __restrict__ s = gmem(nd.cg_q);
  GMEM double* __restrict__ w = gmem(nd.cg_w);
  GMEM double* __restrict__ tp1 = gmem(nd.tprev);
  GMEM double* __restrict__ tp2 = gmem(nd.tprev2);
  if (tid == 0) tab[2 * m] = 0.0;
  // this thread's slots: tid + k WG for k < ks
  const int ks = m > tid ? (m - tid + WG - 1) / WG : 0;
  // reduction with one barrier: every call is separated from the previous
  // one by the barrier that publishes the table, so no thread still reads red
  auto reduce = [&](auto& v) {
    constexpr int NV = sizeof(v) / sizeof(double);
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = wave_sum(v[q]);
    if (lane == 0) {
#pragma unroll
      for (int q = 0; q < NV; ++q) red[q * 16 + wv] = v[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double a = 0.0;
      for (int i = 0; i < WG / 64; ++i) a += red[q * 16 + i];
      v[q] = a;
    }
  };
  auto put = [&](int c, double v) {
    tab[2 * c] = v;
    tab[2 * c + 1] = -v;
  };
  // per-slot operands loaded ahead of use
  struct Ops {
    double a, b, c, d, e, sc, di;
    int k, cut;
  };
  // this thread's slots four at a time: ld(i, o) for all four (a step past
  // the thread's share loads its first slot's operands, unused), then
  // st(i, o) for the real ones
  auto forall4 = [&](auto ld, auto st) {
    for (int k = 0; k < ks; k += 4) {
      Ops o[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) ld(tid + (k + q < ks ? k + q : 0) * WG, o[q]);
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (k + q < ks) st(tid + (k + q) * WG, o[q]);
    }
  };
  // this thread's slots two at a time through the product: ld(i, o) before
  // the row sums, u(o) the slot's u, fin(i, o, acc) after with acc =
  // k ps + the row sum
  auto products = [&](auto ld, auto u, auto fin) {
    for (int k = 0; k < ks; k += 2) {
      const bool two = k + 1 < ks;
      const int i0 = tid + k * WG, i1 = two ? i0 + WG : i0;
      Ops o[2];
      ld(i0, o[0]);
      ld(i1, o[1]);
      const int s0 = i0 >> 6, s1 = i1 >> 6;
      const int f[2] = {xoff[s0] + lane, xoff[s1] + lane};
      const int L[2] = {(xoff[s0 + 1] - xoff[s0]) >> 6,
                        two ? (xoff[s1 + 1] - xoff[s1]) >> 6 : 0};
      double acc[2] = {o[0].k * (o[0].sc * u(o[0])), o[1].k * (o[1].sc * u(o[1]))};
      x_rows2(xe, (const char*)tab, f, L, acc);
      fin(i0, o[0], acc[0]);
      if (two) fin(i1, o[1], acc[1]);
    }
  };
  auto ld_r = [&](int i, Ops& o) {
    o.a = r[i];
    o.sc = tsc[i];
    o.di = tdv[i];
    o.k = tmeta[i];
    o.cut = tcut[i];
  };
  auto u_r = [](const Ops& o) { return o.a * o.di; };
  // starting guess extrapolated from the previous solutions (as k_cg)
  const int nbit = nd.nbit;
  forall4(
      [&](int i, Ops& o) {
        o.cut = tcut[i];
        o.a = t[o.cut];
        o.b = tp1[i];
        o.c = tp2[i];
        o.sc = tsc[i];
      },
      [&](int i, Ops& o) {
        const double tk = o.a, t1 = o.b;
        double xc = tk;
        if (nbit > 1)
          xc = 3.0 * (tk - t1) + o.c;
        else if (nbit > 0)
          xc = tk + (tk - t1);
        x[i] = xc;
        tp2[i] = t1;
        tp1[i] = tk;
        put(o.cut, o.sc * xc);
      });
  __syncthreads();
  double v3[3] = {0, 0, 0};  // (r,u), (r,r), (b,b)
  products(
      [&](int i, Ops& o) {
        o.cut = tcut[i];
        o.a = x[i];
        o.b = b[o.cut];
        o.sc = tsc[i];
        o.di = tdv[i];
        o.k = tmeta[i];
      },
      [](const Ops& o) { return o.a; },
      [&](int i, Ops& o, double acc) {
        const double rc = o.b - (o.a + 0.5 * o.sc * acc);
        r[i] = rc;
        const double uc = rc * o.di;
        v3[0] += rc * uc;
        v3[1] += rc * rc;
        v3[2] += o.b * o.b;
      });
  reduce(v3);
  double gamma = v3[0], rr = v3[1];
  const double bb = v3[2];
  const double rt = cg_tol(nd, rtol, adapt);
  const double thr = rt * rt * bb;
  double alpha = 0.0, beta = 0.0;
  int it = 0;
  bool run = it < maxit && rr > thr && bb > 0.0;
  if (run) {
    // w0 = A u0, alpha0 = (r0,u0) / (w0,u0)
    forall4(ld_r, [&](int i, Ops& o) {
      put(o.cut, o.sc * u_r(o));
      p[i] = 0.0;
      s[i] = 0.0;
    });
    __syncthreads();
    double v1[1] = {0.0};
    products(ld_r, u_r, [&](int i, Ops& o, double acc) {
      const double uc = u_r(o);
      const double wc = uc + 0.5 * o.sc * acc;
      w[i] = wc;
      v1[0] += wc * uc;
    });
    reduce(v1);
    alpha = gamma / v1[0];
  }
  while (run) {
    forall4(
        [&](int i, Ops& o) {
          o.a = r[i];
          o.b = p[i];
          o.c = w[i];
          o.d = s[i];
          o.e = x[i];
          o.sc = tsc[i];
          o.di = tdv[i];
          o.cut = tcut[i];
        },
        [&](int i, Ops& o) {
          const double pc = o.a * o.di + beta * o.b;
          const double sc = o.c + beta * o.d;
          p[i] = pc;
          s[i] = sc;
          x[i] = o.e + alpha * pc;
          const double rc = o.a - alpha * sc;
          r[i] = rc;
          put(o.cut, o.sc * (rc * o.di));
        });
    __syncthreads();
    double v[3] = {0, 0, 0};  // (r,u), (w,u), (r,r)
    products(ld_r, u_r, [&](int i, Ops& o, double acc) {
      const double rc = o.a;
      const double uc = u_r(o);
      const double wc = uc + 0.5 * o.sc * acc;
      w[i] = wc;
      v[0] += rc * uc;
      v[1] += wc * uc;
      v[2] += rc * rc;
    });
    reduce(v);
    ++it;
    rr = v[2];
    beta = v[0] / gamma;
    alpha = v[0] / (v[1] - beta * v[0] / alpha);
    gamma = v[0];
    run = it < maxit && rr > thr;
  }
  // the solution back to cut order
  for (int k = 0; k < ks; ++k) {
    const int i = tid + k * WG;
    t[tcut[i]] = x[i];
  }
  if (tid == 0) {
    nd.cg_iters = it;
    nd.cg_total += it;
  }
}

// Cut duals (admm.c:557-560, 589-591) and the matrix to project
// W = L + X/sigma - Diag(y) - B^T(t) (admm.c:564-568); with use_s the safe
// bound matrix L - Diag(y) - B^T(s) (admm.c:727-729) instead.  Lower
// triangle computed in the reference's order, mirrored to the upper: an
// active pair's entry starts from L (+ X/sigma) and subtracts its cuts'
// terms in cut order, one pair per thread (two at a time); every other
// lower entry (the column pairs (j, n-1-j) of k_post_part, four at a time)
// is L (+ X/sigma), minus y on the diagonal.  The two passes write
// disjoint entries.
__global__ __launch_bounds__(WGS) void k_build_w(NodeView* nodes, int use_s) {
#pragma clang fp contract(off)
  NodeView& nd = nodes[blockIdx.y];
  if (!use_s && nd.done) return;
  const int n = nd.n, m = nd.m, na = nd.nact;
  const double sg = nd.sigma;
  const int gt = blockIdx.x * WGS + threadIdx.x, gs = gridDim.x * WGS;
  const GMEM double* __restrict__ Lm = gmem(nd.L);
  const GMEM double* __restrict__ Xm = gmem(nd.X);
  const GMEM double* __restrict__ cs = gmem(nd.cscale);
  const GMEM int* __restrict__ aptr = gmem(nd.aptr);
  const GMEM int* __restrict__ apair = gmem(nd.apair);
  const GMEM int* __restrict__ p2a = gmem(nd.p2a);
  GMEM double* __restrict__ Wm = gmem(nd.W);
  const GMEM double* __restrict__ vec = gmem(use_s ? nd.s : nd.t);
  if (!use_s) {
    GMEM double* __restrict__ sv = gmem(nd.s);
    GMEM double* __restrict__ uv = gmem(nd.u);
    const GMEM double* __restrict__ tv = gmem(nd.t);
    for (int c = gt; c < m; c += gs) {
      const double v = tv[c] - uv[c] / sg;
      sv[c] = v > 0 ? v : 0;
      uv[c] = v < 0 ? -sg * v : 0;
    }
  }
  // the dual step above must finish before use_s readers run; callers order
  // them with separate launches.
  for (int a0 = gt; a0 < na; a0 += 2 * gs) {
    int lo[2], e0[2], e1[2];
    double w[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int a = a0 + u * gs < na ? a0 + u * gs : a0;
      lo[u] = apair[a];
      e0[u] = aptr[a];
      e1[u] = aptr[a + 1];
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) w[u] = use_s ? Lm[lo[u]] : Lm[lo[u]] + Xm[lo[u]] / sg;
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      if (a0 + u * gs >= na) continue;
      const double wv = signed_sum<true>(nd.aent, e0[u], e1[u], w[u],
                                         [&](int c) { return (0.5 * cs[c]) * vec[c]; });
      const int i = lo[u] % n, j = lo[u] / n;
      Wm[lo[u]] = wv;
      Wm[j + i * n] = wv;
    }
  }
  const GMEM double* __restrict__ yv = gmem(nd.y);
  const int np1 = n + 1, ntot = ((n + 1) >> 1) * np1;
  for (int f0 = gt; f0 < ntot; f0 += 4 * gs) {
    int lo[4], up[4], act[4];
    bool ok[4];
    double l4[4], x4[4], y4[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int f = f0 + u * gs;
      const int pp = f / np1, r = f - pp * np1;
      int i, j;
      if (r < n - pp) {
        j = pp;
        i = pp + r;
      } else {
        j = n - 1 - pp;
        i = j + (r - (n - pp));
      }
      ok[u] = f < ntot && i < n && (r < n - pp || j != pp);
      if (!ok[u]) i = j = 0;
      lo[u] = i + j * n;
      up[u] = j + i * n;
      act[u] = i == j ? -2 : p2a[lo[u]];
      l4[u] = Lm[lo[u]];
      x4[u] = use_s ? 0.0 : Xm[lo[u]];
      y4[u] = yv[i];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (!ok[u] || act[u] >= 0) continue;
      double w = use_s ? l4[u] : l4[u] + x4[u] / sg;
      if (act[u] == -2) w -= y4[u];
      Wm[lo[u]] = w;
      Wm[up[u]] = w;
    }
  }
}

// ------------------------------------------------------ Jacobi eigensolver
//
// Cyclic two-sided Jacobi with round-robin parallel ordering, one workgroup
// per node.  The working matrix A is a packed lower triangle, in LDS when it
// fits (A_LDS) and in the node's S2 scratch otherwise; the eigenvector
// matrix V lives in LDS too when A and V both fit (V_LDS), else in HBM.
// Warm start: when nd.eig_warm is set, A = V^T W V with V the eigenvectors
// of the previous projection (nearly diagonal, one or two sweeps finish it)
// and V accumulates the new rotations; cold start uses A = W, V = I.
// values_only (safe bound) never touches V and always starts cold.
// Sweeps stop when the off-diagonal Frobenius norm is <= 1e-10 |A|_F.
//
// Output: ev[0..n-1] = diag(A) (unsorted), V (unless values_only),
// off_last, sweeps_last; with form_xplus also the projection
// X_+ = V diag(lambda_+) V^T into S1 (full), evp and M, where lambda_+ keeps
// eigenvalues > 1e-8 (the dsyevr range (1e-8, VU] of operators.c:536-579;
// VU = min(|W|_inf, |W|_F) >= lambda_max, so only the lower end matters).
// form_xplus == 3 is form 1 restricted to the nodes whose sign-iteration
// projection failed its guard (nd.ns_fail); the other nodes return at once.

struct PackedOut {
  double* A;
  int n2;
  __device__ void operator()(int i, int j, double v) const {
    if (i >= j) A[pk(i, j, n2)] = v;
  }
};
struct DenseOut {
  double* C;
  int ld;
  __device__ void operator()(int i, int j, double v) const { C[i + (size_t)j * ld] = v; }
};

// Dynamic LDS layout of k_jacobi (in doubles): rc, rs, rt [h each], pair
// table [h doubles = 2h ints], red [64], tA, tB [16*65 each], then A
// (packed, if A_LDS), then V (n*n, if V_LDS).
__host__ __device__ inline size_t jac_fixed_doubles(int h) { return 4 * (size_t)h + 64 + 2 * 16 * 65; }

template <bool A_LDS, bool V_LDS>
__global__ __launch_bounds__(WG) void k_jacobi(NodeView* nodes, int values_only, int max_sweeps,
                                               int ignore_done, int form_xplus, int n_limit) {
  extern __shared__ double smem[];
  NodeView& nd = nodes[blockIdx.x];
  if (!ignore_done && nd.done) return;
  if (nd.n >= n_limit) return;  // solved by the block Jacobi kernels
  if (form_xplus == 3) {
    // fallback behind the sign-iteration projection: only its failed nodes
    if (!nd.ns_fail) return;
    if (threadIdx.x == 0) nd.ns_fallbacks += 1;
    form_xplus = 1;
  }
  const int n = nd.n;
  if (n <= 0) return;
  const int n2 = n + (n & 1), h = n2 >> 1;
  const int tid = threadIdx.x;
  double* rc = smem;
  double* rs = rc + h;
  double* rt = rs + h;
  int* ptab = (int*)(rt + h);  // ptab[k] = p, ptab[h + k] = q
  double* red = rt + 2 * h;
  double* tA = red + 64;
  double* tB = tA + 16 * 65;
  double* lds_tail = tB + 16 * 65;
  const size_t npk = (size_t)n2 * (n2 + 1) / 2;
  double* A = A_LDS ? lds_tail : nd.S2;
  double* V = V_LDS ? (lds_tail + (A_LDS ? npk : 0)) : nd.V;
  const double* __restrict__ Wm = nd.W;
  const bool warm = !values_only && nd.eig_warm;
  const float inv_n = 1.0f / (float)n;

  // ---- form A (and V)
  if (warm) {
    wg_gemm<false, false>(Wm, n, nd.V, n, nullptr, n, n, n, DenseOut{nd.S1, n}, tA, tB);
    __syncthreads();
    wg_gemm<true, false>(nd.V, n, nd.S1, n, nullptr, n, n, n, PackedOut{A, n2}, tA, tB);
    if (V_LDS)
      for (int f = tid; f < n * n; f += WG) V[f] = nd.V[f];
  } else {
    for (int f = tid; f < n * n; f += WG) {
      int j = (int)((f + 0.5f) * inv_n);
      int i = f - j * n;
      if (i < 0) { --j; i += n; } else if (i >= n) { ++j; i -= n; }
      if (i >= j) A[pk(i, j, n2)] = Wm[f];
      if (!values_only) V[f] = (i == j) ? 1.0 : 0.0;
    }
  }
  if (n2 != n)
    for (int j = tid; j < n2; j += WG) A[pk(n, j, n2)] = 0.0;
  __syncthreads();

  // ---- norms over the packed triangle
  double nv[2] = {0.0, 0.0};  // off^2, total^2
  for (int j = 0; j < n; ++j) {
    const int base = pk(j, j, n2);
    for (int i = j + tid; i < n; i += WG) {
      const double a = A[base + (i - j)];
      const double a2 = (i == j) ? a * a : 2.0 * a * a;
      nv[1] += a2;
      if (i != j) nv[0] += a2;
    }
  }
  block_sums<2>(nv, red);
  double off2 = nv[0];
  const double thr = 1e-20 * nv[1];
  const int nblk = h * (h + 1) / 2;
  int sweep = 0;
  for (; sweep < max_sweeps && off2 > thr; ++sweep) {
    for (int r = 0; r < n2 - 1; ++r) {
      for (int k = tid; k < h; k += WG) {
        int p, q;
        rr_pair(r, k, n2, p, q);
        ptab[k] = p;
        ptab[h + k] = q;
        double c = 1.0, s = 0.0, t = 0.0;
        if (p < n && q < n) {
          const double apq = A[pk(p, q, n2)];
          if (apq != 0.0) {
            const double app = A[pk(p, p, n2)], aqq = A[pk(q, q, n2)];
            const double tau = (aqq - app) / (2.0 * apq);
            t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + t * t);
            s = t * c;
          }
        }
        rc[k] = c;
        rs[k] = s;
        rt[k] = t;
      }
      __syncthreads();
      for (int f = tid; f < nblk; f += WG) {
        int k = (int)((sqrtf(8.0f * (float)f + 1.0f) - 1.0f) * 0.5f);
        if ((k + 1) * (k + 2) / 2 <= f) ++k;
        if (k * (k + 1) / 2 > f) --k;
        const int l = f - k * (k + 1) / 2;
        const int p = ptab[k], q = ptab[h + k];
        if (k == l) {
          const int ipq = pk(p, q, n2);
          const double apq = A[ipq], t = rt[k];
          const int ipp = pk(p, p, n2), iqq = pk(q, q, n2);
          A[ipp] = A[ipp] - t * apq;
          A[iqq] = A[iqq] + t * apq;
          A[ipq] = 0.0;
        } else {
          const int p2 = ptab[l], q2 = ptab[h + l];
          const double c1 = rc[k], s1 = rs[k], c2 = rc[l], s2 = rs[l];
          const int i00 = pk(p, p2, n2), i01 = pk(p, q2, n2), i10 = pk(q, p2, n2),
                    i11 = pk(q, q2, n2);
          const double a = A[i00], b = A[i01], cc = A[i10], dd = A[i11];
          const double a1 = c1 * a - s1 * cc, b1 = c1 * b - s1 * dd;
          const double c1v = s1 * a + c1 * cc, d1 = s1 * b + c1 * dd;
          A[i00] = c2 * a1 - s2 * b1;
          A[i01] = s2 * a1 + c2 * b1;
          A[i10] = c2 * c1v - s2 * d1;
          A[i11] = s2 * c1v + c2 * d1;
        }
      }
      if (!values_only)
        for (int f = tid; f < n * h; f += WG) {
          int l = (int)((f + 0.5f) * inv_n);
          int i = f - l * n;
          if (i < 0) { --l; i += n; } else if (i >= n) { ++l; i -= n; }
          const int p = ptab[l], q = ptab[h + l];
          if (p >= n || q >= n) continue;
          const double c = rc[l], s = rs[l];
          const double vp = V[i + (size_t)p * n], vq = V[i + (size_t)q * n];
          V[i + (size_t)p * n] = c * vp - s * vq;
          V[i + (size_t)q * n] = s * vp + c * vq;
        }
      __syncthreads();
    }
    double ov[1] = {0.0};
    for (int j = 0; j < n; ++j) {
      const int base = pk(j, j, n2);
      for (int i = j + 1 + tid; i < n; i += WG) {
        const double a = A[base + (i - j)];
        ov[0] += 2.0 * a * a;
      }
    }
    block_sums<1>(ov, red);
    off2 = ov[0];
  }
  for (int i = tid; i < n; i += WG) nd.ev[i] = A[pk(i, i, n2)];
  if (tid == 0) {
    nd.off_last = sqrt(off2);
    nd.sweeps_last = sweep;
    nd.sweeps_total += sweep;
    if (!values_only) nd.eig_warm = 1;
  }
  if (!values_only && form_xplus) {
    // form_xplus == 2: only lambda_+ (evp) and M, no X_+
    // lambda_+ into the rotation arrays (3h >= n doubles, free after the sweeps)
    double* dpos = rc;
    __syncthreads();
    double mv[1] = {0.0};
    for (int j = tid; j < n; j += WG) {
      const double lam = A[pk(j, j, n2)];
      const double d = lam > 1e-8 ? lam : 0.0;
      dpos[j] = d;
      nd.evp[j] = d;
      mv[0] += d > 0.0 ? 1.0 : 0.0;
    }
    block_sums<1>(mv, red);
    if (tid == 0) nd.M = (int)mv[0];
    if (form_xplus == 2) {
      // eigen-pairs only
    } else if (V_LDS) {
      // register-blocked X_+ (4x4 per thread) straight from LDS V
      const int nb4 = (n + 3) >> 2;
      const int nbl = nb4 * (nb4 + 1) / 2;
      for (int f = tid; f < nbl; f += WG) {
        int bi = (int)((sqrtf(8.0f * (float)f + 1.0f) - 1.0f) * 0.5f);
        if ((bi + 1) * (bi + 2) / 2 <= f) ++bi;
        if (bi * (bi + 1) / 2 > f) --bi;
        const int bj = f - bi * (bi + 1) / 2;
        const int i0 = bi * 4, j0 = bj * 4;
        double acc[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
        for (int kk = 0; kk < n; ++kk) {
          const double dk = dpos[kk];
          if (dk == 0.0) continue;
          const double* col = V + (size_t)kk * n;
          double vi[4], vj[4];
#pragma unroll
          for (int a = 0; a < 4; ++a) {
            vi[a] = (i0 + a < n) ? col[i0 + a] * dk : 0.0;
            vj[a] = (j0 + a < n) ? col[j0 + a] : 0.0;
          }
#pragma unroll
          for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = fma(vi[a], vj[b], acc[a][b]);
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) {
            const int i = i0 + a, j = j0 + b;
            if (i < n && j < n && i >= j) {
              nd.S1[i + (size_t)j * n] = acc[a][b];
              nd.S1[j + (size_t)i * n] = acc[a][b];
            }
          }
      }
    } else {
      wg_gemm<false, true>(V, n, V, n, dpos, n, n, n, DenseOut{nd.S1, n}, tA, tB);
    }
  }
  if (V_LDS && !values_only) {
    __syncthreads();
    for (int f = tid; f < n * n; f += WG) nd.V[f] = V[f];
  }
}

// After the projection formed X_+ in S1: X = sigma X_+, Q = X_+ - W and
// the residuals, sigma update and stop test of admm.c:580-702, in two
// launches: k_post_part spreads the n^2 and cut
// passes over a (chunk, node) grid and leaves per-workgroup partial sums
// (B(X) reads sigma X_+ from S1, the value X receives, so no workgroup
// waits for another's writes); k_post_fin adds them in workgroup order and
// updates the node's scalars.
__global__ __launch_bounds__(WGS) void k_post_part(NodeView* nodes, double* part) {
  __shared__ double red[8 * 16];
  NodeView& nd = nodes[blockIdx.y];
  if (nd.done) return;
  const int n = nd.n, m = nd.m, tid = threadIdx.x;
  const int g = blockIdx.x, G = gridDim.x, stride = G * WGS;
  const double sg = nd.sigma;
  const GMEM double* __restrict__ S1 = gmem(nd.S1);
  const GMEM double* __restrict__ Wm = gmem(nd.W);
  const GMEM double* __restrict__ Lm = gmem(nd.L);
  GMEM double* __restrict__ Xm = gmem(nd.X);
  GMEM double* __restrict__ Qm = gmem(nd.Q);
  double v[7] = {0, 0, 0, 0, 0, 0, 0};  // |X+ - Xold/sigma|^2, |diag X - 1|^2, <L,X>,
                                        // |max(B(X)-1,0)|^2, |s-t|^2, sum t, sum y
  // the lower triangle as ceil(n/2) column pairs (j, n-1-j) of n+1 entries
  // each (rows j..n-1 of column j, then rows n-1-j..n-1 of column n-1-j), so
  // that every index is an entry; four entries' operands are loaded before
  // any store (distinct entries: no store reaches another's loads)
  const int np1 = n + 1, ntot = ((n + 1) >> 1) * np1;
  for (int f0 = g * WGS + tid; f0 < ntot; f0 += 4 * stride) {
    int lo[4], up[4];
    bool ok[4], dg[4];
    double xp[4], xo[4], wv[4], lv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int f = f0 + u * stride;
      const int pp = f / np1, r = f - pp * np1;
      int i, j;
      if (r < n - pp) {
        j = pp;
        i = pp + r;
      } else {
        j = n - 1 - pp;
        i = j + (r - (n - pp));
      }
      ok[u] = f < ntot && i < n && (r < n - pp || j != pp);
      if (!ok[u]) i = j = 0;
      lo[u] = i + j * n;
      up[u] = j + i * n;
      dg[u] = i == j;
      xp[u] = S1[lo[u]];
      xo[u] = Xm[lo[u]];
      wv[u] = Wm[lo[u]];
      lv[u] = Lm[lo[u]];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (!ok[u]) continue;
      const double xn = sg * xp[u];
      const double dd = xp[u] - xo[u] / sg;
      Xm[lo[u]] = xn;
      Xm[up[u]] = xn;
      const double qv = xp[u] - wv[u];
      Qm[lo[u]] = qv;
      Qm[up[u]] = qv;
      if (dg[u]) {
        v[0] += dd * dd;
        v[1] += (xn - 1.0) * (xn - 1.0);
        v[2] += lv[u] * xn;
      } else {
        v[0] += 2.0 * dd * dd;
        v[2] += 2.0 * lv[u] * xn;
      }
    }
  }
  for (int c = g * WGS + tid; c < m; c += stride) {
    const double acc =
        signed_sum<false>(nd.cent, nd.cptr[c], nd.cptr[c + 1], 0.0,
                          [&](int lo) { return sg * nd.S1[lo]; });
    const double bx = -1.0 + nd.cscale[c] * acc;
    if (bx > 0) v[3] += bx * bx;
    const double ds = nd.s[c] - nd.t[c];
    v[4] += ds * ds;
    v[5] += nd.t[c];
  }
  if (g == 0)
    for (int i = tid; i < n; i += WGS) v[6] += nd.y[i];
  block_sums<7>(v, red);
  if (tid == 0)
    for (int q = 0; q < 7; ++q) part[((size_t)blockIdx.y * 64 + g) * 8 + q] = v[q];
}

__global__ __launch_bounds__(64) void k_post_fin(NodeView* nodes, const double* part, int G) {
  __shared__ double pl[64 * 8];
  NodeView& nd = nodes[blockIdx.x];
  if (nd.done) return;
  // the partials into LDS by all lanes at once, then summed in workgroup
  // order by lane 0
  const GMEM double* pg = gmem(part) + (size_t)blockIdx.x * 64 * 8;
  for (int f = threadIdx.x; f < G * 8; f += 64) pl[f] = pg[f];
  __syncthreads();
  if (threadIdx.x != 0) return;
  double v[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int g = 0; g < G; ++g)
    for (int q = 0; q < 7; ++q) v[q] += pl[g * 8 + q];
  const int n = nd.n;
  const double sg = nd.sigma;
  nd.sigma_last = sg;
  const double err_p = (sqrt(v[1]) + sqrt(v[3])) / (1.0 + sqrt((double)n));
  const double err_d = (sqrt(v[4]) + sqrt(v[0])) / (1.0 + nd.normL);
  const double primal = v[2];
  const double dual = v[6] + v[5];
  const double gap = fabs(dual - primal) / (1.0 + fabs(dual) + fabs(primal));
  const double err = err_p > err_d ? (err_p > gap ? err_p : gap) : (err_d > gap ? err_d : gap);
  const double dgp = log10(err_p), dgd = log10(err_d);
  if (dgd - dgp < -0.5)
    nd.sigma = sg / nd.scaleSigma;
  else if (dgd - dgp > 0.5)
    nd.sigma = sg * nd.scaleSigma;
  nd.err_p = err_p;
  nd.err_d = err_d;
  nd.gap = gap;
  nd.err = err;
  nd.primal = primal;
  nd.dual = dual;
  nd.nbit += 1;
  if (err < nd.tol || nd.nbit > nd.itermax) nd.done = 1;
}

// f = n * lambda_max(L - Diag(y) - B^T(s)) + sum y + sum s (admm.c:757-763),
// with lambda_max raised by the Jacobi residual bound so that it is never
// below the true eigenvalue: eigenvalues of D + E lie within |E|_2 <= |E|_F
// of diag(D), plus a rounding allowance of 1e-12 |A|_F.
__global__ __launch_bounds__(WG) void k_safe_bound(NodeView* nodes) {
  __shared__ double red[64];
  NodeView& nd = nodes[blockIdx.x];
  const int n = nd.n, m = nd.m, tid = threadIdx.x;
  double mx = -1e300;
  for (int i = tid; i < n; i += WG) mx = fmax(mx, nd.ev[i]);
  mx = block_max(mx, red);
  double v3[3] = {0, 0, 0};  // sum y, sum s, |A|_F^2
  for (int i = tid; i < n; i += WG) v3[0] += nd.y[i];
  for (int c = tid; c < m; c += WG) v3[1] += nd.s[c];
  for (int f = tid; f < n * n; f += WG) v3[2] += nd.W[f] * nd.W[f];
  block_sums<3>(v3, red);
  if (tid == 0) {
    const double margin = nd.off_last + 1e-12 * sqrt(v3[2]);
    nd.lam = mx + margin;
    nd.lam_margin = margin;
    double f = n * nd.lam;
    f += v3[0];
    f += v3[1];
    nd.f = f;
  }
}

// ------------------------------------------------------ standalone maps

// yout[c] = B(X)_c for node 0's cut set; X dense n x n (op_B,
// operators.c:62-150).
__global__ __launch_bounds__(WGS) void k_op_B(const NodeView* nodes, const double* X,
                                             double* yout) {
#pragma clang fp contract(off)
  const NodeView& nd = nodes[0];
  for (int c = blockIdx.x * WGS + threadIdx.x; c < nd.m; c += gridDim.x * WGS) {
    double acc = 0.0;
    for (int e = nd.cptr[c]; e < nd.cptr[c + 1]; ++e) {
      const int en = nd.cent[e];
      const double v = X[ent_idx(en)];
      acc = en >= 0 ? acc + v : acc - v;
    }
    yout[c] = nd.cscale[c] == 1.0 ? acc : nd.cscale[c] * acc;
  }
}

// Xout = -B^T(t) accumulated from zero in the reference's order (op_Bt,
// operators.c:171-513, applied to a zero matrix).
__global__ __launch_bounds__(WGS) void k_op_Bt(const NodeView* nodes, const double* tin,
                                              double* Xout) {
#pragma clang fp contract(off)
  const NodeView& nd = nodes[0];
  const int n = nd.n;
  for (int f = blockIdx.x * WGS + threadIdx.x; f < n * n; f += gridDim.x * WGS) {
    const int i = f % n, j = f / n;
    if (i < j) continue;
    double w = 0.0;
    if (i != j) {
      const int a = nd.p2a[i + j * n];
      if (a >= 0)
        for (int e = nd.aptr[a]; e < nd.aptr[a + 1]; ++e) {
          const int en = nd.aent[e], c = ent_idx(en);
          const double temp = (0.5 * nd.cscale[c]) * tin[c];
          w = en >= 0 ? w - temp : w + temp;
        }
    }
    Xout[i + j * n] = w;
    Xout[j + i * n] = w;
  }
}

// Projection only (project_SDP, operators.c:521-619): after k_jacobi formed
// X_+ in S1, X = X_+ and Q = X_+ - W (no sigma scaling).
__global__ __launch_bounds__(WGS) void k_project_out(NodeView* nodes) {
  NodeView& nd = nodes[blockIdx.x];
  const int n = nd.n;
  for (int f = threadIdx.x; f < n * n; f += WGS) {
    const double xp = nd.S1[f];
    nd.X[f] = xp;
    nd.Q[f] = xp - nd.W[f];
  }
}

// ------------------------------------------------------------ launchers

static int chunks_for(int work, int per) {
  int c = (work + per - 1) / per;
  if (c < 1) c = 1;
  if (c > 64) c = 64;
  return c;
}

// Batched transfer pieces (staging buffer <-> node buffers): 16-byte
// vector copies where both ends are 16-byte aligned, bytes otherwise.
__global__ __launch_bounds__(256) void k_xfer(const XDesc* __restrict__ d, int nd) {
  typedef unsigned int u4 __attribute__((ext_vector_type(4)));
  for (int q = blockIdx.x; q < nd; q += gridDim.x) {
    const XDesc x = d[q];
    const size_t nb = (size_t)x.bytes;
    size_t done = 0;
    if ((((uintptr_t)x.src | (uintptr_t)x.dst) & 15) == 0) {
      const size_t n16 = nb >> 4;
      u4* o = (u4*)x.dst;
      const u4* in = (const u4*)x.src;
      for (size_t i = threadIdx.x; i < n16; i += blockDim.x)
        o[i] = in ? in[i] : u4{0u, 0u, 0u, 0u};
      done = n16 << 4;
    }
    for (size_t i = done + threadIdx.x; i < nb; i += blockDim.x)
      x.dst[i] = x.src ? x.src[i] : (char)0;
  }
}

void launch_xfer(const XDesc* d, int nd, hipStream_t st) {
  if (nd <= 0) return;
  k_xfer<<<std::min(nd, 4096), 256, 0, st>>>(d, nd);
}

__global__ __launch_bounds__(256) void k_p2a(const P2aJob* __restrict__ jobs) {
  const P2aJob J = jobs[blockIdx.x];
  for (int f = threadIdx.x; f < J.nn; f += blockDim.x) J.p2a[f] = -1;
  __syncthreads();
  for (int a = threadIdx.x; a < J.nact; a += blockDim.x) J.p2a[J.apair[a]] = a;
}

void launch_p2a(const P2aJob* jobs, int njobs, hipStream_t st) {
  if (njobs > 0) k_p2a<<<njobs, 256, 0, st>>>(jobs);
}

void launch_admm_prep(NodeView* nodes, int nb, int mmax, hipStream_t st) {
  // one cut per thread: the per-cut gathers are latency chains
  dim3 grid(chunks_for(mmax, WGS), nb);
  k_admm_prep<<<grid, WGS, 0, st>>>(nodes);
}

CgPlan plan_cg(const NodeView* hv, int nb, size_t lds_limit, int use_x) {
  // per class: is any node in it, and the LDS its largest node asks for
  // (vectors, plus the staged index streams when those fit)
  CgPlan P;
  for (int c = 0; c < 6; ++c) {
    P.need[c] = 0;
    P.any[c] = false;
  }
  for (int b = 0; b < nb; ++b) {
    const NodeView& v = hv[b];
    if (v.done || v.m == 0) continue;
    const int cls = cg_class(v.m, v.nact, use_x ? v.xrows : 0, lds_limit);
    P.any[cls] = true;
    // LDS: the cut vector (and the pair vector where it lives there) and 64
    // reduction doubles, plus for k_cg1 the index streams when they fit
    const size_t base = cls == 2 ? (2 * (size_t)v.m + 66) * sizeof(double)
                                 : ((size_t)v.m + (cls == 4 ? 0 : v.nact) + 64) * sizeof(double);
    size_t staged = base;
    if (cls <= 1)
      staged += ((size_t)v.nact + 1) * sizeof(int) +
                2 * (size_t)v.ne * (v.aent16 ? sizeof(short) : sizeof(int));
    const size_t want = staged <= lds_limit ? staged : base;
    if (cls < 5 && want > P.need[cls]) P.need[cls] = want;
  }
  return P;
}

void launch_cg_plan(NodeView* nodes, int nb, const CgPlan& P, double rtol, double adapt,
                    int maxit, size_t lds_limit, int use_x, hipStream_t st) {
  const size_t* need = P.need;
  const bool* any = P.any;
  const int ux = use_x ? 1 : 0;
  if (any[0])
    k_cg1<2><<<nb, WG, need[0], st>>>(nodes, rtol, adapt, maxit, lds_limit, need[0], ux);
  if (any[1])
    k_cg1<4><<<nb, WG, need[1], st>>>(nodes, rtol, adapt, maxit, lds_limit, need[1], ux);
  // k_cg_x runs only with use_x (any[2] is false otherwise)
  if (any[2]) k_cg_x<<<nb, WG, need[2], st>>>(nodes, rtol, adapt, maxit, lds_limit);
  if (any[3]) k_cg_big<true><<<nb, WG, need[3], st>>>(nodes, rtol, adapt, maxit, lds_limit, ux);
  if (any[4]) k_cg_big<false><<<nb, WG, need[4], st>>>(nodes, rtol, adapt, maxit, lds_limit, ux);
  if (any[5]) k_cg<<<nb, WG, 0, st>>>(nodes, rtol, adapt, maxit, lds_limit, ux);
}

void launch_cg(NodeView* nodes, const NodeView* hv, int nb, double rtol, double adapt, int maxit,
               size_t lds_limit, int use_x, hipStream_t st) {
  launch_cg_plan(nodes, nb, plan_cg(hv, nb, lds_limit, use_x), rtol, adapt, maxit, lds_limit,
                 use_x, st);
}

void launch_build_w(NodeView* nodes, int nb, int nmax, int mmax, int use_s, hipStream_t st) {
  int work = nmax * nmax / 2 > mmax ? nmax * nmax / 2 : mmax;
  // one lower entry per thread (up to 64 chunks a node): each is a chain of
  // dependent gathers, so a lone node finishes sooner with more threads
  dim3 grid(chunks_for(work, WGS), nb);
  k_build_w<<<grid, WGS, 0, st>>>(nodes, use_s);
}

size_t jacobi_lds_bytes(int nmax, bool with_v) {
  const int n2 = nmax + (nmax & 1), h = n2 / 2;
  size_t d = jac_fixed_doubles(h) + (size_t)n2 * (n2 + 1) / 2;
  if (with_v) d += (size_t)nmax * nmax;
  return d * sizeof(double);
}

void launch_jacobi(NodeView* nodes, int nb, int nmax, int values_only, int max_sweeps,
                   int ignore_done, size_t lds_limit, hipStream_t st, int form_xplus,
                   int n_limit) {
  const int h = (nmax + (nmax & 1)) / 2;
  const size_t fixed = jac_fixed_doubles(h) * sizeof(double);
  if (!values_only && jacobi_lds_bytes(nmax, true) <= lds_limit) {
    k_jacobi<true, true><<<nb, WG, jacobi_lds_bytes(nmax, true), st>>>(
        nodes, values_only, max_sweeps, ignore_done, form_xplus, n_limit);
  } else if (jacobi_lds_bytes(nmax, false) <= lds_limit) {
    k_jacobi<true, false><<<nb, WG, jacobi_lds_bytes(nmax, false), st>>>(
        nodes, values_only, max_sweeps, ignore_done, form_xplus, n_limit);
  } else {
    k_jacobi<false, false><<<nb, WG, fixed, st>>>(nodes, values_only, max_sweeps, ignore_done,
                                                  form_xplus, n_limit);
  }
}

void launch_admm_post(NodeView* nodes, int nb, int nmax, int mmax, double* part,
                      hipStream_t st) {
  const int work = nmax * nmax / 2 > mmax ? nmax * nmax / 2 : mmax;
  const int G = chunks_for(work, WGS);  // <= 64
  k_post_part<<<dim3(G, nb), WGS, 0, st>>>(nodes, part);
  k_post_fin<<<nb, 64, 0, st>>>(nodes, part, G);
}

void launch_safe_bound(NodeView* nodes, int nb, hipStream_t st) {
  k_safe_bound<<<nb, WG, 0, st>>>(nodes);
}

void launch_op_B(const NodeView* nodes, int nb, const double* Xin, double* yout, int mmax,
                 hipStream_t st) {
  (void)nb;
  k_op_B<<<chunks_for(mmax, WGS), WGS, 0, st>>>(nodes, Xin, yout);
}

void launch_op_Bt(const NodeView* nodes, int nb, const double* tin, double* Xout, int nmax,
                  hipStream_t st) {
  (void)nb;
  k_op_Bt<<<chunks_for(nmax * nmax, WGS), WGS, 0, st>>>(nodes, tin, Xout);
}

void launch_project(NodeView* nodes, int nb, int nmax, hipStream_t st) {
  (void)nmax;
  k_project_out<<<nb, WGS, 0, st>>>(nodes);
}

}  // namespace madam
