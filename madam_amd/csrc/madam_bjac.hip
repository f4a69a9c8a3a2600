// madam_bjac.hip -- FP64 symmetric eigen-solver of the plain projection
// (project_SDP, operators.c:521-619, whose dsyevr call is operators.c:579)
// and of the plain safe bound's eigenvalues (admm.c:727-747), for orders
// where the one-workgroup Jacobi kernel is slow: a two-sided block Jacobi
// method spread over many workgroups per node.
//
// The matrix is padded to n_p = 16 nbk (nbk even; padding rows and columns
// zero, which no rotation couples to the others) and split into nbk blocks
// of 16.  A sweep is nbk - 1 rounds of the round-robin pairing of the
// blocks; in a round every pair (P, Q) forms the 32 x 32 subproblem
// A[P u Q, P u Q], to which one workgroup applies one sweep of cyclic scalar
// Jacobi (k_bj_solve, G = its accumulated rotation), and then every
// off-diagonal 32 x 32 block pair and the eigenvector rows are transformed
// on FP64 MFMA (k_bj_update: A_IK <- G_I^T A_IK G_K, V_I <- V_I G_I).
// Sweeps run until the off-diagonal sum of squares is at most 1e-24 of
// |A|_F^2.
//
// Warm start: where the node's V holds the eigenvectors of the previous
// projection (eig_warm), the sweeps start from A = V^T W V (two FP64-MFMA
// products) with V as the basis, so that a slowly changing W needs two or
// three sweeps.  The result is W_+ = V diag(lambda_+) V^T (eigenvalues
// above 1e-8, as the reference keeps) in S1, the eigenvalues in ev and the
// eigenvectors in V -- what the one-workgroup Jacobi kernel leaves.
#include "madam_devutil.h"
#include "madam_engine.h"

#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>
#include <vector>

namespace madam {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int BJ_B = 16;            // block width
constexpr int BJ_S = 2 * BJ_B;      // subproblem order
constexpr int BJ_LD = BJ_S + 1;     // LDS row stride of a subproblem
// stop-test partial sums of one node: nA + npair for nbk <= 64 (n <= 1024)
constexpr int BJ_PART_MAX = 32 * 31 / 2 + 32;

struct BjJob {
  double* A;     // n_p x n_p, ld n_p
  double* V;     // n_p x n_p, ld n_p (not used for values only)
  double* G;     // nbk/2 rotations of 32 x 32, column-major
  double* part;  // off-diagonal partial sums, one per block-pair task
  double* part2; // off-diagonal sums of squares of the subproblems, one per pair
  int node;      // index of the node in the NodeView array
  int n, np, nbk;
  int done;      // converged, or a node that has stopped
  int sweeps;
  int values_only;
  int inner;     // inner sweeps summed over this call's subproblem solves
  int skip;      // order below the block Jacobi range: left to k_jacobi
  double off2;   // off-diagonal sum of squares after the latest sweep
  double fro2;
};

__device__ __forceinline__ void bj_pair(int r, int k, int nb, int& p, int& q) {
  rr_pair(r, k, nb, p, q);
}

// The Jacobi rotation J = [c s; -s c] on (p, q) that annihilates a_pq
// (t = s / c = sgn(tau) / (|tau| + sqrt(1 + tau^2)), tau = (a_qq - a_pp) /
// (2 a_pq)), with reciprocals and reciprocal square roots from the hardware
// estimates (v_rcp_f64, v_rsq_f64) refined by Newton steps; c = 1/sqrt(1+t^2)
// to the last bits so that c^2 + s^2 = 1 (the angle itself need only be
// close: the next sweep removes what an inexact angle leaves).
__device__ __forceinline__ double bj_rcp(double x) {
  double y = __builtin_amdgcn_rcp(x);
  y = y * fma(-x, y, 2.0);
  return y * fma(-x, y, 2.0);
}
__device__ __forceinline__ double bj_rsq(double x) {
  double y = __builtin_amdgcn_rsq(x);
  y = y * fma(-0.5 * x * y, y, 1.5);
  return y * fma(-0.5 * x * y, y, 1.5);
}
__device__ __forceinline__ void bj_rot(double app, double aqq, double apq, double& c, double& s,
                                       double& t) {
  c = 1.0;
  s = 0.0;
  t = 0.0;
  // an entry this small is dropped (its reciprocal could overflow)
  if (fabs(apq) < 1e-290) return;
  const double tau = (aqq - app) * bj_rcp(2.0 * apq);
  const double at = fabs(tau);
  double ta;
  if (at > 1e150) {
    ta = 0.5 * bj_rcp(at);  // sqrt(1 + tau^2) = |tau| to working precision
  } else {
    const double u = fma(tau, tau, 1.0);
    ta = bj_rcp(at + u * bj_rsq(u));
  }
  t = tau >= 0.0 ? ta : -ta;
  c = bj_rsq(fma(t, t, 1.0));
  s = t * c;
}

// row of the 32 x 32 subproblem of blocks (P, Q) -> row of A
__device__ __forceinline__ int bj_row(int i, int P, int Q) {
  return i < BJ_B ? P * BJ_B + i : Q * BJ_B + (i - BJ_B);
}

// f -> (k1, k2), k1 >= k2, f = k1(k1+1)/2 + k2 (f < 136: float sqrt exact)
__device__ __forceinline__ void bj_tri(int f, int& k1, int& k2) {
  int a = (int)((sqrtf(8.0f * (float)f + 1.0f) - 1.0f) * 0.5f);
  if ((a + 1) * (a + 2) / 2 <= f) ++a;
  if (a * (a + 1) / 2 > f) --a;
  k1 = a;
  k2 = f - a * (a + 1) / 2;
}

// ---------------------------------------------------------------- set-up
// Cold nodes (and values only): A = W padded, V = I.  Warm nodes: the
// padding of A (the products write A[0:n, 0:n]) and V = [V_prev 0; 0 I].
__global__ __launch_bounds__(256) void k_bj_setup(BjJob* jobs, NodeView* nodes, int ignore_done) {
  BjJob& J = jobs[blockIdx.y];
  if (J.skip) return;
  NodeView& nd = nodes[J.node];
  if (!ignore_done && nd.done) {
    if (blockIdx.x == 0 && threadIdx.x == 0) J.done = 1;
    return;
  }
  const int n = J.n, np = J.np;
  const bool warm = !J.values_only && nd.eig_warm;
  const size_t tot = (size_t)np * np;
  for (size_t f = (size_t)blockIdx.x * 256 + threadIdx.x; f < tot; f += (size_t)gridDim.x * 256) {
    const int j = (int)(f / np), i = (int)(f - (size_t)j * np);
    const bool in = i < n && j < n;
    if (!warm || !in) J.A[f] = in ? nd.W[i + (size_t)j * n] : 0.0;
    if (!J.values_only) J.V[f] = (warm && in) ? nd.V[i + (size_t)j * n] : (i == j ? 1.0 : 0.0);
  }
}

// ------------------------------------------------------- FP64-MFMA products
// One 16 x 16 output tile per workgroup, K split into four contiguous
// quarters (one per wave) summed as ((q0 + q1) + q2) + q3.
//   which 0: S2 = W V              (all tiles; warm nodes)
//   which 1: A[0:n,0:n] = V^T S2   (lower tiles, mirrored; warm nodes)
//   which 2: S1 = V_p diag(evp) V_p^T over the first n rows (lower, mirrored)
// FP64 MFMA lane layout: A operand lane l -> (row l%16, k l/16), B operand
// lane l -> (k l/16, col l%16), result register v -> (row l/16 + 4v, col l%16).
__device__ __forceinline__ void bj_tile_lower(int t, int& i0, int& j0) {
  int ti = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
  if ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
  if (ti * (ti + 1) / 2 > t) --ti;
  i0 = ti << 4;
  j0 = (t - ti * (ti + 1) / 2) << 4;
}

__global__ __launch_bounds__(256) void k_bj_prod(BjJob* jobs, NodeView* nodes, int which,
                                                 int ignore_done) {
  __shared__ f64x4 part[3][64];
  BjJob& J = jobs[blockIdx.y];
  NodeView& nd = nodes[J.node];
  if (which < 2) {
    if (J.done || J.values_only || !nd.eig_warm) return;
  } else if (J.skip || J.values_only || (!ignore_done && nd.done)) {
    return;
  }
  const int n = J.n, np = J.np, nt = (n + 15) >> 4;
  const int t = blockIdx.x;
  int i0, j0;
  if (which == 0) {
    if (t >= nt * nt) return;
    i0 = (t / nt) << 4;
    j0 = (t % nt) << 4;
  } else {
    if (t >= nt * (nt + 1) / 2) return;
    bj_tile_lower(t, i0, j0);
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
  const int kq = (((n + 3) >> 2) + 3) >> 2 << 2;  // k per quarter, a multiple of 4
  const int k0 = w * kq, k1 = min(n, k0 + kq);
  const int ia = i0 + r, jb = j0 + r;
  const int iac = min(ia, n - 1), jbc = min(jb, n - 1);
  const GMEM double* __restrict__ Wm = gmem(nd.W);
  const GMEM double* __restrict__ Vn = gmem(nd.V);
  const GMEM double* __restrict__ S2 = gmem(nd.S2);
  const GMEM double* __restrict__ Vp = gmem(J.V);
  const GMEM double* __restrict__ dp = gmem(nd.evp);
  f64x4 acc = {0.0, 0.0, 0.0, 0.0};
  for (int k = k0; k < k1; k += 32) {
    double a[8], b[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int kk = k + 4 * u + q, kc = min(kk, n - 1);
      double av, bv;
      if (which == 0) {
        av = Wm[iac + (size_t)kc * n];
        bv = Vn[kc + (size_t)jbc * n];
      } else if (which == 1) {
        av = Vn[kc + (size_t)iac * n];
        bv = S2[kc + (size_t)jbc * n];
      } else {
        av = Vp[iac + (size_t)kc * np] * dp[kc];
        bv = Vp[jbc + (size_t)kc * np];
      }
      const bool ok = kk < k1;
      a[u] = (ok && ia < n) ? av : 0.0;
      b[u] = (ok && jb < n) ? bv : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[u], acc, 0, 0, 0);
  }
  if (w > 0) part[w - 1][lane] = acc;
  __syncthreads();
  if (w != 0) return;
  acc = ((acc + part[0][lane]) + part[1][lane]) + part[2][lane];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int i = i0 + q + 4 * v, j = j0 + r;
    if (i >= n || j >= n) continue;
    if (which == 0) {
      nd.S2[i + (size_t)j * n] = acc[v];
    } else if (which == 1) {
      if (i >= j) {
        J.A[i + (size_t)j * np] = acc[v];
        J.A[j + (size_t)i * np] = acc[v];
      }
    } else if (i >= j) {
      nd.S1[i + (size_t)j * n] = acc[v];
      nd.S1[j + (size_t)i * n] = acc[v];
    }
  }
}

// --------------------------------------------------------- subproblems
// One workgroup (four waves) per block pair of round `round`: load the
// 32 x 32 subproblem, apply one cyclic Jacobi sweep to it (31 rounds of 16
// disjoint rotations, the pairs from a round-robin table in LDS; none when
// it is diagonal already), write it back to A[P u Q, P u Q] whole (what the
// sweep leaves off the diagonal is removed by later rounds and sweeps: one
// inner sweep needs no more outer sweeps than a converged inner solve) and
// the accumulated rotation G (S_after = G^T S G) for k_bj_update.  In the
// last round of an outer sweep its off-diagonal sum of squares goes to
// part2 for the stop test.  Per round,
// threads 0..15 form the rotations, then thread x < 136 updates one 2 x 2
// block of S (rotation pair k1 >= k2) and every thread two (row, pair)
// entries of G.
constexpr int BJ_SOLVE_WG = 256;

__device__ __forceinline__ double bj_block_sum(double v, double* red) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return ((red[0] + red[1]) + red[2]) + red[3];
}

__global__ __launch_bounds__(BJ_SOLVE_WG) void k_bj_solve(BjJob* jobs, int round) {
  __shared__ double S[BJ_S * BJ_LD];
  __shared__ double Gl[BJ_S * BJ_LD];
  __shared__ double rc[16], rs[16], rt[16], red[4];
  __shared__ unsigned char rrp[(BJ_S - 1) * 16], rrq[(BJ_S - 1) * 16];
  BjJob& J = jobs[blockIdx.y];
  if (J.done) return;
  const int nbk = J.nbk, k = blockIdx.x;
  if (round >= nbk - 1 || k >= nbk / 2) return;
  const int tid = threadIdx.x, np = J.np;
  int P, Q;
  bj_pair(round, k, nbk, P, Q);
  GMEM double* __restrict__ A = gmem(J.A);
  for (int f = tid; f < (BJ_S - 1) * 16; f += BJ_SOLVE_WG) {
    int p, q;
    rr_pair(f >> 4, f & 15, BJ_S, p, q);
    rrp[f] = (unsigned char)p;
    rrq[f] = (unsigned char)q;
  }
  double nrm = 0.0, off = 0.0;
  for (int f = tid; f < BJ_S * BJ_S; f += BJ_SOLVE_WG) {
    const int i = f & (BJ_S - 1), j = f >> 5;
    const double v = A[bj_row(i, P, Q) + (size_t)bj_row(j, P, Q) * np];
    S[i * BJ_LD + j] = v;
    Gl[i * BJ_LD + j] = i == j ? 1.0 : 0.0;
    nrm += v * v;
    if (i != j) off += v * v;
  }
  nrm = bj_block_sum(nrm, red);
  off = bj_block_sum(off, red);
  // this thread's fixed tasks: block (k1, k2) of S, entries (gi, gk0/gk1) of G
  const bool sb = tid < 136;
  int k1 = 0, k2 = 0;
  bj_tri(sb ? tid : 0, k1, k2);
  const bool dg = k1 == k2;
  const int gi = tid & (BJ_S - 1), gk0 = tid >> 5, gk1 = gk0 + 8;
  // one sweep: what it leaves stays in A (written back whole) for the
  // rounds and sweeps that follow
  for (int sw = 0; sw < 1 && off > 1e-30 * nrm; ++sw) {
    for (int r = 0; r < BJ_S - 1; ++r) {
      const unsigned char* tp = rrp + r * 16;
      const unsigned char* tq = rrq + r * 16;
      if (tid < 16) {
        const int p = tp[tid], q = tq[tid];
        double c, sn, t;
        bj_rot(S[p * BJ_LD + p], S[q * BJ_LD + q], S[p * BJ_LD + q], c, sn, t);
        rc[tid] = c;
        rs[tid] = sn;
        rt[tid] = t;
      }
      __syncthreads();
      // S <- J^T S J on block (k1, k2), J = [c s; -s c] on each pair (p, q);
      // both triangles written
      if (sb) {
        const int p1 = tp[k1], q1 = tq[k1], p2 = tp[k2], q2 = tq[k2];
        const double a = S[p1 * BJ_LD + p2], b = S[p1 * BJ_LD + q2];
        const double cc = S[q1 * BJ_LD + p2], dd = S[q1 * BJ_LD + q2];
        if (dg) {
          // (p, p) -= t a_pq, (q, q) += t a_pq, (p, q) = 0
          const double t = rt[k1];
          S[p1 * BJ_LD + p1] = a - t * b;
          S[q1 * BJ_LD + q1] = dd + t * b;
          S[p1 * BJ_LD + q1] = 0.0;
          S[q1 * BJ_LD + p1] = 0.0;
        } else {
          const double c1 = rc[k1], s1 = rs[k1], c2 = rc[k2], s2 = rs[k2];
          const double a1 = c1 * a - s1 * cc, b1 = c1 * b - s1 * dd;
          const double c1v = s1 * a + c1 * cc, d1 = s1 * b + c1 * dd;
          const double n00 = c2 * a1 - s2 * b1, n01 = s2 * a1 + c2 * b1;
          const double n10 = c2 * c1v - s2 * d1, n11 = s2 * c1v + c2 * d1;
          S[p1 * BJ_LD + p2] = n00;
          S[p1 * BJ_LD + q2] = n01;
          S[q1 * BJ_LD + p2] = n10;
          S[q1 * BJ_LD + q2] = n11;
          S[p2 * BJ_LD + p1] = n00;
          S[q2 * BJ_LD + p1] = n01;
          S[p2 * BJ_LD + q1] = n10;
          S[q2 * BJ_LD + q1] = n11;
        }
      }
      // G <- G J on row gi, pairs gk0 and gk1
      {
        const int pa = tp[gk0], qa = tq[gk0], pb = tp[gk1], qb = tq[gk1];
        const double ca = rc[gk0], sa = rs[gk0], cb = rc[gk1], sbv = rs[gk1];
        const double vpa = Gl[gi * BJ_LD + pa], vqa = Gl[gi * BJ_LD + qa];
        const double vpb = Gl[gi * BJ_LD + pb], vqb = Gl[gi * BJ_LD + qb];
        Gl[gi * BJ_LD + pa] = ca * vpa - sa * vqa;
        Gl[gi * BJ_LD + qa] = sa * vpa + ca * vqa;
        Gl[gi * BJ_LD + pb] = cb * vpb - sbv * vqb;
        Gl[gi * BJ_LD + qb] = sbv * vpb + cb * vqb;
      }
      __syncthreads();
    }
    double o = 0.0;
    for (int f = tid; f < BJ_S * BJ_S; f += BJ_SOLVE_WG) {
      const int i = f & (BJ_S - 1), j = f >> 5;
      const double v = S[i * BJ_LD + j];
      if (i != j) o += v * v;
    }
    off = bj_block_sum(o, red);
    if (tid == 0) atomicAdd(&J.inner, 1);
  }
  if (round == nbk - 2 && tid == 0) J.part2[k] = off;
  __syncthreads();
  GMEM double* __restrict__ Gout = gmem(J.G) + (size_t)k * BJ_S * BJ_S;
  for (int f = tid; f < BJ_S * BJ_S; f += BJ_SOLVE_WG) {
    const int i = f & (BJ_S - 1), j = f >> 5;
    A[bj_row(i, P, Q) + (size_t)bj_row(j, P, Q) * np] = S[i * BJ_LD + j];
    Gout[i + j * BJ_S] = Gl[i * BJ_LD + j];
  }
}

// --------------------------------------------------------- transformation
// Block x < nA(node): the block pair (I, K), I > K, of the round's pairs:
// A_IK <- G_I^T (A_IK G_K), written to A_IK and A_KI (four waves, one
// 16 x 16 quadrant each); with `last`, the block's off-diagonal sum of
// squares (twice, for both triangles) into part[x].  Later blocks: four
// eigenvector row tiles each, V[rows, I] <- V[rows, I] G_I (one wave per
// 16-row tile and pair).
__global__ __launch_bounds__(256) void k_bj_update(BjJob* jobs, int round) {
  __shared__ double Tm[BJ_S * BJ_LD];
  __shared__ double red[4];
  BjJob& J = jobs[blockIdx.y];
  if (J.done) return;
  const int nbk = J.nbk, np = J.np;
  if (round >= nbk - 1) return;
  const int npair = nbk / 2, nA = npair * (npair - 1) / 2;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
  GMEM double* __restrict__ A = gmem(J.A);
  const GMEM double* __restrict__ Gg = gmem(J.G);
  const int x = blockIdx.x;
  if (x < nA) {
    int ti, K;
    bj_tri(x, ti, K);
    const int I = ti + 1;
    int PI, QI, PK, QK;
    bj_pair(round, I, nbk, PI, QI);
    bj_pair(round, K, nbk, PK, QK);
    const GMEM double* GI = Gg + (size_t)I * BJ_S * BJ_S;
    const GMEM double* GK = Gg + (size_t)K * BJ_S * BJ_S;
    const int qi = w & 1, qj = w >> 1;
    // T quadrant = A_IK[qi rows, :] G_K[:, qj cols]
    double a[8], b[8];
    const int arow = bj_row(qi * 16 + r, PI, QI);
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const int kk = 4 * s + q;
      a[s] = A[arow + (size_t)bj_row(kk, PK, QK) * np];
      b[s] = GK[kk + (qj * 16 + r) * BJ_S];
    }
    f64x4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 8; ++s) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b[s], acc, 0, 0, 0);
#pragma unroll
    for (int v = 0; v < 4; ++v) Tm[(qi * 16 + q + 4 * v) * BJ_LD + qj * 16 + r] = acc[v];
    __syncthreads();
    // C quadrant = G_I[:, qi rows]^T T[:, qj cols]
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const int kk = 4 * s + q;
      a[s] = GI[kk + (qi * 16 + r) * BJ_S];
      b[s] = Tm[kk * BJ_LD + qj * 16 + r];
    }
    acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 8; ++s) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b[s], acc, 0, 0, 0);
    double o2 = 0.0;
    const int gj = bj_row(qj * 16 + r, PK, QK);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int gi = bj_row(qi * 16 + q + 4 * v, PI, QI);
      A[gi + (size_t)gj * np] = acc[v];
      A[gj + (size_t)gi * np] = acc[v];
      o2 += acc[v] * acc[v];
    }
    if (round == nbk - 2) {
      o2 = wave_sum(o2);
      if (lane == 0) red[w] = o2;
      __syncthreads();
      if (threadIdx.x == 0) J.part[x] = 2.0 * (((red[0] + red[1]) + red[2]) + red[3]);
    }
    return;
  }
  if (J.values_only) return;
  const int ntr = (J.n + 15) >> 4;
  const int vt = (x - nA) * 4 + w;
  if (vt >= ntr * npair) return;
  const int I = vt % npair, tr = vt / npair;
  int PI, QI;
  bj_pair(round, I, nbk, PI, QI);
  const GMEM double* GI = Gg + (size_t)I * BJ_S * BJ_S;
  GMEM double* __restrict__ Vp = gmem(J.V);
  const int row = tr * 16 + r;
  double a[8], b0[8], b1[8];
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const int kk = 4 * s + q;
    a[s] = Vp[row + (size_t)bj_row(kk, PI, QI) * np];
    b0[s] = GI[kk + r * BJ_S];
    b1[s] = GI[kk + (16 + r) * BJ_S];
  }
  f64x4 c0 = {0.0, 0.0, 0.0, 0.0}, c1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    c0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b0[s], c0, 0, 0, 0);
    c1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b1[s], c1, 0, 0, 0);
  }
  const int g0 = bj_row(r, PI, QI), g1 = bj_row(16 + r, PI, QI);
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int i = tr * 16 + q + 4 * v;
    Vp[i + (size_t)g0 * np] = c0[v];
    Vp[i + (size_t)g1 * np] = c1[v];
  }
}

// End of a sweep: the off-diagonal sum of squares in block-pair order and
// the diagonal's, the stop test off2 <= 1e-24 |A|_F^2.
__global__ __launch_bounds__(64) void k_bj_check(BjJob* jobs) {
  __shared__ double pl[BJ_PART_MAX];
  BjJob& J = jobs[blockIdx.x];
  if (J.done) return;
  const int lane = threadIdx.x, np = J.np;
  const int npair = J.nbk / 2, nA = npair * (npair - 1) / 2;
  double d2 = 0.0;
  for (int i = lane; i < J.n; i += 64) {
    const double d = J.A[i + (size_t)i * np];
    d2 += d * d;
  }
  d2 = wave_sum(d2);
  // the partial sums into LDS by all lanes at once, then added in order by
  // lane 0
  for (int x = lane; x < nA; x += 64) pl[x] = J.part[x];
  for (int x = lane; x < npair; x += 64) pl[nA + x] = J.part2[x];
  __syncthreads();
  if (lane != 0) return;
  double off2 = 0.0;
  for (int x = 0; x < nA; ++x) off2 += pl[x];
  for (int x = 0; x < npair; ++x) off2 += pl[nA + x];
  J.off2 = off2;
  J.fro2 = d2 + off2;
  J.sweeps += 1;
  if (off2 <= 1e-24 * (d2 + off2)) J.done = 1;
}

// Results in the node: eigenvalues (ev), kept eigenvalues (evp) and their
// count (M), the eigenvectors (V, for the next warm start and the rounding
// heuristic), the residual bound off_last = |off(A)|_F and the sweep counts.
__global__ __launch_bounds__(256) void k_bj_finish(BjJob* jobs, NodeView* nodes, int ignore_done) {
  __shared__ int cnt[4];
  BjJob& J = jobs[blockIdx.y];
  if (J.skip) return;
  NodeView& nd = nodes[J.node];
  if (!ignore_done && nd.done) return;
  const int n = J.n, np = J.np;
  if (blockIdx.x == 0) {
    int c = 0;
    for (int i = threadIdx.x; i < n; i += 256) {
      const double lam = J.A[i + (size_t)i * np];
      nd.ev[i] = lam;
      if (!J.values_only) {
        const double d = lam > 1e-8 ? lam : 0.0;
        nd.evp[i] = d;
        c += d > 0.0 ? 1 : 0;
      }
    }
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    if ((threadIdx.x & 63) == 0) cnt[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
      if (!J.values_only) {
        nd.M = cnt[0] + cnt[1] + cnt[2] + cnt[3];
        nd.eig_warm = 1;
      }
      nd.off_last = sqrt(J.off2);
      nd.sweeps_last = J.sweeps;
      nd.sweeps_total += J.sweeps;
    }
  }
  if (J.values_only) return;
  const size_t nn = (size_t)n * n;
  for (size_t f = (size_t)blockIdx.x * 256 + threadIdx.x; f < nn; f += (size_t)gridDim.x * 256) {
    const int j = (int)(f / n), i = (int)(f - (size_t)j * n);
    nd.V[f] = J.V[i + (size_t)j * np];
  }
}

}  // namespace madam

namespace mh {

using madam::BjJob;

static DBuf<BjJob> g_bjobs;
static DBuf<double> g_bjA, g_bjV, g_bjG, g_bjP;
static BjJob* g_hbj = nullptr;
static int g_hbj_cap = 0;
// orders from which the plain projection uses this solver (below: the
// one-workgroup Jacobi kernel)
static int g_bj_min_n = 80;
// diagnostics (madam_bj_stats): calls, sweeps, subproblem solves, inner sweeps
static long g_bj_calls = 0, g_bj_sweeps = 0, g_bj_solves = 0, g_bj_inner = 0;
static int g_bj_stats = 0;

void release_bj() {
  g_bjobs.release();
  g_bjA.release();
  g_bjV.release();
  g_bjG.release();
  g_bjP.release();
  if (g_hbj) (void)hipHostFree(g_hbj);
  g_hbj = nullptr;
  g_hbj_cap = 0;
}

bool bj_wanted(int nmax) { return nmax >= g_bj_min_n; }

int bj_min_order() { return g_bj_min_n; }

void bj_set_min_n(int n) { g_bj_min_n = n; }

int bj_eig(NodeView* d_nodes, const NodeView* hv, int nb, int values_only, int ignore_done,
           int max_sweeps, hipStream_t st) {
  using madam::BJ_B;
  using madam::BJ_S;
  std::vector<BjJob> J(nb);
  size_t offA = 0, offG = 0, offP = 0;
  int nbk_max = 2, ntile_max = 1, upd_max = 1;
  for (int b = 0; b < nb; ++b) {
    const int n = std::max(1, hv[b].n);
    if (n > MADAM_NMAX) {
      set_err("bj_eig: order %d above %d", n, MADAM_NMAX);
      return MADAM_E_ARG;
    }
    int nbk = (n + BJ_B - 1) / BJ_B;
    nbk += nbk & 1;
    const int np = nbk * BJ_B, npair = nbk / 2, nA = npair * (npair - 1) / 2;
    const int ntr = (n + 15) / 16, nt = ntr;
    J[b] = BjJob{};
    J[b].node = b;
    J[b].skip = J[b].done = hv[b].n < g_bj_min_n ? 1 : 0;
    J[b].n = n;
    J[b].np = np;
    J[b].nbk = nbk;
    J[b].values_only = values_only;
    J[b].A = (double*)offA;  // offsets for now
    J[b].V = (double*)offA;
    J[b].G = (double*)offG;
    J[b].part = (double*)offP;
    J[b].part2 = (double*)(offP + (size_t)std::max(1, nA));
    offA += (size_t)np * np;
    offG += (size_t)npair * BJ_S * BJ_S;
    offP += (size_t)std::max(1, nA) + npair;
    nbk_max = std::max(nbk_max, nbk);
    ntile_max = std::max(ntile_max, nt * nt);
    upd_max = std::max(upd_max, nA + (values_only ? 0 : (ntr * npair + 3) / 4));
  }
  int rc;
  if ((rc = g_bjobs.ensure(nb)) || (rc = g_bjA.ensure(offA)) ||
      (!values_only && (rc = g_bjV.ensure(offA))) || (rc = g_bjG.ensure(offG)) ||
      (rc = g_bjP.ensure(offP)))
    return rc;
  for (int b = 0; b < nb; ++b) {
    const size_t oA = (size_t)J[b].A;
    J[b].A = g_bjA.p + oA;
    J[b].V = values_only ? nullptr : g_bjV.p + oA;
    J[b].G = g_bjG.p + (size_t)J[b].G;
    J[b].part = g_bjP.p + (size_t)J[b].part;
    J[b].part2 = g_bjP.p + (size_t)J[b].part2;
  }
  if (g_hbj_cap < nb) {
    if (g_hbj) (void)hipHostFree(g_hbj);
    g_hbj = nullptr;
    g_hbj_cap = 0;
    HIP_TRY(hipHostMalloc((void**)&g_hbj, (size_t)nb * sizeof(BjJob), hipHostMallocDefault));
    g_hbj_cap = nb;
  }
  if ((rc = up(g_bjobs.p, J.data(), (size_t)nb, st))) return rc;
  BjJob* jobs = g_bjobs.p;
  const int np_max = nbk_max * BJ_B;
  const int nsetup = std::min(256, (np_max * np_max + 255) / 256);
  madam::k_bj_setup<<<dim3(nsetup, nb), 256, 0, st>>>(jobs, d_nodes, ignore_done);
  if (!values_only) {
    madam::k_bj_prod<<<dim3(ntile_max, nb), 256, 0, st>>>(jobs, d_nodes, 0, ignore_done);
    madam::k_bj_prod<<<dim3(ntile_max, nb), 256, 0, st>>>(jobs, d_nodes, 1, ignore_done);
  }
  const int rounds = nbk_max - 1;
  for (int sw = 0; sw < max_sweeps; ++sw) {
    for (int r = 0; r < rounds; ++r) {
      madam::k_bj_solve<<<dim3(nbk_max / 2, nb), madam::BJ_SOLVE_WG, 0, st>>>(jobs, r);
      madam::k_bj_update<<<dim3(upd_max, nb), 256, 0, st>>>(jobs, r);
    }
    madam::k_bj_check<<<nb, 64, 0, st>>>(jobs);
    HIP_TRY(hipGetLastError());
    // a warm start typically needs two or three sweeps: the host looks from
    // the second on
    if (sw >= 1) {
      HIP_TRY(hipMemcpyAsync(g_hbj, jobs, (size_t)nb * sizeof(BjJob), hipMemcpyDeviceToHost, st));
      HIP_TRY(hipStreamSynchronize(st));
      bool open = false;
      for (int b = 0; b < nb; ++b) open = open || !g_hbj[b].done;
      if (!open) break;
    }
  }
  if (g_bj_stats) {
    HIP_TRY(hipMemcpyAsync(g_hbj, jobs, (size_t)nb * sizeof(BjJob), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (int b = 0; b < nb; ++b) {
      g_bj_calls += 1;
      g_bj_sweeps += g_hbj[b].sweeps;
      g_bj_solves += (long)g_hbj[b].sweeps * (g_hbj[b].nbk - 1) * (g_hbj[b].nbk / 2);
      g_bj_inner += g_hbj[b].inner;
    }
  }
  const int nfin = std::min(64, std::max(1, (np_max * np_max + 4095) / 4096));
  madam::k_bj_finish<<<dim3(nfin, nb), 256, 0, st>>>(jobs, d_nodes, ignore_done);
  if (!values_only) {
    int nt = 0;
    for (int b = 0; b < nb; ++b) nt = std::max(nt, (J[b].n + 15) / 16);
    madam::k_bj_prod<<<dim3(nt * (nt + 1) / 2, nb), 256, 0, st>>>(jobs, d_nodes, 2, ignore_done);
  }
  HIP_TRY(hipGetLastError());
  return MADAM_OK;
}

}  // namespace mh

extern "C" int madam_bj_stats(int on, long* out4) {
  if (out4) {
    out4[0] = mh::g_bj_calls;
    out4[1] = mh::g_bj_sweeps;
    out4[2] = mh::g_bj_solves;
    out4[3] = mh::g_bj_inner;
  }
  mh::g_bj_stats = on ? 1 : 0;
  mh::g_bj_calls = mh::g_bj_sweeps = mh::g_bj_solves = mh::g_bj_inner = 0;
  return MADAM_OK;
}

extern "C" int madam_set_bj_min_order(int n) {
  if (n < 1) {
    mh::set_err("madam_set_bj_min_order: order %d < 1", n);
    return MADAM_E_ARG;
  }
  mh::bj_set_min_n(n);
  return MADAM_OK;
}
