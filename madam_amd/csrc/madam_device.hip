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

#include <math.h>

namespace madam {

constexpr int WG = 1024;   // per-node kernels (16 waves)
constexpr int WGS = 256;   // streaming kernels

// ------------------------------------------------------------------ utils

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// Sum of NV values over the workgroup; every thread receives the totals.
// `red` must hold NV * (WG/64) doubles.  Deterministic order.
template <int NV>
__device__ __forceinline__ void block_sums(double (&v)[NV], double* red) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = wave_sum(v[q]);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) red[q * 16 + w] = v[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double r = 0.0;
    for (int i = 0; i < nw; ++i) r += red[q * 16 + i];
    v[q] = r;
  }
}

__device__ __forceinline__ double block_max(double v, double* red) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  v = wave_max(v);
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double r = red[0];
  for (int i = 1; i < nw; ++i) r = fmax(r, red[i]);
  return r;
}

__device__ __forceinline__ int ent_idx(int e) { return e >= 0 ? e : ~e; }

// Packed lower-triangular index (column-major) for an order-n2 matrix.
__device__ __forceinline__ int pk(int i, int j, int n2) {
  if (i < j) {
    int t = i;
    i = j;
    j = t;
  }
  return i + j * (n2 - 1) - ((j * (j - 1)) >> 1);
}

// Round-robin (circle method) pairing: round r, pair k, n2 even.
__device__ __forceinline__ void rr_pair(int r, int k, int n2, int& p, int& q) {
  const int N1 = n2 - 1;
  if (k == 0) {
    p = N1;
    q = r;
  } else {
    p = (r + k) % N1;
    q = (r - k + N1) % N1;
  }
}

// f -> (k, l) with k >= l, f = k(k+1)/2 + l.
__device__ __forceinline__ void tri_decode(int f, int& k, int& l) {
  int kk = (int)((sqrt(8.0 * f + 1.0) - 1.0) * 0.5);
  while ((kk + 1) * (kk + 2) / 2 <= f) ++kk;
  while (kk * (kk + 1) / 2 > f) --kk;
  k = kk;
  l = f - kk * (kk + 1) / 2;
}

// -------------------------------------------------------- tiled FP64 GEMM
// C(MxN) = op(A) * diag(d) * op(B) over K, one workgroup of WG threads.
// op(A)[i,k] = TA ? A[k + i*lda] : A[i + k*lda]; op(B)[k,j] = TB ? B[j + k*ldb]
// : B[k + j*ldb]; d may be null.  `out(i, j, v)` stores a result.  Tiles of
// 64x64 outputs, 16-deep k slabs staged in LDS (tA, tB: 16*65 doubles each).
template <bool TA, bool TB, class Out>
__device__ void wg_gemm(const double* __restrict__ A, int lda, const double* __restrict__ B,
                        int ldb, const double* __restrict__ d, int M, int N, int K, Out out,
                        double* tA, double* tB) {
  const int tid = threadIdx.x;
  const int ti = tid & 31, tj = tid >> 5;  // 32 x 32 threads, each 2x2 outputs
  for (int i0 = 0; i0 < M; i0 += 64)
    for (int j0 = 0; j0 < N; j0 += 64) {
      double acc00 = 0, acc01 = 0, acc10 = 0, acc11 = 0;
      for (int k0 = 0; k0 < K; k0 += 16) {
        {  // A slab: 64 (i) x 16 (k)
          int i, kk;
          if (TA) {
            kk = tid & 15;
            i = tid >> 4;
          } else {
            i = tid & 63;
            kk = tid >> 6;
          }
          const int gi = i0 + i, gk = k0 + kk;
          double v = 0.0;
          if (gi < M && gk < K) {
            v = TA ? A[gk + (size_t)gi * lda] : A[gi + (size_t)gk * lda];
            if (d) v *= d[gk];
          }
          tA[kk * 65 + i] = v;
        }
        {  // B slab: 16 (k) x 64 (j)
          int j, kk;
          if (TB) {
            j = tid & 63;
            kk = tid >> 6;
          } else {
            kk = tid & 15;
            j = tid >> 4;
          }
          const int gj = j0 + j, gk = k0 + kk;
          double v = 0.0;
          if (gj < N && gk < K) v = TB ? B[gj + (size_t)gk * ldb] : B[gk + (size_t)gj * ldb];
          tB[kk * 65 + j] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
          const double a0 = tA[kk * 65 + ti], a1 = tA[kk * 65 + ti + 32];
          const double b0 = tB[kk * 65 + tj], b1 = tB[kk * 65 + tj + 32];
          acc00 = fma(a0, b0, acc00);
          acc01 = fma(a0, b1, acc01);
          acc10 = fma(a1, b0, acc10);
          acc11 = fma(a1, b1, acc11);
        }
        __syncthreads();
      }
      const int r0 = i0 + ti, r1 = i0 + ti + 32, c0 = j0 + tj, c1 = j0 + tj + 32;
      if (r0 < M && c0 < N) out(r0, c0, acc00);
      if (r0 < M && c1 < N) out(r0, c1, acc01);
      if (r1 < M && c0 < N) out(r1, c0, acc10);
      if (r1 < M && c1 < N) out(r1, c1, acc11);
    }
}

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
    double acc = 0.0;
    for (int e = nd.cptr[c]; e < nd.cptr[c + 1]; ++e) {
      const int en = nd.cent[e], lo = ent_idx(en);
      const double v = L[lo] + Q[lo] + X[lo] / sg;
      acc = en >= 0 ? acc + v : acc - v;
    }
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
      const double v = w[nd.cact[e]];
      acc = nd.cent[e] >= 0 ? acc + v : acc - v;
    }
    q[c] = p[c] + 0.5 * nd.cscale[c] * acc;
  }
  __syncthreads();
}

// Solve (BB^T + I) t = rhs by conjugate gradients with the diagonal as
// preconditioner, starting from the current t (the previous iterate), to a
// relative residual of rtol.  Replaces the CHOLMOD factor/solve of
// admm.c:486-551.
__global__ __launch_bounds__(WG) void k_cg(NodeView* nodes, double rtol, int maxit) {
  __shared__ double red[4 * 16];
  NodeView& nd = nodes[blockIdx.x];
  if (nd.done) return;
  const int m = nd.m;
  if (m == 0) return;
  double* __restrict__ x = nd.t;
  double* __restrict__ r = nd.cg_r;
  double* __restrict__ p = nd.cg_p;
  double* __restrict__ q = nd.cg_q;
  const double* __restrict__ b = nd.rhs;
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
  const double thr = rtol * rtol * bb;
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

// Cut duals (admm.c:557-560, 589-591) and the matrix to project
// W = L + X/sigma - Diag(y) - B^T(t) (admm.c:564-568); with use_s the safe
// bound matrix L - Diag(y) - B^T(s) (admm.c:727-729) instead.  Lower
// triangle computed in the reference's order, mirrored to the upper.
__global__ __launch_bounds__(WGS) void k_build_w(NodeView* nodes, int use_s) {
#pragma clang fp contract(off)
  NodeView& nd = nodes[blockIdx.y];
  if (!use_s && nd.done) return;
  const int n = nd.n, m = nd.m;
  const double sg = nd.sigma;
  const double* __restrict__ vec = use_s ? nd.s : nd.t;
  if (!use_s)
    for (int c = blockIdx.x * WGS + threadIdx.x; c < m; c += gridDim.x * WGS) {
      const double v = nd.t[c] - nd.u[c] / sg;
      nd.s[c] = v > 0 ? v : 0;
      nd.u[c] = v < 0 ? -sg * v : 0;
    }
  // the dual step above must finish before use_s readers run; callers order
  // them with separate launches.
  const int total = n * n;
  for (int f = blockIdx.x * WGS + threadIdx.x; f < total; f += gridDim.x * WGS) {
    const int i = f % n, j = f / n;
    if (i < j) continue;
    const int lo = i + j * n;
    double w = use_s ? nd.L[lo] : nd.L[lo] + nd.X[lo] / sg;
    if (i == j) {
      w -= nd.y[i];
    } else {
      const int a = nd.p2a[lo];
      if (a >= 0)
        for (int e = nd.aptr[a]; e < nd.aptr[a + 1]; ++e) {
          const int en = nd.aent[e], c = ent_idx(en);
          const double temp = (0.5 * nd.cscale[c]) * vec[c];
          w = en >= 0 ? w - temp : w + temp;
        }
    }
    nd.W[lo] = w;
    nd.W[j + i * n] = w;
  }
}

// ------------------------------------------------------ Jacobi eigensolver
//
// Cyclic two-sided Jacobi with round-robin parallel ordering, one workgroup
// per node.  The working matrix A is kept as a packed lower triangle, in LDS
// when it fits (IN_LDS) and in the node's S2 scratch otherwise.  Warm start:
// when nd.eig_warm is set, A = V^T W V with V the eigenvectors of the
// previous projection (nearly diagonal, so one or two sweeps finish it) and
// V accumulates the new rotations; cold start uses A = W, V = I.
// values_only (safe bound) never touches V and always starts cold.
//
// Output: ev[0..n-1] = diag(A) (unsorted), V (unless values_only),
// off_last = off-diagonal Frobenius norm left, sweeps_last.

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

template <bool IN_LDS>
__global__ __launch_bounds__(WG) void k_jacobi(NodeView* nodes, int values_only, int max_sweeps,
                                               int ignore_done) {
  extern __shared__ double smem[];
  NodeView& nd = nodes[blockIdx.x];
  if (!ignore_done && nd.done) return;
  const int n = nd.n;
  if (n <= 0) return;
  const int n2 = n + (n & 1), h = n2 >> 1;
  const int tid = threadIdx.x;
  double* rc = smem;           // h
  double* rs = rc + 512;       // h
  double* rt = rs + 512;       // h
  double* red = rt + 512;      // 64
  double* tA = red + 64;       // 16*65
  double* tB = tA + 16 * 65;   // 16*65
  double* A = IN_LDS ? (tB + 16 * 65) : nd.S2;
  double* __restrict__ V = nd.V;
  const double* __restrict__ Wm = nd.W;
  const bool warm = !values_only && nd.eig_warm;

  // ---- form A
  if (warm) {
    wg_gemm<false, false>(Wm, n, V, n, nullptr, n, n, n, DenseOut{nd.S1, n}, tA, tB);
    __syncthreads();
    wg_gemm<true, false>(V, n, nd.S1, n, nullptr, n, n, n, PackedOut{A, n2}, tA, tB);
  } else {
    for (int f = tid; f < n * n; f += WG) {
      const int i = f % n, j = f / n;
      if (i >= j) A[pk(i, j, n2)] = Wm[i + (size_t)j * n];
      if (!values_only) V[f] = (i == j) ? 1.0 : 0.0;
    }
  }
  if (n2 != n)
    for (int j = tid; j < n2; j += WG) A[pk(n, j, n2)] = 0.0;
  __syncthreads();

  // ---- norms
  double nv[2] = {0.0, 0.0};  // off^2, total^2
  for (int f = tid; f < n * n; f += WG) {
    const int i = f % n, j = f / n;
    if (i < j) continue;
    const double a = A[pk(i, j, n2)];
    const double a2 = (i == j) ? a * a : 2.0 * a * a;
    nv[1] += a2;
    if (i != j) nv[0] += a2;
  }
  block_sums<2>(nv, red);
  double off2 = nv[0];
  const double fro2 = nv[1];
  const double thr = 1e-28 * fro2;  // relative off-norm 1e-14
  const int nblk = h * (h + 1) / 2;
  int sweep = 0;
  for (; sweep < max_sweeps && off2 > thr; ++sweep) {
    for (int r = 0; r < n2 - 1; ++r) {
      for (int k = tid; k < h; k += WG) {
        int p, q;
        rr_pair(r, k, n2, p, q);
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
        int k, l;
        tri_decode(f, k, l);
        int p, q, p2, q2;
        rr_pair(r, k, n2, p, q);
        if (k == l) {
          const int ipq = pk(p, q, n2);
          const double apq = A[ipq], t = rt[k];
          const int ipp = pk(p, p, n2), iqq = pk(q, q, n2);
          A[ipp] = A[ipp] - t * apq;
          A[iqq] = A[iqq] + t * apq;
          A[ipq] = 0.0;
        } else {
          rr_pair(r, l, n2, p2, q2);
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
          const int i = f % n, l = f / n;
          int p, q;
          rr_pair(r, l, n2, p, q);
          if (p >= n || q >= n) continue;
          const double c = rc[l], s = rs[l];
          const double vp = V[i + (size_t)p * n], vq = V[i + (size_t)q * n];
          V[i + (size_t)p * n] = c * vp - s * vq;
          V[i + (size_t)q * n] = s * vp + c * vq;
        }
      __syncthreads();
    }
    double ov[1] = {0.0};
    for (int f = tid; f < n * n; f += WG) {
      const int i = f % n, j = f / n;
      if (i <= j) continue;
      const double a = A[pk(i, j, n2)];
      ov[0] += 2.0 * a * a;
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
}

// X_+ = V diag(lambda_+) V^T with lambda_+ = lambda for lambda > 1e-8, else 0
// (the dsyevr range (1e-8, VU] of operators.c:536-579; VU >= lambda_max);
// then X = sigma X_+, Q = X_+ - W and the residuals, sigma update and stop
// test of admm.c:580-702.
__global__ __launch_bounds__(WG) void k_admm_post(NodeView* nodes) {
  __shared__ double red[4 * 16];
  __shared__ double tA[16 * 65], tB[16 * 65];
  __shared__ double dpos[1024];
  NodeView& nd = nodes[blockIdx.x];
  if (nd.done) return;
  const int n = nd.n, m = nd.m, tid = threadIdx.x;
  const double sg = nd.sigma;
  // positive eigenvalues (n <= 1024)
  int mloc = 0;
  for (int j = tid; j < n; j += WG) {
    const double lam = nd.ev[j];
    const double d = lam > 1e-8 ? lam : 0.0;
    dpos[j] = d;
    nd.evp[j] = d;
    mloc += d > 0.0;
  }
  double mv[1] = {(double)mloc};
  block_sums<1>(mv, red);
  wg_gemm<false, true>(nd.V, n, nd.V, n, dpos, n, n, n, DenseOut{nd.S1, n}, tA, tB);
  __syncthreads();
  double v3[3] = {0, 0, 0};  // |X+ - Xold/sigma|_F^2, |diag X - 1|^2, <L,X>
  for (int f = tid; f < n * n; f += WG) {
    const int i = f % n, j = f / n;
    if (i < j) continue;
    const int lo = i + j * n, up = j + i * n;
    const double xp = nd.S1[lo];
    const double xo = nd.X[lo];
    const double xn = sg * xp;
    const double dd = xp - xo / sg;
    nd.X[lo] = xn;
    nd.X[up] = xn;
    const double qv = xp - nd.W[lo];
    nd.Q[lo] = qv;
    nd.Q[up] = qv;
    if (i == j) {
      v3[0] += dd * dd;
      v3[1] += (xn - 1.0) * (xn - 1.0);
      v3[2] += nd.L[lo] * xn;
    } else {
      v3[0] += 2.0 * dd * dd;
      v3[2] += 2.0 * nd.L[lo] * xn;
    }
  }
  block_sums<3>(v3, red);
  // cut residuals: |max(B(X)-1,0)|^2, |s-t|^2, sum t ; and sum y
  double v4[4] = {0, 0, 0, 0};
  for (int c = tid; c < m; c += WG) {
    double acc = 0.0;
    for (int e = nd.cptr[c]; e < nd.cptr[c + 1]; ++e) {
      const int en = nd.cent[e];
      const double v = nd.X[ent_idx(en)];
      acc = en >= 0 ? acc + v : acc - v;
    }
    const double bx = -1.0 + nd.cscale[c] * acc;
    if (bx > 0) v4[0] += bx * bx;
    const double ds = nd.s[c] - nd.t[c];
    v4[1] += ds * ds;
    v4[2] += nd.t[c];
  }
  for (int i = tid; i < n; i += WG) v4[3] += nd.y[i];
  block_sums<4>(v4, red);
  if (tid == 0) {
    nd.M = (int)mv[0];
    nd.sigma_last = sg;
    const double err_p = (sqrt(v3[1]) + sqrt(v4[0])) / (1.0 + sqrt((double)n));
    const double err_d = (sqrt(v4[1]) + sqrt(v3[0])) / (1.0 + nd.normL);
    const double primal = v3[2];
    const double dual = v4[3] + v4[2];
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

// Projection only (project_SDP, operators.c:521-619): after k_jacobi on W,
// X = X_+, Q = X_+ - W (no sigma scaling).
__global__ __launch_bounds__(WG) void k_project_out(NodeView* nodes) {
  __shared__ double tA[16 * 65], tB[16 * 65];
  __shared__ double dpos[1024];
  __shared__ double red[64];
  NodeView& nd = nodes[blockIdx.x];
  const int n = nd.n, tid = threadIdx.x;
  int mloc = 0;
  for (int j = tid; j < n; j += WG) {
    const double lam = nd.ev[j];
    dpos[j] = lam > 1e-8 ? lam : 0.0;
    mloc += lam > 1e-8;
  }
  double mv[1] = {(double)mloc};
  block_sums<1>(mv, red);
  wg_gemm<false, true>(nd.V, n, nd.V, n, dpos, n, n, n, DenseOut{nd.S1, n}, tA, tB);
  __syncthreads();
  for (int f = tid; f < n * n; f += WG) {
    const int i = f % n, j = f / n;
    if (i < j) continue;
    const double xp = nd.S1[i + j * n];
    nd.X[i + j * n] = xp;
    nd.X[j + i * n] = xp;
    const double qv = xp - nd.W[i + j * n];
    nd.Q[i + j * n] = qv;
    nd.Q[j + i * n] = qv;
  }
  if (tid == 0) nd.M = (int)mv[0];
}

// ------------------------------------------------------------ launchers

static int chunks_for(int work, int per) {
  int c = (work + per - 1) / per;
  if (c < 1) c = 1;
  if (c > 64) c = 64;
  return c;
}

void launch_admm_prep(NodeView* nodes, int nb, int mmax, hipStream_t st) {
  dim3 grid(chunks_for(mmax, WGS * 4), nb);
  k_admm_prep<<<grid, WGS, 0, st>>>(nodes);
}

void launch_cg(NodeView* nodes, int nb, double rtol, int maxit, hipStream_t st) {
  k_cg<<<nb, WG, 0, st>>>(nodes, rtol, maxit);
}

void launch_build_w(NodeView* nodes, int nb, int nmax, int mmax, int use_s, hipStream_t st) {
  int work = nmax * nmax / 2 > mmax ? nmax * nmax / 2 : mmax;
  dim3 grid(chunks_for(work, WGS * 4), nb);
  k_build_w<<<grid, WGS, 0, st>>>(nodes, use_s);
}

static size_t jacobi_lds_fixed() { return (3 * 512 + 64 + 2 * 16 * 65) * sizeof(double); }

size_t jacobi_lds_bytes(int nmax) {
  const int n2 = nmax + (nmax & 1);
  return jacobi_lds_fixed() + (size_t)n2 * (n2 + 1) / 2 * sizeof(double);
}

void launch_jacobi(NodeView* nodes, int nb, int nmax, int values_only, int max_sweeps,
                   int ignore_done, size_t lds_limit, hipStream_t st) {
  const size_t need = jacobi_lds_bytes(nmax);
  if (need <= lds_limit) {
    k_jacobi<true><<<nb, WG, need, st>>>(nodes, values_only, max_sweeps, ignore_done);
  } else {
    k_jacobi<false><<<nb, WG, jacobi_lds_fixed(), st>>>(nodes, values_only, max_sweeps,
                                                      ignore_done);
  }
}

void launch_admm_post(NodeView* nodes, int nb, int nmax, hipStream_t st) {
  (void)nmax;
  k_admm_post<<<nb, WG, 0, st>>>(nodes);
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
  k_project_out<<<nb, WG, 0, st>>>(nodes);
}

}  // namespace madam
