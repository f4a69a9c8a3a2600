// madam_ipm.hip -- basic-SDP primal-dual predictor-corrector interior-point
// method of ipm_mc_pk.c:28-347 (max <L,X> s.t. diag(X) = e, X psd), one
// persistent workgroup per node running every iteration on the device.
//
// Per node: X (nd.X), Z = Diag(y) - L (nd.Q), y (nd.y); scratch Zi (S1),
// M (S2), dX (W), dX1 (ip1), T1 (ip2), T2 (ip3), vectors in ipv.  Cholesky
// factors are packed lower triangles in LDS: orders whose factor does not
// fit LDS run in the many-workgroup form of madam_ipmt.hip.
// Where the reference exits (ipm_mc_pk.c:124-135, 161-165) the kernel
// stops with ipm_status = 1..3 and the host reports the failure.  Zi is
// formed as the reference does (dpotrf + dpotri, ipm_mc_pk.c:121-135); the
// two dposv solves with the same M = Zi .* X (ipm_mc_pk.c:154-158,
// 208-217) share one factorisation, which is the one dposv would compute
// again.
#include "madam_devutil.h"
#include "madam_internal.h"

namespace madam {

// A^{-1} = L^{-T} L^{-1} from the packed Cholesky factor L of A (dpotri):
// L^{-1} in place in P, column by column from the right (dtrti2 order:
// column j below the diagonal becomes -(1/L_jj) T x with T the inverted
// trailing block, one output row per group of 8 lanes), then out = the
// full symmetric A^{-1} (ld n), each lower entry an independent dot
// product of two columns of L^{-1}.  tmp: n - 1 doubles (the caller's
// 16 x 65 LDS slab: n <= 1041).
__device__ void potri_packed(double* P, int n, int n2, double* out, double* tmp) {
  const int tid = threadIdx.x, g = tid & 7, grp = tid >> 3, ngrp = blockDim.x >> 3;
  for (int j = n - 1; j >= 0; --j) {
    const int base = pk(j, j, n2);
    const double ajj = 1.0 / P[base];
    const int m = n - 1 - j;
    for (int i0 = 0; i0 < m; i0 += ngrp) {
      const int ii = i0 + grp;  // the same in all 8 lanes of a group
      double acc = 0.0;
      if (ii < m) {
        const int i = j + 1 + ii;
        for (int l = j + 1 + g; l <= i; l += 8) acc += P[pk(i, l, n2)] * P[base + (l - j)];
      }
      acc += __shfl_xor(acc, 1, 64);
      acc += __shfl_xor(acc, 2, 64);
      acc += __shfl_xor(acc, 4, 64);
      if (ii < m && g == 0) tmp[ii] = acc;
    }
    __syncthreads();
    for (int ii = tid; ii < m; ii += blockDim.x) P[base + 1 + ii] = -ajj * tmp[ii];
    if (tid == 0) P[base] = ajj;
    __syncthreads();
  }
  const int npair = n * (n + 1) / 2;
  for (int f = tid; f < npair; f += blockDim.x) {
    int i, j;
    tri_decode(f, i, j);  // i >= j
    const int ci = pk(i, i, n2), cj = pk(j, j, n2);
    double acc = 0.0;
    for (int k = i; k < n; ++k) acc += P[ci + (k - i)] * P[cj + (k - j)];
    out[i + (size_t)j * n] = acc;
    out[j + (size_t)i * n] = acc;
  }
  __syncthreads();
}

// P[lower] = A + alpha * Diag(dv) (dv may be null) or A + alpha * B.
__device__ void load_packed(double* P, int n, int n2, const double* A, double alpha,
                            const double* B, const double* dv) {
  for (int f = threadIdx.x; f < n * n; f += blockDim.x) {
    const int j = f / n, i = f - j * n;
    if (i < j) continue;
    double v = A[f];
    if (B) v += alpha * B[f];
    if (dv && i == j) v += alpha * dv[i];
    P[pk(i, j, n2)] = v;
  }
  __syncthreads();
}

// Solve (L L^T) x = b with the packed factor, by wave 0 (b and x may
// alias).  Column-oriented sweeps in LAPACK dtrsv's order (forward: x_k =
// b_k / L_kk, then b_i -= L_ik x_k below; backward: x_k = (z_k - sum over
// i > k of L_ik x_i, i descending) / L_kk): row i lives in a register of
// lane i % 64 and each x_k is broadcast from its lane, so a step costs a
// few LDS reads and no reduction.  n <= 256.
__device__ __forceinline__ double lane_bcast(double v, int src) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
  return __hiloint2double(hi, lo);
}

__device__ void chol_solve_reg(const double* P, int n, int n2, const double* b, double* x) {
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    double v[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = lane + 64 * r;
      v[r] = i < n ? b[i] : 0.0;
    }
    for (int k = 0; k < n; ++k) {
      const int src = k & 63, rk = k >> 6;
      double vk = 0.0;
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (r == rk) vk = v[r];
      const double xk = lane_bcast(vk, src) / P[pk(k, k, n2)];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = lane + 64 * r;
        if (i == k)
          v[r] = xk;
        else if (i > k && i < n)
          v[r] -= P[pk(i, k, n2)] * xk;
      }
    }
    for (int k = n - 1; k >= 0; --k) {
      const int src = k & 63, rk = k >> 6;
      double vk = 0.0;
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (r == rk) vk = v[r];
      const double xk = lane_bcast(vk, src) / P[pk(k, k, n2)];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = lane + 64 * r;
        if (i == k)
          v[r] = xk;
        else if (i < k)
          v[r] -= P[pk(k, i, n2)] * xk;
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = lane + 64 * r;
      if (i < n) x[i] = v[r];
    }
  }
  __syncthreads();
}

// The same for any n: row-oriented sweeps with wave reductions.
__device__ void chol_solve1(const double* P, int n, int n2, const double* b, double* x) {
  if (n <= 256) {
    chol_solve_reg(P, n, n2, b, x);
    return;
  }
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    for (int i = 0; i < n; ++i) {
      double part = 0.0;
      for (int k = lane; k < i; k += 64) part += P[pk(i, k, n2)] * x[k];
      const double sacc = wave_sum(part);
      const double xi = (b[i] - sacc) / P[pk(i, i, n2)];
      if (lane == (i & 63)) x[i] = xi;
      __builtin_amdgcn_wave_barrier();
    }
    for (int i = n - 1; i >= 0; --i) {
      double part = 0.0;
      for (int k = i + 1 + lane; k < n; k += 64) part += P[pk(k, i, n2)] * x[k];
      const double sacc = wave_sum(part);
      // x[i] written only by lane (i & 63), read back by the same lane
      double xi = 0.0;
      if (lane == (i & 63)) xi = (x[i] - sacc) / P[pk(i, i, n2)];
      if (lane == (i & 63)) x[i] = xi;
      __builtin_amdgcn_wave_barrier();
    }
  }
  __syncthreads();
}

struct IpmOut {
  double* C;
  int n;
  __device__ void operator()(int i, int j, double v) const { C[i + (size_t)j * n] = v; }
};

__global__ __launch_bounds__(WG) void k_ipm(NodeView* nodes, int maxit) {
  extern __shared__ double smem[];
  NodeView& nd = nodes[blockIdx.x];
  const int n = nd.n, tid = threadIdx.x;
  const int n2 = n;
  double* red = smem;
  double* tA = red + 64;
  double* tB = tA + 16 * 65;
  __shared__ int flag[2];
  double* P = tB + 16 * 65;  // the packed Cholesky factor, in LDS
  auto chol = [&]() { return chol_packed(P, n, n2, flag); };
  auto solve = [&](const double* b, double* x) { chol_solve1(P, n, n2, b, x); };
  auto inverse = [&](double* outm) { potri_packed(P, n, n2, outm, tA); };
  const double* __restrict__ L = nd.L;
  double* X = nd.X;
  double* Z = nd.Q;
  double* y = nd.y;
  double* Zi = nd.S1;
  double* Mm = nd.S2;
  double* dX = nd.W;
  double* dX1 = nd.ip1;
  double* T1 = nd.ip2;
  double* T2 = nd.ip3;
  double* dy = nd.ipv;
  double* dy1 = nd.ipv + n;
  const size_t nn = (size_t)n * n;

  // y = 1.1 + column sums of |L| (ipm_mc_pk.c:61-68); X = I; Z = Diag(y) - L
  for (int i = tid; i < n; i += WG) {
    double acc = 0.0;
    for (int j = 0; j < n; ++j) acc += fabs(L[j + (size_t)i * n]);
    y[i] = acc + 1.1;
  }
  __syncthreads();
  for (size_t f = tid; f < nn; f += WG) {
    const int j = (int)(f / n), i = (int)(f - (size_t)j * n);
    X[f] = (i == j) ? 1.0 : 0.0;
    Z[f] = ((i == j) ? y[i] : 0.0) - L[f];
  }
  __syncthreads();
  double v3[3] = {0, 0, 0};  // phi, psi, <Z,X>
  for (int i = tid; i < n; i += WG) v3[0] += y[i];
  for (size_t f = tid; f < nn; f += WG) {
    v3[1] += L[f] * X[f];
    v3[2] += Z[f] * X[f];
  }
  block_sums<3>(v3, red);
  double phi = v3[0], psi = v3[1], mu = v3[2] / (2.0 * n);
  int it = 0, status = 0;
  while (phi - psi > 1e-4 && it < maxit) {
    ++it;
    // ---- Zi = Z^{-1}: factor, then invert from the factor
    load_packed(P, n, n2, Z, 0.0, nullptr, nullptr);
    if (chol()) { status = 1; break; }
    inverse(Zi);
    // ---- predictor: dy1 = (Zi .* X) \ (-e)
    for (size_t f = tid; f < nn; f += WG) Mm[f] = Zi[f] * X[f];
    __syncthreads();
    load_packed(P, n, n2, Mm, 0.0, nullptr, nullptr);
    if (chol()) { status = 3; break; }
    for (int i = tid; i < n; i += WG) dy1[i] = -1.0;
    __syncthreads();
    solve(dy1, dy1);  // P keeps M's factor for the corrector
    // dX1 = -Zi Diag(dy1) X - X, symmetrised
    for (size_t f = tid; f < nn; f += WG) {
      const int j = (int)(f / n), i = (int)(f - (size_t)j * n);
      T1[f] = -dy1[i] * X[f];
    }
    __syncthreads();
    wg_gemm<false, false>(Zi, n, T1, n, nullptr, n, n, n, IpmOut{dX1, n}, tA, tB);
    __syncthreads();
    for (size_t f = tid; f < nn; f += WG) dX1[f] -= X[f];
    __syncthreads();
    for (size_t f = tid; f < nn; f += WG) {
      const int j = (int)(f / n), i = (int)(f - (size_t)j * n);
      if (i <= j) {
        const size_t g = j + (size_t)i * n;
        const double a = 0.5 * (dX1[f] + dX1[g]);
        T2[f] = a;
        T2[g] = a;
      }
    }
    __syncthreads();
    for (size_t f = tid; f < nn; f += WG) dX1[f] = T2[f];
    __syncthreads();
    // ---- corrector: dy2 = (Zi .* X) \ (mu diag(Zi) - (Zi .* dX1) dy1)
    for (size_t f = tid; f < nn; f += WG) T1[f] = Zi[f] * dX1[f];
    __syncthreads();
    for (int i = tid; i < n; i += WG) {
      double acc = 0.0;
      for (int j = 0; j < n; ++j) acc += T1[i + (size_t)j * n] * dy1[j];
      dy[i] = -acc + mu * Zi[(size_t)i * (n + 1)];
    }
    __syncthreads();
    solve(dy, dy);  // the same M: its factor is still in P
    // dX2 = mu Zi - Zi (Diag(dy2) X + Diag(dy1) dX1)
    for (size_t f = tid; f < nn; f += WG) {
      const int j = (int)(f / n), i = (int)(f - (size_t)j * n);
      T1[f] = -dy[i] * X[f] + (-dy1[i] * dX1[f]);
    }
    __syncthreads();
    wg_gemm<false, false>(Zi, n, T1, n, nullptr, n, n, n, IpmOut{dX, n}, tA, tB);
    __syncthreads();
    for (size_t f = tid; f < nn; f += WG) dX[f] += mu * Zi[f] + dX1[f];
    for (int i = tid; i < n; i += WG) dy[i] += dy1[i];
    __syncthreads();
    for (size_t f = tid; f < nn; f += WG) {
      const int j = (int)(f / n), i = (int)(f - (size_t)j * n);
      if (i <= j) {
        const size_t g = j + (size_t)i * n;
        const double a = 0.5 * (dX[f] + dX[g]);
        T2[f] = a;
        T2[g] = a;
      }
    }
    __syncthreads();
    for (size_t f = tid; f < nn; f += WG) dX[f] = T2[f];
    __syncthreads();
    // ---- step lengths by Cholesky back-off (ipm_mc_pk.c:273-301)
    double ap = 1.0;
    for (int tries = 0; tries < 400; ++tries) {
      load_packed(P, n, n2, X, ap, dX, nullptr);
      if (!chol()) break;
      ap *= 0.8;
    }
    if (ap < 1.0) ap *= 0.95;
    double ad = 1.0;
    for (int tries = 0; tries < 400; ++tries) {
      load_packed(P, n, n2, Z, ad, nullptr, dy);
      if (!chol()) break;
      ad *= 0.8;
    }
    if (ad < 1.0) ad *= 0.95;
    // ---- update
    for (size_t f = tid; f < nn; f += WG) X[f] += ap * dX[f];
    for (int i = tid; i < n; i += WG) {
      y[i] += ad * dy[i];
      Z[(size_t)i * (n + 1)] += ad * dy[i];
    }
    __syncthreads();
    double w3[3] = {0, 0, 0};
    for (size_t f = tid; f < nn; f += WG) {
      w3[0] += Z[f] * X[f];
      w3[1] += L[f] * X[f];
    }
    for (int i = tid; i < n; i += WG) w3[2] += y[i];
    block_sums<3>(w3, red);
    mu = w3[0] / (2.0 * n);
    if (ap + ad > 1.6) mu *= 0.5;
    if (ap + ad > 1.9) mu *= 0.2;
    phi = w3[2];
    psi = w3[1];
  }
  if (tid == 0) {
    nd.f = phi;
    nd.ipm_iters = it;
    nd.ipm_status = status;
    nd.eig_warm = 0;
  }
}

size_t ipm_lds_bytes(int n) {
  return (64 + 2 * 16 * 65 + (size_t)n * (n + 1) / 2) * sizeof(double);
}

// One workgroup per node; orders whose packed factor does not fit LDS run
// in the many-workgroup form of madam_ipmt.hip instead.
void launch_ipm(NodeView* nodes, int nb, int nmax, hipStream_t st) {
  k_ipm<<<nb, WG, ipm_lds_bytes(nmax), st>>>(nodes, 500);
}

}  // namespace madam
