// madam_ipm.hip -- basic-SDP primal-dual predictor-corrector interior-point
// method of ipm_mc_pk.c:28-347 (max <L,X> s.t. diag(X) = e, X psd), one
// persistent workgroup per node running every iteration on the device.
//
// Per node: X (nd.X), Z = Diag(y) - L (nd.Q), y (nd.y); scratch Zi (S1),
// M (S2), dX (W), dX1 (ip1), T1 (ip2), T2 (ip3), vectors in ipv.  Cholesky
// factors are packed lower triangles, in LDS when they fit, else in ipP.
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

// Solve (L L^T) x = b with the packed factor in device memory (orders
// whose factor does not fit LDS), by wave 0 with the vector in LDS (xs, n
// doubles; b and x may alias): forward sweep column-oriented, b_i -= L_ik
// x_k below x_k = b_k / L_kk with column k read across the lanes
// (contiguous); backward sweep x_k = (z_k - sum_{i>k} L_ik x_i) / L_kk with
// the sum over column k (lanes' partial sums, then a wave reduction).  A
// wave's LDS accesses complete in order, so no barrier is needed between
// its steps.
__device__ void chol_solve_lds(const double* P, int n, int n2, const double* b, double* x,
                               double* xs) {
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    for (int i = lane; i < n; i += 64) xs[i] = b[i];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (int k = 0; k < n; ++k) {
      const int cb = pk(k, k, n2) - k;  // column k: P[cb + i] = L(i, k), i >= k
      const double xk = xs[k] / P[cb + k];
      for (int i = k + 1 + lane; i < n; i += 64) xs[i] -= P[cb + i] * xk;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (lane == 0) xs[k] = xk;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
    for (int k = n - 1; k >= 0; --k) {
      const int cb = pk(k, k, n2) - k;
      double part = 0.0;
      for (int i = k + 1 + lane; i < n; i += 64) part += P[cb + i] * xs[i];
      const double sm = wave_sum(part);
      const double xk = (xs[k] - sm) / P[cb + k];
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (lane == 0) xs[k] = xk;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
    for (int i = lane; i < n; i += 64) x[i] = xs[i];
  }
  __syncthreads();
}

// Blocked right-looking Cholesky of the packed lower triangle P (order n,
// leading order n2) in device memory, for orders whose factor does not fit
// LDS.  Panels of CB = 16 columns: the panel's rows j0..n-1 are staged in
// LDS (Lp, CB x ldp doubles, column k at Lp + k ldp), wave 0 factors them
// (unblocked, one column at a time, no barriers: a wave's LDS accesses are
// in order), the panel goes back to P, and every wave then applies the
// panel's rank-CB update to lower 16 x 16 tiles of the trailing block with
// FP64 MFMA (v_mfma_f64_16x16x4f64: A operand lane l -> (row l % 16, k
// l / 16), B operand lane l -> (k l / 16, col l % 16), result register v of
// lane l -> (row l / 16 + 4 v, col l % 16)), three barriers per panel.
// Returns 0, or c + 1 for the first column c with a non-positive pivot --
// the same value in every thread.
constexpr int CB = 16;
typedef double f64x4 __attribute__((ext_vector_type(4)));

size_t chol_blocked_lds_doubles(int n) { return (size_t)CB * (((n + 15) & ~15) + 1); }

__device__ int chol_blocked(double* P, int n, int n2, int* flag, double* Lp) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
  const int ldp = ((n + 15) & ~15) + 1;  // odd stride: the k groups of a tile read spread banks
  __syncthreads();  // readers of flag in a previous call are done
  if (tid == 0) flag[0] = 0;
  for (int j0 = 0; j0 < n; j0 += CB) {
    const int nb = n - j0 < CB ? n - j0 : CB, m = n - j0;
    // stage the panel: Lp[k ldp + (r - j0)] = P(r, j0 + k), r >= j0 + k (else 0)
    for (int f = tid; f < CB * m; f += blockDim.x) {
      const int k = f / m, rr = f - k * m;
      Lp[k * ldp + rr] = (k < nb && rr >= k) ? P[pk(j0 + rr, j0 + k, n2)] : 0.0;
    }
    __syncthreads();
    if (wv == 0) {
      for (int k = 0; k < nb; ++k) {
        double* col = Lp + k * ldp;
        const double d = col[k];
        if (!(d > 0.0)) {
          if (lane == 0) flag[0] = j0 + k + 1;
          break;  // uniform in the wave
        }
        const double sd = sqrt(d);
        for (int rr = k + 1 + lane; rr < m; rr += 64) col[rr] /= sd;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (lane == 0) col[k] = sd;
        for (int q = k + 1; q < nb; ++q) {
          double* cq = Lp + q * ldp;
          const double lq = col[q];
          for (int rr = q + lane; rr < m; rr += 64) cq[rr] -= col[rr] * lq;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
    }
    __syncthreads();
    if (flag[0]) return flag[0];
    // the factored panel back to P
    for (int f = tid; f < CB * m; f += blockDim.x) {
      const int k = f / m, rr = f - k * m;
      if (k < nb && rr >= k) P[pk(j0 + rr, j0 + k, n2)] = Lp[k * ldp + rr];
    }
    // trailing block rows/columns t0..n-1: lower 16 x 16 tiles, one per wave
    const int t0 = j0 + nb;
    if (t0 < n) {
      const int T = (n - t0 + 15) >> 4;
      const int ntile = T * (T + 1) / 2;
      const int rl = lane & 15, kl = lane >> 4;
      for (int t = wv; t < ntile; t += nw) {
        int ti = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
        if ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
        if (ti * (ti + 1) / 2 > t) --ti;
        const int tj = t - ti * (ti + 1) / 2;
        const int r0 = t0 + 16 * ti, c0 = t0 + 16 * tj;
        f64x4 acc;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = r0 + kl + 4 * v, c = c0 + rl;
          acc[v] = (r < n && c < n && r >= c) ? P[pk(r, c, n2)] : 0.0;
        }
#pragma unroll
        for (int kq = 0; kq < CB / 4; ++kq) {
          const int k = 4 * kq + kl;
          const int ra = r0 + rl - j0, rb = c0 + rl - j0;  // rows of the panel (< ldp - 1)
          const double a = (k < nb && r0 + rl < n) ? -Lp[k * ldp + ra] : 0.0;
          const double b = (k < nb && c0 + rl < n) ? Lp[k * ldp + rb] : 0.0;
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = r0 + kl + 4 * v, c = c0 + rl;
          if (r < n && c < n && r >= c) P[pk(r, c, n2)] = acc[v];
        }
      }
    }
    __syncthreads();
  }
  return 0;
}

// A 16 x 16 lower tile t of a T x T tile grid: (ti, tj) with ti >= tj.
__device__ __forceinline__ void lower_tile(int t, int& ti, int& tj) {
  int a = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
  if ((a + 1) * (a + 2) / 2 <= t) ++a;
  if (a * (a + 1) / 2 > t) --a;
  ti = a;
  tj = t - a * (a + 1) / 2;
}

// A^{-1} = W^T W, W = L^{-1}, from the packed Cholesky factor L of A in
// device memory (dpotri for orders beyond LDS).  W in place, blocks of CB
// columns from the right, with the trailing block W22 = L22^{-1} already
// in place:  W11 = L11^{-1} (wave 0, in LDS: column c from the right,
// W11(r, c) = -(1/L_cc) sum_{c<l<=r} W11(r, l) L(l, c), as dtrti2),
// S = L21 W11 (each thread one row of the panel, in place in LDS), then
// W21 = -W22 S on FP64 MFMA tiles (16 rows x the block's columns; W22 read
// from device memory, masked to its lower triangle, S from LDS).  Then
// out = W^T W (full, leading dimension n) as lower 16 x 16 MFMA tiles,
// out(i, j) = sum_{k >= max(i, j)} W(k, i) W(k, j), mirrored.
// LDS: Lp (CB x ldp doubles, the panel) and W11 (CB x CB, in tB).
__device__ void potri_blocked(double* P, int n, int n2, double* out, double* Lp, double* W11) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
  const int ldp = ((n + 15) & ~15) + 1;
  const int rl = lane & 15, kl = lane >> 4;
  const int nblk = (n + CB - 1) / CB;
  for (int bi = nblk - 1; bi >= 0; --bi) {
    const int j0 = bi * CB, nb = n - j0 < CB ? n - j0 : CB, j1 = j0 + nb, m2 = n - j1;
    // stage L11 (lower, nb x nb) into W11 and L21 (m2 x nb) into the panel
    for (int f = tid; f < CB * CB; f += blockDim.x) {
      const int c = f / CB, r = f - c * CB;
      W11[c * CB + r] = (c < nb && r < nb && r >= c) ? P[pk(j0 + r, j0 + c, n2)] : 0.0;
    }
    for (int f = tid; f < CB * m2; f += blockDim.x) {
      const int k = f / m2, rr = f - k * m2;
      Lp[k * ldp + rr] = k < nb ? P[pk(j1 + rr, j0 + k, n2)] : 0.0;
    }
    __syncthreads();
    if (wv == 0) {  // W11 = L11^{-1}, columns from the right
      for (int c = nb - 1; c >= 0; --c) {
        const double lcc = W11[c * CB + c];
        double v = 0.0;
        const int r = c + 1 + lane;  // lanes 0 .. nb-c-2 each own a row below c
        if (r < nb) {
          for (int l = c + 1; l <= r; ++l) v += W11[l * CB + r] * W11[c * CB + l];
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (r < nb) W11[c * CB + r] = -v / lcc;
        if (lane == 0) W11[c * CB + c] = 1.0 / lcc;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
    }
    __syncthreads();
    // S = L21 W11, row by row in place: S(r, c) = sum_{k=c}^{nb-1} L21(r, k)
    // W11(k, c); columns ascending, so S(r, c) overwrites L21(r, c) only
    // after every column that reads it
    for (int rr = tid; rr < m2; rr += blockDim.x) {
      for (int c = 0; c < nb; ++c) {
        double a = 0.0;
        for (int k = c; k < nb; ++k) a += Lp[k * ldp + rr] * W11[c * CB + k];
        Lp[c * ldp + rr] = a;
      }
    }
    __syncthreads();
    // W21 = -W22 S: row tiles of 16 over rows j1 .. n-1, one wave each
    const int T = (m2 + 15) >> 4;
    for (int t = wv; t < T; t += nw) {
      const int r0 = j1 + 16 * t;
      f64x4 acc = {0.0, 0.0, 0.0, 0.0};
      const int kend = r0 + 16 < n ? r0 + 16 : n;  // W22(r, l) = 0 for l > r
      for (int l0 = j1; l0 < kend; l0 += 4) {
        const int l = l0 + kl, r = r0 + rl;
        const double a = (l < n && r < n && l <= r) ? -P[pk(r, l, n2)] : 0.0;
        const double b = (l < n && rl < nb) ? Lp[rl * ldp + (l - j1)] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = r0 + kl + 4 * v, c = rl;  // result: row l/16 + 4v, column l%16
        if (r < n && c < nb) P[pk(r, j0 + c, n2)] = acc[v];
      }
    }
    // W11 into the diagonal block
    for (int f = tid; f < CB * CB; f += blockDim.x) {
      const int c = f / CB, r = f - c * CB;
      if (c < nb && r < nb && r >= c) P[pk(j0 + r, j0 + c, n2)] = W11[c * CB + r];
    }
    __syncthreads();
  }
  // out = W^T W over lower 16 x 16 tiles
  const int T = (n + 15) >> 4;
  const int ntile = T * (T + 1) / 2;
  for (int t = wv; t < ntile; t += nw) {
    int ti, tj;
    lower_tile(t, ti, tj);
    const int i0 = 16 * ti, c0 = 16 * tj;
    f64x4 acc = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = i0; k0 < n; k0 += 4) {  // k >= max(i, j) >= i0
      const int k = k0 + kl, i = i0 + rl, j = c0 + rl;
      const double a = (k < n && i < n && k >= i) ? P[pk(k, i, n2)] : 0.0;
      const double b = (k < n && j < n && k >= j) ? P[pk(k, j, n2)] : 0.0;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = i0 + kl + 4 * v, j = c0 + rl;
      if (i < n && j < n && i >= j) {
        out[i + (size_t)j * n] = acc[v];
        out[j + (size_t)i * n] = acc[v];
      }
    }
  }
  __syncthreads();
}

// C = A B for n x n column-major A, B in device memory, on FP64 MFMA: one
// wave per 16 x 16 output tile, K in two interleaved chains (independent
// accumulators keep the MFMA pipeline busy), summed at the end; `out(i, j,
// v)` stores a result.
template <class Out>
__device__ void gemm_mfma64(const double* __restrict__ A, const double* __restrict__ B, int n,
                            Out out) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int rl = lane & 15, kl = lane >> 4;
  const int T = (n + 15) >> 4;
  for (int t = wv; t < T * T; t += nw) {
    const int i0 = 16 * (t % T), j0 = 16 * (t / T);
    const int i = i0 + rl, j = j0 + rl;
    f64x4 acc[2] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
    for (int k0 = 0; k0 < n; k0 += 8) {
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int k = k0 + 4 * u + kl;
        const double a = (i < n && k < n) ? A[i + (size_t)k * n] : 0.0;
        const double b = (j < n && k < n) ? B[k + (size_t)j * n] : 0.0;
        acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[u], 0, 0, 0);
      }
    }
    const f64x4 s = acc[0] + acc[1];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = i0 + kl + 4 * v, c = j0 + rl;
      if (r < n && c < n) out(r, c, s[v]);
    }
  }
  __syncthreads();
}

struct IpmOut {
  double* C;
  int n;
  __device__ void operator()(int i, int j, double v) const { C[i + (size_t)j * n] = v; }
};

template <bool P_LDS>
__global__ __launch_bounds__(WG) void k_ipm(NodeView* nodes, int maxit) {
  extern __shared__ double smem[];
  NodeView& nd = nodes[blockIdx.x];
  const int n = nd.n, tid = threadIdx.x;
  const int n2 = n;
  double* red = smem;
  double* tA = red + 64;
  double* tB = tA + 16 * 65;
  __shared__ int flag[2];
  double* P = P_LDS ? (tB + 16 * 65) : nd.ipP;
  // Cholesky and solves with the packed factor P: in LDS one column at a
  // time; in device memory (orders too large for LDS) blocked, with the
  // panel after tB, and the solve's vector in the tA slab
  double* Lp = tB + 16 * 65;
  auto chol = [&]() {
    if constexpr (P_LDS)
      return chol_packed(P, n, n2, flag);
    else
      return chol_blocked(P, n, n2, flag, Lp);
  };
  auto solve = [&](const double* b, double* x) {
    if constexpr (P_LDS)
      chol_solve1(P, n, n2, b, x);
    else
      chol_solve_lds(P, n, n2, b, x, tA);
  };
  auto inverse = [&](double* outm) {
    if constexpr (P_LDS)
      potri_packed(P, n, n2, outm, tA);
    else
      potri_blocked(P, n, n2, outm, Lp, tB);
  };
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
    if constexpr (P_LDS)
      wg_gemm<false, false>(Zi, n, T1, n, nullptr, n, n, n, IpmOut{dX1, n}, tA, tB);
    else
      gemm_mfma64(Zi, T1, n, IpmOut{dX1, n});
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
    if constexpr (P_LDS)
      wg_gemm<false, false>(Zi, n, T1, n, nullptr, n, n, n, IpmOut{dX, n}, tA, tB);
    else
      gemm_mfma64(Zi, T1, n, IpmOut{dX, n});
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

size_t ipm_lds_bytes(int n, bool packed) {
  size_t d = 64 + 2 * 16 * 65;
  d += packed ? (size_t)n * (n + 1) / 2 : chol_blocked_lds_doubles(n);
  return d * sizeof(double);
}

void launch_ipm(NodeView* nodes, int nb, int nmax, size_t lds_limit, hipStream_t st) {
  if (ipm_lds_bytes(nmax, true) <= lds_limit)
    k_ipm<true><<<nb, WG, ipm_lds_bytes(nmax, true), st>>>(nodes, 500);
  else
    k_ipm<false><<<nb, WG, ipm_lds_bytes(nmax, false), st>>>(nodes, 500);
}

}  // namespace madam
