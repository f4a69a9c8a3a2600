// madam_ipmt.hip -- the basic-SDP primal-dual predictor-corrector
// interior-point method of ipm_mc_pk.c:28-347 for orders whose Cholesky
// factor does not fit LDS (the one-workgroup k_ipm<true> of madam_ipm.hip
// serves the smaller orders), spread over many workgroups per node: a
// host-driven sequence of launches per iteration over the whole batch, each
// kernel taking its node (and factorisation job) from blockIdx.y and
// returning at once for nodes that have finished.
//
// Per iteration, as the reference (ipm_mc_pk.c:117-301):
//   Z = L L^T (dpotrf), Zi = L^{-T} L^{-1} (dpotri): blocked right-looking
//     Cholesky in IB-column steps, one launch per step; W = L^{-1} by block
//     columns from the right, one launch each; Zi = W^T W on FP64 MFMA.
//   M = Zi .* X, M = R R^T, dy1 = M \ (-e) (dposv): forward and backward
//     substitution by one workgroup per node.
//   dX1 = sym(Zi (-Diag(dy1) X) - X): one launch of FP64 MFMA tile pairs
//     (i, j) and (j, i) that also forms the right operand and symmetrises.
//   dy = M \ (mu diag(Zi) - (Zi .* dX1) dy1), dX = sym(Zi (-Diag(dy) X -
//     Diag(dy1) dX1) + mu Zi + dX1), dy += dy1.
//   Step lengths (ipm_mc_pk.c:273-301): the reference factors X + a dX for
//     a = 1, 0.8, 0.64, ... until dpotrf succeeds (at most 400 tries here,
//     as k_ipm); here TR candidates of each search are factored at once in
//     one batch of jobs, and the first that succeeds is taken -- the same
//     step length.  Then a *= 0.95 if a < 1.
//   Update, mu = <Z, X> / (2n) with the long-step rules, phi = sum y,
//     psi = <L, X>; stop when phi - psi <= 1e-4.
// The host synchronises twice per iteration: after each line-search round
// (to learn whether any search needs more candidates) and after the update
// (to learn which nodes have converged).
//
// Node buffers (NodeView): X, Z (Q), y, L; Zi in S1; dX in W; dX1 in ip1;
// dy and dy1 in ipv; the factor of Z, then of M, in ipP (n x n, lower
// triangle); W = L^{-1} in ip3; line-search candidates in ip2, ip3 (X) and
// S2, V (Z).
#include "madam_devutil.h"
#include "madam_engine.h"
#include "madam_tiles.h"

namespace madam {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int TR = 2;       // step lengths factored per line search and round
constexpr int UPARTS = 64;  // workgroups per node of the update and its sums
constexpr int IPM_MAXIT = 500;

struct IpmCtl {
  double mu, phi, psi;
  double ap, ad;          // accepted step lengths of the current iteration
  double cand_p, cand_d;  // next untried candidates
  int fails_p, fails_d;   // candidates that failed so far
  int fin_p, fin_d;       // line search done
  int it, done, status, need;
};

__device__ __forceinline__ double rdlane(double v, int src) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ void tri_tile(int t, int& ti, int& tj) {
  int a = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
  if ((a + 1) * (a + 2) / 2 <= t) ++a;
  if (a * (a + 1) / 2 > t) --a;
  ti = a;
  tj = t - a * (a + 1) / 2;
}

// ---------------------------------------------------------------- Cholesky

// Cholesky of the leading nbk x nbk block of Dl (column-major, ld IB+1,
// lower triangle) in place, by one wave: lane r < IB holds row r in
// registers; column k is scaled by the pivot and published in cb (IB
// doubles of LDS) for the rank-1 update of the later columns.  Rows and
// columns past nbk are treated as the identity.  Returns 0 or 1 + the
// first column whose pivot is not positive (the same in every lane).
__device__ int chol_diag(double* Dl, int nbk, double* cb) {
  const int lane = threadIdx.x & 63, r = lane & 31;
  double v[IB];
#pragma unroll
  for (int q = 0; q < IB; ++q) {
    const double a = (r < nbk && q < nbk) ? Dl[q * (IB + 1) + r] : (r == q ? 1.0 : 0.0);
    v[q] = q <= r ? a : 0.0;
  }
  int bad = 0;
#pragma unroll
  for (int k = 0; k < IB; ++k) {
    const double piv = rdlane(v[k], k);
    if (!(piv > 0.0)) {
      bad = k + 1;
      break;
    }
    const double s = sqrt(piv);
    const double lk = r == k ? s : (r > k ? v[k] / s : 0.0);
    v[k] = lk;
    if (lane < IB) cb[lane] = lk;
    wave_sync();
#pragma unroll
    for (int j = k + 1; j < IB; ++j) v[j] -= lk * cb[j];  // meaningful for r >= j only
    wave_sync();
  }
  if (!bad && lane < nbk) {
#pragma unroll
    for (int q = 0; q < IB; ++q)
      if (q <= r && q < nbk) Dl[q * (IB + 1) + r] = v[q];
  }
  return bad;
}

// One step p of the right-looking blocked factorisation of every job,
// with the previous step's update folded in (one launch per step):
//  panel workgroups (blockIdx.x < npan): block column p -- the diagonal
//   block and RP rows below it per workgroup -- receives the update by
//   panel p-1 (FP64 MFMA tiles; the diagonal block redundantly in every
//   workgroup), then wave 0 factors the diagonal block and the rows are
//   solved against it (L_r = A_r L_pp^{-T}); workgroup 0 keeps the
//   factored diagonal block in D;
//  trailing workgroups: the columns right of block p receive the update by
//   panel p-1 (lower 16 x 16 tiles, one per wave), and the first of them
//   writes D (the diagonal block of step p-1) into A.
// Launches p = 0 .. ceil(nmax / IB); the last one only writes D back.
// FP64 MFMA lane layout (v_mfma_f64_16x16x4f64): A operand lane l -> (row
// l % 16, k l / 16), B operand lane l -> (k l / 16, col l % 16), result
// register v of lane l -> (row l / 16 + 4 v, col l % 16).
__global__ __launch_bounds__(256) void k_chol_step(CholJob* jobs, int p, int npan) {
  __shared__ double Dl[IB * (IB + 1)];
  __shared__ double Rl[IB * (RP + 1)];
  __shared__ double cb[IB];
  __shared__ int bad;
  CholJob& J = jobs[blockIdx.y];
  if (J.skip || J.fail || (J.stop && *J.stop)) return;
  const int n = J.n, P = (n + IB - 1) / IB;
  if (p > P) return;
  GMEM double* __restrict__ A = gmem(J.A);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int rl = lane & 15, kl = lane >> 4;
  const int j0 = p * IB, jp = j0 - IB;
  if ((int)blockIdx.x < npan) {
    if (p >= P) return;
    const int x = blockIdx.x;
    const int nbk = min(IB, n - j0);
    const int rbeg = j0 + nbk + x * RP;
    const int nrow = max(0, min(RP, n - rbeg));
    if (x > 0 && nrow == 0) return;
    if (tid == 0) bad = 0;
    if (p == 0) {
      for (int f = tid; f < IB * IB; f += 256) {
        const int c = f / IB, r = f - c * IB;
        Dl[c * (IB + 1) + r] = ld0(A, (j0 + r) + (size_t)(j0 + c) * n, c < nbk && r < nbk && r >= c);
      }
      for (int f = tid; f < IB * RP; f += 256) {
        const int c = f / RP, r = f - c * RP;
        Rl[c * (RP + 1) + r] = ld0(A, (rbeg + r) + (size_t)(j0 + c) * n, c < nbk && r < nrow);
      }
    } else {
      // tiles 0..2: the diagonal block (0,0), (1,0), (1,1); then RP/16 row
      // tiles x 2 column tiles of the workgroup's rows
      const int ntl = 3 + 2 * (RP / 16);
      for (int t = wv; t < ntl; t += 4) {
        const bool dg = t < 3;
        int r0, c0;
        if (dg) {
          r0 = j0 + (t >= 1 ? 16 : 0);
          c0 = j0 + (t == 2 ? 16 : 0);
        } else {
          const int u = t - 3;
          r0 = rbeg + 16 * (u >> 1);
          c0 = j0 + 16 * (u & 1);
        }
        if (!dg && r0 >= rbeg + nrow) continue;
        f64x4 acc;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = r0 + kl + 4 * v, c = c0 + rl;
          const bool ok = c < j0 + nbk && (dg ? (r < j0 + nbk && r >= c) : r < rbeg + nrow);
          acc[v] = ld0(A, r + (size_t)c * n, ok);
        }
        const int ra = r0 + rl, ca = c0 + rl;
        double av[IB / 4], bv[IB / 4];
#pragma unroll
        for (int kq = 0; kq < IB / 4; ++kq) {
          const int k = jp + 4 * kq + kl;  // panel p-1 is IB columns wide
          av[kq] = ld0(A, ra + (size_t)k * n, ra < n);
          bv[kq] = ld0(A, ca + (size_t)k * n, ca < n);
        }
#pragma unroll
        for (int kq = 0; kq < IB / 4; ++kq)
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ra < n ? -av[kq] : 0.0, bv[kq], acc, 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = r0 + kl + 4 * v, c = c0 + rl;
          if (dg) {
            if (r < j0 + nbk && c < j0 + nbk && r >= c) Dl[(c - j0) * (IB + 1) + (r - j0)] = acc[v];
          } else if (r < rbeg + nrow && c < j0 + nbk) {
            Rl[(c - j0) * (RP + 1) + (r - rbeg)] = acc[v];
          }
        }
      }
    }
    __syncthreads();
    if (wv == 0) {
      const int e = chol_diag(Dl, nbk, cb);
      if (e && lane == 0) bad = j0 + e;
    }
    __syncthreads();
    if (bad) {
      if (x == 0 && tid == 0) J.fail = bad;
      return;
    }
    if (x == 0) {
      for (int f = tid; f < IB * IB; f += 256) {
        const int c = f / IB, r = f - c * IB;
        J.D[c * IB + r] = (c < nbk && r < nbk && r >= c) ? Dl[c * (IB + 1) + r] : 0.0;
      }
    }
    // rows below the diagonal block: L(r, c) = (A(r, c) - sum_{q<c} L(r, q)
    // L(c, q)) / L(c, c), one thread per row
    if (tid < nrow) {
      double v[IB];
#pragma unroll
      for (int c = 0; c < IB; ++c) v[c] = Rl[c * (RP + 1) + tid];
#pragma unroll
      for (int c = 0; c < IB; ++c) {
        if (c < nbk) {
          double a = v[c];
#pragma unroll
          for (int q = 0; q < c; ++q) a -= v[q] * Dl[q * (IB + 1) + c];
          v[c] = a / Dl[c * (IB + 1) + c];
        }
      }
#pragma unroll
      for (int c = 0; c < IB; ++c)
        if (c < nbk) A[(rbeg + tid) + (size_t)(j0 + c) * n] = v[c];
    }
    return;
  }
  // trailing workgroups
  if (p < 1) return;
  const int x = blockIdx.x - npan;
  if (x == 0) {  // the diagonal block of step p-1 into A (the last block may be narrower)
    for (int f = tid; f < IB * IB; f += 256) {
      const int c = f / IB, r = f - c * IB;
      if (r >= c && jp + r < n) A[(jp + r) + (size_t)(jp + c) * n] = J.D[c * IB + r];
    }
  }
  if (p >= P) return;
  const int t0 = j0 + IB;  // columns right of block p (empty for the last block)
  if (t0 >= n) return;
  const int T = (n - t0 + 15) >> 4;
  const int t = 4 * x + wv;
  if (t >= T * (T + 1) / 2) return;
  int ti, tj;
  tri_tile(t, ti, tj);
  const int r0 = t0 + 16 * ti, c0 = t0 + 16 * tj;
  f64x4 acc;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int r = r0 + kl + 4 * v, c = c0 + rl;
    acc[v] = ld0(A, r + (size_t)c * n, r < n && c < n && r >= c);
  }
  const int ra = r0 + rl, ca = c0 + rl;
  double av[IB / 4], bv[IB / 4];
#pragma unroll
  for (int kq = 0; kq < IB / 4; ++kq) {
    const int k = jp + 4 * kq + kl;
    av[kq] = ld0(A, ra + (size_t)k * n, ra < n);
    bv[kq] = ld0(A, ca + (size_t)k * n, ca < n);
  }
#pragma unroll
  for (int kq = 0; kq < IB / 4; ++kq)
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ra < n ? -av[kq] : 0.0, bv[kq], acc, 0, 0, 0);
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int r = r0 + kl + 4 * v, c = c0 + rl;
    if (r < n && c < n && r >= c) A[r + (size_t)c * n] = acc[v];
  }
}

// ---------------------------------------------------------------- inverse

// W = L^{-1} (L = the factor in ipP, W into ip3), block column p (launches
// from the last block to the first): every workgroup inverts L_pp in
// registers (wave 0; row r in lane r, columns from the right as dtrti2:
// W(r, c) = -(1/L_cc) sum_{c<l<=r} W(r, l) L(l, c), W(c, c) = 1/L_cc);
// workgroup 0 stores it, and workgroup 1 + t the row tile t below:
// W_21 = -(W_22 L_21) W_pp, W_22 (block columns right of p) done in the
// earlier launches.
__global__ __launch_bounds__(256) void k_trtri_step(NodeView* nodes, const IpmCtl* ctl, int p) {
  __shared__ double Ll[IB * (IB + 1)];
  __shared__ double Wl[IB * (IB + 1)];
  __shared__ double Pl[4][16][17];
  const int b = blockIdx.y;
  if (ctl[b].done) return;
  const NodeView& nd = nodes[b];
  const int n = nd.n, P = (n + IB - 1) / IB;
  if (p >= P) return;
  const int j0 = p * IB, nbk = min(IB, n - j0), j1 = j0 + nbk, m2 = n - j1;
  const int x = blockIdx.x;
  if (x > 0 && 16 * (x - 1) >= m2) return;
  const GMEM double* __restrict__ L = gmem(nd.ipP);
  GMEM double* __restrict__ W = gmem(nd.ip3);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int f = tid; f < IB * IB; f += 256) {
    const int c = f / IB, r = f - c * IB;
    const bool in = c < nbk && r < nbk && r >= c;
    const double lv = ld0(L, (j0 + r) + (size_t)(j0 + c) * n, in);
    Ll[c * (IB + 1) + r] = in ? lv : (r == c ? 1.0 : 0.0);
  }
  __syncthreads();
  if (wv == 0) {
    const int r = lane & 31;
    double w[IB];
#pragma unroll
    for (int q = 0; q < IB; ++q) w[q] = 0.0;
#pragma unroll
    for (int c = IB - 1; c >= 0; --c) {
      const double inv = 1.0 / Ll[c * (IB + 1) + c];
      if (r == c) {
        w[c] = inv;
      } else if (r > c) {
        double s = 0.0;
#pragma unroll
        for (int l = c + 1; l < IB; ++l) s += w[l] * Ll[c * (IB + 1) + l];  // w[l] = 0 for l > r
        w[c] = -inv * s;
      }
    }
    if (lane < IB) {
#pragma unroll
      for (int q = 0; q < IB; ++q) Wl[q * (IB + 1) + lane] = w[q];
    }
  }
  __syncthreads();
  if (x == 0) {
    for (int f = tid; f < IB * IB; f += 256) {
      const int c = f / IB, r = f - c * IB;
      if (c < nbk && r < nbk && r >= c) W[(j0 + r) + (size_t)(j0 + c) * n] = Wl[c * (IB + 1) + r];
    }
    return;
  }
  // Y = W_22(rows r0.., cols j1..) L_21: waves (column tile wv & 1, k half wv >> 1)
  const int r0 = j1 + 16 * (x - 1);
  const int rl = lane & 15, kl = lane >> 4;
  const int kend = min(r0 + 16, n);  // W(r, q) = 0 for q > r
  const int klen = kend - j1;
  const int kh = j1 + (((klen + 7) >> 3) << 2);  // split point, a multiple of 4 past j1
  const int ct = wv & 1, half = wv >> 1;
  const int ka = half ? kh : j1, kb = half ? kend : min(kh, kend);
  f64x4 acc = {0.0, 0.0, 0.0, 0.0};
  const int r = r0 + rl, c = 16 * ct + rl;
  // eight k steps' operands in flight at a time
  for (int k0 = ka; k0 < kb; k0 += 32) {
    double a[8], bq[8];
#pragma unroll
    for (int s8 = 0; s8 < 8; ++s8) {
      const int q = k0 + 4 * s8 + kl;
      a[s8] = ld0(W, r + (size_t)q * n, q < kb && r < n && q <= r);
      bq[s8] = ld0(L, q + (size_t)(j0 + c) * n, q < kb && c < nbk);
    }
#pragma unroll
    for (int s8 = 0; s8 < 8; ++s8) {
      if (k0 + 4 * s8 >= kb) break;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s8], bq[s8], acc, 0, 0, 0);
    }
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) Pl[wv][kl + 4 * v][rl] = acc[v];
  __syncthreads();
  // W_21(rho, c) = -sum_{k>=c} Y(rho, k) W_pp(k, c), two outputs per thread
  for (int o = tid; o < 16 * IB; o += 256) {
    const int rho = o / IB, cc = o - rho * IB;
    if (r0 + rho >= n || cc >= nbk) continue;
    double s = 0.0;
    for (int k = cc; k < nbk; ++k) {
      const int kt = k >> 4, kc = k & 15;
      const double y = Pl[kt][rho][kc] + Pl[2 + kt][rho][kc];
      s += y * Wl[cc * (IB + 1) + k];
    }
    W[(r0 + rho) + (size_t)(j0 + cc) * n] = -s;
  }
}

// Zi = W^T W (dpotri's last step), lower 16 x 16 tiles, one per wave, two
// MFMA chains, mirrored: Zi(i, j) = sum_{k >= max(i, j)} W(k, i) W(k, j).
__global__ __launch_bounds__(256) void k_lauum(NodeView* nodes, const IpmCtl* ctl) {
  const int b = blockIdx.y;
  if (ctl[b].done) return;
  const NodeView& nd = nodes[b];
  const int n = nd.n, T = (n + 15) >> 4;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int t = 4 * blockIdx.x + wv;
  if (t >= T * (T + 1) / 2) return;
  int ti, tj;
  tri_tile(t, ti, tj);
  const int i0 = 16 * ti, c0 = 16 * tj;
  const int rl = lane & 15, kl = lane >> 4;
  const GMEM double* __restrict__ W = gmem(nd.ip3);
  GMEM double* __restrict__ Zi = gmem(nd.S1);
  const int i = i0 + rl, j = c0 + rl;
  f64x4 acc[2] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
  // four pairs of k steps' operands in flight at a time (chain u takes the
  // steps k0 + 4 u + 8 s, as before)
  for (int k0 = i0; k0 < n; k0 += 32) {
    double a[4][2], bq[4][2];
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4)
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int k = k0 + 8 * s4 + 4 * u + kl;
        a[s4][u] = ld0(W, k + (size_t)i * n, k < n && i < n && k >= i);
        bq[s4][u] = ld0(W, k + (size_t)j * n, k < n && j < n && k >= j);
      }
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4) {
      if (k0 + 8 * s4 >= n) break;
#pragma unroll
      for (int u = 0; u < 2; ++u)
        acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s4][u], bq[s4][u], acc[u], 0, 0, 0);
    }
  }
  const f64x4 s = acc[0] + acc[1];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int ii = i0 + kl + 4 * v, jj = c0 + rl;
    if (ii < n && jj < n && ii >= jj) {
      Zi[ii + (size_t)jj * n] = s[v];
      Zi[jj + (size_t)ii * n] = s[v];
    }
  }
}

// ---------------------------------------------------------------- elementwise

// y = 1.1 + column sums of |L| (ipm_mc_pk.c:61-68).
__global__ __launch_bounds__(256) void k_ipmt_init_y(NodeView* nodes) {
  const NodeView& nd = nodes[blockIdx.y];
  const int n = nd.n, i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  for (int j = 0; j < n; ++j) acc += fabs(nd.L[j + (size_t)i * n]);
  nd.y[i] = acc + 1.1;
}

// Workgroup sums of three values into part[(node UPARTS + x) 3 + q].
__device__ __forceinline__ void part_sums(double a, double b, double c, double* part) {
  __shared__ double red[3][4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  a = wave_sum(a);
  b = wave_sum(b);
  c = wave_sum(c);
  if (lane == 0) {
    red[0][wv] = a;
    red[1][wv] = b;
    red[2][wv] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double* o = part + ((size_t)blockIdx.y * UPARTS + blockIdx.x) * 3;
#pragma unroll
    for (int q = 0; q < 3; ++q) o[q] = ((red[q][0] + red[q][1]) + red[q][2]) + red[q][3];
  }
}

// X = I, Z = Diag(y) - L, and the sums <Z, X>, <L, X>, sum y.
__global__ __launch_bounds__(256) void k_ipmt_init_xz(NodeView* nodes, double* part) {
  const NodeView& nd = nodes[blockIdx.y];
  const int n = nd.n;
  const size_t nn = (size_t)n * n, ch = (nn + UPARTS - 1) / UPARTS;
  const size_t f0 = blockIdx.x * ch, f1 = f0 + ch < nn ? f0 + ch : nn;
  double zx = 0.0, lx = 0.0, sy = 0.0;
  for (size_t f = f0 + threadIdx.x; f < f1; f += 256) {
    const int j = (int)(f / n), i = (int)(f - (size_t)j * n);
    const double xv = i == j ? 1.0 : 0.0;
    const double zv = (i == j ? nd.y[i] : 0.0) - nd.L[f];
    nd.X[f] = xv;
    nd.Q[f] = zv;
    zx += zv * xv;
    lx += nd.L[f] * xv;
  }
  const int ych = (n + UPARTS - 1) / UPARTS;
  for (int i = blockIdx.x * ych + threadIdx.x; i < min(n, (int)(blockIdx.x + 1) * ych); i += 256)
    sy += nd.y[i];
  part_sums(zx, lx, sy, part);
}

// The factorisation job of node b: lower triangle of ipP = Z (which 0) or
// Zi .* X (which 1, M of ipm_mc_pk.c:143-145); the job is reset.
__global__ __launch_bounds__(256) void k_ipmt_load(NodeView* nodes, const IpmCtl* ctl, CholJob* jobs,
                                                   int which) {
  const int b = blockIdx.y;
  if (ctl[b].done) return;
  const NodeView& nd = nodes[b];
  const int n = nd.n;
  const size_t nn = (size_t)n * n;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    jobs[b].fail = 0;
    jobs[b].skip = 0;
  }
  double* __restrict__ F = nd.ipP;
  for (size_t f = blockIdx.x * 256 + threadIdx.x; f < nn; f += (size_t)gridDim.x * 256) {
    const int j = (int)(f / n), i = (int)(f - (size_t)j * n);
    if (i < j) continue;
    F[f] = which == 0 ? nd.Q[f] : nd.S1[f] * nd.X[f];
  }
}

// A failed factorisation stops the node with the reference's failure point
// (1: Z, 3: M), as k_ipm.
__global__ void k_ipmt_check(NodeView* nodes, IpmCtl* ctl, const CholJob* jobs, int nb, int code) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  IpmCtl& c = ctl[b];
  if (c.done || !jobs[b].fail) return;
  c.status = code;
  c.done = 1;
  NodeView& nd = nodes[b];
  nd.f = c.phi;
  nd.ipm_iters = c.it + 1;
  nd.ipm_status = code;
  nd.eig_warm = 0;
}

// (R R^T) x = b with the factor R of M in ipP, one workgroup per node:
// forward substitution by blocks of IB rows (wave 0 solves the diagonal
// block in registers, then every thread updates rows below with the
// block's IB columns), backward substitution by blocks from the bottom
// (the sums over the solved rows below, one wave per column, then wave 0
// solves the block).  which 0: b = -e, x -> dy1; which 1: b = dy, x -> dy.
__global__ __launch_bounds__(256) void k_ipmt_solve(NodeView* nodes, const IpmCtl* ctl, int which) {
  __shared__ double z[MADAM_NMAX];
  __shared__ double sv[IB];
  const int b = blockIdx.x;
  if (ctl[b].done) return;
  const NodeView& nd = nodes[b];
  const int n = nd.n, P = (n + IB - 1) / IB;
  const GMEM double* __restrict__ R = gmem(nd.ipP);
  double* dy = nd.ipv;
  double* dy1 = nd.ipv + n;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int i = tid; i < n; i += 256) z[i] = which == 0 ? -1.0 : dy[i];
  __syncthreads();
  for (int p = 0; p < P; ++p) {
    const int j0 = p * IB, nbk = min(IB, n - j0), j1 = j0 + nbk;
    if (wv == 0) {
      const int r = lane & 31;
      double v[IB];
#pragma unroll
      for (int q = 0; q < IB; ++q)
        v[q] = ld0(R, (j0 + r) + (size_t)(j0 + q) * n, r < nbk && q < nbk && q <= r);
      double zr = r < nbk ? z[j0 + r] : 0.0;
#pragma unroll
      for (int c = 0; c < IB; ++c) {
        if (c < nbk) {
          const double zc = rdlane(zr, c) / rdlane(v[c], c);
          if (r == c)
            zr = zc;
          else if (r > c)
            zr -= v[c] * zc;
        }
      }
      if (lane < nbk) z[j0 + lane] = zr;
    }
    __syncthreads();
    for (int i = j1 + tid; i < n; i += 256) {
      // the block's IB columns loaded together, subtracted in column order
      double rv[IB];
#pragma unroll
      for (int c = 0; c < IB; ++c) rv[c] = ld0(R, i + (size_t)(j0 + c) * n, c < nbk);
      double s = z[i];
#pragma unroll
      for (int c = 0; c < IB; ++c)
        if (c < nbk) s -= rv[c] * z[j0 + c];
      z[i] = s;
    }
    __syncthreads();
  }
  for (int p = P - 1; p >= 0; --p) {
    const int j0 = p * IB, nbk = min(IB, n - j0), j1 = j0 + nbk;
    for (int kk = wv; kk < nbk; kk += 4) {
      const int k = j0 + kk;
      double part = 0.0;
      // four rows' loads in flight, added in row order
      for (int i = j1 + lane; i < n; i += 256) {
        double rv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) rv[u] = ld0(R, (i + 64 * u) + (size_t)k * n, i + 64 * u < n);
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (i + 64 * u < n) part += rv[u] * z[i + 64 * u];
      }
      part = wave_sum(part);
      if (lane == 0) sv[kk] = part;
    }
    __syncthreads();
    if (wv == 0) {
      const int r = lane & 31;
      double u[IB];  // column r of the diagonal block: u[q] = R(j0 + q, j0 + r)
#pragma unroll
      for (int q = 0; q < IB; ++q)
        u[q] = ld0(R, (j0 + q) + (size_t)(j0 + r) * n, r < nbk && q < nbk && q >= r);
      double acc = r < nbk ? z[j0 + r] - sv[r] : 0.0;
#pragma unroll
      for (int i = IB - 1; i >= 0; --i) {
        if (i < nbk) {
          const double xi = rdlane(acc, i) / rdlane(u[i], i);
          if (r == i)
            acc = xi;
          else if (r < i)
            acc -= u[i] * xi;
        }
      }
      if (lane < nbk) z[j0 + lane] = acc;
    }
    __syncthreads();
  }
  double* out = which == 0 ? dy1 : dy;
  for (int i = tid; i < n; i += 256) out[i] = z[i];
}

// The two n^3 products of an iteration with their operands formed on the
// fly and the result symmetrised: tile pair (i, j), (j, i) of lower tile t,
// waves (tile, K half).  which 0 (predictor): dX1 = sym(Zi T - X), T(k, c)
// = -dy1_k X(k, c); which 1 (corrector): dX = sym(Zi T + (mu Zi + dX1)),
// T(k, c) = -dy_k X(k, c) + (-dy1_k dX1(k, c)).  X and dX1 are symmetric:
// T's column c is read as row c (contiguous).
__global__ __launch_bounds__(256) void k_ipmt_gemm(NodeView* nodes, const IpmCtl* ctl, int which) {
  __shared__ double Cl[4][16][17];
  const int b = blockIdx.y;
  if (ctl[b].done) return;
  const NodeView& nd = nodes[b];
  const int n = nd.n, T = (n + 15) >> 4;
  const int t = blockIdx.x;
  if (t >= T * (T + 1) / 2) return;
  int ti, tj;
  tri_tile(t, ti, tj);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int rl = lane & 15, kl = lane >> 4;
  const int rb = 16 * (wv < 2 ? ti : tj), cbs = 16 * (wv < 2 ? tj : ti);
  const int kh = ((n + 7) >> 3) << 2;
  const int ka = (wv & 1) ? kh : 0, kb = (wv & 1) ? n : min(kh, n);
  const GMEM double* __restrict__ Zi = gmem(nd.S1);
  const GMEM double* __restrict__ X = gmem(nd.X);
  const GMEM double* __restrict__ dX1 = gmem(nd.ip1);
  const GMEM double* __restrict__ dy = gmem(nd.ipv);
  const GMEM double* __restrict__ dy1 = gmem(nd.ipv + n);
  const int r = rb + rl, c = cbs + rl;
  f64x4 acc = {0.0, 0.0, 0.0, 0.0};
  constexpr int NSB = 8;
  for (int k0 = ka; k0 < kb; k0 += 4 * NSB) {
    double a[NSB], bv[NSB];
#pragma unroll
    for (int s = 0; s < NSB; ++s) {
      const int k = k0 + 4 * s + kl;
      const bool ok = k < kb;
      const bool okc = ok && c < n;
      a[s] = ld0(Zi, r + (size_t)k * n, ok && r < n);
      const size_t g = c + (size_t)k * n;
      const double xg = ld0(X, g, okc), d1 = ld0(dy1, k, okc);
      double v;
      if (which == 0) {
        v = -d1 * xg;
      } else {
        const double d0 = ld0(dy, k, okc), e1 = ld0(dX1, g, okc);
        v = -d0 * xg + (-d1 * e1);
      }
      bv[s] = okc ? v : 0.0;
    }
#pragma unroll
    for (int s = 0; s < NSB; ++s) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], bv[s], acc, 0, 0, 0);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) Cl[wv][kl + 4 * v][rl] = acc[v];
  __syncthreads();
  const int rho = tid >> 4, gam = tid & 15;
  const int i = 16 * ti + rho, j = 16 * tj + gam;
  if (i >= n || j >= n) return;
  double aij = Cl[0][rho][gam] + Cl[1][rho][gam];
  double aji = Cl[2][gam][rho] + Cl[3][gam][rho];
  const size_t fij = i + (size_t)j * n, fji = j + (size_t)i * n;
  double* out;
  if (which == 0) {
    aij -= X[fij];
    aji -= X[fji];
    out = nd.ip1;
  } else {
    const double mu = ctl[b].mu;
    aij += mu * Zi[fij] + dX1[fij];
    aji += mu * Zi[fji] + dX1[fji];
    out = nd.W;
  }
  const double sv = 0.5 * (aij + aji);
  out[fij] = sv;
  out[fji] = sv;
}

// dy = -(Zi .* dX1) dy1 + mu diag(Zi) (dsymv of ipm_mc_pk.c:205), one row
// per wave; Zi and dX1 are symmetric, so row i is read as column i.
__global__ __launch_bounds__(256) void k_ipmt_matvec(NodeView* nodes, const IpmCtl* ctl) {
  const int b = blockIdx.y;
  if (ctl[b].done) return;
  const NodeView& nd = nodes[b];
  const int n = nd.n, lane = threadIdx.x & 63;
  const int i = 4 * blockIdx.x + (threadIdx.x >> 6);
  if (i >= n) return;
  const double* Zi = nd.S1 + (size_t)i * n;
  const double* D1 = nd.ip1 + (size_t)i * n;
  const double* dy1 = nd.ipv + n;
  double part = 0.0;
  for (int j = lane; j < n; j += 64) {
    const double tij = Zi[j] * D1[j];
    part += tij * dy1[j];
  }
  part = wave_sum(part);
  if (lane == 0) nd.ipv[i] = -part + ctl[b].mu * Zi[i];
}

// dy += dy1.
__global__ __launch_bounds__(256) void k_ipmt_dy(NodeView* nodes, const IpmCtl* ctl) {
  const int b = blockIdx.x;
  if (ctl[b].done) return;
  const NodeView& nd = nodes[b];
  for (int i = threadIdx.x; i < nd.n; i += 256) nd.ipv[i] += nd.ipv[nd.n + i];
}

// Line-search candidates of one round: job (b, s, t) = node b, search s
// (0: X + a dX, 1: Z + a Diag(dy)), the t-th untried candidate a = cand
// 0.8^t (the reference's repeated a *= 0.8); lower triangles only.
__global__ __launch_bounds__(256) void k_ipmt_trials(NodeView* nodes, const IpmCtl* ctl,
                                                     CholJob* tjobs) {
  const int jb = blockIdx.y, b = jb / (2 * TR), s = (jb / TR) & 1, t = jb % TR;
  CholJob& J = tjobs[jb];
  const IpmCtl& c = ctl[b];
  const bool idle = c.done || (s == 0 ? c.fin_p : c.fin_d);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    J.fail = 0;
    J.skip = idle ? 1 : 0;
  }
  if (idle) return;
  double a = s == 0 ? c.cand_p : c.cand_d;
  for (int q = 0; q < t; ++q) a *= 0.8;
  const NodeView& nd = nodes[b];
  const int n = nd.n;
  const size_t nn = (size_t)n * n;
  double* __restrict__ A = J.A;
  for (size_t f = blockIdx.x * 256 + threadIdx.x; f < nn; f += (size_t)gridDim.x * 256) {
    const int j = (int)(f / n), i = (int)(f - (size_t)j * n);
    if (i < j) continue;
    if (s == 0) {
      A[f] = nd.X[f] + a * nd.W[f];
    } else {
      double v = nd.Q[f];
      if (i == j) v += a * nd.ipv[i];
      A[f] = v;
    }
  }
}

// The first candidate of each search whose factorisation succeeded (or the
// 401st after 400 failures, as k_ipm); a *= 0.95 if a < 1.  need = some
// search of the node still open.
__global__ void k_ipmt_pick(IpmCtl* ctl, const CholJob* tjobs, int nb) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  IpmCtl& c = ctl[b];
  if (c.done) {
    c.need = 0;
    return;
  }
  for (int s = 0; s < 2; ++s) {
    int& fin = s == 0 ? c.fin_p : c.fin_d;
    if (fin) continue;
    double& cand = s == 0 ? c.cand_p : c.cand_d;
    int& fails = s == 0 ? c.fails_p : c.fails_d;
    for (int t = 0; t < TR && !fin; ++t) {
      if (fails >= 400 || !tjobs[(b * 2 + s) * TR + t].fail) {
        fin = 1;
      } else {
        cand *= 0.8;
        ++fails;
      }
    }
    if (!fin && fails >= 400) fin = 1;
    if (fin) {
      double a = cand;
      if (a < 1.0) a *= 0.95;
      (s == 0 ? c.ap : c.ad) = a;
    }
  }
  c.need = !(c.fin_p && c.fin_d);
}

// X += ap dX, y += ad dy, Z_ii += ad dy_i, and the sums <Z, X>, <L, X>,
// sum y of the updated values.
__global__ __launch_bounds__(256) void k_ipmt_update(NodeView* nodes, const IpmCtl* ctl,
                                                     double* part) {
  const int b = blockIdx.y;
  if (ctl[b].done) return;
  const NodeView& nd = nodes[b];
  const int n = nd.n;
  const double ap = ctl[b].ap, ad = ctl[b].ad;
  const size_t nn = (size_t)n * n, ch = (nn + UPARTS - 1) / UPARTS;
  const size_t f0 = blockIdx.x * ch, f1 = f0 + ch < nn ? f0 + ch : nn;
  double zx = 0.0, lx = 0.0, sy = 0.0;
  for (size_t f = f0 + threadIdx.x; f < f1; f += 256) {
    const double xv = nd.X[f] + ap * nd.W[f];
    nd.X[f] = xv;
    const int j = (int)(f / n), i = (int)(f - (size_t)j * n);
    double zv = nd.Q[f];
    if (i == j) {
      zv += ad * nd.ipv[i];
      nd.Q[f] = zv;
    }
    zx += zv * xv;
    lx += nd.L[f] * xv;
  }
  const int ych = (n + UPARTS - 1) / UPARTS;
  for (int i = blockIdx.x * ych + threadIdx.x; i < min(n, (int)(blockIdx.x + 1) * ych); i += 256) {
    const double yv = nd.y[i] + ad * nd.ipv[i];
    nd.y[i] = yv;
    sy += yv;
  }
  part_sums(zx, lx, sy, part);
}

// mu, phi, psi from the sums (in workgroup order), the long-step rules,
// the stopping test of ipm_mc_pk.c:117, the node's outputs; the line
// search state is reset for the next iteration.
__global__ void k_ipmt_finish(NodeView* nodes, IpmCtl* ctl, const double* part, int nb, int first) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  IpmCtl& c = ctl[b];
  if (!first && c.done) return;
  NodeView& nd = nodes[b];
  double zx = 0.0, lx = 0.0, sy = 0.0;
  const double* pp = part + (size_t)b * UPARTS * 3;
  for (int x = 0; x < UPARTS; ++x) {
    zx += pp[3 * x];
    lx += pp[3 * x + 1];
    sy += pp[3 * x + 2];
  }
  double mu = zx / (2.0 * nd.n);
  if (first) {
    c.it = 0;
    c.status = 0;
  } else {
    if (c.ap + c.ad > 1.6) mu *= 0.5;
    if (c.ap + c.ad > 1.9) mu *= 0.2;
    c.it += 1;
  }
  c.mu = mu;
  c.phi = sy;
  c.psi = lx;
  c.done = (c.phi - c.psi > 1e-4 && c.it < IPM_MAXIT) ? 0 : 1;
  c.fin_p = c.fin_d = 0;
  c.fails_p = c.fails_d = 0;
  c.cand_p = c.cand_d = 1.0;
  c.need = 0;
  nd.f = c.phi;
  nd.ipm_iters = c.it;
  nd.ipm_status = c.status;
  nd.eig_warm = 0;
}

void chol_launches(CholJob* jobs, int nj, int nmax, hipStream_t st) {
  const int P = (nmax + IB - 1) / IB;
  for (int p = 0; p <= P; ++p) {
    const int below = std::max(0, nmax - (p + 1) * IB);
    const int npan = std::max(1, (below + RP - 1) / RP);
    const int T = (below + 15) >> 4;
    const int ntr = std::max(1, (T * (T + 1) / 2 + 3) / 4);
    k_chol_step<<<dim3(npan + ntr, nj), 256, 0, st>>>(jobs, p, npan);
  }
}

}  // namespace madam

namespace mh {

using madam::CholJob;
using madam::IB;
using madam::IpmCtl;
using madam::TR;
using madam::UPARTS;

static DBuf<IpmCtl> g_ictl;
static DBuf<CholJob> g_ijobs;
static DBuf<double> g_ipart, g_iD;
static IpmCtl* g_hctl = nullptr;
static int g_hctl_cap = 0;

void release_ipm_tiles() {
  g_ictl.release();
  g_ijobs.release();
  g_ipart.release();
  g_iD.release();
  if (g_hctl) (void)hipHostFree(g_hctl);
  g_hctl = nullptr;
  g_hctl_cap = 0;
}


int ipm_tiles(NodeView* d_nodes, const NodeView* hv, int nb, int nmax, hipStream_t st) {
  static_assert(TR == 2, "candidate buffers: ip2, ip3 (X) and S2, V (Z)");
  const int njob = nb + 2 * TR * nb;
  int rc;
  if ((rc = g_ictl.ensure(nb)) || (rc = g_ijobs.ensure(njob)) ||
      (rc = g_ipart.ensure((size_t)nb * UPARTS * 3)) || (rc = g_iD.ensure((size_t)njob * IB * IB)))
    return rc;
  if (g_hctl_cap < nb) {
    if (g_hctl) (void)hipHostFree(g_hctl);
    g_hctl = nullptr;
    g_hctl_cap = 0;
    HIP_TRY(hipHostMalloc((void**)&g_hctl, (size_t)nb * sizeof(IpmCtl), hipHostMallocDefault));
    g_hctl_cap = nb;
  }
  std::vector<CholJob> J(njob);
  // a job stops with its node: the address of ctl[b].done on the device
  auto stop = [&](int b) {
    return (const int*)((const char*)(g_ictl.p + b) + offsetof(IpmCtl, done));
  };
  for (int b = 0; b < nb; ++b)
    J[b] = {hv[b].ipP, g_iD.p + (size_t)b * IB * IB, stop(b), hv[b].n, 0, 0, 0};
  for (int b = 0; b < nb; ++b)
    for (int s = 0; s < 2; ++s)
      for (int t = 0; t < TR; ++t) {
        const int q = (b * 2 + s) * TR + t, jq = nb + q;
        double* A = s == 0 ? (t == 0 ? hv[b].ip2 : hv[b].ip3) : (t == 0 ? hv[b].S2 : hv[b].V);
        J[jq] = {A, g_iD.p + (size_t)jq * IB * IB, stop(b), hv[b].n, 0, 1, 0};
      }
  if ((rc = up(g_ijobs.p, J.data(), (size_t)njob, st))) return rc;
  CholJob* mjobs = g_ijobs.p;
  CholJob* tjobs = g_ijobs.p + nb;
  IpmCtl* ctl = g_ictl.p;
  double* part = g_ipart.p;
  const int tpb = 64, nbt = (nb + tpb - 1) / tpb;
  const int T = (nmax + 15) >> 4, npair = T * (T + 1) / 2;
  const int P = (nmax + IB - 1) / IB;
  madam::k_ipmt_init_y<<<dim3((nmax + 255) / 256, nb), 256, 0, st>>>(d_nodes);
  madam::k_ipmt_init_xz<<<dim3(UPARTS, nb), 256, 0, st>>>(d_nodes, part);
  madam::k_ipmt_finish<<<nbt, tpb, 0, st>>>(d_nodes, ctl, part, nb, 1);
  auto read_ctl = [&]() -> int {
    HIP_TRY(hipMemcpyAsync(g_hctl, ctl, (size_t)nb * sizeof(IpmCtl), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return MADAM_OK;
  };
  if ((rc = read_ctl())) return rc;
  const int nload = std::min(256, (nmax * nmax + 255) / 256);
  for (int it = 0; it < madam::IPM_MAXIT; ++it) {
    bool all = true;
    for (int b = 0; b < nb; ++b) all = all && g_hctl[b].done;
    if (all) break;
    // Zi = Z^{-1}
    madam::k_ipmt_load<<<dim3(nload, nb), 256, 0, st>>>(d_nodes, ctl, mjobs, 0);
    madam::chol_launches(mjobs, nb, nmax, st);
    madam::k_ipmt_check<<<nbt, tpb, 0, st>>>(d_nodes, ctl, mjobs, nb, 1);
    for (int p = P - 1; p >= 0; --p) {
      const int m2 = std::max(0, nmax - (p + 1) * IB);
      madam::k_trtri_step<<<dim3(1 + (m2 + 15) / 16, nb), 256, 0, st>>>(d_nodes, ctl, p);
    }
    madam::k_lauum<<<dim3((npair + 3) / 4, nb), 256, 0, st>>>(d_nodes, ctl);
    // predictor and corrector
    madam::k_ipmt_load<<<dim3(nload, nb), 256, 0, st>>>(d_nodes, ctl, mjobs, 1);
    madam::chol_launches(mjobs, nb, nmax, st);
    madam::k_ipmt_check<<<nbt, tpb, 0, st>>>(d_nodes, ctl, mjobs, nb, 3);
    madam::k_ipmt_solve<<<nb, 256, 0, st>>>(d_nodes, ctl, 0);
    madam::k_ipmt_gemm<<<dim3(npair, nb), 256, 0, st>>>(d_nodes, ctl, 0);
    madam::k_ipmt_matvec<<<dim3((nmax + 3) / 4, nb), 256, 0, st>>>(d_nodes, ctl);
    madam::k_ipmt_solve<<<nb, 256, 0, st>>>(d_nodes, ctl, 1);
    madam::k_ipmt_gemm<<<dim3(npair, nb), 256, 0, st>>>(d_nodes, ctl, 1);
    madam::k_ipmt_dy<<<nb, 256, 0, st>>>(d_nodes, ctl);
    // step lengths
    for (int round = 0;; ++round) {
      madam::k_ipmt_trials<<<dim3(nload, 2 * TR * nb), 256, 0, st>>>(d_nodes, ctl, tjobs);
      madam::chol_launches(tjobs, 2 * TR * nb, nmax, st);
      madam::k_ipmt_pick<<<nbt, tpb, 0, st>>>(ctl, tjobs, nb);
      if ((rc = read_ctl())) return rc;
      bool need = false;
      for (int b = 0; b < nb; ++b) need = need || g_hctl[b].need;
      if (!need) break;
      if (round > 400) {
        set_err("interior-point line search did not finish");
        return MADAM_E_HIP;
      }
    }
    madam::k_ipmt_update<<<dim3(UPARTS, nb), 256, 0, st>>>(d_nodes, ctl, part);
    madam::k_ipmt_finish<<<nbt, tpb, 0, st>>>(d_nodes, ctl, part, nb, 0);
    HIP_TRY(hipGetLastError());
    if ((rc = read_ctl())) return rc;
  }
  return MADAM_OK;
}

}  // namespace mh
