// madam_cert.hip -- certified largest eigenvalue for the safe bound of the
// chosen form (admm.c:718-763 computes it with dsyevr RANGE='I', IL=IU=n).
//
// For A = L - Diag(y) - B^T(s) (in nd.W, full, built by k_build_w with
// use_s): an estimate theta of lambda_max(A) -- Householder reduction to
// tridiagonal form and multisection on the Sturm count of the tridiagonal
// -- then a Cholesky factorisation of M = mu I - A with mu = theta + delta,
// which, if it completes with positive pivots, proves lambda_min(M + E) > 0
// for the rounding error E of floating-point Cholesky, |E| <= gamma_{n+1}
// |R^T||R| (Higham, Accuracy and Stability of Numerical Algorithms, Thm
// 10.3), so
//   lambda_max(A) <= mu + gamma_{n+1} trace(R^T R) + 2u (max_j |mu - a_jj| + |mu|),
// trace(R^T R) = the sum of squared entries of the computed factor and the
// last term covering the rounding of M's diagonal and of the final sum; the
// result is rounded up with nextafter.  delta starts at 1e-11 |A|_F and
// grows 16-fold after each failed factorisation; once it passes 2 |A|_F the
// bound |A|_F >= lambda_max is used instead.  Output: ev[0..n) = the
// certified value (what k_safe_bound takes), off_last = 0.  One workgroup
// of 1024 threads per node: k_safe_hh keeps A as a packed triangle in LDS
// (n up to about 195), k_safe_hhg keeps it in the node's S2 scratch.
#include <stdint.h>

#include "madam_devutil.h"
#include "madam_internal.h"

namespace madam {

// Largest eigenvalue of the symmetric tridiagonal (ta[0..n), tb[1..n)):
// bracket [max a_i, Gershgorin], then multisection on the Sturm count --
// each of the workgroup's 1024 threads counts the eigenvalues above its own
// point of the bracket, five rounds shrink it by 1025^5.  Returns the top of
// the final bracket.  Every thread calls it.
__device__ double tri_top_eig(const double* ta, const double* tb, int n, int* sh_best) {
  const int tid = threadIdx.x;
  double lo = ta[0], hi = ta[0];
  for (int i = 0; i < n; ++i) {
    const double rad = fabs(tb[i]) + (i + 1 < n ? fabs(tb[i + 1]) : 0.0);
    lo = fmax(lo, ta[i]);
    hi = fmax(hi, ta[i] + rad);
  }
  for (int round = 0; round < 5; ++round) {
    const double x = lo + (hi - lo) * ((double)(tid + 1) / 1025.0);
    int cnt = 0;  // eigenvalues of T above x: positive pivots of T - x I
    double q = ta[0] - x;
    if (q > 0.0) ++cnt;
    for (int i = 1; i < n; ++i) {
      const double den = fabs(q) < 1e-300 ? (q < 0.0 ? -1e-300 : 1e-300) : q;
      q = (ta[i] - x) - tb[i] * tb[i] / den;
      if (q > 0.0) ++cnt;
    }
    if (tid == 0) *sh_best = -1;
    __syncthreads();
    if (cnt > 0) atomicMax(sh_best, tid);
    __syncthreads();
    const int t = *sh_best;
    const double nlo = t >= 0 ? lo + (hi - lo) * ((double)(t + 1) / 1025.0) : lo;
    const double nhi = t < 1023 ? lo + (hi - lo) * ((double)(t + 2) / 1025.0) : hi;
    lo = nlo;
    hi = nhi;
    __syncthreads();  // sh_best is reset in the next round only after all have read it
  }
  return hi;
}

// The certificate's bound from a completed factorisation (see the file
// comment): rr = sum of squared factor entries, dmx = max_j |mu - a_jj|.
__device__ __forceinline__ double cert_value(double mu, double rr, double dmx, int n) {
  const double u = 1.1102230246251565e-16;  // 2^-53
  const double gam = (n + 1) * u / (1.0 - (n + 1) * u);
  return nextafter(mu + (1.01 * gam * rr + 2.0 * u * (dmx + fabs(mu))), INFINITY);
}

// Global-memory form for nodes whose packed triangle does not fit LDS.
// Tridiagonalisation (Golub and Van Loan, Alg. 8.3.1) on the full
// symmetric copy G of A in S2: reflection j stages v in LDS, p = tau A v
// reads the trailing block by columns (coalesced; for blocks of at most 512
// rows the columns are split over two halves of the workgroup), and the
// update A <- A - v w^T - w v^T is applied to both triangles so that the
// next product reads columns again; four barriers per reflection.
// Certificate: right-looking Cholesky of mu I - A in S1 (lower triangle),
// the scaled pivot column staged in LDS for the rank-1 update of the
// trailing block, two barriers per column.  LDS: 5 n doubles.
size_t safe_hhg_lds_bytes(int n) { return 5 * (size_t)n * sizeof(double); }

__global__ __launch_bounds__(1024) void k_safe_hhg(NodeView* nodes) {
  extern __shared__ double sm[];
  __shared__ double red[64];
  __shared__ int sh_best;
  NodeView& nd = nodes[blockIdx.x];
  const int n = nd.n, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
  if (n <= 0) return;
  const double* __restrict__ Wm = nd.W;
  double* G = nd.S2;
  double* vv = sm;       // Householder vector / scaled pivot column
  double* pv = vv + n;   // tau A v (first half of the columns)
  double* ph = pv + n;   // second half's partial sums
  double* ta = ph + n;   // diagonal of T
  double* tb = ta + n;   // tb[i] = T(i, i-1)
  const size_t nn = (size_t)n * n;
  double s2[1] = {0.0};
  for (size_t f = tid; f < nn; f += 1024) {
    const double a = Wm[f];
    s2[0] += a * a;
    G[f] = a;
  }
  block_sums<1>(s2, red);  // (its barriers also publish G)
  const double fro = sqrt(s2[0]);

  // ---- tridiagonalisation
  for (int j = 0; j + 2 < n; ++j) {
    const int m = n - 1 - j, o = j + 1;  // trailing block: rows/columns o .. n-1
    const double* colj = G + o + (size_t)j * n;
    for (int i = tid; i < m; i += 1024) vv[i] = colj[i];
    if (tid == 0) ta[j] = G[j + (size_t)j * n];
    __syncthreads();
    double sub = 0.0;
    for (int i = 1 + lane; i < m; i += 64) sub += vv[i] * vv[i];
    sub = wave_sum(sub);
    const double x0 = vv[0];
    double alpha = x0, v0 = 0.0, tau = 0.0;
    if (sub > 0.0) {
      const double nrm = sqrt(x0 * x0 + sub);
      alpha = x0 >= 0.0 ? -nrm : nrm;
      v0 = x0 - alpha;
      tau = 2.0 / (v0 * v0 + sub);
    }
    if (tid == 0) tb[j + 1] = alpha;
    if (tau != 0.0) {  // the same in every thread
      const int halves = m <= 512 ? 2 : 1;
      const int rows = 1024 / halves;
      const int h = tid / rows, r0 = tid - h * rows;
      const int kb = (halves == 2 && h == 1) ? m / 2 : 0;
      const int ke = (halves == 2 && h == 0) ? m / 2 : m;
      for (int r = r0; r < m; r += rows) {
        double acc = 0.0;
        for (int k = kb; k < ke; ++k)
          acc += G[(o + r) + (size_t)(o + k) * n] * (k == 0 ? v0 : vv[k]);
        if (h == 0)
          pv[r] = acc;
        else
          ph[r] = acc;
      }
      __syncthreads();
      for (int r = tid; r < m; r += 1024) pv[r] = tau * (halves == 2 ? pv[r] + ph[r] : pv[r]);
      __syncthreads();
      double kk = 0.0;  // (tau / 2) v^T p, the same in every wave
      for (int i = lane; i < m; i += 64) kk += (i == 0 ? v0 : vv[i]) * pv[i];
      kk = 0.5 * tau * wave_sum(kk);
      for (int c = wv; c < m; c += nw) {  // one wave per column of the block
        const double vc = c == 0 ? v0 : vv[c];
        const double wc = pv[c] - kk * vc;
        double* col = G + o + (size_t)(o + c) * n;
        for (int r = lane; r < m; r += 64) {
          const double vr = r == 0 ? v0 : vv[r];
          col[r] -= vr * wc + (pv[r] - kk * vr) * vc;
        }
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    tb[0] = 0.0;
    if (n >= 2) {
      ta[n - 2] = G[(n - 2) + (size_t)(n - 2) * n];
      tb[n - 1] = G[(n - 1) + (size_t)(n - 2) * n];
    }
    ta[n - 1] = G[(size_t)(n - 1) * (n + 1)];
  }
  __syncthreads();
  const double theta = tri_top_eig(ta, tb, n, &sh_best);

  // ---- certificate
  double* R = nd.S1;
  double delta = 1e-11 * fro;
  double cert = 1.01 * fro;  // lambda_max <= |A|_2 <= |A|_F
  bool done = false;
  while (!done && delta <= 2.0 * fro && fro > 0.0) {
    const double mu = theta + delta;
    __syncthreads();  // the previous attempt's readers of R are done
    for (size_t f = tid; f < nn; f += 1024) {
      const int j = (int)(f / n), i = (int)(f - (size_t)j * n);
      if (i >= j) R[f] = i == j ? mu - Wm[f] : -Wm[f];
    }
    __syncthreads();
    double rr = 0.0;  // this thread's share of the squared factor entries
    bool fail = false;
    for (int j = 0; j < n; ++j) {
      // every thread reads the same pivot (R's diagonal is never rewritten)
      const double piv = R[j + (size_t)j * n];
      if (!(piv > 0.0)) {
        fail = true;
        break;  // uniform
      }
      const double ljj = sqrt(piv);
      if (tid == 0) rr += piv;
      const int m = n - 1 - j;
      double* cj = R + (j + 1) + (size_t)j * n;
      for (int i = tid; i < m; i += 1024) {
        const double l = cj[i] / ljj;
        cj[i] = l;
        vv[i] = l;
        rr += l * l;
      }
      __syncthreads();
      for (int c = wv; c < m; c += nw) {  // trailing lower triangle, one wave per column
        const double lc = vv[c];
        double* col = R + (j + 1) + (size_t)(j + 1 + c) * n;
        for (int r = c + lane; r < m; r += 64) col[r] -= vv[r] * lc;
      }
      __syncthreads();
    }
    if (!fail) {
      double rs[1] = {rr};
      block_sums<1>(rs, red);
      double dmx = 0.0;
      for (int i = tid; i < n; i += 1024) dmx = fmax(dmx, fabs(mu - Wm[(size_t)i * (n + 1)]));
      dmx = block_max(dmx, red);
      cert = cert_value(mu, rs[0], dmx, n);
      done = true;
    } else {
      delta *= 16.0;
    }
  }
  for (int i = tid; i < n; i += 1024) nd.ev[i] = cert;
  if (tid == 0) nd.off_last = 0.0;
}

// The same certificate with a direct estimate, everything in LDS, for nodes
// whose packed lower triangle fits there (n up to about 195).
//
// Estimate: A (packed lower triangle, order n) is reduced to tridiagonal
// T = Q^T A Q by n - 2 Householder reflections H = I - tau v v^T (Golub and
// Van Loan, Matrix Computations, Alg. 8.3.1: p = tau A v,
// w = p - (tau/2)(p^T v) v, A <- A - v w^T - w v^T on the trailing block),
// two barriers per reflection: after the product, after the update.  Every
// wave computes the reflection's scalars itself (same data, same order,
// same bits), so they need no barrier.  theta = largest eigenvalue of T by
// multisection on the Sturm count: each of the 1024 threads counts the
// eigenvalues above its own point of the bracket, five rounds shrink it by
// 1025^5.  Certificate: Cholesky of mu I - A (A re-read from nd.W into the
// same LDS triangle; chol_packed, one barrier per column), mu = theta +
// delta, bound mu + 1.01 gamma_{n+1} trace(R^T R) + 2u (max|mu - a_jj| +
// |mu|): the last term covers the rounding of the diagonal of mu I - A and
// of the final sum.  Output as k_safe_hhg.

// dynamic LDS of k_safe_hh; the kernel's static LDS (reductions, flags) adds
// under 1 KB
size_t safe_hh_lds_bytes(int n) {
  return ((size_t)n * (n + 1) / 2 + 3 * (size_t)n) * sizeof(double);
}

__global__ __launch_bounds__(1024) void k_safe_hh(NodeView* nodes) {
  extern __shared__ double sm[];
  __shared__ double red[64];
  __shared__ int flag[2];
  __shared__ int sh_best;
  NodeView& nd = nodes[blockIdx.x];
  const int n = nd.n, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
  if (n <= 0) return;
  const double* __restrict__ Wm = nd.W;
  double* P = sm;                              // packed lower triangle, order n
  double* ta = P + (size_t)n * (n + 1) / 2;    // diagonal of T
  double* tb = ta + n;                         // tb[i] = T(i, i-1)
  double* pv = tb + n;                         // tau A v of the current reflection
  double s2[1] = {0.0};
  for (int f = tid; f < n * n; f += 1024) {
    const int j = f / n, i = f - j * n;
    const double a = Wm[f];
    s2[0] += a * a;
    if (i >= j) P[pk(i, j, n)] = a;
  }
  block_sums<1>(s2, red);  // (its barriers also publish P)
  const double fro = sqrt(s2[0]);

  // ---- tridiagonalisation
  for (int j = 0; j + 2 < n; ++j) {
    const int m = n - 1 - j;             // entries below the diagonal in column j
    const int cj = pk(j + 1, j, n);      // P[cj + i] = A(j+1+i, j) = v_i (i >= 1)
    const int bs = pk(j + 1, j + 1, n);  // trailing block, packed order m, from here
    double sub = 0.0;
    for (int i = 1 + lane; i < m; i += 64) sub += P[cj + i] * P[cj + i];
    sub = wave_sum(sub);
    const double x0 = P[cj];
    double alpha = x0, v0 = 0.0, tau = 0.0;
    if (sub > 0.0) {
      const double nrm = sqrt(x0 * x0 + sub);
      alpha = x0 >= 0.0 ? -nrm : nrm;
      v0 = x0 - alpha;
      tau = 2.0 / (v0 * v0 + sub);
    }
    if (tid == 0) {
      ta[j] = P[pk(j, j, n)];
      tb[j + 1] = alpha;
    }
    if (tau != 0.0) {  // the same in every thread
      const int g = tid & 7;
      for (int r = tid >> 3; r < m; r += 128) {  // 8 lanes per row
        double acc = 0.0;
        for (int k = g; k < m; k += 8) {
          const double vk = k == 0 ? v0 : P[cj + k];
          acc += P[bs + (r > k ? pk(r, k, m) : pk(k, r, m))] * vk;
        }
        acc += __shfl_xor(acc, 1, 64);
        acc += __shfl_xor(acc, 2, 64);
        acc += __shfl_xor(acc, 4, 64);
        if (g == 0) pv[r] = tau * acc;
      }
      __syncthreads();
      double kk = 0.0;
      for (int i = lane; i < m; i += 64) kk += (i == 0 ? v0 : P[cj + i]) * pv[i];
      kk = 0.5 * tau * wave_sum(kk);
      for (int c = wv; c < m; c += nw) {  // one wave per column of the block
        const double vc = c == 0 ? v0 : P[cj + c];
        const double wc = pv[c] - kk * vc;
        double* col = P + bs + pk(c, c, m);
        for (int r = c + lane; r < m; r += 64) {
          const double vr = r == 0 ? v0 : P[cj + r];
          const double wr = pv[r] - kk * vr;
          col[r - c] -= vr * wc + wr * vc;
        }
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    tb[0] = 0.0;
    if (n >= 2) {
      ta[n - 2] = P[pk(n - 2, n - 2, n)];
      tb[n - 1] = P[pk(n - 1, n - 2, n)];
    }
    ta[n - 1] = P[pk(n - 1, n - 1, n)];
  }
  __syncthreads();

  const double theta = tri_top_eig(ta, tb, n, &sh_best);

  // ---- certificate
  double delta = 1e-11 * fro;
  double cert = 1.01 * fro;  // lambda_max <= |A|_2 <= |A|_F
  bool done = false;
  while (!done && delta <= 2.0 * fro && fro > 0.0) {
    const double mu = theta + delta;
    __syncthreads();  // the previous attempt's readers of P are done
    for (int f = tid; f < n * n; f += 1024) {
      const int j = f / n, i = f - j * n;
      if (i >= j) P[pk(i, j, n)] = i == j ? mu - Wm[f] : -Wm[f];
    }
    if (chol_packed(P, n, n, flag) == 0) {
      double rr[1] = {0.0};
      for (int f = tid; f < n * (n + 1) / 2; f += 1024) rr[0] += P[f] * P[f];
      block_sums<1>(rr, red);
      double dmx = 0.0;
      for (int i = tid; i < n; i += 1024) dmx = fmax(dmx, fabs(mu - Wm[(size_t)i * (n + 1)]));
      dmx = block_max(dmx, red);
      cert = cert_value(mu, rr[0], dmx, n);
      done = true;
    } else {
      delta *= 16.0;
    }
  }
  for (int i = tid; i < n; i += 1024) nd.ev[i] = cert;
  if (tid == 0) nd.off_last = 0.0;
}

void launch_safe_hh(NodeView* nodes, int nb, int nmax, hipStream_t st) {
  k_safe_hh<<<nb, 1024, safe_hh_lds_bytes(nmax), st>>>(nodes);
}

void launch_safe_hhg(NodeView* nodes, int nb, int nmax, hipStream_t st) {
  k_safe_hhg<<<nb, 1024, safe_hhg_lds_bytes(nmax), st>>>(nodes);
}

}  // namespace madam
