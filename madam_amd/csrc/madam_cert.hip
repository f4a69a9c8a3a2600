// madam_cert.hip -- certified largest eigenvalue for the safe bound of the
// chosen form (admm.c:718-763 computes it with dsyevr RANGE='I', IL=IU=n).
//
// For A = L - Diag(y) - B^T(s) (in nd.W, full, built by k_build_w with
// use_s): Lanczos (full reorthogonalisation, at most 160 steps) estimates
// theta <= lambda_max(A) (the largest Ritz value); then a
// Cholesky factorisation of M = mu I - A with mu = theta + delta, if it
// completes with positive pivots, proves lambda_min(M + E) > 0 for the
// rounding error E of floating-point Cholesky, |E| <= gamma_{n+1} |R^T||R|
// (Higham, Accuracy and Stability of Numerical Algorithms, Thm 10.3), so
//   lambda_max(A) <= mu + gamma_{n+1} trace(R^T R),  trace(R^T R) = sum of
//   squared entries of the computed factor.
// delta starts at 1e-11 |A|_F and grows 16-fold after each failed
// factorisation; once it passes 2 |A|_F the bound |A|_F >= lambda_max is
// used instead.  Result: ev[0..n) = the certified value (what k_safe_bound
// takes the maximum of), off_last = 0.  One workgroup per node; O(n^2) per
// Lanczos step, O(n^3 / 3) per factorisation, against O(n^3) per Jacobi
// sweep for the eigenvalue route.
#include <stdint.h>

#include "madam_devutil.h"
#include "madam_internal.h"

namespace madam {

constexpr int LZ_MAX = 160;  // Lanczos steps

// largest eigenvalue of the symmetric tridiagonal (a[0..k), b[1..k)) by
// bisection on the Sturm count, to a few ulps
__device__ double tri_max_eig(const double* a, const double* b, int k) {
  double lo = a[0], hi = a[0];
  for (int i = 0; i < k; ++i) {
    const double r = (i > 0 ? fabs(b[i]) : 0.0) + (i + 1 < k ? fabs(b[i + 1]) : 0.0);
    lo = fmin(lo, a[i] - r);
    hi = fmax(hi, a[i] + r);
  }
  for (int it = 0; it < 200 && hi - lo > 1e-15 * fmax(fabs(lo), fabs(hi)); ++it) {
    const double x = 0.5 * (lo + hi);
    // number of eigenvalues > x
    int cnt = 0;
    double q = a[0] - x;
    if (q > 0) ++cnt;
    for (int i = 1; i < k; ++i) {
      const double den = fabs(q) < 1e-300 ? 1e-300 : q;
      q = a[i] - x - b[i] * b[i] / den;
      if (q > 0) ++cnt;
    }
    if (cnt > 0)
      lo = x;
    else
      hi = x;
  }
  return hi;
}

__global__ __launch_bounds__(1024) void k_safe_lc(NodeView* nodes) {
  extern __shared__ double sm[];
  __shared__ double red[64];
  __shared__ double alpha[LZ_MAX], beta[LZ_MAX + 1];
  __shared__ double sh_theta;
  __shared__ int sh_fail;
  NodeView& nd = nodes[blockIdx.x];
  const int n = nd.n, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
  const double* __restrict__ A = nd.W;
  double* v = sm;         // current Lanczos vector
  double* wk = sm + n;
  double s2[1] = {0.0};
  for (int f = tid; f < n * n; f += WG) s2[0] += A[f] * A[f];
  block_sums<1>(s2, red);
  const double fro = sqrt(s2[0]);
  // Lanczos with full reorthogonalisation (classical Gram-Schmidt, twice)
  // against the basis kept in S1 (n x kmax <= n x n), from a pseudo-random
  // start vector (splitmix64 of the index, fixed)
  double* Qb = nd.S1;
  const int kmax = n < LZ_MAX ? n : LZ_MAX;
  __shared__ double cdot[LZ_MAX];
  double nn[1] = {0.0};
  for (int i = tid; i < n; i += WG) {
    uint64_t z = 0x2545F4914F6CDD1Dull + (uint64_t)i * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const double x = (double)(z >> 11) * 0x1.0p-53 - 0.5;
    v[i] = x;
    nn[0] += x * x;
  }
  block_sums<1>(nn, red);
  const double inv0 = 1.0 / sqrt(nn[0]);
  for (int i = tid; i < n; i += WG) {
    v[i] *= inv0;
    Qb[i] = v[i];
  }
  __syncthreads();
  int k = 0;
  for (; k < kmax; ++k) {
    // wk = A v, one wave per row
    for (int i = wv; i < n; i += nw) {
      double a = 0.0;
      for (int j = lane; j < n; j += 64) a += A[i + (size_t)j * n] * v[j];
      a = wave_sum(a);
      if (lane == 0) wk[i] = a;
    }
    __syncthreads();
    double al[1] = {0.0};
    for (int i = tid; i < n; i += WG) al[0] += wk[i] * v[i];
    block_sums<1>(al, red);
    for (int pass = 0; pass < 2; ++pass) {
      // c_j = q_j . wk for j <= k (one wave per basis vector), wk -= Q c
      for (int j = wv; j <= k; j += nw) {
        double a = 0.0;
        for (int i = lane; i < n; i += 64) a += Qb[i + (size_t)j * n] * wk[i];
        a = wave_sum(a);
        if (lane == 0) cdot[j] = a;
      }
      __syncthreads();
      for (int i = tid; i < n; i += WG) {
        double a = wk[i];
        for (int j = 0; j <= k; ++j) a -= cdot[j] * Qb[i + (size_t)j * n];
        wk[i] = a;
      }
      __syncthreads();
    }
    double bb[1] = {0.0};
    for (int i = tid; i < n; i += WG) bb[0] += wk[i] * wk[i];
    block_sums<1>(bb, red);
    const double bk = sqrt(bb[0]);
    if (tid == 0) {
      alpha[k] = al[0];
      beta[k + 1] = bk;
    }
    if (!(bk > 1e-13 * fro) || k + 1 == kmax) {
      ++k;
      break;
    }
    for (int i = tid; i < n; i += WG) {
      v[i] = wk[i] / bk;
      Qb[i + (size_t)(k + 1) * n] = v[i];
    }
    __syncthreads();
  }
  __syncthreads();
  if (tid == 0) sh_theta = tri_max_eig(alpha, beta, k);
  __syncthreads();
  const double theta = sh_theta;
  // certificate: Cholesky of mu I - A (lower triangle, left-looking, in S2)
  double* R = nd.S2;
  const double u = 1.1102230246251565e-16;  // 2^-53
  const double gam = (n + 1) * u / (1.0 - (n + 1) * u);
  double delta = 1e-11 * fro;
  double cert = 1.01 * fro;  // lambda_max <= |A|_2 <= |A|_F
  bool done = false;
  while (!done && delta <= 2.0 * fro && fro > 0.0) {
    const double mu = theta + delta;
    __syncthreads();  // every thread has read the previous attempt's sh_fail
    if (tid == 0) sh_fail = 0;
    double rr[1] = {0.0};  // sum of squares of the factor
    for (int j = 0; j < n; ++j) {
      __syncthreads();
      double d[1] = {0.0};
      for (int t = tid; t < j; t += WG) d[0] += R[j + (size_t)t * n] * R[j + (size_t)t * n];
      block_sums<1>(d, red);
      const double piv = (mu - A[(size_t)j * (n + 1)]) - d[0];
      if (!(piv > 0.0)) {
        if (tid == 0) sh_fail = 1;
        break;  // uniform: every thread has the same piv
      }
      const double ljj = sqrt(piv);
      if (tid == 0) {
        R[(size_t)j * (n + 1)] = ljj;
        rr[0] += piv;
      }
      for (int i = j + 1 + tid; i < n; i += WG) {
        double a = -A[i + (size_t)j * n];
        for (int t = 0; t < j; ++t) a -= R[i + (size_t)t * n] * R[j + (size_t)t * n];
        const double l = a / ljj;
        R[i + (size_t)j * n] = l;
        rr[0] += l * l;
      }
    }
    __syncthreads();
    if (!sh_fail) {
      block_sums<1>(rr, red);
      cert = mu + gam * rr[0] * 1.01;
      done = true;
    } else {
      delta *= 16.0;
    }
  }
  for (int i = tid; i < n; i += WG) nd.ev[i] = cert;
  if (tid == 0) nd.off_last = 0.0;
}

size_t safe_lc_lds_bytes(int nmax) { return 2 * (size_t)nmax * sizeof(double); }

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
// of the final sum.  Output as k_safe_lc.

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

  // ---- largest eigenvalue of T: bracket [max a_i, Gershgorin], multisection
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
    if (tid == 0) sh_best = -1;
    __syncthreads();
    if (cnt > 0) atomicMax(&sh_best, tid);
    __syncthreads();
    const int t = sh_best;
    const double nlo = t >= 0 ? lo + (hi - lo) * ((double)(t + 1) / 1025.0) : lo;
    const double nhi = t < 1023 ? lo + (hi - lo) * ((double)(t + 2) / 1025.0) : hi;
    lo = nlo;
    hi = nhi;
    __syncthreads();  // sh_best is reset in the next round only after all have read it
  }
  const double theta = hi;

  // ---- certificate
  const double u = 1.1102230246251565e-16;  // 2^-53
  const double gam = (n + 1) * u / (1.0 - (n + 1) * u);
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
      const double c = mu + (1.01 * gam * rr[0] + 2.0 * u * (dmx + fabs(mu)));
      cert = nextafter(c, INFINITY);
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

void launch_safe_lc(NodeView* nodes, int nb, int nmax, hipStream_t st) {
  k_safe_lc<<<nb, 1024, safe_lc_lds_bytes(nmax), st>>>(nodes);
}

}  // namespace madam
