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
// certified value (what k_safe_bound takes), off_last = 0.  k_safe_hh: one
// workgroup of 1024 threads per node with A as a packed triangle in LDS (n
// up to about 195); larger orders: the same steps spread over many
// workgroups per node (k_sft_*, at the end of this file).
#include <stdint.h>

#include "madam_devutil.h"
#include "madam_engine.h"
#include "madam_internal.h"
#include "madam_tiles.h"

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
// of the final sum.  Output: ev[0..n) = the certified value, off_last = 0.

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

// ---------------------------------------------------------------------------
// The same certificate for orders beyond LDS, spread over many workgroups
// per node and driven by the host over the batch.  Tridiagonalisation
// (Golub and Van Loan, Alg. 8.3.1) on the full copy G of A in S2, two
// launches per reflection j: k_sft_symv -- every workgroup forms the
// reflection of column j from the same data in the same order (the same
// bits in all of them) and computes its rows of p = tau A v, one wave per
// row (A symmetric: row r read as column r, contiguous); k_sft_update --
// every workgroup forms kk = (tau/2) v^T p and applies A <- A - v w^T -
// w v^T, w = p - kk v, to its columns of the trailing block, both
// triangles.  theta = lambda_max(T) by k_sft_top (tri_top_eig, one
// workgroup per node).  Certificate: the Cholesky factorisation of
// mu I - A, mu = theta + delta, by the blocked many-workgroup
// factorisation (madam_tiles.h) in S1, retried with delta x 16 while it
// fails, the bound as cert_value (the sum of squared factor entries over
// the lower triangle of the factor).  Per-node vectors (v, p, and T's
// diagonal and subdiagonal) in a batch scratch of 4 nmax doubles a node.
struct SafeCtl {
  double fro, theta, delta, tau, cert;
  int state;  // 0: factorisation pending, 1: retry, 2: certified (or the |A|_F bound)
  int pad;
};
constexpr int SFT_PARTS = 64;

__global__ __launch_bounds__(256) void k_sft_init(NodeView* nodes, double* part) {
  __shared__ double red[4];
  const NodeView& nd = nodes[blockIdx.y];
  const int n = nd.n;
  const size_t nn = (size_t)n * n, ch = (nn + SFT_PARTS - 1) / SFT_PARTS;
  const size_t f0 = blockIdx.x * ch, f1 = f0 + ch < nn ? f0 + ch : nn;
  double s = 0.0;
  for (size_t f = f0 + threadIdx.x; f < f1; f += 256) {
    const double a = nd.W[f];
    nd.S2[f] = a;
    s += a * a;
  }
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0)
    part[(size_t)blockIdx.y * SFT_PARTS + blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

__global__ void k_sft_fro(SafeCtl* sc, const double* part, int nb) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  double s = 0.0;
  for (int x = 0; x < SFT_PARTS; ++x) s += part[(size_t)b * SFT_PARTS + x];
  SafeCtl& c = sc[b];
  c.fro = sqrt(s);
  c.cert = 1.01 * c.fro;  // lambda_max <= |A|_2 <= |A|_F
  c.delta = 1e-11 * c.fro;
  c.tau = 0.0;
  c.state = c.fro > 0.0 ? 0 : 2;
}

__global__ __launch_bounds__(256) void k_sft_symv(NodeView* nodes, SafeCtl* sc, double* vec,
                                                  int vstride, int j) {
  __shared__ double vv[MADAM_NMAX];
  const int b = blockIdx.y;
  const NodeView& nd = nodes[b];
  const int n = nd.n;
  if (j + 2 >= n || sc[b].state == 2) return;
  const int m = n - 1 - j, o = j + 1, r0 = 16 * blockIdx.x;
  if (r0 >= m) return;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const double* __restrict__ G = nd.S2;
  for (int i = tid; i < m; i += 256) vv[i] = G[o + i + (size_t)j * n];
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
  double* vg = vec + (size_t)b * vstride;
  double* pg = vg + vstride / 4;
  double* ta = pg + vstride / 4;
  double* tb = ta + vstride / 4;
  if (blockIdx.x == 0) {
    for (int i = tid; i < m; i += 256) vg[i] = i == 0 ? v0 : vv[i];
    if (tid == 0) {
      tb[j + 1] = alpha;
      ta[j] = G[j + (size_t)j * n];
      sc[b].tau = tau;
    }
  }
  if (tau == 0.0) return;  // the same in every thread
  for (int q = 0; q < 4; ++q) {
    const int r = r0 + 4 * wv + q;
    if (r >= m) break;
    const double* col = G + o + (size_t)(o + r) * n;
    double acc = 0.0;
    for (int k = lane; k < m; k += 64) acc += col[k] * (k == 0 ? v0 : vv[k]);
    acc = wave_sum(acc);
    if (lane == 0) pg[r] = tau * acc;
  }
}

__global__ __launch_bounds__(256) void k_sft_update(NodeView* nodes, const SafeCtl* sc,
                                                    const double* vec, int vstride, int j) {
  __shared__ double vs[MADAM_NMAX];
  __shared__ double ps[MADAM_NMAX];
  const int b = blockIdx.y;
  const NodeView& nd = nodes[b];
  const int n = nd.n;
  if (j + 2 >= n || sc[b].state == 2) return;
  const double tau = sc[b].tau;
  if (tau == 0.0) return;
  const int m = n - 1 - j, o = j + 1, c0 = 8 * blockIdx.x;
  if (c0 >= m) return;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const double* vg = vec + (size_t)b * vstride;
  const double* pg = vg + vstride / 4;
  for (int i = tid; i < m; i += 256) {
    vs[i] = vg[i];
    ps[i] = pg[i];
  }
  __syncthreads();
  double kk = 0.0;  // (tau / 2) v^T p, the same in every wave
  for (int i = lane; i < m; i += 64) kk += vs[i] * ps[i];
  kk = 0.5 * tau * wave_sum(kk);
  double* G = nd.S2;
  for (int q = 0; q < 2; ++q) {
    const int c = c0 + 2 * wv + q;
    if (c >= m) break;
    const double vc = vs[c];
    const double wc = ps[c] - kk * vc;
    double* col = G + o + (size_t)(o + c) * n;
    for (int r = lane; r < m; r += 64) {
      const double vr = vs[r];
      col[r] -= vr * wc + (ps[r] - kk * vr) * vc;
    }
  }
}

__global__ __launch_bounds__(1024) void k_sft_top(NodeView* nodes, SafeCtl* sc, double* vec,
                                                  int vstride) {
  __shared__ double ta[MADAM_NMAX];
  __shared__ double tb[MADAM_NMAX];
  __shared__ int sh_best;
  const int b = blockIdx.x;
  if (sc[b].state == 2) return;
  const NodeView& nd = nodes[b];
  const int n = nd.n;
  const double* G = nd.S2;
  const double* tag = vec + (size_t)b * vstride + 2 * (vstride / 4);
  const double* tbg = tag + vstride / 4;
  for (int i = threadIdx.x; i < n; i += 1024) {
    ta[i] = i + 2 < n ? tag[i] : G[i + (size_t)i * n];
    tb[i] = i == 0 ? 0.0 : (i + 1 < n ? tbg[i] : G[i + (size_t)(i - 1) * n]);
  }
  __syncthreads();
  const double theta = tri_top_eig(ta, tb, n, &sh_best);
  if (threadIdx.x == 0) sc[b].theta = theta;
}

// R (lower triangle of S1) = mu I - A for the nodes whose factorisation is
// pending; the job is reset (skipped for certified nodes).
__global__ __launch_bounds__(256) void k_sft_load(NodeView* nodes, const SafeCtl* sc,
                                                  CholJob* jobs) {
  const int b = blockIdx.y;
  const bool idle = sc[b].state == 2;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    jobs[b].fail = 0;
    jobs[b].skip = idle ? 1 : 0;
  }
  if (idle) return;
  const NodeView& nd = nodes[b];
  const int n = nd.n;
  const double mu = sc[b].theta + sc[b].delta;
  const size_t nn = (size_t)n * n;
  for (size_t f = blockIdx.x * 256 + threadIdx.x; f < nn; f += (size_t)gridDim.x * 256) {
    const int j = (int)(f / n), i = (int)(f - (size_t)j * n);
    if (i >= j) nd.S1[f] = i == j ? mu - nd.W[f] : -nd.W[f];
  }
}

// A completed factorisation gives the certified bound; a failed one
// multiplies delta by 16 (while delta <= 2 |A|_F, else the |A|_F bound).
__global__ __launch_bounds__(256) void k_sft_check(NodeView* nodes, SafeCtl* sc,
                                                   const CholJob* jobs) {
  __shared__ double red[8];
  const int b = blockIdx.x;
  SafeCtl& c = sc[b];
  if (c.state == 2) return;
  if (jobs[b].fail) {
    if (threadIdx.x == 0) {
      c.delta *= 16.0;
      c.state = c.delta <= 2.0 * c.fro ? 1 : 2;
    }
    return;
  }
  const NodeView& nd = nodes[b];
  const int n = nd.n;
  const double mu = c.theta + c.delta;
  double rr = 0.0, dmx = 0.0;
  for (int j = 0; j < n; ++j) {  // the factor's lower triangle, column by column
    const double* col = nd.S1 + (size_t)j * n;
    for (int i = j + threadIdx.x; i < n; i += 256) rr += col[i] * col[i];
  }
  for (int i = threadIdx.x; i < n; i += 256) dmx = fmax(dmx, fabs(mu - nd.W[(size_t)i * (n + 1)]));
  rr = wave_sum(rr);
  dmx = wave_max(dmx);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    red[wv] = rr;
    red[4 + wv] = dmx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double r = ((red[0] + red[1]) + red[2]) + red[3];
    const double d = fmax(fmax(red[4], red[5]), fmax(red[6], red[7]));
    c.cert = cert_value(mu, r, d, n);
    c.state = 2;
  }
}

__global__ __launch_bounds__(256) void k_sft_out(NodeView* nodes, const SafeCtl* sc) {
  NodeView& nd = nodes[blockIdx.x];
  const double cert = sc[blockIdx.x].cert;
  for (int i = threadIdx.x; i < nd.n; i += 256) nd.ev[i] = cert;
  if (threadIdx.x == 0) nd.off_last = 0.0;
}

}  // namespace madam

namespace mh {

using madam::CholJob;
using madam::SafeCtl;

static DBuf<SafeCtl> g_sctl;
static DBuf<CholJob> g_sjobs;
static DBuf<double> g_svec, g_spart, g_sD;
static SafeCtl* g_hsctl = nullptr;
static int g_hsctl_cap = 0;

void release_safe_tiles() {
  g_sctl.release();
  g_sjobs.release();
  g_svec.release();
  g_spart.release();
  g_sD.release();
  if (g_hsctl) (void)hipHostFree(g_hsctl);
  g_hsctl = nullptr;
  g_hsctl_cap = 0;
}

int safe_tiles(NodeView* d_nodes, const NodeView* hv, int nb, int nmax, hipStream_t st) {
  const int vstride = 4 * nmax;
  int rc;
  if ((rc = g_sctl.ensure(nb)) || (rc = g_sjobs.ensure(nb)) ||
      (rc = g_svec.ensure((size_t)nb * vstride)) ||
      (rc = g_spart.ensure((size_t)nb * madam::SFT_PARTS)) ||
      (rc = g_sD.ensure((size_t)nb * madam::IB * madam::IB)))
    return rc;
  if (g_hsctl_cap < nb) {
    if (g_hsctl) (void)hipHostFree(g_hsctl);
    g_hsctl = nullptr;
    g_hsctl_cap = 0;
    HIP_TRY(hipHostMalloc((void**)&g_hsctl, (size_t)nb * sizeof(SafeCtl), hipHostMallocDefault));
    g_hsctl_cap = nb;
  }
  std::vector<CholJob> J(nb);
  for (int b = 0; b < nb; ++b)
    J[b] = {hv[b].S1, g_sD.p + (size_t)b * madam::IB * madam::IB, nullptr, hv[b].n, 0, 1, 0};
  if ((rc = up(g_sjobs.p, J.data(), (size_t)nb, st))) return rc;
  SafeCtl* sc = g_sctl.p;
  const int nbt = (nb + 63) / 64;
  madam::k_sft_init<<<dim3(madam::SFT_PARTS, nb), 256, 0, st>>>(d_nodes, g_spart.p);
  madam::k_sft_fro<<<nbt, 64, 0, st>>>(sc, g_spart.p, nb);
  for (int j = 0; j + 2 < nmax; ++j) {
    const int m = nmax - 1 - j;
    madam::k_sft_symv<<<dim3((m + 15) / 16, nb), 256, 0, st>>>(d_nodes, sc, g_svec.p, vstride, j);
    madam::k_sft_update<<<dim3((m + 7) / 8, nb), 256, 0, st>>>(d_nodes, sc, g_svec.p, vstride, j);
  }
  madam::k_sft_top<<<nb, 1024, 0, st>>>(d_nodes, sc, g_svec.p, vstride);
  const int nload = std::min(256, (nmax * nmax + 255) / 256);
  for (int attempt = 0;; ++attempt) {
    madam::k_sft_load<<<dim3(nload, nb), 256, 0, st>>>(d_nodes, sc, g_sjobs.p);
    madam::chol_launches(g_sjobs.p, nb, nmax, st);
    madam::k_sft_check<<<nb, 256, 0, st>>>(d_nodes, sc, g_sjobs.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(g_hsctl, sc, (size_t)nb * sizeof(SafeCtl), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    bool open = false;
    for (int b = 0; b < nb; ++b) open = open || g_hsctl[b].state != 2;
    if (!open) break;
    if (attempt > 40) {
      set_err("safe bound certificate did not finish");
      return MADAM_E_HIP;
    }
  }
  madam::k_sft_out<<<nb, 256, 0, st>>>(d_nodes, sc);
  HIP_TRY(hipGetLastError());
  return MADAM_OK;
}

}  // namespace mh
