// madam_mlp.hip -- the learned warm start: a small MLP trained and run with
// this file's own HIP kernels (no reference counterpart; the reference
// always starts ADMM from the interior-point solution or cold).
//
// The network maps the features of one vertex pair (i, j) of a node's cost
// matrix L to a prediction of the converged ADMM primal entry X*_ij in
// [-1, 1]:  x (nin) -> tanh(x W1 + b1) (64) -> tanh(. W2 + b2) (64) ->
// tanh(. W3 + b3).  Weights are BF16 (rounded from FP32 master copies that
// the optimizer updates), biases FP32, every sum accumulates in FP32.
// Training minimises the mean of (o - t)^2 / 2 with Adam; gradients are
// taken with respect to the BF16 weights and applied to the masters.
//
// Parameter vector layout (FP32, P = nin*64 + 64 + 64*64 + 64 + 64 + 1):
// W1[f*64 + h], b1[h], W2[j*64 + k], b2[k], W3[k], b3.
//
// Kernels: k_mlp_fwd (one thread per sample: activations, output, loss),
// k_mlp_bwd (per-sample deltas), k_mlp_wgrad (one thread per parameter,
// fixed-order sums over the batch: deterministic), k_mlp_adam.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <random>
#include <vector>

#include "madam_engine.h"

namespace madam {

constexpr int MH = 64;       // hidden width
constexpr int MAX_IN = 16;   // feature count limit

struct MlpDev {
  int nin, P;
  int oW1, ob1, oW2, ob2, oW3, ob3;
};

inline MlpDev mlp_layout(int nin) {
  MlpDev L;
  L.nin = nin;
  L.oW1 = 0;
  L.ob1 = nin * MH;
  L.oW2 = L.ob1 + MH;
  L.ob2 = L.oW2 + MH * MH;
  L.oW3 = L.ob2 + MH;
  L.ob3 = L.oW3 + MH;
  L.P = L.ob3 + 1;
  return L;
}

// Round to the nearest BF16, ties to even, returned as float.
__host__ __device__ inline float bf16_round(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return x;  // inf / nan unchanged
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  float r;
  memcpy(&r, &u, 4);
  return r;
}

// Forward pass for samples idx[0..B) of feat (N x nin, row-major): stores
// h1, h2 (B x 64), o (B) and the per-sample loss (o - t)^2 / 2 when tgt.
// pw: parameters with the weights already rounded to BF16.
__global__ __launch_bounds__(256) void k_mlp_fwd(MlpDev L, const float* __restrict__ pw,
                                                 const float* __restrict__ feat,
                                                 const float* __restrict__ tgt,
                                                 const int* __restrict__ idx, int B,
                                                 float* __restrict__ h1o, float* __restrict__ h2o,
                                                 float* __restrict__ out, float* __restrict__ loss) {
  __shared__ float sw[MAX_IN * MH + MH + MH * MH + MH + MH + 1];
  for (int p = threadIdx.x; p < L.P; p += blockDim.x) sw[p] = pw[p];
  __syncthreads();
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B) return;
  const int row = idx ? idx[s] : s;
  float x[MAX_IN];
  for (int f = 0; f < L.nin; ++f) x[f] = feat[(size_t)row * L.nin + f];
  float h1[MH];
#pragma unroll
  for (int h = 0; h < MH; ++h) {
    float a = sw[L.ob1 + h];
    for (int f = 0; f < L.nin; ++f) a += x[f] * sw[L.oW1 + f * MH + h];
    h1[h] = tanhf(a);
  }
  float z = sw[L.ob3];
  for (int k = 0; k < MH; ++k) {
    float a = sw[L.ob2 + k];
#pragma unroll
    for (int j = 0; j < MH; ++j) a += h1[j] * sw[L.oW2 + j * MH + k];
    const float h2 = tanhf(a);
    if (h2o) h2o[(size_t)s * MH + k] = h2;
    z += h2 * sw[L.oW3 + k];
  }
  const float o = tanhf(z);
  out[s] = o;
  if (h1o)
    for (int h = 0; h < MH; ++h) h1o[(size_t)s * MH + h] = h1[h];
  if (tgt && loss) {
    const float e = o - tgt[row];
    loss[s] = 0.5f * e * e;
  }
}

// Per-sample deltas of the loss mean: dz3 (B), da2 (B x 64), da1 (B x 64).
__global__ __launch_bounds__(256) void k_mlp_bwd(MlpDev L, const float* __restrict__ pw,
                                                 const float* __restrict__ tgt,
                                                 const int* __restrict__ idx, int B,
                                                 const float* __restrict__ h1,
                                                 const float* __restrict__ h2,
                                                 const float* __restrict__ out,
                                                 float* __restrict__ dz3, float* __restrict__ da2,
                                                 float* __restrict__ da1) {
  __shared__ float sw[MH * MH + MH];
  for (int p = threadIdx.x; p < MH * MH; p += blockDim.x) sw[p] = pw[L.oW2 + p];
  for (int p = threadIdx.x; p < MH; p += blockDim.x) sw[MH * MH + p] = pw[L.oW3 + p];
  __syncthreads();
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B) return;
  const int row = idx ? idx[s] : s;
  const float o = out[s];
  const float d3 = (o - tgt[row]) / (float)B * (1.0f - o * o);
  dz3[s] = d3;
  float d2[MH];
#pragma unroll
  for (int k = 0; k < MH; ++k) {
    const float h = h2[(size_t)s * MH + k];
    d2[k] = d3 * sw[MH * MH + k] * (1.0f - h * h);
    da2[(size_t)s * MH + k] = d2[k];
  }
  for (int j = 0; j < MH; ++j) {
    float a = 0.f;
#pragma unroll
    for (int k = 0; k < MH; ++k) a += sw[j * MH + k] * d2[k];
    const float h = h1[(size_t)s * MH + j];
    da1[(size_t)s * MH + j] = a * (1.0f - h * h);
  }
}

// One thread per parameter: the gradient summed over the batch in sample
// order; thread P sums the loss.
__global__ __launch_bounds__(256) void k_mlp_wgrad(MlpDev L, const float* __restrict__ feat,
                                                   const int* __restrict__ idx, int B,
                                                   const float* __restrict__ h1,
                                                   const float* __restrict__ h2,
                                                   const float* __restrict__ dz3,
                                                   const float* __restrict__ da2,
                                                   const float* __restrict__ da1,
                                                   const float* __restrict__ loss,
                                                   float* __restrict__ grad,
                                                   float* __restrict__ loss_sum) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > L.P) return;
  float g = 0.f;
  if (p == L.P) {
    for (int s = 0; s < B; ++s) g += loss[s];
    *loss_sum = g / (float)B;
    return;
  }
  if (p < L.ob1) {  // W1[f][h] = sum_s x_f da1_h
    const int f = p / MH, h = p % MH;
    for (int s = 0; s < B; ++s) {
      const int row = idx ? idx[s] : s;
      g += feat[(size_t)row * L.nin + f] * da1[(size_t)s * MH + h];
    }
  } else if (p < L.oW2) {
    const int h = p - L.ob1;
    for (int s = 0; s < B; ++s) g += da1[(size_t)s * MH + h];
  } else if (p < L.ob2) {
    const int j = (p - L.oW2) / MH, k = (p - L.oW2) % MH;
    for (int s = 0; s < B; ++s) g += h1[(size_t)s * MH + j] * da2[(size_t)s * MH + k];
  } else if (p < L.oW3) {
    const int k = p - L.ob2;
    for (int s = 0; s < B; ++s) g += da2[(size_t)s * MH + k];
  } else if (p < L.ob3) {
    const int k = p - L.oW3;
    for (int s = 0; s < B; ++s) g += h2[(size_t)s * MH + k] * dz3[s];
  } else {
    for (int s = 0; s < B; ++s) g += dz3[s];
  }
  grad[p] = g;
}

// Adam on the FP32 masters; pw receives the BF16-rounded weights (biases
// unrounded) for the next forward pass.
__global__ __launch_bounds__(256) void k_mlp_adam(MlpDev L, float* __restrict__ master,
                                                  float* __restrict__ pw, float* __restrict__ m1,
                                                  float* __restrict__ m2,
                                                  const float* __restrict__ grad, float lr,
                                                  float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= L.P) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float g = grad[p];
  const float a = b1 * m1[p] + (1.0f - b1) * g;
  const float v = b2 * m2[p] + (1.0f - b2) * g * g;
  m1[p] = a;
  m2[p] = v;
  const float w = master[p] - lr * (a / c1) / (sqrtf(v / c2) + eps);
  master[p] = w;
  const bool weight = p < L.ob1 || (p >= L.oW2 && p < L.ob2) || (p >= L.oW3 && p < L.ob3);
  pw[p] = weight ? bf16_round(w) : w;
}

__global__ void k_mlp_round(MlpDev L, const float* __restrict__ master, float* __restrict__ pw) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= L.P) return;
  const bool weight = p < L.ob1 || (p >= L.oW2 && p < L.ob2) || (p >= L.oW3 && p < L.ob3);
  pw[p] = weight ? bf16_round(master[p]) : master[p];
}


// ------------------------------------------------------------ pair features
// Features of pair (i, j), i < j, of a cost matrix L (n x n, FP64,
// column-major), A = L - Diag(L), sA = sign(A), D = diag(L):
//   A_ij, sA_ij, (A^2)_ij / n, (|A|^2)_ij / n, (A^3)_ij / n^2, (sA^2)_ij / n,
//   (sA^3)_ij / n^2, (D_i + D_j) / n, |D_i - D_j| / n, [j == n-1]
// (walk counts of the signed graph and of its support, and the row-sum
// vertex flag), standardised as (f - mu) / sd.  Pairs in row order:
// (0,1), (0,2), ..., (0,n-1), (1,2), ...
constexpr int NFEAT = 10;

__global__ __launch_bounds__(256) void k_feat_sq(int n, const double* __restrict__ L,
                                                 double* __restrict__ A2, double* __restrict__ C2,
                                                 double* __restrict__ S2) {
  const int f = blockIdx.x * 256 + threadIdx.x;
  if (f >= n * n) return;
  const int j = f / n, i = f - j * n;
  double a2 = 0.0, c2 = 0.0, s2 = 0.0;
  for (int k = 0; k < n; ++k) {
    const double x = (i == k) ? 0.0 : L[i + (size_t)k * n];
    const double y = (k == j) ? 0.0 : L[k + (size_t)j * n];
    a2 += x * y;
    c2 += fabs(x) * fabs(y);
    s2 += (double)((x > 0) - (x < 0)) * (double)((y > 0) - (y < 0));
  }
  A2[f] = a2;
  C2[f] = c2;
  S2[f] = s2;
}

__global__ __launch_bounds__(256) void k_feat_pairs(int n, const double* __restrict__ L,
                                                    const double* __restrict__ A2,
                                                    const double* __restrict__ C2,
                                                    const double* __restrict__ S2,
                                                    const float* __restrict__ mu,
                                                    const float* __restrict__ sd,
                                                    float* __restrict__ feat) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  const int np = n * (n - 1) / 2;
  if (p >= np) return;
  // pair index p -> (i, j), i < j, row order
  int i = 0, rem = p;
  while (rem >= n - 1 - i) {
    rem -= n - 1 - i;
    ++i;
  }
  const int j = i + 1 + rem;
  double a3 = 0.0, s3 = 0.0;
  for (int k = 0; k < n; ++k) {
    const double x = (i == k) ? 0.0 : L[i + (size_t)k * n];
    a3 += x * A2[k + (size_t)j * n];
    s3 += (double)((x > 0) - (x < 0)) * S2[k + (size_t)j * n];
  }
  const double aij = L[i + (size_t)j * n];
  const double di = L[i + (size_t)i * n], dj = L[j + (size_t)j * n];
  const double dn = (double)n;
  const double f[NFEAT] = {aij, (double)((aij > 0) - (aij < 0)), A2[i + (size_t)j * n] / dn,
                           C2[i + (size_t)j * n] / dn, a3 / (dn * dn), S2[i + (size_t)j * n] / dn,
                           s3 / (dn * dn), (di + dj) / dn, fabs(di - dj) / dn,
                           j == n - 1 ? 1.0 : 0.0};
#pragma unroll
  for (int q = 0; q < NFEAT; ++q) {
    const float m = mu ? mu[q] : 0.f, d = sd ? sd[q] : 1.f;
    feat[(size_t)p * NFEAT + q] = (float)((f[q] - m) / d);
  }
}

}  // namespace madam

using namespace madam;
using namespace mh;

namespace {

struct MlpBufs {
  DBuf<float> master, pw, m1, m2, grad, feat, tgt, h1, h2, out, dz3, da2, da1, loss, lsum, lhist;
  DBuf<int> idx;
};

int check_net(const madam_mlp* net) {
  if (!net || !net->params || net->nin < 1 || net->nin > MAX_IN || net->nh != MH) {
    set_err("madam_mlp: nin must be in [1, %d], nh must be %d, params non-null", MAX_IN, MH);
    return MADAM_E_ARG;
  }
  return MADAM_OK;
}

int grow_batch(MlpBufs& b, int B) {
  int rc;
  if ((rc = b.h1.ensure((size_t)B * MH)) || (rc = b.h2.ensure((size_t)B * MH)) ||
      (rc = b.out.ensure(B)) || (rc = b.dz3.ensure(B)) || (rc = b.da2.ensure((size_t)B * MH)) ||
      (rc = b.da1.ensure((size_t)B * MH)) || (rc = b.loss.ensure(B)) || (rc = b.lsum.ensure(1)) ||
      (rc = b.idx.ensure(B)))
    return rc;
  return MADAM_OK;
}

// grad <- d loss / d params on samples idx[0..B) (or 0..B) of the uploaded
// set; the batch's mean loss goes to *lsum.
void launch_step(MlpBufs& b, const MlpDev& L, const int* idx, int B, float* lsum, hipStream_t st) {
  const int nb = (B + 255) / 256;
  k_mlp_fwd<<<nb, 256, 0, st>>>(L, b.pw.p, b.feat.p, b.tgt.p, idx, B, b.h1.p, b.h2.p, b.out.p,
                                b.loss.p);
  k_mlp_bwd<<<nb, 256, 0, st>>>(L, b.pw.p, b.tgt.p, idx, B, b.h1.p, b.h2.p, b.out.p, b.dz3.p,
                                b.da2.p, b.da1.p);
  k_mlp_wgrad<<<(L.P + 1 + 255) / 256, 256, 0, st>>>(L, b.feat.p, idx, B, b.h1.p, b.h2.p, b.dz3.p,
                                                     b.da2.p, b.da1.p, b.loss.p, b.grad.p, lsum);
}

MlpBufs* g_mlp = nullptr;

}  // namespace

extern "C" {

int madam_mlp_nparams(int nin) { return mlp_layout(nin).P; }

int madam_mlp_forward(const madam_mlp* net, int B, const float* feat, float* out) {
  int rc = ensure_open();
  if (rc || (rc = check_net(net))) return rc;
  if (B < 0 || (B > 0 && (!feat || !out))) {
    set_err("madam_mlp_forward: bad arguments");
    return MADAM_E_ARG;
  }
  if (B == 0) return MADAM_OK;
  if (!g_mlp) g_mlp = new MlpBufs();
  MlpBufs& b = *g_mlp;
  const MlpDev L = mlp_layout(net->nin);
  const hipStream_t st = g_eng->stream;
  if ((rc = b.master.ensure(L.P)) || (rc = b.pw.ensure(L.P)) ||
      (rc = b.feat.ensure((size_t)B * L.nin)) || (rc = b.out.ensure(B)))
    return rc;
  if ((rc = up(b.master.p, net->params, L.P, st)) || (rc = up(b.feat.p, feat, (size_t)B * L.nin, st)))
    return rc;
  k_mlp_round<<<(L.P + 255) / 256, 256, 0, st>>>(L, b.master.p, b.pw.p);
  k_mlp_fwd<<<(B + 255) / 256, 256, 0, st>>>(L, b.pw.p, b.feat.p, nullptr, nullptr, B, nullptr,
                                             nullptr, b.out.p, nullptr);
  HIP_TRY(hipGetLastError());
  if ((rc = down(out, b.out.p, B, st))) return rc;
  HIP_TRY(hipStreamSynchronize(st));
  return MADAM_OK;
}

int madam_mlp_grad(const madam_mlp* net, int B, const float* feat, const float* target,
                   float* grad, float* loss) {
  int rc = ensure_open();
  if (rc || (rc = check_net(net))) return rc;
  if (B < 1 || !feat || !target || !grad) {
    set_err("madam_mlp_grad: bad arguments");
    return MADAM_E_ARG;
  }
  if (!g_mlp) g_mlp = new MlpBufs();
  MlpBufs& b = *g_mlp;
  const MlpDev L = mlp_layout(net->nin);
  const hipStream_t st = g_eng->stream;
  if ((rc = b.master.ensure(L.P)) || (rc = b.pw.ensure(L.P)) || (rc = b.grad.ensure(L.P)) ||
      (rc = b.feat.ensure((size_t)B * L.nin)) || (rc = b.tgt.ensure(B)) || (rc = grow_batch(b, B)))
    return rc;
  if ((rc = up(b.master.p, net->params, L.P, st)) ||
      (rc = up(b.feat.p, feat, (size_t)B * L.nin, st)) || (rc = up(b.tgt.p, target, B, st)))
    return rc;
  k_mlp_round<<<(L.P + 255) / 256, 256, 0, st>>>(L, b.master.p, b.pw.p);
  launch_step(b, L, nullptr, B, b.lsum.p, st);
  HIP_TRY(hipGetLastError());
  if ((rc = down(grad, b.grad.p, L.P, st))) return rc;
  float ls = 0.f;
  if ((rc = down(&ls, b.lsum.p, 1, st))) return rc;
  HIP_TRY(hipStreamSynchronize(st));
  if (loss) *loss = ls;
  return MADAM_OK;
}

int madam_mlp_train(madam_mlp* net, long long N, const float* feat, const float* target,
                    int epochs, int batch, float lr, unsigned long long seed, float* losses) {
  int rc = ensure_open();
  if (rc || (rc = check_net(net))) return rc;
  if (N < 1 || epochs < 1 || batch < 1 || !feat || !target || lr <= 0.f) {
    set_err("madam_mlp_train: bad arguments");
    return MADAM_E_ARG;
  }
  if (!g_mlp) g_mlp = new MlpBufs();
  MlpBufs& b = *g_mlp;
  const MlpDev L = mlp_layout(net->nin);
  const hipStream_t st = g_eng->stream;
  const int B = (int)std::min<long long>(batch, N);
  const long long steps_per_epoch = N / B;
  const long long steps = steps_per_epoch * epochs;
  if ((rc = b.master.ensure(L.P)) || (rc = b.pw.ensure(L.P)) || (rc = b.m1.ensure(L.P)) ||
      (rc = b.m2.ensure(L.P)) || (rc = b.grad.ensure(L.P)) ||
      (rc = b.feat.ensure((size_t)N * L.nin)) || (rc = b.tgt.ensure(N)) ||
      (rc = grow_batch(b, B)) || (rc = b.lhist.ensure(steps)))
    return rc;
  DBuf<int> perm;
  if ((rc = perm.ensure(N))) return rc;
  if ((rc = up(b.master.p, net->params, L.P, st)) ||
      (rc = up(b.feat.p, feat, (size_t)N * L.nin, st)) || (rc = up(b.tgt.p, target, N, st)))
    return rc;
  HIP_TRY(hipMemsetAsync(b.m1.p, 0, L.P * sizeof(float), st));
  HIP_TRY(hipMemsetAsync(b.m2.p, 0, L.P * sizeof(float), st));
  k_mlp_round<<<(L.P + 255) / 256, 256, 0, st>>>(L, b.master.p, b.pw.p);
  std::vector<int> order(N);
  std::iota(order.begin(), order.end(), 0);
  std::mt19937_64 rng(seed);
  long long step = 0;
  double c1 = 1.0, c2 = 1.0;
  for (int e = 0; e < epochs; ++e) {
    std::shuffle(order.begin(), order.end(), rng);
    if ((rc = up(perm.p, order.data(), N, st))) return rc;
    for (long long q = 0; q < steps_per_epoch; ++q, ++step) {
      launch_step(b, L, perm.p + q * B, B, b.lhist.p + step, st);
      c1 *= 0.9;
      c2 *= 0.999;
      k_mlp_adam<<<(L.P + 255) / 256, 256, 0, st>>>(L, b.master.p, b.pw.p, b.m1.p, b.m2.p,
                                                     b.grad.p, lr, (float)(1.0 - c1),
                                                     (float)(1.0 - c2));
    }
    HIP_TRY(hipStreamSynchronize(st));  // the next epoch's order overwrites perm
  }
  HIP_TRY(hipGetLastError());
  if ((rc = down(net->params, b.master.p, L.P, st))) return rc;
  if (losses && (rc = down(losses, b.lhist.p, steps, st))) return rc;
  HIP_TRY(hipStreamSynchronize(st));
  perm.release();
  return MADAM_OK;
}

int madam_pair_features(int n, const double* L, const float* mu, const float* sd, float* feat) {
  int rc = ensure_open();
  if (rc) return rc;
  if (n < 2 || n > MADAM_NMAX || !L || !feat) {
    set_err("madam_pair_features: bad arguments");
    return MADAM_E_ARG;
  }
  const hipStream_t st = g_eng->stream;
  const size_t nn = (size_t)n * n;
  const int np = n * (n - 1) / 2;
  DBuf<double> Ld, A2, C2, S2;
  DBuf<float> ms, fo;
  if ((rc = Ld.ensure(nn)) || (rc = A2.ensure(nn)) || (rc = C2.ensure(nn)) ||
      (rc = S2.ensure(nn)) || (rc = ms.ensure(2 * NFEAT)) || (rc = fo.ensure((size_t)np * NFEAT)))
    return rc;
  if ((rc = up(Ld.p, L, nn, st))) return rc;
  if (mu && sd && ((rc = up(ms.p, mu, NFEAT, st)) || (rc = up(ms.p + NFEAT, sd, NFEAT, st))))
    return rc;
  k_feat_sq<<<(int)((nn + 255) / 256), 256, 0, st>>>(n, Ld.p, A2.p, C2.p, S2.p);
  k_feat_pairs<<<(np + 255) / 256, 256, 0, st>>>(n, Ld.p, A2.p, C2.p, S2.p,
                                                 (mu && sd) ? ms.p : nullptr,
                                                 (mu && sd) ? ms.p + NFEAT : nullptr, fo.p);
  HIP_TRY(hipGetLastError());
  if ((rc = down(feat, fo.p, (size_t)np * NFEAT, st))) return rc;
  HIP_TRY(hipStreamSynchronize(st));
  return MADAM_OK;
}

int madam_warm_start(int n, const double* L, const float* pred, double* X, double* y, double* Q) {
  if (n < 2 || n > MADAM_NMAX || !L || !pred || !X || !y || !Q) {
    set_err("madam_warm_start: bad arguments");
    return MADAM_E_ARG;
  }
  const size_t nn = (size_t)n * n;
  std::vector<double> Xh(nn, 0.0), Xp(nn), Qt(nn), Z(nn), W(nn);
  int p = 0;
  for (int i = 0; i < n; ++i) {
    Xh[i + (size_t)i * n] = 1.0;
    for (int j = i + 1; j < n; ++j, ++p) Xh[i + (size_t)j * n] = Xh[j + (size_t)i * n] = pred[p];
  }
  int M = 0, rc;
  if ((rc = madam_project_sdp(n, Xh.data(), Xp.data(), Qt.data(), Z.data(), &M))) return rc;
  std::vector<double> d(n);
  for (int i = 0; i < n; ++i) d[i] = 1.0 / sqrt(std::max(Xp[i + (size_t)i * n], 1e-12));
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) X[i + (size_t)j * n] = Xp[i + (size_t)j * n] * d[i] * d[j];
  // complementarity (Diag(y) - L) X = 0 on the diagonal: y_i = (L X)_ii
  for (int i = 0; i < n; ++i) {
    double a = 0.0;
    for (int k = 0; k < n; ++k) a += L[i + (size_t)k * n] * X[k + (size_t)i * n];
    y[i] = a;
  }
  for (size_t f = 0; f < nn; ++f) W[f] = L[f];
  for (int i = 0; i < n; ++i) W[i + (size_t)i * n] -= y[i];
  // Q = -(L - Diag(y))_- : the PSD part of Diag(y) - L
  if ((rc = madam_project_sdp(n, W.data(), Xp.data(), Q, Z.data(), &M))) return rc;
  return MADAM_OK;
}

}  // extern "C"
