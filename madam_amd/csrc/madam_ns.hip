// madam_ns.hip -- mixed-precision PSD projection by a matrix sign
// iteration on MFMA (the chosen form of project_SDP, operators.c:521-619).
//
// W_+ = (W + W sign(W)) / 2.  sign(W) is computed from S_0 = W / |W|_F
// (spectrum inside [-1, 1]) by 13 steps S <- a S + b S^3 + c S^5 whose
// coefficients (madam_ns_coef.h) drive every |x| >= 1e-6 to 1 without ever
// squashing a large value towards 0 (which would let FP32 rounding rotate
// its eigenvector); the measured smallest |lambda| / |W|_F of ADMM's W goes
// down to ~1e-6 on the torus instances; eigenvalues below that are not
// driven to +-1, which changes W_+ by at most |lambda| / 2 in their
// eigen-directions.  Each step is three symmetric products on FP32 MFMA
// (v_mfma_f32_16x16x4f32, FP32 accumulate): T = S S, U = T (b I + c T),
// S' = S (a I + U); all iterates are polynomials in W, so every product is
// symmetric and only lower 16x16 tiles are computed and mirrored.  One FP64
// correction step S <- S (3 I - S^2) / 2 on FP64 MFMA
// (v_mfma_f64_16x16x4f64) restores S^2 = I to FP64 precision, and the
// projection itself, W_+ = (W + W S) / 2, is an FP64 product.
//
// Two launch layouts with identical arithmetic: one workgroup per node
// (k_ns_project; the FP32 iterates, 3 n_p^2 floats with n_p = n rounded up
// to 16 and zero padded, live in LDS when n_p <= 112, else in the node's
// ns32 scratch), or one launch per product spread over many workgroups
// (k_ns_gemm32 and friends, below) for batches that cannot fill the CUs.
// FP64 matrices live in the node's S1/S2.  Output: X_+ in S1 (full), M = round(trace((I+S)/2)).
// Guard: an FP32 iterate with an entry above 1.5 in magnitude or not finite
// (a sign matrix has all entries in [-1, 1]) sets nd.ns_fail and leaves S1
// for the FP64 Jacobi projection that the host launches behind this kernel
// (k_jacobi with form 3 runs only on such nodes).
#include "madam_devutil.h"
#include "madam_internal.h"
#include "madam_ns_coef.h"

#include <algorithm>

namespace madam {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

__constant__ float NS_COEF[MADAM_NS_STEPS][3] = MADAM_NS_COEF_INIT;
constexpr int NS_STEPS = MADAM_NS_STEPS;

// C = A (alpha I + beta B) for symmetric n_p x n_p A, B (column-major, ld
// n_p, n_p % 16 == 0) whose product is symmetric; FP32 MFMA, lower tiles.
// Lane layout of v_mfma_f32_16x16x4f32: A operand lane l -> (row l%16,
// k l/16); B operand lane l -> (k l/16, col l%16); result lane l -> rows
// 4(l/16)+v, col l%16.
__device__ inline void tile_decode(int t, int& i0, int& j0) {
  int ti = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
  if ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
  if (ti * (ti + 1) / 2 > t) --ti;
  const int tj = t - ti * (ti + 1) / 2;
  i0 = ti << 4;
  j0 = tj << 4;
}

// Leading dimension of the one-workgroup kernel's LDS operand buffer
// (mode 2): n_p, or n_p + 16 where n_p is 0 or 32 (mod 64).
__host__ __device__ inline int ns_lds_ld(int np) { return (np & 63) == 0 || (np & 63) == 32 ? np + 16 : np; }

// Tile products split K into four contiguous quarters, each accumulated
// by its own MFMA chain and summed as ((q0 + q1) + q2) + q3: one wave can
// run the four chains interleaved (independent accumulators keep the MFMA
// pipeline full), or four waves can run one quarter each (split-K, for
// latency when few tiles are in flight) -- with the same result bits.
//
// Quarter qi of a tile of C = A (alpha I + beta B): k in [qi kq, (qi+1) kq),
// kq = n_p / 4 (a multiple of 4).  Lane layout of v_mfma_f32_16x16x4f32:
// A operand lane l -> (row l%16, k l/16); B operand lane l -> (k l/16,
// col l%16); result lane l -> rows 4(l/16)+v, col l%16.
// NS consecutive k-steps of one quarter: all 2 NS operand loads are issued
// before the first MFMA (one memory latency per NS steps instead of one per
// step; the accumulation order is unchanged).
// P: const float* or const GMEM float* (global loads)
template <int NS, class P>
__device__ __forceinline__ void ksteps_f32(P A, P B, int np, float alpha, float beta, int ia, int jb,
                                           int k, f32x4& acc) {
  float a[NS], b[NS];
#pragma unroll
  for (int t = 0; t < NS; ++t) {
    const int kk = k + 4 * t;
    a[t] = A[ia + (size_t)kk * np];
    // B symmetric: B[k][j] = B[j + k np]
    float bb = beta * B[jb + (size_t)kk * np];
    if (jb == kk) bb += alpha;
    b[t] = bb;
  }
#pragma unroll
  for (int t = 0; t < NS; ++t) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[t], b[t], acc, 0, 0, 0);
}

// (the k steps of a quarter in batches of up to 32, every load of a batch
// in flight together)
template <class P>
__device__ __forceinline__ f32x4 quarter_f32(P A, P B, int np, float alpha, float beta, int i0,
                                             int j0, int qi) {
  const int lane = threadIdx.x & 63;
  const int r = lane & 15, q = lane >> 4;
  const int kq = np >> 2;
  const int ia = i0 + r, jb = j0 + r;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  int k = qi * kq + q, rem = kq >> 2;
  for (; rem >= 32; rem -= 32, k += 128) ksteps_f32<32>(A, B, np, alpha, beta, ia, jb, k, acc);
  if (rem >= 16) {
    ksteps_f32<16>(A, B, np, alpha, beta, ia, jb, k, acc);
    rem -= 16;
    k += 64;
  }
  switch (rem) {
#define MADAM_KS(N) \
  case N:           \
    ksteps_f32<N>(A, B, np, alpha, beta, ia, jb, k, acc); \
    break;
    MADAM_KS(1) MADAM_KS(2) MADAM_KS(3) MADAM_KS(4) MADAM_KS(5) MADAM_KS(6) MADAM_KS(7)
    MADAM_KS(8) MADAM_KS(9) MADAM_KS(10) MADAM_KS(11) MADAM_KS(12) MADAM_KS(13) MADAM_KS(14)
    MADAM_KS(15)
#undef MADAM_KS
    default:
      break;
  }
  return acc;
}

// only the lower element of each (i,j)/(j,i) pair is stored, to both
// places: the iterates stay exactly symmetric, which the transposed operand
// reads rely on (an asymmetric part would grow step by step)
template <class P>
__device__ __forceinline__ void store_tile_f32(P Cm, int np, int i0, int j0, f32x4 acc) {
  const int lane = threadIdx.x & 63;
  const int r = lane & 15, q = lane >> 4;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int i = i0 + 4 * q + v, j = j0 + r;
    if (i >= j) {
      Cm[i + (size_t)j * np] = acc[v];
      Cm[j + (size_t)i * np] = acc[v];
    }
  }
}

// One lower 16x16 tile t, all four quarters by the calling wave.
__device__ __forceinline__ void gemm_tile_f32(const float* __restrict__ A, const float* __restrict__ B,
                                              float* __restrict__ Cm, int np, float alpha,
                                              float beta, int t) {
  const int lane = threadIdx.x & 63;
  const int r = lane & 15, q = lane >> 4;
  int i0, j0;
  tile_decode(t, i0, j0);
  const int kq = np >> 2;
  f32x4 acc[4] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f},
                  {0.f, 0.f, 0.f, 0.f}};
  for (int s = 0; s < kq; s += 4) {
    float a[4], b[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = u * kq + s + q;
      a[u] = A[(i0 + r) + (size_t)k * np];
      b[u] = beta * B[(j0 + r) + (size_t)k * np];
      if (j0 + r == k) b[u] += alpha;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      acc[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u], b[u], acc[u], 0, 0, 0);
  }
  store_tile_f32(Cm, np, i0, j0, ((acc[0] + acc[1]) + acc[2]) + acc[3]);
}

__device__ __forceinline__ void sym_gemm_f32(const float* __restrict__ A, const float* __restrict__ B,
                             float* __restrict__ Cm, int np, float alpha, float beta) {
  const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int nt = np >> 4;
  const int ntile = nt * (nt + 1) / 2;
  for (int t = wave; t < ntile; t += nw) gemm_tile_f32(A, B, Cm, np, alpha, beta, t);
}

// The same product over 2 x 2 blocks of tiles: the calling wave computes
// the (up to four) lower tiles of block t of the ceil(n_p / 32)-square block
// grid, all four K quarters each, so that every operand value loaded feeds
// two MFMAs.  Each tile's quarters run the k steps of gemm_tile_f32 in the
// same order and are summed as ((q0 + q1) + q2) + q3: the same bits.  Used
// where an operand comes from L2/HBM (256 nodes of order 150: 4.4 against
// 4.6 us per node and projection; order 170: 5.9 against 6.0); with all
// iterates in LDS single tiles are faster (order 100: 1.3 against 1.6).
template <class PA, class PB, class PC>
__device__ __forceinline__ void gemm_block2_f32(PA A, PB B, PC Cm, int np, float alpha, float beta,
                                                int t) {
  const int lane = threadIdx.x & 63;
  const int r = lane & 15, q = lane >> 4;
  const int nt = np >> 4;
  int bi, bj;
  tile_decode(t, bi, bj);
  bi >>= 4;
  bj >>= 4;
  bool vr[2], vc[2];
  int ra[2], cb[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    vr[u] = 2 * bi + u < nt;
    vc[u] = 2 * bj + u < nt;
    ra[u] = (vr[u] ? (2 * bi + u) * 16 : 0) + r;  // a missing tile row reads row 0, unused
    cb[u] = (vc[u] ? (2 * bj + u) * 16 : 0) + r;
  }
  bool on[4];
#pragma unroll
  for (int x = 0; x < 4; ++x)
    on[x] = vr[x >> 1] && vc[x & 1] && 2 * bj + (x & 1) <= 2 * bi + (x >> 1);
  const int kq = np >> 2;
  f32x4 acc[4][4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int x = 0; x < 4; ++x) acc[u][x] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int s = 0; s < kq; s += 4) {
    float a[4][2], b[4][2];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = u * kq + s + q;
#pragma unroll
      for (int v = 0; v < 2; ++v) {
        a[u][v] = A[ra[v] + (size_t)k * np];
        float bb = beta * B[cb[v] + (size_t)k * np];
        if (cb[v] == k) bb += alpha;
        b[u][v] = bb;
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int x = 0; x < 4; ++x)
        if (on[x])
          acc[u][x] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u][x >> 1], b[u][x & 1], acc[u][x], 0,
                                                           0, 0);
  }
#pragma unroll
  for (int x = 0; x < 4; ++x)
    if (on[x])
      store_tile_f32(Cm, np, (2 * bi + (x >> 1)) * 16, (2 * bj + (x & 1)) * 16,
                     ((acc[0][x] + acc[1][x]) + acc[2][x]) + acc[3][x]);
}

// The same block with its results kept in registers (the k loop of
// gemm_block2_f32, same bits), stored later by block2_store.
struct Blk2 {
  f32x4 c[4];
  bool on[4];
  int bi, bj;
};

template <class PA, class PB>
__device__ __forceinline__ void block2_compute(PA A, int lda, PB B, int ldb, int np, float alpha,
                                               float beta, int t, Blk2& out) {
  const int lane = threadIdx.x & 63;
  const int r = lane & 15, q = lane >> 4;
  const int nt = np >> 4;
  int bi, bj;
  tile_decode(t, bi, bj);
  bi >>= 4;
  bj >>= 4;
  bool vr[2], vc[2];
  int ra[2], cb[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    vr[u] = 2 * bi + u < nt;
    vc[u] = 2 * bj + u < nt;
    ra[u] = (vr[u] ? (2 * bi + u) * 16 : 0) + r;  // a missing tile row reads row 0, unused
    cb[u] = (vc[u] ? (2 * bj + u) * 16 : 0) + r;
  }
  bool on[4];
#pragma unroll
  for (int x = 0; x < 4; ++x)
    on[x] = vr[x >> 1] && vc[x & 1] && 2 * bj + (x & 1) <= 2 * bi + (x >> 1);
  const int kq = np >> 2;
  f32x4 acc[4][4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int x = 0; x < 4; ++x) acc[u][x] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int s = 0; s < kq; s += 4) {
    float a[4][2], b[4][2];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = u * kq + s + q;
#pragma unroll
      for (int v = 0; v < 2; ++v) {
        a[u][v] = A[ra[v] + (size_t)k * lda];
        float bb = beta * B[cb[v] + (size_t)k * ldb];
        if (cb[v] == k) bb += alpha;
        b[u][v] = bb;
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int x = 0; x < 4; ++x)
        if (on[x])
          acc[u][x] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u][x >> 1], b[u][x & 1], acc[u][x], 0,
                                                           0, 0);
  }
#pragma unroll
  for (int x = 0; x < 4; ++x) {
    out.on[x] = on[x];
    out.c[x] = ((acc[0][x] + acc[1][x]) + acc[2][x]) + acc[3][x];
  }
  out.bi = bi;
  out.bj = bj;
}


template <class PC>
__device__ __forceinline__ void block2_store(PC Cm, int ldc, const Blk2& b) {
#pragma unroll
  for (int x = 0; x < 4; ++x)
    if (b.on[x]) store_tile_f32(Cm, ldc, (2 * b.bi + (x >> 1)) * 16, (2 * b.bj + (x & 1)) * 16, b.c[x]);
}

template <class PA, class PB, class PC>
__device__ __forceinline__ void sym_gemm_f32_b2(PA A, PB B, PC Cm, int np, float alpha,
                                                float beta) {
  const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int ng = ((np >> 4) + 1) >> 1;
  const int nblk = ng * (ng + 1) / 2;
  for (int t = wave; t < nblk; t += nw) gemm_block2_f32(A, B, Cm, np, alpha, beta, t);
}

// Same in FP64 with v_mfma_f64_16x16x4f64, on n x n matrices of any n
// (k and tiles past n are masked), quarters of kq = 4 ceil(n / 16).  The
// result register v of lane l holds row (l/16) + 4v, column l%16
// (measured on gfx950).
template <int NS, class P>
__device__ __forceinline__ void ksteps_f64(P A, P B, int n, double alpha, double beta, int ia,
                                           int jb, int k, f64x4& acc) {
  // every load from an in-range address (row, k clamped to n - 1), its value
  // masked afterwards: no load waits behind a branch
  double a[NS], b[NS];
  const int iac = ia < n ? ia : n - 1, jbc = jb < n ? jb : n - 1;
#pragma unroll
  for (int t = 0; t < NS; ++t) {
    const int kk = k + 4 * t, kc = kk < n ? kk : n - 1;
    a[t] = A[iac + (size_t)kc * n];
    b[t] = B[jbc + (size_t)kc * n];
  }
#pragma unroll
  for (int t = 0; t < NS; ++t) {
    const int kk = k + 4 * t;
    a[t] = (ia < n && kk < n) ? a[t] : 0.0;
    double bb = (jb < n && kk < n) ? beta * b[t] : 0.0;
    if (jb == kk && kk < n) bb += alpha;
    b[t] = bb;
  }
#pragma unroll
  for (int t = 0; t < NS; ++t) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], b[t], acc, 0, 0, 0);
}

template <class P>
__device__ __forceinline__ f64x4 quarter_f64(P A, P B, int n, double alpha, double beta, int i0,
                                             int j0, int qi) {
  const int lane = threadIdx.x & 63;
  const int r = lane & 15, q = lane >> 4;
  const int kq = ((n + 15) >> 4) << 2;
  const int ia = i0 + r, jb = j0 + r;
  f64x4 acc = {0.0, 0.0, 0.0, 0.0};
  int k = qi * kq + q, rem = kq >> 2;
  for (; rem >= 16; rem -= 16, k += 64) ksteps_f64<16>(A, B, n, alpha, beta, ia, jb, k, acc);
  switch (rem) {
#define MADAM_KS(N) \
  case N:           \
    ksteps_f64<N>(A, B, n, alpha, beta, ia, jb, k, acc); \
    break;
    MADAM_KS(1) MADAM_KS(2) MADAM_KS(3) MADAM_KS(4) MADAM_KS(5) MADAM_KS(6) MADAM_KS(7)
    MADAM_KS(8) MADAM_KS(9) MADAM_KS(10) MADAM_KS(11) MADAM_KS(12) MADAM_KS(13) MADAM_KS(14)
    MADAM_KS(15)
#undef MADAM_KS
    default:
      break;
  }
  return acc;
}

template <class P>
__device__ __forceinline__ void store_tile_f64(P Cm, int n, int i0, int j0, f64x4 acc) {
  const int lane = threadIdx.x & 63;
  const int r = lane & 15, q = lane >> 4;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int i = i0 + q + 4 * v, j = j0 + r;
    if (i < n && j < n && i >= j) {
      Cm[i + (size_t)j * n] = acc[v];
      Cm[j + (size_t)i * n] = acc[v];
    }
  }
}

__device__ __forceinline__ void gemm_tile_f64(const double* __restrict__ A, const double* __restrict__ B,
                                              double* __restrict__ Cm, int n, double alpha,
                                              double beta, int t) {
  const int lane = threadIdx.x & 63;
  const int r = lane & 15, q = lane >> 4;
  int i0, j0;
  tile_decode(t, i0, j0);
  const int kq = ((n + 15) >> 4) << 2;
  const int ia = i0 + r, jb = j0 + r;
  f64x4 acc[4] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0},
                  {0.0, 0.0, 0.0, 0.0}};
  for (int s = 0; s < kq; s += 4) {
    double a[4], b[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = u * kq + s + q;
      a[u] = (ia < n && k < n) ? A[ia + (size_t)k * n] : 0.0;
      b[u] = (jb < n && k < n) ? beta * B[jb + (size_t)k * n] : 0.0;
      if (jb == k && k < n) b[u] += alpha;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[u], acc[u], 0, 0, 0);
  }
  store_tile_f64(Cm, n, i0, j0, ((acc[0] + acc[1]) + acc[2]) + acc[3]);
}

__device__ __forceinline__ void sym_gemm_f64(const double* __restrict__ A, const double* __restrict__ B,
                             double* __restrict__ Cm, int n, double alpha, double beta) {
  const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int nt = (n + 15) >> 4;
  const int ntile = nt * (nt + 1) / 2;
  for (int t = wave; t < ntile; t += nw) gemm_tile_f64(A, B, Cm, n, alpha, beta, t);
}

// Projection of nd.W (FP64) -> X_+ in nd.S1, one workgroup per node.
// MODE 1: the three FP32 iterates live in LDS (3 n_p^2 floats, n_p <=
// 112); MODE 2: one n_p^2 LDS buffer (n_p <= 192) holding the matrix both
// operands of the next product read -- S for T = S^2, T for U, U for S' --
// and S also in the node's ns32 scratch, from where S' = S (a I + U) reads
// it.  A product's results stay in registers until every wave has
// finished reading the buffer, then overwrite it (S' also goes to ns32):
// no round trip through L2/HBM for T and U.  MODE 0: all three in ns32.
// The products and their bits are the same in every mode.
template <int MODE>
__global__ __launch_bounds__(1024) void k_ns_project(NodeView* nodes, int ignore_done) {
  extern __shared__ float sm32[];
  __shared__ double red[64];
  NodeView& nd = nodes[blockIdx.x];
  if (!ignore_done && nd.done) return;
  const int n = nd.n, np = (n + 15) & ~15, tid = threadIdx.x;
  const size_t npp = (size_t)np * np;
  float* Sb = MODE == 1 ? sm32 : nd.ns32;
  float* Tb = Sb + npp;
  float* Ub = Tb + npp;
  float* Lb = sm32;  // MODE 2, 3: the LDS operand buffer
  // MODE 2: its leading dimension, = 16 or 48 (mod 64) so that the four
  // k rows an MFMA operand load spans fall in different banks
  const int ldl = MODE == 2 ? ns_lds_ld(np) : np;
  const double* __restrict__ Wm = nd.W;
  double fv[1] = {0.0};
  for (int f = tid; f < n * n; f += WG) fv[0] += Wm[f] * Wm[f];
  block_sums<1>(fv, red);
  const double fro = sqrt(fv[0]);
  const float inv = fro > 0.0 ? (float)(1.0 / fro) : 0.f;
  for (size_t f = tid; f < npp; f += WG) {
    const int j = (int)(f / np), i = (int)(f - (size_t)j * np);
    const float v = (i < n && j < n) ? (float)Wm[i + (size_t)j * n] * inv : 0.f;
    Sb[f] = v;
    if (MODE >= 2) Lb[i + (size_t)j * ldl] = v;
  }
  __syncthreads();
  float* S = Sb;
  float* T = Tb;
  float* U = Ub;
  // MODE 2: C = A (alpha I + beta B) into the LDS buffer (and into G when
  // given); each wave holds one 2 x 2 block (n_p <= 160: at most 15
  // blocks for the 16 waves)
  auto product_lds = [&](const float* A, float alpha, float beta, float* G) {
    const int wave = tid >> 6;
    const int ng = ((np >> 4) + 1) >> 1;
    const bool mine = wave < ng * (ng + 1) / 2;
    Blk2 res;
    if (mine) block2_compute(A, A == Lb ? ldl : np, Lb, ldl, np, alpha, beta, wave, res);
    __syncthreads();  // every wave has read the buffer
    if (mine) {
      block2_store(Lb, ldl, res);
      if (G) block2_store(G, np, res);
    }
    __syncthreads();
  };
  // MODE 3: the LDS buffer refilled from ns32 by a coalesced copy after
  // each product (n_p 176..192, where a wave would hold two blocks)
  auto stage = [&](const float* G) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    for (size_t f = (size_t)tid * 4; f < npp; f += (size_t)WG * 4)
      *(f4*)(Lb + f) = *(const f4*)(G + f);
    __syncthreads();
  };
  for (int s = 0; s < NS_STEPS; ++s) {
    const float a = NS_COEF[s][0], b = NS_COEF[s][1], c = NS_COEF[s][2];
    if (MODE == 2) {
      product_lds(Lb, 0.f, 1.f, nullptr);  // T = S^2
      product_lds(Lb, b, c, nullptr);      // U = T (b I + c T)
      product_lds(Sb, a, 1.f, Sb);         // S' = S (a I + U), S from ns32
      continue;
    } else if (MODE == 3) {
      sym_gemm_f32_b2(Lb, Lb, T, np, 0.f, 1.f);  // T = S^2 (S staged)
      __syncthreads();
      stage(T);
      sym_gemm_f32_b2(Lb, Lb, U, np, b, c);      // U = T (b I + c T)
      __syncthreads();
      stage(U);
      sym_gemm_f32_b2(S, Lb, T, np, a, 1.f);     // S' = S (a I + U), into T
      __syncthreads();
      stage(T);
    } else {
      // (all in LDS, MODE 1: single tiles, measured faster there; MODE 0:
      // 2 x 2 blocks)
      if (MODE == 1) {
        sym_gemm_f32(S, S, T, np, 0.f, 1.f);  // T = S^2
        __syncthreads();
        sym_gemm_f32(T, T, U, np, b, c);      // U = T (b I + c T)
        __syncthreads();
        sym_gemm_f32(S, U, T, np, a, 1.f);    // S' = S (a I + U), into T
      } else {
        sym_gemm_f32_b2(S, S, T, np, 0.f, 1.f);
        __syncthreads();
        sym_gemm_f32_b2(T, T, U, np, b, c);
        __syncthreads();
        sym_gemm_f32_b2(S, U, T, np, a, 1.f);
      }
      __syncthreads();
    }
    float* tmp = S;
    S = T;
    T = tmp;
  }
  // FP64 correction S <- S (3 I - S^2) / 2 and W_+ = (W + W S) / 2
  double* S64 = nd.S2;
  double* T64 = nd.S1;
  __shared__ int bad;
  if (tid == 0) bad = 0;
  __syncthreads();
  for (int f = tid; f < n * n; f += WG) {
    const int j = f / n, i = f - j * n;
    const float v = S[i + (size_t)j * np];
    if (!(fabsf(v) <= 1.5f)) bad = 1;
    S64[f] = (double)v;
  }
  __syncthreads();
  if (bad) {
    if (tid == 0) nd.ns_fail = 1;
    return;
  }
  if (tid == 0) nd.ns_fail = 0;
  sym_gemm_f64(S64, S64, T64, n, 0.0, 1.0);  // T = S^2
  __syncthreads();
  sym_gemm_f64(S64, T64, nd.ip1, n, 1.5, -0.5);  // S' = S (1.5 I - 0.5 S^2)
  __syncthreads();
  double tr[1] = {0.0};
  for (int i = tid; i < n; i += WG) tr[0] += nd.ip1[(size_t)i * (n + 1)];
  block_sums<1>(tr, red);
  // X_+ = (W + W S) / 2 = W (0.5 I + 0.5 S)
  sym_gemm_f64(Wm, nd.ip1, nd.S1, n, 0.5, 0.5);
  if (tid == 0) nd.M = (int)lrint(0.5 * (n + tr[0]));
}

size_t ns_lds_bytes(int nmax) {
  const size_t np = (size_t)((nmax + 15) & ~15);
  return 3 * np * np * sizeof(float);
}

void launch_ns_project(NodeView* nodes, int nb, int nmax, int ignore_done, size_t lds_limit,
                       hipStream_t st) {
  const size_t lds = ns_lds_bytes(nmax);
  if (lds + 1024 <= lds_limit)
    k_ns_project<1><<<nb, 1024, lds, st>>>(nodes, ignore_done);
  else if (lds / 3 + 1024 <= lds_limit && ((nmax + 15) & ~15) <= 160) {
    // (every node's buffer, n_p x ns_lds_ld(n_p), fits the largest's)
    const size_t np = (size_t)((nmax + 15) & ~15);
    const size_t l2 = np * (np + 16) * sizeof(float);
    k_ns_project<2><<<nb, 1024, l2, st>>>(nodes, ignore_done);
  }
  else if (lds / 3 + 1024 <= lds_limit)
    k_ns_project<3><<<nb, 1024, lds / 3, st>>>(nodes, ignore_done);
  else
    k_ns_project<0><<<nb, 1024, 0, st>>>(nodes, ignore_done);
}

// ---------------------------------------------------------------------------
// Tile-parallel form of the same projection: one launch per product, one
// wave per lower 16x16 tile, grid (tile groups, nodes), so that a node's
// product is spread over many CUs.  Used when the batch has fewer nodes
// than the chip has CUs (a single node, or the stragglers of a batch); the
// arithmetic is the one of k_ns_project, tile for tile.

// S_0 = W / |W|_F into buffer 0 of ns32 (zero padded), clear ns_fail, in
// two launches spread over many workgroups with the arithmetic of
// k_ns_project: |W|_F^2 is summed as there by 1024 virtual threads (thread
// v adds W[f]^2 for f = v, v + 1024, ... in order), each 64 of them
// combined by wave_sum -- here one 64-thread workgroup per virtual wave,
// its sum stored after the three iterates in ns32 -- and the 16 wave sums
// added in order (block_sums).  Same bits as the one-workgroup layout.
constexpr int NS_NORM_WAVES = WG / 64;

__device__ __forceinline__ double* ns_norm_parts(const NodeView& nd) {
  const size_t np = (size_t)((nd.n + 15) & ~15);
  return (double*)(nd.ns32 + 3 * np * np);
}

// Workgroup-to-work maps of the tile-parallel kernels.  Batches of 8 or
// more nodes: grid (8 x parts, ceil(nodes / 8)), workgroup (x, y) serves
// node 8 y + x % 8 and part x / 8; workgroups go to the 8 XCDs round-robin
// in dispatch order, so every part of a node runs on one XCD and the node's
// operands are fetched into one L2.  Fewer nodes (`spread`): grid (parts,
// nodes), so that a node's parts go round-robin to all 8 XCDs instead of
// leaving seven of them idle.
__device__ __forceinline__ int ns_node(int nb, int spread, int& part, int& nparts) {
  if (spread) {
    part = blockIdx.x;
    nparts = gridDim.x;
    return blockIdx.y;
  }
  part = blockIdx.x >> 3;
  nparts = gridDim.x >> 3;
  const int b = blockIdx.y * 8 + (blockIdx.x & 7);
  return b < nb ? b : -1;
}

__global__ __launch_bounds__(64) void k_ns_norm(NodeView* nodes) {
  NodeView& nd = nodes[blockIdx.y];
  if (nd.done) return;
  const int n = nd.n;
  const GMEM double* __restrict__ Wm = gmem(nd.W);
  const int nn = n * n;
  double fv = 0.0;
  int f = blockIdx.x * 64 + threadIdx.x;
  // the terms of each virtual thread in order, their loads issued 128 at a
  // time (a latency per 128 terms instead of per term)
  constexpr int NB = 128;
  for (; f + (NB - 1) * WG < nn; f += NB * WG) {
    double v[NB];
#pragma unroll
    for (int t = 0; t < NB; ++t) v[t] = Wm[f + t * WG];
#pragma unroll
    for (int t = 0; t < NB; ++t) fv += v[t] * v[t];
  }
  for (; f < nn; f += WG) fv += Wm[f] * Wm[f];
  fv = wave_sum(fv);
  if (threadIdx.x == 0) ns_norm_parts(nd)[blockIdx.x] = fv;
}

__global__ __launch_bounds__(256) void k_ns_fill(NodeView* nodes, int nb, int spread) {
  int part, nparts;
  const int b = ns_node(nb, spread, part, nparts);
  if (b < 0) return;
  NodeView& nd = nodes[b];
  if (nd.done) return;
  const int n = nd.n, np = (n + 15) & ~15;
  const size_t npp = (size_t)np * np;
  const double* parts = ns_norm_parts(nd);
  double r = 0.0;
  for (int i = 0; i < NS_NORM_WAVES; ++i) r += parts[i];
  const double fro = sqrt(r);
  const float inv = fro > 0.0 ? (float)(1.0 / fro) : 0.f;
  const GMEM double* __restrict__ Wm = gmem(nd.W);
  GMEM float* Sb = gmem(nd.ns32);
  for (size_t f = (size_t)part * 256 + threadIdx.x; f < npp; f += (size_t)nparts * 256) {
    const int j = (int)(f / np), i = (int)(f - (size_t)j * np);
    Sb[f] = (i < n && j < n) ? (float)Wm[i + (size_t)j * n] * inv : 0.f;
  }
  if (part == 0 && threadIdx.x == 0) nd.ns_fail = 0;
}

// C = A (alpha I + beta B) on buffers ia, ib, ic of ns32: one workgroup of
// four waves per tile, wave w accumulating K quarter w (split K: a quarter
// of the latency of one wave doing the whole tile), the partial tiles summed
// in LDS in the quarter order of gemm_tile_f32.
__global__ __launch_bounds__(256) void k_ns_gemm32(NodeView* nodes, int nb, int spread, int ia,
                                                   int ib, int ic, float alpha, float beta) {
  __shared__ f32x4 part[3][64];
  int t, nt_unused;
  const int b = ns_node(nb, spread, t, nt_unused);
  if (b < 0) return;
  NodeView& nd = nodes[b];
  if (nd.done) return;
  const int np = (nd.n + 15) & ~15, nt = np >> 4, ntile = nt * (nt + 1) / 2;
  if (t >= ntile) return;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int i0, j0;
  tile_decode(t, i0, j0);
  const size_t npp = (size_t)np * np;
  GMEM float* base = gmem(nd.ns32);
  const f32x4 acc = quarter_f32((const GMEM float*)(base + ia * npp),
                                (const GMEM float*)(base + ib * npp), np, alpha, beta, i0, j0, w);
  if (w > 0) part[w - 1][lane] = acc;
  __syncthreads();
  if (w == 0)
    store_tile_f32(base + ic * npp, np, i0, j0,
                   ((acc + part[0][lane]) + part[1][lane]) + part[2][lane]);
}

// The same product on 2 x 2 blocks of tiles: workgroup t covers lower
// block t of the ceil(n_p / 32)-square block grid with two waves, wave w
// accumulating K quarters 2w and 2w + 1 of all four tiles, so that every
// operand value loaded feeds two MFMAs (half the loads per flop of
// k_ns_gemm32).  Each tile's quarters run the same k steps in the same
// order and are summed as ((q0 + q1) + q2) + q3: the same bits.  Faster
// for batches of large orders (n = 512, 32 nodes: 78 against 114 us per
// product), slower for one node (fewer, longer waves).
__global__ __launch_bounds__(128) void k_ns_gemm32b(NodeView* nodes, int nb, int spread, int ia,
                                                    int ib, int ic, float alpha, float beta) {
  constexpr int NSB = 8;
  __shared__ f32x4 red[2][4][64];
  int t, nt_unused;
  const int b = ns_node(nb, spread, t, nt_unused);
  if (b < 0) return;
  NodeView& nd = nodes[b];
  if (nd.done) return;
  const int np = (nd.n + 15) & ~15, nt = np >> 4, ng = (nt + 1) >> 1;
  if (t >= ng * (ng + 1) / 2) return;
  int bi, bj;
  tile_decode(t, bi, bj);
  bi >>= 4;
  bj >>= 4;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = lane & 15, q = lane >> 4;
  const size_t npp = (size_t)np * np;
  const GMEM float* __restrict__ A = gmem(nd.ns32) + ia * npp;
  const GMEM float* __restrict__ B = gmem(nd.ns32) + ib * npp;
  bool vr[2], vc[2];
  int ra[2], cb[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    vr[u] = 2 * bi + u < nt;
    vc[u] = 2 * bj + u < nt;
    ra[u] = (vr[u] ? (2 * bi + u) * 16 : 0) + r;  // a missing tile row reads row 0, unused
    cb[u] = (vc[u] ? (2 * bj + u) * 16 : 0) + r;
  }
  bool on[4];
#pragma unroll
  for (int x = 0; x < 4; ++x)
    on[x] = vr[x >> 1] && vc[x & 1] && 2 * bj + (x & 1) <= 2 * bi + (x >> 1);
  // quarter c = k steps [c nt, (c + 1) nt) of 4 k each (lane group q holds k = 4s + q)
  f32x4 acc[2][4];
#pragma unroll
  for (int e = 0; e < 2; ++e)
#pragma unroll
    for (int x = 0; x < 4; ++x) acc[e][x] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int s0 = 0; s0 < nt; s0 += NSB) {
    float a[NSB][2][2], bq[NSB][2][2];
#pragma unroll
    for (int s = 0; s < NSB; ++s)
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const int st = s0 + s < nt ? s0 + s : 0;  // past the quarter: in bounds, unused
        const int k = 4 * ((2 * w + e) * nt + st) + q;
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          a[s][e][u] = A[ra[u] + (size_t)k * np];
          float bb = beta * B[cb[u] + (size_t)k * np];
          if (cb[u] == k) bb += alpha;
          bq[s][e][u] = bb;
        }
      }
#pragma unroll
    for (int s = 0; s < NSB; ++s) {
      if (s0 + s >= nt) break;
#pragma unroll
      for (int e = 0; e < 2; ++e)
#pragma unroll
        for (int x = 0; x < 4; ++x)
          if (on[x])
            acc[e][x] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s][e][x >> 1], bq[s][e][x & 1],
                                                             acc[e][x], 0, 0, 0);
    }
  }
  if (w == 1) {
#pragma unroll
    for (int e = 0; e < 2; ++e)
#pragma unroll
      for (int x = 0; x < 4; ++x) red[e][x][lane] = acc[e][x];
  }
  __syncthreads();
  if (w != 0) return;
  GMEM float* Cm = gmem(nd.ns32) + ic * npp;
#pragma unroll
  for (int x = 0; x < 4; ++x)
    if (on[x])
      store_tile_f32(Cm, np, (2 * bi + (x >> 1)) * 16, (2 * bj + (x & 1)) * 16,
                     ((acc[0][x] + acc[1][x]) + red[0][x][lane]) + red[1][x][lane]);
}

// FP32 iterate (buffer is) -> FP64 n x n in S2, with the guard.
__global__ __launch_bounds__(256) void k_ns_conv(NodeView* nodes, int nb, int spread, int is) {
  int part, nparts;
  const int b = ns_node(nb, spread, part, nparts);
  if (b < 0) return;
  NodeView& nd = nodes[b];
  if (nd.done) return;
  const int n = nd.n, np = (n + 15) & ~15;
  const GMEM float* S = gmem(nd.ns32) + (size_t)is * np * np;
  GMEM double* S2 = gmem(nd.S2);
  int bad = 0;
  for (int f = part * 256 + threadIdx.x; f < n * n; f += nparts * 256) {
    const int j = f / n, i = f - j * n;
    const float v = S[i + (size_t)j * np];
    if (!(fabsf(v) <= 1.5f)) bad = 1;
    S2[f] = (double)v;
  }
  if (bad) nd.ns_fail = 1;  // every writer stores the same value
}

// The FP64 finish, one product per launch (split K as above): 0: S1 = S2^2;
// 1: ip1 = S2 (1.5 I - 0.5 S1); 2: S1 = W (0.5 I + 0.5 ip1) = X_+.
__global__ __launch_bounds__(256) void k_ns_gemm64(NodeView* nodes, int nb, int spread,
                                                   int which) {
  __shared__ f64x4 part[3][64];
  int t, nt_unused;
  const int b = ns_node(nb, spread, t, nt_unused);
  if (b < 0) return;
  NodeView& nd = nodes[b];
  if (nd.done || nd.ns_fail) return;
  const int n = nd.n, nt = (n + 15) >> 4, ntile = nt * (nt + 1) / 2;
  if (t >= ntile) return;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int i0, j0;
  tile_decode(t, i0, j0);
  const GMEM double* A = gmem(which == 2 ? nd.W : nd.S2);
  const GMEM double* B = gmem(which == 0 ? nd.S2 : (which == 1 ? nd.S1 : nd.ip1));
  GMEM double* Cm = gmem(which == 1 ? nd.ip1 : nd.S1);
  const double alpha = which == 0 ? 0.0 : (which == 1 ? 1.5 : 0.5);
  const double beta = which == 0 ? 1.0 : (which == 1 ? -0.5 : 0.5);
  const f64x4 acc = quarter_f64(A, B, n, alpha, beta, i0, j0, w);
  if (w > 0) part[w - 1][lane] = acc;
  __syncthreads();
  if (w == 0) store_tile_f64(Cm, n, i0, j0, ((acc + part[0][lane]) + part[1][lane]) + part[2][lane]);
}

void launch_ns_tiles(NodeView* nodes, int nb, int nmax, hipStream_t st, int blocks2) {
  static const float coef[MADAM_NS_STEPS][3] = MADAM_NS_COEF_INIT;
  const int np = (nmax + 15) & ~15, nt = np >> 4, ntile = nt * (nt + 1) / 2;
  const int ng = (nt + 1) >> 1, nblk = ng * (ng + 1) / 2;
  // spread a node over all XCDs only when its tiles outnumber one XCD's 32
  // CUs (n = 512: 528 tiles); below that one L2 serving every tile is
  // faster (n = 100, 28 tiles: 0.17 ms per projection spread, 0.135 local)
  const int spread = (nb < 8 && ntile > 48) ? 1 : 0;
  const int ny = (nb + 7) / 8;
  auto grid = [&](int parts) { return spread ? dim3(parts, nb) : dim3(8 * parts, ny); };
  const dim3 g32 = grid(blocks2 ? nblk : ntile);
  const int wg32 = blocks2 ? 128 : 256;
  auto gemm = [&](int ia, int ib, int ic, float alpha, float beta) {
    if (blocks2)
      k_ns_gemm32b<<<g32, wg32, 0, st>>>(nodes, nb, spread, ia, ib, ic, alpha, beta);
    else
      k_ns_gemm32<<<g32, wg32, 0, st>>>(nodes, nb, spread, ia, ib, ic, alpha, beta);
  };
  k_ns_norm<<<dim3(NS_NORM_WAVES, nb), 64, 0, st>>>(nodes);
  const int nfill = std::min(256, (np * np + 255) / 256);
  k_ns_fill<<<grid(nfill), 256, 0, st>>>(nodes, nb, spread);
  int sb = 0, tb = 1, ub = 2;
  for (int k = 0; k < MADAM_NS_STEPS; ++k) {
    gemm(sb, sb, tb, 0.f, 1.f);                   // T = S^2
    gemm(tb, tb, ub, coef[k][1], coef[k][2]);     // U = T (b I + c T)
    gemm(sb, ub, tb, coef[k][0], 1.f);            // S' = S (a I + U)
    const int x = sb;
    sb = tb;
    tb = x;
  }
  const int nconv = std::min(256, (nmax * nmax + 255) / 256);
  k_ns_conv<<<grid(nconv), 256, 0, st>>>(nodes, nb, spread, sb);
  const int nt64 = (nmax + 15) >> 4, ntile64 = nt64 * (nt64 + 1) / 2;
  for (int w = 0; w < 3; ++w) k_ns_gemm64<<<grid(ntile64), 256, 0, st>>>(nodes, nb, spread, w);
}

// A synthetic symmetric W for timing the projection layouts: entry (i, j)
// a hash of the unordered pair and the node, uniform in [-1, 1).
__global__ __launch_bounds__(256) void k_ns_synth(NodeView* nodes) {
  NodeView& nd = nodes[blockIdx.y];
  const int n = nd.n;
  const size_t nn = (size_t)n * n;
  for (size_t f = blockIdx.x * 256 + threadIdx.x; f < nn; f += (size_t)gridDim.x * 256) {
    const int j = (int)(f / n), i = (int)(f - (size_t)j * n);
    const unsigned long long lo = i < j ? i : j, hi = i < j ? j : i;
    unsigned long long h = lo * 0x9E3779B97F4A7C15ull + hi * 0xC2B2AE3D27D4EB4Full + blockIdx.y;
    h ^= h >> 30;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 27;
    h *= 0x94D049BB133111EBull;
    h ^= h >> 31;
    nd.W[f] = (double)(h >> 11) * 0x1.0p-52 - 1.0;
  }
}

void launch_ns_synth(NodeView* nodes, int nb, int nmax, hipStream_t st) {
  k_ns_synth<<<dim3(std::min(256, (nmax * nmax + 255) / 256), nb), 256, 0, st>>>(nodes);
}

// ---------------------------------------------------------------------------
// The rounding heuristic's factor in the sign-iteration form, which has no
// eigenvectors: a pivoted Cholesky X = Z Z^T of the node's primal X (PSD),
// Z stored as V's first r columns with evp = 1 / sigma_last there and 0
// elsewhere, so that the heuristic's V diag(sqrt(evp sigma_last)) is Z and
// M = r.  Pivots: largest remaining diagonal, first index among equals;
// stop when it falls below 1e-10 of the largest diagonal of X.  Any Z with
// Z Z^T = X serves the heuristic, which rounds sign(Z v).
__global__ __launch_bounds__(1024) void k_zfac_chol(NodeView* nodes) {
  __shared__ double rv[16];
  __shared__ int ri[16];
  __shared__ double red[64];
  NodeView& nd = nodes[blockIdx.x];
  const int n = nd.n, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  double* __restrict__ Z = nd.V;
  double* __restrict__ d = nd.ev;  // remaining diagonal; -1 marks a pivoted row
  const double* __restrict__ X = nd.X;
  double mx = 0.0;
  for (int i = tid; i < n; i += WG) {
    d[i] = X[(size_t)i * (n + 1)];
    nd.evp[i] = 0.0;
    mx = fmax(mx, d[i]);
  }
  const double dmax0 = block_max(mx, red);
  const double tol = 1e-10 * dmax0;
  int r = 0;
  for (int q = 0; q < n && dmax0 > 0.0; ++q) {
    __syncthreads();
    double bv = -2.0;
    int bi = n;
    for (int i = tid; i < n; i += WG)
      if (d[i] > bv) {
        bv = d[i];
        bi = i;
      }
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_xor(bv, off, 64);
      const int oi = __shfl_xor(bi, off, 64);
      if (ov > bv || (ov == bv && oi < bi)) {
        bv = ov;
        bi = oi;
      }
    }
    if (lane == 0) {
      rv[w] = bv;
      ri[w] = bi;
    }
    __syncthreads();
    bv = rv[0];
    bi = ri[0];
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k)
      if (rv[k] > bv || (rv[k] == bv && ri[k] < bi)) {
        bv = rv[k];
        bi = ri[k];
      }
    if (!(bv > tol)) break;  // uniform: every thread holds the same bv
    const int p = bi;
    const double sp = sqrt(bv);
    for (int i = tid; i < n; i += WG) {
      double z = 0.0;
      if (i == p) {
        z = sp;
      } else if (d[i] >= 0.0) {
        double a = X[i + (size_t)p * n];
        for (int t = 0; t < q; ++t) a -= Z[i + (size_t)t * n] * Z[p + (size_t)t * n];
        z = a / sp;
      }
      Z[i + (size_t)q * n] = z;
    }
    __syncthreads();
    for (int i = tid; i < n; i += WG) {
      if (i == p)
        d[i] = -1.0;
      else if (d[i] >= 0.0)
        d[i] = fmax(0.0, d[i] - Z[i + (size_t)q * n] * Z[i + (size_t)q * n]);
    }
    ++r;
  }
  __syncthreads();
  const double e = nd.sigma_last > 0.0 ? 1.0 / nd.sigma_last : 1.0;
  for (int q = tid; q < r; q += WG) nd.evp[q] = e;
  if (tid == 0) nd.M = r;
}

void launch_zfac_chol(NodeView* nodes, int nb, hipStream_t st) {
  k_zfac_chol<<<nb, 1024, 0, st>>>(nodes);
}

// Self-test of the MFMA operand/result layouts: C32 = A (alpha I + beta B)
// in FP32 MFMA (np x np), C64 the same in FP64 MFMA (n x n).
__global__ __launch_bounds__(1024) void k_mfma_test(const float* A32, const float* B32, float* C32,
                                                   int np, const double* A64, const double* B64,
                                                   double* C64, int n, double alpha, double beta) {
  sym_gemm_f32(A32, B32, C32, np, (float)alpha, (float)beta);
  sym_gemm_f64(A64, B64, C64, n, alpha, beta);
}

void launch_mfma_test(const float* A32, const float* B32, float* C32, int np, const double* A64,
                      const double* B64, double* C64, int n, double alpha, double beta,
                      hipStream_t st) {
  k_mfma_test<<<1, 1024, 0, st>>>(A32, B32, C32, np, A64, B64, C64, n, alpha, beta);
}

}  // namespace madam
