// madam_devutil.h -- device helpers shared by the HIP kernels: workgroup
// reductions, packed-triangle indexing, round-robin pairing and an LDS-tiled
// FP64 GEMM for one workgroup.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace madam {

constexpr int WG = 1024;   // per-node kernels (16 waves)
constexpr int WGS = 256;   // streaming kernels

// Pointers known to address global memory.  Loads and stores through a
// pointer read out of a NodeView compile to flat_* instructions, whose
// completion the LDS wait counter tracks too (so a wait for an LDS read also
// drains every global load in flight) and which take 64-bit VGPR addresses;
// through a GMEM pointer they are global_* (SGPR base + VGPR offset).
#define GMEM __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ GMEM T* gmem(T* p) {
  return (GMEM T*)p;
}

// ok ? p[i] : 0, with the load issued unconditionally (from p[0] when not
// ok) and its value masked.  A load under a condition compiles to a branch
// whose block waits for the load before its value is used, so a loop of
// them pays one memory latency per iteration; this form lets a batch of
// loads be in flight together.  p must hold at least one element.
template <class P>
__device__ __forceinline__ double ld0(P p, size_t i, bool ok) {
  const double v = p[ok ? i : 0];
  return ok ? v : 0.0;
}

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

// acc (+/-) val(ent_idx(en)) over the entries en of ent[e0..e1), in entry
// order: + for en >= 0 and - otherwise (the reverse when NEG).  The entries
// of a group of 8, then their values, are loaded without branches before
// any is added (a group past the end repeats the last entry, whose value is
// then not added), so a list of cnt entries waits for about 2 ceil(cnt/8)
// memory round trips instead of 2 cnt; the sum is the same bits as the
// one-at-a-time loop.  Requires e0 < e1 or an empty loop.
template <bool NEG, class F>
__device__ __forceinline__ double signed_sum(const int* __restrict__ ent, int e0, int e1,
                                             double acc, F val) {
#pragma clang fp contract(off)
  for (int e = e0; e < e1; e += 8) {
    int en[8];
    double v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) en[q] = ent[e + q < e1 ? e + q : e1 - 1];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = val(ent_idx(en[q]));
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const double t = (en[q] >= 0) != NEG ? acc + v[q] : acc - v[q];
      acc = e + q < e1 ? t : acc;
    }
  }
  return acc;
}

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

// Right-looking Cholesky of the packed lower triangle P (order n, leading
// order n2) in place, one barrier per column: in step j wave w subtracts
// column j's outer product from trailing columns k = j+1+w, j+1+w+nw, ...
// (lanes down the column: contiguous LDS); the wave that owns column j+1
// then finishes it (square root of its diagonal, column scaled), so it is
// the pivot column of step j+1.  Returns 0, or c+1 for the first column c
// with a non-positive pivot -- the same value in every thread.  flag: two
// shared ints; a failure in step j is posted in flag[(j+1) & 1], which is
// not written again before every thread has read it.
__device__ inline int chol_packed(double* P, int n, int n2, int* flag) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
  __syncthreads();  // readers of flag in a previous call are done
  if (wv == 0) {
    if (lane == 0) flag[0] = flag[1] = 0;
    const double d = P[0];
    if (d > 0.0) {
      const double sd = sqrt(d);
      for (int i = 1 + lane; i < n; i += 64) P[i] /= sd;  // pk(i, 0) = i
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (lane == 0) P[0] = sd;
    } else if (lane == 0) {
      flag[0] = 1;
    }
  }
  __syncthreads();
  if (flag[0]) return 1;
  for (int j = 0; j + 1 < n; ++j) {
    const int base = pk(j, j, n2);
    for (int k = j + 1 + wv; k < n; k += nw) {
      const double lk = P[base + (k - j)];
      const int ck = pk(k, k, n2);
      for (int i = k + lane; i < n; i += 64) P[ck + (i - k)] -= P[base + (i - j)] * lk;
      if (k == j + 1) {
        // the wave's own LDS writes are seen by its later reads (in order)
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const double d = P[ck];
        if (d > 0.0) {
          const double sd = sqrt(d);
          for (int i = k + 1 + lane; i < n; i += 64) P[ck + (i - k)] /= sd;
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          __builtin_amdgcn_wave_barrier();
          if (lane == 0) P[ck] = sd;
        } else if (lane == 0) {
          flag[(j + 1) & 1] = k + 1;
        }
      }
    }
    __syncthreads();
    const int f = flag[(j + 1) & 1];
    if (f) return f;
  }
  return 0;
}

}  // namespace madam
