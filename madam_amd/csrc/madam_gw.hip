// madam_gw.hip -- Goemans-Williamson hyperplane rounding with 1-opt
// (heuristic.c:24-201) for a batch of nodes.
//
// Per node: Zc = V_+ diag(sqrt(lambda_+ sigma)) (the Z of admm.c:583-585),
// `num` random hyperplanes v (components uniform in [-1/2, 1/2) from a
// splitmix64 stream), z = sign(Zc v) (>= 0 -> +1), 1-opt on the subproblem
// matrix exactly as mc_1opt does it (same incremental updates, first index
// among equal gains), then the value of the 0/1 cut that heuristic.c:76-84
// builds from z -- the homogenising coordinate is dropped there, i.e. the
// cut is z' = (z_0..z_{n-2}, +1) -- computed as z'^T L z' from the 1-opt's
// own L z in O(n).  One wave per hyperplane; the best hyperplane per node
// (first among equal values) is copied out.
#include "madam_devutil.h"
#include "madam_internal.h"
#include "madam_gw.h"

namespace madam {

__device__ __forceinline__ uint64_t gw_sm64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Zc (n x M, compacted positive eigenvector columns, scaled) into S1.
__global__ __launch_bounds__(1024) void k_gw_prep(NodeView* nodes, GwNode* gw) {
  __shared__ int cols[1024];
  __shared__ int cnt;
  NodeView& nd = nodes[blockIdx.x];
  GwNode& g = gw[blockIdx.x];
  const int n = nd.n;
  if (threadIdx.x == 0) {
    int m = 0;
    if (!g.zero_planes)
      for (int j = 0; j < n; ++j)
        if (nd.evp[j] > 0.0) cols[m++] = j;
    cnt = m;
    g.M = m;
  }
  __syncthreads();
  const int M = cnt;
  const double sg = nd.sigma_last;
  for (int f = threadIdx.x; f < n * M; f += blockDim.x) {
    const int q = f / n, i = f - q * n;
    const int j = cols[q];
    nd.S1[i + (size_t)q * n] = nd.V[i + (size_t)j * n] * sqrt(nd.evp[j] * sg);
  }
}

constexpr int GW_WAVES = 8;

__global__ __launch_bounds__(64 * GW_WAVES) void k_gw(NodeView* nodes, GwNode* gw) {
#pragma clang fp contract(off)
  extern __shared__ double gsm[];
  NodeView& nd = nodes[blockIdx.y];
  GwNode& g = gw[blockIdx.y];
  const int n = nd.n, M = g.M, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int h = blockIdx.x * GW_WAVES + w;
  double* dg = gsm;                       // n
  double* vh = dg + n + (size_t)w * 2 * n;  // n (M used)
  double* Lx = vh + n;                    // n
  signed char* xs = (signed char*)(gsm + n + (size_t)GW_WAVES * 2 * n) + (size_t)w * n;
  const double* __restrict__ L = nd.L;
  const double* __restrict__ Zc = nd.S1;
  for (int i = threadIdx.x; i < n; i += 64 * GW_WAVES) dg[i] = L[i + (size_t)i * n];
  __syncthreads();
  if (h >= g.num) return;
  // hyperplane
  const uint64_t key = gw_sm64(g.seed ^ gw_sm64(((uint64_t)g.stream << 24) + (uint64_t)h));
  for (int j = lane; j < M; j += 64)
    vh[j] = (double)(gw_sm64(key + (uint64_t)j) >> 11) * (1.0 / 9007199254740992.0) - 0.5;
  __builtin_amdgcn_wave_barrier();
  for (int i = lane; i < n; i += 64) {
    double sca = 0.0;
    for (int j = 0; j < M; ++j) sca += vh[j] * Zc[i + (size_t)j * n];
    xs[i] = sca < 0 ? -1 : 1;
  }
  __builtin_amdgcn_wave_barrier();
  // 1-opt (heuristic.c:100-201)
  for (int i = lane; i < n; i += 64) {
    double acc = 0.0;
    for (int j = 0; j < n; ++j) acc += L[j + (size_t)i * n] * xs[j];
    Lx[i] = acc;
  }
  __builtin_amdgcn_wave_barrier();
  for (int guard = 0; guard < 64 * n + 64; ++guard) {
    double best = -1e9;
    int at = 0x7fffffff;
    for (int i = lane; i < n; i += 64) {
      const double gain = dg[i] - xs[i] * Lx[i];
      if (gain > best) {
        best = gain;
        at = i;
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double ob = __shfl_xor(best, off, 64);
      const int oa = __shfl_xor(at, off, 64);
      if (ob > best || (ob == best && oa < at)) {
        best = ob;
        at = oa;
      }
    }
    if (!(best > 0.001)) break;
    const double sgn = xs[at] > 0 ? -2.0 : 2.0;
    __builtin_amdgcn_wave_barrier();
    for (int j = lane; j < n; j += 64) {
      const double lv = L[j + (size_t)at * n];
      if (fabs(lv) > 0.001) Lx[j] += sgn * lv;
    }
    if (lane == (at & 63)) xs[at] = -xs[at];
    __builtin_amdgcn_wave_barrier();
  }
  // value of the cut with the homogenising coordinate set to +1
  const bool fl = xs[n - 1] < 0;
  double part = 0.0;
  for (int i = lane; i < n; i += 64) {
    const double zi = (i == n - 1) ? 1.0 : (double)xs[i];
    const double lz = Lx[i] + (fl ? 2.0 * L[i + (size_t)(n - 1) * n] : 0.0);
    part += zi * lz;
  }
  const double val = wave_sum(part);
  signed char* zout = g.zall + (size_t)h * n;
  for (int i = lane; i < n; i += 64) zout[i] = xs[i];
  if (lane == 0) g.val[h] = val + g.fixedvalue;
}

// One workgroup per node: best hyperplane (first among equal values).
__global__ __launch_bounds__(256) void k_gw_best(NodeView* nodes, GwNode* gw) {
  NodeView& nd = nodes[blockIdx.x];
  GwNode& g = gw[blockIdx.x];
  __shared__ int bh;
  if (threadIdx.x == 0) {
    int b = 0;
    for (int h = 1; h < g.num; ++h)
      if (g.val[h] > g.val[b]) b = h;
    bh = b;
    g.best_val[0] = g.val[b];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nd.n; i += 256) g.best_z[i] = g.zall[(size_t)bh * nd.n + i];
}

size_t gw_lds_bytes(int nmax) {
  return (size_t)nmax * sizeof(double) * (1 + 2 * GW_WAVES) + (size_t)GW_WAVES * nmax + 64;
}

void launch_gw_given_z(NodeView* nodes, GwNode* gw, int nb, int nmax, int nummax,
                       hipStream_t st) {
  dim3 grid((nummax + GW_WAVES - 1) / GW_WAVES, nb);
  k_gw<<<grid, 64 * GW_WAVES, gw_lds_bytes(nmax), st>>>(nodes, gw);
  k_gw_best<<<nb, 256, 0, st>>>(nodes, gw);
}

void launch_gw(NodeView* nodes, GwNode* gw, int nb, int nmax, int nummax, hipStream_t st) {
  k_gw_prep<<<nb, 1024, 0, st>>>(nodes, gw);
  dim3 grid((nummax + GW_WAVES - 1) / GW_WAVES, nb);
  k_gw<<<grid, 64 * GW_WAVES, gw_lds_bytes(nmax), st>>>(nodes, gw);
  k_gw_best<<<nb, 256, 0, st>>>(nodes, gw);
}

}  // namespace madam
