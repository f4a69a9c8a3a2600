// madam_sep.hip -- cutting-plane separation and primal heuristic kernels
// (cutting_planes.c:20-127, qap_simulated_annealing.c:12-206,
// heuristic.c:24-201) for gfx950.
//
// Triangle separation (getViolated_TriangleInequalities) enumerates the
// 4 * C(n,3) rows B_c(X) - 1 of every node in parallel: one workgroup per
// (node, i) handles all j < i, k < j with row i of X staged in LDS and the
// (j,k) entries read as contiguous columns of the symmetric X.  The top-K
// selection is exact: pass 0 counts violated rows (> thresh) and takes the
// maximum, pass 1 histograms the violated values over [thresh, max], a scan
// finds the bin holding the K-th largest value, pass 2 emits every row above
// that bin and collects the boundary bin, and a last kernel ranks the
// boundary rows by (value desc, enumeration order asc), keeps what is
// needed and sorts the list into enumeration order.  Keys pack
// (type-1, i, j, k) as 2+10+10+10 bits, so key order is the reference's
// enumeration order (type outermost, then i, j, k ascending).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "madam_sep.h"

namespace madam {

constexpr int SEP_WG = 256;

__device__ __forceinline__ unsigned long long ord_bits(double v) {
  unsigned long long b = __double_as_longlong(v);
  return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double from_ord(unsigned long long o) {
  unsigned long long b = (o & 0x8000000000000000ull) ? (o & 0x7fffffffffffffffull) : ~o;
  return __longlong_as_double(b);
}

// the four triangle rows minus one, in the reference's operation order
// (cutting_planes.c:24-36)
__device__ __forceinline__ double tri_val(int type, double a, double b, double c) {
#pragma clang fp contract(off)
  switch (type) {
    case 1: return -a - b - c - 1.0;
    case 2: return -a + b + c - 1.0;
    case 3: return a - b + c - 1.0;
    default: return a + b - c - 1.0;
  }
}

__device__ __forceinline__ uint32_t tri_key(int type, int i, int j, int k) {
  return ((uint32_t)(type - 1) << 30) | ((uint32_t)i << 20) | ((uint32_t)j << 10) | (uint32_t)k;
}

__device__ __forceinline__ void pair_decode(int p, int& j, int& k) {
  // p = j(j-1)/2 + k, 0 <= k < j
  int jj = (int)((1.0f + sqrtf(1.0f + 8.0f * (float)p)) * 0.5f);
  if (jj * (jj - 1) / 2 > p) --jj;
  if ((jj + 1) * jj / 2 <= p) ++jj;
  j = jj;
  k = p - jj * (jj - 1) / 2;
}

// PASS 0: count + max; PASS 1: histogram; PASS 2: emit.
template <int PASS>
__global__ __launch_bounds__(SEP_WG) void k_tri_pass(SepView* sv) {
  __shared__ double xi[1024];
  __shared__ uint32_t hist[SEP_NBIN];
  __shared__ unsigned long long smax;
  __shared__ uint32_t scount;
  SepView& s = sv[blockIdx.y];
  const int n = s.n, i = blockIdx.x;
  if (i >= n || i < 2) return;
  if (PASS > 0 && s.ctl[CTL_COUNT] <= (uint32_t)s.K) {
    if (PASS == 1) return;
  }
  const double* __restrict__ X = s.X;
  const int tid = threadIdx.x;
  for (int j = tid; j < i; j += SEP_WG) xi[j] = X[i + (size_t)j * n];
  if (PASS == 1)
    for (int b = tid; b < SEP_NBIN; b += SEP_WG) hist[b] = 0;
  if (tid == 0) {
    smax = 0;
    scount = 0;
  }
  __syncthreads();
  const double thr = s.thresh;
  double lo = 0.0, scale = 0.0;
  int emit_all = 0, bstar = 0;
  if (PASS >= 1) {
    const double vmax = from_ord(s.maxbits[0]);
    lo = thr;
    scale = (vmax > thr) ? (double)SEP_NBIN / (vmax - thr) : 0.0;
    emit_all = s.ctl[CTL_COUNT] <= (uint32_t)s.K;
    bstar = (int)s.ctl[CTL_BSTAR];
  }
  double vmax_loc = -1e300;
  uint32_t cnt_loc = 0;
  const int npairs = i * (i - 1) / 2;
  for (int p = tid; p < npairs; p += SEP_WG) {
    int j, k;
    pair_decode(p, j, k);
    const double a = xi[j], b = xi[k], c = X[k + (size_t)j * n];
#pragma unroll
    for (int type = 1; type <= 4; ++type) {
      const double v = tri_val(type, a, b, c);
      if (PASS == 0) {
        vmax_loc = fmax(vmax_loc, v);
        cnt_loc += v > thr;
      } else if (v > thr) {
        int bin = (int)((v - lo) * scale);
        bin = bin < 0 ? 0 : (bin >= SEP_NBIN ? SEP_NBIN - 1 : bin);
        if (PASS == 1) {
          atomicAdd(&hist[bin], 1u);
        } else {
          const uint32_t key = tri_key(type, i, j, k);
          if (emit_all || bin > bstar) {
            const uint32_t slot = atomicAdd(&s.ctl[CTL_NOUT], 1u);
            if (slot < (uint32_t)s.K) s.okey[slot] = key;
          } else if (bin == bstar) {
            const uint32_t slot = atomicAdd(&s.ctl[CTL_NBOUND], 1u);
            if (slot < (uint32_t)SEP_BCAP) {
              s.bkey[slot] = key;
              s.bval[slot] = v;
            }
          }
        }
      }
    }
  }
  if (PASS == 0) {
    atomicMax(&smax, ord_bits(vmax_loc));
    atomicAdd(&scount, cnt_loc);
    __syncthreads();
    if (tid == 0) {
      atomicMax(&s.maxbits[0], smax);
      atomicAdd(&s.ctl[CTL_COUNT], scount);
    }
  } else if (PASS == 1) {
    __syncthreads();
    for (int b = tid; b < SEP_NBIN; b += SEP_WG)
      if (hist[b]) atomicAdd(&s.hist[b], hist[b]);
  }
}

// One workgroup per node: find the bin holding the K-th largest violated value.
__global__ __launch_bounds__(SEP_WG) void k_tri_scan(SepView* sv) {
  SepView& s = sv[blockIdx.x];
  if (threadIdx.x != 0) return;
  const uint32_t count = s.ctl[CTL_COUNT];
  if (count <= (uint32_t)s.K) {
    s.ctl[CTL_BSTAR] = 0;
    s.ctl[CTL_NEED] = 0;
    return;
  }
  uint32_t above = 0;
  int b = SEP_NBIN - 1;
  for (; b >= 0; --b) {
    if (above + s.hist[b] >= (uint32_t)s.K) break;
    above += s.hist[b];
  }
  s.ctl[CTL_BSTAR] = (uint32_t)b;
  s.ctl[CTL_NEED] = (uint32_t)s.K - above;
}

// One workgroup per node: finish the list.  Boundary rows are ranked by
// (value desc, key asc) and the first `need` kept; the list is then written
// in key (= enumeration) order with the row values recomputed from X.
// A boundary bin with more rows than SEP_BCAP (many equal or crowded
// values) is not ranked here: the node is flagged (CTL_OVER = 1), its
// refinement state is set up, and the list is finished by the AFTER
// instance once k_tri_refine_* have selected the boundary exactly.
template <bool AFTER>
__global__ __launch_bounds__(1024) void k_tri_finish(SepView* sv) {
  SepView& s = sv[blockIdx.x];
  const int tid = threadIdx.x, NT = blockDim.x;
  if (AFTER && s.ctl[CTL_OVER] != 1) return;
  const uint32_t count = s.ctl[CTL_COUNT];
  uint32_t nout = s.ctl[CTL_NOUT];
  if (nout > (uint32_t)s.K) nout = s.K;
  if (count > (uint32_t)s.K) {
    uint32_t nb = s.ctl[CTL_NBOUND];
    if (!AFTER && nb > (uint32_t)SEP_BCAP) {
      // the pass-1 histogram is spent: it becomes the refinement's
      for (int b = tid; b < SEP_NBIN; b += NT) s.hist[b] = 0;
      if (tid == 0) {
        s.ctl[CTL_OVER] = 1;
        s.ctl[CTL_RMODE] = 1;
        s.ctl[CTL_RL] = 0;
        s.ctl[CTL_RNEED] = s.ctl[CTL_NEED];
        s.ctl[CTL_NBOUND] = 0;
        s.ctl[CTL_NFINAL] = 0;
      }
      return;
    }
    if (nb > SEP_BCAP) nb = SEP_BCAP;  // not reached: the refinement collects at most SEP_BCAP
    const uint32_t need = s.ctl[CTL_NEED];
    for (uint32_t q = tid; q < nb; q += NT) {
      const double v = s.bval[q];
      const uint32_t key = s.bkey[q];
      uint32_t rank = 0;
      for (uint32_t r = 0; r < nb; ++r) {
        const double w = s.bval[r];
        rank += (w > v) || (w == v && s.bkey[r] < key);
      }
      if (rank < need) s.okey[nout + rank] = key;
    }
    __syncthreads();
    nout += need < nb ? need : nb;
    if (nout > (uint32_t)s.K) nout = s.K;
  }
  // sort keys ascending by rank counting (keys are distinct)
  const int n = s.n;
  for (uint32_t q = tid; q < nout; q += NT) {
    const uint32_t key = s.okey[q];
    uint32_t rank = 0;
    for (uint32_t r = 0; r < nout; ++r) rank += s.okey[r] < key;
    const int type = (int)(key >> 30) + 1, i = (key >> 20) & 1023, j = (key >> 10) & 1023,
              k = key & 1023;
    s.fkey[rank] = key;
    s.fval[rank] = tri_val(type, s.X[i + (size_t)j * n], s.X[i + (size_t)k * n],
                           s.X[j + (size_t)k * n]);
  }
  if (tid == 0) {
    s.ctl[CTL_NFINAL] = nout;
    if (AFTER) s.ctl[CTL_OVER] = 2;
  }
}

// ---- exact selection inside an overflowing boundary bin
//
// The rows of the boundary bin are ordered by the 96-bit key
// C = (ord(value) : 64 bits, ~key : 32 bits), so that C descending is
// (value descending, enumeration order ascending), the order the list
// keeps.  A most-significant-digit radix selection on 12-bit digits fixes
// C's digits one level at a time: each level histograms the next digit of
// the rows that match the digits fixed so far, the scan picks the digit d*
// holding the `need`-th largest, the rows above d* join the list, and the
// rows at d* are either refined further or -- once they number at most
// SEP_BCAP (at the latest after all 8 digits, when one row is left) --
// collected for k_tri_finish<true>, which ranks them as usual.
// Refinement modes (CTL_RMODE): 1 histogram digit 0; 2 emit the rows above
// the last fixed digit and histogram the next one; 3 emit and collect;
// 4 done.

// digit p (0 = most significant) of C
__device__ __forceinline__ uint32_t ref_digit(unsigned long long o, uint32_t nk, int p) {
  const int off = 84 - 12 * p;  // lowest bit of digit p in C
  if (off >= 32) return (uint32_t)(o >> (off - 32)) & 0xFFFu;
  if (off + 12 <= 32) return (nk >> off) & 0xFFFu;
  return (uint32_t)((o << (32 - off)) | (unsigned long long)(nk >> off)) & 0xFFFu;
}

__global__ __launch_bounds__(SEP_WG) void k_tri_refine_pass(SepView* sv) {
  __shared__ double xi[1024];
  __shared__ uint32_t hist[SEP_NBIN];
  SepView& s = sv[blockIdx.y];
  const uint32_t mode = s.ctl[CTL_RMODE];
  if (mode == 0 || mode >= 4) return;
  const int n = s.n, i = blockIdx.x;
  if (i >= n || i < 2) return;
  const int L = (int)s.ctl[CTL_RL];
  uint32_t rd[SEP_RLEVELS];
#pragma unroll
  for (int p = 0; p < SEP_RLEVELS; ++p) rd[p] = p < L ? s.ctl[CTL_RD0 + p] : 0u;
  const double* __restrict__ X = s.X;
  const int tid = threadIdx.x;
  for (int j = tid; j < i; j += SEP_WG) xi[j] = X[i + (size_t)j * n];
  const bool counting = mode <= 2;
  if (counting)
    for (int b = tid; b < SEP_NBIN; b += SEP_WG) hist[b] = 0;
  __syncthreads();
  const double thr = s.thresh, lo = thr;
  const double vmax = from_ord(s.maxbits[0]);
  const double scale = (vmax > thr) ? (double)SEP_NBIN / (vmax - thr) : 0.0;
  const int bstar = (int)s.ctl[CTL_BSTAR];
  const int npairs = i * (i - 1) / 2;
  for (int p = tid; p < npairs; p += SEP_WG) {
    int j, k;
    pair_decode(p, j, k);
    const double a = xi[j], b = xi[k], c = X[k + (size_t)j * n];
#pragma unroll
    for (int type = 1; type <= 4; ++type) {
      const double v = tri_val(type, a, b, c);
      if (!(v > thr)) continue;
      int bin = (int)((v - lo) * scale);
      bin = bin < 0 ? 0 : (bin >= SEP_NBIN ? SEP_NBIN - 1 : bin);
      if (bin != bstar) continue;
      const uint32_t key = tri_key(type, i, j, k);
      const unsigned long long o = ord_bits(v);
      const uint32_t nk = ~key;
      if (mode == 1) {
        atomicAdd(&hist[ref_digit(o, nk, 0)], 1u);
        continue;
      }
      bool match = true;
      for (int q = 0; q + 1 < L; ++q) match = match && ref_digit(o, nk, q) == rd[q];
      if (!match) continue;
      const uint32_t d = ref_digit(o, nk, L - 1);
      const uint32_t dstar = rd[L - 1];
      if (d > dstar) {
        const uint32_t slot = atomicAdd(&s.ctl[CTL_NOUT], 1u);
        if (slot < (uint32_t)s.K) s.okey[slot] = key;
      } else if (d == dstar) {
        if (mode == 2) {
          atomicAdd(&hist[ref_digit(o, nk, L)], 1u);
        } else {
          const uint32_t slot = atomicAdd(&s.ctl[CTL_NBOUND], 1u);
          if (slot < (uint32_t)SEP_BCAP) {
            s.bkey[slot] = key;
            s.bval[slot] = v;
          }
        }
      }
    }
  }
  if (counting) {
    __syncthreads();
    for (int b = tid; b < SEP_NBIN; b += SEP_WG)
      if (hist[b]) atomicAdd(&s.hist[b], hist[b]);
  }
}

// One workgroup per node: fix the next digit from the histogram, or end
// the refinement after the collecting pass.
__global__ __launch_bounds__(SEP_WG) void k_tri_refine_scan(SepView* sv) {
  SepView& s = sv[blockIdx.x];
  const uint32_t mode = s.ctl[CTL_RMODE];
  if (mode == 0 || mode >= 4) return;
  if (mode == 3) {
    if (threadIdx.x == 0) {
      s.ctl[CTL_NEED] = s.ctl[CTL_RNEED];
      s.ctl[CTL_RMODE] = 4;
    }
    return;
  }
  if (threadIdx.x == 0) {
    const uint32_t need = s.ctl[CTL_RNEED];
    uint32_t above = 0;
    int d = SEP_NBIN - 1;
    for (; d > 0; --d) {
      if (above + s.hist[d] >= need) break;
      above += s.hist[d];
    }
    const uint32_t at = s.hist[d];
    const uint32_t L = s.ctl[CTL_RL];
    s.ctl[CTL_RD0 + L] = (uint32_t)d;
    s.ctl[CTL_RL] = L + 1;
    s.ctl[CTL_RNEED] = need - above;
    s.ctl[CTL_RMODE] = (at <= (uint32_t)SEP_BCAP || L + 1 == (uint32_t)SEP_RLEVELS) ? 3u : 2u;
  }
  __syncthreads();
  for (int b = threadIdx.x; b < SEP_NBIN; b += SEP_WG) s.hist[b] = 0;
}

void launch_tri_separate(SepView* sv, int nb, int nmax, hipStream_t st) {
  dim3 grid(nmax, nb);
  k_tri_pass<0><<<grid, SEP_WG, 0, st>>>(sv);
  k_tri_pass<1><<<grid, SEP_WG, 0, st>>>(sv);
  k_tri_scan<<<nb, SEP_WG, 0, st>>>(sv);
  k_tri_pass<2><<<grid, SEP_WG, 0, st>>>(sv);
  k_tri_finish<false><<<nb, 1024, 0, st>>>(sv);
}

void launch_tri_refine(SepView* sv, int nb, int nmax, hipStream_t st) {
  dim3 grid(nmax, nb);
  // mode 1, then at most SEP_RLEVELS - 1 levels in mode 2, then mode 3
  for (int l = 0; l <= SEP_RLEVELS; ++l) {
    k_tri_refine_pass<<<grid, SEP_WG, 0, st>>>(sv);
    k_tri_refine_scan<<<nb, SEP_WG, 0, st>>>(sv);
  }
  k_tri_finish<true><<<nb, 1024, 0, st>>>(sv);
}

// ------------------------------------------------------------------ SA
//
// qap_simulated_annealing (qap_simulated_annealing.c:12-206), one lane per
// run.  Run r draws from its own stream rand31(seed, r, 0), (..., 1), ...
// (splitmix64 based, values in [0, 2^31-1] like glibc rand()).  The
// permutation lives in LDS (16-bit entries); X is read from HBM.  Same
// arithmetic order as the reference (contraction off), so with the same
// stream the run is identical draw for draw.

__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ int rand31(uint64_t base, uint64_t& ctr) {
  return (int)(splitmix64(base + ctr++) >> 33);
}

// Jobs are laid out node by node, `per` runs each; grid (ceil(per / SA_WG),
// nodes), so that a workgroup's runs share one X.  XM = 1: the workgroup
// first copies X into LDS (n^2 doubles after the permutation buffers) and
// every run reads it there: same values, same trajectory, LDS latency.
// XM = 2 (orders whose X fits LDS only by half): the lower triangle only,
// row by row (entry (a, b), a >= b, at a (a + 1) / 2 + b).  Every mode reads
// entry (a, b) from the lower triangle, as the reference does after copying
// X's lower triangle over its upper (qap_simulated_annealing.c:14-17).
// x mod n for 0 <= x < 2^31 by a double-precision reciprocal and one
// correction (the quotient estimate is off by at most one): exact, and
// cheaper than the integer division of a run-time divisor.
__device__ __forceinline__ int mod_n(int x, int n, double inv_n) {
  const int q = (int)((double)x * inv_n);
  int r = x - q * n;
  if (r < 0)
    r += n;
  else if (r >= n)
    r -= n;
  return r;
}

// entry (a, b) of a symmetric matrix from its lower triangle: packed row
// by row, or in place in the column-major n x n matrix
__device__ __forceinline__ double xp_at(const double* __restrict__ P, int a, int b) {
  const int hi = a > b ? a : b, lo = a > b ? b : a;
  return P[((hi * (hi + 1)) >> 1) + lo];
}
__device__ __forceinline__ double x_low(const double* __restrict__ X, int n, int a, int b) {
  const int hi = a > b ? a : b, lo = a > b ? b : a;
  return X[(size_t)hi + (size_t)lo * n];
}

// K: the cut order of every run of the launch (5 or 7; 0 = read from the
// job), so that the loops over the permutation's first K entries unroll.
template <int XM, int K>
__global__ __launch_bounds__(SA_WG) void k_qap_sa(const SaJob* jobs, int per, SaOut* out) {
#pragma clang fp contract(off)
  extern __shared__ unsigned short permbuf[];
  const int rr0 = blockIdx.x * SA_WG + threadIdx.x;
  const int r = blockIdx.y * per + rr0;
  const int first = blockIdx.y * per + blockIdx.x * SA_WG;  // a valid job of this workgroup
  const int nst = jobs[first].nstride;
  const double* __restrict__ X = jobs[first].X;
  const double* __restrict__ Xp = nullptr;  // XM = 2: packed lower triangle
  if (XM == 1) {
    const int nx = jobs[first].n;
    double* Xl = (double*)(permbuf + (((size_t)SA_WG * nst + 3) & ~(size_t)3));
    for (int f = threadIdx.x; f < nx * nx; f += SA_WG) Xl[f] = X[f];
    __syncthreads();
    X = Xl;
  } else if (XM == 2) {
    const int nx = jobs[first].n;
    double* Xl = (double*)(permbuf + (((size_t)SA_WG * nst + 3) & ~(size_t)3));
    for (int f = threadIdx.x; f < nx * nx; f += SA_WG) {
      const int b = f / nx, a = f - b * nx;  // X[f] = entry (a, b)
      if (a >= b) Xl[((a * (a + 1)) >> 1) + b] = X[f];
    }
    __syncthreads();
    Xp = Xl;
  }
  if (rr0 >= per) return;
  const SaJob J = jobs[r];
  const int n = J.n, k = K ? K : J.k;
  const double inv_n = 1.0 / (double)n;
  unsigned short* perm = permbuf + (size_t)threadIdx.x * nst;
  uint64_t base = splitmix64(J.seed ^ splitmix64((uint64_t)J.stream_id));
  uint64_t ctr = 0;
  // H = e e^T with e_a = -1 for a < neg (cutting_planes.c:242-244, 438-441)
#define SG(a) ((a) < J.neg ? -1 : 1)
#define XV(a, b) (XM == 2 ? xp_at(Xp, a, b) : x_low(X, n, a, b))
  const double sumX = *J.sumX;
  double temp0 = (double)(k * k) * sumX / (double)(n * (n - 1));
  double cost = 1e9;
  int best[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int i = 0; i < n; ++i) perm[i] = (unsigned short)i;
  for (int i = n - 1; i > 0; --i) {
    const int rr = rand31(base, ctr) % i;
    const unsigned short sw = perm[i];
    perm[i] = perm[rr];
    perm[rr] = sw;
  }
  double sol = 0.0;
  for (int i = 0; i < k; ++i) {
    sol += XV(perm[i], perm[i]);
    for (int j = i + 1; j < k; ++j) sol += 2 * (SG(j) * SG(i)) * XV(perm[j], perm[i]);
  }
  double T = temp0;
  int m1 = n;
  int changed = 1;
  long moves = 0;
  while (changed > 0) {
    changed = 0;
    for (int it = 0; it < m1; ++it) {
      int i1 = rand31(base, ctr) % k;
      int i2 = mod_n(rand31(base, ctr), n, inv_n);
      if (i2 < i1) {
        const int sw = i1;
        i1 = i2;
        i2 = sw;
      }
      const int p1 = perm[i1], p2 = perm[i2];
      double delta = 0.0;
#pragma unroll
      for (int i = 0; i < (K ? K : k); ++i) {
        const int pi = perm[i];
        delta += (SG(i) * SG(i1)) * (XV(pi, p2) - XV(pi, p1));
      }
      if (i2 < k) {
#pragma unroll
        for (int i = 0; i < (K ? K : k); ++i) {
          const int pi = perm[i];
          delta += (SG(i) * SG(i2)) * (XV(pi, p1) - XV(pi, p2));
        }
      }
      delta *= 2.0;
      const double d1 = XV(p1, p1) + XV(p2, p2) - 2.0 * XV(p2, p1);
      double d2 = 1.0;
      if (i2 < k) d2 += -2 * (SG(i1) * SG(i2)) + 1.0;
      delta += d1 * d2;
      const double cand = sol + delta;
      int accept;
      if (delta > 0) {
        const double rt = delta / T;
        if (rt > 5)
          accept = 0;
        else
          accept = ((double)rand31(base, ctr) / ((double)2147483647)) < exp(-rt);
      } else {
        accept = 1;
      }
      if (accept) {
        if (fabs(delta) > 0.0001) changed = 1;
        perm[i1] = (unsigned short)p2;
        perm[i2] = (unsigned short)p1;
        sol = cand;
        if (sol < cost) {
          cost = sol;
          for (int a = 0; a < k; ++a) best[a] = perm[a];
        }
      }
      ++moves;
    }
    T *= 0.6;
    m1 = (int)round(m1 * 1.1);
  }
#undef XV
#undef SG
  out[r].cost = cost;
  for (int a = 0; a < 7; ++a) out[r].perm[a] = a < k ? best[a] : -1;
  out[r].draws = (long long)ctr;
  out[r].moves = moves;
}

// sum_i X_ii + 2 sum_{j>i} |X_ji| in the reference's order
// (qap_simulated_annealing.c:28-35), one workgroup per node: the terms of a
// run of columns are staged in LDS in that order (coalesced, all threads),
// then one thread adds them in sequence -- the same bits as the plain loop.
constexpr int SUMX_WG = 256, SUMX_BUF = 8192;

__global__ __launch_bounds__(SUMX_WG) void k_sumx(const double* const* Xs, const int* ns,
                                                 double* out) {
#pragma clang fp contract(off)
  __shared__ double buf[SUMX_BUF];
  const int b = blockIdx.x, tid = threadIdx.x;
  const double* __restrict__ X = Xs[b];
  const int n = ns[b];
  double acc = 0.0;
  for (int i0 = 0; i0 < n;) {
    // columns [i0, i1): their n - i terms each fill at most SUMX_BUF slots
    int i1 = i0, cnt = 0;
    while (i1 < n && cnt + (n - i1) <= SUMX_BUF) cnt += n - i1++;
    __syncthreads();  // the previous run's terms have been added
    for (int f = i0 * n + tid; f < i1 * n; f += SUMX_WG) {
      const int i = f / n, j = f - i * n;
      if (j < i) continue;
      // slot of (j, i): the terms of columns i0..i-1, then j - i
      const int base = (i - i0) * n - ((i * (i - 1) - i0 * (i0 - 1)) >> 1);
      buf[base + (j - i)] = j == i ? X[f] : 2 * fabs(X[f]);
    }
    __syncthreads();
    if (tid == 0)
      for (int t = 0; t < cnt; ++t) acc += buf[t];
    i0 = i1;
  }
  if (tid == 0) out[b] = acc;
}

void launch_sumx(const double* const* X, const int* n, double* out, int nb, hipStream_t st) {
  k_sumx<<<nb, SUMX_WG, 0, st>>>(X, n, out);
}

// Layout: X in LDS (whole, else its lower triangle) where that still lets
// every workgroup of the launch be resident at once -- each is one wave of
// serial chains, and a CU that runs several of them (one per SIMD) does
// more than one LDS-fed wave; else X is read from global memory.
// pw05_150.0 proof, 1024 nodes in flight: 207 s with X in LDS wherever it
// fits whole, 214 s with the lower triangle in LDS as well, 201 s (annealing
// phase 34 -> 28 s) with this rule.
void launch_qap_sa(const SaJob* jobs, int nnodes, int per, int nmax, int kk, SaOut* out,
                   size_t lds_limit, int cu_count, hipStream_t st) {
  const dim3 grid((per + SA_WG - 1) / SA_WG, nnodes);
  const size_t lperm = (((size_t)SA_WG * sa_nstride(nmax) + 3) & ~(size_t)3) * sizeof(unsigned short);
  const size_t lx = (size_t)nmax * nmax * sizeof(double);
  const size_t lxp = (size_t)nmax * (nmax + 1) / 2 * sizeof(double);
  const long wgs = (long)grid.x * grid.y;
  const long per_cu = std::max(1L, (wgs + cu_count - 1) / std::max(1, cu_count));
  const size_t room = lds_limit / (size_t)std::min(per_cu, 8L);
  const int xm = lperm + lx <= room ? 1 : (lperm + lxp <= room ? 2 : 0);
  const size_t lds = lperm + (xm == 1 ? lx : (xm == 2 ? lxp : 0));
#define MADAM_SA(XMV, KV) k_qap_sa<XMV, KV><<<grid, SA_WG, lds, st>>>(jobs, per, out)
#define MADAM_SA_K(XMV)   \
  if (kk == 5)            \
    MADAM_SA(XMV, 5);     \
  else if (kk == 7)       \
    MADAM_SA(XMV, 7);     \
  else                    \
    MADAM_SA(XMV, 0);
  if (xm == 1) {
    MADAM_SA_K(1)
  } else if (xm == 2) {
    MADAM_SA_K(2)
  } else {
    MADAM_SA_K(0)
  }
#undef MADAM_SA_K
#undef MADAM_SA
}

}  // namespace madam
