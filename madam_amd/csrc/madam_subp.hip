// madam_subp.hip -- node subproblems on the device and the rudy reader.
//
// createSubproblem (evaluate.c:39-116) for a batch of nodes: one workgroup
// per node reads SP->L (resident in HBM for the whole search) and the
// node's fixings, and writes PP->L straight into the node's slot, so that a
// node costs an upload of its 2 (N-1) fixing flags instead of its n^2
// matrix.  Every sum runs in the order of the host build (the reference's
// loops), so the matrix is bit-identical to madam_create_subproblem's:
//   frow[f] = sum over fixed i with x_i = 1, i ascending, of SP->L(idx_f, i)
//   column f of PP->L = the free rows of SP->L's column idx_f, its free
//     row sum (rows ascending) in the last row, mirrored to the last column
//   the diagonal, last row and last column get + 2 frow, the corner the sum
//     of the last column in order, then everything x 1/4.
// |PP->L|_F (the ADMM's dual-error scale) is summed by one thread in
// column-major order with contraction off, as the host does.
//
// readData (process_input.c:182-312): the rudy text format "n m" then m
// lines "i j w" (1-based vertices) into SP->L = [Lap, Lap e; e^T Lap,
// e^T Lap e] with the Laplacian of the first n-1 vertices, the same sums in
// the same order; a repeated edge overwrites the earlier weight.
#include <hip/hip_runtime.h>
#include <stdio.h>

#include <vector>

#include "madam_engine.h"

namespace madam {

// Output of one node: its matrix (device, n x n, leading dimension n),
// order and Frobenius norm.
struct SubpJob {
  double* PPL;
  int n;
  double normL;
};

__global__ __launch_bounds__(256) void k_subproblem(const double* __restrict__ SPL, int N,
                                                    const int* __restrict__ fix_all,
                                                    SubpJob* jobs) {
#pragma clang fp contract(off)
  __shared__ int idx[MADAM_NMAX];     // free vertices, ascending
  __shared__ int fixed1[MADAM_NMAX];  // fixed vertices on side 1, ascending
  __shared__ int xf[MADAM_NMAX];      // xfixed, then solX
  __shared__ int sx[MADAM_NMAX];
  __shared__ double frow[MADAM_NMAX];
  __shared__ int cnt[2];
  const int tid = threadIdx.x, bs = N - 1;
  const int* fix = fix_all + (size_t)blockIdx.x * 2 * bs;
  SubpJob& J = jobs[blockIdx.x];
  for (int i = tid; i < bs; i += 256) {
    xf[i] = fix[i];
    sx[i] = fix[bs + i];
  }
  __syncthreads();
  if (tid == 0) {
    int nf = 0, n1 = 0;
    for (int i = 0; i < bs; ++i) {
      if (!xf[i])
        idx[nf++] = i;
      else if (sx[i] == 1)
        fixed1[n1++] = i;
    }
    cnt[0] = nf;
    cnt[1] = n1;
  }
  __syncthreads();
  const int nfree = cnt[0], n1 = cnt[1], n = nfree + 1;
  double* P = J.PPL;
  for (int f = tid; f < nfree; f += 256) {
    double a = 0.0;
    for (int q = 0; q < n1; ++q) a += SPL[idx[f] + (size_t)fixed1[q] * N];
    frow[f] = a;
  }
  // column f: the free rows of SP->L's column idx_f and their sum
  for (int f = tid; f < nfree; f += 256) {
    const double* src = SPL + (size_t)idx[f] * N;
    double* dst = P + (size_t)f * n;
    double rs = 0.0;
    for (int r = 0; r < nfree; ++r) {
      const double v = src[idx[r]];
      dst[r] = v;
      rs += v;
    }
    dst[n - 1] = rs;
  }
  __syncthreads();
  // last column = last row, then the fixed rows folded in
  for (int i = tid; i < nfree; i += 256) {
    const double fr2 = 2 * frow[i];
    P[i + (size_t)(n - 1) * n] = P[(n - 1) + (size_t)i * n] + fr2;
    P[(n - 1) + (size_t)i * n] += fr2;
    P[i + (size_t)i * n] += fr2;
  }
  __syncthreads();
  if (tid == 0) {
    double corner = 0.0;
    for (int i = 0; i < nfree; ++i) corner += P[i + (size_t)(n - 1) * n];
    P[(n - 1) + (size_t)(n - 1) * n] = corner;
  }
  __syncthreads();
  const size_t nn = (size_t)n * n;
  for (size_t e = tid; e < nn; e += 256) P[e] *= 0.25;
  __syncthreads();
  // |L|_F in frob_norm_cols's order: column sums (one thread per column,
  // top to bottom), then their sum left to right
  for (int j = tid; j < n; j += 256) {
    const double* c = P + (size_t)j * n;
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += c[i] * c[i];
    frow[j] = s;
  }
  __syncthreads();
  if (tid == 0) {
    double fro = 0.0;
    for (int j = 0; j < n; ++j) fro += frow[j];
    J.n = n;
    J.normL = sqrt(fro);
  }
}

void launch_subproblem(const double* SPL, int N, const int* fix_all, SubpJob* jobs, int k,
                       hipStream_t st) {
  k_subproblem<<<k, 256, 0, st>>>(SPL, N, fix_all, jobs);
}

}  // namespace madam

namespace mh {

// The instance's SP->L in HBM for the subproblem kernel: uploaded at the
// start of every call that builds subproblems (the caller owns the host
// matrix and may change it between calls).
static DBuf<double> g_spl;
static int g_spl_N = 0;
static DBuf<int> g_fix;                 // the nodes' fixings, 2 (N-1) ints each
static DBuf<madam::SubpJob> g_jobs;

int upload_spl(int N, const double* SPL) {
  int rc;
  if ((rc = g_spl.ensure((size_t)N * N)) || (rc = up(g_spl.p, SPL, (size_t)N * N, g_eng->stream)))
    return rc;
  g_spl_N = N;
  return MADAM_OK;
}

int device_subproblems(int N, int k, const int* const* xfixed, const int* const* solX,
                       double* const* dst, int* n_out, double* norm_out) {
  Engine& E = *g_eng;
  const hipStream_t st = E.stream;
  int rc;
  if (g_spl_N != N || !g_spl.p) {
    set_err("device_subproblems: SP->L of order %d not uploaded", N);
    return MADAM_E_ARG;
  }
  const int bs = N - 1;
  std::vector<int> fix((size_t)k * 2 * bs);
  std::vector<madam::SubpJob> jobs(k);
  for (int q = 0; q < k; ++q) {
    for (int i = 0; i < bs; ++i) {
      fix[(size_t)q * 2 * bs + i] = xfixed[q][i] ? 1 : 0;
      fix[(size_t)q * 2 * bs + bs + i] = solX[q][i];
    }
    jobs[q].PPL = dst[q];
  }
  if ((rc = g_fix.ensure(fix.size())) || (rc = g_jobs.ensure(k)) ||
      (rc = up(g_fix.p, fix.data(), fix.size(), st)) || (rc = up(g_jobs.p, jobs.data(), k, st)))
    return rc;
  madam::launch_subproblem(g_spl.p, N, g_fix.p, g_jobs.p, k, st);
  HIP_TRY(hipGetLastError());
  if ((rc = down(jobs.data(), g_jobs.p, k, st))) return rc;
  HIP_TRY(hipStreamSynchronize(st));
  for (int q = 0; q < k; ++q) {
    if (n_out) n_out[q] = jobs[q].n;
    if (norm_out) norm_out[q] = jobs[q].normL;
  }
  return MADAM_OK;
}

void release_spl() {
  g_spl.release();
  g_fix.release();
  g_jobs.release();
  g_spl_N = 0;
}

}  // namespace mh

using namespace mh;

extern "C" {

int madam_create_subproblem_batch(int N, const double* SPL, int k, const int* xfixed,
                                  const int* solX, double* PPL, int* n_out) {
  int rc = ensure_open();
  if (rc) return rc;
  if (N < 2 || N > MADAM_NMAX || !SPL || k <= 0 || !xfixed || !solX || !PPL || !n_out) {
    set_err("madam_create_subproblem_batch: bad arguments (N=%d, k=%d)", N, k);
    return MADAM_E_ARG;
  }
  const size_t NN = (size_t)N * N;
  struct Out {
    DBuf<double> b;
    ~Out() { b.release(); }
  } out;
  if ((rc = out.b.ensure((size_t)k * NN))) return rc;
  std::vector<const int*> xf(k), sx(k);
  std::vector<double*> dst(k);
  for (int q = 0; q < k; ++q) {
    xf[q] = xfixed + (size_t)q * (N - 1);
    sx[q] = solX + (size_t)q * (N - 1);
    dst[q] = out.b.p + (size_t)q * NN;
  }
  if ((rc = upload_spl(N, SPL)) ||
      (rc = device_subproblems(N, k, xf.data(), sx.data(), dst.data(), n_out, nullptr)))
    return rc;
  for (int q = 0; q < k; ++q)
    if ((rc = down(PPL + (size_t)q * NN, out.b.p + (size_t)q * NN, (size_t)n_out[q] * n_out[q],
                   g_eng->stream)))
      return rc;
  HIP_TRY(hipStreamSynchronize(g_eng->stream));
  return MADAM_OK;
}

int madam_read_rudy(const char* path, int capN, int* N, double* SPL) {
  if (!path || !N || !SPL || capN < 2) {
    set_err("madam_read_rudy: bad arguments");
    return MADAM_E_ARG;
  }
  FILE* f = fopen(path, "r");
  if (!f) {
    set_err("madam_read_rudy: cannot open %s", path);
    return MADAM_E_ARG;
  }
  int nv = 0, ne = 0;
  if (fscanf(f, "%d %d", &nv, &ne) != 2 || nv <= 0 || ne < 0) {
    fclose(f);
    set_err("madam_read_rudy: %s: bad header (vertices and edges)", path);
    return MADAM_E_ARG;
  }
  if (nv < 2 || nv > capN || nv > MADAM_NMAX) {
    fclose(f);
    set_err("madam_read_rudy: %s: %d vertices (2..%d)", path, nv,
            capN < MADAM_NMAX ? capN : MADAM_NMAX);
    return MADAM_E_ARG;
  }
  std::vector<double> adj((size_t)nv * nv, 0.0);
  for (int e = 0; e < ne; ++e) {
    int i = 0, j = 0;
    double w = 0.0;
    if (fscanf(f, "%d %d %lf", &i, &j, &w) != 3) {
      fclose(f);
      set_err("madam_read_rudy: %s: edge %d unreadable", path, e + 1);
      return MADAM_E_ARG;
    }
    if (i < 1 || i > nv || j < 1 || j > nv) {
      fclose(f);
      set_err("madam_read_rudy: %s: edge %d (%d, %d) has a vertex outside 1..%d", path, e + 1, i,
              j, nv);
      return MADAM_E_ARG;
    }
    adj[(size_t)nv * (j - 1) + (i - 1)] = w;
    adj[(size_t)nv * (i - 1) + (j - 1)] = w;
  }
  fclose(f);
  std::vector<double> deg(nv, 0.0);
  for (int a = 0; a < nv; ++a)
    for (int b = 0; b < nv; ++b) deg[a] += adj[b + (size_t)a * nv];
  double sum = 0.0;
  for (int a = 0; a < nv; ++a) {
    double row = 0.0;
    for (int b = 0; b < nv; ++b) {
      double& L = SPL[b + (size_t)a * nv];
      if (a < nv - 1 && b < nv - 1) {
        L = (a == b ? deg[a] : 0.0) - adj[b + (size_t)a * nv];
        row += L;
      } else if (b == nv - 1 && a != nv - 1) {
        L = row;
        sum += row;
      } else if (a == nv - 1 && b != nv - 1) {
        L = SPL[a + (size_t)b * nv];
      } else {
        L = sum;
      }
    }
  }
  *N = nv;
  return MADAM_OK;
}

}  // extern "C"
