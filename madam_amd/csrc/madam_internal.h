// madam_internal.h -- device-side node layout shared by the HIP kernels and
// the host engine.  Not part of the C ABI (include/madam_bound.h is).
//
// One B&B node's ADMM state lives in HBM as a NodeView: FP64 column-major
// n x n matrices with leading dimension n, the cut rows of B in two CSR
// forms, and the per-node scalars of the ADMM loop.  A batch is an array of
// NodeViews; every kernel takes the array and a node index from blockIdx.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace madam {

// Cut row encoding.  A cut c with vertex tuple v[0..k-1] and sign vector b
// has the row B(X)_c = scale_c * sum_{a<b} sgn_ab * X[v_a, v_b] with
// sgn_ab = -b_a b_b and scale_c = 1, 1/2, 1/3 for k = 3, 5, 7
// (operators.c:85-148).  Each (a,b) term is one CSR entry holding the lower
// linear index lo = max(v_a,v_b) + n*min(v_a,v_b), stored as lo for
// sgn = +1 and ~lo for sgn = -1, in the reference's pair order
// (0,1),(0,2),...,(1,2),...  The transpose groups the same entries by
// "active pair" (a distinct vertex pair touched by at least one cut): each
// active pair lists (cut, sign) in increasing cut order, stored as c or ~c.
struct NodeView {
  int n;       // matrix order (PP->n)
  int m;       // cuts (triangles + pentagonals + heptagonals, in that order)
  int nact;    // active pairs
  int pad0;

  double *L, *X, *Q, *W, *V, *ev, *y;  // n*n (L, X, Q, W, V), n (ev, y)
  double *S1, *S2;                     // n*n scratch
  double *evp;                         // n: kept eigenvalues of the latest projection (0 if dropped)
  double *ip1, *ip2, *ip3;             // n*n IPM scratch
  double *ipv;                         // 4n IPM vectors
  double *ipP;                         // n(n+1)/2 packed Cholesky factor (when not in LDS)
  float *ns32;                         // 3 n_p^2 FP32 sign-iteration scratch (n_p = n rounded up to 16)

  const int *cptr;      // m+1
  const int *cent;      // cptr[m]: lo or ~lo
  const int *cact;      // cptr[m]: active-pair id a of the entry, as a or ~a like cent
  const double *cscale; // m
  const int *aptr;      // nact+1
  const int *aent;      // aptr[nact]: c or ~c
  const short *aent16;  // aent and cact in 16 bits (null unless m, nact < 2^15)
  const short *cact16;
  // sliced Gram rows of BB^T for the one-pass cut solve (null when not
  // built): slot i = cut tcut[i], of thread i % XCUT_THREADS; slice s =
  // slots 64s..64s+63 with word q of lane l at xent[xoff[s] + 64 q + l]
  const int *xoff;      // m/64 + 1
  const unsigned *xent; // words: LDS byte offset 8 (2 c' + neg) of +-ps[c'] (16 m: zero)
  const int *tcut;      // m: cut of each slot (rows longest first)
  const int *tmeta;     // m: pair entries of each slot's cut
  const double *tsc;    // m: scale of each slot's cut
  const double *tdinv;  // m: diagonal preconditioner of each slot
  double *xs;           // m: the cut solve's iterate in slot order
  const int *apair;     // nact: lower linear index of the active pair
  const int *p2a;       // n*n: active-pair id of lower index lo, or -1

  double *t, *s, *u, *rhs;             // m
  double *tprev, *tprev2;              // m: the cut solve's solutions one and two iterations
                                       // back (in slot order for the one-pass solve)
  double *cg_r, *cg_p, *cg_q, *cg_w;   // m, m, m, nact

  // per-node ADMM scalars
  double sigma, tol, scaleSigma, normL;
  double f, lam, lam_margin, off_last;
  double err, err_p, err_d, gap;
  double primal, dual;
  double sigma_last;   // sigma used by the latest projection (X = sigma_last * X_+)
  int itermax, nbit, done, M;
  int cg_iters, cg_total, sweeps_last, sweeps_total;
  int eig_warm;   // V holds a usable basis from the previous projection
  int ipm_iters;  // IPM iterations of the last k_ipm
  int ipm_status; // 0 ok, else the reference's failure point (1..4)
  int ns_fail;    // the latest sign-iteration projection failed its guard
  int ns_fallbacks;  // projections redone by FP64 Jacobi after a failed guard
  int ne;         // cptr[m]: pair entries of all cuts (= aptr[nact])
  int xrows;      // the sliced Gram rows are built
};

// threads of the one-pass cut solve (one workgroup per node)
constexpr int XCUT_THREADS = 1024;

// Launch helpers implemented in madam_device.hip.  All take the device array
// of nodes and the batch size, and run on `stream`.
void launch_admm_prep(NodeView* nodes, int nb, int mmax, hipStream_t stream);
// hv: host copies of the views (per-node choice of the cut-solve kernel);
// use_x: the one-pass kernel may serve nodes with expanded rows
void launch_cg(NodeView* nodes, const NodeView* hv, int nb, double rtol, double adapt, int maxit,
               size_t lds_limit, int use_x, hipStream_t stream);
// The same in two parts: which cut-solve kernels a batch needs and the LDS
// each asks for (from the host views), then their launches.  A plan fixes
// every launch argument, so equal plans issue identical launches.
struct CgPlan {
  size_t need[6];
  bool any[6];
};
CgPlan plan_cg(const NodeView* hv, int nb, size_t lds_limit, int use_x);
void launch_cg_plan(NodeView* nodes, int nb, const CgPlan& plan, double rtol, double adapt,
                    int maxit, size_t lds_limit, int use_x, hipStream_t stream);
void launch_build_w(NodeView* nodes, int nb, int nmax, int mmax, int use_s, hipStream_t stream);
// n_limit: nodes of that order or more are skipped (the block Jacobi
// solver's)
void launch_jacobi(NodeView* nodes, int nb, int nmax, int values_only, int max_sweeps,
                   int ignore_done, size_t lds_limit, hipStream_t stream, int form_xplus,
                   int n_limit = 1 << 30);
size_t jacobi_lds_bytes(int nmax, bool with_v);
// part: nb x 64 x 8 doubles of scratch for the per-workgroup partial sums
void launch_admm_post(NodeView* nodes, int nb, int nmax, int mmax, double* part,
                      hipStream_t stream);
void launch_safe_bound(NodeView* nodes, int nb, hipStream_t stream);
void launch_op_B(const NodeView* nodes, int nb, const double* Xin, double* yout, int mmax,
                 hipStream_t stream);
void launch_op_Bt(const NodeView* nodes, int nb, const double* tin, double* Xout, int nmax,
                  hipStream_t stream);
void launch_project(NodeView* nodes, int nb, int nmax, hipStream_t stream);
void launch_ipm(NodeView* nodes, int nb, int nmax, hipStream_t stream);
size_t ipm_lds_bytes(int n);
void launch_ns_project(NodeView* nodes, int nb, int nmax, int ignore_done, size_t lds_limit,
                       hipStream_t stream);
size_t ns_lds_bytes(int nmax);
// blocks2: the 2 x 2-tile product kernel (same bits)
void launch_ns_tiles(NodeView* nodes, int nb, int nmax, hipStream_t stream, int blocks2);
void launch_ns_synth(NodeView* nodes, int nb, int nmax, hipStream_t stream);
void launch_zfac_chol(NodeView* nodes, int nb, hipStream_t stream);
size_t safe_hh_lds_bytes(int nmax);
void launch_safe_hh(NodeView* nodes, int nb, int nmax, hipStream_t stream);
void launch_mfma_test(const float* A32, const float* B32, float* C32, int np, const double* A64,
                      const double* B64, double* C64, int n, double alpha, double beta,
                      hipStream_t stream);

// One piece of a batched host-device transfer: `bytes` from src to dst,
// both device addresses (src null: fill dst with zeros).
struct XDesc {
  const char* src;
  char* dst;
  unsigned long long bytes;
};
// Every piece of d[0..nd) (a device array), one workgroup per piece.
void launch_xfer(const XDesc* d, int nd, hipStream_t stream);
// The pair map of a cut structure built on the device: p2a[0 .. nn) = -1,
// then p2a[apair[a]] = a for a < nact (each active pair's lower entry ->
// its index), one workgroup per job.
struct P2aJob {
  int* p2a;
  const int* apair;
  int nn, nact;
};
void launch_p2a(const P2aJob* jobs, int njobs, hipStream_t stream);

}  // namespace madam
