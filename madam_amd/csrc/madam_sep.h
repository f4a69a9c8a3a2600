// madam_sep.h -- device structures of the separation and heuristic kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace madam {

constexpr int SEP_NBIN = 4096;   // histogram bins over [thresh, max]
constexpr int SEP_BCAP = 16384;  // boundary-bin capacity per node
constexpr int SEP_RLEVELS = 8;   // 12-bit digits of the 96-bit refinement key
// Control words of one node's separation.  CTL_OVER: 1 when the boundary
// bin held more than SEP_BCAP rows (the exact refinement then runs), 2 once
// the refinement has resolved it.  CTL_RMODE/RL/RNEED/RD0..: refinement
// state (mode, digits fixed, rows still needed, the fixed digits).
enum {
  CTL_COUNT = 0, CTL_NOUT, CTL_NBOUND, CTL_BSTAR, CTL_NEED, CTL_NFINAL, CTL_OVER, CTL_RMODE,
  CTL_RL, CTL_RNEED, CTL_RD0, CTL_N = CTL_RD0 + SEP_RLEVELS
};

// Per-node triangle separation state (all device pointers).
struct SepView {
  int n, K;
  double thresh;
  const double* X;              // n x n, symmetric
  uint32_t* ctl;                // CTL_N counters (zeroed before a pass)
  unsigned long long* maxbits;  // ordered bits of the max row value
  uint32_t* hist;               // SEP_NBIN
  uint32_t* okey;               // K
  uint32_t* bkey;               // SEP_BCAP
  double* bval;                 // SEP_BCAP
  uint32_t* fkey;               // K, final keys in enumeration order
  double* fval;                 // K, their values
};

void launch_tri_separate(SepView* sv, int nb, int nmax, hipStream_t st);
// the exact refinement for nodes whose boundary bin overflowed (CTL_OVER
// = 1 after launch_tri_separate; a no-op for the others), then the list
void launch_tri_refine(SepView* sv, int nb, int nmax, hipStream_t st);

// One simulated-annealing run (qap_simulated_annealing.c).
constexpr int SA_WG = 64;
inline int sa_nstride(int n) { return (n + 1) & ~1; }
struct SaJob {
  const double* X;
  const double* sumX;    // per-node precomputed sum of diag + 2 |offdiag|
  int n, k, neg, nstride;
  unsigned long long seed;
  long long stream_id;
};
struct SaOut {
  double cost;
  int perm[7];
  int pad;
  long long draws, moves;
};
// jobs: nnodes groups of `per` runs, the runs of a group on one X
void launch_qap_sa(const SaJob* jobs, int nnodes, int per, int nmax, int kk, SaOut* out,
                   size_t lds_limit, int cu_count, hipStream_t st);
void launch_sumx(const double* const* X, const int* n, double* out, int nb, hipStream_t st);

}  // namespace madam
