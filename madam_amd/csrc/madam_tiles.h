// madam_tiles.h -- the blocked Cholesky factorisation spread over many
// workgroups per matrix (madam_ipmt.hip), shared by the interior-point
// method and the safe bound's certificate for orders beyond LDS.  Not part
// of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

namespace madam {

constexpr int IB = 32;  // block width of the factorisation
constexpr int RP = 64;  // panel rows per workgroup of a factorisation step

// One Cholesky factorisation: A (n x n, ld n) factored in place, lower
// triangle (the strict upper triangle is neither read nor written).  D
// (IB x IB doubles) holds the diagonal block factored in the latest step
// until the next step writes it into A.
struct CholJob {
  double* A;
  double* D;
  const int* stop;  // nonzero *stop: the job's node has finished (null: never)
  int n;
  int fail;  // 0, or 1 + the first column with a non-positive pivot
  int skip;  // 1: nothing to factor in this sequence
  int pad;
};

// All ceil(nmax / IB) + 1 launches of a factorisation of the nj jobs
// (orders <= nmax), on stream st.  Reset fail and skip before.
void chol_launches(CholJob* jobs, int nj, int nmax, hipStream_t st);

}  // namespace madam
