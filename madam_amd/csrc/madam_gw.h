// madam_gw.h -- per-node data of the GW rounding kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "madam_internal.h"

namespace madam {

struct GwNode {
  int num;                 // hyperplanes (SP->n, heuristic.c:17)
  int M;                   // set by k_gw_prep
  int zero_planes;         // 1: use no eigenvectors (all-ones start), see bounding.c:98
  int pad;
  unsigned long long seed;
  long long stream;        // distinct per (node, call)
  double fixedvalue;       // getFixedValue of the node
  double* val;             // num
  signed char* zall;       // num * n
  signed char* best_z;     // n
  double* best_val;        // 1
};

void launch_gw(NodeView* nodes, GwNode* gw, int nb, int nmax, int nummax, hipStream_t st);
size_t gw_lds_bytes(int nmax);
// same, with Zc (n x M) already in S1 and GwNode.M set by the caller
void launch_gw_given_z(NodeView* nodes, GwNode* gw, int nb, int nmax, int nummax, hipStream_t st);

}  // namespace madam
