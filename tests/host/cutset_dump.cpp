// Host-only harness for the cut-structure build of libmadam_bound
// (mh::build_cutset, the CSR form of op_B / op_Bt's cut rows): reads n and
// the cuts from stdin, prints the structure (or ERR and the message).
// Compiled and run by tests/test_host_cutset.py on the CPU; no device call.
#include <iostream>
#include <vector>

#include "../../madam_amd/csrc/madam_engine.h"

namespace mh {
char g_err[512] = "";
}

template <class T>
static void put(const char* name, const std::vector<T>& v) {
  std::cout << name << ' ' << v.size();
  for (const auto& x : v) std::cout << ' ' << x;
  std::cout << '\n';
}

int main() {
  int n, nt, np, nh;
  if (!(std::cin >> n >> nt >> np >> nh)) return 2;
  std::vector<madam_tri> tri(nt);
  std::vector<madam_pent> pent(np);
  std::vector<madam_hept> hept(nh);
  for (auto& c : tri) {
    std::cin >> c.type >> c.i >> c.j >> c.k;
    c.value = c.y = 0.0;
  }
  for (auto& c : pent) {
    std::cin >> c.type;
    for (int a = 0; a < 5; ++a) std::cin >> c.perm[a];
    c.value = c.y = 0.0;
  }
  for (auto& c : hept) {
    std::cin >> c.type;
    for (int a = 0; a < 7; ++a) std::cin >> c.perm[a];
    c.value = c.y = 0.0;
  }
  std::cout.precision(17);
  std::cout << "valid " << (mh::cuts_valid(n, tri, pent, hept) ? 1 : 0) << '\n';
  mh::CutSet cs;
  const int rc = mh::build_cutset(n, nt, tri.data(), np, pent.data(), nh, hept.data(), cs);
  if (rc != MADAM_OK) {
    std::cout << "ERR " << rc << ' ' << mh::g_err << '\n';
    return 0;
  }
  std::cout << "m " << cs.m << "\nnact " << cs.nact << '\n';
  put("cptr", cs.cptr);
  put("cent", cs.cent);
  put("cact", cs.cact);
  put("aptr", cs.aptr);
  put("aent", cs.aent);
  put("apair", cs.apair);
  put("cscale", cs.cscale);
  put("aent16", cs.aent16);
  put("cact16", cs.cact16);
  return 0;
}
