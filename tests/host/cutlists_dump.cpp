// Host-only harness for the cut-list rules of the separation rounds
// (madam_cutlists.h): the top-K list of simulated-annealing results and the
// purge / duplicate check / append of a cut set.  Reads a case from stdin,
// prints the result.  Compiled and run by tests/test_host_cutlists.py on
// the CPU.
#include <iostream>
#include <vector>

#include "../../include/madam_bound.h"
#include "../../madam_amd/csrc/madam_cutlists.h"

template <class Cut, int KK>
static int run_list() {
  int K, R;
  std::cin >> K >> R;
  std::vector<double> cost(R);
  std::vector<int> types(R), perms((size_t)R * KK);
  for (int r = 0; r < R; ++r) {
    std::cin >> cost[r] >> types[r];
    for (int a = 0; a < KK; ++a) std::cin >> perms[(size_t)r * KK + a];
  }
  std::vector<Cut> list;
  double best = 0;
  madam::kgon_list<Cut, KK>(cost, perms, types, K, list, best);
  std::cout << "best " << best << "\n" << list.size() << "\n";
  for (const Cut& c : list) {
    std::cout << c.type << ' ' << c.value;
    for (int a = 0; a < KK; ++a) std::cout << ' ' << c.perm[a];
    std::cout << '\n';
  }
  return 0;
}

template <class Cut, int KK>
static int run_purge() {
  int M, L;
  std::cin >> M;
  std::vector<Cut> cuts(M);
  std::vector<double> dual(M);
  for (int c = 0; c < M; ++c) {
    std::cin >> cuts[c].type >> dual[c];
    for (int a = 0; a < KK; ++a) std::cin >> cuts[c].perm[a];
    cuts[c].y = 0.0;
  }
  std::cin >> L;
  std::vector<Cut> list(L);
  for (int c = 0; c < L; ++c) {
    std::cin >> list[c].type;
    for (int a = 0; a < KK; ++a) std::cin >> list[c].perm[a];
    list[c].y = 0.0;
  }
  madam::purge_append(cuts, dual.data(), list, [](const Cut& a, const Cut& b) {
    if (a.type != b.type) return false;
    for (int x = 0; x < KK; ++x)
      if (a.perm[x] != b.perm[x]) return false;
    return true;
  });
  std::cout << cuts.size() << "\n";
  for (const Cut& c : cuts) {
    std::cout << c.type << ' ' << c.y;
    for (int a = 0; a < KK; ++a) std::cout << ' ' << c.perm[a];
    std::cout << '\n';
  }
  return 0;
}

int main() {
  std::string mode;
  int k;
  if (!(std::cin >> mode >> k)) return 2;
  std::cout.precision(17);
  if (mode == "list") return k == 5 ? run_list<madam_pent, 5>() : run_list<madam_hept, 7>();
  if (mode == "purge") return k == 5 ? run_purge<madam_pent, 5>() : run_purge<madam_hept, 7>();
  return 2;
}
