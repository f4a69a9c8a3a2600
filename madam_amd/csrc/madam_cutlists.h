// madam_cutlists.h -- host-side cut list rules of the separation rounds:
// the top-K list of simulated-annealing results and the purge / duplicate
// check / append of a node's cut set.  Plain C++ (no device code), shared
// by the bounding driver (madam_bab.hip) and its CPU test harness.
#pragma once

#include <stdint.h>

#include <unordered_map>
#include <vector>

namespace madam {

// Top-K list semantics of getViolated_Pentagonal/HeptagonalInequalities
// (cutting_planes.c:247-303, 444-501) applied to SA results in run order.
template <class Cut, int KK>
void kgon_list(const std::vector<double>& cost, const std::vector<int>& perms,
               const std::vector<int>& types, int K, std::vector<Cut>& list, double& best) {
  list.clear();
  int slot = 0;
  double slotval = -1e9;
  best = 1e9;
  for (size_t r = 0; r < cost.size(); ++r) {
    const double val = cost[r];
    if (val < best) best = val;
    if (!(val < 1.0)) continue;
    Cut cut{};
    cut.type = types[r];
    cut.value = val;
    for (int a = 0; a < KK; ++a) cut.perm[a] = perms[r * KK + a];
    if ((int)list.size() < K) {
      if (list.empty() || val > slotval) {
        slotval = val;
        slot = (int)list.size();
      }
      list.push_back(cut);
    } else if (val < slotval) {
      list[slot] = cut;
      slotval = 1e9;
      for (int q = 0; q < (int)list.size(); ++q)
        if (list[q].value > slotval) {
          slotval = list[q].value;
          slot = q;
        }
    }
  }
}

// FNV-1a over a pentagonal/heptagonal cut's type and vertex tuple.
template <class Cut>
uint64_t kgon_hash(const Cut& c) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](uint32_t v) {
    h ^= v;
    h *= 1099511628211ull;
  };
  mix((uint32_t)c.type);
  for (size_t a = 0; a < sizeof(c.perm) / sizeof(c.perm[0]); ++a) mix((uint32_t)c.perm[a]);
  return h;
}

// Purge (dual < 1e-5) and append non-duplicates (cutting_planes.c:132-214,
// 316-409, 512-605).  Duplicates are checked against the survivors only.
template <class Cut, class Same>
void purge_append(std::vector<Cut>& cuts, const double* dual, const std::vector<Cut>& list,
                  Same same) {
  std::vector<Cut> keep;
  keep.reserve(cuts.size() + list.size());
  for (size_t c = 0; c < cuts.size(); ++c) {
    Cut x = cuts[c];
    x.y = dual[c];
    if (x.y < 1e-5) continue;
    keep.push_back(x);
  }
  const size_t survivors = keep.size();
  // survivors indexed by a hash of (type, perm); equal hashes are compared
  // in full, so the outcome is the reference's linear scan
  std::unordered_multimap<uint64_t, size_t> index;
  index.reserve(survivors * 2);
  for (size_t c = 0; c < survivors; ++c) index.emplace(kgon_hash(keep[c]), c);
  for (const Cut& cand : list) {
    if (keep.size() == 50000) break;
    bool dup = false;
    const auto range = index.equal_range(kgon_hash(cand));
    for (auto it = range.first; it != range.second && !dup; ++it) dup = same(keep[it->second], cand);
    if (dup) continue;
    Cut x = cand;
    x.y = 0.0;
    keep.push_back(x);
  }
  cuts.swap(keep);
}

}  // namespace madam
