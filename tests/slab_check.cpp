// Randomised check of SlabSet (maxlq_amd/csrc/mlq_slab.hpp) against a model
// that tracks every 64 KiB grain of every slab: blocks never overlap and stay
// inside their slab, each take picks the smallest free run that holds the
// block, the free ranges after every step are exactly the maximal runs of free
// grains (returned blocks merge with their neighbours), and with every block
// returned each slab is one free range and is released.  Prints "ok" and
// exits 0, else names the first failed check and exits 1.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <iterator>
#include <map>
#include <random>
#include <vector>

#include "mlq_slab.hpp"

using mlq::SlabSet;

namespace {
constexpr size_t G = SlabSet::GRAIN;
int fails = 0;
void check(bool ok, const char* what, long step) {
  if (!ok && fails++ < 5) std::printf("FAIL step %ld: %s\n", step, what);
}

struct Model {
  std::vector<char*> base;
  std::vector<std::vector<int>> owner;   // per slab, per grain: block id or -1
  // maximal runs of free grains of slab s: start -> length (grains)
  std::map<size_t, size_t> runs(int s) const {
    std::map<size_t, size_t> r;
    const auto& o = owner[s];
    for (size_t i = 0; i < o.size();) {
      if (o[i] >= 0) { ++i; continue; }
      size_t j = i;
      while (j < o.size() && o[j] < 0) ++j;
      r[i] = j - i;
      i = j;
    }
    return r;
  }
};
}  // namespace

int main() {
  std::mt19937_64 rng(12345);
  SlabSet ss;
  Model m;
  std::vector<std::vector<char>> mem;   // backing store of the slabs
  struct Live { char* p; int slab; size_t g0, ng; };
  std::map<int, Live> live;
  int next_id = 0;
  long step = 0;
  const size_t min_slab = 9;   // grains, as the engine's 9 real fields
  for (; step < 20000; ++step) {
    // the live set hovers around 30 blocks: slabs are reused far more often
    // than new ones are added
    const bool do_take = live.empty() || (int)(rng() % 100) < (live.size() < 30 ? 60 : 40);
    if (do_take) {
      const size_t ng = 1 + rng() % 9;
      // smallest free run that holds ng grains, over all slabs
      size_t best = 0;
      for (size_t s = 0; s < m.base.size(); ++s)
        for (const auto& r : m.runs((int)s))
          if (r.second >= ng && (best == 0 || r.second < best)) best = r.second;
      char* p = ss.take(ng * G);
      check((p == nullptr) == (best == 0), "take fails exactly when no free run holds the block", step);
      if (!p) {
        const size_t sz = std::max(ng, min_slab) + rng() % 4;
        mem.emplace_back(sz * G);
        ss.add(mem.back().data(), sz * G);
        m.base.push_back(mem.back().data());
        m.owner.emplace_back(sz, -1);
        best = sz;
        p = ss.take(ng * G);
        check(p != nullptr, "take from a new slab", step);
        if (!p) break;
      }
      int s = -1;
      for (size_t k = 0; k < m.base.size(); ++k)
        if (m.base[k] && p >= m.base[k] && p < m.base[k] + m.owner[k].size() * G) s = (int)k;
      check(s >= 0, "block inside a slab", step);
      if (s < 0) break;
      const size_t off = (size_t)(p - m.base[s]);
      check(off % G == 0, "block on a grain boundary", step);
      const size_t g0 = off / G;
      check(g0 + ng <= m.owner[s].size(), "block ends inside its slab", step);
      // the run the block starts has the best-fit length
      const auto runs = m.runs(s);
      const auto it = runs.find(g0);
      check(it != runs.end() && it->second == best, "best fit: the smallest free run that holds it",
            step);
      for (size_t g = g0; g < g0 + ng && g < m.owner[s].size(); ++g) {
        check(m.owner[s][g] < 0, "no overlap with a live block", step);
        m.owner[s][g] = next_id;
      }
      live[next_id++] = Live{p, s, g0, ng};
    } else {
      auto it = live.begin();
      std::advance(it, (long)(rng() % live.size()));
      const Live b = it->second;
      const size_t len = ss.give(b.p);
      check(len == b.ng * G, "give returns the block length", step);
      for (size_t g = b.g0; g < b.g0 + b.ng; ++g) m.owner[b.slab][g] = -1;
      live.erase(it);
    }
    for (size_t s = 0; s < m.base.size(); ++s) {
      std::map<size_t, size_t> want;
      for (const auto& r : m.runs((int)s)) want[r.first * G] = r.second * G;
      check(ss.free_ranges((int)s) == want, "free ranges are the maximal free runs", step);
    }
    if (fails) break;
  }
  char other[16] = {};
  check(ss.give(other) == 0, "give of a block not carved returns 0", step);
  for (const auto& kv : live) ss.give(kv.second.p);
  size_t total = 0, released = 0;
  for (const auto& o : m.owner) total += o.size() * G;
  for (size_t s = 0; s < ss.slab_count(); ++s)
    check(ss.free_ranges((int)s).size() == 1 && ss.free_ranges((int)s).begin()->second == ss.slab_size((int)s),
          "all returned: one free range per slab", step);
  released = ss.release_free([&](char*, size_t) {});
  check(released == total, "every slab released", step);
  std::printf("%s: %ld steps, %zu slabs\n", fails ? "FAILED" : "ok", step, m.base.size());
  return fails ? 1 : 0;
}
