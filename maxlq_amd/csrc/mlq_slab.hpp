// Carving of device temporaries out of slabs (host bookkeeping only: no HIP
// calls, so that it builds and is tested with the host compiler alone).
//
// A block is taken from the smallest free range of any slab that holds it
// (best fit) and comes back merged with the free ranges either side.  The
// temporaries of one phase of an iteration (9-component gradients,
// 3-component fields, single components, spectra) take other shapes than
// those of the next; carving lets each phase reuse the memory of the last.
#pragma once
#include <cstddef>
#include <iterator>
#include <map>
#include <utility>
#include <vector>

namespace mlq {

class SlabSet {
 public:
  static constexpr size_t GRAIN = 65536;   // blocks start on 64 KiB boundaries
  static size_t round(size_t bytes) { return (bytes + GRAIN - 1) / GRAIN * GRAIN; }

  // best fit over the free ranges of all slabs; nullptr when none is large
  // enough.  bytes: a multiple of GRAIN.
  char* take(size_t bytes) {
    int bs = -1;
    size_t bo = 0, bl = 0;
    for (int s = 0; s < (int)slabs_.size(); ++s)
      for (const auto& r : slabs_[s].fr)
        if (r.second >= bytes && (bs < 0 || r.second < bl)) {
          bs = s;
          bo = r.first;
          bl = r.second;
        }
    if (bs < 0) return nullptr;
    Slab& sl = slabs_[bs];
    sl.fr.erase(bo);
    if (bl > bytes) sl.fr[bo + bytes] = bl - bytes;
    char* p = sl.base + bo;
    carved_[p] = {bs, bytes};
    return p;
  }

  // a new slab, all of it free
  void add(char* base, size_t size) {
    Slab sl;
    sl.base = base;
    sl.size = size;
    sl.fr[0] = size;
    for (auto& s : slabs_)
      if (!s.base) {
        s = std::move(sl);
        return;
      }
    slabs_.push_back(std::move(sl));
  }

  // return a block taken by take(); its length, 0 if p is not a carved block
  size_t give(const void* p) {
    const auto it = carved_.find((const char*)p);
    if (it == carved_.end()) return 0;
    Slab& sl = slabs_[it->second.first];
    const size_t len0 = it->second.second;
    size_t off = (size_t)((const char*)p - sl.base), len = len0;
    carved_.erase(it);
    auto nx = sl.fr.lower_bound(off);
    if (nx != sl.fr.end() && off + len == nx->first) {
      len += nx->second;
      nx = sl.fr.erase(nx);
    }
    if (nx != sl.fr.begin()) {
      const auto pv = std::prev(nx);
      if (pv->first + pv->second == off) {
        pv->second += len;
        return len0;
      }
    }
    sl.fr[off] = len;
    return len0;
  }

  // hand each slab with nothing carved to release(base, size) and forget it;
  // the bytes released
  template <class F> size_t release_free(F&& release) {
    size_t had = 0;
    for (auto& sl : slabs_) {
      if (!sl.base || sl.fr.size() != 1 || sl.fr.begin()->second != sl.size) continue;
      had += sl.size;
      release(sl.base, sl.size);
      sl = Slab{};
    }
    return had;
  }

  // every slab's base (for the owner's teardown)
  template <class F> void for_each_slab(F&& f) const {
    for (const auto& sl : slabs_)
      if (sl.base) f(sl.base, sl.size);
  }

  // the free ranges of slab s (offset -> length), for checks
  size_t slab_count() const { return slabs_.size(); }
  const std::map<size_t, size_t>& free_ranges(int s) const { return slabs_[s].fr; }
  size_t slab_size(int s) const { return slabs_[s].size; }
  char* slab_base(int s) const { return slabs_[s].base; }

 private:
  struct Slab {
    char* base = nullptr;   // nullptr: released
    size_t size = 0;
    std::map<size_t, size_t> fr;   // free ranges: offset -> length
  };
  std::vector<Slab> slabs_;
  std::map<const char*, std::pair<int, size_t>> carved_;   // block -> slab, length
};

}  // namespace mlq
