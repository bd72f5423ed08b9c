// Host helpers shared by the eval-proofs (prove.cpp: LigeSIS, basefold.cpp: standalone Basefold):
// small eq tables, rs_merkle multiproof node lists, Basefold query positions, the staged gather
// table of proof openings, and the canonical proof writer.
#pragma once
#include <algorithm>
#include <cstring>
#include <vector>

#include "prove_kernels.hpp"
#include "runtime.hpp"

namespace lgs {
namespace pu {

typedef std::vector<ef2> EfV;

inline EfV rev(const EfV& v) { return EfV(v.rbegin(), v.rend()); }
inline EfV cat(const EfV& a, const EfV& b) {
  EfV r = a;
  r.insert(r.end(), b.begin(), b.end());
  return r;
}

// new_eq on the host for the short points of the eq-compression (<= 64 entries).
inline EfV host_eq(const EfV& p) {
  EfV e{ef_one()};
  for (size_t k = p.size(); k-- > 0;) {
    EfV n;
    ef2 q = ef_sub(ef_one(), p[k]);
    for (auto& x : e) {
      n.push_back(ef_mul(q, x));
      n.push_back(ef_mul(p[k], x));
    }
    e.swap(n);
  }
  return e;
}

inline int ilog2(uint64_t x) {
  int l = 0;
  while ((uint64_t(1) << l) < x) l++;
  return l;
}

// Nodes of an rs_merkle multiproof (utils/src/merkle.rs:76-78) for sorted distinct leaf
// positions, as node indices of the layered tree storage, in proof order.
inline std::vector<uint64_t> multiproof_nodes(uint64_t leaves, const std::vector<uint64_t>& pos) {
  std::vector<uint64_t> out, cur = pos;
  uint64_t off = 0, size = leaves;
  while (true) {
    for (size_t a = 0; a < cur.size(); a++) {
      uint64_t sib = cur[a] ^ 1;
      bool known = (a + 1 < cur.size() && cur[a + 1] == sib) || (a > 0 && cur[a - 1] == sib);
      if (!known && sib < size) out.push_back(off + sib);
    }
    if (size <= 1) break;
    std::vector<uint64_t> par;
    for (uint64_t i : cur)
      if (par.empty() || par.back() != i / 2) par.push_back(i / 2);
    cur = par;
    off += size;
    size >>= 1;
  }
  return out;
}

inline std::vector<uint64_t> level_positions(const std::vector<uint64_t>& base, uint64_t half) {
  std::vector<uint64_t> p;
  for (uint64_t b : base) {
    p.push_back(b % half);
    p.push_back(b % half + half);
  }
  std::sort(p.begin(), p.end());
  p.erase(std::unique(p.begin(), p.end()), p.end());
  return p;
}

static_assert(sizeof(GatherPart) % 8 == 0, "gather table is uploaded as words");

// Staged gathers from several device sources into one staging buffer, described by positions
// (GatherPart): a part is a list of positions, each giving `rows` runs of `width` words.
struct Gather {
  std::vector<GatherPart> parts;
  std::vector<uint32_t> pos;
  uint64_t total = 0;
  // returns the part's first word in the staging buffer
  template <class V>
  uint64_t add(const uint64_t* src, const V& p, uint64_t pmul, uint32_t width, uint32_t rows = 1,
               uint64_t rstride = 0) {
    const uint64_t at = total;
    if (p.empty()) return at;
    parts.push_back(GatherPart{src, at, pmul, rstride, (uint32_t)pos.size(), rows, width, 0});
    for (auto v : p) {
      if ((uint64_t)v > UINT32_MAX) throw Error(1, "prove: opening position out of range");
      pos.push_back((uint32_t)v);
    }
    total += (uint64_t)p.size() * rows * width;
    return at;
  }
  // 4-word digests of tree nodes (node indices of the layered storage)
  uint64_t add_nodes(const DevTree& t, const std::vector<uint64_t>& nodes) {
    return add((const uint64_t*)t.nodes.ptr, nodes, 4, 4);
  }
  uint64_t add_root(const DevTree& t) { return add_nodes(t, {t.node_count() - 1}); }
  uint64_t add_range(const uint64_t* src, uint64_t words) {  // words consecutive words of src
    return add(src, std::vector<uint32_t>{0}, 0, (uint32_t)words);
  }
};

static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__, "proof encoding copies host words as little-endian");

// Appends little-endian words to a ProofOut; past a fixed buffer's end it only counts, so the
// final n is the proof's size either way.
struct Writer {
  ProofOut o;
  size_t n = 0;
  explicit Writer(ProofOut out) : o(out) {}
  // the host is little-endian (x86-64), so words are copied as they are
  void u64(uint64_t v) { words(&v, 1); }
  void words(const uint64_t* w, size_t k) {
    const size_t need = n + 8 * k;
    if (o.vec) {
      if (need > o.vec->size()) o.vec->resize(std::max(need, 2 * o.vec->size()));
      std::memcpy(o.vec->data() + n, w, 8 * k);
    } else if (need <= o.cap) {
      std::memcpy(o.dst + n, w, 8 * k);
    }
    n = need;
  }
  void raw(const uint64_t* w, size_t n) {  // n words as raw little-endian bytes (digests)
    words(w, n);
  }
  void ef(const uint64_t* w) { words(w, 2); }
};

}  // namespace pu
}  // namespace lgs
