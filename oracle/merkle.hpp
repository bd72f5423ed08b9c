// Oracle (test infrastructure only): BLAKE3 Merkle tree with rs_merkle semantics.
//
// Restates utils/src/merkle.rs:5-116 and the crate rs_merkle 1.5.0 (reference
// Cargo.lock:559-565) that it calls:
//   - leaf hash = blake3(leaf bytes)                         (merkle.rs:49-55)
//   - parent    = blake3(left || right); an unpaired last node is promoted unchanged
//   - multiproof (`proof(indices).to_bytes()`, merkle.rs:76-78): walking up from the leaves,
//     for each known index (sorted) its sibling index^1, if not itself known and present in
//     the layer, contributes its 32-byte hash, in that order; the known set then becomes the
//     deduplicated parents i/2.
//   - field serialization = canonical little-endian u64 per base element, EF as (c0, c1)
//     (merkle.rs:24-35; pinned by the reference test merkle.rs:146-156).
// The multiproof layout is restated from rs_merkle's documented algorithm; no reference test
// pins its bytes (parity unpinned against the real crate, pinned here by prove/verify round
// trips and tamper tests).
#pragma once
#include <algorithm>
#include <map>
#include <stdexcept>
#include <vector>

#include "blake3.hpp"
#include "field.hpp"

namespace lgso {

inline void put_u64(std::vector<uint8_t>& out, uint64_t v) {
  for (int b = 0; b < 8; b++) out.push_back((uint8_t)(v >> (8 * b)));
}
inline void ser_field(std::vector<uint8_t>& out, uint64_t v) { put_u64(out, canon(v)); }
inline void ser_field(std::vector<uint8_t>& out, const Ef& v) {
  put_u64(out, canon(v.c0));
  put_u64(out, canon(v.c1));
}
template <class T>
inline std::vector<uint8_t> serialize_fields(const std::vector<T>& v) {
  std::vector<uint8_t> out;
  for (auto& x : v) ser_field(out, x);
  return out;
}

inline H32 hash_pair(const H32& l, const H32& r) {
  uint8_t buf[64];
  std::memcpy(buf, l.data(), 32);
  std::memcpy(buf + 32, r.data(), 32);
  return blake3(buf, 64);
}

struct MerkleTree {
  std::vector<std::vector<H32>> layers;  // layers[0] = leaf hashes, back() = {root}

  static MerkleTree from_leaf_hashes(std::vector<H32> leaves) {
    MerkleTree t;
    t.layers.push_back(std::move(leaves));
    while (t.layers.back().size() > 1) {
      const auto& cur = t.layers.back();
      std::vector<H32> nxt;
      for (size_t i = 0; i < cur.size(); i += 2)
        nxt.push_back(i + 1 < cur.size() ? hash_pair(cur[i], cur[i + 1]) : cur[i]);
      t.layers.push_back(std::move(nxt));
    }
    return t;
  }
  static MerkleTree from_leaves(const std::vector<std::vector<uint8_t>>& leaves) {
    std::vector<H32> h;
    h.reserve(leaves.size());
    for (auto& l : leaves) h.push_back(blake3(l));
    return from_leaf_hashes(std::move(h));
  }
  H32 root() const { return layers.back()[0]; }

  std::vector<uint8_t> open(const std::vector<size_t>& leaf_indices) const {
    std::vector<size_t> cur = leaf_indices;
    std::vector<uint8_t> out;
    for (const auto& layer : layers) {
      for (size_t i : cur) {
        size_t sib = i ^ 1;
        if (std::find(cur.begin(), cur.end(), sib) != cur.end()) continue;
        if (sib < layer.size()) out.insert(out.end(), layer[sib].begin(), layer[sib].end());
      }
      std::vector<size_t> par;
      for (size_t i : cur) {
        size_t p = i / 2;
        if (par.empty() || par.back() != p) par.push_back(p);
      }
      cur = par;
    }
    return out;
  }
};

// Recompute the root from a multiproof for sorted distinct `indices` of a tree with
// `n_leaves` leaves; false when the proof is malformed or the root differs.
inline bool merkle_verify(const H32& root, size_t n_leaves, const std::vector<uint8_t>& proof,
                          const std::vector<size_t>& indices, const std::vector<H32>& leaf_hashes) {
  if (proof.size() % 32 != 0 || indices.size() != leaf_hashes.size() || indices.empty())
    return false;
  std::vector<std::pair<size_t, H32>> known;
  for (size_t k = 0; k < indices.size(); k++) {
    if (indices[k] >= n_leaves) return false;
    known.push_back({indices[k], leaf_hashes[k]});
  }
  std::sort(known.begin(), known.end(),
            [](const std::pair<size_t, H32>& a, const std::pair<size_t, H32>& b) {
              return a.first < b.first;
            });
  size_t cursor = 0, layer_size = n_leaves;
  while (layer_size > 1) {
    std::map<size_t, H32> have;
    for (auto& kv : known) {
      auto it = have.find(kv.first);
      if (it != have.end() && it->second != kv.second) return false;
      have[kv.first] = kv.second;
    }
    std::vector<size_t> cur;
    for (auto& kv : have) cur.push_back(kv.first);
    for (size_t i : cur) {
      size_t sib = i ^ 1;
      if (have.count(sib) || sib >= layer_size) continue;
      if (cursor + 32 > proof.size()) return false;
      H32 h;
      std::memcpy(h.data(), proof.data() + cursor, 32);
      cursor += 32;
      have[sib] = h;
    }
    std::vector<std::pair<size_t, H32>> nxt;
    for (size_t i : cur) {
      size_t l = i & ~size_t(1), r = l + 1;
      size_t p = i / 2;
      if (!nxt.empty() && nxt.back().first == p) continue;
      H32 h = r < layer_size ? hash_pair(have.at(l), have.at(r)) : have.at(l);
      nxt.push_back({p, h});
    }
    known = nxt;
    layer_size = (layer_size + 1) / 2;
  }
  return known.size() == 1 && known[0].second == root;
}

}  // namespace lgso
