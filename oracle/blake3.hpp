// Oracle (test infrastructure only): BLAKE3 hash (32-byte output, unkeyed), written from the
// public BLAKE3 specification (O'Connor, Aumasson, Neves, Wilcox-O'Hearn, 2020): 7-round
// compression on a 16-word state, 1024-byte chunks of 64-byte blocks, CHUNK_START / CHUNK_END /
// PARENT / ROOT flags, left-balanced binary chunk tree.
//
// Restates crate blake3 1.8.5 (reference Cargo.lock:36-46) as called by
// utils/src/merkle.rs:10-12 (`blake3::hash(data)`). Pinned by the published digests of ""
// and "abc" (tests/test_oracle_blake3.py). Multi-block and multi-chunk inputs follow the
// specification; no reference test pins them.
#pragma once
#include <array>
#include <cstdint>
#include <cstring>
#include <vector>

namespace lgso {

typedef std::array<uint8_t, 32> H32;

namespace b3 {

constexpr uint32_t IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                            0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
constexpr int PERM[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
constexpr uint32_t CHUNK_START = 1, CHUNK_END = 2, PARENT = 4, ROOT = 8;
constexpr size_t BLOCK_LEN = 64, CHUNK_LEN = 1024;

inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

inline void g(uint32_t* s, int a, int b, int c, int d, uint32_t mx, uint32_t my) {
  s[a] = s[a] + s[b] + mx;
  s[d] = rotr(s[d] ^ s[a], 16);
  s[c] = s[c] + s[d];
  s[b] = rotr(s[b] ^ s[c], 12);
  s[a] = s[a] + s[b] + my;
  s[d] = rotr(s[d] ^ s[a], 8);
  s[c] = s[c] + s[d];
  s[b] = rotr(s[b] ^ s[c], 7);
}

inline void round_fn(uint32_t* s, const uint32_t* m) {
  g(s, 0, 4, 8, 12, m[0], m[1]);
  g(s, 1, 5, 9, 13, m[2], m[3]);
  g(s, 2, 6, 10, 14, m[4], m[5]);
  g(s, 3, 7, 11, 15, m[6], m[7]);
  g(s, 0, 5, 10, 15, m[8], m[9]);
  g(s, 1, 6, 11, 12, m[10], m[11]);
  g(s, 2, 7, 8, 13, m[12], m[13]);
  g(s, 3, 4, 9, 14, m[14], m[15]);
}

// Returns the full 16-word output; the first 8 words are the chaining value.
inline void compress(const uint32_t cv[8], const uint32_t block[16], uint64_t counter,
                     uint32_t block_len, uint32_t flags, uint32_t out[16]) {
  uint32_t s[16] = {cv[0], cv[1], cv[2], cv[3], cv[4], cv[5], cv[6], cv[7],
                    IV[0], IV[1], IV[2], IV[3], (uint32_t)counter, (uint32_t)(counter >> 32),
                    block_len, flags};
  uint32_t m[16];
  std::memcpy(m, block, 64);
  for (int r = 0; r < 7; r++) {
    round_fn(s, m);
    if (r < 6) {
      uint32_t t[16];
      for (int i = 0; i < 16; i++) t[i] = m[PERM[i]];
      std::memcpy(m, t, 64);
    }
  }
  for (int i = 0; i < 8; i++) {
    out[i] = s[i] ^ s[i + 8];
    out[i + 8] = s[i + 8] ^ cv[i];
  }
}

inline void load_block(const uint8_t* p, size_t len, uint32_t w[16]) {
  uint8_t buf[64] = {0};
  std::memcpy(buf, p, len);
  for (int i = 0; i < 16; i++)
    w[i] = (uint32_t)buf[4 * i] | ((uint32_t)buf[4 * i + 1] << 8) |
           ((uint32_t)buf[4 * i + 2] << 16) | ((uint32_t)buf[4 * i + 3] << 24);
}

// A pending output: the inputs of the last compression of a node, so the caller can add ROOT.
struct Output {
  uint32_t cv[8];
  uint32_t block[16];
  uint64_t counter;
  uint32_t block_len, flags;
  void chaining(uint32_t out8[8]) const {
    uint32_t o[16];
    compress(cv, block, counter, block_len, flags, o);
    std::memcpy(out8, o, 32);
  }
  H32 root_hash() const {
    uint32_t o[16];
    compress(cv, block, 0, block_len, flags | ROOT, o);
    H32 h;
    for (int i = 0; i < 8; i++)
      for (int b = 0; b < 4; b++) h[4 * i + b] = (uint8_t)(o[i] >> (8 * b));
    return h;
  }
};

inline Output chunk_output(const uint8_t* data, size_t len, uint64_t chunk_counter) {
  uint32_t cv[8];
  std::memcpy(cv, IV, 32);
  size_t nblocks = len == 0 ? 1 : (len + BLOCK_LEN - 1) / BLOCK_LEN;
  for (size_t b = 0; b + 1 < nblocks; b++) {
    uint32_t w[16], o[16];
    load_block(data + b * BLOCK_LEN, BLOCK_LEN, w);
    compress(cv, w, chunk_counter, BLOCK_LEN, b == 0 ? CHUNK_START : 0, o);
    std::memcpy(cv, o, 32);
  }
  Output out;
  std::memcpy(out.cv, cv, 32);
  size_t last = (nblocks - 1) * BLOCK_LEN;
  size_t last_len = len - last;
  load_block(data + last, last_len, out.block);
  out.counter = chunk_counter;
  out.block_len = (uint32_t)last_len;
  out.flags = CHUNK_END | (nblocks == 1 ? CHUNK_START : 0);
  return out;
}

inline Output parent_output(const uint32_t l[8], const uint32_t r[8]) {
  Output out;
  std::memcpy(out.cv, IV, 32);
  std::memcpy(out.block, l, 32);
  std::memcpy(out.block + 8, r, 32);
  out.counter = 0;
  out.block_len = BLOCK_LEN;
  out.flags = PARENT;
  return out;
}

// Output of the subtree covering `len` bytes starting at chunk index `chunk0`.
inline Output subtree(const uint8_t* data, size_t len, uint64_t chunk0) {
  if (len <= CHUNK_LEN) return chunk_output(data, len, chunk0);
  size_t nchunks = (len + CHUNK_LEN - 1) / CHUNK_LEN;
  size_t left_chunks = 1;  // largest power of two strictly below nchunks
  while (left_chunks * 2 < nchunks) left_chunks *= 2;
  size_t left_len = left_chunks * CHUNK_LEN;
  uint32_t lcv[8], rcv[8];
  subtree(data, left_len, chunk0).chaining(lcv);
  subtree(data + left_len, len - left_len, chunk0 + left_chunks).chaining(rcv);
  return parent_output(lcv, rcv);
}

}  // namespace b3

inline H32 blake3(const uint8_t* data, size_t len) { return b3::subtree(data, len, 0).root_hash(); }
inline H32 blake3(const std::vector<uint8_t>& v) { return blake3(v.data(), v.size()); }

}  // namespace lgso
