// BLAKE3 compression for device code, written from the public BLAKE3 specification (7 rounds
// of the G mixing function over a 16-word state; message word permutation
// 2,6,3,10,7,0,4,13,1,11,12,5,9,14,15,8 between rounds; IV = the SHA-256 initial words). Every
// message hashed on this path is at most 1024 bytes, i.e. one chunk: its blocks are compressed in
// order with chunk counter 0, CHUNK_START on the first block and CHUNK_END | ROOT on the last,
// which yields blake3(msg) as called by utils/src/merkle.rs:10-12.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace lgs {

constexpr uint32_t B3_CHUNK_START = 1, B3_CHUNK_END = 2, B3_ROOT = 8;

__device__ __forceinline__ uint32_t b3_rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

// ---- four lanes per compression ----
// The lanes of a quad (lanes 4k .. 4k + 3 of a wave, c = lane & 3) share one compression: lane c
// holds column c of the state (s[c], s[4 + c], s[8 + c], s[12 + c]) and runs the column mixing
// G(c) itself; for the diagonal mixings the rows rotate across the quad (DPP quad permutations),
// lane c runs G on (s[c], s[4 + (c+1)%4], s[8 + (c+2)%4], s[12 + (c+3)%4]), and the rows rotate
// back. Every lane holds the whole message block and picks its words by its column.

// word i of the message after r permutations is word b3_sched(r, i) of the block
__host__ __device__ constexpr int b3_sched(int r, int i) {
  constexpr int perm[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
  for (int k = 0; k < r; k++) i = perm[i];
  return i;
}

__device__ __forceinline__ uint32_t b3_sel4(uint32_t c, uint32_t v0, uint32_t v1, uint32_t v2, uint32_t v3) {
  return c < 2 ? (c == 0 ? v0 : v1) : (c == 2 ? v2 : v3);
}

// quad permutation: lane i of the quad receives lane p_i (ctrl = p0 | p1 << 2 | p2 << 4 | p3 << 6)
template <int CTRL>
__device__ __forceinline__ uint32_t b3_quad(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}

#define LGS_B3_G4(x, y)           \
  do {                            \
    a = a + b + (x);              \
    d = b3_rotr(d ^ a, 16);       \
    cc = cc + d;                  \
    b = b3_rotr(b ^ cc, 12);      \
    a = a + b + (y);              \
    d = b3_rotr(d ^ a, 8);        \
    cc = cc + d;                  \
    b = b3_rotr(b ^ cc, 7);       \
  } while (0)

// Lane c's share of cv <- compress(cv, m, counter 0, block_len, flags): h0 = cv[c], h1 = cv[4 + c].
// All four lanes of the quad must be active.
__device__ __forceinline__ void b3_compress4(uint32_t& h0, uint32_t& h1, const uint32_t m[16], uint32_t block_len,
                                             uint32_t flags, uint32_t c) {
  uint32_t a = h0, b = h1;
  uint32_t cc = b3_sel4(c, 0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au);
  uint32_t d = b3_sel4(c, 0u, 0u, block_len, flags);
#pragma unroll
  for (int r = 0; r < 7; r++) {
    LGS_B3_G4(b3_sel4(c, m[b3_sched(r, 0)], m[b3_sched(r, 2)], m[b3_sched(r, 4)], m[b3_sched(r, 6)]),
              b3_sel4(c, m[b3_sched(r, 1)], m[b3_sched(r, 3)], m[b3_sched(r, 5)], m[b3_sched(r, 7)]));
    b = b3_quad<0x39>(b);   // [1, 2, 3, 0]
    cc = b3_quad<0x4E>(cc); // [2, 3, 0, 1]
    d = b3_quad<0x93>(d);   // [3, 0, 1, 2]
    LGS_B3_G4(b3_sel4(c, m[b3_sched(r, 8)], m[b3_sched(r, 10)], m[b3_sched(r, 12)], m[b3_sched(r, 14)]),
              b3_sel4(c, m[b3_sched(r, 9)], m[b3_sched(r, 11)], m[b3_sched(r, 13)], m[b3_sched(r, 15)]));
    b = b3_quad<0x93>(b);
    cc = b3_quad<0x4E>(cc);
    d = b3_quad<0x39>(d);
  }
  h0 = a ^ cc;
  h1 = b ^ d;
}

#undef LGS_B3_G4

__device__ __forceinline__ void b3_init4(uint32_t& h0, uint32_t& h1, uint32_t c) {
  h0 = b3_sel4(c, 0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au);
  h1 = b3_sel4(c, 0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u);
}

}  // namespace lgs
