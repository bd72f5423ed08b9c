// BLAKE3 compression for device code, written from the public BLAKE3 specification
// (7 rounds of the G mixing function over a 16-word state; message word permutation
// 2,6,3,10,7,0,4,13,1,11,12,5,9,14,15,8 between rounds). Every message hashed on this path is
// at most 1024 bytes, i.e. one chunk: its blocks are compressed in order with chunk counter 0,
// CHUNK_START on the first block and CHUNK_END | ROOT on the last, which yields blake3(msg)
// as called by utils/src/merkle.rs:10-12.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace lgs {

constexpr uint32_t B3_CHUNK_START = 1, B3_CHUNK_END = 2, B3_ROOT = 8;

__host__ __device__ __forceinline__ uint32_t b3_iv(int i) {
  switch (i) {
    case 0: return 0x6A09E667u;
    case 1: return 0xBB67AE85u;
    case 2: return 0x3C6EF372u;
    case 3: return 0xA54FF53Au;
    case 4: return 0x510E527Fu;
    case 5: return 0x9B05688Cu;
    case 6: return 0x1F83D9ABu;
    default: return 0x5BE0CD19u;
  }
}

__device__ __forceinline__ uint32_t b3_rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

#define LGS_B3_G(a, b, c, d, x, y)   \
  do {                               \
    s[a] = s[a] + s[b] + (x);        \
    s[d] = b3_rotr(s[d] ^ s[a], 16); \
    s[c] = s[c] + s[d];              \
    s[b] = b3_rotr(s[b] ^ s[c], 12); \
    s[a] = s[a] + s[b] + (y);        \
    s[d] = b3_rotr(s[d] ^ s[a], 8);  \
    s[c] = s[c] + s[d];              \
    s[b] = b3_rotr(s[b] ^ s[c], 7);  \
  } while (0)

// cv <- first 8 words of compress(cv, m, counter = 0, block_len, flags). The message
// schedule is applied by renaming (fully unrolled), so no permutation moves are executed.
__device__ __forceinline__ void b3_compress(uint32_t cv[8], const uint32_t m_in[16],
                                            uint32_t block_len, uint32_t flags) {
  uint32_t s[16];
#pragma unroll
  for (int i = 0; i < 8; i++) s[i] = cv[i];
  s[8] = 0x6A09E667u;
  s[9] = 0xBB67AE85u;
  s[10] = 0x3C6EF372u;
  s[11] = 0xA54FF53Au;
  s[12] = 0;
  s[13] = 0;
  s[14] = block_len;
  s[15] = flags;
  uint32_t m[16];
#pragma unroll
  for (int i = 0; i < 16; i++) m[i] = m_in[i];
#pragma unroll
  for (int r = 0; r < 7; r++) {
    LGS_B3_G(0, 4, 8, 12, m[0], m[1]);
    LGS_B3_G(1, 5, 9, 13, m[2], m[3]);
    LGS_B3_G(2, 6, 10, 14, m[4], m[5]);
    LGS_B3_G(3, 7, 11, 15, m[6], m[7]);
    LGS_B3_G(0, 5, 10, 15, m[8], m[9]);
    LGS_B3_G(1, 6, 11, 12, m[10], m[11]);
    LGS_B3_G(2, 7, 8, 13, m[12], m[13]);
    LGS_B3_G(3, 4, 9, 14, m[14], m[15]);
    if (r < 6) {
      uint32_t t[16] = {m[2], m[6], m[3], m[10], m[7], m[0], m[4], m[13],
                        m[1], m[11], m[12], m[5], m[9], m[14], m[15], m[8]};
#pragma unroll
      for (int i = 0; i < 16; i++) m[i] = t[i];
    }
  }
#pragma unroll
  for (int i = 0; i < 8; i++) cv[i] = s[i] ^ s[i + 8];
}

#undef LGS_B3_G

__device__ __forceinline__ void b3_init(uint32_t cv[8]) {
#pragma unroll
  for (int i = 0; i < 8; i++) cv[i] = b3_iv(i);
}

// blake3(left || right) for two 32-byte digests.
__device__ __forceinline__ void b3_parent(const uint32_t l[8], const uint32_t r[8], uint32_t out[8]) {
  uint32_t m[16];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    m[i] = l[i];
    m[i + 8] = r[i];
  }
  b3_init(out);
  b3_compress(out, m, 64, B3_CHUNK_START | B3_CHUNK_END | B3_ROOT);
}

}  // namespace lgs
