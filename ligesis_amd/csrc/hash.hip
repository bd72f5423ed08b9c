// Fused bit-decompose + subset-sum hash H = A . B (pcs/src/ligesis.rs:288-302,
// pcs/src/subset_sum.rs:43-67): H[r][j] = sum over i < W with bit (i mod 64) of F'[i / 64][j]
// set, of A[r][i], mod p. B (W x 2n bits) is never materialised: its bits are read straight out
// of the encoded rows F' in registers.
//
// Two exact kernels:
//  * VALU: one lane per column; A rows are wave-uniform (scalar loads from the transposed copy
//    At[i][r]); 96-bit accumulation without reduction, one reduction per output.
//  * MFMA: A is split at setup into 8 signed base-256 digits, d_l in [-128, 127] with
//    sum_l d_l 256^l = A (mod p) (A itself when A <= 0x7F7F7F7F7F7F7F7F, else A - p). Then
//    S_l = D_l . B is an exact int8 x {0,1} product on v_mfma_i32_16x16x64_i8 with int32 sums
//    (|S_l| <= 128 W < 2^31), and H = sum_l 256^l S_l mod p. M = 8 digits x 32 rows = 256.
// Both write per-K-split partials that a finalize kernel sums mod p.
#include <map>
#include <mutex>

#include <hip/hip_ext.h>

#include "kernels.hpp"

namespace lgs {

typedef int v4i __attribute__((ext_vector_type(4)));

// ---------------- VALU ----------------
// grid (ncols / 256, ksplit); at = A transposed, W x 32.
__global__ __launch_bounds__(256) void hash_valu_kernel(const uint64_t* __restrict__ enc,
                                                        uint64_t ld, const uint64_t* __restrict__ at,
                                                        uint32_t rows_per_split, uint32_t ncols,
                                                        uint64_t* __restrict__ part) {
  const uint32_t j = blockIdx.x * 256 + threadIdx.x;
  const uint32_t split = blockIdx.y;
  const uint32_t i0 = split * rows_per_split;
  uint64_t lo[32], hi[32];
#pragma unroll
  for (int r = 0; r < 32; r++) lo[r] = hi[r] = 0;
  const bool live = j < ncols;
  for (uint32_t i = i0; i < i0 + rows_per_split; i++) {
    const uint64_t x = live ? enc[(uint64_t)i * ld + j] : 0;
    const uint64_t* arow = at + (uint64_t)i * 64 * 32;
    for (int k = 0; k < 64; k++) {
      const uint64_t mask = 0 - ((x >> k) & 1);
      const uint64_t* a = arow + k * 32;
#pragma unroll
      for (int r = 0; r < 32; r++) {
        const uint64_t v = a[r];
        lo[r] += (v & 0xffffffffull) & mask;
        hi[r] += (v >> 32) & mask;
      }
    }
  }
  if (!live) return;
#pragma unroll
  for (int r = 0; r < 32; r++) {
    // value = lo + hi * 2^32 (< 2^96)
    uint64_t sh = hi[r] << 32;
    uint64_t L = lo[r] + sh;
    uint64_t H = (hi[r] >> 32) + (L < sh ? 1 : 0);
    part[((uint64_t)split * 32 + r) * ncols + j] = gl_reduce(L, H);
  }
}

// H = the sum of the K-split partials, mod p, or (ring) mod 2^64 for the Z_2^64 hash.
__global__ void hash_valu_finalize_kernel(const uint64_t* __restrict__ part, uint32_t ksplit,
                                          uint32_t ncols, uint64_t* __restrict__ h, int ring) {
  uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (uint64_t)32 * ncols) return;
  uint64_t acc = 0;
  for (uint32_t s = 0; s < ksplit; s++) {
    const uint64_t v = part[(uint64_t)s * 32 * ncols + idx];
    acc = ring ? acc + v : gl_add(acc, v);
  }
  h[idx] = acc;
}

// ---------------- MFMA ----------------
// Digit layout (prepared at setup): for F' row i ("slab", 64 values of the K index), M-tile mt
// (rows m = 16 mt .. 16 mt + 15, m = 32 l + r for digit l and hash row r), lane L = 16 g + c:
// 16 bytes at ((i * 16 + mt) * 64 + L) * 16, byte 4 s + t = digit l of A[r][64 i + kb(g) + s + 8 t]
// with r = 16 (mt & 1) + c, l = mt >> 1, kb(g) = 32 (g >> 1) + 4 (g & 1). The B operand of lane
// L holds the same (s, t) bits of F'[i][col] shifted right by kb(g), so A and B agree on the
// K order inside every lane whatever order the instruction assigns to the 16 bytes.
constexpr int HM_NT = 128;       // columns per workgroup
constexpr int HM_SPS = 2;        // slabs staged in LDS per step (16 KiB each)

__global__ __launch_bounds__(256) void hash_mfma_kernel(const uint64_t* __restrict__ enc, uint64_t ld,
                                                        const v4i* __restrict__ adig,
                                                        uint32_t slabs_per_split, uint32_t ntiles,
                                                        uint32_t ncols, int32_t* __restrict__ part,
                                                        uint32_t split0) {
  __shared__ v4i lds_a[HM_SPS * 1024];
  // XCD-aware mapping: consecutive block ids land on different XCDs; give each XCD a
  // contiguous run of logical blocks so the blocks sharing a K-split share an L2.
  const uint32_t nb = gridDim.x;
  const uint32_t per_xcd = nb / 8;
  const uint32_t bid = blockIdx.x;
  const uint32_t logical = (nb % 8 == 0) ? (bid % 8) * per_xcd + bid / 8 : bid;
  const uint32_t split = split0 + logical / ntiles;
  const uint32_t ntile = logical % ntiles;

  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t wm = wave >> 1, wn = wave & 1;
  const uint32_t g = lane >> 4, c = lane & 15;
  const uint32_t kb = 32 * (g >> 1) + 4 * (g & 1);
  const uint32_t col_base = ntile * HM_NT + wn * 64 + c;

  v4i acc[8][4];
#pragma unroll
  for (int a = 0; a < 8; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) acc[a][b] = v4i{0, 0, 0, 0};

  const uint32_t slab0 = split * slabs_per_split;
  for (uint32_t s0 = 0; s0 < slabs_per_split; s0 += HM_SPS) {
    __syncthreads();
    const v4i* src = adig + (uint64_t)(slab0 + s0) * 1024;
#pragma unroll
    for (int q = 0; q < HM_SPS * 4; q++) lds_a[q * 256 + threadIdx.x] = src[q * 256 + threadIdx.x];
    __syncthreads();
#pragma unroll
    for (int ss = 0; ss < HM_SPS; ss++) {
      const uint64_t i = slab0 + s0 + ss;
      v4i b[4];
#pragma unroll
      for (int nt = 0; nt < 4; nt++) {
        const uint64_t x = enc[i * ld + col_base + nt * 16];
        const uint32_t y = (uint32_t)(x >> kb);
        b[nt] = v4i{(int)(y & 0x01010101u), (int)((y >> 1) & 0x01010101u),
                    (int)((y >> 2) & 0x01010101u), (int)((y >> 3) & 0x01010101u)};
      }
#pragma unroll
      for (int mt = 0; mt < 8; mt++) {
        const v4i a = lds_a[ss * 1024 + (wm * 8 + mt) * 64 + lane];
#pragma unroll
        for (int nt = 0; nt < 4; nt++)
          acc[mt][nt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b[nt], acc[mt][nt], 0, 0, 0);
      }
    }
  }
  // D[row][col] of a 16x16 tile: lane L, register v -> row 4 (L >> 4) + v, col L & 15.
#pragma unroll
  for (int mt = 0; mt < 8; mt++) {
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const uint32_t m = (wm * 8 + mt) * 16 + 4 * g + v;
        const uint32_t col = ntile * HM_NT + wn * 64 + nt * 16 + c;
        part[((uint64_t)split * 256 + m) * ncols + col] = acc[mt][nt][v];
      }
    }
  }
}

__device__ __forceinline__ uint32_t xcd_logical_block() {
  // consecutive block ids are dealt round-robin to the 8 XCDs; give each XCD a contiguous run
  // of logical tiles so the tiles sharing a K-split (and its A digits) share an L2
  const uint32_t nb = gridDim.x, bid = blockIdx.x;
  return (nb % 8 == 0) ? (bid % 8) * (nb / 8) + bid / 8 : bid;
}

// Pipelined variant: A digits double-buffered in LDS (one barrier per stage of SPS slabs), the
// next stage's digits and F' words loaded into registers while the current stage computes,
// and at most 256 registers per lane so two workgroups share a CU.
// WN waves along N: a workgroup has 2 x WN waves (128 WN threads) and a tile of 256 rows x
// 64 WN columns; every wave owns 128 rows x 64 columns.
template <int SPS, int WN, bool PF = false>
__device__ __forceinline__ void hash_pipe_body(const uint64_t* __restrict__ enc, uint64_t ld,
                                               const v4i* __restrict__ adig, uint32_t slabs_per_split,
                                               uint32_t split, uint32_t ntile, uint32_t ncols,
                                               int32_t* __restrict__ part, v4i (*lds_a)[SPS * 1024]);

template <int SPS, int WN, bool PF = false>
__global__ __launch_bounds__(128 * WN, 2) void hash_mfma_pipe_kernel(const uint64_t* __restrict__ enc,
                                                                     uint64_t ld, const v4i* __restrict__ adig,
                                                                     uint32_t slabs_per_split, uint32_t ntiles,
                                                                     uint32_t ncols, int32_t* __restrict__ part,
                                                                     uint32_t split0) {
  __shared__ v4i lds_a[2][SPS * 1024];
  const uint32_t logical = xcd_logical_block();
  hash_pipe_body<SPS, WN, PF>(enc, ld, adig, slabs_per_split, split0 + logical / ntiles, logical % ntiles, ncols,
                              part, lds_a);
}

template <int SPS, int WN, bool PF>
__device__ __forceinline__ void hash_pipe_body(const uint64_t* __restrict__ enc, uint64_t ld,
                                               const v4i* __restrict__ adig, uint32_t slabs_per_split,
                                               uint32_t split, uint32_t ntile, uint32_t ncols,
                                               int32_t* __restrict__ part, v4i (*lds_a)[SPS * 1024]) {
  constexpr int T = 128 * WN;           // threads
  constexpr int NA = SPS * 1024 / T;    // 16-byte A words staged per thread per stage
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t wm = wave / WN, wn = wave % WN;
  const uint32_t g = lane >> 4, c = lane & 15;
  const uint32_t kb = 32 * (g >> 1) + 4 * (g & 1);
  const uint64_t col_base = (uint64_t)ntile * (64 * WN) + wn * 64 + c;

  v4i acc[8][4];
#pragma unroll
  for (int a = 0; a < 8; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) acc[a][b] = v4i{0, 0, 0, 0};

  const uint64_t slab0 = (uint64_t)split * slabs_per_split;
  const uint32_t nstages = slabs_per_split / SPS;
  v4i pa[NA];
  uint64_t px[SPS][4];
  {
    const v4i* src = adig + slab0 * 1024;
#pragma unroll
    for (int q = 0; q < NA; q++) pa[q] = src[q * T + threadIdx.x];
#pragma unroll
    for (int ss = 0; ss < SPS; ss++)
#pragma unroll
      for (int nt = 0; nt < 4; nt++) px[ss][nt] = enc[(slab0 + ss) * ld + col_base + nt * 16];
#pragma unroll
    for (int q = 0; q < NA; q++) lds_a[0][q * T + threadIdx.x] = pa[q];
  }
  __syncthreads();
  for (uint32_t st = 0; st < nstages; st++) {
    const uint32_t buf = st & 1;
    uint64_t cx[SPS][4];
#pragma unroll
    for (int ss = 0; ss < SPS; ss++)
#pragma unroll
      for (int nt = 0; nt < 4; nt++) cx[ss][nt] = px[ss][nt];
    const bool more = st + 1 < nstages;
    if (more) {
      const uint64_t s1 = slab0 + (uint64_t)(st + 1) * SPS;
      const v4i* src = adig + s1 * 1024;
#pragma unroll
      for (int q = 0; q < NA; q++) pa[q] = src[q * T + threadIdx.x];
#pragma unroll
      for (int ss = 0; ss < SPS; ss++)
#pragma unroll
        for (int nt = 0; nt < 4; nt++) px[ss][nt] = enc[(s1 + ss) * ld + col_base + nt * 16];
    }
#pragma unroll
    for (int ss = 0; ss < SPS; ss++) {
      v4i b[4];
#pragma unroll
      for (int nt = 0; nt < 4; nt++) {
        const uint32_t y = (uint32_t)(cx[ss][nt] >> kb);
        b[nt] = v4i{(int)(y & 0x01010101u), (int)((y >> 1) & 0x01010101u),
                    (int)((y >> 2) & 0x01010101u), (int)((y >> 3) & 0x01010101u)};
      }
      if (PF) {
        // A fragment mt + 1 is read while the four products of fragment mt issue, so the LDS
        // latency hides behind MFMA work; the schedule is pinned (one LDS read, then four MFMAs)
        v4i a = lds_a[buf][ss * 1024 + (wm * 8) * 64 + lane];
#pragma unroll
        for (int mt = 0; mt < 8; mt++) {
          v4i an = a;
          if (mt < 7) an = lds_a[buf][ss * 1024 + (wm * 8 + mt + 1) * 64 + lane];
#pragma unroll
          for (int nt = 0; nt < 4; nt++)
            acc[mt][nt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b[nt], acc[mt][nt], 0, 0, 0);
          a = an;
        }
        // groups in issue order: reads of fragments 0 and 1, then (4 MFMAs, the next read) x 6,
        // then the last 8 MFMAs
        __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
#pragma unroll
        for (int mt = 0; mt < 6; mt++) {
          __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);
          __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
        }
        __builtin_amdgcn_sched_group_barrier(0x008, 8, 0);
      } else {
#pragma unroll
        for (int mt = 0; mt < 8; mt++) {
          const v4i a = lds_a[buf][ss * 1024 + (wm * 8 + mt) * 64 + lane];
#pragma unroll
          for (int nt = 0; nt < 4; nt++)
            acc[mt][nt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b[nt], acc[mt][nt], 0, 0, 0);
        }
      }
    }
    if (more) {
#pragma unroll
      for (int q = 0; q < NA; q++) lds_a[buf ^ 1][q * T + threadIdx.x] = pa[q];
    }
    __syncthreads();
  }
#pragma unroll
  for (int mt = 0; mt < 8; mt++) {
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const uint32_t m = (wm * 8 + mt) * 16 + 4 * g + v;
        const uint64_t col = (uint64_t)ntile * (64 * WN) + wn * 64 + nt * 16 + c;
        part[((uint64_t)split * 256 + m) * ncols + col] = acc[mt][nt][v];
      }
    }
  }
}

// Same tile as hash_pipe_body (one slab per stage), with the global loads issued D stages ahead
// into a ring of D register slots, so D - 1 stages of MFMA work cover HBM latency instead of
// one. Slot u of an unrolled group holds stage st0 + u: its B words are expanded first, then the
// slot is refilled with stage st0 + u + D (clamped to the last slab: those loads are unused).
// A(st + 1) moves from its slot into the other LDS buffer after the stage's MFMAs.
// Needs slabs_per_split % D == 0.
template <int WN, int D>
__global__ __launch_bounds__(128 * WN, 2) void hash_mfma_ring_kernel(const uint64_t* __restrict__ enc,
                                                                     uint64_t ld, const v4i* __restrict__ adig,
                                                                     uint32_t slabs_per_split, uint32_t ntiles,
                                                                     uint32_t ncols, int32_t* __restrict__ part,
                                                                     uint32_t split0) {
  constexpr int T = 128 * WN;
  constexpr int NA = 1024 / T;
  __shared__ v4i lds_a[2][1024];
  const uint32_t logical = xcd_logical_block();
  const uint32_t split = split0 + logical / ntiles, ntile = logical % ntiles;
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t wm = wave / WN, wn = wave % WN;
  const uint32_t g = lane >> 4, c = lane & 15;
  const uint32_t kb = 32 * (g >> 1) + 4 * (g & 1);
  const uint64_t col_base = (uint64_t)ntile * (64 * WN) + wn * 64 + c;

  v4i acc[8][4];
#pragma unroll
  for (int a = 0; a < 8; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) acc[a][b] = v4i{0, 0, 0, 0};

  const uint32_t ns = slabs_per_split;
  const uint64_t slab0 = (uint64_t)split * ns;
  v4i ra[D][NA];
  uint64_t rb[D][4];
#pragma unroll
  for (int u = 0; u < D; u++) {
    const uint64_t sl = slab0 + (u < (int)ns ? u : ns - 1);
#pragma unroll
    for (int q = 0; q < NA; q++) ra[u][q] = adig[sl * 1024 + q * T + threadIdx.x];
#pragma unroll
    for (int nt = 0; nt < 4; nt++) rb[u][nt] = enc[sl * ld + col_base + nt * 16];
  }
#pragma unroll
  for (int q = 0; q < NA; q++) lds_a[0][q * T + threadIdx.x] = ra[0][q];
  __syncthreads();

  for (uint32_t st0 = 0; st0 < ns; st0 += D) {
#pragma unroll
    for (int u = 0; u < D; u++) {
      const uint32_t st = st0 + u;
      v4i b[4];
#pragma unroll
      for (int nt = 0; nt < 4; nt++) {
        const uint32_t y = (uint32_t)(rb[u][nt] >> kb);
        b[nt] = v4i{(int)(y & 0x01010101u), (int)((y >> 1) & 0x01010101u), (int)((y >> 2) & 0x01010101u),
                    (int)((y >> 3) & 0x01010101u)};
      }
      {
        const uint32_t nx = st + D;
        const uint64_t sl = slab0 + (nx < ns ? nx : ns - 1);
#pragma unroll
        for (int q = 0; q < NA; q++) ra[u][q] = adig[sl * 1024 + q * T + threadIdx.x];
#pragma unroll
        for (int nt = 0; nt < 4; nt++) rb[u][nt] = enc[sl * ld + col_base + nt * 16];
        // keep the refill here: without a scheduling barrier the compiler sinks these loads to
        // the loop latch, next to their first use, and the ring hides no latency
        asm volatile("" ::: "memory");
      }
      const uint32_t buf = st & 1;
#pragma unroll
      for (int mt = 0; mt < 8; mt++) {
        const v4i a = lds_a[buf][(wm * 8 + mt) * 64 + lane];
#pragma unroll
        for (int nt = 0; nt < 4; nt++)
          acc[mt][nt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b[nt], acc[mt][nt], 0, 0, 0);
      }
      if (st + 1 < ns) {
#pragma unroll
        for (int q = 0; q < NA; q++) lds_a[buf ^ 1][q * T + threadIdx.x] = ra[(u + 1) % D][q];
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int mt = 0; mt < 8; mt++) {
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const uint32_t m = (wm * 8 + mt) * 16 + 4 * g + v;
        const uint64_t col = (uint64_t)ntile * (64 * WN) + wn * 64 + nt * 16 + c;
        part[((uint64_t)split * 256 + m) * ncols + col] = acc[mt][nt][v];
      }
    }
  }
}

// Smaller wave tiles for more waves per SIMD: every wave owns MT x NT 16 x 16 tiles (MT * 16
// rows of the 256, NT * 16 columns), so its accumulators take MT * NT * 4 registers; a workgroup
// has (16 / MT) x WN waves covering all 256 rows and WN * NT * 16 columns. With MT = NT = 4 a wave
// needs 64 accumulator registers and four waves fit on a SIMD, which hides the per-stage latency
// (loads, LDS staging, barrier) that bounds the 128 x 64 tiles at two waves per SIMD.
template <int MT, int NT, int WN, int MINB>
__global__ __launch_bounds__(64 * (16 / MT) * WN, MINB) void hash_mfma_tile_kernel(
    const uint64_t* __restrict__ enc, uint64_t ld, const v4i* __restrict__ adig, uint32_t slabs_per_split,
    uint32_t ntiles, uint32_t ncols, int32_t* __restrict__ part, uint32_t split0) {
  constexpr int WM = 16 / MT;
  constexpr int T = 64 * WM * WN;
  constexpr int NA = 1024 / T;  // 16-byte A words staged per thread per slab
  static_assert(NA >= 1 && 1024 % T == 0, "A staging needs T | 1024");
  __shared__ v4i lds_a[2][1024];
  const uint32_t logical = xcd_logical_block();
  const uint32_t split = split0 + logical / ntiles, ntile = logical % ntiles;
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t wm = wave / WN, wn = wave % WN;
  const uint32_t g = lane >> 4, c = lane & 15;
  const uint64_t col_base = (uint64_t)ntile * (WN * NT * 16) + wn * (NT * 16) + c;

  v4i acc[MT][NT];
#pragma unroll
  for (int a = 0; a < MT; a++)
#pragma unroll
    for (int b = 0; b < NT; b++) acc[a][b] = v4i{0, 0, 0, 0};

  const uint32_t ns = slabs_per_split;
  const uint64_t slab0 = (uint64_t)split * ns;
  // a lane uses bits kb .. kb + 27 of its F' word, all inside 32-bit half g >> 1: it loads that
  // half only, and y = half >> (kb mod 32)
  const uint32_t* enc32 = (const uint32_t*)enc + (g >> 1);
  const uint32_t ksh = 4 * (g & 1);
  v4i pa[NA];
  uint32_t px[NT];
#pragma unroll
  for (int q = 0; q < NA; q++) pa[q] = adig[slab0 * 1024 + q * T + threadIdx.x];
#pragma unroll
  for (int nt = 0; nt < NT; nt++) px[nt] = enc32[2 * (slab0 * ld + col_base + nt * 16)];
#pragma unroll
  for (int q = 0; q < NA; q++) lds_a[0][q * T + threadIdx.x] = pa[q];
  __syncthreads();
  for (uint32_t st = 0; st < ns; st++) {
    const uint32_t buf = st & 1;
    uint32_t cx[NT];
#pragma unroll
    for (int nt = 0; nt < NT; nt++) cx[nt] = px[nt];
    const bool more = st + 1 < ns;
    if (more) {
      const uint64_t s1 = slab0 + st + 1;
#pragma unroll
      for (int q = 0; q < NA; q++) pa[q] = adig[s1 * 1024 + q * T + threadIdx.x];
#pragma unroll
      for (int nt = 0; nt < NT; nt++) px[nt] = enc32[2 * (s1 * ld + col_base + nt * 16)];
    }
    v4i b[NT];
#pragma unroll
    for (int nt = 0; nt < NT; nt++) {
      const uint32_t y = cx[nt] >> ksh;
      b[nt] = v4i{(int)(y & 0x01010101u), (int)((y >> 1) & 0x01010101u), (int)((y >> 2) & 0x01010101u),
                  (int)((y >> 3) & 0x01010101u)};
    }
#pragma unroll
    for (int mt = 0; mt < MT; mt++) {
      const v4i a = lds_a[buf][(wm * MT + mt) * 64 + lane];
#pragma unroll
      for (int nt = 0; nt < NT; nt++)
        acc[mt][nt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b[nt], acc[mt][nt], 0, 0, 0);
    }
    if (more) {
#pragma unroll
      for (int q = 0; q < NA; q++) lds_a[buf ^ 1][q * T + threadIdx.x] = pa[q];
    }
    __syncthreads();
  }
#pragma unroll
  for (int mt = 0; mt < MT; mt++) {
#pragma unroll
    for (int nt = 0; nt < NT; nt++) {
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const uint32_t m = (wm * MT + mt) * 16 + 4 * g + v;
        const uint64_t col = (uint64_t)ntile * (WN * NT * 16) + wn * (NT * 16) + nt * 16 + c;
        part[((uint64_t)split * 256 + m) * ncols + col] = acc[mt][nt][v];
      }
    }
  }
}

// Asynchronous-copy variant: a 256 x 64 WN tile (2 x WN waves of 128 x 64; WN = 2 for 128 and
// WN = 4 for 256 columns). Each slab's A digits (16 KiB, shared by every wave of the workgroup)
// and F' words (64 WN columns, 512 WN bytes) go from global memory straight into an LDS ring of
// NB slots (global_load_lds, no staging registers), NB - 1 slabs ahead of the slab being
// multiplied. Every wave issues the same copies per slab (dma_copies: 16 KiB of A over the
// workgroup's waves in 1 KiB pieces, and one 256-byte piece of F'), so "this wave's copies of slab
// s are done" is vmcnt <= copies (NB - 2) once the copies of the later slabs are issued; a barrier
// then makes every wave's copies visible. The copy into a slot is issued after the barrier that
// follows the last reads of its previous slab. Near the end the copies repeat the last slab
// (unused), so the count stays uniform.
//
// Copies into LDS must not share a CU with workgroups of other kernels: on this device such
// copies have corrupted the LDS of co-resident workgroups of other kernels (never this kernel's
// own results), so every launch of these kernels requests the whole LDS of a CU (dma_lds_bytes).
// WN = 4 keeps two waves per SIMD in one workgroup per CU; WN = 2 then runs at one.
#define LGS_AS1 __attribute__((address_space(1)))
#define LGS_AS3 __attribute__((address_space(3)))
constexpr int vmcnt_imm(int n) { return (n & 0xF) | (((n >> 4) & 3) << 14) | (7 << 4) | (0xF << 8); }
constexpr int kLgkmZero = 0xC07F;  // lgkmcnt(0), vmcnt and expcnt unconstrained

// One slot of the ring: separate LDS arrays per slot, so the compiler can tell a slot's reads from
// the copies in flight into the other slots (a shared array with a dynamic slot index makes it
// wait for every copy before each read).
struct DmaSlot {
  v4i* a;
  uint32_t* b;
};

struct DmaTile {
  const uint64_t* enc;
  uint64_t ld;
  const v4i* adig;
  uint64_t slab0, col0;
  uint32_t ns, tid, lane, wave, wm, wn, g, c, ksh;
};

template <int WN>
__device__ __forceinline__ DmaTile make_dma_tile(const uint64_t* enc, uint64_t ld, const v4i* adig, uint32_t ns,
                                                 uint32_t split, uint32_t ntile) {
  DmaTile t;
  t.enc = enc;
  t.ld = ld;
  t.adig = adig;
  t.ns = ns;
  t.tid = threadIdx.x;
  t.lane = t.tid & 63;
  t.wave = t.tid >> 6;
  t.wm = t.wave / WN;
  t.wn = t.wave % WN;
  t.g = t.lane >> 4;
  t.c = t.lane & 15;
  t.ksh = 4 * (t.g & 1);
  t.col0 = (uint64_t)ntile * (64 * WN);
  t.slab0 = (uint64_t)split * ns;
  return t;
}

// Copies per wave and slab: 16 KiB of A in 1 KiB pieces over the 2 WN waves, plus one of F'.
template <int WN>
constexpr int dma_copies() { return 16 / (2 * WN) + 1; }

template <int WN>
__device__ __forceinline__ void dma_copy_slab(const DmaTile& t, uint32_t sl, DmaSlot d) {
  constexpr int T = 128 * WN;  // threads
  const v4i* src = t.adig + (t.slab0 + sl) * 1024;
#pragma unroll
  for (int q = 0; q < 1024 / T; q++)
    __builtin_amdgcn_global_load_lds((const LGS_AS1 void*)(src + q * T + t.tid),
                                     (LGS_AS3 void*)(d.a + q * T + t.wave * 64), 16, 0, 0);
  const uint32_t* bs = (const uint32_t*)(t.enc + (t.slab0 + sl) * t.ld + t.col0) + t.tid;
  __builtin_amdgcn_global_load_lds((const LGS_AS1 void*)bs, (LGS_AS3 void*)(d.b + t.wave * 64), 4, 0, 0);
}

// Slab sl from slot cur; the copy of slab sl + NB - 1 goes to slot nxt.
template <int NB, int WN>
__device__ __forceinline__ void dma_step(const DmaTile& t, uint32_t sl, DmaSlot cur, DmaSlot nxt, v4i (&acc)[8][4]) {
  __builtin_amdgcn_s_waitcnt(vmcnt_imm(dma_copies<WN>() * (NB - 2)));
  asm volatile("s_barrier" ::: "memory");
  const uint32_t nx = sl + NB - 1;
  dma_copy_slab<WN>(t, nx < t.ns ? nx : t.ns - 1, nxt);
  // the products issue at raised wave priority, so a SIMD's other wave (copies, barrier) does
  // not take issue slots from them
  __builtin_amdgcn_sched_barrier(0);
  __builtin_amdgcn_s_setprio(1);
  v4i b[4];
#pragma unroll
  for (int nt = 0; nt < 4; nt++) {
    const uint32_t y = cur.b[2 * (t.wn * 64 + nt * 16 + t.c) + (t.g >> 1)] >> t.ksh;
    b[nt] = v4i{(int)(y & 0x01010101u), (int)((y >> 1) & 0x01010101u), (int)((y >> 2) & 0x01010101u),
                (int)((y >> 3) & 0x01010101u)};
  }
  // A fragments in pairs: the pair after next is read while the current pair's eight products
  // issue (with copies into LDS in flight the compiler's LDS waits are all-or-nothing, so a read
  // is waited for one product group after it is issued)
  const v4i* A = cur.a + (t.wm * 8) * 64 + t.lane;
  v4i a0 = A[0], a1 = A[64];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    v4i n0 = a0, n1 = a1;
    if (k < 3) {
      n0 = A[(2 * k + 2) * 64];
      n1 = A[(2 * k + 3) * 64];
    }
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
      acc[2 * k][nt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a0, b[nt], acc[2 * k][nt], 0, 0, 0);
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
      acc[2 * k + 1][nt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a1, b[nt], acc[2 * k + 1][nt], 0, 0, 0);
    a0 = n0;
    a1 = n1;
  }
  // issue order: the B reads (two paired reads) and A pairs 0 and 1; pair 0's products; pair 2's
  // reads; pair 1's products; one product of pair 2 (after the wait for pair 2, so pair 3's reads
  // are not waited for with it); pair 3's reads; the remaining products
  __builtin_amdgcn_sched_group_barrier(0x100, 6, 0);
  __builtin_amdgcn_sched_group_barrier(0x008, 8, 0);
  __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
  __builtin_amdgcn_sched_group_barrier(0x008, 9, 0);
  __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
  __builtin_amdgcn_sched_group_barrier(0x008, 15, 0);
  __builtin_amdgcn_sched_barrier(0);
  __builtin_amdgcn_s_setprio(0);
}

// Epilogue of a 256 x 64 WN tile: the split's contribution to H, reduced mod p. Row m = 32 l + r of
// the tile is digit plane l of hash row r; a lane holds, for r = 16 h + 4 g + v and column
// nt * 16 + c, the sums of planes l = 4 wm + k (k < 4) in acc[2 k + h][nt][v]. Each wave folds its
// four planes exactly in 64 bits (|sum_k 256^k S| < 2^55); the upper planes' fold (wm = 1) moves
// through LDS (xch: 32 x 64 words of a ring slot, free once every copy and read has finished) to
// the lower planes' wave, which adds 2^32 times it mod p and writes one word per (split, row,
// column).
template <bool RING>
__device__ __forceinline__ void dma_epilogue(const DmaTile& t, v4i (&acc)[8][4], uint64_t* xch, uint32_t split,
                                             uint32_t ncols, uint64_t* __restrict__ part) {
  __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));  // no copy into LDS outlives the workgroup
  __builtin_amdgcn_s_waitcnt(kLgkmZero);
  asm volatile("s_barrier" ::: "memory");
  // fold first, so the accumulators die here: f[(h * 4 + nt) * 4 + v]
  uint64_t f[32];
#pragma unroll
  for (int h = 0; h < 2; h++)
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        int64_t x = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) x += (int64_t)acc[2 * k + h][nt][v] * ((int64_t)1 << (8 * k));
        // mod p, or mod 2^64 for the Z_2^64 hash (two's complement wrap)
        f[(h * 4 + nt) * 4 + v] = RING ? (uint64_t)x : x >= 0 ? (uint64_t)x : GL_P - (uint64_t)(-x);
      }
  if (t.wm == 1) {
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
      for (int nt = 0; nt < 4; nt++)
#pragma unroll
        for (int v = 0; v < 4; v++) xch[(16 * h + 4 * t.g + v) * 64 + nt * 16 + t.c] = f[(h * 4 + nt) * 4 + v];
  }
  __builtin_amdgcn_s_waitcnt(kLgkmZero);
  asm volatile("s_barrier" ::: "memory");
  if (t.wm == 0) {
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
      for (int nt = 0; nt < 4; nt++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const uint32_t r = 16 * h + 4 * t.g + v;
          const uint64_t hi = xch[r * 64 + nt * 16 + t.c];
          const uint64_t col = t.col0 + t.wn * 64 + nt * 16 + t.c;
          part[((uint64_t)split * 32 + r) * ncols + col] =
              RING ? f[(h * 4 + nt) * 4 + v] + (hi << 32) : gl_add(f[(h * 4 + nt) * 4 + v], gl_mul(hi, uint64_t(1) << 32));
        }
  }
}

// One 256 x 64 WN tile over slabs [split * ns, (split + 1) * ns) with a ring of NB slots (NB = 4
// needs ns % 4 == 0).
template <int NB, int WN, bool RING = false>
__device__ __forceinline__ void hash_dma_body(const uint64_t* __restrict__ enc, uint64_t ld,
                                              const v4i* __restrict__ adig, uint32_t ns, uint32_t split,
                                              uint32_t ntile, uint32_t ncols, uint64_t* __restrict__ part,
                                              DmaSlot s0, DmaSlot s1, DmaSlot s2, DmaSlot s3) {
  static_assert(NB == 3 || NB == 4, "ring of 3 or 4 slabs");
  static_assert(WN == 2 || WN == 4, "2 or 4 waves along N");
  // the epilogue exchanges through slot wn's A array: with WN = 4 all four slots hold 16 KiB,
  // also for a 3-slot ring
  const DmaTile t = make_dma_tile<WN>(enc, ld, adig, ns, split, ntile);

  v4i acc[8][4];
#pragma unroll
  for (int a = 0; a < 8; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) acc[a][b] = v4i{0, 0, 0, 0};

  dma_copy_slab<WN>(t, 0, s0);
  dma_copy_slab<WN>(t, 1 < ns ? 1 : ns - 1, s1);
  if (NB == 4) dma_copy_slab<WN>(t, 2 < ns ? 2 : ns - 1, s2);
  for (uint32_t sl = 0; sl < ns; sl += NB) {
    if (NB == 3) {
      dma_step<3, WN>(t, sl, s0, s2, acc);
      if (sl + 1 < ns) dma_step<3, WN>(t, sl + 1, s1, s0, acc);
      if (sl + 2 < ns) dma_step<3, WN>(t, sl + 2, s2, s1, acc);
    } else {  // one straight-line body
      dma_step<4, WN>(t, sl, s0, s3, acc);
      dma_step<4, WN>(t, sl + 1, s1, s0, acc);
      dma_step<4, WN>(t, sl + 2, s2, s1, acc);
      dma_step<4, WN>(t, sl + 3, s3, s2, acc);
    }
  }
  dma_epilogue<RING>(t, acc, (uint64_t*)(t.wn == 0 ? s0.a : t.wn == 1 ? s1.a : t.wn == 2 ? s2.a : s3.a), split, ncols,
                     part);
}

template <int NB, int WN, bool RING = false>
__global__ __launch_bounds__(128 * WN, 2) void hash_mfma_dma_kernel(const uint64_t* __restrict__ enc, uint64_t ld,
                                                                   const v4i* __restrict__ adig, uint32_t ns,
                                                                   uint32_t ntiles, uint32_t ncols,
                                                                   int32_t* __restrict__ part, uint32_t split0) {
  __shared__ v4i la0[1024], la1[1024], la2[1024], la3[NB == 4 ? 1024 : 1];
  __shared__ uint32_t lb0[128 * WN], lb1[128 * WN], lb2[128 * WN], lb3[NB == 4 ? 128 * WN : 1];
  const uint32_t logical = xcd_logical_block();
  hash_dma_body<NB, WN, RING>(enc, ld, adig, ns, split0 + logical / ntiles, logical % ntiles, ncols, (uint64_t*)part,
                        DmaSlot{la0, lb0}, DmaSlot{la1, lb1}, DmaSlot{la2, lb2}, DmaSlot{la3, lb3});
}

// Tiles of many commitments in one launch (batched commit): logical tile -> job by binary
// search over the jobs' first tiles; a job whose split has a multiple of 4 slabs takes the 4-slot
// ring. Tiles are 64 WN columns wide.
template <int WN>
__global__ __launch_bounds__(128 * WN, 2) void hash_mfma_dma_group_kernel(const HashJob* __restrict__ jobs, int njobs) {
  __shared__ v4i la0[1024], la1[1024], la2[1024], la3[1024];
  __shared__ uint32_t lb0[128 * WN], lb1[128 * WN], lb2[128 * WN], lb3[128 * WN];
  const uint32_t logical = xcd_logical_block();
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (jobs[mid].first_block <= logical) lo = mid;
    else hi = mid - 1;
  }
  const HashJob& jb = jobs[lo];
  const uint32_t local = logical - jb.first_block;
  const DmaSlot s0{la0, lb0}, s1{la1, lb1}, s2{la2, lb2}, s3{la3, lb3};
  if (jb.slabs_per_split % 4 != 0)
    hash_dma_body<3, WN>(jb.enc, jb.ld, (const v4i*)jb.adig, jb.slabs_per_split, local / jb.ntiles, local % jb.ntiles,
                         jb.ncols, (uint64_t*)jb.part, s0, s1, s2, s3);
  else
    hash_dma_body<4, WN>(jb.enc, jb.ld, (const v4i*)jb.adig, jb.slabs_per_split, local / jb.ntiles,
                         local % jb.ntiles, jb.ncols, (uint64_t*)jb.part, s0, s1, s2, s3);
}

// H[r][j] = sum_l 256^l (sum_split part[split][32 l + r][j]) mod p.
// Software-pipelined variant of the 8-wave tile (256 x 256, WN = 4): the F' bits of slab s + 1 are
// expanded, and its first A pair read, while slab s's 32 products issue, one expansion instruction
// after each product, so the matrix pipe is fed without a gap for the expansion or the first LDS
// reads. The barrier before slab s waits for slab s + 1's copies (vmcnt <= copies (NB - 3)): slab
// s + 1 is then visible to every wave, and every wave has finished slab s - 1, whose slot takes the
// copies of slab s + NB - 1. Those copies issue one at a time between slab s's products, so no
// wave waits behind a burst of copies before its products. ns % NB == 0.
__device__ __forceinline__ v4i expand_bits(uint32_t y) {
  return v4i{(int)(y & 0x01010101u), (int)((y >> 1) & 0x01010101u), (int)((y >> 2) & 0x01010101u),
             (int)((y >> 3) & 0x01010101u)};
}

// Slab sl (slot cur, B operands bc and A pair 0 in p0, p1); the next slab's B operands and A pair 0
// (slot nxt) come back in bc, p0, p1; slab sl + NB - KS (clamped) is copied into slot dst.
// KS slabs share one barrier (SYNC on the first of them). Before slab s = KS u the slabs up to
// s + KS must be visible (slab s + KS - 1 reads slab s + KS's first operands) and the copies of
// slabs up to s + NB - KS - 1 have been issued, so slabs s + KS + 1 .. s + NB - KS - 1 may still be
// in flight: vmcnt <= copies (NB - 2 KS - 1). Slab s + j copies slab s + j + NB - KS into the slot
// of slab s + j - KS, which every wave finished before this barrier.
template <int NB, int KS, bool SYNC>
__device__ __forceinline__ void dma_pipe_slab(const DmaTile& t, uint32_t sl, DmaSlot cur, DmaSlot nxt, DmaSlot dst,
                                              v4i (&bc)[4], v4i& p0, v4i& p1, v4i (&acc)[8][4]) {
  static_assert(NB - 2 * KS - 1 >= 0, "ring too short for KS slabs per barrier");
  if (SYNC) {
    __builtin_amdgcn_s_waitcnt(vmcnt_imm(dma_copies<4>() * (NB - 2 * KS - 1)));
    asm volatile("s_barrier" ::: "memory");
  }
  __builtin_amdgcn_sched_barrier(0);
  const uint32_t nx = sl + NB - KS;
  dma_copy_slab<4>(t, nx < t.ns ? nx : t.ns - 1, dst);
  uint32_t y[4];
#pragma unroll
  for (int nt = 0; nt < 4; nt++) y[nt] = nxt.b[2 * (t.wn * 64 + nt * 16 + t.c) + (t.g >> 1)];
  const v4i* A = cur.a + (t.wm * 8) * 64 + t.lane;
  const v4i* An = nxt.a + (t.wm * 8) * 64 + t.lane;
  v4i a0 = p0, a1 = p1, bn[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    v4i n0, n1;
    if (k < 3) {
      n0 = A[(2 * k + 2) * 64];
      n1 = A[(2 * k + 3) * 64];
    } else {
      n0 = An[0];
      n1 = An[64];
    }
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
      acc[2 * k][nt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a0, bc[nt], acc[2 * k][nt], 0, 0, 0);
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
      acc[2 * k + 1][nt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a1, bc[nt], acc[2 * k + 1][nt], 0, 0, 0);
    bn[k] = expand_bits(y[k] >> t.ksh);
    a0 = n0;
    a1 = n1;
  }
  p0 = a0;
  p1 = a1;
#pragma unroll
  for (int nt = 0; nt < 4; nt++) bc[nt] = bn[nt];
  // issue order: B words and A pair 1; four products; then one expansion instruction after each
  // product. A pair k + 1 is read right after pair k's first product (products 8, 16, 24 for
  // pairs 2, 3 and the next slab's pair 0): an LDS wait here covers every read in flight, so
  // a read issued just before the wait for the previous pair would be waited for at once. The
  // three copies go after products 5, 13, 21.
  __builtin_amdgcn_sched_group_barrier(0x100, 4, 0);
  __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);
#pragma unroll
  for (int i = 4; i < 32; i++) {
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
    if (i % 8 == 0) __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
    if (i % 8 == 5) __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
    __builtin_amdgcn_sched_group_barrier(0x002, 1, 0);
  }
  __builtin_amdgcn_sched_barrier(0);
  // the next slab's pair 0 has had eight products to arrive; waiting here (rather than at the
  // next slab's first product, behind that slab's new reads) keeps the products back to back
  __builtin_amdgcn_s_waitcnt(kLgkmZero);
}

// One 256 x 256 tile of the pipelined 8-wave kernel over slabs [split * ns, (split + 1) * ns)
// (ns % NB == 0), slots in s[0 .. NB).
template <int NB, bool RING, int KS>
__device__ __forceinline__ void hash_pipe8_body(const uint64_t* __restrict__ enc, uint64_t ld,
                                                const v4i* __restrict__ adig, uint32_t ns, uint32_t split,
                                                uint32_t ntile, uint32_t ncols, int32_t* __restrict__ part,
                                                const DmaSlot (&s)[8]) {
  const DmaTile t = make_dma_tile<4>(enc, ld, adig, ns, split, ntile);

  v4i acc[8][4];
#pragma unroll
  for (int a = 0; a < 8; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) acc[a][b] = v4i{0, 0, 0, 0};

  // slabs 0 .. NB - KS - 1 in flight; wait for slab 0
#pragma unroll
  for (int j = 0; j < NB - KS; j++) dma_copy_slab<4>(t, (uint32_t)j < ns ? j : ns - 1, s[j]);
  __builtin_amdgcn_s_waitcnt(vmcnt_imm(dma_copies<4>() * (NB - KS - 1)));
  asm volatile("s_barrier" ::: "memory");
  v4i bc[4];
#pragma unroll
  for (int nt = 0; nt < 4; nt++) bc[nt] = expand_bits(s[0].b[2 * (t.wn * 64 + nt * 16 + t.c) + (t.g >> 1)] >> t.ksh);
  v4i p0 = s[0].a[(t.wm * 8) * 64 + t.lane], p1 = s[0].a[(t.wm * 8 + 1) * 64 + t.lane];
  __builtin_amdgcn_s_waitcnt(kLgkmZero);
  for (uint32_t sl = 0; sl < ns; sl += NB) {
#pragma unroll
    for (int j = 0; j < NB; j++) {
      if (j % KS == 0)
        dma_pipe_slab<NB, KS, true>(t, sl + j, s[j], s[(j + 1) % NB], s[(j + NB - KS) % NB], bc, p0, p1, acc);
      else
        dma_pipe_slab<NB, KS, false>(t, sl + j, s[j], s[(j + 1) % NB], s[(j + NB - KS) % NB], bc, p0, p1, acc);
    }
  }
  dma_epilogue<RING>(t, acc, (uint64_t*)(t.wn == 0 ? s[0].a : t.wn == 1 ? s[1].a : t.wn == 2 ? s[2].a : s[3].a),
                    split, ncols, (uint64_t*)part);
}

#define LGS_PIPE8_SLOTS                                                                                              \
  __shared__ v4i la0[1024], la1[1024], la2[1024], la3[1024], la4[NB > 4 ? 1024 : 1], la5[NB > 5 ? 1024 : 1],      \
      la6[NB > 6 ? 1024 : 1], la7[NB > 7 ? 1024 : 1];                                                                \
  __shared__ uint32_t lb0[512], lb1[512], lb2[512], lb3[512], lb4[NB > 4 ? 512 : 1], lb5[NB > 5 ? 512 : 1],        \
      lb6[NB > 6 ? 512 : 1], lb7[NB > 7 ? 512 : 1];                                                                  \
  const DmaSlot s[8] = {DmaSlot{la0, lb0}, DmaSlot{la1, lb1}, DmaSlot{la2, lb2}, DmaSlot{la3, lb3},                 \
                        DmaSlot{la4, lb4}, DmaSlot{la5, lb5}, DmaSlot{la6, lb6}, DmaSlot{la7, lb7}};

template <int NB, bool RING = false, int KS = 1>
__global__ __launch_bounds__(512, 2) void hash_mfma_pipe8_kernel(const uint64_t* __restrict__ enc, uint64_t ld,
                                                                const v4i* __restrict__ adig, uint32_t ns,
                                                                uint32_t ntiles, uint32_t ncols,
                                                                int32_t* __restrict__ part, uint32_t split0) {
  static_assert(NB >= 4 && NB <= 8, "4 to 8 slots");
  LGS_PIPE8_SLOTS
  const uint32_t logical = xcd_logical_block();
  hash_pipe8_body<NB, RING, KS>(enc, ld, adig, ns, split0 + logical / ntiles, logical % ntiles, ncols, part, s);
}

// Tiles of many commitments in one launch (batched commit) with the pipelined 8-wave body: every
// job has a multiple of 256 columns and of 8 slabs per K-split; logical tile -> job by binary
// search over the jobs' first tiles.
__global__ __launch_bounds__(512, 2) void hash_mfma_pipe8_group_kernel(const HashJob* __restrict__ jobs, int njobs) {
  constexpr int NB = 8;
  LGS_PIPE8_SLOTS
  const uint32_t logical = xcd_logical_block();
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (jobs[mid].first_block <= logical) lo = mid;
    else hi = mid - 1;
  }
  const HashJob& jb = jobs[lo];
  const uint32_t local = logical - jb.first_block;
  hash_pipe8_body<NB, false, 2>(jb.enc, jb.ld, (const v4i*)jb.adig, jb.slabs_per_split, local / jb.ntiles,
                                local % jb.ntiles, jb.ncols, jb.part, s);
}

__global__ void hash_mfma_finalize_kernel(const int32_t* __restrict__ part, uint32_t ksplit,
                                          uint32_t ncols, uint64_t* __restrict__ h) {
  uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (uint64_t)32 * ncols) return;
  const uint32_t r = (uint32_t)(idx / ncols), j = (uint32_t)(idx % ncols);
  uint64_t acc = 0;
#pragma unroll
  for (int l = 0; l < 8; l++) {
    int64_t s = 0;
    for (uint32_t sp = 0; sp < ksplit; sp++) s += part[((uint64_t)sp * 256 + 32 * l + r) * ncols + j];
    uint64_t sm = s >= 0 ? (uint64_t)s : gl_neg((uint64_t)(-s));
    acc = gl_add(acc, gl_mul(sm, (uint64_t)1 << (8 * l)));
  }
  h[idx] = acc;
}

// Setup-time layout transforms of A (C = 32 rows x W, row-major, canonical).
__global__ void a_transpose_kernel(const uint64_t* __restrict__ a, uint64_t W, uint64_t* __restrict__ at) {
  uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= 32 * W) return;
  uint64_t r = idx / W, i = idx % W;
  at[i * 32 + r] = a[idx];
}

// ring = 0: digits of A or A - p (both = A mod p); ring = 1 (the Z_2^64 hash): digits of the
// representative of A mod 2^64 in [-0x8080808080808080, 0x7F7F7F7F7F7F7F7F] (the loop's 64-bit
// arithmetic shift gives exactly that one).
__global__ void a_digits_kernel(const uint64_t* __restrict__ a, uint64_t W, int8_t* __restrict__ dig, int ring) {
  // one thread per output byte position (slab, mt, lane, byte)
  uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= 256 * W) return;
  const uint32_t byte = idx & 15;
  const uint32_t lane = (idx >> 4) & 63;
  const uint32_t mt = (idx >> 10) & 15;
  const uint64_t slab = idx >> 14;
  const uint32_t g = lane >> 4, c = lane & 15;
  const uint32_t s = byte >> 2, t = byte & 3;
  const uint32_t r = 16 * (mt & 1) + c, l = mt >> 1;
  const uint64_t i = 64 * slab + 32 * (g >> 1) + 4 * (g & 1) + s + 8 * t;
  uint64_t v = a[(uint64_t)r * W + i];
  // balanced base-256 digits of v or v - p
  int64_t x = (ring || v <= 0x7F7F7F7F7F7F7F7Full) ? (int64_t)v : -(int64_t)(GL_P - v);
  int d = 0;
  for (uint32_t q = 0; q <= l; q++) {
    d = (int)(x & 0xff);
    if (d >= 128) d -= 256;
    x = (x - d) >> 8;
  }
  dig[idx] = (int8_t)d;
}

hipError_t launch_a_layouts(const uint64_t* a, uint64_t W, uint64_t* at, int8_t* dig, hipStream_t st, bool ring) {
  uint64_t n = 32 * W;
  hipLaunchKernelGGL(a_transpose_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, a, W, at);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || dig == nullptr) return e;
  uint64_t nd = 256 * W;
  hipLaunchKernelGGL(a_digits_kernel, dim3((uint32_t)((nd + 255) / 256)), dim3(256), 0, st, a, W, dig,
                     ring ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_hash_valu(const uint64_t* enc, uint64_t ld, uint32_t m_rows, uint32_t ncols,
                            const uint64_t* at, uint32_t ksplit, uint64_t* part, uint64_t* h,
                            hipStream_t st, hipEvent_t mid) {
  uint32_t rps = m_rows / ksplit;
  hipLaunchKernelGGL(hash_valu_kernel, dim3((ncols + 255) / 256, ksplit), dim3(256), 0, st, enc, ld,
                     at, rps, ncols, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (mid && (e = hipEventRecord(mid, st)) != hipSuccess) return e;
  if (!h) return hipSuccess;
  uint64_t n = (uint64_t)32 * ncols;
  hipLaunchKernelGGL(hash_valu_finalize_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st,
                     part, ksplit, ncols, h, 0);
  return hipGetLastError();
}

// Dynamic LDS that makes a launch of `kernel` hold all of a CU's LDS: the CU's LDS minus the
// kernel's static LDS (hipFuncGetAttributes), looked up once per device and kernel.
static uint32_t full_cu_pad(const void* kernel) {
  static std::mutex mu;
  static std::map<std::pair<int, const void*>, uint32_t> known;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  std::lock_guard<std::mutex> g(mu);
  auto it = known.find({dev, kernel});
  if (it != known.end()) return it->second;
  int cu_lds = 0;
  hipFuncAttributes fa{};
  uint32_t pad = 0;
  if (hipDeviceGetAttribute(&cu_lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) == hipSuccess &&
      hipFuncGetAttributes(&fa, kernel) == hipSuccess && (size_t)cu_lds > fa.sharedSizeBytes)
    pad = (uint32_t)(cu_lds - fa.sharedSizeBytes);
  (void)hipGetLastError();
  known[{dev, kernel}] = pad;
  return pad;
}

// A kernel launch whose start and stop times go to e0 / e1 when given (hipExtLaunchKernel: the
// timestamps of the dispatch itself, no marker packets on the stream around it).
template <typename K, typename... A>
static void launch_k(K kernel, dim3 grid, dim3 block, uint32_t lds, hipStream_t st, hipEvent_t e0, hipEvent_t e1,
                     A... args) {
  if (e0 || e1)
    hipExtLaunchKernelGGL(kernel, grid, block, lds, st, e0, e1, 0, args...);
  else
    hipLaunchKernelGGL(kernel, grid, block, lds, st, args...);
}

hipError_t launch_hash_mfma_splits(const uint64_t* enc, uint64_t ld, uint32_t m_rows, uint32_t ncols,
                                  const int8_t* dig, uint32_t ksplit, uint32_t split0, uint32_t nsplits,
                                  int32_t* part, hipStream_t st, int impl, bool ring, hipEvent_t k0, hipEvent_t k1) {
  const uint32_t ntiles = ncols / HM_NT;
  const uint32_t sps = m_rows / ksplit;
  const v4i* ad = (const v4i*)dig;
  if (ring) {  // the Z_2^64 hash: the LDS-copy kernels that reduce in their epilogue only
    if (!hash_impl_reduces(impl)) return hipErrorInvalidValue;
    if (impl == 17 && ncols % 256 == 0 && sps % 8 == 0)
      launch_k((hash_mfma_pipe8_kernel<8, true, 2>), dim3((ncols / 256) * nsplits), dim3(512),
                         full_cu_pad((const void*)hash_mfma_pipe8_kernel<8, true, 2>), st, k0, k1, enc, ld, ad, sps,
                         ncols / 256, ncols, part, split0);
    else if (impl >= 16 && ncols % 256 == 0 && sps % 8 == 0)
      launch_k((hash_mfma_pipe8_kernel<8, true>), dim3((ncols / 256) * nsplits), dim3(512),
                         full_cu_pad((const void*)hash_mfma_pipe8_kernel<8, true>), st, k0, k1, enc, ld, ad, sps, ncols / 256,
                         ncols, part, split0);
    else if (impl >= 15 && ncols % 256 == 0 && sps % 4 == 0)
      launch_k((hash_mfma_pipe8_kernel<4, true>), dim3((ncols / 256) * nsplits), dim3(512),
                         full_cu_pad((const void*)hash_mfma_pipe8_kernel<4, true>), st, k0, k1, enc, ld, ad, sps, ncols / 256,
                         ncols, part, split0);
    else if (impl >= 14 && ncols % 256 == 0 && sps % 4 == 0)
      launch_k((hash_mfma_dma_kernel<4, 4, true>), dim3((ncols / 256) * nsplits), dim3(512),
                         full_cu_pad((const void*)hash_mfma_dma_kernel<4, 4, true>), st, k0, k1, enc, ld, ad, sps, ncols / 256,
                         ncols, part, split0);
    else if (impl == 11 || sps % 4)
      launch_k((hash_mfma_dma_kernel<3, 2, true>), dim3(ntiles * nsplits), dim3(256),
                         full_cu_pad((const void*)hash_mfma_dma_kernel<3, 2, true>), st, k0, k1, enc, ld, ad, sps, ntiles,
                         ncols, part, split0);
    else
      launch_k((hash_mfma_dma_kernel<4, 2, true>), dim3(ntiles * nsplits), dim3(256),
                         full_cu_pad((const void*)hash_mfma_dma_kernel<4, 2, true>), st, k0, k1, enc, ld, ad, sps, ntiles,
                         ncols, part, split0);
    return hipGetLastError();
  }
  if (impl == 4 || impl == 5) {
    const int wn = impl == 5 ? 4 : 2;
    if (sps % 2 == 0 && ncols % (64 * wn) == 0) {
      if (impl == 4)
        launch_k((hash_mfma_ring_kernel<2, 2>), dim3(ntiles * nsplits), dim3(256), 0, st, k0, k1, enc, ld, ad, sps,
                           ntiles, ncols, part, split0);
      else
        launch_k((hash_mfma_ring_kernel<4, 2>), dim3((ncols / 256) * nsplits), dim3(512), 0, st, k0, k1, enc, ld,
                           ad, sps, ncols / 256, ncols, part, split0);
    } else {
      launch_k((hash_mfma_pipe_kernel<1, 2>), dim3(ntiles * nsplits), dim3(256), 0, st, k0, k1, enc, ld, ad, sps,
                         ntiles, ncols, part, split0);
    }
  } else if (impl == 6 && ncols % 128 == 0) {  // 64 x 64 wave tiles, 8 waves, 256 x 128 workgroup tile
    launch_k((hash_mfma_tile_kernel<4, 4, 2, 2>), dim3((ncols / 128) * nsplits), dim3(512), 0, st, k0, k1, enc, ld,
                       ad, sps, ncols / 128, ncols, part, split0);
  } else if (impl == 17 && ncols % 256 == 0 && sps % 8 == 0) {  // 8 slots, one barrier per two slabs
    launch_k((hash_mfma_pipe8_kernel<8, false, 2>), dim3((ncols / 256) * nsplits), dim3(512),
                       full_cu_pad((const void*)hash_mfma_pipe8_kernel<8, false, 2>), st, k0, k1, enc, ld, ad, sps,
                       ncols / 256, ncols, part, split0);
  } else if (impl >= 16 && ncols % 256 == 0 && sps % 8 == 0) {  // pipelined 8-wave tile, 8 slots
    launch_k((hash_mfma_pipe8_kernel<8>), dim3((ncols / 256) * nsplits), dim3(512),
                       full_cu_pad((const void*)hash_mfma_pipe8_kernel<8>), st, k0, k1, enc, ld, ad, sps, ncols / 256, ncols,
                       part, split0);
  } else if (impl >= 15 && ncols % 256 == 0 && sps % 4 == 0) {  // pipelined, 4 slots
    launch_k((hash_mfma_pipe8_kernel<4>), dim3((ncols / 256) * nsplits), dim3(512),
                       full_cu_pad((const void*)hash_mfma_pipe8_kernel<4>), st, k0, k1, enc, ld, ad, sps, ncols / 256, ncols,
                       part, split0);
  } else if (impl >= 14 && ncols % 256 == 0 && sps % 4 == 0) {  // 8 waves sharing the A ring
    launch_k((hash_mfma_dma_kernel<4, 4>), dim3((ncols / 256) * nsplits), dim3(512),
                       full_cu_pad((const void*)hash_mfma_dma_kernel<4, 4>), st, k0, k1, enc, ld, ad, sps, ncols / 256, ncols,
                       part, split0);
  } else if (hash_impl_reduces(impl)) {  // asynchronous copies into an LDS ring of 3 / 4 slabs
    if (impl == 11 || sps % 4)
      launch_k((hash_mfma_dma_kernel<3, 2>), dim3(ntiles * nsplits), dim3(256),
                         full_cu_pad((const void*)hash_mfma_dma_kernel<3, 2>), st, k0, k1, enc,
                         ld, ad, sps, ntiles, ncols, part, split0);
    else
      launch_k((hash_mfma_dma_kernel<4, 2>), dim3(ntiles * nsplits), dim3(256),
                         full_cu_pad((const void*)hash_mfma_dma_kernel<4, 2>), st, k0, k1, enc,
                         ld, ad, sps, ntiles, ncols, part, split0);
  } else if (impl == 10) {  // as 0, A fragments read one product group ahead
    launch_k((hash_mfma_pipe_kernel<2, 2, true>), dim3(ntiles * nsplits), dim3(256), 0, st, k0, k1, enc, ld, ad,
                       sps, ntiles, ncols, part, split0);
  } else if (impl == 8 && ncols % 128 == 0) {  // 64 x 128 wave tiles, 4 waves, 256 x 128 workgroup tile
    launch_k((hash_mfma_tile_kernel<4, 8, 1, 2>), dim3((ncols / 128) * nsplits), dim3(256), 0, st, k0, k1, enc, ld,
                       ad, sps, ncols / 128, ncols, part, split0);
  } else if (impl == 9 && ncols % 256 == 0) {  // 64 x 128 wave tiles, 8 waves, 256 x 256 workgroup tile
    launch_k((hash_mfma_tile_kernel<4, 8, 2, 1>), dim3((ncols / 256) * nsplits), dim3(512), 0, st, k0, k1, enc, ld,
                       ad, sps, ncols / 256, ncols, part, split0);
  } else if (impl == 7 && ncols % 256 == 0) {  // 64 x 64 wave tiles, 16 waves, 256 x 256 workgroup tile
    launch_k((hash_mfma_tile_kernel<4, 4, 4, 1>), dim3((ncols / 256) * nsplits), dim3(1024), 0, st, k0, k1, enc,
                       ld, ad, sps, ncols / 256, ncols, part, split0);
  } else if (impl == 3 && ncols % 256 == 0) {
    launch_k((hash_mfma_pipe_kernel<2, 4>), dim3((ncols / 256) * nsplits), dim3(512), 0, st, k0, k1, enc, ld, ad,
                       sps, ncols / 256, ncols, part, split0);
  } else if (impl == 2) {
    launch_k(hash_mfma_kernel, dim3(ntiles * nsplits), dim3(256), 0, st, k0, k1, enc, ld, ad, sps, ntiles, ncols,
                       part, split0);
  } else if (impl == 1) {
    launch_k((hash_mfma_pipe_kernel<1, 2>), dim3(ntiles * nsplits), dim3(256), 0, st, k0, k1, enc, ld, ad, sps,
                       ntiles, ncols, part, split0);
  } else {
    launch_k((hash_mfma_pipe_kernel<2, 2>), dim3(ntiles * nsplits), dim3(256), 0, st, k0, k1, enc, ld, ad, sps,
                       ntiles, ncols, part, split0);
  }
  return hipGetLastError();
}

bool hash_impl_reduces(int impl) { return impl == 11 || impl == 12 || (impl >= 14 && impl <= 17); }

hipError_t launch_hash_mfma_finalize(const int32_t* part, uint32_t ksplit, uint32_t ncols, uint64_t* h,
                                     hipStream_t st, int impl, bool ring) {
  const uint64_t n = (uint64_t)32 * ncols;
  if (ring && !hash_impl_reduces(impl)) return hipErrorInvalidValue;
  if (hash_impl_reduces(impl))  // one word per (split, row, column), already reduced
    hipLaunchKernelGGL(hash_valu_finalize_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st,
                       (const uint64_t*)part, ksplit, ncols, h, ring ? 1 : 0);
  else
    hipLaunchKernelGGL(hash_mfma_finalize_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, part, ksplit,
                       ncols, h);
  return hipGetLastError();
}

hipError_t launch_hash_mfma(const uint64_t* enc, uint64_t ld, uint32_t m_rows, uint32_t ncols,
                            const int8_t* dig, uint32_t ksplit, int32_t* part, uint64_t* h,
                            hipStream_t st, hipEvent_t mid, int impl, bool ring, hipEvent_t k0, hipEvent_t k1) {
  hipError_t e = launch_hash_mfma_splits(enc, ld, m_rows, ncols, dig, ksplit, 0, ksplit, part, st, impl, ring, k0, k1);
  if (e != hipSuccess) return e;
  if (mid && (e = hipEventRecord(mid, st)) != hipSuccess) return e;
  if (!h) return hipSuccess;
  return launch_hash_mfma_finalize(part, ksplit, ncols, h, st, impl, ring);
}

__global__ void hash_finalize_group_kernel(const FinalizeJob* __restrict__ jobs, int njobs) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (jobs[mid].first_block <= blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const FinalizeJob& jb = jobs[lo];
  const uint64_t idx = (uint64_t)(blockIdx.x - jb.first_block) * blockDim.x + threadIdx.x;
  if (idx >= (uint64_t)32 * jb.ncols) return;
  // the batched hash kernel leaves one reduced word per (split, row, column)
  const uint64_t* part = (const uint64_t*)jb.part;
  uint64_t acc = 0;
  for (uint32_t sp = 0; sp < jb.ksplit; sp++) acc = gl_add(acc, part[(uint64_t)sp * 32 * jb.ncols + idx]);
  jb.h[idx] = acc;
}

hipError_t launch_hash_group(const HashJob* jobs, int njobs, uint32_t total_tiles, const FinalizeJob* fjobs,
                             uint32_t total_fblocks, hipStream_t st, int tile_cols, bool pipe8) {
  if (total_tiles) {
    if (tile_cols == 256 && pipe8)
      hipLaunchKernelGGL(hash_mfma_pipe8_group_kernel, dim3(total_tiles), dim3(512),
                         full_cu_pad((const void*)hash_mfma_pipe8_group_kernel), st, jobs, njobs);
    else if (tile_cols == 256)
      hipLaunchKernelGGL((hash_mfma_dma_group_kernel<4>), dim3(total_tiles), dim3(512),
                         full_cu_pad((const void*)hash_mfma_dma_group_kernel<4>), st, jobs,
                         njobs);
    else
      hipLaunchKernelGGL((hash_mfma_dma_group_kernel<2>), dim3(total_tiles), dim3(256),
                         full_cu_pad((const void*)hash_mfma_dma_group_kernel<2>), st, jobs,
                         njobs);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (total_fblocks)
    hipLaunchKernelGGL(hash_finalize_group_kernel, dim3(total_fblocks), dim3(256), 0, st, fjobs, njobs);
  return hipGetLastError();
}

}  // namespace lgs
