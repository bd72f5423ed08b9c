// Batched natural-order radix-2 DFT / Reed-Solomon low-degree extension over Goldilocks.
//
// Restates the per-row encoding of pcs/src/ligesis.rs:231-235 (and the chunk encodings of
// pcs/src/basefold.rs:351-358, 371-378): each row of length n_in is zero-padded to N and
// transformed to out[j] = sum_k c_k w^(jk), w = two_adic_generator(log2 N).
//
// Layout: input element k of row r is in[r * rs + k * es] (row-major rows: rs = n, es = 1;
// Basefold chunks f[j :: 2^li]: rs = 1, es = 2^li). Output rows are dense (row stride N).
//
// Small transforms (N <= LGS_LDS_MAX) run entirely in LDS, one workgroup per row: the
// bit-reversed load is a coalesced global read scattered into LDS, then log2 N butterfly
// stages, then a coalesced store. Larger transforms run the first log2 T stages in LDS per
// T-element block of the bit-reversed array and the remaining stages as global passes.
#include "kernels.hpp"

namespace lgs {

__device__ __forceinline__ uint32_t brev_bits(uint32_t x, int bits) {
  return __brev(x) >> (32 - bits);
}

// One workgroup per (row, block). Each block holds T = 2^lt consecutive positions of the
// bit-reversed array of one row and runs stages 1..lt on them.
__global__ __launch_bounds__(256) void ntt_lds_kernel(const uint64_t* __restrict__ in, uint64_t rs,
                                                      uint64_t es, uint32_t n_in, int logN, int lt,
                                                      const uint64_t* __restrict__ tw,
                                                      uint64_t* __restrict__ out) {
  extern __shared__ uint64_t sm[];
  const uint32_t T = 1u << lt;
  const uint32_t nblk = 1u << (logN - lt);
  const uint32_t row = blockIdx.x / nblk;
  const uint32_t blk = blockIdx.x % nblk;
  const uint64_t* src = in + (uint64_t)row * rs;
  if (lt == logN) {
    // whole row: coalesced read of the n_in inputs, scatter to bit-reversed LDS slots
    for (uint32_t k = threadIdx.x; k < T; k += blockDim.x) {
      uint64_t v = k < n_in ? src[(uint64_t)k * es] : 0;
      sm[brev_bits(k, logN)] = v;
    }
  } else {
    // position q = blk*T + t of the bit-reversed array holds x[brev(q)]
    for (uint32_t t = threadIdx.x; t < T; t += blockDim.x) {
      uint32_t k = brev_bits(blk * T + t, logN);
      sm[t] = k < n_in ? src[(uint64_t)k * es] : 0;
    }
  }
  __syncthreads();
  for (int s = 1; s <= lt; s++) {
    const uint32_t half = 1u << (s - 1);
    const int tshift = logN - s;
    for (uint32_t b = threadIdx.x; b < T / 2; b += blockDim.x) {
      uint32_t j = b & (half - 1);
      uint32_t i0 = ((b >> (s - 1)) << s) | j;
      uint32_t i1 = i0 + half;
      uint64_t u = sm[i0];
      uint64_t v = sm[i1];
      if (j) v = gl_mul(v, tw[(uint64_t)j << tshift]);
      sm[i0] = gl_add(u, v);
      sm[i1] = gl_sub(u, v);
    }
    __syncthreads();
  }
  uint64_t* dst = out + (uint64_t)row * ((uint64_t)1 << logN) + (uint64_t)blk * T;
  for (uint32_t t = threadIdx.x; t < T; t += blockDim.x) dst[t] = sm[t];
}

// ---------------- radix-8 row transform (whole row in LDS) ----------------
// Roots of unity of order <= 64 are powers of two in Goldilocks: w_64 = 2^39 and
// w_{2^k} = 2^(39 * 2^(6-k) mod 192) (2^96 = -1, 2^192 = 1). Multiplying by them is a shift and
// a reduction, so the internal twiddles of an 8-point DFT (fixed exponents) cost no general
// multiplication.

// (hi 2^64 + lo) mod p for hi < 2^32: lo + hi (2^32 - 1), one carry fix, then canonical.
__device__ __forceinline__ uint64_t gl_reduce_small(uint64_t lo, uint64_t hi) {
  const uint64_t t = (hi << 32) - hi;
  uint64_t r = lo + t;
  r += (r < t) ? GL_EPS : 0;  // a wrap is 2^64 = 2^32 - 1; no second wrap (t < 2^64 - 2^33)
  return gl_canon(r);
}

// x * 2^e mod p for 0 <= e < 192 (e is a compile-time constant where this is inlined), by the
// cheapest exact form: 2^96 = -1 gives 2^e = -2^(e - 96); for 64 <= e < 96, 2^e = -2^-(96 - e),
// a division by at most 2^32 (gl_mul_inv_pow2); for 32 < e < 64 a word rotation and one
// two-word subtraction (gl_mul_pow2_mid); no 128-bit reduction in any case.
__device__ __forceinline__ uint64_t gl_mul_pow2(uint64_t x, uint32_t e) {
  bool neg = e >= 96;
  if (neg) e -= 96;
  uint64_t r;
  if (e == 0) {
    r = x;
  } else if (e <= 32) {
    r = gl_reduce_small(x << e, x >> (64 - e));
  } else if (e < 64) {
    r = gl_mul_pow2_mid(x, e);
  } else {
    r = gl_mul_inv_pow2(x, 96 - e);
    neg = !neg;
  }
  return neg ? gl_neg(r) : r;
}

// w_N^k for k < N from the half table tw[k] = w^k (k < N/2), w^(k + N/2) = -w^k. The table may
// belong to a larger transform 2^tsh N (entries taken with stride 2^tsh).
__device__ __forceinline__ uint64_t tw_full(const uint64_t* __restrict__ tw, uint32_t k, uint32_t half,
                                            int tsh) {
  const uint64_t v = tw[(uint64_t)(k & (half - 1)) << tsh];  // half is a power of two
  return (k & half) ? gl_neg(v) : v;
}

// In-place radix-2 DIT network on 8 values held in bit-reversed order; output natural order.
// ODD_ZERO: x[1], x[3], x[5], x[7] are known to be zero (the upper half of a zero-padded row),
// so the first level is copies.
template <bool ODD_ZERO = false>
__device__ __forceinline__ void dft8(uint64_t (&x)[8]) {
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const uint64_t a = x[2 * u], b = x[2 * u + 1];
    x[2 * u] = ODD_ZERO ? a : gl_add64(a, b);
    x[2 * u + 1] = ODD_ZERO ? a : gl_sub(a, b);
  }
#pragma unroll
  for (int h = 0; h < 8; h += 4) {
    uint64_t a = x[h], b = x[h + 2];
    x[h] = gl_add64(a, b);
    x[h + 2] = gl_sub(a, b);
    a = x[h + 1];
    b = gl_mul_pow2(x[h + 3], 48);  // w_4
    x[h + 1] = gl_add64(a, b);
    x[h + 3] = gl_sub(a, b);
  }
  {
    uint64_t a = x[0], b = x[4];
    x[0] = gl_add64(a, b);
    x[4] = gl_sub(a, b);
    a = x[1];
    b = gl_mul_pow2(x[5], 24);  // w_8 = 2^120 = -2^24: the sign swaps the butterfly's outputs
    x[1] = gl_sub(a, b);
    x[5] = gl_add64(a, b);
    a = x[2];
    b = gl_mul_pow2(x[6], 48);  // w_8^2
    x[2] = gl_add64(a, b);
    x[6] = gl_sub(a, b);
    a = x[3];
    b = gl_mul_pow2(x[7], 168);  // w_8^3 = 2^-24
    x[3] = gl_add64(a, b);
    x[7] = gl_sub(a, b);
  }
}

// Multiply x[t] by c^rev(t) where c = w_{2^(s+lg)}^j and rev is the lg-bit reversal. Always a
// table twiddle and a full multiply: the roots of order <= 64 are powers of two, but their
// exponent depends on j, so a shift-based product would branch differently in every lane.
template <int LG>
__device__ __forceinline__ void pretwiddle(uint64_t (&x)[1 << LG], uint32_t j, int s, int logN,
                                           const uint64_t* __restrict__ tw, int tsh) {
  constexpr int R = 1 << LG;
  const int lsz = s + LG;  // log2 of this pass's transform size
  const uint32_t half = 1u << (logN - 1);
#pragma unroll
  for (int t = 1; t < R; t++) {
    const uint32_t r = __brev((uint32_t)t) >> (32 - LG);
    x[t] = gl_mul(x[t], tw_full(tw, (j * r) << (logN - lsz), half, tsh));
  }
}

// LDS word of row position w: the low 5 bits XOR the top 5 bits of the row index (sh = logN - 5)
// and bits 3-4 also XOR bits 6-7 (sh = 31 leaves w alone). The first pass scatters to
// bit-reversed blocks, where 32 consecutive lanes differ only in the top bits; the pass over
// 8-point groups (stride 8) has lanes 8 apart in bits 6-7; both would otherwise hit few banks.
// Only bits 0-4 change, as a function of bits >= 5, so this is a permutation, and runs of 32
// consecutive positions (the other passes) stay a permutation of one bank row.
__device__ __forceinline__ uint32_t lds_pos(uint32_t w, int sh) {
  return sh == 31 ? w : w ^ ((w >> sh) & 31u) ^ (((w >> 6) & 3u) << 3);
}
__host__ __device__ constexpr size_t ntt_lds_bytes(uint32_t N) { return (size_t)N * sizeof(uint64_t); }

// One row: src[k * es] for k < n_in (zero above), transform size 2^logN, result to dst; tw is
// the twiddle table of a transform of size 2^(logN + tsh).
__device__ __forceinline__ void ntt_r8_row(const uint64_t* __restrict__ src, uint64_t es, uint32_t n_in,
                                           int logN, const uint64_t* __restrict__ tw, int tsh,
                                           uint64_t* __restrict__ dst, uint64_t* sm) {
  const uint32_t N = 1u << logN;
  const int sh = logN >= 10 ? logN - 5 : 31;
  int s = 0;
  if (n_in <= N / 2 && logN >= 7 && (logN - 1) % 3 == 0) {
    // Zero upper half (RS padding): DIT stage 1 pairs a[2i] = x[rev(2i)] with a[2i + 1], a value
    // of the zero half, so it leaves a[2i] = a[2i + 1] = x[rev(2i)]. The first pass then does
    // stages 2-4: group (g, j) holds positions 16 g + 2 t + j (t < 8), whose stage-1 values are
    // x[rev(16 g + 2 t)] = x[k + rev3(t) N / 16] with k = rev_{logN-4}(g) -- consecutive threads
    // load consecutive words -- pretwiddled by w_16^(j rev3(t)), a power of two. Stages 5..logN
    // are the radix-8 passes from s = 4, the last one writing dst: one pass fewer than
    // starting from stage 1, and no radix-2 pass.
    const uint32_t E = N / 16;
    for (uint32_t gi = threadIdx.x; gi < N / 8; gi += blockDim.x) {
      const uint32_t k = gi & (E - 1), j = gi / E;  // j uniform per half of the threads
      uint64_t x[8];
#pragma unroll
      for (int t = 0; t < 8; t++) {
        const uint32_t u = __brev((uint32_t)t) >> 29;
        const uint32_t idx = k + u * E;
        x[t] = idx < n_in ? src[(uint64_t)idx * es] : 0;
      }
      if (j) {
#pragma unroll
        for (int t = 1; t < 8; t++) {
          const uint32_t u = __brev((uint32_t)t) >> 29;
          x[t] = gl_mul_pow2(x[t], (156u * u) % 192u);  // w_16 = 2^156
        }
      }
      dft8(x);
      const uint32_t g = brev_bits(k, logN - 4);
#pragma unroll
      for (int t = 0; t < 8; t++) sm[lds_pos(16 * g + 2 * t + j, sh)] = x[t];
    }
    s = 4;
  } else if (logN >= 3) {
    // first pass fused with the load: thread k takes x[k + r N/8] (r = 0..7), which sit at
    // positions rev3(r) of block rev_{logN-3}(k) of the bit-reversed array
    const uint32_t E = N / 8;
    const bool last = logN == 3;
    // RS encoding pads every row to twice its length: then the odd slots (r = rev3(t) >= 4,
    // idx >= N / 2) are all zero
    const bool half_zero = n_in <= N / 2;
    for (uint32_t k = threadIdx.x; k < E; k += blockDim.x) {
      uint64_t x[8];
#pragma unroll
      for (int t = 0; t < 8; t++) {
        const uint32_t r = __brev((uint32_t)t) >> 29;
        const uint32_t idx = k + r * E;
        x[t] = (idx < n_in && !(half_zero && (t & 1))) ? src[(uint64_t)idx * es] : 0;
      }
      if (half_zero)
        dft8<true>(x);
      else
        dft8(x);
      const uint32_t b = logN > 3 ? brev_bits(k, logN - 3) : 0;
#pragma unroll
      for (int t = 0; t < 8; t++) {
        if (last) dst[8 * b + t] = x[t];
        else sm[lds_pos(8 * b + t, sh)] = x[t];
      }
    }
    s = 3;
  } else {
    for (uint32_t k = threadIdx.x; k < N; k += blockDim.x)
      sm[lds_pos(brev_bits(k, logN), sh)] = k < n_in ? src[(uint64_t)k * es] : 0;
  }
  __syncthreads();
  while (logN - s >= 3) {
    const bool last = logN - s == 3;
    const uint32_t M = 1u << s;
    for (uint32_t gi = threadIdx.x; gi < N / 8; gi += blockDim.x) {
      const uint32_t j = gi & (M - 1);
      const uint32_t base = ((gi >> s) << (s + 3)) + j;
      uint64_t x[8];
#pragma unroll
      for (int t = 0; t < 8; t++) x[t] = sm[lds_pos(base + t * M, sh)];
      pretwiddle<3>(x, j, s, logN, tw, tsh);  // j = 0: factors of one (no divergent branch)
      dft8(x);
#pragma unroll
      for (int t = 0; t < 8; t++) {
        if (last) dst[base + t * M] = x[t];
        else sm[lds_pos(base + t * M, sh)] = x[t];
      }
    }
    if (!last) __syncthreads();
    s += 3;
  }
  if (logN - s == 2) {
    const uint32_t M = 1u << s;
    for (uint32_t gi = threadIdx.x; gi < N / 4; gi += blockDim.x) {
      const uint32_t j = gi & (M - 1);
      const uint32_t base = ((gi >> s) << (s + 2)) + j;
      uint64_t x[4];
#pragma unroll
      for (int t = 0; t < 4; t++) x[t] = sm[lds_pos(base + t * M, sh)];
      if (j) pretwiddle<2>(x, j, s, logN, tw, tsh);
      uint64_t a0 = gl_add(x[0], x[1]), a1 = gl_sub(x[0], x[1]);
      uint64_t a2 = gl_add(x[2], x[3]), a3 = gl_mul_pow2(gl_sub(x[2], x[3]), 48);
      dst[base] = gl_add(a0, a2);
      dst[base + 2 * M] = gl_sub(a0, a2);
      dst[base + M] = gl_add(a1, a3);
      dst[base + 3 * M] = gl_sub(a1, a3);
    }
  } else if (logN - s == 1) {
    const uint32_t M = 1u << s;
    for (uint32_t j = threadIdx.x; j < M; j += blockDim.x) {
      uint64_t a = sm[lds_pos(j, sh)], b = sm[lds_pos(j + M, sh)];
      if (j) b = gl_mul(b, tw[(uint64_t)j << (logN - s - 1 + tsh)]);
      dst[j] = gl_add(a, b);
      dst[j + M] = gl_sub(a, b);
    }
  }
}

__global__ __launch_bounds__(1024) void ntt_r8_kernel(const uint64_t* __restrict__ in, uint64_t rs,
                                                     uint64_t es, uint32_t n_in, int logN,
                                                     const uint64_t* __restrict__ tw,
                                                     uint64_t* __restrict__ out) {
  extern __shared__ uint64_t sm[];
  const uint64_t row = blockIdx.x;
  ntt_r8_row(in + row * rs, es, n_in, logN, tw, 0, out + (row << logN), sm);
}

// Pass twiddles of a 2^L-point transform whose radix-8 passes start at s = 4 (L = 1 mod 3):
// for s = 4, 7, ..., L - 3, section s holds tp[off_s + (t - 1) 2^s + j] = w_{2^(s+3)}^(j rev3(t))
// (t = 1..7, j < 2^s) -- the factors pretwiddle<3> takes from tw -- in Montgomery form (times
// 2^64 mod p, for gl_mul_mont), so that the 64 lanes of a wave (consecutive j) read 512
// contiguous bytes per factor.
__global__ void pass_twiddle_kernel(uint64_t* __restrict__ tp, const uint64_t* __restrict__ tw, int L) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t off = 0;
  for (int s = 4; L - s >= 3; s += 3) {
    const uint32_t M = 1u << s;
    if (i < off + 7 * M) {
      const uint32_t q = i - off, t = q / M + 1, j = q & (M - 1);
      const uint32_t r = __brev(t) >> 29;
      tp[i] = gl_mul(tw_full(tw, (j * r) << (L - s - 3), 1u << (L - 1), 0), GL_EPS);  // 2^64 = 2^32 - 1
      return;
    }
    off += 7 * M;
  }
}

static uint32_t pass_twiddle8_entries(int L) {
  uint32_t n = 0;
  if ((L - 1) % 3 == 0)
    for (int s = 4; L - s >= 3; s += 3) n += 7u << s;
  return n;
}

// LDS word at byte address b (a word's position times 8): the fixed-shape kernels keep a group's
// base position and its offsets as byte addresses, so an access is one XOR (or none, where the
// compiler folds an offset into the instruction) instead of an XOR and a shift.
__device__ __forceinline__ uint64_t& lds_word(uint64_t* sm, uint32_t b) {
  return *(uint64_t*)((char*)sm + b);
}

// The row LDE of rows of 2^(L-1) values (row stride rs, unit element stride) zero-padded to
// 2^L points, one workgroup of 2^(L-3) threads per row: ntt_r8_row's zero-half schedule (stages
// 2-4 fused with the load, then radix-8 passes from s = 4) with every size a compile-time
// constant, one butterfly group per thread and pass, Montgomery-form twiddles from the pass
// table tp (one Montgomery reduction per product instead of a 128-bit one), and each
// LDS address one XOR of byte addresses: lds_pos is linear over GF(2), and a group's base and its
// offsets t 2^s occupy disjoint bits, so lds_pos(base + t 2^s) = lds_pos(base) ^ lds_pos(t 2^s).
template <int L>
__device__ __forceinline__ void ntt_r8_fixed_row(const uint64_t* __restrict__ src, const uint64_t* __restrict__ tp,
                                                 uint64_t* __restrict__ dst, uint64_t* sm, uint64_t ss = 1) {
  static_assert((L - 1) % 3 == 0 && L >= 10 && L <= LGS_LDS_LOG, "radix-8 passes from s = 4");
  constexpr uint32_t N = 1u << L, E = N / 16;
  constexpr int sh = L - 5;
  const uint32_t gi = threadIdx.x;
  {
    const uint32_t k = gi & (E - 1), j = gi / E;  // j uniform per half of the workgroup
    uint64_t x[8];
#pragma unroll
    for (int t = 0; t < 8; t++) x[t] = src[(k + (__brev((uint32_t)t) >> 29) * E) * ss];
    if (j) {
#pragma unroll
      for (int t = 1; t < 8; t++) x[t] = gl_mul_pow2(x[t], (156u * (__brev((uint32_t)t) >> 29)) % 192u);
    }
    dft8(x);
    const uint32_t pb = lds_pos(16 * brev_bits(k, L - 4) + j, sh) * 8;
#pragma unroll
    for (int t = 0; t < 8; t++) lds_word(sm, pb ^ (lds_pos(2 * t, sh) * 8)) = x[t];
  }
  __syncthreads();
  uint32_t off = 0;
#pragma unroll
  for (int s = 4; s < L; s += 3) {
    const uint32_t M = 1u << s;
    const uint32_t j = gi & (M - 1);
    const uint32_t base = ((gi >> s) << (s + 3)) | j;
    const uint32_t pb = lds_pos(base, sh) * 8;
    uint64_t x[8];
#pragma unroll
    for (int t = 0; t < 8; t++) x[t] = lds_word(sm, pb ^ (lds_pos(t << s, sh) * 8));
#pragma unroll
    for (int t = 1; t < 8; t++) x[t] = gl_mul_mont(x[t], tp[off + (t - 1) * M + j]);
    dft8(x);
    if (s == L - 3) {
#pragma unroll
      for (int t = 0; t < 8; t++) dst[base + t * M] = x[t];
    } else {
#pragma unroll
      for (int t = 0; t < 8; t++) lds_word(sm, pb ^ (lds_pos(t << s, sh) * 8)) = x[t];
      __syncthreads();
    }
    off += 7 * M;
  }
}

template <int L>
__global__ __launch_bounds__(1 << (L - 3)) void ntt_r8_fixed_kernel(const uint64_t* __restrict__ in, uint64_t rs,
                                                                   const uint64_t* __restrict__ tp,
                                                                   uint64_t* __restrict__ out) {
  extern __shared__ uint64_t sm[];
  ntt_r8_fixed_row<L>(in + (uint64_t)blockIdx.x * rs, tp, out + ((uint64_t)blockIdx.x << L), sm);
}

// The first pass of a long row transform (2^logN points, the first 2^(logN - 1) of them values)
// with 2^L-point subsequence transforms: block b is subsequence c = rev(b mod 2^(logN - L)) of
// row b / 2^(logN - L), whose 2^(L-1) values sit at stride 2^(logN - L) es; its DFT goes to block
// b mod 2^(logN - L) of the bit-reversed-block layout that the column pass completes. The
// compile-time-shaped row kernel with its pass table, in place of the runtime-shaped one.
template <int L>
__global__ __launch_bounds__(1 << (L - 3)) void ntt_r8_fixed_sub_kernel(const uint64_t* __restrict__ in, uint64_t rs,
                                                                       uint64_t es, int logN,
                                                                       const uint64_t* __restrict__ tp,
                                                                       uint64_t* __restrict__ out) {
  extern __shared__ uint64_t sm[];
  const uint32_t nblk = 1u << (logN - L);
  const uint64_t row = blockIdx.x / nblk;
  const uint32_t blk = blockIdx.x % nblk;
  const uint32_t c = brev_bits(blk, logN - L);
  ntt_r8_fixed_row<L>(in + row * rs + (uint64_t)c * es, tp, out + (row << logN) + ((uint64_t)blk << L), sm,
                      (uint64_t)nblk * es);
}

// The rows of many commitments whose rows hold 2^(L-1) values at unit stride, in one launch
// (batched commit): block b is row (b - first_block) of the job whose range holds b.
template <int L>
__global__ __launch_bounds__(1 << (L - 3)) void ntt_r8_fixed_group_kernel(const NttJob* __restrict__ jobs, int njobs,
                                                                         const uint64_t* __restrict__ tp) {
  extern __shared__ uint64_t sm[];
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (jobs[mid].first_block <= blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const NttJob& jb = jobs[lo];
  const uint64_t row = blockIdx.x - jb.first_block;
  ntt_r8_fixed_row<L>(jb.in + row * jb.rs, tp, jb.out + (row << L), sm);
}

// ---------------- radix-16 row transform for rows of 2^(L-1) values (L = 1 mod 4) ----------------
// b * 2^E folded into the butterfly (a, b) -> (a + b 2^E, a - b 2^E) for a compile-time E < 192:
// 2^96 = -1 turns E >= 96 into a swap of the two outputs, and 64 <= E < 96 into a division by
// 2^(96 - E) <= 2^32 with the opposite sign.
template <uint32_t E>
__device__ __forceinline__ void bfly_pow2(uint64_t& a, uint64_t& b) {
  constexpr bool neg0 = E >= 96;
  constexpr uint32_t e = neg0 ? E - 96 : E;
  constexpr bool neg = e >= 64 ? !neg0 : neg0;
  uint64_t c;
  if constexpr (e == 0) c = b;
  else if constexpr (e <= 32) c = gl_reduce_small(b << e, b >> (64 - e));
  else if constexpr (e < 64) c = gl_mul_pow2_mid(b, e);
  else c = gl_mul_inv_pow2(b, 96 - e);
  const uint64_t s = gl_add64(a, c), d = gl_sub(a, c);
  a = neg ? d : s;
  b = neg ? s : d;
}

// 16-point DFT of values held in bit-reversed order (radix-2 DIT network, w_16 = 2^156 of the
// order-2^13 generator's powers: w_2^k = 2^(39 2^(6-k) mod 192)); output in natural order.
__device__ __forceinline__ void dft16(uint64_t (&x)[16]) {
#pragma unroll
  for (int u = 0; u < 8; u++) bfly_pow2<0>(x[2 * u], x[2 * u + 1]);
#pragma unroll
  for (int h = 0; h < 16; h += 4) {
    bfly_pow2<0>(x[h], x[h + 2]);
    bfly_pow2<48>(x[h + 1], x[h + 3]);  // w_4 = 2^48
  }
#pragma unroll
  for (int h = 0; h < 16; h += 8) {
    bfly_pow2<0>(x[h], x[h + 4]);
    bfly_pow2<120>(x[h + 1], x[h + 5]);  // w_8 = 2^120
    bfly_pow2<48>(x[h + 2], x[h + 6]);
    bfly_pow2<168>(x[h + 3], x[h + 7]);
  }
  bfly_pow2<0>(x[0], x[8]);
  bfly_pow2<156>(x[1], x[9]);  // w_16 = 2^156
  bfly_pow2<120>(x[2], x[10]);
  bfly_pow2<84>(x[3], x[11]);
  bfly_pow2<48>(x[4], x[12]);
  bfly_pow2<12>(x[5], x[13]);
  bfly_pow2<168>(x[6], x[14]);
  bfly_pow2<132>(x[7], x[15]);
}

// Pass twiddles of a 2^L-point transform whose radix-16 passes start at s = 5 (L = 1 mod 4): for
// s = 5, 9, ..., L - 4, section s holds tp[off_s + (t - 1) 2^s + j] = w_{2^(s+4)}^(j rev4(t))
// (t = 1..15, j < 2^s) in Montgomery form.
__global__ void pass_twiddle16_kernel(uint64_t* __restrict__ tp, const uint64_t* __restrict__ tw, int L) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t off = 0;
  for (int s = 5; L - s >= 4; s += 4) {
    const uint32_t M = 1u << s;
    if (i < off + 15 * M) {
      const uint32_t q = i - off, t = q / M + 1, j = q & (M - 1);
      const uint32_t r = __brev(t) >> 28;
      tp[i] = gl_mul(tw_full(tw, (j * r) << (L - s - 4), 1u << (L - 1), 0), GL_EPS);  // 2^64 = 2^32 - 1
      return;
    }
    off += 15 * M;
  }
}

static uint32_t pass_twiddle16_entries(int L) {
  uint32_t n = 0;
  if ((L - 1) % 4 == 0)
    for (int s = 5; L - s >= 4; s += 4) n += 15u << s;
  return n;
}

uint32_t pass_twiddle_entries(int L, int impl) {
  return impl == 4 ? pass_twiddle8_entries(L) : impl == 5 ? pass_twiddle16_entries(L) : 0;
}

hipError_t launch_pass_twiddles(uint64_t* tp, const uint64_t* tw, int L, int impl, hipStream_t st) {
  const uint32_t n = pass_twiddle_entries(L, impl);
  if (n == 0) return hipErrorInvalidValue;
  if (impl == 4)
    hipLaunchKernelGGL(pass_twiddle_kernel, dim3((n + 255) / 256), dim3(256), 0, st, tp, tw, L);
  else
    hipLaunchKernelGGL(pass_twiddle16_kernel, dim3((n + 255) / 256), dim3(256), 0, st, tp, tw, L);
  return hipGetLastError();
}

// The row LDE of rows of 2^(L-1) values zero-padded to 2^L points, one workgroup of 2^(L-4)
// threads per row, 16 values per thread and pass. Stage 1 pairs each value with one of the zero
// half (a copy); the first pass does stages 2-5: group (g, j) holds positions 32 g + 2 t + j
// (t < 16), whose stage-1 values are x[k + rev4(t) N / 32] with k = rev_{L-5}(g), pretwiddled by
// w_32^(j rev4(t)) = 2^(78 j rev4(t) mod 192); then radix-16 passes from s = 5 with Montgomery-form
// table twiddles, the last one writing dst. LDS addressing as in ntt_r8_fixed_kernel.
template <int L>
__global__ __launch_bounds__(1 << (L - 4)) void ntt_r16_fixed_kernel(const uint64_t* __restrict__ in, uint64_t rs,
                                                                    const uint64_t* __restrict__ tp,
                                                                    uint64_t* __restrict__ out) {
  static_assert((L - 1) % 4 == 0 && L >= 13 && L <= LGS_LDS_LOG, "radix-16 passes from s = 5");
  constexpr uint32_t N = 1u << L, E = N / 32;
  constexpr int sh = L - 5;
  extern __shared__ uint64_t sm[];
  const uint64_t* __restrict__ src = in + (uint64_t)blockIdx.x * rs;
  uint64_t* __restrict__ dst = out + ((uint64_t)blockIdx.x << L);
  const uint32_t gi = threadIdx.x;
  {
    const uint32_t k = gi & (E - 1), j = gi / E;  // j uniform per half of the workgroup
    uint64_t x[16];
#pragma unroll
    for (int t = 0; t < 16; t++) x[t] = src[k + (__brev((uint32_t)t) >> 28) * E];
    if (j) {
      x[1] = gl_mul_pow2(x[1], 78u * 8 % 192);
      x[2] = gl_mul_pow2(x[2], 78u * 4 % 192);
      x[3] = gl_mul_pow2(x[3], 78u * 12 % 192);
      x[4] = gl_mul_pow2(x[4], 78u * 2 % 192);
      x[5] = gl_mul_pow2(x[5], 78u * 10 % 192);
      x[6] = gl_mul_pow2(x[6], 78u * 6 % 192);
      x[7] = gl_mul_pow2(x[7], 78u * 14 % 192);
      x[8] = gl_mul_pow2(x[8], 78u * 1 % 192);
      x[9] = gl_mul_pow2(x[9], 78u * 9 % 192);
      x[10] = gl_mul_pow2(x[10], 78u * 5 % 192);
      x[11] = gl_mul_pow2(x[11], 78u * 13 % 192);
      x[12] = gl_mul_pow2(x[12], 78u * 3 % 192);
      x[13] = gl_mul_pow2(x[13], 78u * 11 % 192);
      x[14] = gl_mul_pow2(x[14], 78u * 7 % 192);
      x[15] = gl_mul_pow2(x[15], 78u * 15 % 192);
    }
    dft16(x);
    const uint32_t pb = lds_pos(32 * brev_bits(k, L - 5) + j, sh) * 8;
#pragma unroll
    for (int t = 0; t < 16; t++) lds_word(sm, pb ^ (lds_pos(2 * t, sh) * 8)) = x[t];
  }
  __syncthreads();
  uint32_t off = 0;
#pragma unroll
  for (int s = 5; s < L; s += 4) {
    const uint32_t M = 1u << s;
    const uint32_t j = gi & (M - 1);
    const uint32_t base = ((gi >> s) << (s + 4)) | j;
    const uint32_t pb = lds_pos(base, sh) * 8;
    uint64_t x[16];
#pragma unroll
    for (int t = 0; t < 16; t++) x[t] = lds_word(sm, pb ^ (lds_pos(t << s, sh) * 8));
#pragma unroll
    for (int t = 1; t < 16; t++) x[t] = gl_mul_mont(x[t], tp[off + (t - 1) * M + j]);
    dft16(x);
    if (s == L - 4) {
#pragma unroll
      for (int t = 0; t < 16; t++) dst[base + t * M] = x[t];
    } else {
#pragma unroll
      for (int t = 0; t < 16; t++) lds_word(sm, pb ^ (lds_pos(t << s, sh) * 8)) = x[t];
      __syncthreads();
    }
    off += 15 * M;
  }
}

// Transforms larger than one LDS tile: block (row, blk) computes the size-2^lt transform of the
// subsequence x[c + k 2^(logN-lt)], c = rev_{logN-lt}(blk), which is what the first lt stages of
// an in-place radix-2 DIT leave in positions [blk 2^lt, (blk + 1) 2^lt) of the bit-reversed
// array; the remaining logN - lt stages run as global passes.
__global__ __launch_bounds__(1024) void ntt_r8_blocks_kernel(const uint64_t* __restrict__ in, uint64_t rs,
                                                            uint64_t es, uint32_t n_in, int logN, int lt,
                                                            const uint64_t* __restrict__ tw,
                                                            uint64_t* __restrict__ out) {
  extern __shared__ uint64_t sm[];
  const uint32_t nblk = 1u << (logN - lt);
  const uint64_t row = blockIdx.x / nblk;
  const uint32_t blk = blockIdx.x % nblk;
  const uint32_t c = brev_bits(blk, logN - lt);
  const uint64_t S = uint64_t(1) << (logN - lt);
  const uint32_t n_sub = n_in > c ? (uint32_t)((n_in - c + S - 1) / S) : 0;
  ntt_r8_row(in + row * rs + (uint64_t)c * es, es * S, n_sub, lt, tw, logN - lt,
             out + (row << logN) + ((uint64_t)blk << lt), sm);
}

// Many transforms of different sizes in one launch: block b is row (b - first_block) of the job
// whose [first_block, next first_block) range holds b.
__global__ __launch_bounds__(256) void ntt_r8_group_kernel(const NttJob* __restrict__ jobs, int njobs) {
  extern __shared__ uint64_t sm[];
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (jobs[mid].first_block <= blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const NttJob& jb = jobs[lo];
  const uint64_t row = blockIdx.x - jb.first_block;
  ntt_r8_row(jb.in + row * jb.rs, jb.es, jb.n_in, jb.logN, jb.tw, 0, jb.out + (row << jb.logN), sm);
}

hipError_t launch_ntt_fixed_group(const NttJob* jobs, int njobs, uint32_t total_blocks, const uint64_t* tp,
                                  hipStream_t st) {
  if (total_blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(ntt_r8_fixed_group_kernel<13>, dim3(total_blocks), dim3(1024), ntt_lds_bytes(1u << 13), st, jobs,
                     njobs, tp);
  return hipGetLastError();
}

hipError_t launch_ntt_group(const NttJob* jobs, int njobs, uint32_t total_blocks, int max_logN, hipStream_t st) {
  if (total_blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(ntt_r8_group_kernel, dim3(total_blocks), dim3(256),
                     ntt_lds_bytes(1u << max_logN), st, jobs, njobs);
  return hipGetLastError();
}

// The last K (<= 7) stages of a transform of 2^logN points. The first logN - K stages of the
// in-place DIT left, in block b of T = 2^(logN - K) positions, the T-point DFT Y_c (c = rev_K(b))
// of the subsequence x[c + k 2^K]; the result is X[j + T q] = sum_c w_S^(c q) (w_N^(c j) Y_c[j])
// with S = 2^K: for every j an S-point DFT of twiddled values. A workgroup takes kColJ
// consecutive j of one row: it loads the kColJ x S values (runs of kColJ words), twiddles them,
// keeps each j's sequence in LDS in bit-reversed order, runs K radix-2 DIT stages there and
// stores the natural-order result at stride T (in place: it reads and writes the same words).
constexpr uint32_t kColJ = 8;
__global__ __launch_bounds__(256) void ntt_cols_kernel(uint64_t* __restrict__ data, int logN, int K,
                                                      const uint64_t* __restrict__ tw) {
  __shared__ uint64_t sm[kColJ * 128];
  const uint32_t S = 1u << K;
  const uint64_t T = uint64_t(1) << (logN - K);
  const uint32_t jblocks = (uint32_t)(T / kColJ);
  const uint64_t row = blockIdx.x / jblocks;
  const uint64_t j0 = (uint64_t)(blockIdx.x % jblocks) * kColJ;
  uint64_t* d = data + (row << logN);
  const uint32_t half = 1u << (logN - 1);
  const uint32_t n = kColJ * S;
  for (uint32_t e = threadIdx.x; e < n; e += blockDim.x) {
    const uint32_t t = e / kColJ, jj = e % kColJ;
    uint64_t v = d[(uint64_t)t * T + j0 + jj];
    const uint64_t cj = (uint64_t)brev_bits(t, K) * (j0 + jj);  // < 2^logN
    if (cj) v = gl_mul(v, tw_full(tw, (uint32_t)cj, half, 0));
    sm[jj * S + t] = v;
  }
  __syncthreads();
  for (int s = 1; s <= K; s++) {
    const uint32_t hs = 1u << (s - 1);
    for (uint32_t b = threadIdx.x; b < n / 2; b += blockDim.x) {
      const uint32_t jj = b >> (K - 1), w = b & (S / 2 - 1);
      const uint32_t k = w & (hs - 1), at = jj * S + ((w >> (s - 1)) << s) + k;
      const uint64_t u = sm[at];
      uint64_t v = sm[at + hs];
      if (k) v = gl_mul(v, tw[(uint64_t)k << (logN - s)]);  // w_{2^s}^k
      sm[at] = gl_add(u, v);
      sm[at + hs] = gl_sub(u, v);
    }
    __syncthreads();
  }
  for (uint32_t e = threadIdx.x; e < n; e += blockDim.x) {
    const uint32_t q = e / kColJ, jj = e % kColJ;
    d[(uint64_t)q * T + j0 + jj] = sm[jj * S + q];
  }
}

// The same last K stages for K = 3, 6 or 7 (S = 8, 64 or 128) as radix-8 / radix-16 passes in
// registers: a workgroup of 256 threads takes J = 256 / (S / 8) consecutive j of one row.
// Stages 1-3: a thread loads the 8 values of one group g (t = 8 g .. 8 g + 7, runs of J
// consecutive words across the lanes), twiddles them by w_N^(rev_K(t) j), and runs dft8 on them
// (they sit in bit-reversed order within the group); for K = 3 that is the result. Otherwise the
// groups go through LDS, and stages 4..K are one more DIT pass per (j, group residue r) over
// positions r + 8 u: a dft8 (K = 6) or dft16 (K = 7) pretwiddled by w_{2^K}^(r rev(u)), written
// straight to position r + 8 u of the natural-order result. One general product per value plus
// 7/8 or 15/16 of one, against one per value and per radix-2 stage, and one barrier instead of K.
template <int K>
__global__ __launch_bounds__(256) void ntt_cols8_kernel(uint64_t* __restrict__ data, int logN,
                                                       const uint64_t* __restrict__ tw) {
  static_assert(K == 3 || K == 6 || K == 7, "radix-8 pass, then none, a radix-8 or a radix-16 pass");
  constexpr uint32_t S = 1u << K, G = S / 8, J = 256 / G;
  constexpr uint32_t SP = S + 1;  // a column's LDS stride: odd, so the J columns fall in distinct banks
  constexpr int U = K == 7 ? 16 : 8;  // the second pass's radix
  __shared__ uint64_t sm[K == 3 ? 1 : J * SP];
  const uint64_t T = uint64_t(1) << (logN - K);
  const uint32_t jblocks = (uint32_t)(T / J);
  const uint64_t row = blockIdx.x / jblocks;
  const uint64_t j0 = (uint64_t)(blockIdx.x % jblocks) * J;
  uint64_t* d = data + (row << logN);
  const uint32_t half = 1u << (logN - 1);
  const uint32_t jj = threadIdx.x % J, g = threadIdx.x / J;
  const uint64_t j = j0 + jj;
  uint64_t x[8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const uint32_t t = 8 * g + i;
    const uint64_t v = d[(uint64_t)t * T + j];
    const uint64_t cj = (uint64_t)brev_bits(t, K) * j;  // < 2^logN
    x[i] = cj ? gl_mul(v, tw_full(tw, (uint32_t)cj, half, 0)) : v;
  }
  dft8(x);
  if constexpr (K == 3) {
#pragma unroll
    for (int i = 0; i < 8; i++) d[(uint64_t)i * T + j] = x[i];
  } else {
#pragma unroll
    for (int i = 0; i < 8; i++) sm[jj * SP + 8 * g + i] = x[i];
    __syncthreads();
    if (g < 8) {  // K = 7: the 8 residues of a column take half of its 16 threads
      const uint32_t r = g;  // positions r + 8 u
      uint64_t y[U];
#pragma unroll
      for (int u = 0; u < U; u++) y[u] = sm[jj * SP + r + 8 * u];
#pragma unroll
      for (int u = 1; u < U; u++) {
        const uint32_t e = r * (__brev((uint32_t)u) >> (32 - (K - 3)));  // w_{2^K}^(r rev(u)), e < 2^K
        if (e) y[u] = gl_mul(y[u], tw_full(tw, e << (logN - K), half, 0));
      }
      if constexpr (K == 7) dft16(y);
      else dft8(y);
#pragma unroll
      for (int u = 0; u < U; u++) d[(uint64_t)(r + 8 * u) * T + j] = y[u];
    }
  }
}

// One global radix-2 stage s (> lt) over every row, in place.
__global__ __launch_bounds__(256) void ntt_global_stage_kernel(uint64_t* __restrict__ data,
                                                               uint32_t rows, int logN, int s,
                                                               const uint64_t* __restrict__ tw) {
  const uint64_t halfN = (uint64_t)1 << (logN - 1);
  uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (uint64_t)rows * halfN) return;
  uint64_t row = gid / halfN, b = gid % halfN;
  const uint64_t half = (uint64_t)1 << (s - 1);
  uint64_t j = b & (half - 1);
  uint64_t i0 = ((b >> (s - 1)) << s) | j;
  uint64_t* d = data + (row << logN);
  uint64_t u = d[i0], v = d[i0 + half];
  if (j) v = gl_mul(v, tw[j << (logN - s)]);
  d[i0] = gl_add(u, v);
  d[i0 + half] = gl_sub(u, v);
}

__global__ void twiddle_kernel(uint64_t* tw, int logN) {
  // tw[k] = w^k for k < N/2, computed by square-and-multiply per entry
  uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ((uint64_t)1 << (logN - 1))) return;
  tw[k] = gl_pow(gl_root(logN), k);
}

hipError_t launch_twiddles(uint64_t* tw, int logN, hipStream_t st) {
  uint64_t half = (uint64_t)1 << (logN - 1);
  uint32_t blocks = (uint32_t)((half + 255) / 256);
  hipLaunchKernelGGL(twiddle_kernel, dim3(blocks), dim3(256), 0, st, tw, logN);
  return hipGetLastError();
}

// Canonical field elements from SplitMix64 of (seed + index): timing input for the autotuner.
__global__ void fill_field_kernel(uint64_t* __restrict__ p, uint64_t n, uint64_t seed) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t z = seed + (i + 1) * 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    p[i] = gl_canon(z);
  }
}

hipError_t launch_fill_field(uint64_t* p, uint64_t n, uint64_t seed, hipStream_t st) {
  hipLaunchKernelGGL(fill_field_kernel, dim3(1024), dim3(256), 0, st, p, n, seed);
  return hipGetLastError();
}

hipError_t launch_ntt(const uint64_t* in, uint64_t rs, uint64_t es, uint32_t n_in, uint32_t rows,
                      int logN, const uint64_t* tw, uint64_t* out, hipStream_t st, int impl,
                      const uint64_t* tp, const uint64_t* tp13) {
  if (rows == 0) return hipSuccess;
  if (impl == 4 || impl == 5) {
    if (logN == 13 && es == 1 && n_in == (1u << 12) && tp) {
      if (impl == 4)
        hipLaunchKernelGGL(ntt_r8_fixed_kernel<13>, dim3(rows), dim3(1024), ntt_lds_bytes(1u << 13), st, in, rs, tp,
                           out);
      else
        hipLaunchKernelGGL(ntt_r16_fixed_kernel<13>, dim3(rows), dim3(512), ntt_lds_bytes(1u << 13), st, in, rs, tp,
                           out);
      return hipGetLastError();
    }
    impl = 3;
  }
  if (impl != 1 && logN >= 1 && logN <= LGS_LDS_LOG) {
    const uint32_t N = 1u << logN;
    const uint32_t want = impl == 2 ? 512 : impl == 3 ? 1024 : 256;
    const uint32_t threads = N / 8 >= want ? want : (N / 8 >= 256 ? 256 : 64);
    hipLaunchKernelGGL(ntt_r8_kernel, dim3(rows), dim3(threads), ntt_lds_bytes(N), st, in, rs, es,
                       n_in, logN, tw, out);
    return hipGetLastError();
  }
  if (impl == 1) {  // radix-2: LDS stages, then one global pass per remaining stage
    const int lt = logN < LGS_LDS_LOG ? logN : LGS_LDS_LOG;
    const uint32_t T = 1u << lt, nblk = 1u << (logN - lt);
    const uint32_t threads = T / 2 < 256 ? (T / 2 < 64 ? 64 : T / 2) : 256;
    hipLaunchKernelGGL(ntt_lds_kernel, dim3(rows * nblk), dim3(threads), T * sizeof(uint64_t), st, in, rs, es,
                       n_in, logN, lt, tw, out);
    hipError_t e = hipGetLastError();
    const uint32_t blocks = (uint32_t)((((uint64_t)rows << (logN - 1)) + 255) / 256);
    for (int s = lt + 1; s <= logN && e == hipSuccess; s++) {
      hipLaunchKernelGGL(ntt_global_stage_kernel, dim3(blocks), dim3(256), 0, st, out, rows, logN, s, tw);
      e = hipGetLastError();
    }
    return e;
  }
  // Longer than one LDS tile: 2^lt-point transforms of strided subsequences in LDS, then the last
  // K stages in one column pass. For 2^14 .. 2^19 points and enough rows to give at least 512
  // subsequence workgroups, K = 3 (2^14 .. 2^16), 6 (2^17, 2^18) or 7 (2^19): one column pass in
  // registers (ntt_cols8_kernel), the subsequence transforms at most 2^12 points past 2^16 (at
  // 2^19, 2^13-point ones and 6 column stages measured slower than 2^12 and 7); otherwise
  // lt = 10 while that leaves at most 7 stages (so that few rows still give many workgroups),
  // global radix-2 stages while more than 7 remain, and the last K <= 7 stages in the radix-2
  // column pass.
  const int K8 = logN <= 16 ? 3 : logN <= 18 ? 6 : 7;
  if (logN >= 14 && logN <= 19 && ((uint64_t)rows << K8) >= 512) {
    const int lt8 = logN - K8;
    const uint32_t T = 1u << lt8, nblk = 1u << K8;
    const uint32_t bt = T / 8;  // one radix-8 butterfly per thread and pass
    if (tp13 && lt8 == 13 && n_in == (1u << (logN - 1)))  // zero upper half: the compile-time kernel
      hipLaunchKernelGGL(ntt_r8_fixed_sub_kernel<13>, dim3(rows * nblk), dim3(1024), ntt_lds_bytes(T), st, in, rs, es,
                         logN, tp13, out);
    else
      hipLaunchKernelGGL(ntt_r8_blocks_kernel, dim3(rows * nblk), dim3(bt), ntt_lds_bytes(T), st, in, rs, es, n_in,
                         logN, lt8, tw, out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const uint32_t cblocks = (uint32_t)(((uint64_t)rows << (logN - K8)) / (256 >> (K8 - 3)));
    if (K8 == 3)
      hipLaunchKernelGGL(ntt_cols8_kernel<3>, dim3(cblocks), dim3(256), 0, st, out, logN, tw);
    else if (K8 == 6)
      hipLaunchKernelGGL(ntt_cols8_kernel<6>, dim3(cblocks), dim3(256), 0, st, out, logN, tw);
    else
      hipLaunchKernelGGL(ntt_cols8_kernel<7>, dim3(cblocks), dim3(256), 0, st, out, logN, tw);
    return hipGetLastError();
  }
  const int K = logN - 10 < 7 ? logN - 10 : 7;
  const int lt = logN - K > LGS_LDS_LOG ? LGS_LDS_LOG : logN - K;
  const uint32_t T = 1u << lt, nblk = 1u << (logN - lt);
  uint32_t bt = T / 8;  // one radix-8 butterfly per thread and pass
  if (bt > 256 && (uint64_t)rows * nblk >= 512) bt = 256;
  hipLaunchKernelGGL(ntt_r8_blocks_kernel, dim3(rows * nblk), dim3(bt), ntt_lds_bytes(T), st, in, rs, es, n_in,
                     logN, lt, tw, out);
  hipError_t e = hipGetLastError();
  const uint32_t blocks = (uint32_t)((((uint64_t)rows << (logN - 1)) + 255) / 256);
  for (int s = lt + 1; s <= logN - K && e == hipSuccess; s++) {
    hipLaunchKernelGGL(ntt_global_stage_kernel, dim3(blocks), dim3(256), 0, st, out, rows, logN, s, tw);
    e = hipGetLastError();
  }
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(ntt_cols_kernel, dim3((uint32_t)(((uint64_t)rows << (logN - K)) / kColJ)), dim3(256), 0, st, out,
                     logN, K, tw);
  return hipGetLastError();
}


}  // namespace lgs
