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

hipError_t launch_ntt(const uint64_t* in, uint64_t rs, uint64_t es, uint32_t n_in, uint32_t rows,
                      int logN, const uint64_t* tw, uint64_t* out, hipStream_t st) {
  if (rows == 0) return hipSuccess;
  int lt = logN < LGS_LDS_LOG ? logN : LGS_LDS_LOG;
  uint32_t T = 1u << lt;
  uint32_t nblk = 1u << (logN - lt);
  uint32_t threads = T / 2 < 256 ? (T / 2 < 64 ? 64 : T / 2) : 256;
  hipLaunchKernelGGL(ntt_lds_kernel, dim3(rows * nblk), dim3(threads), T * sizeof(uint64_t), st,
                     in, rs, es, n_in, logN, lt, tw, out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  uint64_t work = (uint64_t)rows << (logN - 1);
  uint32_t blocks = (uint32_t)((work + 255) / 256);
  for (int s = lt + 1; s <= logN; s++) {
    hipLaunchKernelGGL(ntt_global_stage_kernel, dim3(blocks), dim3(256), 0, st, out, rows, logN, s,
                       tw);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace lgs
