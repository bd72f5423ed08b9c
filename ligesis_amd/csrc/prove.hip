// Eval-prove kernels (pcs/src/ligesis.rs:347-549, utils/src/poly.rs, utils/src/sumcheck.rs,
// pcs/src/basefold.rs:644-856). Every challenge of the transcript is an input known before the
// first launch (the reference's RandomOracle is a pre-drawn stream, utils/src/oracle.rs:10-64),
// so every kernel reads its challenges from a device array and the whole prove is enqueued
// without host round trips. Extension elements are stored interleaved (c0, c1).
#include "prove_kernels.hpp"

namespace lgs {

__device__ __forceinline__ ef2 ld_ef(const ef2* p, uint64_t i) { return p[i]; }
__device__ __forceinline__ ef2 ef_one_minus(ef2 a) { return ef2{gl_sub(1, a.c0), gl_neg(a.c1)}; }

// Block-wide sum of W extension values per thread; result valid in thread 0.
template <int W>
__device__ __forceinline__ void block_sum(ef2 (&v)[W], ef2* sh /* blockDim x W */) {
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < W; q++) sh[t * W + q] = v[q];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (t < s) {
#pragma unroll
      for (int q = 0; q < W; q++) sh[t * W + q] = ef_add(sh[t * W + q], sh[(t + s) * W + q]);
    }
    __syncthreads();
  }
  if (t == 0) {
#pragma unroll
    for (int q = 0; q < W; q++) v[q] = sh[q];
  }
}

// One deferred reduction (RedJob) by one block: strided partial sums, then a tree in LDS (sh
// holds blockDim x 4 values).
__device__ __forceinline__ void red_job_body(const RedJob& r, ef2* sh) {
  ef2 acc[4] = {ef_zero(), ef_zero(), ef_zero(), ef_zero()};
  for (uint32_t b = threadIdx.x; b < r.nblk; b += blockDim.x)
    for (uint32_t q = 0; q < r.w; q++) acc[q] = ef_add(acc[q], r.part[(uint64_t)b * r.w + q]);
  block_sum<4>(acc, sh);
  if (threadIdx.x == 0)
    for (uint32_t q = 0; q < r.w; q++) r.out[q] = acc[q];
}

__global__ __launch_bounds__(256) void reduce_jobs_kernel(RedJobs r) {
  __shared__ ef2 sh[256 * 4];
  red_job_body(r.j[blockIdx.x], sh);
}

hipError_t launch_reduce_jobs(const RedJobs& r, hipStream_t st) {
  if (r.n > 0) hipLaunchKernelGGL(reduce_jobs_kernel, dim3(r.n), dim3(256), 0, st, r);
  return hipGetLastError();
}

// ---------------- eq tables (MlPoly::new_eq, poly.rs:44-53) ----------------
// Many eq tables in two launches. Parts: one job per (point, half) pair of every term; expand:
// out[b] = sum_t scale_t * lo_t[b mod 2^klo] * hi_t[b >> klo] over the table's terms (a table
// with several terms is a random linear combination of eq tables, ligesis.rs:513-522).
__global__ void eq_parts_group_kernel(const EqTerm* __restrict__ terms, int nterms) {
  int lo = 0, hi = nterms - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (terms[mid].first_block <= blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const EqTerm& t = terms[lo];
  const uint32_t b = (blockIdx.x - t.first_block) * blockDim.x + threadIdx.x;
  const int klo = (int)t.klo, khi = (int)(t.k - t.klo);
  if (b < (1u << klo)) {
    ef2 acc = ef_one();
    for (int q = 0; q < klo; q++) acc = ef_mul(acc, ((b >> q) & 1) ? t.pt[q] : ef_one_minus(t.pt[q]));
    t.lo[b] = acc;
  }
  if (b < (1u << khi)) {
    ef2 acc = t.scale ? *t.scale : ef_one();
    for (int q = 0; q < khi; q++) acc = ef_mul(acc, ((b >> q) & 1) ? t.pt[klo + q] : ef_one_minus(t.pt[klo + q]));
    t.hi[b] = acc;
  }
}

__global__ void eq_expand_group_kernel(const EqTable* __restrict__ tabs, int ntabs) {
  int lo = 0, hi = ntabs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (tabs[mid].first_block <= blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const EqTable& t = tabs[lo];
  const uint64_t b = (uint64_t)(blockIdx.x - t.first_block) * blockDim.x + threadIdx.x;
  if (b >= t.n) return;
  const uint64_t bl = b & ((uint64_t(1) << t.klo) - 1), bh = b >> t.klo;
  ef2 v = ef_mul(t.lo[0][bl], t.hi[0][bh]);
  for (uint32_t q = 1; q < t.nterms; q++) v = ef_add(v, ef_mul(t.lo[q][bl], t.hi[q][bh]));
  t.out[b] = v;
}

hipError_t launch_eq_group(const EqTerm* terms, int nterms, uint32_t term_blocks, const EqTable* tabs, int ntabs,
                           uint32_t tab_blocks, hipStream_t st) {
  if (term_blocks) hipLaunchKernelGGL(eq_parts_group_kernel, dim3(term_blocks), dim3(256), 0, st, terms, nterms);
  if (tab_blocks) hipLaunchKernelGGL(eq_expand_group_kernel, dim3(tab_blocks), dim3(256), 0, st, tabs, ntabs);
  return hipGetLastError();
}

// ---------------- weighted row sums: out[j] = sum_i w[i] * M[i][j] (M base field) -----------
// ā = eq_{z1}^T F (ligesis.rs:372-378) and g1 = eq_{alpha2}^T A (ligesis.rs:436-444).
__global__ void wrows_kernel(const uint64_t* __restrict__ M, uint64_t cols, const ef2* __restrict__ w,
                             uint32_t rows_per_chunk, ef2* __restrict__ part) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cols) return;
  const uint32_t i0 = blockIdx.y * rows_per_chunk;
  ef2 acc = ef_zero();
  for (uint32_t i = i0; i < i0 + rows_per_chunk; i++) acc = ef_add(acc, ef_mul_base(w[i], M[i * cols + j]));
  part[(uint64_t)blockIdx.y * cols + j] = acc;
}

// out (and the optional copies out2, out3, which later kernels fold in place) = sum of chunks
__global__ void sum_chunks_kernel(const ef2* __restrict__ part, uint32_t chunks, uint64_t cols,
                                  ef2* __restrict__ out, ef2* __restrict__ out2, ef2* __restrict__ out3) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cols) return;
  ef2 acc = ef_zero();
  for (uint32_t c = 0; c < chunks; c++) acc = ef_add(acc, part[(uint64_t)c * cols + j]);
  out[j] = acc;
  if (out2) out2[j] = acc;
  if (out3) out3[j] = acc;
}

hipError_t launch_wrows(const uint64_t* M, uint32_t rows, uint64_t cols, const ef2* w, ef2* part,
                        uint32_t chunks, ef2* out, hipStream_t st, ef2* out2, ef2* out3) {
  const uint32_t bx = (uint32_t)((cols + 255) / 256);
  hipLaunchKernelGGL(wrows_kernel, dim3(bx, chunks), dim3(256), 0, st, M, cols, w, rows / chunks, part);
  hipLaunchKernelGGL(sum_chunks_kernel, dim3(bx), dim3(256), 0, st, part, chunks, cols, out, out2, out3);
  return hipGetLastError();
}

// ---------------- dot products: out = sum_i a[i] * b[i] ----------------
__global__ __launch_bounds__(256) void dot_kernel(const ef2* __restrict__ a, const ef2* __restrict__ b,
                                                  uint64_t n, ef2* __restrict__ part) {
  __shared__ ef2 sh[256];
  ef2 acc[1] = {ef_zero()};
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    acc[0] = ef_add(acc[0], ef_mul(a[i], b[i]));
  block_sum<1>(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

// out[q] = sum_{b < nb} part[b * width + q], for q < width <= 4 (one block of 256 threads:
// strided partial sums, then a tree in LDS)

// The last rounds of a product sumcheck in one workgroup: the factors (at most 2 * 512 entries
// each) are folded in LDS round after round; round polynomials go to `rounds` (D + 1 each) and
// the final evaluations to `finals`.
template <int D, int CAP>
__device__ __forceinline__ void sc_tail_body(const ef2* __restrict__ p0, const ef2* __restrict__ p1,
                                             const ef2* __restrict__ p2, uint32_t len,
                                             const ef2* __restrict__ chal, ef2* __restrict__ rounds,
                                             ef2* __restrict__ finals, ef2 (*buf)[CAP], ef2* sh) {
  const ef2* src[3] = {p0, p1, p2};
#pragma unroll
  for (int j = 0; j < D; j++)
    for (uint32_t i = threadIdx.x; i < len; i += blockDim.x) buf[j][i] = src[j][i];
  __syncthreads();
  uint32_t t = 0;
  for (uint32_t cur = len; cur > 1; cur >>= 1, t++) {
    const uint32_t half = cur / 2;
    const ef2 r = chal[t];
    ef2 acc[D + 1];
#pragma unroll
    for (int q = 0; q <= D; q++) acc[q] = ef_zero();
    ef2 folded[D][2];
    for (uint32_t k = 0; k < 2; k++) {
      const uint32_t i = threadIdx.x + k * blockDim.x;
      if (i >= half) continue;
      ef2 lo[D], diff[D], x[D];
#pragma unroll
      for (int j = 0; j < D; j++) {
        lo[j] = buf[j][i];
        diff[j] = ef_sub(buf[j][i + half], lo[j]);
        folded[j][k] = ef_add(lo[j], ef_mul(r, diff[j]));
        x[j] = lo[j];
      }
#pragma unroll
      for (int q = 0; q <= D; q++) {
        ef2 prod = x[0];
#pragma unroll
        for (int j = 1; j < D; j++) prod = ef_mul(prod, x[j]);
        acc[q] = ef_add(acc[q], prod);
#pragma unroll
        for (int j = 0; j < D; j++) x[j] = ef_add(x[j], diff[j]);
      }
    }
    block_sum<D + 1>(acc, sh);  // ends with a barrier: every read of buf is done
    for (uint32_t k = 0; k < 2; k++) {
      const uint32_t i = threadIdx.x + k * blockDim.x;
      if (i < half) {
#pragma unroll
        for (int j = 0; j < D; j++) buf[j][i] = folded[j][k];
      }
    }
    if (threadIdx.x == 0) {
#pragma unroll
      for (int q = 0; q <= D; q++) rounds[(uint64_t)t * (D + 1) + q] = acc[q];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < D; j++) finals[j] = buf[j][0];
  }
}

// The tails of several sumchecks, one wave each (blockIdx.x = job).
__global__ __launch_bounds__(64) void sc_tail_group_kernel(const ScTailJob* __restrict__ jobs, uint32_t njobs,
                                                           const RedJob* __restrict__ reds) {
  __shared__ ef2 buf[3][128];
  __shared__ ef2 sh[64 * 4];
  if (blockIdx.x >= njobs) {
    red_job_body(reds[blockIdx.x - njobs], sh);
    return;
  }
  const ScTailJob& j = jobs[blockIdx.x];
  if (j.d == 3)
    sc_tail_body<3, 128>(j.p[0], j.p[1], j.p[2], j.len, j.chal, j.rounds, j.finals, buf, sh);
  else
    sc_tail_body<2, 128>(j.p[0], j.p[1], nullptr, j.len, j.chal, j.rounds, j.finals, buf, sh);
}

hipError_t launch_sc_tail_group(const ScTailJob* jobs, uint32_t njobs, const RedJob* reds, uint32_t nreds,
                                hipStream_t st) {
  if (njobs + nreds)
    hipLaunchKernelGGL(sc_tail_group_kernel, dim3(njobs + nreds), dim3(64), 0, st, jobs, njobs, reds);
  return hipGetLastError();
}


uint32_t dot_blocks(uint64_t n) {
  const uint64_t nb = (n + 255) / 256;
  return (uint32_t)(nb > 256 ? 256 : nb == 0 ? 1 : nb);
}

hipError_t launch_dot(const ef2* a, const ef2* b, uint64_t n, ef2* part, hipStream_t st) {
  hipLaunchKernelGGL(dot_kernel, dim3(dot_blocks(n)), dim3(256), 0, st, a, b, n, part);
  return hipGetLastError();
}

// ---------------- LigeSIS-specific vectors ----------------
// B_I (ligesis.rs:403-410): bi[i * s + k] = bit (i mod 64) of F'[i / 64][idx[k]], as 0/1
// extension elements; bi_m1 = bi - 1 (ligesis.rs:424-425).
__global__ void bits_gather_kernel(const uint64_t* __restrict__ enc, uint64_t ld, const uint32_t* __restrict__ idx,
                                   uint32_t s, uint64_t W, uint64_t* __restrict__ bi) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= W * s) return;
  const uint64_t i = t / s;
  const uint32_t k = (uint32_t)(t % s);
  bi[t] = (enc[(i >> 6) * ld + idx[k]] >> (i & 63)) & 1;
}

// g2[i] = sum_k eq_r2[k] * bi[i * s + k] (ligesis.rs:445-455): one wave per i, its lanes
// reading consecutive words of the row (coalesced), then a butterfly sum across the wave.
__device__ __forceinline__ ef2 wave_sum(ef2 v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const ef2 o{(uint64_t)__shfl_xor((unsigned long long)v.c0, off), (uint64_t)__shfl_xor((unsigned long long)v.c1, off)};
    v = ef_add(v, o);
  }
  return v;
}

__global__ __launch_bounds__(256) void g2_kernel(const uint64_t* __restrict__ bi, uint32_t s, uint64_t W,
                                                 const ef2* __restrict__ eq_r2, ef2* __restrict__ g2,
                                                 ef2* __restrict__ g2b) {
  const uint64_t i = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t lane = threadIdx.x & 63;
  if (i >= W) return;  // whole waves leave together
  ef2 acc = ef_zero();
  for (uint32_t k = lane; k < s; k += 64)
    if (bi[i * s + k]) acc = ef_add(acc, eq_r2[k]);
  acc = wave_sum(acc);
  if (lane == 0) {
    g2[i] = acc;
    if (g2b) g2b[i] = acc;
  }
}

// v[i] = eqz1[i / 64] * 2^(i mod 64) (ligesis.rs:474-476)
__global__ void vvec_kernel(const ef2* __restrict__ eqz1, uint64_t W, ef2* __restrict__ v) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W) return;
  const uint32_t k = i & 63;
  const uint64_t p2 = k < 64 ? gl_canon(uint64_t(1) << k) : 0;
  v[i] = ef_mul_base(eqz1[i >> 6], p2);
}

// sel[idx[k]] = eq_r2[k] over a zeroed vector (ligesis.rs:488-491)
__global__ void sel_kernel(const uint32_t* __restrict__ idx, const ef2* __restrict__ eq_r2, uint32_t s,
                           ef2* __restrict__ sel) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < s) sel[idx[k]] = eq_r2[k];
}

__global__ void zero_kernel(uint64_t* __restrict__ p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// w[k] = prod_t (1 - r3[t] + r3[t] * g^((2^t k) mod 2n)) (rs_weights, ligesis.rs:553-567);
// tw[e] = g^e for e < n, and g^(e + n) = -g^e.
__global__ void rs_weights_kernel(const ef2* __restrict__ r3, int log2n, uint64_t n,
                                  const uint64_t* __restrict__ tw, ef2* __restrict__ w) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint64_t mask = 2 * n - 1;
  ef2 prod = ef_one();
  for (int t = 0; t < log2n; t++) {
    const uint64_t e = (k << t) & mask;
    const uint64_t gp = e < n ? tw[e] : gl_neg(tw[e - n]);
    const ef2 r = r3[t];
    const ef2 f = ef_add(ef_one_minus(r), ef_mul_base(r, gp));
    prod = ef_mul(prod, f);
  }
  w[k] = prod;
}

// Planar (c0 row, c1 row) -> interleaved extension vector.
__global__ void planar_to_ef_kernel(const uint64_t* __restrict__ c0, const uint64_t* __restrict__ c1,
                                    uint64_t n, ef2* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = ef2{c0[i], c1[i]};
}

// ---------------- product sumcheck round (sumcheck.rs:37-84) ----------------
// Round over pairs (lo = p[i], hi = p[i + half]): g(t) = sum_i prod_j ((1 - t) lo_j + t hi_j) for
// t = 0..D, then p[i] <- lo + r (hi - lo) in place (the high variable is bound first).
template <int D>
__device__ __forceinline__ void sc_round_body(ef2* __restrict__ p0, ef2* __restrict__ p1, ef2* __restrict__ p2,
                                              uint64_t half, const ef2* __restrict__ r_ptr, ef2* __restrict__ part,
                                              uint32_t blk, uint32_t nblk, ef2* sh) {
  ef2* ps[3] = {p0, p1, p2};
  const ef2 r = *r_ptr;
  ef2 acc[D + 1];
#pragma unroll
  for (int q = 0; q <= D; q++) acc[q] = ef_zero();
  for (uint64_t i = (uint64_t)blk * blockDim.x + threadIdx.x; i < half; i += (uint64_t)nblk * blockDim.x) {
    ef2 lo[D], diff[D];
#pragma unroll
    for (int j = 0; j < D; j++) {
      lo[j] = ps[j][i];
      const ef2 hi = ps[j][i + half];
      diff[j] = ef_sub(hi, lo[j]);
      ps[j][i] = ef_add(lo[j], ef_mul(r, diff[j]));
    }
    // evaluations at t = 0..D by stepping x_j(t + 1) = x_j(t) + diff_j
    ef2 x[D];
#pragma unroll
    for (int j = 0; j < D; j++) x[j] = lo[j];
#pragma unroll
    for (int q = 0; q <= D; q++) {
      ef2 prod = x[0];
#pragma unroll
      for (int j = 1; j < D; j++) prod = ef_mul(prod, x[j]);
      acc[q] = ef_add(acc[q], prod);
#pragma unroll
      for (int j = 0; j < D; j++) x[j] = ef_add(x[j], diff[j]);
    }
  }
  block_sum<D + 1>(acc, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q <= D; q++) part[(uint64_t)blk * (D + 1) + q] = acc[q];
  }
}

// First round of the binary-check sumcheck sum_x eq(x) b(x) (b(x) - 1) (ligesis.rs:405-409)
// with b read as 0/1 words. Per pair (lo, hi) of the top variable, b(t) = lo + t (hi - lo):
// b(t) (b(t) - 1) vanishes when lo = hi, and is t (t - 1) otherwise, so the round values are
// s(0) = s(1) = 0, s(2) = 2 sum (2 E_hi - E_lo), s(3) = 6 sum (3 E_hi - 2 E_lo) over the pairs
// with lo != hi -- the same field values the general product rule gives, from additions only.
// The folded bit lo + r (hi - lo) is one of 0, 1, r, 1 - r.
__device__ __forceinline__ void sc_bits3_body(const uint64_t* __restrict__ bits, ef2* __restrict__ b_out,
                                              ef2* __restrict__ m1_out, ef2* __restrict__ e, uint64_t half,
                                              const ef2* __restrict__ r_ptr, ef2* __restrict__ part, uint32_t blk,
                                              uint32_t nblk, ef2* sh) {
  const ef2 r = *r_ptr, one = ef_one();
  const ef2 omr = ef_sub(one, r), rm1 = ef_sub(r, one), nr = ef_neg(r), m1 = ef_neg(one);
  ef2 L = ef_zero(), H = ef_zero();
  for (uint64_t i = (uint64_t)blk * blockDim.x + threadIdx.x; i < half; i += (uint64_t)nblk * blockDim.x) {
    const uint64_t blo = bits[i], bhi = bits[i + half];
    const ef2 elo = e[i], ehi = e[i + half];
    e[i] = ef_add(elo, ef_mul(r, ef_sub(ehi, elo)));
    ef2 bf, bm;
    if (blo == bhi) {
      bf = blo ? one : ef_zero();
      bm = blo ? ef_zero() : m1;
    } else {
      bf = bhi ? r : omr;
      bm = bhi ? rm1 : nr;
      L = ef_add(L, elo);
      H = ef_add(H, ehi);
    }
    b_out[i] = bf;
    m1_out[i] = bm;
  }
  // 2 (2H - L) and 6 (3H - 2L), linear in the block's sums, so partials add up
  const ef2 h2 = ef_add(H, H), l2 = ef_add(L, L);
  const ef2 s2 = ef_sub(h2, L);
  const ef2 s3 = ef_sub(ef_add(h2, H), l2);
  const ef2 s3x2 = ef_add(s3, s3);
  ef2 acc[4] = {ef_zero(), ef_zero(), ef_add(s2, s2), ef_add(s3x2, ef_add(s3x2, s3x2))};
  block_sum<4>(acc, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < 4; q++) part[(uint64_t)blk * 4 + q] = acc[q];
  }
}

// First round of sum_x b(x) G(x) with b read as 0/1 words (the batched bit opening,
// ligesis.rs:513-522): s(0) = sum_{lo = 1} G_lo, s(1) = sum_{hi = 1} G_hi and
// s(2) = sum (2 hi - lo)(2 G_hi - G_lo) = 4 s(1) - 2 C + s(0) with
// C = sum_{hi = 1} G_lo + sum_{lo = 1} G_hi.
__device__ __forceinline__ void sc_bits2_body(const uint64_t* __restrict__ bits, ef2* __restrict__ b_out,
                                              ef2* __restrict__ g, uint64_t half, const ef2* __restrict__ r_ptr,
                                              ef2* __restrict__ part, uint32_t blk, uint32_t nblk, ef2* sh) {
  const ef2 r = *r_ptr, one = ef_one();
  const ef2 omr = ef_sub(one, r);
  ef2 A0 = ef_zero(), A1 = ef_zero(), C = ef_zero();
  for (uint64_t i = (uint64_t)blk * blockDim.x + threadIdx.x; i < half; i += (uint64_t)nblk * blockDim.x) {
    const uint64_t blo = bits[i], bhi = bits[i + half];
    const ef2 glo = g[i], ghi = g[i + half];
    g[i] = ef_add(glo, ef_mul(r, ef_sub(ghi, glo)));
    if (blo) {
      A0 = ef_add(A0, glo);
      C = ef_add(C, ghi);
    }
    if (bhi) {
      A1 = ef_add(A1, ghi);
      C = ef_add(C, glo);
    }
    b_out[i] = blo == bhi ? (blo ? one : ef_zero()) : (bhi ? r : omr);
  }
  const ef2 a1x2 = ef_add(A1, A1), cx2 = ef_add(C, C);
  ef2 acc[3] = {A0, A1, ef_add(ef_sub(ef_add(a1x2, a1x2), cx2), A0)};
  block_sum<3>(acc, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < 3; q++) part[(uint64_t)blk * 3 + q] = acc[q];
  }
}

// One round of several independent sumchecks in one launch (sumchecks run in lockstep): block
// -> job by binary search over the jobs' first blocks.
__global__ __launch_bounds__(256) void sc_round_group_kernel(const ScRoundJob* __restrict__ jobs, int njobs,
                                                             uint32_t round_blocks, const RedJob* __restrict__ reds) {
  __shared__ ef2 sh[256 * 4];
  if (blockIdx.x >= round_blocks) {
    red_job_body(reds[blockIdx.x - round_blocks], sh);
    return;
  }
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (jobs[mid].first_block <= blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const ScRoundJob& j = jobs[lo];
  const uint32_t blk = blockIdx.x - j.first_block;
  if (j.bits) {
    if (j.d == 3)
      sc_bits3_body(j.bits, j.p[0], j.p[1], j.p[2], j.half, j.r, j.part, blk, j.nblk, sh);
    else
      sc_bits2_body(j.bits, j.p[0], j.p[1], j.half, j.r, j.part, blk, j.nblk, sh);
  } else if (j.d == 3)
    sc_round_body<3>(j.p[0], j.p[1], j.p[2], j.half, j.r, j.part, blk, j.nblk, sh);
  else
    sc_round_body<2>(j.p[0], j.p[1], nullptr, j.half, j.r, j.part, blk, j.nblk, sh);
}

hipError_t launch_sc_round_group(const ScRoundJob* jobs, int njobs, uint32_t total_blocks, const RedJob* reds,
                                 uint32_t nreds, hipStream_t st) {
  if (total_blocks + nreds)
    hipLaunchKernelGGL(sc_round_group_kernel, dim3(total_blocks + nreds), dim3(256), 0, st, jobs, njobs,
                       total_blocks, reds);
  return hipGetLastError();
}

uint32_t sc_round_blocks(uint64_t half) {
  uint64_t nb = (half + 255) / 256;
  return (uint32_t)(nb > 1024 ? 1024 : nb);
}



// ---------------- Basefold batch (basefold.rs:648-856) ----------------
// combined[i] = sum over channels c, chunks j of wt_c[j] * codes_c[j][i], where
// wt_c[j] = coeffs[c] * eq4_c[j]; an extension channel keeps c0 / c1 in rows 2j / 2j+1.
__global__ void bf_combine_kernel(BfChannels ch, uint64_t D, ef2* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= D) return;
  ef2 acc = ef_zero();
  for (int c = 0; c < ch.count; c++) {
    const uint64_t* codes = ch.codes[c];
    const ef2* wt = ch.wt[c];
    for (uint32_t j = 0; j < ch.chunks[c]; j++) {
      if (ch.ext[c]) {
        const ef2 v{codes[(2 * j) * D + i], codes[(2 * j + 1) * D + i]};
        acc = ef_add(acc, ef_mul(wt[j], v));
      } else {
        acc = ef_add(acc, ef_mul_base(wt[j], codes[j * D + i]));
      }
    }
  }
  out[i] = acc;
}

// m[x] = sum_{b < 2^li} eq4[b] * poly[x * 2^li + b] (MlPoly::fold over the low li variables)
__global__ void fold_low_kernel(const uint64_t* __restrict__ poly, int ext, int li, uint64_t nout,
                                const ef2* __restrict__ eq4, ef2* __restrict__ m) {
  const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= nout) return;
  ef2 acc = ef_zero();
  const uint64_t base = x << li;
  for (uint32_t b = 0; b < (1u << li); b++) {
    if (ext) {
      const ef2 v{poly[2 * (base + b)], poly[2 * (base + b) + 1]};
      acc = ef_add(acc, ef_mul(eq4[b], v));
    } else {
      acc = ef_add(acc, ef_mul_base(eq4[b], poly[base + b]));
    }
  }
  m[x] = acc;
}

// One reduction round: partial sums of m[2i] * eq_rest[i] (l0, basefold.rs:752-758) and the
// low-variable fold m'[i] = m[2i] + r (m[2i+1] - m[2i]) into a second buffer.

// Fold of a codeword over the subgroup of order 2h at challenge r: with x = v[i], nx = v[i + h]
// (the values at g^i and -g^i), out[i] = (1 - r) (x + nx) / 2 + r (x - nx) / (2 g^i).
__device__ __forceinline__ void eq_fold_codeword_at(const ef2* __restrict__ v, uint64_t h, ef2 r,
                                                    const uint64_t* __restrict__ tw, ef2* __restrict__ out,
                                                    uint64_t i) {
  const ef2 x = v[i], nx = v[i + h];
  const uint64_t ginv = i == 0 ? 1 : gl_neg(tw[h - i]);
  const ef2 pe = ef_mul_base(ef_add(x, nx), GL_INV2);
  const ef2 po = ef_mul_base(ef_sub(x, nx), gl_mul(ginv, GL_INV2));
  out[i] = ef_add(ef_mul(ef_one_minus(r), pe), ef_mul(r, po));
}


// All channels of one reduction round in one launch (blockIdx.y = channel): l0 partials with
// the last block of each channel reducing them into out[c], the low-variable fold of m, and the
// eq table of the next round's residual point (pair sums).
__global__ __launch_bounds__(256) void bf_round3_kernel(BfRound a, uint64_t half, const ef2* __restrict__ r_ptr) {
  __shared__ ef2 sh[256 * 4];
  const int c = blockIdx.y;
  if (c == 4) {  // the previous round's l0 reductions
    if (blockIdx.x < 3 && a.prev[blockIdx.x].nblk) red_job_body(a.prev[blockIdx.x], sh);
    return;
  }
  if (c == 3) {  // codeword fold, grid-stride over its half length
    if (!a.cw_out) return;
    const ef2 r = *r_ptr;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.cw_half;
         i += (uint64_t)gridDim.x * blockDim.x)
      eq_fold_codeword_at(a.cw_in, a.cw_half, r, a.cw_tw, a.cw_out, i);
    return;
  }
  const ef2* m = a.m_in[c];
  const ef2* eq = a.eq[c];
  ef2* m_out = a.m_out[c];
  ef2* eq_next = a.eq_next[c];
  const ef2 r = *r_ptr;
  ef2 acc[1] = {ef_zero()};
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < half; i += (uint64_t)gridDim.x * blockDim.x) {
    const ef2 e = m[2 * i], o = m[2 * i + 1];
    acc[0] = ef_add(acc[0], ef_mul(e, eq[i]));
    m_out[i] = ef_add(e, ef_mul(r, ef_sub(o, e)));
    if (half > 1 && i < half / 2) eq_next[i] = ef_add(eq[2 * i], eq[2 * i + 1]);
  }
  block_sum<1>(acc, sh);
  if (threadIdx.x == 0) a.part[(uint64_t)c * gridDim.x + blockIdx.x] = acc[0];
}

uint32_t bf_round3_blocks(uint64_t half, uint64_t cw_half) {
  // blocks of the three channels; the codeword fold (y = 3) grid-strides over the same count,
  // at least 64 blocks while it has work for them; at least 3 for the reductions (y = 4)
  uint32_t nb = sc_round_blocks(half);
  const uint64_t fold_blocks = (cw_half + 255) / 256;
  if (nb < 64 && fold_blocks > nb) nb = (uint32_t)(fold_blocks < 64 ? fold_blocks : 64);
  return nb < 3 ? 3 : nb;
}

hipError_t launch_bf_round3(const BfRound& a, uint64_t half, const ef2* r, hipStream_t st) {
  hipLaunchKernelGGL(bf_round3_kernel, dim3(bf_round3_blocks(half, a.cw_out ? a.cw_half : 0), 5), dim3(256), 0,
                     st, a, half, r);
  return hipGetLastError();
}


// eq_fold_codeword (basefold.rs:201-218): out[i] = (1 - r)(x + nx)/2 + r (x - nx) g^-i / 2 with
// x = v[i], nx = v[i + h]; g^-i = -g^(h - i) for 0 < i < h, tw[e] = g^e (e < h).


// ---------------- gathers for openings ----------------
__global__ void gather_parts_kernel(const GatherPart* __restrict__ parts, uint32_t nparts,
                                    const uint64_t* __restrict__ off, uint64_t total, uint64_t* __restrict__ dst) {
  for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < total; w += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t lo = 0, hi = nparts - 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) / 2;
      if (parts[mid].first <= w) lo = mid;
      else hi = mid - 1;
    }
    dst[w] = parts[lo].src[off[w]];
  }
}

#define LGS_GRID(n) dim3((uint32_t)(((n) + 255) / 256)), dim3(256)

hipError_t launch_bits_gather(const uint64_t* enc, uint64_t ld, const uint32_t* idx, uint32_t s, uint64_t W,
                              uint64_t* bi, hipStream_t st) {
  hipLaunchKernelGGL(bits_gather_kernel, LGS_GRID(W * s), 0, st, enc, ld, idx, s, W, bi);
  return hipGetLastError();
}
hipError_t launch_g2(const uint64_t* bi, uint32_t s, uint64_t W, const ef2* eq_r2, ef2* g2, hipStream_t st,
                     ef2* g2b) {
  hipLaunchKernelGGL(g2_kernel, LGS_GRID(W * 64), 0, st, bi, s, W, eq_r2, g2, g2b);
  return hipGetLastError();
}
hipError_t launch_vvec(const ef2* eqz1, uint64_t W, ef2* v, hipStream_t st) {
  hipLaunchKernelGGL(vvec_kernel, LGS_GRID(W), 0, st, eqz1, W, v);
  return hipGetLastError();
}
hipError_t launch_sel(const uint32_t* idx, const ef2* eq_r2, uint32_t s, ef2* sel, uint64_t len, hipStream_t st) {
  hipLaunchKernelGGL(zero_kernel, dim3(256), dim3(256), 0, st, (uint64_t*)sel, 2 * len);
  hipLaunchKernelGGL(sel_kernel, LGS_GRID(s), 0, st, idx, eq_r2, s, sel);
  return hipGetLastError();
}
hipError_t launch_rs_weights(const ef2* r3, int log2n, uint64_t n, const uint64_t* tw, ef2* w, hipStream_t st) {
  hipLaunchKernelGGL(rs_weights_kernel, LGS_GRID(n), 0, st, r3, log2n, n, tw, w);
  return hipGetLastError();
}
hipError_t launch_planar_to_ef(const uint64_t* c0, const uint64_t* c1, uint64_t n, ef2* out, hipStream_t st) {
  hipLaunchKernelGGL(planar_to_ef_kernel, LGS_GRID(n), 0, st, c0, c1, n, out);
  return hipGetLastError();
}
hipError_t launch_bf_combine(const BfChannels& ch, uint64_t D, ef2* out, hipStream_t st) {
  hipLaunchKernelGGL(bf_combine_kernel, LGS_GRID(D), 0, st, ch, D, out);
  return hipGetLastError();
}
hipError_t launch_fold_low(const uint64_t* poly, int ext, int li, uint64_t nout, const ef2* eq4, ef2* m,
                           hipStream_t st) {
  hipLaunchKernelGGL(fold_low_kernel, LGS_GRID(nout), 0, st, poly, ext, li, nout, eq4, m);
  return hipGetLastError();
}
hipError_t launch_gather_parts(const GatherPart* parts, uint32_t nparts, const uint64_t* off, uint64_t total,
                               uint64_t* dst, hipStream_t st) {
  if (total == 0 || nparts == 0) return hipSuccess;
  uint64_t blocks = (total + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(gather_parts_kernel, dim3((uint32_t)blocks), dim3(256), 0, st, parts, nparts, off, total, dst);
  return hipGetLastError();
}
hipError_t launch_zero(uint64_t* p, uint64_t n, hipStream_t st) {
  hipLaunchKernelGGL(zero_kernel, dim3(256), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

}  // namespace lgs
