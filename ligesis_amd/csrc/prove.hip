// Eval-prove kernels (pcs/src/ligesis.rs:347-549, utils/src/poly.rs, utils/src/sumcheck.rs,
// pcs/src/basefold.rs:644-856). Every challenge of the transcript is an input known before the
// first launch (the reference's RandomOracle is a pre-drawn stream, utils/src/oracle.rs:10-64),
// so every kernel reads its challenges from a device array and the whole prove is enqueued
// without host round trips. Extension elements are stored interleaved (c0, c1).
#include "prove_kernels.hpp"

namespace lgs {

__device__ __forceinline__ ef2 ld_ef(const ef2* p, uint64_t i) { return p[i]; }
__device__ __forceinline__ ef2 ef_one_minus(ef2 a) { return ef2{gl_sub(1, a.c0), gl_neg(a.c1)}; }

// The value of the lane that DPP control CTRL selects, 0 in rows outside ROW_MASK.
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ ef2 ef_dpp(ef2 v) {
  const uint32_t w0 = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)v.c0, CTRL, ROW_MASK, 0xF, false);
  const uint32_t w1 = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v.c0 >> 32), CTRL, ROW_MASK, 0xF, false);
  const uint32_t w2 = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)v.c1, CTRL, ROW_MASK, 0xF, false);
  const uint32_t w3 = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v.c1 >> 32), CTRL, ROW_MASK, 0xF, false);
  return ef2{((uint64_t)w1 << 32) | w0, ((uint64_t)w3 << 32) | w2};
}

// Sum over the 64 lanes of a wave, returned in every lane; every lane must be active. Row sums
// by DPP (quad permutations [1,0,3,2] and [2,3,0,1], half-row and row mirrors: every lane of a
// row of 16 then holds the row's sum), then row 0 into row 1 and row 2 into row 3 (row_bcast15),
// rows 0-1 into rows 2-3 (row_bcast31): lane 63 holds the total, read back as a scalar. Each step
// reads a neighbour's registers directly; a shuffle through LDS (ds_bpermute) per step made a
// reduction several times slower where a kernel waits for it.
__device__ __forceinline__ ef2 wave_sum(ef2 v) {
  v = ef_add(v, ef_dpp<0xB1>(v));
  v = ef_add(v, ef_dpp<0x4E>(v));
  v = ef_add(v, ef_dpp<0x141>(v));
  v = ef_add(v, ef_dpp<0x140>(v));
  v = ef_add(v, ef_dpp<0x142, 0xA>(v));
  v = ef_add(v, ef_dpp<0x143, 0xC>(v));
  const uint32_t w0 = __builtin_amdgcn_readlane((int)(uint32_t)v.c0, 63);
  const uint32_t w1 = __builtin_amdgcn_readlane((int)(uint32_t)(v.c0 >> 32), 63);
  const uint32_t w2 = __builtin_amdgcn_readlane((int)(uint32_t)v.c1, 63);
  const uint32_t w3 = __builtin_amdgcn_readlane((int)(uint32_t)(v.c1 >> 32), 63);
  return ef2{((uint64_t)w1 << 32) | w0, ((uint64_t)w3 << 32) | w2};
}

// Block-wide sum of W extension values per thread (blockDim a multiple of 64, sh at least
// blockDim / 64 x W values); result valid in thread 0. Contains barriers before and after the
// exchange through sh, so every thread's earlier LDS reads are done when it returns.
template <int W>
__device__ __forceinline__ void block_sum(ef2 (&v)[W], ef2* sh) {
#pragma unroll
  for (int q = 0; q < W; q++) v[q] = wave_sum(v[q]);
  const uint32_t wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int q = 0; q < W; q++) sh[wave * W + q] = v[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < W; q++) {
      ef2 t = sh[q];
      for (uint32_t w = 1; w < nw; w++) t = ef_add(t, sh[w * W + q]);
      v[q] = t;
    }
  }
}

// One deferred reduction (RedJob, w <= 16 values) by one block: strided partial sums, then a
// block sum per group of four values.
__device__ __forceinline__ void red_job_body(const RedJob& r, ef2* sh) {
  for (uint32_t q0 = 0; q0 < r.w; q0 += 4) {
    const uint32_t wq = r.w - q0 < 4 ? r.w - q0 : 4;
    ef2 acc[4] = {ef_zero(), ef_zero(), ef_zero(), ef_zero()};
    for (uint32_t b = threadIdx.x; b < r.nblk; b += blockDim.x) {
#pragma unroll
      for (uint32_t q = 0; q < 4; q++)
        if (q < wq) acc[q] = ef_add(acc[q], r.part[(uint64_t)b * r.w + q0 + q]);
    }
    block_sum<4>(acc, sh);
    if (threadIdx.x == 0) {
#pragma unroll
      for (uint32_t q = 0; q < 4; q++)
        if (q < wq) r.out[q0 + q] = acc[q];
    }
  }
}

// The same reduction by one wave when the partials fit it (nblk <= 64): lane b takes block b's
// values; no barrier. Every lane of the wave must be active.
__device__ __forceinline__ void red_job_wave(const RedJob& r, uint32_t lane) {
  for (uint32_t q = 0; q < r.w; q++) {
    const ef2 v = wave_sum(lane < r.nblk ? r.part[(uint64_t)lane * r.w + q] : ef_zero());
    if (lane == 0) r.out[q] = v;
  }
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

// Entry x of an eq table from its half tables (what eq_expand_group_kernel would store).
// (constant term indices: a loop over nterms would index the table's pointer arrays at run time
// and put a copy of the table in scratch memory)
__device__ __forceinline__ ef2 eq_table_at(const EqTable& t, uint64_t x) {
  const uint64_t bl = x & ((uint64_t(1) << t.klo) - 1), bh = x >> t.klo;
  ef2 v = ef_mul(t.lo[0][bl], t.hi[0][bh]);
  if (t.nterms > 1) v = ef_add(v, ef_mul(t.lo[1][bl], t.hi[1][bh]));
  if (t.nterms > 2) v = ef_add(v, ef_mul(t.lo[2][bl], t.hi[2][bh]));
  return v;
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
  const uint32_t i0 = blockIdx.y * rows_per_chunk, i1 = i0 + rows_per_chunk;
  // unreduced sums of the two components, eight rows' loads in flight at a time
  WideAcc a0 = wide_zero(), a1 = wide_zero();
  uint32_t i = i0;
  for (; i + 8 <= i1; i += 8) {
    uint64_t x[8];
#pragma unroll
    for (int u = 0; u < 8; u++) x[u] = M[(uint64_t)(i + u) * cols + j];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const ef2 wi = w[i + u];
      wide_mul_add(a0, wi.c0, x[u]);
      wide_mul_add(a1, wi.c1, x[u]);
    }
  }
  for (; i < i1; i++) {
    const uint64_t x = M[(uint64_t)i * cols + j];
    wide_mul_add(a0, w[i].c0, x);
    wide_mul_add(a1, w[i].c1, x);
  }
  part[(uint64_t)blockIdx.y * cols + j] = ef2{wide_reduce(a0), wide_reduce(a1)};
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

// The last rounds (at most 2^kScTailLog entries) of one product sumcheck in one workgroup of 1024
// threads, the tables in LDS. Four threads share a pair (lo, hi): thread q evaluates the
// product at t = q (q <= D) and folds table q (q < D); the folded values are written back in
// place after a barrier. The round sums do not feed later rounds (every challenge is an input),
// so each wave keeps per-round partials (a sum over its lanes of one q) and the workgroup adds
// them up once, after the last round. The challenges are read into LDS with the tables, so no
// round waits for a global load.
constexpr int kTailThreads = 1024;
template <int D>
__device__ __forceinline__ void sc_tail_body(const ScTailJob& jb, ef2 (*tab)[1 << kScTailLog],
                                             ef2 (*wpart)[kTailThreads / 64][4], ef2* rch) {
  const uint32_t len = jb.len, tid = threadIdx.x, q = tid & 3, wave = tid >> 6, lane = tid & 63;
  const bool m1 = jb.m1 != 0;  // table 1 = table 0 - 1, not stored
#pragma unroll
  for (int j = 0; j < D; j++) {
    if (m1 && j == 1) continue;
    for (uint32_t i = tid; i < len; i += blockDim.x) tab[j][i] = jb.p[j][i];
  }
  if (len > 1 && tid < 31u - __clz(len)) rch[tid] = jb.chal[tid];  // log2(len) rounds
  __syncthreads();
  uint32_t k = 0;
  for (uint32_t cur = len; cur > 1; cur >>= 1, k++) {
    const uint32_t half = cur / 2;
    const ef2 r = rch[k];
    ef2 acc = ef_zero();
    ef2 fv[(4 << kScTailLog) / 2 / kTailThreads];  // folded values held across the barrier
#pragma unroll
    for (uint32_t it = 0; it < (4u << kScTailLog) / 2 / kTailThreads; it++) {
      const uint32_t e = tid + it * kTailThreads, i = e >> 2;
      if (i >= half) break;
      ef2 y[D], flo = ef_zero(), fdiff = ef_zero();
#pragma unroll
      for (int j = 0; j < D; j++) {
        if (m1 && j == 1) {
          y[1] = ef_sub(y[0], ef_one());
          continue;
        }
        const ef2 lo = tab[j][i], hi = tab[j][i + half];
        const ef2 diff = ef_sub(hi, lo);
        // y_j = lo + q diff
        ef2 v = q == 0 ? lo : hi;
        if (q >= 2) v = ef_add(v, diff);
        if (q == 3) v = ef_add(v, diff);
        y[j] = v;
        if ((int)q == j) {  // table q's pair, folded below: one product per thread, not one per table
          flo = lo;
          fdiff = diff;
        }
      }
      if ((int)q < D) fv[it] = ef_add(flo, ef_mul(r, fdiff));
      if ((int)q <= D) {
        ef2 prod = y[0];
#pragma unroll
        for (int j = 1; j < D; j++) prod = ef_mul(prod, y[j]);
        acc = ef_add(acc, prod);
      }
    }
    __syncthreads();  // every read of this round's tables is done
#pragma unroll
    for (uint32_t it = 0; it < (4u << kScTailLog) / 2 / kTailThreads; it++) {
      const uint32_t e = tid + it * kTailThreads, i = e >> 2;
      if (i >= half) break;
      if ((int)q < D && !(m1 && q == 1)) tab[q][i] = fv[it];
    }
    // lanes with equal q (stride 4) summed: offsets 32, 16, 8, 4
#pragma unroll
    for (int off = 32; off >= 4; off >>= 1) {
      const ef2 o{(uint64_t)__shfl_xor((unsigned long long)acc.c0, off),
                  (uint64_t)__shfl_xor((unsigned long long)acc.c1, off)};
      acc = ef_add(acc, o);
    }
    if (lane < 4) wpart[k][wave][lane] = acc;
    __syncthreads();  // folded tables and partials visible
  }
  // round k, value q: sum over the waves
  const uint32_t nw = blockDim.x >> 6;
  if (tid < k * (D + 1)) {
    const uint32_t kk = tid / (D + 1), qq = tid % (D + 1);
    ef2 t = ef_zero();
    for (uint32_t w = 0; w < nw; w++) t = ef_add(t, wpart[kk][w][qq]);
    jb.rounds[tid] = t;
  }
  if (tid < (uint32_t)D) jb.finals[tid] = (m1 && tid == 1) ? ef_sub(tab[0][0], ef_one()) : tab[tid][0];
}

// The tails of several sumchecks, one workgroup each (blockIdx.x = job); blocks past njobs run
// the deferred reductions `reds`.
__global__ __launch_bounds__(kTailThreads) void sc_tail_group_kernel(ScTail a) {
  __shared__ ef2 tab[3][1 << kScTailLog];
  __shared__ ef2 wpart[kScTailLog][kTailThreads / 64][4];
  __shared__ ef2 rch[kScTailLog];
  if (blockIdx.x >= a.nj) {
    red_job_body(a.r[blockIdx.x - a.nj], &wpart[0][0][0]);
    return;
  }
  const ScTailJob& j = a.j[blockIdx.x];
  if (j.d == 3)
    sc_tail_body<3>(j, tab, wpart, rch);
  else
    sc_tail_body<2>(j, tab, wpart, rch);
}

hipError_t launch_sc_tail_group(const ScTail& tail, hipStream_t st) {
  if (tail.nj + tail.nr) hipLaunchKernelGGL(sc_tail_group_kernel, dim3(tail.nj + tail.nr), dim3(kTailThreads), 0, st, tail);
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
                                   uint32_t s, uint64_t W, uint64_t* __restrict__ bi, uint8_t* __restrict__ bi8) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= W * s) return;
  const uint64_t i = t / s;
  const uint32_t k = (uint32_t)(t % s);
  const uint64_t bit = (enc[(i >> 6) * ld + idx[k]] >> (i & 63)) & 1;
  bi[t] = bit;
  bi8[t] = (uint8_t)bit;
}

// g2[i] = sum_k eq_r2[k] * bi[i * s + k] (ligesis.rs:445-455): one wave per i, its lanes
// reading consecutive words of the row (coalesced), then a butterfly sum across the wave.

__device__ __forceinline__ void g2_body(const uint64_t* __restrict__ bi, uint32_t s, uint64_t W,
                                        const ef2* __restrict__ eq_r2, ef2* __restrict__ g2, ef2* __restrict__ g2b,
                                        uint32_t blk) {
  const uint64_t i = ((uint64_t)blk * blockDim.x + threadIdx.x) >> 6;
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
__device__ __forceinline__ void vvec_body(const ef2* __restrict__ eqz1, uint64_t W, ef2* __restrict__ v,
                                          uint32_t blk) {
  const uint64_t i = (uint64_t)blk * blockDim.x + threadIdx.x;
  if (i >= W) return;
  const uint32_t k = i & 63;
  v[i] = ef_mul_base(eqz1[i >> 6], gl_canon(uint64_t(1) << k));
}

// w[k] = prod_t (1 - r3[t] + r3[t] * g^((2^t k) mod 2n)) (rs_weights, ligesis.rs:553-567);
// tw[e] = g^e for e < n, and g^(e + n) = -g^e. Four lanes share k: lane q multiplies the
// factors t = q mod 4, then two butterfly steps multiply the four partial products.
__device__ __forceinline__ void rs_weights_body(const ef2* __restrict__ r3, int log2n, uint64_t n,
                                                const uint64_t* __restrict__ tw, ef2* __restrict__ w,
                                                uint32_t blk) {
  const uint64_t gid = (uint64_t)blk * blockDim.x + threadIdx.x;
  const uint64_t k = gid >> 2;
  const int q = (int)(gid & 3);
  const bool live = k < n;  // the same for the four lanes of k; every lane takes the shuffles
  const uint64_t mask = 2 * n - 1;
  ef2 prod = ef_one();
  if (live) {
    for (int t = q; t < log2n; t += 4) {
      const uint64_t e = (k << t) & mask;
      const uint64_t gp = e < n ? tw[e] : gl_neg(tw[e - n]);
      const ef2 r = r3[t];
      prod = ef_mul(prod, ef_add(ef_one_minus(r), ef_mul_base(r, gp)));
    }
  }
#pragma unroll
  for (int off = 1; off < 4; off <<= 1) {
    const ef2 o{(uint64_t)__shfl_xor((unsigned long long)prod.c0, off),
                (uint64_t)__shfl_xor((unsigned long long)prod.c1, off)};
    prod = ef_mul(prod, o);
  }
  if (live && q == 0) w[k] = prod;
}

// The elementwise inputs of the Stage A sumchecks in one launch, block ranges in order: g2 (and
// its copy), the zeroed sel vector, v, the RS weights, RS(a-bar) interleaved from its planes.
__global__ __launch_bounds__(256) void prep_group_kernel(PrepJobs a) {
  uint32_t b = blockIdx.x;
  if (b < a.b_g2) return g2_body(a.bi, a.s, a.W, a.eq_r2, a.g2, a.g2b, b);
  b -= a.b_g2;
  if (b < a.b_zero) {
    for (uint64_t i = (uint64_t)b * blockDim.x + threadIdx.x; i < a.sel_n; i += (uint64_t)a.b_zero * blockDim.x)
      a.sel[i] = ef_zero();
    return;
  }
  b -= a.b_zero;
  if (b < a.b_vv) return vvec_body(a.eqz1, a.W, a.vv, b);
  b -= a.b_vv;
  if (b < a.b_rs) return rs_weights_body(a.r3, a.log2n, a.n, a.tw, a.wv, b);
  b -= a.b_rs;
  const uint64_t i = (uint64_t)b * blockDim.x + threadIdx.x;
  if (i < a.pn) a.ra[i] = ef2{a.c0[i], a.c1[i]};
}

// sel[idx[k]] = eq_r2[k] (ligesis.rs:488-491) and y4 = sum_k eq_r2[k] ra[idx[k]] (ligesis.rs:456),
// the only nonzero terms of <sel, ra>; one workgroup.
__global__ __launch_bounds__(256) void sel_y4_kernel(const uint32_t* __restrict__ idx, const ef2* __restrict__ eq_r2,
                                                     uint32_t s, const ef2* __restrict__ ra, ef2* __restrict__ sel,
                                                     ef2* __restrict__ y4) {
  __shared__ ef2 sh[4];
  ef2 acc[1] = {ef_zero()};
  for (uint32_t k = threadIdx.x; k < s; k += blockDim.x) {
    sel[idx[k]] = eq_r2[k];
    acc[0] = ef_add(acc[0], ef_mul(eq_r2[k], ra[idx[k]]));
  }
  block_sum<1>(acc, sh);
  if (threadIdx.x == 0) *y4 = acc[0];
}

// ---------------- product sumcheck round (sumcheck.rs:37-84) ----------------
// One round of one product sumcheck over its pairs (lo, hi) = (p[i], p[i + half]):
// g(t) = sum_i prod_j ((1 - t) lo_j + t hi_j) for t = 0..D goes to the block's partial sums, and
// every table folds in place with the round's challenge, p[i] = lo + r (hi - lo) (the high
// variable is bound first).
// M1: the second table is the first minus one and is not stored (its pair and fold follow from
// the first table's).
// A value that is the same in every lane, moved to scalar registers.
__device__ __forceinline__ ef2 ef_uniform(ef2 x) {
  auto u = [](uint64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
  };
  return ef2{u(x.c0), u(x.c1)};
}

template <int D, bool M1>
__device__ __forceinline__ void sc_round_body(ef2* __restrict__ p0, ef2* __restrict__ p1, ef2* __restrict__ p2,
                                              uint64_t half, const ef2* __restrict__ r_ptr, ef2* __restrict__ part,
                                              uint32_t blk, uint32_t nblk, ef2* sh) {
  ef2* ps[3] = {p0, p1, p2};
  const ef2 r = *r_ptr;
  ef2 acc[D + 1];
#pragma unroll
  for (int q = 0; q <= D; q++) acc[q] = ef_zero();
  for (uint64_t i = (uint64_t)blk * blockDim.x + threadIdx.x; i < half; i += (uint64_t)nblk * blockDim.x) {
    ef2 y[D], diff[D];
#pragma unroll
    for (int j = 0; j < D; j++) {
      if (M1 && j == 1) {
        y[1] = ef_sub(y[0], ef_one());
        diff[1] = diff[0];
        continue;
      }
      y[j] = ps[j][i];
      diff[j] = ef_sub(ps[j][i + half], y[j]);
      ps[j][i] = ef_add(y[j], ef_mul(r, diff[j]));
    }
    // products at t = 0..D, stepping each factor by its difference
#pragma unroll
    for (int q = 0; q <= D; q++) {
      ef2 prod = y[0];
#pragma unroll
      for (int j = 1; j < D; j++) prod = ef_mul(prod, y[j]);
      acc[q] = ef_add(acc[q], prod);
      if (q < D) {
#pragma unroll
        for (int j = 0; j < D; j++) y[j] = ef_add(y[j], diff[j]);
      }
    }
  }
  block_sum<D + 1>(acc, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q <= D; q++) part[(uint64_t)blk * (D + 1) + q] = acc[q];
  }
}

// Rounds tau = 0..3 of B_I's two sumchecks, from the 0/1 bytes (ligesis.rs:405-409, 513-522).
// After tau rounds the first table at x is b(x) = sum_y W(y) bits[y 2^(nv - tau) + x] (y the
// bound tau-bit prefix), looked up in an LDS copy of the 2^(2^tau) subset sums of W (wts)
// indexed by the 2^tau bits. The eq table is sum_q lo_q[x mod 2^klo] hi_q[x >> klo] (hi_q already folded
// over the bound variables): a block covers 256 kpt consecutive pairs inside one 2^klo run, so
// the high factors of both halves, Ha_q and Hb_q, are the same for the whole block and come out
// of its sums. Per pair only the low factors L_q multiply:
//  * binary check, sum eq b (b - 1) (F = 0, one term): with b(t) = b_lo + t db,
//    b(t) (b(t) - 1) = c0 + t c1 + t^2 c2 (c0 = b_lo (b_lo - 1), c1 = db (2 b_lo - 1),
//    c2 = db^2) and eq(t) = L H(t), H(t) = Ha + t (Hb - Ha), so s(t) = H(t) (C0 + t C1 + t^2 C2)
//    with C_k = sum L c_k. In round 0 (bits) the product vanishes unless b_lo != b_hi, where it
//    is t (t - 1): s = 0, 0, 2 (2 Hb - Ha) U, 6 (3 Hb - 2 Ha) U with U the sum of L over those
//    pairs;
//  * batched opening, sum b G (F = 1, up to three terms): s(t) = sum_q H_q(t) (A_q + t (B_q -
//    A_q)) with A_q = sum L_q b_lo, B_q = sum L_q b_hi; in round 0 A_q, B_q are sums of L_q over
//    the set bits.
// The values are the ones the product rule gives entry by entry. With p[0] set (the last such
// round) both tables are also folded with the round's challenge for the rounds that read tables.
template <int TAU, int F>
__device__ __forceinline__ void sc_bits_body(const ScRoundJob& j, uint32_t blk, ef2* wl, ef2* sh) {
  constexpr int NY = 1 << TAU, NM = 1 << NY;
  constexpr int D = F ? 2 : 3;
  // binary check, rounds 1 and 2: the quadratic's coefficients (c0, c1, c2) for every pair of
  // prefix-bit patterns follow the subset sums in wts (host-made), so a pair costs three products
  constexpr bool CT = F == 0 && (TAU == 1 || TAU == 2);
  constexpr int NWL = NM + (CT ? 3 * NM * NM : 0);
  static_assert(NWL <= 1024, "LDS table size");
  if (TAU > 0) {  // one coalesced copy of the host-made tables
    for (uint32_t m = threadIdx.x; m < (uint32_t)NWL; m += blockDim.x) wl[m] = j.wts[m];
    __syncthreads();
  }
  // terms: the binary check's eq table has one, the batched opening's three (checked on the
  // host), so every load below is unconditional
  constexpr int NT = F ? 3 : 1;
  const EqTable* __restrict__ T = j.eqt;
  const uint32_t klo = T->klo;
  const uint64_t half = j.half, stride = 2 * half, lmask = (uint64_t(1) << klo) - 1;
  const uint64_t base = (uint64_t)blk * blockDim.x * j.kpt;
  const uint64_t bh = base >> klo, boff = half >> klo;
  // The high factors are the same for the whole block: they are read again after the loop
  // rather than held in vector registers across it, and their fold with r lives in scalar ones.
  const ef2* lo_tab[NT];
#pragma unroll
  for (int q = 0; q < NT; q++) lo_tab[q] = T->lo[q];
  const ef2 one = ef_one();
  ef2* __restrict__ bout = j.p[0];
  ef2* __restrict__ eout = j.p[D - 1];
  ef2 Hr[NT];  // high factors folded with r (the next round's eq table; the last bits round)
#pragma unroll
  for (int q = 0; q < NT; q++) Hr[q] = ef_zero();
  ef2 r = ef_zero();
  if (bout) {
    r = ef_uniform(*j.r);
#pragma unroll
    for (int q = 0; q < NT; q++) {
      const ef2 ha = T->hi[q][bh], hb = T->hi[q][bh + boff];
      Hr[q] = ef_uniform(ef_add(ha, ef_mul(r, ef_sub(hb, ha))));
    }
  }
  // accumulators: F = 0: [C0, C1, C2] (round 0: [U]); F = 1: [A_q, B_q] per term
  constexpr int NA = F ? 2 * NT : (TAU ? 3 : 1);
  ef2 acc[NA];
#pragma unroll
  for (int q = 0; q < NA; q++) acc[q] = ef_zero();
  for (uint32_t k = 0; k < j.kpt; k++) {
    const uint64_t i = base + (uint64_t)k * blockDim.x + threadIdx.x;
    ef2 L[NT];
#pragma unroll
    for (int q = 0; q < NT; q++) L[q] = lo_tab[q][i & lmask];
    ef2 blo, bhi;
    if (TAU == 0) {
      const uint32_t xlo = j.bits[i], xhi = j.bits[i + half];
      if (F == 0) {
        if (xlo != xhi) acc[0] = ef_add(acc[0], L[0]);
      } else {
#pragma unroll
        for (int q = 0; q < NT; q++) {
          if (xlo) acc[2 * q] = ef_add(acc[2 * q], L[q]);
          if (xhi) acc[2 * q + 1] = ef_add(acc[2 * q + 1], L[q]);
        }
      }
      blo = xlo ? one : ef_zero();
      bhi = xhi ? one : ef_zero();
    } else {
      uint32_t mlo = 0, mhi = 0;
#pragma unroll
      for (int y = 0; y < NY; y++) {
        mlo |= (uint32_t)j.bits[(uint64_t)y * stride + i] << y;
        mhi |= (uint32_t)j.bits[(uint64_t)y * stride + i + half] << y;
      }
      blo = wl[mlo];
      bhi = wl[mhi];
      const ef2 db = ef_sub(bhi, blo);
      if (CT) {
        const ef2* ce = &wl[NM + 3 * (mlo * NM + mhi)];
        acc[0] = ef_add(acc[0], ef_mul(L[0], ce[0]));
        acc[1] = ef_add(acc[1], ef_mul(L[0], ce[1]));
        acc[2] = ef_add(acc[2], ef_mul(L[0], ce[2]));
      } else if (F == 0) {
        const ef2 c0 = ef_mul(blo, ef_sub(blo, one));
        const ef2 c1 = ef_mul(db, ef_sub(ef_add(blo, blo), one));
        const ef2 c2 = ef_mul(db, db);
        acc[0] = ef_add(acc[0], ef_mul(L[0], c0));
        acc[1] = ef_add(acc[1], ef_mul(L[0], c1));
        acc[2] = ef_add(acc[2], ef_mul(L[0], c2));
      } else {
#pragma unroll
        for (int q = 0; q < NT; q++) {
          acc[2 * q] = ef_add(acc[2 * q], ef_mul(L[q], blo));
          acc[2 * q + 1] = ef_add(acc[2 * q + 1], ef_mul(L[q], bhi));
        }
      }
    }
    if (bout) {
      bout[i] = ef_add(blo, ef_mul(r, ef_sub(bhi, blo)));
      ef2 e = ef_mul(L[0], Hr[0]);
#pragma unroll
      for (int q = 1; q < NT; q++) e = ef_add(e, ef_mul(L[q], Hr[q]));
      eout[i] = e;
    }
  }
  block_sum<NA>(acc, sh);
  if (threadIdx.x == 0) {
    ef2 Ha[NT], Hb[NT];
#pragma unroll
    for (int q = 0; q < NT; q++) {
      Ha[q] = T->hi[q][bh];
      Hb[q] = T->hi[q][bh + boff];
    }
    ef2 s[4];
    if (F == 0 && TAU == 0) {
      const ef2 L = ef_mul(Ha[0], acc[0]), H = ef_mul(Hb[0], acc[0]);
      const ef2 h2 = ef_add(H, H), l2 = ef_add(L, L);
      const ef2 s2 = ef_sub(h2, L), s3 = ef_sub(ef_add(h2, H), l2);
      const ef2 s3x2 = ef_add(s3, s3);
      s[0] = ef_zero();
      s[1] = ef_zero();
      s[2] = ef_add(s2, s2);
      s[3] = ef_add(s3x2, ef_add(s3x2, s3x2));
    } else if (F == 0) {
      const ef2 dH = ef_sub(Hb[0], Ha[0]);
      ef2 H = Ha[0], tt = ef_zero();  // H(t), t as a field element
#pragma unroll
      for (int t = 0; t < 4; t++) {
        // C0 + t C1 + t^2 C2
        const ef2 poly = ef_add(acc[0], ef_mul(tt, ef_add(acc[1], ef_mul(tt, acc[2]))));
        s[t] = ef_mul(H, poly);
        H = ef_add(H, dH);
        tt = ef_add(tt, one);
      }
    } else if (TAU == 0) {
      ef2 A0 = ef_zero(), A1 = ef_zero(), C = ef_zero();
#pragma unroll
      for (int q = 0; q < NT; q++) {
        A0 = ef_add(A0, ef_mul(Ha[q], acc[2 * q]));
        A1 = ef_add(A1, ef_mul(Hb[q], acc[2 * q + 1]));
        C = ef_add(C, ef_add(ef_mul(Ha[q], acc[2 * q + 1]), ef_mul(Hb[q], acc[2 * q])));
      }
      const ef2 a1x2 = ef_add(A1, A1), cx2 = ef_add(C, C);
      s[0] = A0;
      s[1] = A1;
      s[2] = ef_add(ef_sub(ef_add(a1x2, a1x2), cx2), A0);
      s[3] = ef_zero();
    } else {
#pragma unroll
      for (int t = 0; t < 3; t++) s[t] = ef_zero();
      s[3] = ef_zero();
#pragma unroll
      for (int q = 0; q < NT; q++) {
        const ef2 dH = ef_sub(Hb[q], Ha[q]), dA = ef_sub(acc[2 * q + 1], acc[2 * q]);
        ef2 H = Ha[q], A = acc[2 * q];
#pragma unroll
        for (int t = 0; t < 3; t++) {
          s[t] = ef_add(s[t], ef_mul(H, A));
          H = ef_add(H, dH);
          A = ef_add(A, dA);
        }
      }
    }
#pragma unroll
    for (int q = 0; q <= D; q++) j.part[(uint64_t)blk * (D + 1) + q] = s[q];
  }
}

// The bits rounds of one step (every job in the launch is at round TAU): a kernel of their own,
// so that each launch carries two small bodies instead of the whole round code.
template <int TAU>
__global__ __launch_bounds__(256) void sc_bits_group_kernel(ScStep a) {
  __shared__ ef2 sh[256 * 4];
  __shared__ ef2 wl[1024];  // subset sums of the prefix weights (and the binary check's table)
  if (blockIdx.x >= a.round_blocks) {
    red_job_body(a.r[blockIdx.x - a.round_blocks], sh);
    return;
  }
  int lo = 0, hi = (int)a.nj - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (a.j[mid].first_block <= blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  lo = __builtin_amdgcn_readfirstlane(lo);  // uniform: the job's fields load as scalars
  const ScRoundJob& j = a.j[lo];
  const uint32_t blk = blockIdx.x - j.first_block;
  if (j.d == 3) sc_bits_body<TAU, 0>(j, blk, wl, sh);
  else sc_bits_body<TAU, 1>(j, blk, wl, sh);
}

hipError_t launch_sc_bits_group(const ScStep& step, uint32_t tau, hipStream_t st) {
  const dim3 grid(step.round_blocks + step.nr), blk(256);
  switch (tau) {
    case 0: hipLaunchKernelGGL(sc_bits_group_kernel<0>, grid, blk, 0, st, step); break;
    case 1: hipLaunchKernelGGL(sc_bits_group_kernel<1>, grid, blk, 0, st, step); break;
    case 2: hipLaunchKernelGGL(sc_bits_group_kernel<2>, grid, blk, 0, st, step); break;
    default: hipLaunchKernelGGL(sc_bits_group_kernel<3>, grid, blk, 0, st, step); break;
  }
  return hipGetLastError();
}

// One round of several independent sumchecks in one launch (sumchecks run in lockstep): block
// -> job by binary search over the jobs' first blocks.
__global__ __launch_bounds__(256) void sc_round_group_kernel(ScStep a) {
  __shared__ ef2 sh[256 * 4];
  if (blockIdx.x >= a.round_blocks) {
    red_job_body(a.r[blockIdx.x - a.round_blocks], sh);
    return;
  }
  int lo = 0, hi = (int)a.nj - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (a.j[mid].first_block <= blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const ScRoundJob& j = a.j[lo];
  const uint32_t blk = blockIdx.x - j.first_block;
  if (j.d == 3) {
    if (j.m1) sc_round_body<3, true>(j.p[0], nullptr, j.p[2], j.half, j.r, j.part, blk, j.nblk, sh);
    else sc_round_body<3, false>(j.p[0], j.p[1], j.p[2], j.half, j.r, j.part, blk, j.nblk, sh);
  } else {
    sc_round_body<2, false>(j.p[0], j.p[1], nullptr, j.half, j.r, j.part, blk, j.nblk, sh);
  }
}

hipError_t launch_sc_round_group(const ScStep& step, hipStream_t st) {
  if (step.round_blocks + step.nr)
    hipLaunchKernelGGL(sc_round_group_kernel, dim3(step.round_blocks + step.nr), dim3(256), 0, st, step);
  return hipGetLastError();
}

uint32_t sc_round_blocks(uint64_t half) {
  uint64_t nb = (half + 255) / 256;
  return (uint32_t)(nb > 1024 ? 1024 : nb);
}




// ---------------- Basefold batch (basefold.rs:648-856) ----------------
// combined[i] = sum over channels c, chunks j of wt_c[j] * codes_c[j][i], where
// wt_c[j] = coeffs[c] * eq4_c[j]; an extension channel keeps c0 / c1 in rows 2j / 2j+1.
// A workgroup takes 64 consecutive positions i; its four waves split the rows (row j to wave
// j mod 4), accumulate unreduced, and add their four sums through LDS.
__global__ __launch_bounds__(256) void bf_combine_kernel(BfChannels ch, uint64_t D, ef2* __restrict__ out) {
  __shared__ ef2 sh[3][64];
  const uint32_t lane = threadIdx.x & 63, sub = threadIdx.x >> 6;
  const uint64_t i = (uint64_t)blockIdx.x * 64 + lane;
  WideAcc a0 = wide_zero(), a1 = wide_zero();
  if (i < D) {
    for (int c = 0; c < ch.count; c++) {
      const uint64_t* codes = ch.codes[c];
      const ef2* wt = ch.wt[c];
      for (uint32_t j = sub; j < ch.chunks[c]; j += 4) {
        if (ch.ext[c]) {
          const ef2 e = ef_mul(wt[j], ef2{codes[(2 * j) * D + i], codes[(2 * j + 1) * D + i]});
          wide_add(a0, e.c0);
          wide_add(a1, e.c1);
        } else {
          const uint64_t x = codes[j * D + i];
          wide_mul_add(a0, wt[j].c0, x);
          wide_mul_add(a1, wt[j].c1, x);
        }
      }
    }
  }
  const ef2 r{wide_reduce(a0), wide_reduce(a1)};
  if (sub) sh[sub - 1][lane] = r;
  __syncthreads();
  if (sub == 0 && i < D) out[i] = ef_add(ef_add(r, sh[0][lane]), ef_add(sh[1][lane], sh[2][lane]));
}

// m[x] = sum_{b < 2^li} eq4[b] * poly[x * 2^li + b] (MlPoly::fold over the low li variables); a
// workgroup takes 64 consecutive x, its four waves split b (b mod 4) as in bf_combine_kernel.
__global__ __launch_bounds__(256) void fold_low_kernel(const uint64_t* __restrict__ poly, int ext, int li,
                                                       uint64_t nout, const ef2* __restrict__ eq4,
                                                       ef2* __restrict__ m) {
  __shared__ ef2 sh[3][64];
  const uint32_t lane = threadIdx.x & 63, sub = threadIdx.x >> 6;
  const uint64_t x = (uint64_t)blockIdx.x * 64 + lane;
  WideAcc a0 = wide_zero(), a1 = wide_zero();
  if (x < nout) {
    const uint64_t base = x << li;
    for (uint32_t b = sub; b < (1u << li); b += 4) {
      if (ext) {
        const ef2 e = ef_mul(eq4[b], ef2{poly[2 * (base + b)], poly[2 * (base + b) + 1]});
        wide_add(a0, e.c0);
        wide_add(a1, e.c1);
      } else {
        const uint64_t v = poly[base + b];
        wide_mul_add(a0, eq4[b].c0, v);
        wide_mul_add(a1, eq4[b].c1, v);
      }
    }
  }
  const ef2 r{wide_reduce(a0), wide_reduce(a1)};
  if (sub) sh[sub - 1][lane] = r;
  __syncthreads();
  if (sub == 0 && x < nout) m[x] = ef_add(ef_add(r, sh[0][lane]), ef_add(sh[1][lane], sh[2][lane]));
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

// Offset of round k's sum shares in bf_tail_kernel's share table: rounds j < k hold 3 half_j
// shares each (half_j = len0 / 2^(j+2)), so sum_{j<k} 3 half_j = 3 (len0 / 2 - len0 / 2^(k+1)).
__device__ __forceinline__ uint32_t bf_tail_share_off(uint32_t len0, uint32_t k) {
  return 3 * (len0 / 2 - (len0 >> (k + 1)));
}

// One round of bf_tail_kernel (level of len entries, h = len / 2 codeword pairs, half = len / 4
// pairs of each channel's m). The round's work items -- codeword pairs, then each channel's
// pairs, every kind starting at a multiple of 64 so that a wave takes one kind -- are spread
// over the workgroup (item x to thread x mod blockDim, up to three items per thread in the
// first rounds), so a thread's chain of products is one item long once the level is small.
// Codeword item i folds pair (i, i + h); channel item i stores its share m_c[2i] eq_c[i] of the
// round's sum into `share`, folds m_c and (i < half / 2) eq_c. Reads before the barrier, writes
// after it.
__device__ __forceinline__ void bf_tail_round(uint32_t k, uint32_t len0, uint32_t tid, ef2 r, ef2* cw,
                                              const uint64_t* tws, ef2 (*m)[1 << (kBfTailLog - 1)],
                                              ef2 (*eq)[1 << (kBfTailLog - 2)], ef2* lvl, ef2* share) {
  const uint32_t len = len0 >> k, h = len / 2, half = len / 4;
  const uint32_t rc = (h + 63) & ~63u, rh = (half + 63) & ~63u, items = rc + 3 * rh;
  constexpr int kPass = 3;  // items <= 512 + 3 * 256 = 1280 <= 3 * 512
  ef2 v0[kPass], v1[kPass];
#pragma unroll
  for (int p = 0; p < kPass; p++) {
    const uint32_t x = tid + (uint32_t)p * blockDim.x;
    v0[p] = ef_zero();
    v1[p] = ef_zero();
    if (x >= items) continue;
    if (x < rc) {
      if (x < h) {  // eq_fold_codeword_at over the subgroup of order len
        const ef2 a = cw[x], b = cw[x + h];
        const uint64_t ginv = x == 0 ? 1 : gl_neg(tws[(h - x) << k]);
        const ef2 pe = ef_mul_base(ef_add(a, b), GL_INV2);
        const ef2 po = ef_mul_base(ef_sub(a, b), gl_mul(ginv, GL_INV2));
        v0[p] = ef_add(ef_mul(ef_one_minus(r), pe), ef_mul(r, po));
      }
    } else {
      const uint32_t y = x - rc, c = y / rh, i = y % rh;
      if (i < half) {
        const ef2 e = m[c][2 * i], o = m[c][2 * i + 1];
        share[c * half + i] = ef_mul(e, eq[c][i]);
        v0[p] = ef_add(e, ef_mul(r, ef_sub(o, e)));
        if (i < half / 2) v1[p] = ef_add(eq[c][2 * i], eq[c][2 * i + 1]);
      }
    }
  }
  __syncthreads();  // every read of this round's tables is done
#pragma unroll
  for (int p = 0; p < kPass; p++) {
    const uint32_t x = tid + (uint32_t)p * blockDim.x;
    if (x >= items) continue;
    if (x < rc) {
      if (x < h) {
        cw[x] = v0[p];
        lvl[x] = v0[p];
      }
    } else {
      const uint32_t y = x - rc, c = y / rh, i = y % rh;
      if (i < half) m[c][i] = v0[p];
      if (i < half / 2) eq[c][i] = v1[p];
    }
  }
  __syncthreads();
}

// The last Basefold reduction rounds t0..R-1 (level t0 has len0 <= 2^kBfTailLog entries) in one
// workgroup of 512 threads, with the codeword, the three channels' m and eq tables in LDS
// (bf_tail_round per round). The round sums are outputs only (no later round reads them): the
// rounds store their shares in LDS, and the sums are taken after the last round, in two levels
// (chunks of 16 shares, then a chunk sum per round and channel) instead of a workgroup reduction
// inside every round. The challenges and the twiddles of every round (tws[i] = tw[i 2^t0]; round
// t = t0 + k takes g^((h - i) 2^t) = tws[(h - i) 2^k]) are read into LDS with the tables, so no
// round waits for a global load. R - t0 <= kBfTailLog (the code rate is at least 1).
__global__ __launch_bounds__(512) void bf_tail_kernel(BfTail a) {
  constexpr uint32_t kShares = 3u << (kBfTailLog - 1), kChunk = 16;
  __shared__ ef2 cw[1 << kBfTailLog];
  __shared__ ef2 m[3][1 << (kBfTailLog - 1)];
  __shared__ ef2 eq[3][1 << (kBfTailLog - 2)];
  __shared__ ef2 share[kShares];
  __shared__ ef2 csum[kShares / kChunk + 3 * kBfTailLog];
  __shared__ ef2 wp[8][4];
  __shared__ ef2 rch[32];
  __shared__ uint64_t tws[1 << (kBfTailLog - 1)];
  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  // the last grid round's sums: a wave per channel where its partials fit one wave (the usual
  // case: a few blocks per channel), else one channel after the other on the whole block
  if (a.prev[0].nblk <= 64 && a.prev[1].nblk <= 64 && a.prev[2].nblk <= 64) {
    if (wave < 3 && a.prev[wave].nblk) red_job_wave(a.prev[wave], lane);
  } else {
    for (int c = 0; c < 3; c++) {
      if (a.prev[c].nblk) red_job_body(a.prev[c], &wp[0][0]);
      __syncthreads();
    }
  }
  const uint32_t len0 = 1u << (a.logD - a.t0), rounds = a.R - a.t0;
  for (uint32_t i = tid; i < len0; i += blockDim.x) cw[i] = a.cw_in[i];
  for (uint32_t i = tid; i < len0 / 2; i += blockDim.x) tws[i] = a.tw[(uint64_t)i << a.t0];
  if (tid < rounds) rch[tid] = a.chal[a.t0 + tid];
  for (int c = 0; c < 3; c++) {
    for (uint32_t i = tid; i < len0 / 2; i += blockDim.x) m[c][i] = a.m_in[c][i];
    for (uint32_t i = tid; i < len0 / 4; i += blockDim.x) eq[c][i] = a.eq[c][i];
  }
  __syncthreads();
  ef2* lvl = a.lvl_out;
  for (uint32_t k = 0; k < rounds; k++) {
    bf_tail_round(k, len0, tid, rch[k], cw, tws, m, eq, lvl, share + bf_tail_share_off(len0, k));
    lvl += len0 >> (k + 1);
  }
  // level 1: thread x sums chunk j of round k, channel c (chunks of kChunk shares)
  uint32_t x = tid, k = 0, nch = 0, base = 0;
  for (; k < rounds; k++) {
    const uint32_t half = len0 >> (k + 2);
    nch = (half + kChunk - 1) / kChunk;
    if (x < 3 * nch) break;
    x -= 3 * nch;
    base += 3 * nch;
  }
  if (k < rounds) {
    const uint32_t half = len0 >> (k + 2), c = x / nch, j = x % nch;
    const ef2* sp = share + bf_tail_share_off(len0, k) + c * half + j * kChunk;
    const uint32_t n = half - j * kChunk < kChunk ? half - j * kChunk : kChunk;
    ef2 t = ef_zero();
    for (uint32_t q = 0; q < n; q++) t = ef_add(t, sp[q]);
    csum[base + c * nch + j] = t;
  }
  __syncthreads();
  // level 2: thread 3 k + c sums round k's chunk sums of channel c
  if (tid < 3 * rounds) {
    const uint32_t kk = tid / 3, c = tid % 3;
    uint32_t b = 0;
    for (uint32_t q = 0; q < kk; q++) b += 3 * (((len0 >> (q + 2)) + kChunk - 1) / kChunk);
    const uint32_t n = ((len0 >> (kk + 2)) + kChunk - 1) / kChunk;
    ef2 t = ef_zero();
    for (uint32_t q = 0; q < n; q++) t = ef_add(t, csum[b + c * n + q]);
    a.l0[(uint64_t)c * a.R + a.t0 + kk] = t;
  }
}

hipError_t launch_bf_tail(const BfTail& a, hipStream_t st) {
  if (a.t0 < a.R) hipLaunchKernelGGL(bf_tail_kernel, dim3(1), dim3(512), 0, st, a);
  return hipGetLastError();
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


// ---------------- Basefold, one polynomial (basefold.rs:384-516) ----------------
// One reduction round: partial sums of l0 = sum_i m[2i] eq[i] and l1 = sum_i m[2i+1] eq[i]
// (basefold.rs:449-456; both are sent), the low-variable fold m'[i] = m[2i] + r (m[2i+1] - m[2i]),
// the next round's eq table (pair sums), and (row y = 1) the codeword fold of the same round;
// row y = 2 reduces the previous round's partials.
__global__ __launch_bounds__(256) void bf1_round_kernel(Bf1Round a, uint64_t half, const ef2* __restrict__ r_ptr) {
  __shared__ ef2 sh[256 * 4];
  if (blockIdx.y == 2) {
    if (blockIdx.x == 0 && a.prev.nblk) red_job_body(a.prev, sh);
    return;
  }
  const ef2 r = *r_ptr;
  if (blockIdx.y == 1) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.cw_half;
         i += (uint64_t)gridDim.x * blockDim.x)
      eq_fold_codeword_at(a.cw_in, a.cw_half, r, a.cw_tw, a.cw_out, i);
    return;
  }
  ef2 acc[2] = {ef_zero(), ef_zero()};
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < half; i += (uint64_t)gridDim.x * blockDim.x) {
    const ef2 e = a.m_in[2 * i], o = a.m_in[2 * i + 1], q = a.eq[i];
    acc[0] = ef_add(acc[0], ef_mul(e, q));
    acc[1] = ef_add(acc[1], ef_mul(o, q));
    a.m_out[i] = ef_add(e, ef_mul(r, ef_sub(o, e)));
    if (half > 1 && i < half / 2) a.eq_next[i] = ef_add(a.eq[2 * i], a.eq[2 * i + 1]);
  }
  block_sum<2>(acc, sh);
  if (threadIdx.x == 0) {
    a.part[2 * (uint64_t)blockIdx.x] = acc[0];
    a.part[2 * (uint64_t)blockIdx.x + 1] = acc[1];
  }
}

uint32_t bf1_round_blocks(uint64_t half, uint64_t cw_half) { return bf_round3_blocks(half, cw_half); }

hipError_t launch_bf1_round(const Bf1Round& a, uint64_t half, const ef2* r, hipStream_t st) {
  hipLaunchKernelGGL(bf1_round_kernel, dim3(bf1_round_blocks(half, a.cw_half), 3), dim3(256), 0, st, a, half, r);
  return hipGetLastError();
}

// Rounds t0..R-1 in one workgroup of 512 threads: level t0 (at most 2^kBfTailLog entries), m
// and eq in LDS. Per round thread i folds codeword pair (i, i + len/2) (kept in LDS and written
// to the level's place in global memory), and for i < half adds m[2i] eq[i] and m[2i+1] eq[i]
// to the round's sums and folds m and eq; a barrier separates reads from writes. The twiddles
// of every round come from tws[j] = tw[j 2^t0] (round t0 + k takes tws[(h - i) 2^k]).
__global__ __launch_bounds__(512) void bf1_tail_kernel(Bf1Tail a) {
  __shared__ ef2 cw[1 << kBfTailLog];
  __shared__ ef2 m[1 << (kBfTailLog - 1)];
  __shared__ ef2 eq[1 << (kBfTailLog - 2)];
  __shared__ ef2 wp[kBfTailLog][8][2];
  __shared__ ef2 rch[kBfTailLog];
  __shared__ uint64_t tws[1 << (kBfTailLog - 1)];
  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  if (a.prev.nblk) red_job_body(a.prev, &wp[0][0][0]);
  __syncthreads();
  const uint32_t len0 = 1u << (a.logD - a.t0), mlen0 = 1u << (a.R - a.t0);
  for (uint32_t i = tid; i < len0; i += blockDim.x) cw[i] = a.cw_in[i];
  for (uint32_t i = tid; i < len0 / 2; i += blockDim.x) tws[i] = a.tw[(uint64_t)i << a.t0];
  if (tid < a.R - a.t0) rch[tid] = a.chal[a.t0 + tid];
  for (uint32_t i = tid; i < mlen0; i += blockDim.x) m[i] = a.m_in[i];
  for (uint32_t i = tid; i < mlen0 / 2; i += blockDim.x) eq[i] = a.eq[i];
  __syncthreads();
  ef2* lvl = a.lvl_out;
  for (uint32_t t = a.t0; t < a.R; t++) {
    const uint32_t k = t - a.t0, len = len0 >> k, h = len / 2, half = (mlen0 >> k) / 2;
    const ef2 r = rch[k];
    ef2 cwv = ef_zero(), mv = ef_zero(), ev = ef_zero(), acc0 = ef_zero(), acc1 = ef_zero();
    if (tid < h) {
      const ef2 x = cw[tid], nx = cw[tid + h];
      const uint64_t ginv = tid == 0 ? 1 : gl_neg(tws[(h - tid) << k]);
      const ef2 pe = ef_mul_base(ef_add(x, nx), GL_INV2);
      const ef2 po = ef_mul_base(ef_sub(x, nx), gl_mul(ginv, GL_INV2));
      cwv = ef_add(ef_mul(ef_one_minus(r), pe), ef_mul(r, po));
    }
    if (tid < half) {
      const ef2 e = m[2 * tid], o = m[2 * tid + 1], q = eq[tid];
      acc0 = ef_mul(e, q);
      acc1 = ef_mul(o, q);
      mv = ef_add(e, ef_mul(r, ef_sub(o, e)));
      if (tid < half / 2) ev = ef_add(eq[2 * tid], eq[2 * tid + 1]);
    }
    __syncthreads();  // every read of this round's tables is done
    if (tid < h) {
      cw[tid] = cwv;
      lvl[tid] = cwv;
    }
    if (tid < half) m[tid] = mv;
    if (tid < half / 2) eq[tid] = ev;
    acc0 = wave_sum(acc0);
    acc1 = wave_sum(acc1);
    if (lane == 0) {
      wp[k][wave][0] = acc0;
      wp[k][wave][1] = acc1;
    }
    lvl += h;
    __syncthreads();
  }
  const uint32_t rounds = a.R - a.t0, nw = blockDim.x >> 6;
  if (tid < 2 * rounds) {
    const uint32_t k = tid / 2, q = tid % 2;
    ef2 t = ef_zero();
    for (uint32_t w = 0; w < nw; w++) t = ef_add(t, wp[k][w][q]);
    a.lpair[2 * (uint64_t)(a.t0 + k) + q] = t;
  }
}

hipError_t launch_bf1_tail(const Bf1Tail& a, hipStream_t st) {
  if (a.t0 < a.R) hipLaunchKernelGGL(bf1_tail_kernel, dim3(1), dim3(512), 0, st, a);
  return hipGetLastError();
}


// eq_fold_codeword (basefold.rs:201-218): out[i] = (1 - r)(x + nx)/2 + r (x - nx) g^-i / 2 with
// x = v[i], nx = v[i + h]; g^-i = -g^(h - i) for 0 < i < h, tw[e] = g^e (e < h).


// ---------------- gathers for openings ----------------
__global__ void gather_parts_kernel(const GatherPart* __restrict__ parts, uint32_t nparts,
                                    const uint32_t* __restrict__ pos, uint64_t total, uint64_t* __restrict__ dst) {
  for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < total; w += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t lo = 0, hi = nparts - 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) / 2;
      if (parts[mid].first <= w) lo = mid;
      else hi = mid - 1;
    }
    const GatherPart& g = parts[lo];
    const uint64_t t = w - g.first, run = (uint64_t)g.rows * g.width;
    const uint64_t k = t / run, rem = t - k * run, rr = rem / g.width, q = rem - rr * g.width;
    dst[w] = g.src[(uint64_t)pos[g.pos_at + k] * g.pmul + rr * g.rstride + q];
  }
}

#define LGS_GRID(n) dim3((uint32_t)(((n) + 255) / 256)), dim3(256)

hipError_t launch_bits_gather(const uint64_t* enc, uint64_t ld, const uint32_t* idx, uint32_t s, uint64_t W,
                              uint64_t* bi, uint8_t* bi8, hipStream_t st) {
  hipLaunchKernelGGL(bits_gather_kernel, LGS_GRID(W * s), 0, st, enc, ld, idx, s, W, bi, bi8);
  return hipGetLastError();
}
hipError_t launch_bf_combine(const BfChannels& ch, uint64_t D, ef2* out, hipStream_t st) {
  hipLaunchKernelGGL(bf_combine_kernel, dim3((uint32_t)((D + 63) / 64)), dim3(256), 0, st, ch, D, out);
  return hipGetLastError();
}
hipError_t launch_fold_low(const uint64_t* poly, int ext, int li, uint64_t nout, const ef2* eq4, ef2* m,
                           hipStream_t st) {
  hipLaunchKernelGGL(fold_low_kernel, dim3((uint32_t)((nout + 63) / 64)), dim3(256), 0, st, poly, ext, li, nout,
                     eq4, m);
  return hipGetLastError();
}
hipError_t launch_prep_group(PrepJobs a, hipStream_t st) {
  a.b_g2 = (uint32_t)((a.W * 64 + 255) / 256);
  a.b_zero = 256;
  a.b_vv = (uint32_t)((a.W + 255) / 256);
  a.b_rs = (uint32_t)((4 * a.n + 255) / 256);  // four lanes per weight
  const uint32_t b_pl = (uint32_t)((a.pn + 255) / 256);
  hipLaunchKernelGGL(prep_group_kernel, dim3(a.b_g2 + a.b_zero + a.b_vv + a.b_rs + b_pl), dim3(256), 0, st, a);
  return hipGetLastError();
}
hipError_t launch_sel_y4(const uint32_t* idx, const ef2* eq_r2, uint32_t s, const ef2* ra, ef2* sel, ef2* y4,
                         hipStream_t st) {
  hipLaunchKernelGGL(sel_y4_kernel, dim3(1), dim3(256), 0, st, idx, eq_r2, s, ra, sel, y4);
  return hipGetLastError();
}
hipError_t launch_gather_parts(const GatherPart* parts, uint32_t nparts, const uint32_t* pos, uint64_t total,
                               uint64_t* dst, hipStream_t st) {
  if (total == 0 || nparts == 0) return hipSuccess;
  uint64_t blocks = (total + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(gather_parts_kernel, dim3((uint32_t)blocks), dim3(256), 0, st, parts, nparts, pos, total, dst);
  return hipGetLastError();
}

}  // namespace lgs
