// BLAKE3 column-Merkle trees with the semantics of utils/src/merkle.rs:49-74 (rs_merkle over
// pre-hashed leaves): leaf p = blake3(serialize(column p)), parent = blake3(left || right).
//
// Tree storage: all layers of digests (8 x u32, little-endian words = the 32 digest bytes)
// concatenated, leaves first: layer k starts at sum_{t<k} L >> t, the root is the last node.
// Leaf counts are powers of two on this path (2n, Basefold domains, folded levels).
//
// One workgroup of 4 x 2^g threads (g = 6..8, 8 in grouped launches) reduces up to 2^g nodes of
// a layer by up to g levels in LDS, so a tree over L leaves takes ceil(log2(L) / g) launches; the
// first launch also hashes the leaves.
#include "blake3_dev.hpp"
#include "kernels.hpp"

namespace lgs {

// Every compression runs on the four lanes of a quad (b3_compress4): a workgroup of 4 W threads
// holds W nodes, thread t works on node t / 4 as column t % 4.

// Leaf p = words M[r * rs + p * cs] for r < rows (rows * 8 <= 1024 bytes: a single chunk);
// lane c of the quad gets h0 = digest word c, h1 = word 4 + c.
__device__ __forceinline__ void leaf_digest4(const uint64_t* __restrict__ M, uint32_t rows, uint64_t rs, uint64_t cs,
                                             uint64_t p, uint32_t c, uint32_t& h0, uint32_t& h1) {
  b3_init4(h0, h1, c);
  const uint32_t nblocks = (rows + 7) / 8;
  for (uint32_t b = 0; b < nblocks; b++) {
    uint32_t m[16];
    const uint32_t wcount = (rows - 8 * b) < 8 ? (rows - 8 * b) : 8;
#pragma unroll
    for (uint32_t w = 0; w < 8; w++) {
      const uint64_t v = w < wcount ? M[(uint64_t)(8 * b + w) * rs + p * cs] : 0;
      m[2 * w] = (uint32_t)v;
      m[2 * w + 1] = (uint32_t)(v >> 32);
    }
    const uint32_t flags = (b == 0 ? B3_CHUNK_START : 0) | (b + 1 == nblocks ? (B3_CHUNK_END | B3_ROOT) : 0);
    b3_compress4(h0, h1, m, wcount * 8, flags, c);
  }
}

__device__ __forceinline__ void merkle_reduce_body(const uint64_t* __restrict__ M, uint32_t rows, uint64_t rs,
                                                   uint64_t cs, int from_leaves, uint32_t* __restrict__ nodes,
                                                   uint64_t child_off, uint64_t child_count, int levels,
                                                   uint32_t blk, uint32_t* sh, uint32_t* root_out = nullptr) {
  const uint32_t node = threadIdx.x >> 2, c = threadIdx.x & 3, width = blockDim.x >> 2;
  const uint64_t base = (uint64_t)blk * width;
  uint32_t local_n = (uint32_t)((child_count - base) < width ? (child_count - base) : width);
  if (node < local_n) {  // whole quads: a quad's four lanes share the node
    uint32_t h0, h1;
    const uint64_t at = (child_off + base + node) * 8;
    if (from_leaves) {
      leaf_digest4(M, rows, rs, cs, base + node, c, h0, h1);
      nodes[at + c] = h0;
      nodes[at + 4 + c] = h1;
    } else {
      h0 = nodes[at + c];
      h1 = nodes[at + 4 + c];
    }
    sh[node * 8 + c] = h0;
    sh[node * 8 + 4 + c] = h1;
  }
  uint64_t off = child_off, count = child_count, blk_base = base;
  for (int lv = 0; lv < levels && local_n > 1; lv++) {
    __syncthreads();
    const uint32_t half = local_n / 2;
    uint32_t h0 = 0, h1 = 0;
    if (node < half) {  // parent = blake3(left || right)
      uint32_t m[16];
#pragma unroll
      for (int i = 0; i < 16; i++) m[i] = sh[(2 * node) * 8 + i];
      b3_init4(h0, h1, c);
      b3_compress4(h0, h1, m, 64, B3_CHUNK_START | B3_CHUNK_END | B3_ROOT, c);
    }
    __syncthreads();
    const uint64_t poff = off + count;
    blk_base >>= 1;
    if (node < half) {
      sh[node * 8 + c] = h0;
      sh[node * 8 + 4 + c] = h1;
      const uint64_t at = (poff + blk_base + node) * 8;
      nodes[at + c] = h0;
      nodes[at + 4 + c] = h1;
    }
    off = poff;
    count >>= 1;
    local_n = half;
  }
  if (root_out && count == 1 && node == 0) {  // this launch reached the root (quad of node 0)
    root_out[c] = sh[c];
    root_out[4 + c] = sh[4 + c];
  }
}

__global__ __launch_bounds__(1024) void merkle_reduce_kernel(
    const uint64_t* __restrict__ M, uint32_t rows, uint64_t rs, uint64_t cs, int from_leaves,
    uint32_t* __restrict__ nodes, uint64_t child_off, uint64_t child_count, int levels, uint32_t* root_out) {
  __shared__ uint32_t sh[256 * 8];
  merkle_reduce_body(M, rows, rs, cs, from_leaves, nodes, child_off, child_count, levels, blockIdx.x, sh, root_out);
}

// One reduction round of many trees in one launch (batched commit): 256 nodes per workgroup.
__global__ __launch_bounds__(1024) void merkle_group_kernel(const MerkleJob* __restrict__ jobs, int njobs) {
  __shared__ uint32_t sh[256 * 8];
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (jobs[mid].first_block <= blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const MerkleJob& jb = jobs[lo];
  merkle_reduce_body(jb.M, jb.rows, jb.rs, jb.cs, jb.from_leaves, jb.nodes, jb.child_off, jb.child_count,
                     jb.levels, blockIdx.x - jb.first_block, sh);
}

hipError_t launch_merkle_group(const MerkleJob* jobs, int njobs, uint32_t total_blocks, hipStream_t st) {
  if (total_blocks) hipLaunchKernelGGL(merkle_group_kernel, dim3(total_blocks), dim3(1024), 0, st, jobs, njobs);
  return hipGetLastError();
}

static int ilog2_u64(uint64_t x) {
  int l = 0;
  while (((uint64_t)1 << l) < x) l++;
  return l;
}

hipError_t launch_merkle(const uint64_t* M, uint32_t rows, uint64_t rs, uint64_t cs,
                         uint64_t nleaves, uint32_t* nodes, hipStream_t st, int lg_block, uint32_t* root_out) {
  if (lg_block < 6 || lg_block > 8) return hipErrorInvalidValue;
  const uint32_t width = 1u << lg_block;
  uint64_t off = 0, count = nleaves;
  int total_levels = ilog2_u64(nleaves);
  int done = 0;
  bool first = true;
  while (first || count > 1) {
    int lv = total_levels - done < lg_block ? total_levels - done : lg_block;
    uint32_t blocks = (uint32_t)((count + width - 1) / width);
    const bool last = total_levels - done <= lg_block;  // this launch writes the root
    hipLaunchKernelGGL(merkle_reduce_kernel, dim3(blocks), dim3(4 * width), 0, st, M, rows, rs, cs,
                       first ? 1 : 0, nodes, off, count, lv, last ? root_out : nullptr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    for (int i = 0; i < lv; i++) {
      off += count;
      count >>= 1;
    }
    done += lv;
    first = false;
  }
  return hipSuccess;
}

}  // namespace lgs
