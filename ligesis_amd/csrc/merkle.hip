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
#include "gl.hpp"
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

// The levels of one workgroup's subtree: sh holds the digests of local_n nodes (node t at
// sh[8 t .. 8 t + 7]) of a layer of `count` nodes stored at nodes[8 off ..], the first of them node
// blk_base of that layer. Reduces by up to `levels` levels, writing every parent to its layer, and
// root_out (if set) receives the root from the quad that computes it, followed by root_seq (if not
// zero) in root_out[8] once the root is visible to the host: a caller may wait for that word
// instead of for the stream.
__device__ __forceinline__ void merkle_block_levels(uint32_t* __restrict__ nodes, uint32_t* sh, uint64_t off,
                                                    uint64_t count, uint64_t blk_base, uint32_t local_n, int levels,
                                                    uint32_t* root_out, uint32_t root_seq = 0) {
  const uint32_t node = threadIdx.x >> 2, c = threadIdx.x & 3;
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
    // the root's stores, then (a system-scope release) the completion word: the last store of
    // the commit's last launch, the nodes of every level having been stored before the level's
    // barriers
    if (root_seq && c == 0) __hip_atomic_store(root_out + 8, root_seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

__device__ __forceinline__ void merkle_reduce_body(const uint64_t* __restrict__ M, uint32_t rows, uint64_t rs,
                                                   uint64_t cs, int from_leaves, uint32_t* __restrict__ nodes,
                                                   uint64_t child_off, uint64_t child_count, int levels,
                                                   uint32_t blk, uint32_t* sh, uint32_t* root_out = nullptr,
                                                   uint32_t root_seq = 0) {
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
  merkle_block_levels(nodes, sh, child_off, child_count, base, local_n, levels, root_out, root_seq);
}

__global__ __launch_bounds__(1024) void merkle_reduce_kernel(
    const uint64_t* __restrict__ M, uint32_t rows, uint64_t rs, uint64_t cs, int from_leaves,
    uint32_t* __restrict__ nodes, uint64_t child_off, uint64_t child_count, int levels, uint32_t* root_out,
    uint32_t root_seq) {
  __shared__ uint32_t sh[256 * 8];
  merkle_reduce_body(M, rows, rs, cs, from_leaves, nodes, child_off, child_count, levels, blockIdx.x, sh, root_out,
                     root_seq);
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

// Launches over the layer of `count` nodes at nodes[8 off ..] (or the leaves, from_leaves) until
// the root, with `done` of the tree's total_levels levels already reduced.
static hipError_t merkle_rounds(const uint64_t* M, uint32_t rows, uint64_t rs, uint64_t cs, uint32_t* nodes,
                                hipStream_t st, int lg_block, uint32_t* root_out, uint32_t root_seq, uint64_t off,
                                uint64_t count, int total_levels, int done, bool from_leaves) {
  const uint32_t width = 1u << lg_block;
  bool first = from_leaves;
  while (first || count > 1) {
    int lv = total_levels - done < lg_block ? total_levels - done : lg_block;
    uint32_t blocks = (uint32_t)((count + width - 1) / width);
    const bool last = total_levels - done <= lg_block;  // this launch writes the root
    hipLaunchKernelGGL(merkle_reduce_kernel, dim3(blocks), dim3(4 * width), 0, st, M, rows, rs, cs,
                       first ? 1 : 0, nodes, off, count, lv, last ? root_out : nullptr, root_seq);
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

hipError_t launch_merkle(const uint64_t* M, uint32_t rows, uint64_t rs, uint64_t cs,
                         uint64_t nleaves, uint32_t* nodes, hipStream_t st, int lg_block, uint32_t* root_out,
                         uint32_t root_seq) {
  if (lg_block < 6 || lg_block > 8) return hipErrorInvalidValue;
  return merkle_rounds(M, rows, rs, cs, nodes, st, lg_block, root_out, root_seq, 0, nleaves, ilog2_u64(nleaves), 0,
                       true);
}

// ---------------- K-split sum, leaf digests and subtrees in one launch ----------------
// The commit's H (32 rows of ncols words) is the sum of its K-split partials part[s][r][col]
// (mod p, or mod 2^64 for the Z_2^64 hash); its Merkle leaves are the columns of H. Workgroup b
// takes the 2^LGW columns from b 2^LGW: the quad of column p sums rows 8 k + 2 c and 8 k + 2 c + 1
// (k < 4) in lane c -- every partial read once, the loads of four splits in flight together --
// writes them to H and to the column's LDS copy, then hashes the column's four 64-byte blocks
// from LDS and reduces the workgroup's subtree. One launch instead of a sum kernel, whose H the
// leaf kernel read back, and a leaf kernel of a quarter as many workgroups.
constexpr int TREE_LGW = 5;
constexpr uint32_t TREE_COL_WORDS = 66;  // 64 words of a column + 2: columns start 2 banks apart

template <int LGW>
__global__ __launch_bounds__(4 << LGW) void commit_tree_kernel(const uint64_t* __restrict__ part, uint32_t ksplit,
                                                              uint32_t ncols, int ring, uint64_t* __restrict__ h,
                                                              uint32_t* __restrict__ nodes, uint32_t* root_out,
                                                              uint32_t root_seq) {
  constexpr uint32_t W = 1u << LGW;
  __shared__ uint32_t colw[W * TREE_COL_WORDS];
  __shared__ uint32_t sh[W * 8];
  const uint32_t q = threadIdx.x >> 2, c = threadIdx.x & 3;
  const uint64_t base = (uint64_t)blockIdx.x * W, p = base + q;
  const uint32_t local_n = (uint32_t)((ncols - base) < W ? (ncols - base) : W);
  if (q < local_n) {
    uint64_t acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t s0 = 0; s0 < ksplit; s0 += 4) {
      uint64_t v[4][8];
#pragma unroll
      for (int j = 0; j < 4; j++)
#pragma unroll
        for (int i = 0; i < 8; i++) {
          const uint32_t r = 8 * (i >> 1) + 2 * c + (i & 1);
          v[j][i] = s0 + j < ksplit ? part[((uint64_t)(s0 + j) * 32 + r) * ncols + p] : 0;
        }
#pragma unroll
      for (int j = 0; j < 4; j++)
#pragma unroll
        for (int i = 0; i < 8; i++) acc[i] = ring ? acc[i] + v[j][i] : gl_add(acc[i], v[j][i]);
    }
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const uint32_t r = 8 * (i >> 1) + 2 * c + (i & 1);
      h[(uint64_t)r * ncols + p] = acc[i];
      colw[q * TREE_COL_WORDS + 2 * r] = (uint32_t)acc[i];
      colw[q * TREE_COL_WORDS + 2 * r + 1] = (uint32_t)(acc[i] >> 32);
    }
  }
  __syncthreads();
  if (q < local_n) {  // leaf = blake3(serialize(column p)): 32 words = 4 blocks of 64 bytes
    uint32_t h0, h1;
    b3_init4(h0, h1, c);
#pragma unroll
    for (uint32_t b = 0; b < 4; b++) {
      uint32_t m[16];
#pragma unroll
      for (int i = 0; i < 16; i++) m[i] = colw[q * TREE_COL_WORDS + 16 * b + i];
      b3_compress4(h0, h1, m, 64, (b == 0 ? B3_CHUNK_START : 0) | (b == 3 ? (B3_CHUNK_END | B3_ROOT) : 0), c);
    }
    nodes[p * 8 + c] = h0;
    nodes[p * 8 + 4 + c] = h1;
    sh[q * 8 + c] = h0;
    sh[q * 8 + 4 + c] = h1;
  }
  merkle_block_levels(nodes, sh, 0, ncols, base, local_n, LGW, root_out, root_seq);
}

hipError_t launch_commit_tree(const uint64_t* part, uint32_t ksplit, uint32_t ncols, bool ring, uint64_t* h,
                              uint32_t* nodes, hipStream_t st, int lg_block, uint32_t* root_out,
                              uint32_t root_seq) {
  if (lg_block < 6 || lg_block > 8 || ncols == 0 || (ncols & (ncols - 1)) || ksplit == 0)
    return hipErrorInvalidValue;
  constexpr uint32_t W = 1u << TREE_LGW;
  const int total = ilog2_u64(ncols);
  const bool one = ncols <= W;  // the first launch reaches the root
  hipLaunchKernelGGL(commit_tree_kernel<TREE_LGW>, dim3((ncols + W - 1) / W), dim3(4 * W), 0, st, part, ksplit,
                     ncols, ring ? 1 : 0, h, nodes, one ? root_out : nullptr, root_seq);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || one) return e;
  uint64_t off = 0, count = ncols;
  for (int i = 0; i < TREE_LGW; i++) {
    off += count;
    count >>= 1;
  }
  // the rest in workgroups wide enough for all of it where it fits one (2^8 nodes)
  const int rest = total - TREE_LGW, lg_top = rest > lg_block ? (rest < 8 ? rest : 8) : lg_block;
  return merkle_rounds(nullptr, 0, 0, 0, nodes, st, lg_top, root_out, root_seq, off, count, total, TREE_LGW, false);
}

}  // namespace lgs
