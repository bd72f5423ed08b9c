// Host-side launchers of the device kernels (all enqueue on the given stream, no sync).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "gl.hpp"

#define LGS_LDS_LOG 13  // transforms up to 2^13 elements (64 KiB) run wholly in LDS

namespace lgs {

// Job tables of the batched (grouped) launches: block / tile ids [first_block, next job's
// first_block) belong to a job; jobs are sorted by first_block.
struct NttJob {
  const uint64_t* in;
  uint64_t rs, es;
  uint32_t n_in;
  int logN;
  const uint64_t* tw;
  uint64_t* out;
  uint32_t first_block;
};
struct HashJob {
  const uint64_t* enc;
  uint64_t ld;
  const void* adig;
  uint32_t slabs_per_split, ntiles, ncols;
  int32_t* part;
  uint32_t first_block;
};
struct FinalizeJob {
  const int32_t* part;
  uint32_t ksplit, ncols;
  uint64_t* h;
  uint32_t first_block;
};
struct MerkleJob {
  const uint64_t* M;
  uint32_t rows;
  uint64_t rs, cs;
  int from_leaves;
  uint32_t* nodes;
  uint64_t child_off, child_count;
  int levels;
  uint32_t first_block;
};

// ntt.hip
hipError_t launch_twiddles(uint64_t* tw, int logN, hipStream_t st);
hipError_t launch_fill_field(uint64_t* p, uint64_t n, uint64_t seed, hipStream_t st);
// impl 0 / 2 / 3: radix-8 passes with the whole row in LDS when 2^logN <= 2^LGS_LDS_LOG, with 256 /
// 512 / 1024 threads per row; larger rows run radix-8 sub-blocks then global radix-2 passes.
// impl 1: radix-2 stages, the first LGS_LDS_LOG of them in LDS and the rest as global passes.
// impl 4 / 5: for 2^13-point rows of 2^12 values at unit stride, the compile-time-shaped radix-8 /
// radix-16 kernel with the pass twiddle table tp (launch_pass_twiddles for that impl); other
// shapes, or tp null, run impl 3.
// tp13 (optional): the pass table of a 2^13-point transform (pass_twiddle_entries(13, 4)); with
// it, rows of 2^16 points whose upper half is zero take their 2^13-point subsequence transforms
// on the compile-time-shaped kernel (at 2^19 points, 2^13-point subsequences measured slower
// than 2^12-point ones whatever the kernel).
hipError_t launch_ntt(const uint64_t* in, uint64_t rs, uint64_t es, uint32_t n_in, uint32_t rows,
                      int logN, const uint64_t* tw, uint64_t* out, hipStream_t st, int impl = 0,
                      const uint64_t* tp = nullptr, const uint64_t* tp13 = nullptr);
// Entries of the pass twiddle table of a 2^L-point transform for row-transform impl 4 (radix-8
// passes from s = 4, L = 1 mod 3) or 5 (radix-16 passes from s = 5, L = 1 mod 4), 0 for other
// impls and shapes, and its fill from the transform's twiddle table tw (entries in Montgomery
// form, c 2^64 mod p).
uint32_t pass_twiddle_entries(int L, int impl);
hipError_t launch_pass_twiddles(uint64_t* tp, const uint64_t* tw, int L, int impl, hipStream_t st);

// hash.hip (`mid`, when not null, is recorded between the main kernel and the finalize; k0 / k1,
// when not null, receive the main kernel's own start and stop times; h null: the partials only,
// no finalize).
// ring: the Z_2^64 subset-sum hash (sums mod 2^64 instead of mod p): A's digits represent A mod
// 2^64, and only the MFMA kernels that reduce in their epilogue (11, 12, 14-17) take it.
hipError_t launch_a_layouts(const uint64_t* a, uint64_t W, uint64_t* at, int8_t* dig, hipStream_t st,
                            bool ring = false);
hipError_t launch_hash_valu(const uint64_t* enc, uint64_t ld, uint32_t m_rows, uint32_t ncols,
                            const uint64_t* at, uint32_t ksplit, uint64_t* part, uint64_t* h,
                            hipStream_t st, hipEvent_t mid);
hipError_t launch_hash_mfma(const uint64_t* enc, uint64_t ld, uint32_t m_rows, uint32_t ncols,
                            const int8_t* dig, uint32_t ksplit, int32_t* part, uint64_t* h,
                            hipStream_t st, hipEvent_t mid, int impl, bool ring = false, hipEvent_t k0 = nullptr,
                            hipEvent_t k1 = nullptr);
// The main MFMA kernel for K-splits [split0, split0 + nsplits) only (partials of those splits),
// and the finalize over all ksplit partials: a commit may issue its splits in several launches.
hipError_t launch_hash_mfma_splits(const uint64_t* enc, uint64_t ld, uint32_t m_rows, uint32_t ncols,
                                  const int8_t* dig, uint32_t ksplit, uint32_t split0, uint32_t nsplits,
                                  int32_t* part, hipStream_t st, int impl, bool ring = false,
                                  hipEvent_t k0 = nullptr, hipEvent_t k1 = nullptr);
// Kernels 11 and 12 (and the batched hash) leave one word per (split, row, column), reduced mod p,
// in the first quarter of the same buffer; the others 8 digit-plane int32 sums.
bool hash_impl_reduces(int impl);
hipError_t launch_hash_mfma_finalize(const int32_t* part, uint32_t ksplit, uint32_t ncols, uint64_t* h,
                                     hipStream_t st, int impl, bool ring = false);

// Batched launches over job tables that live in device memory.
hipError_t launch_ntt_group(const NttJob* jobs, int njobs, uint32_t total_blocks, int max_logN, hipStream_t st);
// Jobs whose rows hold 2^12 values at unit stride (2^13-point transforms), with the compile-time
// shaped radix-8 kernel and its pass twiddle table tp (ntt_impl 4's).
hipError_t launch_ntt_fixed_group(const NttJob* jobs, int njobs, uint32_t total_blocks, const uint64_t* tp,
                                  hipStream_t st);
// pipe8: the pipelined 8-wave body (every job with 256-column tiles and a multiple of 8 slabs per
// K-split), else the 4-slot ring body.
hipError_t launch_hash_group(const HashJob* jobs, int njobs, uint32_t total_tiles, const FinalizeJob* fjobs,
                             uint32_t total_fblocks, hipStream_t st, int tile_cols = 128, bool pipe8 = false);
hipError_t launch_merkle_group(const MerkleJob* jobs, int njobs, uint32_t total_blocks, hipStream_t st);

// merkle.hip: tree over nleaves (a power of two) leaves, leaf p = M[r*rs + p*cs], r < rows.
// lg_block: log2 nodes per workgroup (6..8; 4 threads per node) = levels reduced per launch.
// root_out (optional, device-accessible, e.g. pinned host memory) receives the root's 8 words
// from the launch that computes it, and then, if root_seq is not zero, root_seq in root_out[8]
// (stored with system-scope release ordering after the root: the tree's last store).
hipError_t launch_merkle(const uint64_t* M, uint32_t rows, uint64_t rs, uint64_t cs,
                         uint64_t nleaves, uint32_t* nodes, hipStream_t st, int lg_block = 7,
                         uint32_t* root_out = nullptr, uint32_t root_seq = 0);
// A commit's finalize and column tree together: H = the sum of the ksplit partials that the
// reducing hash kernels leave (one word per split, row and column; mod 2^64 when ring), written
// to h, and the tree over H's ncols columns (a power of two), the root also to root_out.
hipError_t launch_commit_tree(const uint64_t* part, uint32_t ksplit, uint32_t ncols, bool ring, uint64_t* h,
                              uint32_t* nodes, hipStream_t st, int lg_block = 7, uint32_t* root_out = nullptr,
                              uint32_t root_seq = 0);

}  // namespace lgs
