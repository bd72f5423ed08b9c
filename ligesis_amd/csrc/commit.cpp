// Commit pipeline (pcs/src/ligesis.rs:277-325) and GPU Basefold commitments
// (pcs/src/basefold.rs:257-307, 341-382), enqueued on the context stream.
#include <cstring>

#include "runtime.hpp"

namespace lgs {

static int ilog2(uint64_t x) {
  int l = 0;
  while ((uint64_t(1) << l) < x) l++;
  return l;
}

const uint64_t* Context::twiddle(int logN) {
  auto it = twiddles.find(logN);
  if (it != twiddles.end()) return it->second.ptr;
  DevBuf<uint64_t> t(&pool, logN > 0 ? (size_t(1) << (logN - 1)) : 1);
  if (logN > 0) hip_check(launch_twiddles(t.ptr, logN, stream), "twiddles");
  const uint64_t* p = t.ptr;
  twiddles.emplace(logN, std::move(t));
  return p;
}

uint8_t* Context::staging(size_t bytes) {
  if (bytes > pinned_size) {
    if (pinned) (void)hipHostFree(pinned);
    pinned = nullptr;
    pinned_size = 0;
    hip_check(hipHostMalloc((void**)&pinned, bytes, hipHostMallocDefault), "hipHostMalloc");
    pinned_size = bytes;
  }
  return pinned;
}

void enqueue_basefold_commit(Context& ctx, DevBasefold& st, const uint64_t* poly, bool ext,
                             uint32_t num_vars, uint32_t li, uint32_t code_rate) {
  if (li > num_vars || li > 6) throw Error(1, "basefold: log_interleave out of range");
  st.ext = ext;
  st.num_vars = num_vars;
  st.li = li;
  st.code_rate = code_rate;
  uint64_t len_c = uint64_t(1) << (num_vars - li);
  st.D = len_c << code_rate;
  int logD = ilog2(st.D);
  uint32_t rows = st.code_rows();
  st.codes = DevBuf<uint64_t>(&ctx.pool, (size_t)rows * st.D);
  st.poly_ptr = poly;
  // chunk j of the poly (basefold.rs:346, MlPoly::split) is poly[j :: 2^li]; an extension
  // element is two words, so word row rho = 2 j + c walks with stride 2^(li+1).
  hip_check(launch_ntt(poly, 1, rows, (uint32_t)len_c, rows, logD, ctx.twiddle(logD), st.codes.ptr,
                       ctx.stream),
            "basefold encode");
  st.tree.leaves = st.D;
  st.tree.nodes = DevBuf<uint32_t>(&ctx.pool, (size_t)st.tree.node_count() * 8);
  hip_check(launch_merkle(st.codes.ptr, rows, st.D, 1, st.D, st.tree.nodes.ptr, ctx.stream),
            "basefold merkle");
}

void setup_key(Context& ctx, Key& key, const uint64_t* host_a, bool commit_a) {
  key.ctx = &ctx;
  key.W = uint64_t(64) << key.log_m;
  key.a = DevBuf<uint64_t>(&ctx.pool, 32 * key.W);
  hip_check(hipMemcpyAsync(key.a.ptr, host_a, 32 * key.W * 8, hipMemcpyHostToDevice, ctx.stream),
            "upload A");
  key.at = DevBuf<uint64_t>(&ctx.pool, 32 * key.W);
  key.dig = DevBuf<int8_t>(&ctx.pool, 256 * key.W);
  hip_check(launch_a_layouts(key.a.ptr, key.W, key.at.ptr, key.dig.ptr, ctx.stream), "A layouts");
  if (commit_a) {
    uint32_t nv = (uint32_t)ilog2(32 * key.W);
    uint32_t li = shared_li(nv, key.log_m, key.log_n);
    enqueue_basefold_commit(ctx, key.a_state, key.a.ptr, false, nv, li, 1);
    uint8_t* h = ctx.staging(32);
    hip_check(hipMemcpyAsync(h, key.a_state.tree.nodes.ptr + (key.a_state.tree.node_count() - 1) * 8,
                             32, hipMemcpyDeviceToHost, ctx.stream),
              "read com_a");
    hip_check(hipStreamSynchronize(ctx.stream), "setup sync");
    std::memcpy(key.com_a, h, 32);
    std::memcpy(key.a_state.root, h, 32);
    key.has_state = true;
  } else {
    hip_check(hipStreamSynchronize(ctx.stream), "setup sync");
  }
}

static uint32_t pow2_at_least(uint64_t x) {
  uint32_t p = 1;
  while (p < x) p <<= 1;
  return p;
}

void enqueue_commit(Context& ctx, const Key& key, Commitment& c) {
  const uint64_t m = uint64_t(1) << key.log_m, n = uint64_t(1) << key.log_n, two_n = 2 * n;
  const int logN = key.log_n + 1;
  c.log_m = key.log_m;
  c.log_n = key.log_n;
  c.enc = DevBuf<uint64_t>(&ctx.pool, m * two_n);
  c.h = DevBuf<uint64_t>(&ctx.pool, 32 * two_n);
  c.tree.leaves = two_n;
  c.tree.nodes = DevBuf<uint32_t>(&ctx.pool, (size_t)c.tree.node_count() * 8);
  if (ctx.profile) hip_check(hipEventRecord(ctx.ev[0], ctx.stream), "event");
  // F'[i] = RS(F[i]) (ligesis.rs:289-291)
  hip_check(launch_ntt(c.f, n, 1, (uint32_t)n, (uint32_t)m, logN, ctx.twiddle(logN), c.enc.ptr,
                       ctx.stream, ctx.ntt_impl),
            "row encode");
  if (ctx.profile) hip_check(hipEventRecord(ctx.ev[1], ctx.stream), "event");
  // H = A . B with B read bitwise from F' (ligesis.rs:292-302)
  bool mfma_ok = (two_n % 128 == 0) && m >= 2;
  int variant = ctx.hash_variant == 0 ? (mfma_ok ? 1 : 2) : ctx.hash_variant;
  if (variant == 1 && !mfma_ok) throw Error(1, "MFMA hash needs 2n % 128 == 0 and m >= 2");
  if (variant == 1) {
    uint64_t ntiles = two_n / 128;
    uint32_t ks = ctx.hash_ksplit ? (uint32_t)ctx.hash_ksplit : pow2_at_least((512 + ntiles - 1) / ntiles);
    if (ks > m / 2) ks = (uint32_t)(m / 2);
    if (ks == 0 || (m % (2 * ks)) != 0) throw Error(1, "hash_ksplit must divide m / 2");
    DevBuf<int32_t> part(&ctx.pool, (size_t)ks * 256 * two_n);
    hip_check(launch_hash_mfma(c.enc.ptr, two_n, (uint32_t)m, (uint32_t)two_n, key.dig.ptr, ks, part.ptr,
                               c.h.ptr, ctx.stream, ctx.profile ? ctx.ev[2] : nullptr, ctx.hash_impl),
              "hash (MFMA)");
  } else {
    uint64_t nb = (two_n + 255) / 256;
    uint32_t ks = ctx.hash_ksplit ? (uint32_t)ctx.hash_ksplit : pow2_at_least((1024 + nb - 1) / nb);
    if (ks > m) ks = (uint32_t)m;
    if (m % ks) throw Error(1, "hash_ksplit must divide m");
    DevBuf<uint64_t> part(&ctx.pool, (size_t)ks * 32 * two_n);
    hip_check(launch_hash_valu(c.enc.ptr, two_n, (uint32_t)m, (uint32_t)two_n, key.at.ptr, ks, part.ptr,
                               c.h.ptr, ctx.stream, ctx.profile ? ctx.ev[2] : nullptr),
              "hash (VALU)");
  }
  if (ctx.profile) hip_check(hipEventRecord(ctx.ev[3], ctx.stream), "event");
  // column-wise Merkle tree over serialize(H[:, j]) (ligesis.rs:306-313)
  hip_check(launch_merkle(c.h.ptr, 32, two_n, 1, two_n, c.tree.nodes.ptr, ctx.stream), "merkle");
  if (ctx.profile) hip_check(hipEventRecord(ctx.ev[4], ctx.stream), "event");
}

}  // namespace lgs
