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

const uint64_t* Context::pass_twiddle(int logN, int impl) {
  const int key = impl << 8 | logN;
  auto it = pass_twiddles.find(key);
  if (it != pass_twiddles.end()) return it->second.ptr;
  const uint32_t n = pass_twiddle_entries(logN, impl);
  if (n == 0) return nullptr;
  const uint64_t* tw = twiddle(logN);
  DevBuf<uint64_t> t(&pool, n);
  hip_check(launch_pass_twiddles(t.ptr, tw, logN, impl, stream), "pass twiddles");
  const uint64_t* p = t.ptr;
  pass_twiddles.emplace(key, std::move(t));
  return p;
}

hipStream_t Context::side_stream() {
  if (!stream2) hip_check(hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking), "side stream");
  return stream2;
}

void Context::drain() {
  for (hipStream_t s : {stream, stream2})
    if (s) (void)hipStreamSynchronize(s);
}

void Context::destroy_streams() {
  for (hipStream_t* s : {&stream, &stream2})
    if (*s) {
      (void)hipStreamDestroy(*s);
      *s = nullptr;
    }
}

hipEvent_t Context::pipe_event(int i) {
  if (!pipe_ev[i]) hip_check(hipEventCreateWithFlags(&pipe_ev[i], hipEventDisableTiming), "pipeline event");
  return pipe_ev[i];
}

hipEvent_t Context::timing_event(int i) {
  if (!tev[i]) hip_check(hipEventCreateWithFlags(&tev[i], hipEventDisableSystemFence), "timing event");
  return tev[i];
}

hipError_t Context::create_events() {
  for (int i = 0; i < 10; i++) {
    const hipError_t e = hipEventCreateWithFlags(&ev[i], i < 5 ? hipEventDisableSystemFence : hipEventDefault);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
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

void PinnedArena::reset() {
  size_t total = 0;
  for (auto& c : chunks_) total += c.second;
  if (chunks_.size() > 1) {  // merge into one chunk of the combined size
    release();
    uint8_t* p = nullptr;
    hip_check(hipHostMalloc((void**)&p, total, hipHostMallocDefault), "hipHostMalloc");
    chunks_.emplace_back(p, total);
  }
  used_ = 0;
}

void* PinnedArena::upload(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes == 0) return nullptr;
  size_t at = (used_ + 255) & ~size_t(255);
  if (chunks_.empty() || at + bytes > chunks_.back().second) {  // new chunk; earlier ones stay
    size_t sz = std::max<size_t>(bytes, size_t(1) << 20);
    if (!chunks_.empty()) sz = std::max(sz, 2 * chunks_.back().second);
    uint8_t* p = nullptr;
    hip_check(hipHostMalloc((void**)&p, sz, hipHostMallocDefault), "hipHostMalloc");
    chunks_.emplace_back(p, sz);
    at = 0;
  }
  uint8_t* h = chunks_.back().first + at;
  std::memcpy(h, src, bytes);
  used_ = at + bytes;
  hip_check(hipMemcpyAsync(dst, h, bytes, hipMemcpyHostToDevice, s), "upload");
  return h;
}

void PinnedArena::release() {
  for (auto& c : chunks_) (void)hipHostFree(c.first);
  chunks_.clear();
  used_ = 0;
}

std::unique_ptr<Context> make_context(int device, bool two_streams) {
  std::unique_ptr<Context> c(new Context());
  c->device = device;
  hip_check(hipSetDevice(device), "set device");
  hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "stream");
  if (two_streams) (void)c->side_stream();
  hip_check(c->create_events(), "event");
  return c;
}

void teardown_context(Context& c) {
  (void)hipSetDevice(c.device);
  for (auto& w : c.workers) teardown_context(*w);
  c.workers.clear();
  c.drain();
  c.twiddles.clear();
  c.pass_twiddles.clear();
  c.pool.trim();
  for (auto& e : c.ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : c.pipe_ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : c.tev)
    if (e) (void)hipEventDestroy(e);
  c.free_host();
  c.destroy_streams();
}

void Context::free_host() {
  if (pinned) (void)hipHostFree(pinned);
  pinned = nullptr;
  pinned_size = 0;
  arena.release();
  prove_arena.release();
}

void enqueue_basefold_commit(Context& ctx, DevBasefold& st, const uint64_t* poly, bool ext,
                             uint32_t num_vars, uint32_t li, uint32_t code_rate, hipStream_t s,
                             hipEvent_t encoded) {
  if (!s) s = ctx.stream;
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
  hip_check(launch_ntt(poly, 1, rows, (uint32_t)len_c, rows, logD, ctx.twiddle(logD), st.codes.ptr, s, 0, nullptr,
                       logD > LGS_LDS_LOG ? ctx.pass_twiddle(13, 4) : nullptr),
            "basefold encode");
  if (encoded) hip_check(hipEventRecord(encoded, s), "basefold encoded");
  st.tree.leaves = st.D;
  st.tree.nodes = DevBuf<uint32_t>(&ctx.pool, (size_t)st.tree.node_count() * 8);
  hip_check(launch_merkle(st.codes.ptr, rows, st.D, 1, st.D, st.tree.nodes.ptr, s), "basefold merkle");
}

void setup_key(Context& ctx, Key& key, const uint64_t* host_a, bool commit_a) {
  key.ctx = &ctx;
  key.W = uint64_t(64) << key.log_m;
  key.a = DevBuf<uint64_t>(&ctx.pool, 32 * key.W);
  hip_check(hipMemcpyAsync(key.a.ptr, host_a, 32 * key.W * 8, hipMemcpyHostToDevice, ctx.stream),
            "upload A");
  key.at = DevBuf<uint64_t>(&ctx.pool, 32 * key.W);
  key.dig = DevBuf<int8_t>(&ctx.pool, 256 * key.W);
  hip_check(launch_a_layouts(key.a.ptr, key.W, key.at.ptr, key.dig.ptr, ctx.stream, key.ring64), "A layouts");
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

Tuning builtin_tuning(uint32_t log_m, uint32_t log_n) {
  (void)log_m;
  Tuning t;
  // one thread per radix-8 butterfly of a pass measured fastest at m = 2^12 (1024 threads for
  // rows of 2^13 points, 512 for 2^12); at 2^13 points the compile-time-shaped kernel with
  // per-pass twiddle tables (185 us against 233 us for 2^12 rows)
  t.ntt_impl = log_n + 1 == 13 ? 4 : log_n + 1 > 13 ? 3 : log_n + 1 == 12 ? 2 : 0;
  return t;
}

Tuning Context::resolve(uint32_t log_m, uint32_t log_n) const {
  Tuning t = builtin_tuning(log_m, log_n);
  TuneEntry e;
  if (tune_lookup(log_m, log_n, e)) t = e.t;
  if (hash_variant) t.hash_variant = hash_variant;
  if (hash_ksplit) t.hash_ksplit = hash_ksplit;
  if (hash_impl >= 0) t.hash_impl = hash_impl;
  if (ntt_impl >= 0) t.ntt_impl = ntt_impl;
  if (merkle_lg) t.merkle_lg = merkle_lg;
  if (pipeline) t.pipeline = pipeline;
  return t;
}

void enqueue_commit(Context& ctx, const Key& key, Commitment& c, const Tuning* force) {
  const uint64_t m = uint64_t(1) << key.log_m, n = uint64_t(1) << key.log_n, two_n = 2 * n;
  const int logN = key.log_n + 1;
  const Tuning tu = force ? *force : ctx.resolve(key.log_m, key.log_n);
  c.log_m = key.log_m;
  c.log_n = key.log_n;
  c.enc = DevBuf<uint64_t>(&ctx.pool, m * two_n);
  c.h = DevBuf<uint64_t>(&ctx.pool, 32 * two_n);
  c.tree.leaves = two_n;
  c.tree.nodes = DevBuf<uint32_t>(&ctx.pool, (size_t)c.tree.node_count() * 8);
  const uint64_t* tw = ctx.twiddle(logN);
  const uint64_t* tp = ctx.pass_twiddle(logN, tu.ntt_impl);
  const HashPlan plan = hash_plan(key, tu);
  const int variant = plan.variant;
  const uint32_t ks = plan.ks;
  const uint32_t chunks = (uint32_t)tu.pipeline;
  if (variant == 1 && chunks > 1 && ks % chunks == 0 && chunks <= 8) {
    // Pipelined: row chunk q (the rows of K-splits [q ks/C, (q+1) ks/C)) is encoded on the
    // context stream, then hashed on the second stream while chunk q + 1 is encoded, so the
    // VALU-bound transform and the MFMA-bound hash share the device.
    hipStream_t sh = ctx.side_stream();
    if (ctx.profile) hip_check(hipEventRecord(ctx.ev[0], ctx.stream), "event");
    const uint64_t rows = m / chunks;
    const uint32_t kc = ks / chunks;
    DevBuf<int32_t> part(&ctx.pool, (size_t)ks * 256 * two_n);
    auto mark = [&](int i, hipStream_t s) {
      if (ctx.profile) hip_check(hipEventRecord(ctx.timing_event(i), s), "event");
    };
    for (uint32_t q = 0; q < chunks; q++) {
      mark(4 * q, ctx.stream);
      hip_check(launch_ntt(c.f + q * rows * n, n, 1, (uint32_t)n, (uint32_t)rows, logN, tw,
                           c.enc.ptr + q * rows * two_n, ctx.stream, tu.ntt_impl, tp),
                "row encode");
      mark(4 * q + 1, ctx.stream);
      hip_check(hipEventRecord(ctx.pipe_event(q), ctx.stream), "pipeline event");
      hip_check(hipStreamWaitEvent(sh, ctx.pipe_event(q), 0), "pipeline wait");
      mark(4 * q + 2, sh);
      hip_check(launch_hash_mfma_splits(c.enc.ptr, two_n, (uint32_t)m, (uint32_t)two_n, key.dig.ptr, ks, q * kc, kc,
                                        part.ptr, sh, plan.impl, key.ring64),
                "hash (MFMA)");
      mark(4 * q + 3, sh);
    }
    ctx.prof_chunks = ctx.profile ? (int)chunks : 0;  // pipelined: every event at any level
    if (ctx.profile) {
      hip_check(hipEventRecord(ctx.ev[1], ctx.stream), "event");  // every chunk encoded
      hip_check(hipEventRecord(ctx.ev[2], sh), "event");          // every split hashed
    }
    hip_check(launch_hash_mfma_finalize(part.ptr, ks, (uint32_t)two_n, c.h.ptr, sh, plan.impl, key.ring64),
              "hash finalize");
    if (ctx.profile) hip_check(hipEventRecord(ctx.ev[3], sh), "event");
    hip_check(hipEventRecord(ctx.pipe_event(9), sh), "pipeline join");
    hip_check(hipStreamWaitEvent(ctx.stream, ctx.pipe_event(9), 0), "pipeline join");
    hip_check(launch_merkle(c.h.ptr, 32, two_n, 1, two_n, c.tree.nodes.ptr, ctx.stream, tu.merkle_lg, c.root_host,
                            c.root_seq),
              "merkle");
    if (ctx.profile) hip_check(hipEventRecord(ctx.ev[4], ctx.stream), "event");
    return;
  }
  ctx.prof_chunks = 0;
  CommitIO io;
  io.f = c.f;
  io.enc = c.enc.ptr;
  io.h = c.h.ptr;
  io.nodes = c.tree.nodes.ptr;
  io.root_host = c.root_host;
  io.root_seq = c.root_seq;
  io.tw = tw;
  io.tp = tp;
  // freed when this returns: later allocations on the context reuse it in stream order
  DevBuf<uint64_t> part(&ctx.pool, (hash_partial_bytes(plan, two_n) + 7) / 8);
  io.part = part.ptr;
  enqueue_commit_stages(ctx, key, io, plan, tu, ctx.stream, ctx.profile);
}

HashPlan hash_plan(const Key& key, const Tuning& tu) {
  const uint64_t m = uint64_t(1) << key.log_m, two_n = uint64_t(2) << key.log_n;
  HashPlan hp;
  const bool mfma_ok = (two_n % 128 == 0) && m >= 2;
  hp.variant = tu.hash_variant == 0 ? (mfma_ok ? 1 : 2) : tu.hash_variant;
  hp.impl = tu.hash_impl;
  if (key.ring64) {  // the Z_2^64 hash runs on the reducing LDS-copy kernels only
    hp.variant = 1;
    if (!hash_impl_reduces(hp.impl)) hp.impl = Tuning{}.hash_impl;
  }
  if (hp.variant == 1 && !mfma_ok) throw Error(1, "MFMA hash needs 2n % 128 == 0 and m >= 2");
  if (hp.variant == 1) {
    const uint64_t ntiles = two_n / 128;
    hp.ks = tu.hash_ksplit ? (uint32_t)tu.hash_ksplit : pow2_at_least((512 + ntiles - 1) / ntiles);
    if (hp.ks > m / 2) hp.ks = (uint32_t)(m / 2);
    if (hp.ks == 0 || (m % (2 * hp.ks)) != 0) throw Error(1, "hash_ksplit must divide m / 2");
  } else {
    const uint64_t nb = (two_n + 255) / 256;
    hp.ks = tu.hash_ksplit ? (uint32_t)tu.hash_ksplit : pow2_at_least((1024 + nb - 1) / nb);
    if (hp.ks > m) hp.ks = (uint32_t)m;
    if (hp.ks == 0 || m % hp.ks) throw Error(1, "hash_ksplit must divide m");
  }
  return hp;
}

uint64_t hash_partial_bytes(const HashPlan& hp, uint64_t two_n) {
  // MFMA kernels that reduce in their epilogue and the VALU kernel leave one word per
  // (split, row, column); the others eight int32 digit-plane sums per (split, row, column)
  if (hp.variant == 1 && !hash_impl_reduces(hp.impl)) return (uint64_t)hp.ks * 256 * two_n * 4;
  return (uint64_t)hp.ks * 32 * two_n * 8;
}

void enqueue_commit_stages(Context& ctx, const Key& key, const CommitIO& io, const HashPlan& hp, const Tuning& tu,
                           hipStream_t st, int profile) {
  const uint64_t m = uint64_t(1) << key.log_m, n = uint64_t(1) << key.log_n, two_n = 2 * n;
  const int logN = key.log_n + 1;
  // level 1: events between all stages; level 2: the MFMA hash kernel's own start and stop times
  // (hipExtLaunchKernel), so nothing is added to the stream around it
  const bool every = profile == 1;
  const bool kernel_only = profile == 2 && hp.variant == 1;
  if (every) hip_check(hipEventRecord(ctx.ev[0], st), "event");
  // F'[i] = RS(F[i]) (ligesis.rs:289-291)
  hip_check(launch_ntt(io.f, n, 1, (uint32_t)n, (uint32_t)m, logN, io.tw, io.enc, st, tu.ntt_impl, io.tp),
            "row encode");
  if (profile && !kernel_only) hip_check(hipEventRecord(ctx.ev[1], st), "event");
  // H = A . B with B read bitwise from F' (ligesis.rs:292-302). Kernels that leave one reduced
  // word per (split, row, column) get the K-split sum fused into the tree's first launch (the
  // finalize stage then times nothing); the others sum in a finalize kernel of their own.
  const bool fused = hp.variant == 2 || hash_impl_reduces(hp.impl);
  uint64_t* h_fin = fused ? nullptr : io.h;
  if (hp.variant == 1) {
    hip_check(launch_hash_mfma(io.enc, two_n, (uint32_t)m, (uint32_t)two_n, key.dig.ptr, hp.ks, (int32_t*)io.part,
                               h_fin, st, every ? ctx.ev[2] : nullptr, hp.impl, key.ring64,
                               kernel_only ? ctx.ev[1] : nullptr, kernel_only ? ctx.ev[2] : nullptr),
              "hash (MFMA)");
  } else {
    hip_check(launch_hash_valu(io.enc, two_n, (uint32_t)m, (uint32_t)two_n, key.at.ptr, hp.ks, io.part, h_fin, st,
                               profile ? ctx.ev[2] : nullptr),
              "hash (VALU)");
  }
  if (every) hip_check(hipEventRecord(ctx.ev[3], st), "event");
  // column-wise Merkle tree over serialize(H[:, j]) (ligesis.rs:306-313)
  if (fused)
    hip_check(launch_commit_tree(io.part, hp.ks, (uint32_t)two_n, key.ring64, io.h, io.nodes, st, tu.merkle_lg,
                                 io.root_host, io.root_seq),
              "hash sum and merkle");
  else
    hip_check(launch_merkle(io.h, 32, two_n, 1, two_n, io.nodes, st, tu.merkle_lg, io.root_host, io.root_seq),
              "merkle");
  if (every) hip_check(hipEventRecord(ctx.ev[4], st), "event");
}

void enqueue_merkle_trees(Context& ctx, const std::vector<MerkleSource>& trees,
                          std::vector<std::vector<MerkleJob>>& host_jobs, std::vector<DevBuf<uint64_t>>& tables,
                          hipStream_t st, PinnedArena* arena) {
  // rounds of up to 8 levels; every tree takes part until its root is written
  const size_t nt = trees.size();
  std::vector<uint64_t> off(nt, 0), count(nt);
  std::vector<int> left(nt);
  for (size_t i = 0; i < nt; i++) {
    count[i] = trees[i].leaves;
    int l = 0;
    while ((uint64_t(1) << l) < count[i]) l++;
    left[i] = l;
  }
  std::vector<std::vector<MerkleJob>> rounds;
  std::vector<uint32_t> blocks;
  for (bool first = true;; first = false) {
    std::vector<MerkleJob> mj;
    uint32_t mblocks = 0;
    for (size_t i = 0; i < nt; i++) {
      if (!first && left[i] == 0) continue;
      const int lv = left[i] < 8 ? left[i] : 8;
      const MerkleSource& t = trees[i];
      mj.push_back(MerkleJob{t.M, t.rows, t.rs, t.cs, first ? 1 : 0, t.nodes, off[i], count[i], lv, mblocks});
      mblocks += (uint32_t)((count[i] + 255) / 256);
      for (int q = 0; q < lv; q++) {
        off[i] += count[i];
        count[i] >>= 1;
      }
      left[i] -= lv;
    }
    if (mj.empty()) break;
    rounds.push_back(std::move(mj));
    blocks.push_back(mblocks);
  }
  // one job table for all rounds, one upload
  std::vector<MerkleJob> all;
  std::vector<size_t> at;
  for (auto& r : rounds) {
    at.push_back(all.size());
    all.insert(all.end(), r.begin(), r.end());
  }
  if (all.empty()) return;
  host_jobs.push_back(std::move(all));
  const std::vector<MerkleJob>& h = host_jobs.back();
  tables.emplace_back(&ctx.pool, (h.size() * sizeof(MerkleJob) + 7) / 8);
  const MerkleJob* d = (const MerkleJob*)tables.back().ptr;
  (arena ? arena : &ctx.arena)->upload((void*)d, h.data(), h.size() * sizeof(MerkleJob), st);
  for (size_t r = 0; r < rounds.size(); r++)
    hip_check(launch_merkle_group(d + at[r], (int)rounds[r].size(), blocks[r], st), "grouped merkle");
}

// Batched commit: every stage of every eligible commitment runs in one launch per stage (row
// transforms, MFMA hash tiles, finalize, and one launch per Merkle reduction round), driven by
// job tables in device memory. Results are identical to enqueue_commit per polynomial.
bool group_eligible(const Key& key) {
  const uint64_t m = uint64_t(1) << key.log_m, two_n = uint64_t(2) << key.log_n;
  return !key.ring64 && two_n % 128 == 0 && m >= 2 && key.log_n + 1 <= LGS_LDS_LOG;
}

void enqueue_commit_group(Context& ctx, const std::vector<std::pair<const Key*, Commitment*>>& items,
                          GroupHold& hold) {
  if (items.empty()) return;
  ctx.upload_reset();  // every earlier call on this context has been synchronized
  // rows of 2^12 values (2^13-point transforms) go to the compile-time-shaped kernel, the rest
  // to the runtime-shaped one, sized by the largest transform among them
  std::vector<NttJob> nj, nj13;
  std::vector<HashJob> hj;
  std::vector<FinalizeJob> fj;
  uint32_t nblocks = 0, nblocks13 = 0, ntiles_total = 0, fblocks = 0;
  int max_logN = 1;
  bool pipe8 = true;  // every job fits the pipelined 8-wave hash body
  // 256-column tiles (8 waves sharing the A ring, one workgroup per CU) when every job has a
  // multiple of 256 columns, else 128
  uint32_t tile_cols = 256;
  for (auto& it : items)
    if ((uint64_t(2) << it.first->log_n) % 256) tile_cols = 128;
  for (auto& it : items) {
    const Key& key = *it.first;
    Commitment& c = *it.second;
    const uint64_t m = uint64_t(1) << key.log_m, n = uint64_t(1) << key.log_n, two_n = 2 * n;
    const int logN = key.log_n + 1;
    c.log_m = key.log_m;
    c.log_n = key.log_n;
    c.enc = DevBuf<uint64_t>(&ctx.pool, m * two_n);
    c.h = DevBuf<uint64_t>(&ctx.pool, 32 * two_n);
    c.tree.leaves = two_n;
    c.tree.nodes = DevBuf<uint32_t>(&ctx.pool, (size_t)c.tree.node_count() * 8);
    if (logN == 13) {
      nj13.push_back(NttJob{c.f, n, 1, (uint32_t)n, logN, ctx.twiddle(logN), c.enc.ptr, nblocks13});
      nblocks13 += (uint32_t)m;
    } else {
      nj.push_back(NttJob{c.f, n, 1, (uint32_t)n, logN, ctx.twiddle(logN), c.enc.ptr, nblocks});
      nblocks += (uint32_t)m;
      max_logN = logN > max_logN ? logN : max_logN;
    }
    const uint32_t ntiles = (uint32_t)(two_n / tile_cols);
    const Tuning tu = ctx.resolve(key.log_m, key.log_n);
    uint32_t ks = tu.hash_ksplit ? (uint32_t)tu.hash_ksplit : (uint32_t)(m >= 1024 ? m / 512 : 1);
    if (ks > m / 2) ks = (uint32_t)(m / 2);
    if (ks == 0 || m % (2 * ks)) throw Error(1, "hash_ksplit must divide m / 2");
    hold.parts.emplace_back(&ctx.pool, (size_t)ks * 256 * two_n);
    hj.push_back(HashJob{c.enc.ptr, two_n, key.dig.ptr, (uint32_t)(m / ks), ntiles, (uint32_t)two_n,
                         hold.parts.back().ptr, ntiles_total});
    if ((m / ks) % 8) pipe8 = false;
    ntiles_total += ntiles * ks;
    fj.push_back(FinalizeJob{hold.parts.back().ptr, ks, (uint32_t)two_n, c.h.ptr, fblocks});
    fblocks += (uint32_t)((32 * two_n + 255) / 256);
  }
  const int njobs = (int)items.size();
  auto upload = [&](const void* src, size_t bytes) {
    hold.tables.emplace_back(&ctx.pool, (bytes + 7) / 8);
    ctx.upload(hold.tables.back().ptr, src, bytes, ctx.stream);
    return (const void*)hold.tables.back().ptr;
  };
  const NttJob* d_nj = nj.empty() ? nullptr : (const NttJob*)upload(nj.data(), nj.size() * sizeof(NttJob));
  const NttJob* d_nj13 = nj13.empty() ? nullptr : (const NttJob*)upload(nj13.data(), nj13.size() * sizeof(NttJob));
  const uint64_t* tp13 = nj13.empty() ? nullptr : ctx.pass_twiddle(13, 4);
  const HashJob* d_hj = (const HashJob*)upload(hj.data(), hj.size() * sizeof(HashJob));
  const FinalizeJob* d_fj = (const FinalizeJob*)upload(fj.data(), fj.size() * sizeof(FinalizeJob));
  if (ctx.profile) hip_check(hipEventRecord(ctx.ev[0], ctx.stream), "event");
  if (!nj13.empty())
    hip_check(launch_ntt_fixed_group(d_nj13, (int)nj13.size(), nblocks13, tp13, ctx.stream), "batched row encode");
  if (!nj.empty()) hip_check(launch_ntt_group(d_nj, (int)nj.size(), nblocks, max_logN, ctx.stream), "batched row encode");
  if (ctx.profile) hip_check(hipEventRecord(ctx.ev[1], ctx.stream), "event");
  hip_check(launch_hash_group(d_hj, njobs, ntiles_total, nullptr, 0, ctx.stream, (int)tile_cols, tile_cols == 256 && pipe8),
            "batched hash");
  if (ctx.profile) hip_check(hipEventRecord(ctx.ev[2], ctx.stream), "event");
  hip_check(launch_hash_group(nullptr, njobs, 0, d_fj, fblocks, ctx.stream), "batched finalize");
  if (ctx.profile) hip_check(hipEventRecord(ctx.ev[3], ctx.stream), "event");
  std::vector<MerkleSource> trees;
  for (auto& it : items) {
    Commitment& c = *it.second;
    trees.push_back(MerkleSource{c.h.ptr, 32, c.tree.leaves, 1, c.tree.leaves, c.tree.nodes.ptr});
  }
  enqueue_merkle_trees(ctx, trees, hold.host_merkle, hold.tables, ctx.stream);
  if (ctx.profile) hip_check(hipEventRecord(ctx.ev[4], ctx.stream), "event");
  nj.insert(nj.end(), nj13.begin(), nj13.end());
  hold.host_ntt = std::move(nj);
  hold.host_hash = std::move(hj);
  hold.host_fin = std::move(fj);
}

}  // namespace lgs
