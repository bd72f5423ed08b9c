// extern "C" boundary of libligesis_hip.so (declared in include/ligesis_hip.h).
#include "../../include/ligesis_hip.h"

#include <cstring>
#include <exception>
#include <string>
#include <thread>
#include <vector>

#include "handles.hpp"

namespace {

void ctx_retain(lgs::Context* c) { c->refs.fetch_add(1); }

// Drops one reference; the last one tears the context down.
void ctx_unref(lgs::Context* c) {
  if (!c || c->refs.fetch_sub(1) != 1) return;
  lgs::teardown_context(*c);
  delete (lgs_ctx*)c->owner;
}

template <class F>
int guarded(lgs::Context* ctx, F&& f) {
  try {
    f();
    return LGS_OK;
  } catch (const lgs::Error& e) {
    if (ctx) ctx->err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    if (ctx) ctx->err = "host allocation failed";
    return LGS_ERR_NOMEM;
  } catch (const std::exception& e) {
    if (ctx) ctx->err = e.what();
    return LGS_ERR_ARG;
  }
}

void finish_profile(lgs::Context& ctx) {
  if (!ctx.profile) return;
  if (ctx.profile == 2 && ctx.prof_chunks == 0) {  // the hash kernel only
    float ms = 0;
    lgs::hip_check(hipEventElapsedTime(&ms, ctx.ev[1], ctx.ev[2]), "event time");
    for (int i = 0; i < 5; i++) ctx.stage_ms[i] = i == 1 ? ms : 0;
    return;
  }
  for (int i = 0; i < 4; i++) {
    float ms = 0;
    lgs::hip_check(hipEventElapsedTime(&ms, ctx.ev[i], ctx.ev[i + 1]), "event time");
    ctx.stage_ms[i] = ms;
  }
  if (ctx.prof_chunks > 0) {  // pipelined: sum the launches of each kind (they overlap in time)
    double ntt = 0, hash = 0;
    for (int q = 0; q < ctx.prof_chunks; q++) {
      float a = 0, b = 0;
      lgs::hip_check(hipEventElapsedTime(&a, ctx.tev[4 * q], ctx.tev[4 * q + 1]), "event time");
      lgs::hip_check(hipEventElapsedTime(&b, ctx.tev[4 * q + 2], ctx.tev[4 * q + 3]), "event time");
      ntt += a;
      hash += b;
    }
    ctx.stage_ms[0] = ntt;
    ctx.stage_ms[1] = hash;
  }
  float all = 0;
  lgs::hip_check(hipEventElapsedTime(&all, ctx.ev[0], ctx.ev[4]), "event time");
  ctx.stage_ms[4] = all;
}

uint32_t next_root_seq(lgs::Context& ctx) {
  ctx.root_seq = ctx.root_seq == 0xffffffffu ? 1 : ctx.root_seq + 1;  // never 0 (no completion word)
  return ctx.root_seq;
}

// Waits until the commit's tree has stored `seq` into the completion word (after the root, with
// system-scope release ordering: the commit's last memory access) -- a few microseconds sooner
// than a stream synchronization, which waits for the kernel's completion signal. The stream is
// polled now and then so that a failed launch raises its error instead of waiting forever.
void wait_root(lgs::Context& ctx, const uint32_t* word, uint32_t seq) {
  for (uint32_t i = 1;; i++) {
    if (__atomic_load_n(word, __ATOMIC_ACQUIRE) == seq) return;
    if ((i & 4095) == 0) {
      const hipError_t e = hipStreamQuery(ctx.stream);
      if (e == hipSuccess) {
        if (__atomic_load_n(word, __ATOMIC_ACQUIRE) == seq) return;
        throw lgs::Error(LGS_ERR_DEVICE, "commit: the stream finished without the root's completion word");
      }
      if (e != hipErrorNotReady) lgs::hip_check(e, "commit");
    }
  }
}

void read_root(lgs::Context& ctx, lgs::Commitment* const* cs, uint32_t count) {
  uint8_t* h = ctx.staging(32 * (size_t)count);
  for (uint32_t i = 0; i < count; i++) {
    const lgs::DevTree& t = cs[i]->tree;
    lgs::hip_check(hipMemcpyAsync(h + 32 * (size_t)i, t.nodes.ptr + (t.node_count() - 1) * 8, 32,
                                  hipMemcpyDeviceToHost, ctx.stream),
                   "read root");
  }
  lgs::hip_check(hipStreamSynchronize(ctx.stream), "commit sync");
  for (uint32_t i = 0; i < count; i++) std::memcpy(cs[i]->root, h + 32 * (size_t)i, 32);
}

void prepare_commitment(lgs::Context& ctx, const lgs::Key& key, const uint64_t* f, uint64_t len,
                        uint32_t flags, lgs::Commitment& c) {
  uint64_t need = uint64_t(1) << (key.log_m + key.log_n);
  if (!f) throw lgs::Error(LGS_ERR_ARG, "commit: null polynomial");
  if (len != need)
    throw lgs::Error(LGS_ERR_ARG, "commit: expected 2^(log_m+log_n) = " + std::to_string(need) +
                                      " evaluations, got " + std::to_string(len));
  c.ctx = &ctx;
  c.key = &key;
  if (flags & LGS_INPUT_DEVICE) {
    c.f = f;
  } else {
    c.f_owned = lgs::DevBuf<uint64_t>(&ctx.pool, len);
    lgs::hip_check(hipMemcpyAsync(c.f_owned.ptr, f, len * 8, hipMemcpyHostToDevice, ctx.stream),
                   "upload f");
    c.f = c.f_owned.ptr;
  }
}

}  // namespace

extern "C" {

int lgs_init(int device, lgs_ctx** out) {
  if (!out) return LGS_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return LGS_ERR_DEVICE;
  if (hipSetDevice(device) != hipSuccess) return LGS_ERR_DEVICE;
  try {
    lgs::scratch_init(device);  // device address of the no-allocation entry point's scratch
  } catch (const lgs::Error&) {
    return LGS_ERR_DEVICE;
  }
  auto* ctx = new lgs_ctx();
  ctx->c.owner = ctx;
  ctx->c.device = device;
  // the second stream right after the first, so the two get different hardware queues (see
  // make_context)
  if (hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->c.stream2, hipStreamNonBlocking) != hipSuccess) {
    lgs::teardown_context(ctx->c);
    delete ctx;
    return LGS_ERR_DEVICE;
  }
  if (ctx->c.create_events() != hipSuccess) {
    lgs::teardown_context(ctx->c);
    delete ctx;
    return LGS_ERR_DEVICE;
  }
  *out = ctx;
  return LGS_OK;
}

void lgs_shutdown(lgs_ctx* ctx) {
  if (!ctx) return;
  ctx_unref(&ctx->c);
}

const char* lgs_last_error(const lgs_ctx* ctx) { return ctx ? ctx->c.err.c_str() : "null context"; }

int lgs_set_option(lgs_ctx* ctx, const char* name, int64_t value) {
  if (!ctx || !name) return LGS_ERR_ARG;
  std::string n(name);
  if (n == "hash" && value >= 0 && value <= 2) {
    ctx->c.hash_variant = (int)value;
  } else if (n == "hash_ksplit" && value >= 0) {
    ctx->c.hash_ksplit = (int)value;
  } else if (n == "hash_impl" && ((value >= -1 && value <= 12) || (value >= 14 && value <= 17))) {
    ctx->c.hash_impl = (int)value;
  } else if (n == "ntt_impl" && value >= -1 && value <= 5) {
    ctx->c.ntt_impl = (int)value;
  } else if (n == "merkle_lg" && (value == 0 || (value >= 6 && value <= 8))) {
    ctx->c.merkle_lg = (int)value;
  } else if (n == "pipeline" && (value == 0 || value == 1 || value == 2 || value == 4 || value == 8)) {
    ctx->c.pipeline = (int)value;
  } else if (n == "prove_graph" && (value == 0 || value == 1)) {
    ctx->c.prove_graph = value != 0;
  } else if (n == "profile" && value >= 0 && value <= 2) {
    ctx->c.profile = (int)value;
  } else {
    ctx->c.err = "unknown option or value: " + n;
    return LGS_ERR_ARG;
  }
  return LGS_OK;
}

static void put_tuning(const lgs::Tuning& t, int32_t* cfg) {
  if (!cfg) return;
  cfg[0] = t.hash_variant;
  cfg[1] = t.hash_impl;
  cfg[2] = t.hash_ksplit;
  cfg[3] = t.ntt_impl;
  cfg[4] = t.merkle_lg;
  cfg[5] = t.pipeline;
}

int lgs_autotune(lgs_ctx* ctx, const lgs_key* key, uint32_t reps, int32_t cfg[LGS_TUNE_FIELDS], double* ms) {
  if (!ctx || !key) return LGS_ERR_ARG;
  return guarded(&ctx->c, [&] {
    (void)hipSetDevice(ctx->c.device);
    lgs::TuneEntry best;
    try {
      lgs::autotune(ctx->c, key->k, reps, best);
    } catch (...) {
      ctx->c.drain();
      throw;
    }
    put_tuning(best.t, cfg);
    if (ms) *ms = best.ms;
  });
}

int lgs_tune_get(lgs_ctx* ctx, uint32_t log_m, uint32_t log_n, int32_t cfg[LGS_TUNE_FIELDS]) {
  if (!ctx || !cfg) return LGS_ERR_ARG;
  put_tuning(ctx->c.resolve(log_m, log_n), cfg);
  return LGS_OK;
}

int lgs_tune_save(lgs_ctx* ctx, const char* path, uint64_t* count) {
  if (!path) return LGS_ERR_ARG;
  return guarded(ctx ? &ctx->c : nullptr, [&] {
    const size_t k = lgs::tune_save(path);
    if (count) *count = k;
  });
}

int lgs_tune_load(lgs_ctx* ctx, const char* path, uint64_t* count) {
  if (!path) return LGS_ERR_ARG;
  return guarded(ctx ? &ctx->c : nullptr, [&] {
    const size_t k = lgs::tune_load(path);
    if (count) *count = k;
  });
}

void lgs_tune_clear(void) { lgs::tune_clear(); }

int lgs_stage_times(const lgs_ctx* ctx, double* out_ms, uint32_t n) {
  if (!ctx || !out_ms) return LGS_ERR_ARG;
  for (uint32_t i = 0; i < n && i < 28; i++) out_ms[i] = ctx->c.stage_ms[i];
  return LGS_OK;
}

int lgs_setup(lgs_ctx* ctx, uint32_t log_m, uint32_t log_n, const uint64_t* a, uint32_t flags,
              lgs_key** out) {
  if (!ctx || !out || !a) return LGS_ERR_ARG;
  *out = nullptr;
  return guarded(&ctx->c, [&] {
    const bool ring64 = (flags & LGS_SETUP_RING64) != 0;
    const bool commit_only = ring64 || (flags & LGS_SETUP_COMMIT_ONLY) != 0;
    if (log_m > 20 || log_n > 26 || log_n < 1)
      throw lgs::Error(LGS_ERR_ARG, "setup: shape out of supported range");
    if (ring64 && (log_m < 1 || log_n < 6))
      throw lgs::Error(LGS_ERR_SHAPE, "setup: the Z_2^64 hash runs on the MFMA kernels (m >= 2, 2n >= 128)");
    if (!ring64) {  // A is C x 64 m field elements (subset_sum.rs:31-39): canonical words
      const uint64_t words = uint64_t(32) * (uint64_t(64) << log_m);
      for (uint64_t i = 0; i < words; i++)
        if (a[i] >= lgs::GL_P) throw lgs::Error(LGS_ERR_ARG, "setup: A holds a word >= p (not a field element)");
    }
    if (!commit_only && !lgs::shape_ok(log_m, log_n))
      throw lgs::Error(LGS_ERR_SHAPE,
                       "batched LigeSIS needs the {A-bar, B_I, a-bar} var counts within "
                       "LOG_INTERLEAVE of each other");
    (void)hipSetDevice(ctx->c.device);
    auto* k = new lgs_key();
    k->k.log_m = log_m;
    k->k.log_n = log_n;
    k->k.ring64 = ring64;
    try {
      lgs::setup_key(ctx->c, k->k, a, !commit_only);
    } catch (...) {
      ctx->c.drain();
      delete k;
      throw;
    }
    ctx_retain(&ctx->c);
    *out = k;
  });
}

int lgs_key_com_a(const lgs_key* key, uint8_t com_a[32]) {
  if (!key || !com_a) return LGS_ERR_ARG;
  if (!key->k.has_state) return LGS_ERR_STATE;
  std::memcpy(com_a, key->k.com_a, 32);
  return LGS_OK;
}

// Commits on the context stream into `st` (uploading a host polynomial into st.poly); waits and
// writes the root.
static void basefold_commit_into(lgs::Context& c, const uint64_t* poly, uint32_t num_vars, uint32_t flags,
                                 uint32_t log_interleave, uint32_t code_rate, lgs::DevBasefold& st, uint8_t root[32]) {
  if (num_vars > 30 || log_interleave > 6 || log_interleave > num_vars || code_rate < 1 || code_rate > 4)
    throw lgs::Error(LGS_ERR_ARG, "basefold commit: num_vars <= 30, log_interleave <= min(6, num_vars), "
                                  "1 <= code_rate <= 4");
  (void)hipSetDevice(c.device);
  const bool ext = (flags & LGS_BASEFOLD_EXT) != 0;
  const uint64_t words = (uint64_t(1) << num_vars) * (ext ? 2 : 1);
  const uint64_t* p = poly;
  if (!(flags & LGS_INPUT_DEVICE)) {
    st.poly = lgs::DevBuf<uint64_t>(&c.pool, words);
    lgs::hip_check(hipMemcpyAsync(st.poly.ptr, poly, words * 8, hipMemcpyHostToDevice, c.stream), "upload poly");
    p = st.poly.ptr;
  }
  lgs::enqueue_basefold_commit(c, st, p, ext, num_vars, log_interleave, code_rate);
  uint8_t* h = c.staging(32);
  lgs::hip_check(hipMemcpyAsync(h, st.tree.nodes.ptr + (st.tree.node_count() - 1) * 8, 32, hipMemcpyDeviceToHost,
                                c.stream),
                 "read basefold root");
  lgs::hip_check(hipStreamSynchronize(c.stream), "basefold sync");
  std::memcpy(root, h, 32);
  std::memcpy(st.root, h, 32);
}

int lgs_basefold_commit(lgs_ctx* ctx, const uint64_t* poly, uint32_t num_vars, uint32_t flags,
                        uint32_t log_interleave, uint32_t code_rate, uint8_t root[32]) {
  if (!ctx || !poly || !root) return LGS_ERR_ARG;
  return guarded(&ctx->c, [&] {
    lgs::DevBasefold st;
    try {
      basefold_commit_into(ctx->c, poly, num_vars, flags, log_interleave, code_rate, st, root);
    } catch (...) {
      ctx->c.drain();
      throw;
    }
  });
}

int lgs_basefold_commit_state(lgs_ctx* ctx, const uint64_t* poly, uint32_t num_vars, uint32_t flags,
                              uint32_t log_interleave, uint32_t code_rate, uint8_t root[32], lgs_basefold** out) {
  if (!ctx || !poly || !root || !out) return LGS_ERR_ARG;
  *out = nullptr;
  return guarded(&ctx->c, [&] {
    auto* b = new lgs_basefold();
    b->ctx = &ctx->c;
    try {
      basefold_commit_into(ctx->c, poly, num_vars, flags, log_interleave, code_rate, b->st, root);
    } catch (...) {
      ctx->c.drain();
      delete b;
      throw;
    }
    ctx_retain(&ctx->c);
    *out = b;
  });
}

int lgs_basefold_prove(lgs_ctx* ctx, const lgs_basefold* st, const uint64_t* point, uint64_t point_len,
                       const uint64_t* fields, uint64_t n_fields, const uint64_t* ints, uint64_t n_ints,
                       uint8_t* proof, uint64_t* proof_len, uint64_t used[2]) {
  if (!ctx || !st || !point || !proof_len || (n_fields && !fields) || (n_ints && !ints)) return LGS_ERR_ARG;
  return guarded(&ctx->c, [&] {
    if (st->ctx->device != ctx->c.device)
      throw lgs::Error(LGS_ERR_ARG, "basefold prove: the commitment lives on another device");
    (void)hipSetDevice(ctx->c.device);
    if (st->ctx != &ctx->c) lgs::hip_check(hipStreamSynchronize(st->ctx->stream), "basefold commit sync");
    const uint64_t cap = proof ? *proof_len : 0;
    uint64_t u[2] = {0, 0}, size = 0;
    try {
      size = (uint64_t)lgs::basefold_prove_impl(ctx->c, st->st, point, point_len, fields, n_fields, ints, n_ints,
                                                lgs::ProofOut{nullptr, proof, cap}, u);
    } catch (...) {
      ctx->c.drain();
      throw;
    }
    if (used) {
      used[0] = u[0];
      used[1] = u[1];
    }
    *proof_len = size;
    if (cap < size) throw lgs::Error(LGS_ERR_BUFFER, "basefold prove: proof needs " + std::to_string(size) + " bytes");
  });
}

void lgs_basefold_release(lgs_basefold* st) {
  if (!st) return;
  lgs::Context* c = st->ctx;
  if (c) (void)hipStreamSynchronize(c->stream);
  delete st;
  ctx_unref(c);
}

void lgs_key_release(lgs_key* key) {
  if (!key) return;
  lgs::Context* c = key->k.ctx;
  if (c) (void)hipStreamSynchronize(c->stream);
  delete key;
  ctx_unref(c);
}

int lgs_commit(lgs_ctx* ctx, const lgs_key* key, const uint64_t* f, uint64_t len, uint32_t flags,
               uint8_t root[32], lgs_commitment** out) {
  if (!ctx || !key || !root || !out) return LGS_ERR_ARG;
  *out = nullptr;
  return guarded(&ctx->c, [&] {
    (void)hipSetDevice(ctx->c.device);
    auto* c = new lgs_commitment();
    try {
      prepare_commitment(ctx->c, key->k, f, len, flags, c->c);
      // the tree's last launch writes the root straight into pinned host memory, then (unprofiled)
      // a completion word that is waited for instead of the stream
      uint8_t* h = ctx->c.staging(64);
      c->c.root_host = (uint32_t*)h;
      c->c.root_seq = ctx->c.profile ? 0 : next_root_seq(ctx->c);
      lgs::enqueue_commit(ctx->c, key->k, c->c);
      c->c.root_host = nullptr;
      if (c->c.root_seq) {
        wait_root(ctx->c, (const uint32_t*)h + 8, c->c.root_seq);
        c->c.device_done = true;
      } else {
        lgs::hip_check(hipStreamSynchronize(ctx->c.stream), "commit sync");
      }
      std::memcpy(c->c.root, h, 32);
      finish_profile(ctx->c);
    } catch (...) {
      ctx->c.drain();
      delete c;
      throw;
    }
    std::memcpy(root, c->c.root, 32);
    ctx_retain(&ctx->c);
    *out = c;
  });
}

int lgs_commit_batch(lgs_ctx* ctx, uint32_t count, const lgs_key* const* keys,
                     const uint64_t* const* fs, const uint64_t* lens, uint32_t flags,
                     uint8_t* roots, lgs_commitment** out) {
  if (!ctx || !keys || !fs || !lens || !roots || !out) return LGS_ERR_ARG;
  for (uint32_t i = 0; i < count; i++) out[i] = nullptr;
  std::vector<lgs_commitment*> made;
  int rc = guarded(&ctx->c, [&] {
    (void)hipSetDevice(ctx->c.device);
    std::vector<lgs::Commitment*> cs;
    std::vector<std::pair<const lgs::Key*, lgs::Commitment*>> grouped, single;
    for (uint32_t i = 0; i < count; i++) {
      if (!keys[i]) throw lgs::Error(LGS_ERR_ARG, "commit_batch: null key");
      auto* c = new lgs_commitment();
      made.push_back(c);
      prepare_commitment(ctx->c, keys[i]->k, fs[i], lens[i], flags, c->c);
      cs.push_back(&c->c);
      // the grouped kernels are the MFMA hash and whole-row transforms; other shapes (and a
      // forced VALU hash) take the per-polynomial path
      bool group = ctx->c.hash_variant != 2 && lgs::group_eligible(keys[i]->k);
      (group ? grouped : single).push_back({&keys[i]->k, &c->c});
    }
    lgs::GroupHold hold;
    const int prof = ctx->c.profile;
    ctx->c.profile = single.empty() ? (prof ? 1 : 0) : 0;  // the grouped path times every stage
    try {
      lgs::enqueue_commit_group(ctx->c, grouped, hold);
      ctx->c.profile = 0;
      for (auto& s : single) lgs::enqueue_commit(ctx->c, *s.first, *s.second);
      read_root(ctx->c, cs.data(), count);
      ctx->c.profile = prof ? 1 : 0;
      if (prof && single.empty() && !grouped.empty()) finish_profile(ctx->c);
      ctx->c.profile = prof;
    } catch (...) {
      ctx->c.profile = prof;
      throw;
    }
  });
  if (rc != LGS_OK) {
    ctx->c.drain();
    for (auto* c : made) delete c;
    return rc;
  }
  for (uint32_t i = 0; i < count; i++) {
    std::memcpy(roots + 32 * (size_t)i, made[i]->c.root, 32);
    ctx_retain(&ctx->c);
    out[i] = made[i];
  }
  return LGS_OK;
}

int lgs_prove(lgs_ctx* ctx, const lgs_key* key, const lgs_commitment* c, const uint64_t* z,
              uint64_t z_len, const uint64_t* fields, uint64_t n_fields, const uint64_t* ints,
              uint64_t n_ints, uint8_t* proof, uint64_t* proof_len, uint64_t used[2]) {
  if (!ctx || !key || !c || !z || !proof_len || (n_fields && !fields) || (n_ints && !ints))
    return LGS_ERR_ARG;
  return guarded(&ctx->c, [&] {
    (void)hipSetDevice(ctx->c.device);
    // encoded straight into the caller's buffer when it is large enough
    const uint64_t cap = proof ? *proof_len : 0;
    uint64_t u[2] = {0, 0}, size = 0;
    try {
      size = (uint64_t)lgs::prove_impl(ctx->c, key->k, c->c, z, z_len, fields, n_fields, ints, n_ints,
                                       lgs::ProofOut{nullptr, proof, cap}, u);
    } catch (...) {
      // no buffer released by the failed prove may be reused before every stream drains
      ctx->c.drain();
      throw;
    }
    if (used) {
      used[0] = u[0];
      used[1] = u[1];
    }
    *proof_len = size;
    if (cap < size) throw lgs::Error(LGS_ERR_BUFFER, "prove: proof needs " + std::to_string(size) + " bytes");
  });
}

int lgs_prove_batch(lgs_ctx* ctx, uint32_t count, const lgs_key* const* keys, const lgs_commitment* const* cs,
                    const uint64_t* const* zs, const uint64_t* z_lens, const uint64_t* fields, uint64_t n_fields,
                    const uint64_t* ints, uint64_t n_ints, uint8_t* const* proofs, uint64_t* proof_lens,
                    uint64_t used[2]) {
  if (!ctx || (count && (!keys || !cs || !zs || !z_lens || !proof_lens)) || (n_fields && !fields) ||
      (n_ints && !ints))
    return LGS_ERR_ARG;
  return guarded(&ctx->c, [&] {
    lgs::Context& c = ctx->c;
    (void)hipSetDevice(c.device);
    // stream positions: proof i starts where proof i - 1's draws end
    std::vector<uint64_t> f_off(count + 1, 0), i_off(count + 1, 0);
    for (uint32_t i = 0; i < count; i++) {
      if (!keys[i] || !cs[i] || !zs[i]) throw lgs::Error(LGS_ERR_ARG, "prove_batch: null key, commitment or point");
      uint64_t nf = 0, ni = 0;
      lgs::prove_draws(keys[i]->k.log_m, keys[i]->k.log_n, nf, ni);
      f_off[i + 1] = f_off[i] + nf;
      i_off[i + 1] = i_off[i] + ni;
    }
    if (f_off[count] > n_fields) throw lgs::Error(LGS_ERR_ORACLE, "prove_batch: randomness oracle field stream exhausted");
    if (i_off[count] > n_ints) throw lgs::Error(LGS_ERR_ORACLE, "prove_batch: randomness oracle integer stream exhausted");
    // worker 0 is this context; the others are its worker contexts, each with its own streams
    const uint32_t nw = count < 3 ? (count ? count : 1) : 3;
    std::vector<lgs::Context*> wk{&c};
    {
      std::lock_guard<std::mutex> g(c.workers_mu);
      while (c.workers.size() + 1 < nw) c.workers.push_back(lgs::make_context(c.device, true));
      for (uint32_t w = 1; w < nw; w++) wk.push_back(c.workers[w - 1].get());
    }
    std::vector<uint64_t> sizes(count, 0);
    std::vector<int> codes(nw, LGS_OK);
    std::vector<std::string> msgs(nw);
    auto run = [&](uint32_t w) {
      (void)hipSetDevice(c.device);
      for (uint32_t i = w; i < count && codes[w] == LGS_OK; i += nw) {
        try {
          const uint64_t cap = proofs && proofs[i] ? proof_lens[i] : 0;
          uint64_t u[2];
          sizes[i] = (uint64_t)lgs::prove_impl(*wk[w], keys[i]->k, cs[i]->c, zs[i], z_lens[i], fields + 2 * f_off[i],
                                               n_fields - f_off[i], ints + i_off[i], n_ints - i_off[i],
                                               lgs::ProofOut{nullptr, cap ? proofs[i] : nullptr, cap}, u);
        } catch (const lgs::Error& e) {
          codes[w] = e.code;
          msgs[w] = e.what();
        } catch (const std::bad_alloc&) {
          codes[w] = LGS_ERR_NOMEM;
          msgs[w] = "host allocation failed";
        } catch (const std::exception& e) {
          codes[w] = LGS_ERR_ARG;
          msgs[w] = e.what();
        }
      }
      if (codes[w] != LGS_OK) wk[w]->drain();
    };
    std::vector<std::thread> th;
    for (uint32_t w = 1; w < nw; w++) th.emplace_back(run, w);
    run(0);
    for (auto& t : th) t.join();
    for (uint32_t w = 0; w < nw; w++)
      if (codes[w] != LGS_OK) throw lgs::Error(codes[w], msgs[w]);
    bool fit = true;
    for (uint32_t i = 0; i < count; i++) {
      const uint64_t cap = proofs && proofs[i] ? proof_lens[i] : 0;
      fit = fit && cap >= sizes[i];
      proof_lens[i] = sizes[i];
    }
    if (used) {
      used[0] = f_off[count];
      used[1] = i_off[count];
    }
    if (!fit) throw lgs::Error(LGS_ERR_BUFFER, "prove_batch: a proof buffer is too small (sizes reported)");
  });
}

void lgs_commitment_release(lgs_commitment* c) {
  if (!c) return;
  lgs::Context* cx = c->c.ctx;
  if (cx && !c->c.device_done) (void)hipStreamSynchronize(cx->stream);
  delete c;
  ctx_unref(cx);
}

int lgs_commitment_read(const lgs_commitment* c, uint32_t what, void* out, uint64_t cap) {
  if (!c || !out) return LGS_ERR_ARG;
  lgs::Context* ctx = c->c.ctx;
  return guarded(ctx, [&] {
    const void* src = nullptr;
    uint64_t bytes = 0;
    uint64_t m = uint64_t(1) << c->c.log_m, two_n = uint64_t(2) << c->c.log_n;
    if (what == LGS_READ_ENCODED) {
      src = c->c.enc.ptr;
      bytes = m * two_n * 8;
    } else if (what == LGS_READ_DIGESTS) {
      src = c->c.h.ptr;
      bytes = 32 * two_n * 8;
    } else if (what == LGS_READ_TREE) {
      src = c->c.tree.nodes.ptr;
      bytes = c->c.tree.node_count() * 32;
    } else {
      throw lgs::Error(LGS_ERR_ARG, "commitment_read: unknown selector");
    }
    if (cap < bytes) throw lgs::Error(LGS_ERR_BUFFER, "commitment_read: need " + std::to_string(bytes) + " bytes");
    lgs::hip_check(hipMemcpyAsync(out, src, bytes, hipMemcpyDeviceToHost, ctx->stream), "read");
    lgs::hip_check(hipStreamSynchronize(ctx->stream), "read sync");
  });
}

// ---- serving loop ----
static thread_local std::string g_server_err;

const char* lgs_server_last_error(void) { return g_server_err.c_str(); }

int lgs_server_start(int device, uint64_t hbm_budget, uint32_t prove_workers, uint32_t max_batch,
                     lgs_server** out) {
  if (!out || hbm_budget == 0) return LGS_ERR_ARG;
  *out = nullptr;
  try {
    auto* w = new lgs_server();
    w->s = lgs::server_start(device, hbm_budget, prove_workers, max_batch);
    *out = w;
    return LGS_OK;
  } catch (const lgs::Error& e) {
    g_server_err = e.what();
    return e.code;
  }
}

int lgs_server_submit(lgs_server* srv, const lgs_key* key, const uint64_t* f, uint64_t len, uint32_t flags,
                      const uint64_t* z, uint64_t z_len, const uint64_t* fields, uint64_t n_fields,
                      const uint64_t* ints, uint64_t n_ints, uint64_t* ticket) {
  if (!srv || !key || !f || !z || !ticket || (n_fields && !fields) || (n_ints && !ints)) return LGS_ERR_ARG;
  try {
    *ticket = lgs::server_submit(srv->s, &key->k, f, len, flags, z, z_len, fields, n_fields, ints, n_ints);
    return LGS_OK;
  } catch (const lgs::Error& e) {
    g_server_err = e.what();
    return e.code;
  }
}

int lgs_server_submit_commit(lgs_server* srv, const lgs_key* key, const uint64_t* f, uint64_t len, uint32_t flags,
                             uint64_t* ticket) {
  if (!srv || !key || !f || !ticket) return LGS_ERR_ARG;
  try {
    *ticket = lgs::server_submit_commit(srv->s, &key->k, f, len, flags);
    return LGS_OK;
  } catch (const lgs::Error& e) {
    g_server_err = e.what();
    return e.code;
  }
}

int lgs_server_submit_open(lgs_server* srv, uint64_t commitment, const uint64_t* z, uint64_t z_len,
                           const uint64_t* fields, uint64_t n_fields, const uint64_t* ints, uint64_t n_ints,
                           uint64_t* ticket) {
  if (!srv || !z || !ticket || (n_fields && !fields) || (n_ints && !ints)) return LGS_ERR_ARG;
  try {
    *ticket = lgs::server_submit_open(srv->s, commitment, z, z_len, fields, n_fields, ints, n_ints);
    return LGS_OK;
  } catch (const lgs::Error& e) {
    g_server_err = e.what();
    return e.code;
  }
}

int lgs_server_release(lgs_server* srv, uint64_t commitment) {
  if (!srv) return LGS_ERR_ARG;
  try {
    lgs::server_release(srv->s, commitment);
    return LGS_OK;
  } catch (const lgs::Error& e) {
    g_server_err = e.what();
    return e.code;
  }
}

int lgs_server_poll(lgs_server* srv, uint64_t ticket, int wait, uint8_t root[32], uint8_t* proof,
                    uint64_t* proof_len, double times_ms[3]) {
  if (!srv || !root || !proof_len) return LGS_ERR_ARG;
  try {
    std::string err;
    int rc = lgs::server_poll(srv->s, ticket, wait != 0, root, proof, proof_len, times_ms, err);
    if (rc != LGS_OK && rc != LGS_PENDING && rc != LGS_ERR_BUFFER) g_server_err = err;
    return rc;
  } catch (const lgs::Error& e) {
    g_server_err = e.what();
    return e.code;
  }
}

int lgs_server_stats(lgs_server* srv, uint64_t out[4]) {
  if (!srv || !out) return LGS_ERR_ARG;
  lgs::server_stats(srv->s, out);
  return LGS_OK;
}

void lgs_server_stop(lgs_server* srv) {
  if (!srv) return;
  lgs::server_stop(srv->s);
  delete srv;
}

}  // extern "C"
