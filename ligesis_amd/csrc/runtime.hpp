// Host-side runtime objects behind the C ABI: context (stream, caching device allocator,
// twiddle tables, scratch), keys and commitments.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <map>
#include <memory>
#include <cstdlib>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "kernels.hpp"

namespace lgs {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw Error(3 /*LGS_ERR_DEVICE*/, std::string(what) + ": " + hipGetErrorString(e));
}

// Caching device allocator: freed blocks are kept per size class and reused, so steady-state
// commits and proofs do not call hipMalloc.
class DevicePool {
 public:
  ~DevicePool() { trim(); }
  // Test aid: with poison = b (0..255), every block handed out is first filled with byte b,
  // after the whole device has drained (blocks are reused in stream order, so a block may still
  // be read by queued work when it is freed). A kernel that reads memory it did not write then
  // sees the pattern instead of stale or zeroed contents. Slow; set from LGS_POOL_POISON.
  int poison = poison_from_env();
  static int poison_from_env() {
    const char* v = std::getenv("LGS_POOL_POISON");
    return v && *v ? (std::atoi(v) & 255) : -1;
  }
  void* alloc(size_t bytes) {
    void* p = take(bytes);
    if (poison >= 0) {
      (void)hipDeviceSynchronize();
      (void)hipMemset(p, poison, size_class(bytes));
      (void)hipDeviceSynchronize();
    }
    return p;
  }
  void* take(size_t bytes) {
    size_t cls = size_class(bytes);
    {
      std::lock_guard<std::mutex> g(mu_);
      auto& fl = free_[cls];
      if (!fl.empty()) {
        void* p = fl.back();
        fl.pop_back();
        return p;
      }
    }
    void* p = nullptr;
    if (hipMalloc(&p, cls) != hipSuccess) {
      (void)hipGetLastError();
      trim();
      if (hipMalloc(&p, cls) != hipSuccess) {
        (void)hipGetLastError();
        throw Error(4 /*LGS_ERR_NOMEM*/, "device allocation of " + std::to_string(cls) + " bytes failed");
      }
    }
    std::lock_guard<std::mutex> g(mu_);
    sizes_[p] = cls;
    live_ += cls;
    return p;
  }
  void release(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> g(mu_);
    auto it = sizes_.find(p);
    if (it == sizes_.end()) return;  // not from this pool: never throw from a destructor path
    free_[it->second].push_back(p);
  }
  void trim() {
    std::lock_guard<std::mutex> g(mu_);
    for (auto& kv : free_)
      for (void* p : kv.second) {
        (void)hipFree(p);
        live_ -= sizes_[p];
        sizes_.erase(p);
      }
    free_.clear();
  }
  size_t reserved() const { return live_; }
  static size_t size_class(size_t b) {
    if (b <= 256) return 256;
    size_t c = 256;
    while (c < b) c <<= 1;  // powers of two up to 2 MiB, then 2 MiB multiples
    if (c > (size_t(2) << 20)) c = (b + (size_t(2) << 20) - 1) & ~((size_t(2) << 20) - 1);
    return c;
  }

 private:
  std::mutex mu_;
  std::map<size_t, std::vector<void*>> free_;
  std::map<void*, size_t> sizes_;
  size_t live_ = 0;
};

template <class T>
struct DevBuf {
  DevicePool* pool = nullptr;
  T* ptr = nullptr;
  size_t count = 0;
  DevBuf() = default;
  DevBuf(DevicePool* p, size_t n) : pool(p), ptr((T*)p->alloc(n * sizeof(T) ? n * sizeof(T) : 1)), count(n) {}
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept { *this = std::move(o); }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      reset();
      pool = o.pool;
      ptr = o.ptr;
      count = o.count;
      o.ptr = nullptr;
      o.count = 0;
    }
    return *this;
  }
  ~DevBuf() { reset(); }
  void reset() {
    if (ptr && pool) pool->release(ptr);
    ptr = nullptr;
    count = 0;
  }
  size_t bytes() const { return count * sizeof(T); }
};

// A Merkle tree resident in HBM: all layers, leaves first (see merkle.hip).
struct DevTree {
  DevBuf<uint32_t> nodes;
  uint64_t leaves = 0;
  uint64_t node_count() const { return 2 * leaves - 1; }
  uint64_t layer_offset(int k) const {
    uint64_t off = 0, c = leaves;
    for (int i = 0; i < k; i++) {
      off += c;
      c >>= 1;
    }
    return off;
  }
};

// Launch schedule of one commit shape. Every choice yields the same bytes; only speed differs.
struct Tuning {
  int hash_variant = 0;  // 1 MFMA, 2 VALU (0: MFMA where eligible)
  int hash_impl = 17;    // MFMA kernel variant (17: 8 waves sharing an 8-slab LDS ring, copies and bit
                         // expansion interleaved with the products, one barrier per two slabs)
  int hash_ksplit = 0;   // slabs of the K (= 64 m) dimension; 0: heuristic
  int ntt_impl = 0;      // row transform variant
  int merkle_lg = 7;     // log2 nodes per Merkle workgroup (four threads per node)
  int pipeline = 1;      // row chunks whose transform overlaps the hash of the previous chunk (1: off)
};

// Process-wide table of measured schedules, keyed by (log_m, log_n) (tune.cpp). Shared by every
// context, so server workers use what an autotune run on another context found.
struct TuneEntry {
  Tuning t;
  double ms = 0;  // commit time measured with this schedule
};
void tune_store(uint32_t log_m, uint32_t log_n, const TuneEntry& e);
bool tune_lookup(uint32_t log_m, uint32_t log_n, TuneEntry& e);
void tune_clear();
size_t tune_save(const char* path);  // entries written
size_t tune_load(const char* path);  // entries read; throws on a malformed file
Tuning builtin_tuning(uint32_t log_m, uint32_t log_n);

// Basefold prover state on the GPU (pcs/src/basefold.rs:36-54): interleaved codewords
// (rows of length D; extension-field polys use two rows per codeword, c0 then c1), their
// column Merkle tree, and the polynomial itself.
struct DevBasefold {
  bool ext = false;
  uint32_t num_vars = 0, li = 0, code_rate = 1;
  uint64_t D = 0;
  DevBuf<uint64_t> codes;  // (ext ? 2 : 1) * 2^li rows x D
  DevTree tree;
  DevBuf<uint64_t> poly;  // 2^num_vars base values, or 2^num_vars (c0, c1) pairs
  const uint64_t* poly_ptr = nullptr;  // poly.ptr, or borrowed storage
  uint8_t root[32];
  uint32_t code_rows() const { return (ext ? 2u : 1u) << li; }
};

// Pinned host staging for uploads: a copy from pageable memory starts only once its stream has
// drained, which holds the issuing thread behind the device. Bump-allocated in chunks (256-byte
// aligned); reset() may run only when no copy from the arena is pending (after the streams that
// read it were synchronized) and merges the chunks into one.
class PinnedArena {
 public:
  PinnedArena() = default;
  PinnedArena(const PinnedArena&) = delete;
  PinnedArena& operator=(const PinnedArena&) = delete;
  ~PinnedArena() { release(); }
  void reset();
  // copies src into the arena and enqueues the copy to dst; returns the staged host copy
  void* upload(void* dst, const void* src, size_t bytes, hipStream_t s);
  void release();
  void take(PinnedArena& o) {  // o's memory (o must have no pending copy), reset
    release();
    chunks_.swap(o.chunks_);
    reset();
  }

 private:
  std::vector<std::pair<uint8_t*, size_t>> chunks_;
  size_t used_ = 0;  // bytes used in the last chunk
};

struct Context {
  // Keys and commitments keep their context alive: lgs_shutdown drops the caller's reference
  // and the context is torn down when the last key / commitment made on it is released.
  std::atomic<int> refs{1};
  void* owner = nullptr;  // the lgs_ctx that embeds this context
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;  // second stream (default priority) for work that overlaps `stream`
  std::string err;
  DevicePool pool;
  std::map<int, DevBuf<uint64_t>> twiddles;
  std::map<int, DevBuf<uint64_t>> pass_twiddles;  // per-pass factor tables (ntt_impl 4, 5), key impl << 8 | logN
  // Explicit launch-schedule options (lgs_set_option); unset fields fall back to the process
  // tuning table, then to the built-in defaults (see resolve).
  int hash_variant = 0;  // 0 auto, 1 MFMA, 2 VALU
  int hash_ksplit = 0;   // 0 auto
  int hash_impl = -1;    // -1 auto; MFMA kernel: 0 pipelined 2x2 waves, 1 one slab per stage, 2 basic,
                         // 3 2x4 waves, 4 loads issued 2 slabs ahead, 5 as 4 with 2x4 waves,
                         // 6 / 7 64x64 wave tiles in 8 / 16-wave workgroups,
                         // 8 / 9 64x128 wave tiles in 4 / 8-wave workgroups, 10 as 0 with A
                         // fragments read ahead, 11 / 12 asynchronous copies into an LDS ring,
                         // 14 8 waves sharing a 4-slab ring, 15 / 16 pipelined 8 waves with a 4- /
                         // 8-slab ring, 17 as 16 with one barrier per two slabs
  int ntt_impl = -1;     // -1 auto; row transform: 0/2/3 radix-8 with 256/512/1024 threads, 1 radix-2
  int merkle_lg = 0;     // 0 auto; log2 nodes per Merkle workgroup = levels per launch (6..8)
  int pipeline = 0;      // 0 auto; 1 off, 2 / 4 / 8 row chunks (transform on stream, hash on stream2)
  // replay repeated proofs of a commitment as a captured graph (off by default: on this ROCm the
  // graph ran the two streams' steps in capture order, slower than issuing them directly)
  bool prove_graph = false;
  // 1: time every stage with HIP events; 2: a plain commit records only the two events around
  // its hash kernel (stage 1), everything else as 1
  int profile = 0;
  // proofs on this context issue both halves on `stream` (server workers: fewer streams than
  // the device has hardware queues, so no stream shares a queue with an upload)
  bool single_stream_prove = false;
  double stage_ms[28] = {0};
  // ev[0..4]: stage timestamps of a profiled commit (and the autotuner's); ev[5..9]: the
  // prove's cross-stream fork / join. Created by create_events.
  hipEvent_t ev[10] = {};
  // Timing events carry no system-scope fence: recording one with the fence writes back the
  // L2 caches, which stalled the next kernel's start by 5-7 us after a kernel that wrote tens
  // of MiB (a visible share of a 0.5 ms commit). They order nothing; the fork / join events
  // keep the default fence.
  hipError_t create_events();
  uint8_t* pinned = nullptr;  // small pinned staging buffer for roots
  size_t pinned_size = 0;
  // Pinned staging of the job tables and constants a batched commit or a prove uploads.
  PinnedArena arena;
  PinnedArena prove_arena;  // staging a proof borrows for its uploads (see prove.cpp)
  void upload_reset() { arena.reset(); }
  void upload(void* dst, const void* src, size_t bytes, hipStream_t s) { arena.upload(dst, src, bytes, s); }
  void free_host();  // pinned staging and arena

  const uint64_t* twiddle(int logN);
  const uint64_t* pass_twiddle(int logN, int impl);  // null where the transform has no such table
  uint8_t* staging(size_t bytes);
  hipStream_t side_stream();  // stream2, created on first use
  void drain();                // waits for every stream, ignoring errors (error and teardown paths)
  void destroy_streams();
  hipEvent_t pipe_ev[10] = {};  // cross-stream events of a pipelined commit, created on first use
  hipEvent_t pipe_event(int i);
  // profiled pipelined commit: events around chunk q's transform (4q, 4q + 1) and hash (4q + 2,
  // 4q + 3); prof_chunks = number of chunks of the last profiled commit (0: not pipelined)
  hipEvent_t tev[48] = {};  // [32..42]: phases of a profiled prove
  int prof_chunks = 0;
  uint32_t root_seq = 0;  // last completion word handed to a commit's tree (wait_root)
  hipEvent_t timing_event(int i);
  // launch schedule of a commit of shape 2^log_m x 2^log_n
  Tuning resolve(uint32_t log_m, uint32_t log_n) const;
  // contexts of the batched prove's worker threads (lgs_prove_batch), made on first use
  std::vector<std::unique_ptr<Context>> workers;
  std::mutex workers_mu;
};

struct Key {
  Context* ctx = nullptr;
  uint32_t log_m = 0, log_n = 0;
  uint64_t W = 0;
  DevBuf<uint64_t> a;    // 32 x W row-major (also A-bar flat, subset_sum.rs:73-79)
  DevBuf<uint64_t> at;   // W x 32 (VALU hash layout)
  DevBuf<int8_t> dig;    // MFMA digit layout
  bool has_state = false;
  // the Z_2^64 subset-sum hash (commit only): H = A . B with sums mod 2^64, A any 64-bit words
  bool ring64 = false;
  DevBasefold a_state;
  uint8_t com_a[32] = {0};
};

struct ProvePlan;  // prove.cpp

struct Commitment {
  Context* ctx = nullptr;
  const Key* key = nullptr;
  uint32_t log_m = 0, log_n = 0;
  DevBuf<uint64_t> f_owned;
  const uint64_t* f = nullptr;  // f_owned.ptr or a borrowed device pointer
  DevBuf<uint64_t> enc;         // m x 2n
  DevBuf<uint64_t> h;           // 32 x 2n
  DevTree tree;
  uint8_t root[32] = {0};
  // when set (host-mapped pinned memory), the tree's last launch also writes the root there,
  // so reading it back needs no copy after the tree; and root_seq, if not zero, after it in
  // root_host[8]
  uint32_t* root_host = nullptr;
  uint32_t root_seq = 0;
  // the commit's kernels have made their last memory access (its tree wrote the completion word
  // and nothing was enqueued on the commitment since): release needs no stream wait
  bool device_done = false;
  // device state of the last proof on ctx, replayed as a graph by later proofs (see prove.cpp)
  mutable std::shared_ptr<ProvePlan> plan;
  mutable uint32_t prove_calls = 0;
};

// Buffers and host job tables of a batched commit, kept until the caller synchronises.
struct GroupHold {
  std::vector<DevBuf<int32_t>> parts;
  std::vector<DevBuf<uint64_t>> tables;
  std::vector<NttJob> host_ntt;
  std::vector<HashJob> host_hash;
  std::vector<FinalizeJob> host_fin;
  std::vector<std::vector<MerkleJob>> host_merkle;
};

// commit.cpp: a context with its stream and events on `device`; the release of everything a
// context holds (its workers' contexts, streams, events, tables, pool and pinned staging)
// A context with its stream; with two_streams also its second stream, created right after the
// first: HIP maps streams to a few hardware queues in creation order, and a second stream made
// later (after other contexts' streams) can land on the first one's queue, which serializes the
// prove's two streams (0.82 ms instead of 0.57 ms at mu = 24).
std::unique_ptr<Context> make_context(int device, bool two_streams = false);
void teardown_context(Context& c);
// commit.cpp; `force` replaces the resolved schedule (autotune candidates)
void enqueue_commit(Context& ctx, const Key& key, Commitment& c, const Tuning* force = nullptr);
// The hash kernel of a commit: variant (1 MFMA, 2 VALU), MFMA kernel and K-split.
struct HashPlan {
  int variant = 1;
  int impl = 16;
  uint32_t ks = 1;
};
HashPlan hash_plan(const Key& key, const Tuning& tu);
uint64_t hash_partial_bytes(const HashPlan& hp, uint64_t two_n);  // K-split partials of one commit
// Device buffers of one commit: f (m x n), F' (m x 2n), H (32 x 2n), the tree's nodes
// ((4n - 1) x 8 words), the hash partials (hash_partial_bytes), the transform's twiddle table and
// (ntt_impl 4 only) its pass table; root_host, if set, receives the root from the tree's last launch
// (and root_seq, if not zero, in root_host[8] after it).
struct CommitIO {
  const uint64_t* f = nullptr;
  uint64_t* enc = nullptr;
  uint64_t* h = nullptr;
  uint32_t* nodes = nullptr;
  uint64_t* part = nullptr;
  const uint64_t* tw = nullptr;
  const uint64_t* tp = nullptr;
  uint32_t* root_host = nullptr;
  uint32_t root_seq = 0;
};
// Row transform, hash, finalize and Merkle tree of one commit on stream st, into io's buffers;
// allocates nothing.
void enqueue_commit_stages(Context& ctx, const Key& key, const CommitIO& io, const HashPlan& hp, const Tuning& tu,
                           hipStream_t st, int profile);
bool group_eligible(const Key& key);
void enqueue_commit_group(Context& ctx, const std::vector<std::pair<const Key*, Commitment*>>& items,
                          GroupHold& hold);
// Merkle trees over many sources (leaf p = rows words M[r * rs + p * cs]), one launch per
// reduction round of up to 8 levels across all trees; the job tables stay in host_jobs / tables
// until the caller synchronises.
struct MerkleSource {
  const uint64_t* M;
  uint32_t rows;
  uint64_t rs, cs, leaves;
  uint32_t* nodes;
};
void enqueue_merkle_trees(Context& ctx, const std::vector<MerkleSource>& trees,
                          std::vector<std::vector<MerkleJob>>& host_jobs, std::vector<DevBuf<uint64_t>>& tables,
                          hipStream_t st, PinnedArena* arena = nullptr);
// On stream `s` (default: the context stream); `encoded`, if set, is recorded once the codewords
// are written (before the Merkle tree). The twiddle table of the codeword length must already
// exist when `s` is not the context stream.
void enqueue_basefold_commit(Context& ctx, DevBasefold& st, const uint64_t* poly, bool ext,
                             uint32_t num_vars, uint32_t li, uint32_t code_rate, hipStream_t s = nullptr,
                             hipEvent_t encoded = nullptr);
void setup_key(Context& ctx, Key& key, const uint64_t* host_a, bool commit_a);
// tune.cpp: times candidate schedules of key's shape on a synthetic polynomial and stores the
// fastest in the process table; returns the best commit time in ms
double autotune(Context& ctx, const Key& key, uint32_t reps, TuneEntry& best);
// prove.cpp
// Where a proof's bytes go: a vector (resized to the proof), or a caller's buffer of cap bytes,
// written only when the whole proof fits (the returned size says whether it did).
struct ProofOut {
  std::vector<uint8_t>* vec = nullptr;
  uint8_t* dst = nullptr;
  uint64_t cap = 0;
};
// The numbers of field and integer draws one proof at (log_m, log_n) takes from the oracle.
void prove_draws(uint32_t log_m, uint32_t log_n, uint64_t& n_fields, uint64_t& n_ints);
// returns the proof's size in bytes
int64_t prove_impl(Context& ctx, const Key& key, const Commitment& cm, const uint64_t* z, uint64_t z_len,
                   const uint64_t* fields, uint64_t n_fields, const uint64_t* ints, uint64_t n_ints,
                   ProofOut out, uint64_t used[2]);

// noalloc.hip: with `device` current, looks up the device address of the no-allocation entry
// point's scratch array on that device (once per device; throws on failure)
void scratch_init(int device);

// basefold.cpp: Basefold::prove (basefold.rs:384-516) of a resident commitment; returns the
// proof's size in bytes
int64_t basefold_prove_impl(Context& ctx, const DevBasefold& bf, const uint64_t* point, uint64_t point_len,
                            const uint64_t* fields, uint64_t n_fields, const uint64_t* ints, uint64_t n_ints,
                            ProofOut out, uint64_t used[2]);

// server.cpp
struct Server;
uint64_t request_footprint(uint32_t log_m, uint32_t log_n, bool host_input);
Server* server_start(int device, uint64_t budget, uint32_t workers, uint32_t max_batch);
uint64_t server_submit(Server* s, const Key* key, const uint64_t* f, uint64_t len, uint32_t flags,
                       const uint64_t* z, uint64_t z_len, const uint64_t* fields, uint64_t nf,
                       const uint64_t* ints, uint64_t ni);
// Commit-only request (its commitment stays resident until server_release) and an open of a
// resident commitment, by the commit-only request's ticket.
uint64_t server_submit_commit(Server* s, const Key* key, const uint64_t* f, uint64_t len, uint32_t flags);
uint64_t server_submit_open(Server* s, uint64_t commitment, const uint64_t* z, uint64_t z_len, const uint64_t* fields,
                            uint64_t nf, const uint64_t* ints, uint64_t ni);
void server_release(Server* s, uint64_t commitment);
int server_poll(Server* s, uint64_t ticket, bool wait, uint8_t root[32], uint8_t* proof, uint64_t* proof_len,
                double times[3], std::string& err);
void server_stop(Server* s);
void server_stats(Server* s, uint64_t out[4]);

// Shape helpers restating pcs/src/ligesis.rs:175-203.
inline void committed_vars(uint32_t lm, uint32_t ln, uint32_t out[3]) {
  out[0] = 11 + lm;  // A-bar: C * eta * m
  out[1] = 13 + lm;  // B_I: eta * m * s
  out[2] = ln;       // a-bar: n
}
inline uint32_t max_vars(uint32_t lm, uint32_t ln) {
  uint32_t v[3];
  committed_vars(lm, ln, v);
  return v[0] > v[1] ? (v[0] > v[2] ? v[0] : v[2]) : (v[1] > v[2] ? v[1] : v[2]);
}
inline uint32_t min_vars(uint32_t lm, uint32_t ln) {
  uint32_t v[3];
  committed_vars(lm, ln, v);
  return v[0] < v[1] ? (v[0] < v[2] ? v[0] : v[2]) : (v[1] < v[2] ? v[1] : v[2]);
}
inline bool shape_ok(uint32_t lm, uint32_t ln) { return max_vars(lm, ln) - min_vars(lm, ln) <= 6; }
inline uint32_t batch_rounds(uint32_t lm, uint32_t ln) { return max_vars(lm, ln) - 6; }
inline uint32_t shared_li(uint32_t nv, uint32_t lm, uint32_t ln) { return nv - batch_rounds(lm, ln); }

}  // namespace lgs
