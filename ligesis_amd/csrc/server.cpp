// Serving loop: commit-then-open requests of mixed sizes, continuously batched.
//
// Three kinds of threads share one device:
//  - the commit thread admits queued requests in arrival order while their HBM footprint fits
//    the budget (up to max_batch at a time) and enqueues, for each, the upload of f on the upload
//    stream and its commit on the commit stream. A large f is uploaded in row chunks and each
//    chunk's rows are transformed and hashed (its K-splits) as soon as the chunk lands, so the
//    commit ends shortly after the upload. The thread never waits for the device: it records one
//    event per admitted group and goes back to admitting;
//  - the finisher thread waits for those events in order, reads the roots (copied into pinned
//    memory by the commit stream) and hands the resident commitments (f, F', H, Merkle tree) to
//    the prove workers;
//  - P prove workers prove on their own contexts, one stream each, so proofs run beside each
//    other, the uploads and the next commits.
// Memory returns to the budget when a proof is done; requests that arrive meanwhile are admitted
// as soon as they fit (continuous batching). Results are identical to lgs_commit + lgs_prove.
// Besides commit-then-open requests the loop takes commit-only requests, whose commitment stays
// resident (f, F', H, tree in HBM) until released, and open requests on such a commitment, queued
// like the others and proved by the workers (several opens of one commitment may run at once).
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <thread>

#include "runtime.hpp"

namespace lgs {

namespace {
double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

constexpr uint64_t kChunkMinBytes = uint64_t(64) << 20;   // f of at least this size is uploaded in chunks
constexpr uint64_t kChunkPartsMax = uint64_t(256) << 20;  // cap on a chunked commit's hash partials
constexpr uint32_t kRootSlots = 8192;                     // pinned root slots = requests in flight
}  // namespace

// Row chunks of a served commit of a host-resident f: each chunk's rows are hashed by their own
// K-splits, so the count divides m / 2 and the partials (one 32 x 2n word plane per split) stay
// under kChunkPartsMax.
static uint32_t serve_chunks(uint32_t log_m, uint32_t log_n, bool host_input) {
  const uint64_t m = uint64_t(1) << log_m, two_n = uint64_t(2) << log_n;
  const uint64_t fbytes = (m << log_n) * 8;
  if (!host_input || fbytes < kChunkMinBytes || two_n % 128 != 0 || m < 4) return 1;
  uint32_t c = fbytes >= (uint64_t(256) << 20) ? 8 : 4;
  while (c > 1 && ((uint64_t)c * 32 * two_n * 8 > kChunkPartsMax || m % (2 * c) != 0)) c /= 2;
  return c;
}

// Conservative device footprints at (log_m, log_n), in bytes: a commit (with its hash partials:
// at most the larger of the built-in K-split, 8 int32 planes a split, and the chunked commit's
// splits, one word plane a split) and one open of it.
static uint64_t commit_footprint(uint32_t log_m, uint32_t log_n, bool host_input) {
  const uint64_t m = uint64_t(1) << log_m, n = uint64_t(1) << log_n, two_n = 2 * n;
  const uint64_t ks = m >= 1024 ? m / 512 : 1;
  const uint64_t parts = std::max<uint64_t>(ks * 256 * two_n * 4,
                                            (uint64_t)serve_chunks(log_m, log_n, host_input) * 32 * two_n * 8);
  return (host_input ? 8 * m * n : 0) + 16 * m * n + 32 * two_n * 8 + 2 * two_n * 32 + parts;
}

static uint64_t open_footprint(uint32_t log_m, uint32_t log_n) {
  const uint64_t m = uint64_t(1) << log_m, n = uint64_t(1) << log_n, W = 64 * m, s = 128;
  const uint64_t two_n = 2 * n, bi = W * s;
  const uint64_t D = uint64_t(2) << batch_rounds(log_m, log_n);
  return bi * (8 + 4 * 16) + 64 * D * 8 + 2 * D * 32 + 4 * D * 16 + 24 * n * 16 + W * 6 * 16 + two_n * 64 +
         (uint64_t(1) << 24);
}

uint64_t request_footprint(uint32_t log_m, uint32_t log_n, bool host_input) {
  return commit_footprint(log_m, log_n, host_input) + open_footprint(log_m, log_n);
}

enum Kind { kCommitOpen = 0, kCommitOnly = 1, kOpen = 2 };

struct Request {
  uint64_t ticket = 0;
  int kind = kCommitOpen;
  uint64_t of = 0;  // kOpen: the ticket of the resident commitment
  const Key* key = nullptr;
  const uint64_t* f = nullptr;
  uint64_t len = 0;
  uint32_t flags = 0;
  std::vector<uint64_t> z, fields, ints;
  uint64_t footprint = 0;
  double t_submit = 0, t_committed = 0, t_done = 0;
  std::shared_ptr<Commitment> cm;  // kOpen: the resident commitment, shared with other opens
  uint32_t slot = 0;  // pinned root slot
  uint8_t root[32] = {0};
  std::vector<uint8_t> proof;
  int status = -1;        // -1 pending, else final status
  int enqueue_code = 0;   // set when enqueueing its commit failed (reported after the group's event)
  std::string err;
};

// An admitted group whose uploads and commits are enqueued; the buffers a commit needs only
// while it runs (hash partials) and the upload events stay here until the group's event.
struct Group {
  std::vector<std::shared_ptr<Request>> reqs;
  std::vector<DevBuf<uint64_t>> parts;
  std::vector<hipEvent_t> events;
  hipEvent_t done = nullptr;
  int rc = 0;
  std::string msg;
};

// A commitment kept in HBM by a commit-only request until released. `footprint` is its commit
// footprint, returned to the budget once it is released and no open of it runs.
struct Resident {
  const Key* key = nullptr;
  std::shared_ptr<Commitment> cm;  // set when the commit is done
  uint64_t footprint = 0;
  uint32_t opens = 0;  // opens admitted and not yet finished
  bool released = false, failed = false;
  std::vector<std::shared_ptr<Request>> waiting;  // opens admitted before the commit was done
};

struct Server {
  int device = 0;
  uint64_t budget = 0, reserved = 0;
  uint32_t max_batch = 16;
  std::mutex mu;
  std::condition_variable cv_commit, cv_finish, cv_prove, cv_done;
  std::deque<std::shared_ptr<Request>> pending, committed;
  std::deque<std::unique_ptr<Group>> finishing;
  std::map<uint64_t, std::shared_ptr<Request>> all;
  std::map<uint64_t, Resident> resident;  // by the commit-only request's ticket
  uint64_t next_ticket = 1;
  uint32_t in_flight = 0;  // admitted requests not yet handed to the prove workers
  bool stopping = false;
  bool commit_running = true, finish_running = true;
  std::vector<std::unique_ptr<Context>> ctxs;  // [0] commit, [1..] prove
  std::vector<std::thread> threads;
  uint64_t batches = 0, batched_requests = 0;

  hipStream_t upload_stream = nullptr;  // host-to-device copies of request inputs
  uint8_t* roots = nullptr;             // kRootSlots x 32 bytes of pinned memory
  uint32_t next_slot = 0;
  std::vector<hipEvent_t> free_events;  // events whose last use has completed (guarded by mu)

  hipEvent_t event();
  void drop_resident_locked(uint64_t ticket);
  void enqueue_request(Group& g, Request& r);
  void commit_loop();
  void finish_loop();
  void prove_loop(Context* ctx);
};

hipEvent_t Server::event() {
  {
    std::lock_guard<std::mutex> lk(mu);
    if (!free_events.empty()) {
      hipEvent_t e = free_events.back();
      free_events.pop_back();
      return e;
    }
  }
  hipEvent_t e = nullptr;
  hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "server event");
  return e;
}

// Frees a released resident commitment once no open of it runs (mu held).
void Server::drop_resident_locked(uint64_t ticket) {
  auto it = resident.find(ticket);
  if (it == resident.end() || !it->second.released || it->second.opens) return;
  if (!it->second.cm && !it->second.failed) return;  // its commit is still in flight
  reserved -= it->second.footprint;
  resident.erase(it);
  cv_commit.notify_all();
}

// Upload (if f is in host memory) and commit of one request, enqueued; nothing waits.
void Server::enqueue_request(Group& g, Request& r) {
  Context& ctx = *ctxs[0];
  const Key& key = *r.key;
  r.cm.reset(new Commitment());
  Commitment& c = *r.cm;
  c.ctx = &ctx;
  c.key = r.key;
  c.log_m = key.log_m;
  c.log_n = key.log_n;
  const uint64_t m = uint64_t(1) << key.log_m, n = uint64_t(1) << key.log_n, two_n = 2 * n;
  const int logN = key.log_n + 1;
  const bool host = !(r.flags & 1u);
  Tuning tu = ctx.resolve(key.log_m, key.log_n);
  tu.pipeline = 1;
  HashPlan hp = hash_plan(key, tu);
  const uint32_t chunks = hp.variant == 1 ? serve_chunks(key.log_m, key.log_n, host) : 1;
  if (chunks > 1) {  // one word plane per split, and whole splits per chunk
    if (!hash_impl_reduces(hp.impl)) hp.impl = Tuning{}.hash_impl;
    if (hp.ks < chunks) hp.ks = chunks;
  }
  c.enc = DevBuf<uint64_t>(&ctx.pool, m * two_n);
  c.h = DevBuf<uint64_t>(&ctx.pool, 32 * two_n);
  c.tree.leaves = two_n;
  c.tree.nodes = DevBuf<uint32_t>(&ctx.pool, (size_t)c.tree.node_count() * 8);
  g.parts.emplace_back(&ctx.pool, (hash_partial_bytes(hp, two_n) + 7) / 8);
  CommitIO io;
  io.enc = c.enc.ptr;
  io.h = c.h.ptr;
  io.nodes = c.tree.nodes.ptr;
  io.part = g.parts.back().ptr;
  io.tw = ctx.twiddle(logN);
  io.tp = ctx.pass_twiddle(logN, tu.ntt_impl);
  if (!host) {
    c.f = r.f;
    io.f = c.f;
    enqueue_commit_stages(ctx, key, io, hp, tu, ctx.stream, false);
  } else {
    c.f_owned = DevBuf<uint64_t>(&ctx.pool, r.len);
    c.f = c.f_owned.ptr;
    io.f = c.f;
    const uint64_t rows = m / chunks;
    const uint32_t kc = hp.ks / chunks;
    for (uint32_t q = 0; q < chunks; q++) {
      const uint64_t off = q * rows * n;
      hip_check(hipMemcpyAsync(c.f_owned.ptr + off, r.f + off, rows * n * 8, hipMemcpyHostToDevice, upload_stream),
                "upload f");
      hipEvent_t e = event();
      g.events.push_back(e);
      hip_check(hipEventRecord(e, upload_stream), "uploaded");
      hip_check(hipStreamWaitEvent(ctx.stream, e, 0), "wait upload");
      if (chunks == 1) break;
      // this chunk's rows are encoded and hashed (its K-splits) while the next chunk uploads
      hip_check(launch_ntt(c.f + off, n, 1, (uint32_t)n, (uint32_t)rows, logN, io.tw, c.enc.ptr + q * rows * two_n,
                           ctx.stream, tu.ntt_impl, io.tp),
                "row encode");
      hip_check(launch_hash_mfma_splits(c.enc.ptr, two_n, (uint32_t)m, (uint32_t)two_n, key.dig.ptr, hp.ks, q * kc, kc,
                                        (int32_t*)io.part, ctx.stream, hp.impl, key.ring64),
                "hash (MFMA)");
    }
    if (chunks == 1) {
      enqueue_commit_stages(ctx, key, io, hp, tu, ctx.stream, false);
    } else {
      hip_check(launch_hash_mfma_finalize((const int32_t*)io.part, hp.ks, (uint32_t)two_n, c.h.ptr, ctx.stream,
                                          hp.impl, key.ring64),
                "hash finalize");
      hip_check(launch_merkle(c.h.ptr, 32, two_n, 1, two_n, c.tree.nodes.ptr, ctx.stream, tu.merkle_lg), "merkle");
    }
  }
  hip_check(hipMemcpyAsync(roots + 32 * (size_t)r.slot, c.tree.nodes.ptr + (c.tree.node_count() - 1) * 8, 32,
                           hipMemcpyDeviceToHost, ctx.stream),
            "read root");
}

// Admission and enqueueing; never waits for the device.
void Server::commit_loop() {
  (void)hipSetDevice(device);
  while (true) {
    std::unique_ptr<Group> g(new Group());
    {
      std::unique_lock<std::mutex> lk(mu);
      auto admissible = [&] {
        return !pending.empty() && reserved + pending.front()->footprint <= budget && in_flight < kRootSlots;
      };
      cv_commit.wait(lk, [&] { return stopping || admissible(); });
      if (stopping) {  // queued requests were cancelled by stop; admitted ones finish
        commit_running = false;
        cv_finish.notify_all();
        return;
      }
      while (admissible() && g->reqs.size() < max_batch) {
        auto r = pending.front();
        pending.pop_front();
        reserved += r->footprint;
        if (r->kind == kOpen) {  // no commit: to the provers once its commitment is done
          Resident& rs = resident.at(r->of);
          if (rs.cm) {
            r->cm = rs.cm;
            r->t_committed = now_ms();
            committed.push_back(r);
            cv_prove.notify_all();
          } else {
            rs.waiting.push_back(r);
          }
          continue;
        }
        in_flight++;
        r->slot = next_slot;
        next_slot = (next_slot + 1) % kRootSlots;
        g->reqs.push_back(r);
      }
      if (g->reqs.empty()) continue;
    }
    std::vector<std::shared_ptr<Request>> ok;
    for (auto& r : g->reqs) {
      try {
        enqueue_request(*g, *r);
        ok.push_back(r);
      } catch (const std::exception& e) {
        // work already enqueued for this request may still read its buffers, which stay with
        // it until the group's event; only this request fails
        const Error* le = dynamic_cast<const Error*>(&e);
        r->enqueue_code = le ? le->code : (dynamic_cast<const std::bad_alloc*>(&e) ? 4 : 1);
        r->err = e.what();
        ok.push_back(r);
      }
    }
    try {
      g->done = event();
      hip_check(hipEventRecord(g->done, ctxs[0]->stream), "committed");
    } catch (const std::exception& e) {
      g->rc = 3;
      g->msg = e.what();
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      g->reqs.swap(ok);
      finishing.push_back(std::move(g));
    }
    cv_finish.notify_all();
  }
}

// Waits for admitted groups in order and hands their commitments to the prove workers.
void Server::finish_loop() {
  (void)hipSetDevice(device);
  Context& ctx = *ctxs[0];
  while (true) {
    std::unique_ptr<Group> g;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_finish.wait(lk, [&] { return !finishing.empty() || !commit_running; });
      if (finishing.empty()) {  // the commit thread has exited and nothing is left
        finish_running = false;
        cv_prove.notify_all();
        return;
      }
      g = std::move(finishing.front());
      finishing.pop_front();
    }
    if (!g->rc) {
      const hipError_t e = hipEventSynchronize(g->done);
      if (e != hipSuccess) {
        g->rc = 3;
        g->msg = std::string("commit sync: ") + hipGetErrorString(e);
      }
    }
    if (g->rc) {
      (void)hipStreamSynchronize(upload_stream);
      ctx.drain();
    }
    g->parts.clear();  // the commits are done (or the streams drained)
    {
      std::lock_guard<std::mutex> lk(mu);
      for (hipEvent_t e : g->events) free_events.push_back(e);
      if (g->done) free_events.push_back(g->done);
      batches++;
      batched_requests += g->reqs.size();
      const double t = now_ms();
      for (auto& r : g->reqs) {
        in_flight--;
        if (g->rc || r->enqueue_code) {
          r->status = r->enqueue_code ? r->enqueue_code : g->rc;
          if (!r->enqueue_code) r->err = g->msg;
          r->cm.reset();
          reserved -= r->footprint;
          if (r->kind == kCommitOnly) {  // its waiting opens fail with it
            Resident& rs = resident.at(r->ticket);
            for (auto& o : rs.waiting) {
              o->status = r->status;
              o->err = "open of a failed commitment: " + r->err;
              reserved -= o->footprint;
              rs.opens--;
            }
            rs.waiting.clear();
            rs.failed = true;
            rs.footprint = 0;
            rs.released = true;
            drop_resident_locked(r->ticket);
          }
          continue;
        }
        std::memcpy(r->root, roots + 32 * (size_t)r->slot, 32);
        std::memcpy(r->cm->root, r->root, 32);
        r->t_committed = t;
        if (r->kind == kCommitOnly) {  // done: the commitment stays resident for its opens
          Resident& rs = resident.at(r->ticket);
          rs.cm = r->cm;
          rs.footprint = r->footprint;  // reserved until the commitment is released
          r->footprint = 0;
          r->cm.reset();
          r->t_done = t;
          r->status = 0;
          for (auto& o : rs.waiting) {
            o->cm = rs.cm;
            o->t_committed = t;
            committed.push_back(o);
          }
          rs.waiting.clear();
          drop_resident_locked(r->ticket);  // released before its commit was done
          continue;
        }
        committed.push_back(r);
      }
    }
    cv_prove.notify_all();
    cv_done.notify_all();
    cv_commit.notify_all();
  }
}

void Server::prove_loop(Context* ctx) {
  (void)hipSetDevice(device);
  while (true) {
    std::shared_ptr<Request> r;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_prove.wait(lk, [&] { return !committed.empty() || !finish_running; });
      if (committed.empty()) return;  // the finisher has exited and nothing is left
      r = committed.front();
      committed.pop_front();
    }
    int rc = 0;
    std::string msg;
    std::vector<uint8_t> proof;
    try {
      uint64_t used[2];
      prove_impl(*ctx, *r->key, *r->cm, r->z.data(), r->z.size() / 2, r->fields.data(), r->fields.size() / 2,
                 r->ints.data(), r->ints.size(), ProofOut{&proof}, used);
    } catch (const std::exception& e) {
      const Error* le = dynamic_cast<const Error*>(&e);
      rc = le ? le->code : (dynamic_cast<const std::bad_alloc*>(&e) ? 4 : 1);
      msg = e.what();
      ctx->drain();
    }
    if (r->kind == kOpen) std::memcpy(r->root, r->cm->root, 32);
    r->cm.reset();  // f, F', H and the tree go back to the commit context's pool (unless resident)
    std::lock_guard<std::mutex> lk(mu);
    r->proof.swap(proof);
    r->status = rc;
    r->err = msg;
    r->t_done = now_ms();
    reserved -= r->footprint;
    if (r->kind == kOpen) {
      auto it = resident.find(r->of);
      if (it != resident.end()) {
        it->second.opens--;
        drop_resident_locked(r->of);
      }
    }
    cv_done.notify_all();
    cv_commit.notify_all();
  }
}

Server* server_start(int device, uint64_t budget, uint32_t workers, uint32_t max_batch) {
  std::unique_ptr<Server> s(new Server());
  s->device = device;
  s->budget = budget;
  s->max_batch = max_batch ? max_batch : 16;
  if (workers == 0) workers = 2;
  hip_check(hipSetDevice(device), "set device");
  for (uint32_t i = 0; i <= workers; i++) {
    s->ctxs.push_back(make_context(device));
    s->ctxs.back()->single_stream_prove = true;
  }
  hip_check(hipStreamCreateWithFlags(&s->upload_stream, hipStreamNonBlocking), "stream");
  hip_check(hipHostMalloc((void**)&s->roots, 32 * (size_t)kRootSlots, hipHostMallocDefault), "root slots");
  s->threads.emplace_back(&Server::commit_loop, s.get());
  s->threads.emplace_back(&Server::finish_loop, s.get());
  for (uint32_t i = 1; i <= workers; i++) s->threads.emplace_back(&Server::prove_loop, s.get(), s->ctxs[i].get());
  return s.release();
}

uint64_t server_submit(Server* s, const Key* key, const uint64_t* f, uint64_t len, uint32_t flags,
                       const uint64_t* z, uint64_t z_len, const uint64_t* fields, uint64_t nf,
                       const uint64_t* ints, uint64_t ni) {
  if (!key->has_state) throw Error(7, "server: requests need a key with the A-bar commitment");
  if (len != (uint64_t(1) << (key->log_m + key->log_n))) throw Error(1, "server: polynomial length mismatch");
  if (z_len != key->log_m + key->log_n) throw Error(1, "server: z must have log_m + log_n entries");
  auto r = std::make_shared<Request>();
  r->key = key;
  r->f = f;
  r->len = len;
  r->flags = flags;
  r->z.assign(z, z + 2 * z_len);
  r->fields.assign(fields, fields + 2 * nf);
  r->ints.assign(ints, ints + ni);
  r->footprint = request_footprint(key->log_m, key->log_n, !(flags & 1u));
  if (r->footprint > s->budget) throw Error(4, "server: request needs more HBM than the whole budget");
  r->t_submit = now_ms();
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stopping) throw Error(7, "server is stopping");
  r->ticket = s->next_ticket++;
  s->all[r->ticket] = r;
  s->pending.push_back(r);
  s->cv_commit.notify_all();
  return r->ticket;
}

uint64_t server_submit_commit(Server* s, const Key* key, const uint64_t* f, uint64_t len, uint32_t flags) {
  if (len != (uint64_t(1) << (key->log_m + key->log_n))) throw Error(1, "server: polynomial length mismatch");
  auto r = std::make_shared<Request>();
  r->kind = kCommitOnly;
  r->key = key;
  r->f = f;
  r->len = len;
  r->flags = flags;
  // the whole commit footprint stays reserved while the commitment is resident
  r->footprint = commit_footprint(key->log_m, key->log_n, !(flags & 1u));
  if (r->footprint > s->budget) throw Error(4, "server: request needs more HBM than the whole budget");
  r->t_submit = now_ms();
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stopping) throw Error(7, "server is stopping");
  r->ticket = s->next_ticket++;
  s->all[r->ticket] = r;
  s->resident[r->ticket].key = key;  // its footprint moves to the entry when the commit is done
  s->pending.push_back(r);
  s->cv_commit.notify_all();
  return r->ticket;
}

uint64_t server_submit_open(Server* s, uint64_t commitment, const uint64_t* z, uint64_t z_len, const uint64_t* fields,
                            uint64_t nf, const uint64_t* ints, uint64_t ni) {
  auto r = std::make_shared<Request>();
  r->kind = kOpen;
  r->of = commitment;
  r->z.assign(z, z + 2 * z_len);
  r->fields.assign(fields, fields + 2 * nf);
  r->ints.assign(ints, ints + ni);
  r->t_submit = now_ms();
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stopping) throw Error(7, "server is stopping");
  auto it = s->resident.find(commitment);
  if (it == s->resident.end() || it->second.released)
    throw Error(1, "server: no resident commitment with this ticket (unknown, failed or released)");
  const Key* key = it->second.key;
  if (!key->has_state) throw Error(7, "server: opens need a key with the A-bar commitment");
  if (z_len != key->log_m + key->log_n) throw Error(1, "server: z must have log_m + log_n entries");
  r->key = key;
  r->footprint = open_footprint(key->log_m, key->log_n);
  if (r->footprint > s->budget) throw Error(4, "server: request needs more HBM than the whole budget");
  it->second.opens++;
  r->ticket = s->next_ticket++;
  s->all[r->ticket] = r;
  s->pending.push_back(r);
  s->cv_commit.notify_all();
  return r->ticket;
}

void server_release(Server* s, uint64_t commitment) {
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->resident.find(commitment);
  if (it == s->resident.end() || it->second.released)
    throw Error(1, "server: no resident commitment with this ticket (unknown or already released)");
  it->second.released = true;
  s->drop_resident_locked(commitment);
}

// 0: done, result copied out and released; 8: still running; 6: proof buffer too small
// (*proof_len = size needed, result kept); other codes are the request's own failure.
int server_poll(Server* s, uint64_t ticket, bool wait, uint8_t root[32], uint8_t* proof, uint64_t* proof_len,
                double times[3], std::string& err) {
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->all.find(ticket);
  if (it == s->all.end()) throw Error(1, "server: unknown ticket");
  auto r = it->second;
  if (wait) s->cv_done.wait(lk, [&] { return r->status >= 0; });
  if (r->status < 0) return 8;
  if (r->status != 0) {
    err = r->err;
    s->all.erase(it);
    return r->status;
  }
  const uint64_t cap = *proof_len;
  *proof_len = r->proof.size();
  if (!r->proof.empty() && (!proof || cap < r->proof.size())) return 6;
  std::memcpy(root, r->root, 32);
  if (!r->proof.empty()) std::memcpy(proof, r->proof.data(), r->proof.size());
  if (times) {
    times[0] = r->t_committed - r->t_submit;
    times[1] = r->t_done - r->t_committed;
    times[2] = r->t_done - r->t_submit;
  }
  s->all.erase(it);
  return 0;
}

void server_stop(Server* s) {
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stopping = true;
    for (auto& r : s->pending) {  // not admitted yet: cancelled
      r->status = 7;
      r->err = "server stopped";
    }
    s->pending.clear();
    s->cv_done.notify_all();
  }
  s->cv_commit.notify_all();
  s->cv_finish.notify_all();
  s->cv_prove.notify_all();
  for (auto& t : s->threads) t.join();
  for (auto& kv : s->all) kv.second->cm.reset();
  s->all.clear();
  s->resident.clear();
  if (s->upload_stream) {
    (void)hipStreamSynchronize(s->upload_stream);
    (void)hipStreamDestroy(s->upload_stream);
  }
  for (auto& c : s->ctxs) teardown_context(*c);
  for (hipEvent_t e : s->free_events) (void)hipEventDestroy(e);
  if (s->roots) (void)hipHostFree(s->roots);
  delete s;
}

void server_stats(Server* s, uint64_t out[4]) {
  std::lock_guard<std::mutex> lk(s->mu);
  out[0] = s->batches;
  out[1] = s->batched_requests;
  out[2] = s->reserved;
  out[3] = s->pending.size();
}

}  // namespace lgs
