// Serving loop: commit-then-open requests of mixed sizes, continuously batched.
//
// One commit worker admits queued requests in arrival order while their HBM footprint fits the
// budget, uploads their inputs on its own stream while the previous batch is still committing,
// commits every admitted request in one batched pass (grouped launches where the shape allows),
// and hands the resident commitments (f, F', H, Merkle tree) to P prove workers. Each worker
// proves on its own context and stream, so proofs of one batch run concurrently with each other
// and with the uploads and commits of the next batches. Memory is returned to the budget when a
// proof is done; requests that arrive meanwhile join the next batch as soon as they fit.
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
}  // namespace

// Conservative device footprint of one commit + open at (log_m, log_n), in bytes.
uint64_t request_footprint(uint32_t log_m, uint32_t log_n, bool host_input) {
  const uint64_t m = uint64_t(1) << log_m, n = uint64_t(1) << log_n, W = 64 * m, s = 128;
  const uint64_t two_n = 2 * n;
  uint64_t ks = m >= 1024 ? m / 512 : 1;
  uint64_t commit = (host_input ? 8 * m * n : 0) + 16 * m * n + 32 * two_n * 8 + 2 * two_n * 32 + ks * 256 * two_n * 4;
  uint64_t bi = W * s;
  uint64_t D = uint64_t(2) << batch_rounds(log_m, log_n);
  uint64_t prove = bi * (8 + 4 * 16) + 64 * D * 8 + 2 * D * 32 + 4 * D * 16 + 24 * n * 16 + W * 6 * 16 +
                   two_n * 64 + (uint64_t(1) << 24);
  return commit + prove;
}

struct Request {
  uint64_t ticket = 0;
  const Key* key = nullptr;
  const uint64_t* f = nullptr;
  uint64_t len = 0;
  uint32_t flags = 0;
  std::vector<uint64_t> z, fields, ints;
  uint64_t footprint = 0;
  double t_submit = 0, t_committed = 0, t_done = 0;
  std::unique_ptr<Commitment> cm;
  uint8_t root[32] = {0};
  std::vector<uint8_t> proof;
  int status = -1;  // -1 pending, else final status
  std::string err;
};

struct Server {
  int device = 0;
  uint64_t budget = 0, reserved = 0;
  uint32_t max_batch = 16;
  std::mutex mu;
  std::condition_variable cv_commit, cv_prove, cv_done;
  std::deque<std::shared_ptr<Request>> pending, committed;
  std::map<uint64_t, std::shared_ptr<Request>> all;
  uint64_t next_ticket = 1;
  bool stopping = false;
  bool commit_running = true;
  std::vector<std::unique_ptr<Context>> ctxs;  // [0] commit, [1..] prove
  std::vector<std::thread> threads;
  uint64_t batches = 0, batched_requests = 0;

  hipStream_t upload_stream = nullptr;  // host-to-device copies of request inputs

  struct Batch;
  std::unique_ptr<Batch> start_batch(std::vector<std::shared_ptr<Request>>& batch);
  void commit_batch(Batch& b);
  void finish_batch(Batch& b);
  void commit_loop();
  void prove_loop(Context* ctx);
};

// A batch between admission and hand-over: its inputs go to the device on the upload stream,
// its commits run on the commit context's stream after them.
struct Server::Batch {
  std::vector<std::shared_ptr<Request>> ok;
  GroupHold hold;
  hipEvent_t uploaded = nullptr, done = nullptr;
  int rc = 0;
  std::string msg;
  ~Batch() {
    if (uploaded) (void)hipEventDestroy(uploaded);
    if (done) (void)hipEventDestroy(done);
  }
};

// Uploads of one admitted batch: a commitment per request, f copied on the upload stream.
std::unique_ptr<Server::Batch> Server::start_batch(std::vector<std::shared_ptr<Request>>& batch) {
  Context& ctx = *ctxs[0];
  std::unique_ptr<Batch> b(new Batch());
  for (auto& r : batch) {
    try {
      r->cm.reset(new Commitment());
      Commitment& c = *r->cm;
      c.ctx = &ctx;
      c.key = r->key;
      if (r->flags & 1u) {
        c.f = r->f;
      } else {
        c.f_owned = DevBuf<uint64_t>(&ctx.pool, r->len);
        hip_check(hipMemcpyAsync(c.f_owned.ptr, r->f, r->len * 8, hipMemcpyHostToDevice, upload_stream),
                  "upload f");
        c.f = c.f_owned.ptr;
      }
      b->ok.push_back(r);
    } catch (const std::exception& e) {
      r->cm.reset();
      std::lock_guard<std::mutex> lk(mu);
      r->status = 4;
      r->err = e.what();
      reserved -= r->footprint;
      cv_done.notify_all();
    }
  }
  hip_check(hipEventCreateWithFlags(&b->uploaded, hipEventDisableTiming), "batch event");
  hip_check(hipEventCreateWithFlags(&b->done, hipEventDisableTiming), "batch event");
  hip_check(hipEventRecord(b->uploaded, upload_stream), "uploaded");
  return b;
}

// Commits of a batch whose uploads are enqueued: grouped launches where the shape allows, then
// the roots to the host.
void Server::commit_batch(Batch& b) {
  Context& ctx = *ctxs[0];
  try {
    hip_check(hipStreamWaitEvent(ctx.stream, b.uploaded, 0), "wait uploads");
    std::vector<std::pair<const Key*, Commitment*>> grouped, single;
    for (auto& r : b.ok) (group_eligible(*r->key) ? grouped : single).push_back({r->key, r->cm.get()});
    enqueue_commit_group(ctx, grouped, b.hold);
    for (auto& s : single) enqueue_commit(ctx, *s.first, *s.second);
    for (auto& r : b.ok) {
      const DevTree& t = r->cm->tree;
      hip_check(hipMemcpyAsync(r->root, t.nodes.ptr + (t.node_count() - 1) * 8, 32, hipMemcpyDeviceToHost,
                               ctx.stream),
                "read root");
    }
    hip_check(hipEventRecord(b.done, ctx.stream), "committed");
  } catch (const std::exception& e) {
    const Error* le = dynamic_cast<const Error*>(&e);
    b.rc = le ? le->code : (dynamic_cast<const std::bad_alloc*>(&e) ? 4 : 1);
    b.msg = e.what();
  }
}

// Waits for a batch's commits and hands its commitments to the prove workers.
void Server::finish_batch(Batch& b) {
  Context& ctx = *ctxs[0];
  if (!b.rc) {
    const hipError_t e = hipEventSynchronize(b.done);
    if (e != hipSuccess) {
      b.rc = 3;
      b.msg = std::string("commit sync: ") + hipGetErrorString(e);
    }
  }
  if (b.rc) {
    (void)hipStreamSynchronize(upload_stream);
    ctx.drain();
  }
  std::lock_guard<std::mutex> lk(mu);
  batches++;
  batched_requests += b.ok.size();
  const double t = now_ms();
  for (auto& r : b.ok) {
    if (b.rc) {
      r->status = b.rc;
      r->err = b.msg;
      r->cm.reset();
      reserved -= r->footprint;
      continue;
    }
    std::memcpy(r->cm->root, r->root, 32);
    r->t_committed = t;
    committed.push_back(r);
  }
  cv_prove.notify_all();
  cv_done.notify_all();
  cv_commit.notify_all();
}

// Admission and commits, pipelined: the uploads of a newly admitted batch are enqueued before the
// batch in flight is waited for, so they overlap its commit kernels (and the proofs running on
// the workers' streams); its commits follow once the batch in flight is handed over.
void Server::commit_loop() {
  (void)hipSetDevice(device);
  std::unique_ptr<Batch> inflight;
  while (true) {
    std::vector<std::shared_ptr<Request>> batch;
    {
      std::unique_lock<std::mutex> lk(mu);
      auto admissible = [&] { return !pending.empty() && reserved + pending.front()->footprint <= budget; };
      if (!inflight) cv_commit.wait(lk, [&] { return stopping || admissible(); });
      if (stopping && !inflight) {  // queued requests were cancelled by stop; admitted ones finish
        commit_running = false;
        cv_prove.notify_all();
        return;
      }
      while (!stopping && admissible() && batch.size() < max_batch) {
        reserved += pending.front()->footprint;
        batch.push_back(pending.front());
        pending.pop_front();
      }
    }
    if (batch.empty()) {  // nothing more fits now: hand over the batch in flight
      finish_batch(*inflight);
      inflight.reset();
      continue;
    }
    std::unique_ptr<Batch> next;
    try {
      next = start_batch(batch);
    } catch (const std::exception& e) {  // event creation failed: fail the batch's requests
      std::lock_guard<std::mutex> lk(mu);
      for (auto& r : batch) {
        if (r->status >= 0) continue;
        r->cm.reset();
        r->status = 3;
        r->err = e.what();
        reserved -= r->footprint;
      }
      cv_done.notify_all();
      continue;
    }
    if (inflight) {
      finish_batch(*inflight);
      inflight.reset();
    }
    commit_batch(*next);
    inflight = std::move(next);
  }
}

void Server::prove_loop(Context* ctx) {
  (void)hipSetDevice(device);
  while (true) {
    std::shared_ptr<Request> r;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_prove.wait(lk, [&] { return !committed.empty() || !commit_running; });
      if (committed.empty()) return;  // the commit worker has exited and nothing is left
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
    r->cm.reset();  // f, F', H and the tree go back to the commit context's pool
    std::lock_guard<std::mutex> lk(mu);
    r->proof.swap(proof);
    r->status = rc;
    r->err = msg;
    r->t_done = now_ms();
    reserved -= r->footprint;
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
  for (uint32_t i = 0; i <= workers; i++) s->ctxs.push_back(make_context(device));
  hip_check(hipStreamCreateWithFlags(&s->upload_stream, hipStreamNonBlocking), "stream");
  s->threads.emplace_back(&Server::commit_loop, s.get());
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
  if (!proof || cap < r->proof.size()) return 6;
  std::memcpy(root, r->root, 32);
  std::memcpy(proof, r->proof.data(), r->proof.size());
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
  s->cv_prove.notify_all();
  for (auto& t : s->threads) t.join();
  for (auto& kv : s->all) kv.second->cm.reset();
  s->all.clear();
  if (s->upload_stream) {
    (void)hipStreamSynchronize(s->upload_stream);
    (void)hipStreamDestroy(s->upload_stream);
  }
  for (auto& c : s->ctxs) teardown_context(*c);
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
