// lgs_commit_noalloc: the commit of lgs_commit into caller-owned device buffers, with every
// scratch buffer (the hash's K-split partials, the row transform's twiddle tables) in one
// file-scope device array. It allocates nothing, for callers that cannot call hipMalloc (code
// under HIP graph capture, serving loops that forbid allocation). No counterpart in the
// reference, whose commit (pcs/src/ligesis.rs:277-325) allocates its outputs itself.
//
// The array is a host-side symbol in host code: its device address is taken once per device with
// hipGetSymbolAddress (lgs_init calls scratch_init with the device current) and only that
// address is handed to kernels. Every call on one device shares the array; calls that overlap
// on one device race (the second overwrites the first's partials and tables).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/ligesis_hip.h"
#include "handles.hpp"

namespace lgs {

static constexpr uint64_t kScratchBytes = LGS_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static uint8_t* g_scratch_ptr[kMaxDevices] = {};

// With `device` current: the device address of this device's copy of the scratch array.
void scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) throw Error(LGS_ERR_DEVICE, "scratch: device index out of range");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || p == nullptr)
    throw Error(LGS_ERR_DEVICE, std::string("scratch: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
  g_scratch_ptr[device] = (uint8_t*)p;
}

static uint8_t* scratch_ptr(int device) {
  std::lock_guard<std::mutex> g(g_scratch_mu);
  return device >= 0 && device < kMaxDevices ? g_scratch_ptr[device] : nullptr;
}

static uint64_t align256(uint64_t b) { return (b + 255) & ~uint64_t(255); }

// Scratch layout of one commit: [twiddles | pass twiddles | K-split partials], each 256-aligned.
struct NoallocLayout {
  uint64_t tw_off = 0, tp_off = 0, part_off = 0, total = 0;
};

static NoallocLayout layout(int logN, bool pass_table, int impl, const HashPlan& hp, uint64_t two_n) {
  NoallocLayout L;
  const uint64_t tw_bytes = (uint64_t(1) << (logN - 1)) * 8;
  L.tw_off = 0;
  L.tp_off = align256(tw_bytes);
  const uint64_t tp_bytes = pass_table ? (uint64_t)pass_twiddle_entries(logN, impl) * 8 : 0;
  L.part_off = align256(L.tp_off + tp_bytes);
  L.total = L.part_off + hash_partial_bytes(hp, two_n);
  return L;
}

}  // namespace lgs

extern "C" int lgs_commit_noalloc(lgs_ctx* ctx, const lgs_key* key, const uint64_t* f, uint64_t len,
                                  uint64_t* encoded, uint64_t* digests, uint8_t* tree, void* stream,
                                  uint8_t root[32]) {
  if (!ctx || !key || !f || !encoded || !digests || !tree) return LGS_ERR_ARG;
  lgs::Context& c = ctx->c;
  const lgs::Key& k = key->k;
  try {
    const uint64_t need = uint64_t(1) << (k.log_m + k.log_n);
    if (len != need)
      throw lgs::Error(LGS_ERR_ARG, "commit_noalloc: expected 2^(log_m+log_n) = " + std::to_string(need) +
                                        " evaluations, got " + std::to_string(len));
    if (k.ctx != &c) throw lgs::Error(LGS_ERR_ARG, "commit_noalloc: the key belongs to another context");
    uint8_t* base = lgs::scratch_ptr(c.device);
    if (!base) throw lgs::Error(LGS_ERR_STATE, "commit_noalloc: scratch not set up on this device");
    (void)hipSetDevice(c.device);
    const uint64_t two_n = uint64_t(2) << k.log_n;
    const int logN = (int)k.log_n + 1;
    if (logN > 20) throw lgs::Error(LGS_ERR_ARG, "commit_noalloc: rows longer than 2^19 values need more scratch");
    hipStream_t st = stream ? (hipStream_t)stream : c.stream;
    lgs::Tuning tu = c.resolve(k.log_m, k.log_n);
    tu.pipeline = 1;  // one stream, no second stream to create
    lgs::HashPlan hp = lgs::hash_plan(k, tu);
    if (hp.variant == 1 && !lgs::hash_impl_reduces(hp.impl)) hp.impl = lgs::Tuning{}.hash_impl;  // one word per partial
    const bool pass_table = lgs::pass_twiddle_entries(logN, tu.ntt_impl) != 0;
    lgs::NoallocLayout L = layout(logN, pass_table, tu.ntt_impl, hp, two_n);
    // fewer K-splits until the partials fit (the split count never changes the result)
    while (L.total > lgs::kScratchBytes && hp.ks > 1) {
      hp.ks /= 2;
      L = layout(logN, pass_table, tu.ntt_impl, hp, two_n);
    }
    if (L.total > lgs::kScratchBytes)
      throw lgs::Error(LGS_ERR_ARG, "commit_noalloc: shape needs " + std::to_string(L.total) +
                                        " scratch bytes, above LGS_NOALLOC_SCRATCH_BYTES = " +
                                        std::to_string(lgs::kScratchBytes));
    lgs::CommitIO io;
    io.f = f;
    io.enc = encoded;
    io.h = digests;
    io.nodes = (uint32_t*)tree;
    io.tw = (const uint64_t*)(base + L.tw_off);
    io.tp = pass_table ? (const uint64_t*)(base + L.tp_off) : nullptr;
    io.part = (uint64_t*)(base + L.part_off);
    // the tables are rebuilt on every call: another call on this device may have overwritten them
    lgs::hip_check(lgs::launch_twiddles((uint64_t*)io.tw, logN, st), "twiddles");
    if (pass_table)
      lgs::hip_check(lgs::launch_pass_twiddles((uint64_t*)io.tp, io.tw, logN, tu.ntt_impl, st), "pass twiddles");
    lgs::enqueue_commit_stages(c, k, io, hp, tu, st, false);
    if (root) {
      lgs::hip_check(hipMemcpyAsync(root, tree + (2 * two_n - 2) * 32, 32, hipMemcpyDeviceToHost, st), "read root");
      lgs::hip_check(hipStreamSynchronize(st), "commit sync");
    }
    return LGS_OK;
  } catch (const lgs::Error& e) {
    c.err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    c.err = e.what();
    return LGS_ERR_ARG;
  }
}
