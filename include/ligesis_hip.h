/*
 * ligesis_hip.h - C ABI of the MI355X LigeSIS commit / eval-prove library
 * (ligesis_amd/libligesis_hip.so).
 *
 * Drop-in boundary for the reference's LigeSIS API (pcs/src/ligesis.rs:205-756):
 *   Ligesis::setup(log_m, log_n, rng)      ligesis.rs:239  -> lgs_setup   (A passed in, not sampled)
 *   Ligesis::commit(&pk, MlPoly f)         ligesis.rs:277  -> lgs_commit, lgs_commit_batch
 *   Ligesis::prove(&pk, &data, z, oracle)  ligesis.rs:347  -> lgs_prove   (oracle streams passed in)
 *   CommitData (dropped by the caller)     ligesis.rs:88   -> lgs_commitment_release
 *   ProvingKey (dropped by the caller)     ligesis.rs:67   -> lgs_key_release
 * Ligesis::verify (ligesis.rs:569) stays on the CPU with the caller.
 *
 * Conventions
 *  - Every function returns an int status (LGS_OK = 0) except the release functions,
 *    lgs_last_error and lgs_shutdown. Shape and length violations, which the reference turns
 *    into assert! panics (ligesis.rs:248, 284, 343, 362), return LGS_ERR_SHAPE / LGS_ERR_ARG;
 *    the message is in lgs_last_error(ctx).
 *  - Field elements are canonical little-endian uint64 (< p = 2^64 - 2^32 + 1); an extension
 *    element of F_p[X]/(X^2 - 7) is two consecutive uint64 (c0, c1).
 *  - Buffers are caller-owned host memory unless a flag says device memory.
 *  - One context per host thread; a context owns one HIP stream on one device.
 */
#ifndef LIGESIS_HIP_H
#define LIGESIS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  LGS_OK = 0,
  LGS_ERR_ARG = 1,    /* bad argument or length (reference: assert_eq! panic) */
  LGS_ERR_SHAPE = 2,  /* (log_m, log_n) outside the shared-domain window (ligesis.rs:248-251) */
  LGS_ERR_DEVICE = 3, /* HIP runtime error */
  LGS_ERR_NOMEM = 4,  /* device allocation failed or the HBM budget is exhausted */
  LGS_ERR_ORACLE = 5, /* the randomness-oracle stream ran out (reference: index panic) */
  LGS_ERR_BUFFER = 6, /* output buffer too small; the required size is reported */
  LGS_ERR_STATE = 7,  /* handle cannot do this (e.g. proving with a commit-only key) */
  LGS_PENDING = 8     /* lgs_server_poll: the request is still running */
};

/* lgs_setup flags */
#define LGS_SETUP_COMMIT_ONLY 1u /* build the hash only: no A-bar commitment, no shape window */
/* The Z_2^64 subset-sum hash (a paper-level variant the reference leaves out, subset_sum.rs:10-12):
 * H = A . B with sums mod 2^64 and A any 64-bit words; its commitments are a different scheme
 * (wire record version 2). Implies LGS_SETUP_COMMIT_ONLY; needs m >= 2 and 2n >= 128. */
#define LGS_SETUP_RING64 2u
/* lgs_commit flags */
#define LGS_INPUT_DEVICE 1u /* f is a device pointer; the commitment borrows it until release */

/* lgs_commitment_read selectors */
#define LGS_READ_ENCODED 0u /* F' = RS rows, m x 2n uint64 */
#define LGS_READ_DIGESTS 1u /* H, 32 x 2n uint64 row-major (h_matrix[r][j], ligesis.rs:302) */
#define LGS_READ_TREE 2u    /* all Merkle layers, (4n - 1) x 32 bytes, leaves first */

typedef struct lgs_ctx lgs_ctx;
typedef struct lgs_key lgs_key;
typedef struct lgs_commitment lgs_commitment;

/* Context on HIP device `device` (one stream, default priority). */
int lgs_init(int device, lgs_ctx** out);
void lgs_shutdown(lgs_ctx* ctx);
const char* lgs_last_error(const lgs_ctx* ctx);

/* Tuning knobs that never change results: "hash" (0 auto, 1 MFMA, 2 VALU), "hash_ksplit"
 * (0 auto), "hash_impl" (-1 auto; MFMA kernel: 0 pipelined, 1 pipelined one slab per stage,
 * 2 basic, 3 pipelined with 256 x 256 workgroup tiles, 4 global loads issued 2 slabs ahead, 5 as
 * 4 with 256 x 256 tiles, 6 / 7 64 x 64 wave tiles in 256 x 128 / 256 x 256 workgroup tiles,
 * 8 / 9 64 x 128 wave tiles in 256 x 128 / 256 x 256 workgroup tiles, 10 as 0 with each A
 * fragment read from LDS one product group ahead, 11 / 12 A digits and F' words copied straight
 * into an LDS ring of 3 / 4 slabs, 14 as 12 with 8 waves and 256 x 256 tiles sharing one ring,
 * 15 / 16 as 14 with a 4 / 8-slab ring, each slab's bit expansion, A reads and copies
 * interleaved with the previous slab's products, 17 as 16 with one barrier per two slabs;
 * shapes a kernel does not fit take the next one down),
 * "ntt_impl" (-1 auto; 0/2/3 radix-8 passes with 256/512/1024 threads per row, 1 radix-2
 * stages, 4 / 5 for rows of 2^12 values a 2^13-point radix-8 / radix-16 kernel with every size
 * fixed at compile time and per-pass twiddle tables, elsewhere as 3), "merkle_lg" (0 auto; 6..8 = log2 nodes per Merkle workgroup = tree levels per
 * launch), "pipeline" (0 auto; 1 off; 2 / 4 / 8: the rows are encoded in that many chunks and
 * each chunk's K-splits are hashed on a second stream while the next chunk is encoded),
 * "prove_graph" (0 default; 1: the second proof of a commitment on its own context captures
 * its launches into a HIP graph, kept with the commitment, and later proofs replay it with their
 * own constants), "profile" (1: time every stage with HIP events; pipelined stages report the
 * summed durations of their launches; proofs are then issued without a graph; 2: as 1, except
 * that a plain commit times only its hash kernel and reports stage [1] alone: the MFMA kernel
 * from its dispatch's own start and stop times, the VALU hash from events around it). An option left
 * at auto takes the autotuned schedule of the commit's shape when the process tuning table has
 * one (lgs_autotune, lgs_tune_load), else a built-in default. */
int lgs_set_option(lgs_ctx* ctx, const char* name, int64_t value);

/* Stage times of the last profiled call, in milliseconds. Commit: [0] row NTT, [1] hash main
 * kernel, [2] hash finalize (no kernel between its two events where the K-split sum runs inside
 * the tree's first launch: the default kernels), [3] Merkle tree, [4] whole commit. Prove: [5]
 * host time to issue the work, [6] host time to the proof's download; device timestamps from
 * the first upload:
 * [7] first sumcheck round starts, [8] last sumcheck launch ends, [9] the Basefold open on the
 * second stream ends, [10] openings gathered, [11] proof downloaded, and on the second stream
 * [12] its work starts, [13] B_I's commitment done, [14] a-bar's commitment done, [15] the
 * batched codeword combined, [16] the single-workgroup rounds start; [17 + k] the host time
 * (from the call's start) at which the work before device timestamp [7 + k] had been issued
 * ([17] for the first upload). Up to `n` (<= 28) values are written. */
int lgs_stage_times(const lgs_ctx* ctx, double* out_ms, uint32_t n);

/* Commit launch schedule: {hash_variant, hash_impl, hash_ksplit, ntt_impl, merkle_lg, pipeline}
 * (the option names above). No counterpart in the reference, whose schedule is fixed at compile
 * time (pcs/src/ligesis.rs:53-59 holds its only constants); results never depend on it. */
#define LGS_TUNE_FIELDS 6

/* Autotuner: times candidate schedules for the key's shape on a synthetic polynomial (each stage
 * separately: row-transform variant, hash kernel variant and K-split, Merkle workgroup width,
 * then transform / hash overlap on the whole commit;
 * `reps` timed commits per candidate, 0 = 5) and stores the fastest in the process-wide table,
 * which every context (server workers included) consults. Writes the chosen schedule to `cfg`
 * and its measured commit time to `ms` (either may be NULL). */
int lgs_autotune(lgs_ctx* ctx, const lgs_key* key, uint32_t reps, int32_t cfg[LGS_TUNE_FIELDS], double* ms);

/* The schedule a commit of shape 2^log_m x 2^log_n on this context would use now. */
int lgs_tune_get(lgs_ctx* ctx, uint32_t log_m, uint32_t log_n, int32_t cfg[LGS_TUNE_FIELDS]);

/* Process tuning table as text, one shape per line:
 * "log_m log_n hash_variant hash_impl hash_ksplit ntt_impl merkle_lg pipeline ms", '#' starts a
 * comment.
 * Load merges into the table and rejects a malformed file as a whole (LGS_ERR_ARG). `count`
 * (may be NULL) receives the number of entries written / read. Errors are reported on `ctx`,
 * which may be NULL (the table needs no device). */
int lgs_tune_save(lgs_ctx* ctx, const char* path, uint64_t* count);
int lgs_tune_load(lgs_ctx* ctx, const char* path, uint64_t* count);
void lgs_tune_clear(void);

/* Setup (ligesis.rs:239): A is the subset-sum matrix, 32 rows x (64 << log_m) uint64,
 * row-major (subset_sum.rs:23-27), canonical (< p; LGS_ERR_ARG otherwise) unless LGS_SETUP_RING64.
 * Unless LGS_SETUP_COMMIT_ONLY, the shape must pass the reference's shared-domain window and
 * A-bar is Basefold-committed on the GPU. */
int lgs_setup(lgs_ctx* ctx, uint32_t log_m, uint32_t log_n, const uint64_t* a, uint32_t flags,
              lgs_key** out);
/* com_a of the key (ligesis.rs:71); LGS_ERR_STATE for a commit-only key. */
int lgs_key_com_a(const lgs_key* key, uint8_t com_a[32]);

#define LGS_BASEFOLD_EXT 2u /* Basefold commitments: the polynomial is (c0, c1) extension pairs */
/* Standalone Basefold commitment root (Basefold::commit_{base,ext}_on_domain,
 * pcs/src/basefold.rs:341-382; commit_base / commit_ext are log_interleave 6, :309-339): the 2^li
 * chunks f[j :: 2^li] are zero-padded 2^code_rate-fold and transformed, and leaf p hashes column
 * p of the codewords into a BLAKE3 Merkle tree. poly holds 2^num_vars canonical base elements,
 * or pairs with LGS_BASEFOLD_EXT; a device pointer with LGS_INPUT_DEVICE. Limits: num_vars <= 30,
 * log_interleave <= min(6, num_vars), 1 <= code_rate <= 4. Returns the root only; the prover state
 * is freed (lgs_basefold_commit_state keeps it). */
int lgs_basefold_commit(lgs_ctx* ctx, const uint64_t* poly, uint32_t num_vars, uint32_t flags,
                        uint32_t log_interleave, uint32_t code_rate, uint8_t root[32]);

/* The same commitment keeping the prover state (BasefoldProverState, basefold.rs:36-54) in HBM:
 * the interleaved codewords, their Merkle tree and the polynomial (copied from host memory, or
 * borrowed until release with LGS_INPUT_DEVICE). */
typedef struct lgs_basefold lgs_basefold;
int lgs_basefold_commit_state(lgs_ctx* ctx, const uint64_t* poly, uint32_t num_vars, uint32_t flags,
                              uint32_t log_interleave, uint32_t code_rate, uint8_t root[32], lgs_basefold** out);
/* Basefold::prove (basefold.rs:384-516): opens the committed polynomial at `point` (point_len =
 * num_vars extension elements, 2 * point_len uint64). The RandomOracle is its two pre-drawn
 * streams, read from the start: one field per reduction round (num_vars - log_interleave), then
 * query_num(code_rate) = ceil(100 / code_rate) integers. The proof is written in the canonical
 * encoding (field order of BasefoldProof, basefold.rs:84-103) as for lgs_prove; *proof_len and
 * used[] as there. Verification (basefold.rs:518-642) stays on the CPU with the caller. */
int lgs_basefold_prove(lgs_ctx* ctx, const lgs_basefold* st, const uint64_t* point, uint64_t point_len,
                       const uint64_t* fields, uint64_t n_fields, const uint64_t* ints, uint64_t n_ints,
                       uint8_t* proof, uint64_t* proof_len, uint64_t used[2]);
void lgs_basefold_release(lgs_basefold* st);
void lgs_key_release(lgs_key* key);

/* Commit (ligesis.rs:277): f = 2^(log_m + log_n) canonical evaluations, F[i][j] = f[i*n + j].
 * Writes the 32-byte root and returns a handle that keeps f, F', H and the Merkle tree in HBM
 * for lgs_prove. */
int lgs_commit(lgs_ctx* ctx, const lgs_key* key, const uint64_t* f, uint64_t len, uint32_t flags,
               uint8_t root[32], lgs_commitment** out);
/* Commit `count` polynomials (each with its own key and length) in one batched pass; roots is
 * count x 32 bytes. Results are identical to `count` lgs_commit calls. */
int lgs_commit_batch(lgs_ctx* ctx, uint32_t count, const lgs_key* const* keys,
                     const uint64_t* const* fs, const uint64_t* lens, uint32_t flags,
                     uint8_t* roots, lgs_commitment** out);
void lgs_commitment_release(lgs_commitment* c);

/* Commit without allocating: no counterpart in the reference (its commit, ligesis.rs:277-325,
 * allocates its outputs); for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops that forbid hipMalloc). The commit of lgs_commit, with f and every
 * output in caller-owned device memory: f = 2^(log_m + log_n) values, encoded = F' (m x 2n
 * uint64), digests = H (32 x 2n uint64), tree = all Merkle layers ((4n - 1) x 32 bytes, leaves
 * first, the root last). Scratch (hash K-split partials, twiddle tables) lives in one static
 * device array of LGS_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; shapes that
 * need more (rows of more than 2^19 values, or 32 x 2n x 8 bytes of partials above the limit)
 * return LGS_ERR_ARG. The work is enqueued on `stream` (a hipStream_t; NULL: the context's
 * stream); with `root` non-NULL the call also copies the root to host memory and waits for that
 * stream, with `root` NULL it returns at once (the root is then the last 32 bytes of `tree` once
 * the stream reaches it). Every call on one GPU shares
 * that GPU's scratch: calls that overlap on one GPU (two contexts, or two host threads) race.
 * Never calls hipMalloc or hipFree. The key must come from the same context. */
#define LGS_NOALLOC_SCRATCH_BYTES (40ull << 20)
int lgs_commit_noalloc(lgs_ctx* ctx, const lgs_key* key, const uint64_t* f, uint64_t len, uint64_t* encoded,
                       uint64_t* digests, uint8_t* tree, void* stream, uint8_t root[32]);

/* Copy a resident stage (LGS_READ_*) to host memory; cap is in bytes. */
int lgs_commitment_read(const lgs_commitment* c, uint32_t what, void* out, uint64_t cap);

/* Eval-prove (ligesis.rs:347): proves f(z) = y for the committed f. z holds z_len =
 * log_m + log_n extension elements (2 * z_len uint64). The randomness oracle
 * (utils/src/oracle.rs:10-64) is passed as its two pre-drawn streams: `fields` holds n_fields
 * extension elements, `ints` n_ints uint64; both are read from the start, as after
 * RandomOracle::restart. The proof is written in the canonical encoding (field order of
 * LigesisProof, ligesis.rs:99-119; u64 little-endian lengths, canonical u64 field words) into
 * `proof` when *proof_len (in: capacity in bytes) suffices; *proof_len is set to the size either
 * way (LGS_ERR_BUFFER if it did not fit; the buffer's contents are then unspecified). used[0], used[1] receive the numbers of field and
 * integer draws, as RandomOracle::drawn() (oracle.rs:57-59). */
int lgs_prove(lgs_ctx* ctx, const lgs_key* key, const lgs_commitment* c, const uint64_t* z,
              uint64_t z_len, const uint64_t* fields, uint64_t n_fields, const uint64_t* ints,
              uint64_t n_ints, uint8_t* proof, uint64_t* proof_len, uint64_t used[2]);

/* Several eval-proofs on one randomness oracle (PolyCommitmentScheme::batch_prove, noop.rs:36-40,
 * over LigeSIS): proof i is lgs_prove of commitment cs[i] (made with keys[i]) at zs[i]
 * (z_lens[i] extension elements), reading the two streams from where proof i - 1's draws ended,
 * exactly as `count` lgs_prove calls on one RandomOracle continuing its cursors. The proofs'
 * device work overlaps on worker contexts owned by ctx (each with its own streams). proofs[i]
 * receives proof i when proof_lens[i] (in: capacity) suffices; every proof_lens[i] is set to its
 * size (LGS_ERR_BUFFER if any did not fit). used receives the total draws. */
int lgs_prove_batch(lgs_ctx* ctx, uint32_t count, const lgs_key* const* keys, const lgs_commitment* const* cs,
                    const uint64_t* const* zs, const uint64_t* z_lens, const uint64_t* fields, uint64_t n_fields,
                    const uint64_t* ints, uint64_t n_ints, uint8_t* const* proofs, uint64_t* proof_lens,
                    uint64_t used[2]);

/* ---- Serving loop (no counterpart in the single-threaded reference; it drives the calls
 * above). Requests are commit-then-open jobs (lgs_commit followed by lgs_prove on the resident
 * commitment) of any shape. One commit worker admits queued requests in arrival order while
 * their estimated HBM footprint fits `hbm_budget` bytes and commits each admitted group in one
 * batched pass; `prove_workers` threads prove on their own streams, so proofs overlap each
 * other and the next commit. Results are identical to the unserved calls. */
typedef struct lgs_server lgs_server;
int lgs_server_start(int device, uint64_t hbm_budget, uint32_t prove_workers, uint32_t max_batch,
                     lgs_server** out);
/* Queue a request. f (host memory, or device memory with LGS_INPUT_DEVICE) must stay valid
 * until the request completes; z / fields / ints are copied. The key must not be a commit-only
 * key. LGS_ERR_NOMEM when the request alone exceeds the budget. */
int lgs_server_submit(lgs_server* srv, const lgs_key* key, const uint64_t* f, uint64_t len, uint32_t flags,
                      const uint64_t* z, uint64_t z_len, const uint64_t* fields, uint64_t n_fields,
                      const uint64_t* ints, uint64_t n_ints, uint64_t* ticket);
/* Commit-only request: f is committed as by lgs_server_submit, and the commitment (f, F', H and
 * the tree) then stays resident in HBM, its commit footprint reserved, until lgs_server_release
 * of this ticket. Polling the ticket returns its root (proof_len 0). The key may be commit-only
 * (LGS_SETUP_COMMIT_ONLY) if no open of it follows. */
int lgs_server_submit_commit(lgs_server* srv, const lgs_key* key, const uint64_t* f, uint64_t len, uint32_t flags,
                             uint64_t* ticket);
/* Open request (Ligesis::prove, ligesis.rs:347) of the resident commitment of commit-only ticket
 * `commitment`, at z (z_len extension elements) with the given oracle streams (copied), queued
 * and admitted like other requests; it waits for the commitment's commit if that is not done.
 * Polling returns the proof (and the commitment's root). Opens of one commitment may run at once.
 * LGS_ERR_ARG for an unknown, failed or released commitment. */
int lgs_server_submit_open(lgs_server* srv, uint64_t commitment, const uint64_t* z, uint64_t z_len,
                           const uint64_t* fields, uint64_t n_fields, const uint64_t* ints, uint64_t n_ints,
                           uint64_t* ticket);
/* Releases a resident commitment: its memory returns to the budget once no open of it runs.
 * Opens submitted earlier still complete; later ones are refused. */
int lgs_server_release(lgs_server* srv, uint64_t commitment);
/* Result of a request: LGS_PENDING while running (wait = 0), LGS_ERR_BUFFER with *proof_len set
 * when `proof` is too small (the result is kept), LGS_OK with root, proof and
 * times_ms = {submit->committed, committed->proved, submit->proved} (then the ticket is
 * retired), or the request's own error code. */
int lgs_server_poll(lgs_server* srv, uint64_t ticket, int wait, uint8_t root[32], uint8_t* proof,
                    uint64_t* proof_len, double times_ms[3]);
/* {batches committed, requests committed, HBM bytes reserved now, requests queued}. */
int lgs_server_stats(lgs_server* srv, uint64_t out[4]);
/* Cancels queued requests, finishes admitted ones, joins the workers, frees everything. */
void lgs_server_stop(lgs_server* srv);
const char* lgs_server_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* LIGESIS_HIP_H */
