// Launchers of the eval-prove kernels (prove.hip). All enqueue on `st` without synchronising.
#pragma once
#include "kernels.hpp"

namespace lgs {

// Deferred reduction of per-block partial sums: out[q] = sum_{b < nblk} part[b * w + q] for
// q < w <= 16. Kernels that produce partials leave them in memory; a block of a later launch on
// the same stream adds them up, after the producing launch has completed, so no cross-workgroup
// fence or completion counter is needed (an agent-scope fence writes back and invalidates the
// L2 of the workgroup's XCD).
struct RedJob {
  const ef2* part;
  ef2* out;
  uint32_t nblk, w;
};
// Up to four reductions in one small launch (one block each).
struct RedJobs {
  RedJob j[4];
  int n;
};
hipError_t launch_reduce_jobs(const RedJobs& r, hipStream_t st);

struct EqTable;

// One round of one sumcheck inside a grouped launch (sc_round_group_kernel): blocks
// [first_block, first_block + nblk) of the launch. p[2] is unused when d == 2.
struct ScRoundJob {
  ef2* p[3];
  uint64_t half;
  const ef2* r;
  ef2* part;  // nblk x (d + 1) partial sums, reduced into the round's d + 1 values later
  // bits: round tau (< 4) of a sumcheck whose first table is 0/1 valued, read from these bytes
  // (2^(tau + 1) half entries, before any fold), and whose last table is an eq table evaluated
  // from the factors eqt (for round tau: its high factor already carries the fold of the tau
  // bound variables; 2^klo >= 256 kpt), never read from memory. After tau rounds the first table
  // at x is sum_y W(y) bits[y 2^(nv - tau) + x] over the tau-bit prefixes y, W(y) the products
  // of c_t or 1 - c_t; wts[m] is the sum of W(y) over the set bits y of m (m < 2^(2^tau)). When p[0] is set the round also folds both tables into p[0]
  // and p[d - 1] for the rounds that read tables. d = 3: the binary check (one eq term); d = 2:
  // the batched opening (up to three). Null bits: an ordinary round.
  const uint8_t* bits;
  const EqTable* eqt;
  const ef2* wts;
  // m1: d = 3 and the second table is the first minus one (not stored: p[1] unused)
  // kpt: pairs per thread of a bits round (its blocks cover 256 kpt consecutive pairs)
  uint32_t d, first_block, nblk, m1, tau, kpt;
};
// One grouped round step, passed by value as the kernel's argument (no upload): jobs j[0..nj)
// own blocks [0, round_blocks); the blocks after them run the reductions r[0..nr) (one each).
constexpr int kScMaxJobs = 7, kScMaxReds = 8;
struct ScStep {
  ScRoundJob j[kScMaxJobs];
  RedJob r[kScMaxReds];
  uint32_t nj, nr, round_blocks, pad;
};
hipError_t launch_sc_round_group(const ScStep& step, hipStream_t st);  // no bits jobs
// A step of bits jobs only, all at round tau.
hipError_t launch_sc_bits_group(const ScStep& step, uint32_t tau, hipStream_t st);

// The remaining rounds (len <= 2^kScTailLog entries) of one sumcheck, one workgroup per job.
constexpr int kScTailLog = 10;
struct ScTailJob {
  const ef2* p[3];
  const ef2* chal;
  ef2* rounds;
  ef2* finals;
  uint32_t d, len, m1, pad;  // m1 as in ScRoundJob
};
// The tails of every sumcheck (one workgroup each), then the reductions r[0..nr); by value.
struct ScTail {
  ScTailJob j[kScMaxJobs];
  RedJob r[kScMaxReds];
  uint32_t nj, nr;
};
hipError_t launch_sc_tail_group(const ScTail& tail, hipStream_t st);

// One Basefold reduction round over its three channels (bf_round3_kernel, basefold.rs:748-775).
struct BfRound {
  const ef2* m_in[3];
  const ef2* eq[3];
  ef2* m_out[3];
  ef2* eq_next[3];
  ef2* part;  // 3 x blocks partial sums of l0 (channel-major)
  RedJob prev[3];  // the previous round's reductions, run by blocks (c, 4) of this launch
  // the same round's fold of the combined codeword (eq_fold_codeword), in the same launch
  const ef2* cw_in;
  uint64_t cw_half;
  const uint64_t* cw_tw;
  ef2* cw_out;
};
// Grid: bf_round3_blocks(half, cw_half) x 5 (rows 0-2 channels, 3 codeword fold, 4 the previous
// round's reductions); the partials of this round are that many per channel.
uint32_t bf_round3_blocks(uint64_t half, uint64_t cw_half);
hipError_t launch_bf_round3(const BfRound& a, uint64_t half, const ef2* r, hipStream_t st);

// Rounds t0..R-1 of the Basefold reduction in one workgroup (bf_tail_kernel); level t0 has
// D >> t0 <= 2^kBfTailLog entries.
constexpr int kBfTailLog = 10;
struct BfTail {
  const ef2* m_in[3];  // round t0's m (D >> (t0 + 1) entries) and eq (D >> (t0 + 2)) per channel
  const ef2* eq[3];
  RedJob prev[3];       // sums of the last grid round (nblk 0: none)
  const ef2* cw_in;     // level t0
  ef2* lvl_out;         // level t0 + 1; each later level follows the previous one
  ef2* l0;              // round t's sum of channel c goes to l0[c * R + t]
  const ef2* chal;      // round t's challenge at chal[t]
  const uint64_t* tw;   // tw[k] = g^k (k < D / 2), g of order D
  uint32_t R, t0, logD, pad;
};
hipError_t launch_bf_tail(const BfTail& a, hipStream_t st);

// One reduction round of the single-polynomial Basefold opening (bf1_round_kernel,
// basefold.rs:445-467): partial sums of l0 and l1 (two per block, block-major), the fold of m and
// of the eq table, the codeword fold (cw_out), and the previous round's reduction (prev).
struct Bf1Round {
  const ef2* m_in;
  const ef2* eq;
  ef2* m_out;
  ef2* eq_next;
  ef2* part;
  RedJob prev;
  const ef2* cw_in;
  uint64_t cw_half;
  const uint64_t* cw_tw;
  ef2* cw_out;
};
uint32_t bf1_round_blocks(uint64_t half, uint64_t cw_half);
hipError_t launch_bf1_round(const Bf1Round& a, uint64_t half, const ef2* r, hipStream_t st);
// Rounds t0..R-1 of the single-polynomial opening in one workgroup (bf1_tail_kernel): level t0
// has D >> t0 <= 2^kBfTailLog entries; lpair[2 t], lpair[2 t + 1] receive round t's l0, l1.
struct Bf1Tail {
  const ef2* m_in;  // round t0's m (2^(R - t0) entries) and eq (half of that)
  const ef2* eq;
  RedJob prev;        // the last grid round's sums (nblk 0: none)
  const ef2* cw_in;   // level t0
  ef2* lvl_out;       // level t0 + 1; each later level follows the previous one
  ef2* lpair;
  const ef2* chal;    // round t's challenge at chal[t]
  const uint64_t* tw; // tw[k] = g^k (k < D / 2), g of order D
  uint32_t R, t0, logD, pad;
};
hipError_t launch_bf1_tail(const Bf1Tail& a, hipStream_t st);

// Up to three Basefold channels combined in one pass (basefold.rs:732-741).
struct BfChannels {
  int count;
  const uint64_t* codes[3];
  const ef2* wt[3];  // coeffs[c] * eq4_c[j]
  uint32_t chunks[3];
  int ext[3];
};

// Half tables of one eq term: lo over pt[0..klo), hi over pt[klo..k) times *scale (if set).
struct EqTerm {
  const ef2* pt;
  const ef2* scale;
  ef2* lo;
  ef2* hi;
  uint32_t k, klo, first_block, pad;
};
// out[b] = sum over nterms of lo[q][b mod 2^klo] * hi[q][b >> klo], b < n.
struct EqTable {
  ef2* out;
  uint64_t n;
  const ef2* lo[3];
  const ef2* hi[3];
  uint32_t klo, nterms, first_block, pad;
};
hipError_t launch_eq_group(const EqTerm* terms, int nterms, uint32_t term_blocks, const EqTable* tabs, int ntabs,
                           uint32_t tab_blocks, hipStream_t st);
// out[j] = sum_i w[i] M[i][j] over `rows` rows of `cols` words: `chunks` row ranges summed into
// part (chunks x cols), then a second launch adds the chunks (copies into out2 / out3 if set).
hipError_t launch_wrows(const uint64_t* M, uint32_t rows, uint64_t cols, const ef2* w, ef2* part,
                        uint32_t chunks, ef2* out, hipStream_t st, ef2* out2 = nullptr, ef2* out3 = nullptr);
// Partial sums of sum_i a[i] b[i], one per block (dot_blocks(n) blocks), for a RedJob.
uint32_t dot_blocks(uint64_t n);
hipError_t launch_dot(const ef2* a, const ef2* b, uint64_t n, ef2* part, hipStream_t st);
// B_I as 0/1 words and bytes: bi[i * s + k] = bi8[i * s + k] = bit i of column idx[k] of the
// encoded rows.
hipError_t launch_bits_gather(const uint64_t* enc, uint64_t ld, const uint32_t* idx, uint32_t s, uint64_t W,
                              uint64_t* bi, uint8_t* bi8, hipStream_t st);
// Elementwise inputs of the Stage A sumchecks in one launch (block ranges filled in by the
// launcher): g2[i] = sum_k eq_r2[k] bi[i s + k] (and its copy g2b), sel zeroed (sel_n entries),
// v[i] = eqz1[i / 64] 2^(i mod 64), the RS weights w (n entries, twiddles of order 2n), and
// RS(a-bar) interleaved from planes c0, c1 (pn entries; pn = 0 skips).
struct PrepJobs {
  const uint64_t* bi;
  uint32_t s;
  uint64_t W;
  const ef2* eq_r2;
  ef2* g2;
  ef2* g2b;
  ef2* sel;
  uint64_t sel_n;
  const ef2* eqz1;
  ef2* vv;
  const ef2* r3;
  int log2n;
  uint64_t n;
  const uint64_t* tw;
  ef2* wv;
  const uint64_t* c0;
  const uint64_t* c1;
  uint64_t pn;
  ef2* ra;
  uint32_t b_g2, b_zero, b_vv, b_rs;
};
hipError_t launch_prep_group(PrepJobs a, hipStream_t st);
// sel[idx[k]] = eq_r2[k] into the zeroed sel, and y4 = sum_k eq_r2[k] ra[idx[k]].
hipError_t launch_sel_y4(const uint32_t* idx, const ef2* eq_r2, uint32_t s, const ef2* ra, ef2* sel, ef2* y4,
                         hipStream_t st);
uint32_t sc_round_blocks(uint64_t half);
hipError_t launch_bf_combine(const BfChannels& ch, uint64_t D, ef2* out, hipStream_t st);
hipError_t launch_fold_low(const uint64_t* poly, int ext, int li, uint64_t nout, const ef2* eq4, ef2* m,
                           hipStream_t st);
// One part of a staged gather: for each listed position k (in order), `rows` runs of `width`
// consecutive words, run rr starting at word pos[k] * pmul + rr * rstride of src. Part words
// land at dst[first ...]; parts are sorted by first and tile [0, total).
struct GatherPart {
  const uint64_t* src;
  uint64_t first;
  uint64_t pmul, rstride;
  uint32_t pos_at, rows, width, pad;
};
hipError_t launch_gather_parts(const GatherPart* parts, uint32_t nparts, const uint32_t* pos, uint64_t total,
                               uint64_t* dst, hipStream_t st);

}  // namespace lgs
