// Eval-prove on the GPU (pcs/src/ligesis.rs:347-549 with Basefold::batch_prove,
// pcs/src/basefold.rs:679-856). The transcript is fixed before any launch: the reference's
// RandomOracle hands out pre-drawn values in order (utils/src/oracle.rs:32-55), so the schedule
// below assigns every challenge its position in the field / integer streams, the host derives
// every evaluation point from them, and the device work is enqueued back to back. Proof values
// are gathered on the device into one staging buffer, copied to the host once, and encoded.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <memory>
#include <vector>

#include "prove_kernels.hpp"
#include "prove_util.hpp"
#include "runtime.hpp"

namespace lgs {

namespace {

using namespace pu;

// Split of B_I's 2^Lb-entry eq tables into low (2^klo) and high factors: blocks of the bits
// rounds take up to 2^klo consecutive pairs sharing the high factor (the larger klo, the fewer
// blocks and the more pairs per thread), and round tau <= 3 pairs entries 2^(Lb - 1 - tau) apart,
// which must be a multiple of 2^klo: 8 <= klo <= Lb - 4.
uint32_t bits_klo(uint32_t Lb) { return std::max<uint32_t>(8, std::min<uint32_t>(12, Lb - 4)); }

struct Sumcheck {
  int d = 0;
  uint32_t nv = 0;
  size_t rounds_at = 0, finals_at = 0;  // offsets (in ef2) into the results buffer
};

}  // namespace

// The device side of one proof up to the openings: every scratch buffer, the Basefold states of
// B_I and a-bar, the folded levels' trees and job tables, and the pinned staging of every upload
// (blob_host / idx_host: the per-proof constants and column indices). A plan stays with its
// commitment after a proof on the commitment's context; the next proof captures its launches
// into a graph, and later ones refill the two staging slots and replay the graph.
struct ProvePlan {
  DevicePool* pool = nullptr;
  std::deque<DevBuf<uint8_t>> bufs;
  PinnedArena arena;
  void* blob_host = nullptr;
  void* idx_host = nullptr;
  DevBasefold av_state, bi_state;
  std::vector<DevTree> ftrees;
  std::vector<std::vector<MerkleJob>> fjobs;
  std::vector<DevBuf<uint64_t>> ftables;
  ef2* res = nullptr;     // round polynomials, final evaluations, y, y4, l0 (layout in prove_impl)
  ef2* levels = nullptr;  // the batched codeword and its folds, level after level
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  template <class T>
  T* get(size_t n) {
    bufs.emplace_back(pool, n * sizeof(T));
    return (T*)bufs.back().ptr;
  }
  ~ProvePlan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

void prove_draws(uint32_t log_m, uint32_t log_n, uint64_t& n_fields, uint64_t& n_ints) {
  // the schedule below: alpha1, alpha2, the binary check, r2, hash, recon, table, RS link,
  // gamma_bi, B_I batch, gamma_a, a-bar batch, the Basefold coefficients and rounds
  const uint64_t log_w = 6 + log_m, log_s = 7, Lb = log_w + log_s;
  n_fields = Lb + 5 + Lb + log_s + log_w + log_w + (log_n + 1) + log_n + 1 + Lb + 1 + log_n + 3 +
             batch_rounds(log_m, log_n);
  n_ints = 4 * 128 + 100;  // column samples (ligesis.rs:329-345), then the Basefold queries
}

int64_t prove_impl(Context& ctx, const Key& key, const Commitment& cm, const uint64_t* z, uint64_t z_len,
                   const uint64_t* fields, uint64_t n_fields, const uint64_t* ints, uint64_t n_ints,
                   ProofOut out, uint64_t used[2]) {
  if (!key.has_state) throw Error(7, "prove: the key was built commit-only (no A-bar commitment)");
  if (cm.key != &key) throw Error(1, "prove: commitment was made with another key");
  const auto t_start = std::chrono::steady_clock::now();
  const uint32_t log_m = key.log_m, log_n = key.log_n;
  const uint64_t m = uint64_t(1) << log_m, n = uint64_t(1) << log_n, two_n = 2 * n;
  const uint64_t W = key.W, s = 128;
  const uint32_t log_w = 6 + log_m, log_s = 7, Lb = log_w + log_s;
  const uint32_t R = batch_rounds(log_m, log_n);
  if (z_len != log_m + log_n) throw Error(1, "prove: z must have log_m + log_n entries");

  // ---- transcript schedule (ligesis.rs:390-530, basefold.rs:732, 761, 788) ----
  size_t fc = 0;
  auto take = [&](size_t k) {
    if (fc + k > n_fields) throw Error(5, "prove: randomness oracle field stream exhausted");
    EfV v(k);
    for (size_t i = 0; i < k; i++) v[i] = ef2{fields[2 * (fc + i)], fields[2 * (fc + i) + 1]};
    fc += k;
    return v;
  };
  EfV alpha1 = take(Lb), alpha2 = take(5), ch_bin = take(Lb), r2 = take(log_s);
  EfV ch_hash = take(log_w), ch_recon = take(log_w), ch_table = take(log_n + 1), ch_rs = take(log_n);
  EfV gamma_bi = take(1), ch_bibatch = take(Lb), gamma_a = take(1), ch_abatch = take(log_n);
  EfV coeffs = take(3), ch_batch = take(R);
  if (n_ints < 4 * s + 100) throw Error(5, "prove: randomness oracle integer stream exhausted");
  std::vector<uint32_t> idx;
  {
    std::vector<uint8_t> seen(two_n, 0);
    for (uint64_t k = 0; k < 4 * s && idx.size() < s; k++) {
      uint64_t j = ints[k] % two_n;
      if (!seen[j]) {
        seen[j] = 1;
        idx.push_back((uint32_t)j);
      }
    }
    if (idx.size() != s) throw Error(5, "not enough distinct columns sampled from oracle");
  }
  std::vector<uint64_t> sorted(idx.begin(), idx.end());
  std::sort(sorted.begin(), sorted.end());
  used[0] = fc;
  used[1] = 4 * s + 100;
  {
    uint64_t nf = 0, ni = 0;
    prove_draws(log_m, log_n, nf, ni);
    if (nf != used[0] || ni != used[1]) throw Error(1, "prove: transcript length does not match prove_draws");
  }

  EfV zz(z_len);
  for (size_t i = 0; i < z_len; i++) zz[i] = ef2{z[2 * i], z[2 * i + 1]};
  EfV z2(zz.begin(), zz.begin() + log_n), z1(zz.begin() + log_n, zz.end());
  EfV r1 = rev(ch_bin), r4 = rev(ch_hash), r5 = rev(ch_recon), r3 = rev(ch_table), r6 = rev(ch_rs);
  EfV r_bi = rev(ch_bibatch), r_a = rev(ch_abatch);
  EfV p_a = cat(r4, alpha2), p_bi_hash = cat(r2, r4), p_bi_recon = cat(r2, r5);
  EfV gbi{ef_one(), gamma_bi[0], ef_mul(gamma_bi[0], gamma_bi[0])}, ga{ef_one(), gamma_a[0]};

  const uint32_t nv_a = 11 + log_m, nv_bi = Lb;
  const uint32_t li_A = shared_li(nv_a, log_m, log_n), li_bi = shared_li(nv_bi, log_m, log_n),
                 li_a = shared_li(log_n, log_m, log_n);
  const EfV* bpoint[3] = {&p_a, &r_bi, &r_a};
  const uint32_t bli[3] = {li_A, li_bi, li_a};
  EfV wt[3], eq4[3], zp[3];
  for (int c = 0; c < 3; c++) {
    EfV head(bpoint[c]->begin(), bpoint[c]->begin() + bli[c]);
    eq4[c] = host_eq(head);
    for (auto& e : eq4[c]) wt[c].push_back(ef_mul(coeffs[c], e));
    zp[c].assign(bpoint[c]->begin() + bli[c], bpoint[c]->end());
    if (zp[c].size() != R) throw Error(1, "prove: Basefold point length mismatch");
  }

  // ---- one constants blob: every point / scalar the kernels read ----
  EfV blob;
  auto put = [&](const EfV& v) {
    size_t at = blob.size();
    blob.insert(blob.end(), v.begin(), v.end());
    return at;
  };
  size_t o_z1 = put(z1), o_z2 = put(z2), o_alpha1 = put(alpha1), o_alpha2 = put(alpha2);
  size_t o_chbin = put(ch_bin), o_r2 = put(r2), o_chhash = put(ch_hash), o_chrecon = put(ch_recon);
  size_t o_chtable = put(ch_table), o_chrs = put(ch_rs), o_r3 = put(r3), o_r1 = put(r1);
  size_t o_pbh = put(p_bi_hash), o_pbr = put(p_bi_recon), o_r6 = put(r6), o_gbi = put(gbi), o_ga = put(ga);
  size_t o_chbib = put(ch_bibatch), o_chab = put(ch_abatch), o_chbatch = put(ch_batch);
  size_t o_wt[3], o_eq4[3], o_zp1[3];
  for (int c = 0; c < 3; c++) {
    o_wt[c] = put(wt[c]);
    o_eq4[c] = put(eq4[c]);
    o_zp1[c] = put(EfV(zp[c].begin() + (R ? 1 : 0), zp[c].end()));
  }
  // Rounds 1..3 of B_I's two sumchecks (binary check, f = 0; batched opening, f = 1) read the
  // 0/1 bytes: the prefix weights W_tau(y) = prod_t (y_t ? c_t : 1 - c_t) (y_0 the top bit, bound
  // first), and for every term q of their eq tables the fold of its high factor over the bound
  // variables, scale_q prod_{t < tau} eq1(pt_q[Lb - 1 - t], c_t), eq1(a, c) = a c + (1 - a)(1 - c).
  const EfV* bits_ch[2] = {&ch_bin, &ch_bibatch};
  const EfV* bits_pt[2][3] = {{&alpha1, nullptr, nullptr}, {&r1, &p_bi_hash, &p_bi_recon}};
  const ef2 bits_scale[2][3] = {{ef_one(), ef_one(), ef_one()}, {gbi[0], gbi[1], gbi[2]}};
  const size_t bits_offs[2][3] = {{o_alpha1, 0, 0}, {o_r1, o_pbh, o_pbr}};
  const uint32_t bits_nterms[2] = {1, 3};
  size_t o_bw[2][4] = {}, o_bs[2][4][3] = {};
  for (int f = 0; f < 2; f++) {
    const EfV& c = *bits_ch[f];
    for (uint32_t tau = 1; tau < 4 && tau < Lb; tau++) {
      EfV w(size_t(1) << tau);
      for (size_t y = 0; y < w.size(); y++) {
        ef2 v = ef_one();
        for (uint32_t t = 0; t < tau; t++)
          v = ef_mul(v, ((y >> (tau - 1 - t)) & 1) ? c[t] : ef_sub(ef_one(), c[t]));
        w[y] = v;
      }
      // the kernel looks b up by the 2^tau prefix bits: store every subset sum of w
      EfV sums(size_t(1) << w.size());
      for (size_t mm = 0; mm < sums.size(); mm++) {
        ef2 acc = ef_zero();
        for (size_t y = 0; y < w.size(); y++)
          if ((mm >> y) & 1) acc = ef_add(acc, w[y]);
        sums[mm] = acc;
      }
      if (f == 0 && tau <= 2) {  // binary check: (c0, c1, c2) for every (lo, hi) bit pattern
        const size_t nm = sums.size();
        for (size_t ml = 0; ml < nm; ml++)
          for (size_t mh = 0; mh < nm; mh++) {
            const ef2 bl = sums[ml], db = ef_sub(sums[mh], bl);
            sums.push_back(ef_mul(bl, ef_sub(bl, ef_one())));
            sums.push_back(ef_mul(db, ef_sub(ef_add(bl, bl), ef_one())));
            sums.push_back(ef_mul(db, db));
          }
      }
      o_bw[f][tau] = put(sums);
      for (uint32_t q = 0; q < bits_nterms[f]; q++) {
        const EfV& pt = *bits_pt[f][q];
        ef2 v = bits_scale[f][q];
        for (uint32_t t = 0; t < tau; t++) {
          const ef2 a = pt[Lb - 1 - t];
          v = ef_mul(v, ef_add(ef_mul(a, c[t]), ef_mul(ef_sub(ef_one(), a), ef_sub(ef_one(), c[t]))));
        }
        o_bs[f][tau][q] = put(EfV{v});
      }
    }
  }

  // shapes and layouts (the same for every prove of this key)
  const uint64_t D = key.a_state.D;  // shared Basefold domain
  const int logN = log_n + 1;
  const uint64_t nbi = W * s;  // entries of B_I
  const uint64_t M = uint64_t(1) << R;
  // results buffer layout
  Sumcheck sc[7];
  const int sc_d[7] = {3, 2, 2, 2, 2, 2, 2};
  const uint32_t sc_nv[7] = {Lb, log_w, log_w, log_n + 1, log_n, Lb, log_n};
  size_t res_n = 0;
  for (int i = 0; i < 7; i++) {
    sc[i].d = sc_d[i];
    sc[i].nv = sc_nv[i];
    sc[i].rounds_at = res_n;
    res_n += (size_t)sc[i].nv * (sc[i].d + 1);
    sc[i].finals_at = res_n;
    res_n += sc[i].d;
  }
  const size_t o_y = res_n++, o_y4 = res_n++;
  const size_t o_l0 = res_n;
  res_n += 3 * (size_t)R;
  // levels 0..R stored back to back: level k has D >> k entries
  std::vector<uint64_t> lvl_off(R + 1);
  uint64_t tot = 0;
  for (uint32_t k = 0; k <= R; k++) {
    lvl_off[k] = tot;
    tot += D >> k;
  }

  hipStream_t st = ctx.stream;
  // the second stream; with single_stream_prove the same stream (every step is issued after the
  // steps it reads, so one stream runs them in a valid order)
  hipStream_t sb = ctx.single_stream_prove ? st : ctx.side_stream();
  DevicePool* pool = &ctx.pool;
  ctx.upload_reset();  // every earlier call on this context has been synchronized
  // profile mode: device timestamps of the phases (stage_ms[7..16], from the first upload)
  auto mark = [&](int i, hipStream_t s) {
    if (!ctx.profile) return;
    hip_check(hipEventRecord(ctx.timing_event(32 + i), s), "profile event");
    // host time at which the work up to this mark was issued
    ctx.stage_ms[17 + i] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  };

  // ---- every device step of the proof up to the openings, on st and sb, into plan P's
  //      buffers; the blob and idx are uploaded from P's staging slots ----
  auto issue = [&](ProvePlan& P) {
  DevBasefold& av_state = P.av_state;
  DevBasefold& bi_state = P.bi_state;
  mark(0, st);
  // the constants blob and the column indices, in one upload
  const size_t blob_bytes = blob.size() * sizeof(ef2);
  uint8_t* dconst = P.get<uint8_t>(blob_bytes + s * 4);
  {
    std::vector<uint8_t> h(blob_bytes + s * 4);
    std::memcpy(h.data(), blob.data(), blob_bytes);
    std::memcpy(h.data() + blob_bytes, idx.data(), s * 4);
    uint8_t* hp = (uint8_t*)P.arena.upload(dconst, h.data(), h.size(), st);
    P.blob_host = hp;
    P.idx_host = hp + blob_bytes;
  }
  const ef2* B = (const ef2*)dconst;
  uint32_t* didx = (uint32_t*)(dconst + blob_bytes);

  // ---- the second stream: B_I, then (issued after st's launches up to the sumchecks) B_I's and
  // a-bar's Basefold commitments, then Stage B ----
  // B_I (ligesis.rs:402-410) needs only F' and the column indices, so the stream forks here;
  // only Stage B reads the commitments (and the proof their roots), and Stage B joins st before
  // the openings are gathered. Every twiddle table either stream needs is made on st first.
  for (uint32_t t = 0; t < R; t++) (void)ctx.twiddle(ilog2(D >> t));
  (void)ctx.twiddle(ilog2(D));
  (void)ctx.twiddle(logN);
  (void)ctx.pass_twiddle(13, 4);  // the long Basefold encodings' subsequence transforms
  hip_check(hipEventRecord(ctx.ev[5], st), "fork");
  hip_check(hipStreamWaitEvent(sb, ctx.ev[5], 0), "fork");
  mark(6, sb);
  // The two sumchecks over B_I read these 0/1 words in their first round, which writes the folded
  // extension-field tables (half the length) that their later rounds work on; st waits for them
  // (ev[8]) before its first use.
  uint64_t* bi = P.get<uint64_t>(nbi);
  uint8_t* bi8 = P.get<uint8_t>(nbi);
  ef2* bi_ef = P.get<ef2>(nbi / 2);
  ef2* bi_ef2 = P.get<ef2>(nbi / 2);
  hip_check(launch_bits_gather(cm.enc.ptr, two_n, didx, (uint32_t)s, W, bi, bi8, sb), "B_I gather");
  hip_check(hipEventRecord(ctx.ev[8], sb), "B_I ready");

  ef2* res = P.get<ef2>(res_n);
  P.res = res;

  // partial sums of the dot product y; the first sumcheck launch reduces them
  ef2* part = P.get<ef2>(256);
  // Every buffer below lives until the end of the prove: the second stream runs beside this
  // one, so no buffer may go back to the pool (and be handed out again) while work is queued.

  // ---- every eq table of the proof, in two launches (MlPoly::new_eq, poly.rs:44-53) ----
  // big and big2 (2^Lb entries) are never expanded: the first round of their sumchecks evaluates
  // them from their half tables and stores the folded halves here
  ef2* eqz1 = P.get<ef2>(m);
  ef2* eqz2 = P.get<ef2>(n);
  ef2* big = P.get<ef2>(nbi / 2);
  ef2* big2 = P.get<ef2>(nbi / 2);
  ef2* eqa2 = P.get<ef2>(32);
  ef2* eqr2 = P.get<ef2>(s);
  ef2* g_a = P.get<ef2>(n);
  ef2* mbuf[3][2];
  ef2* eqr[3][2];
  for (int c = 0; c < 3; c++) {
    mbuf[c][0] = P.get<ef2>(M);
    mbuf[c][1] = P.get<ef2>(M / 2 + 1);
    eqr[c][0] = P.get<ef2>(M / 2 + 1);
    eqr[c][1] = P.get<ef2>(M / 4 + 1);
  }
  std::vector<EqTerm> eq_terms;
  std::vector<EqTable> all_tabs, eq_tabs, eq_fact;  // all in `want` order; expanded; factored
  std::vector<bool> factored;
  uint64_t eq_half_words = 0;
  {
    // a table: out (2^k entries) = sum over (point, scale) terms of scale * eq(point, .)
    struct Want {
      ef2* out;
      uint32_t k;
      std::vector<std::pair<size_t, const ef2*>> terms;  // blob offset of the point, scale
      bool factored;
    };
    std::vector<Want> want = {
        {eqz1, log_m, {{o_z1, nullptr}}, false},
        {eqz2, log_n, {{o_z2, nullptr}}, false},
        {big, Lb, {{o_alpha1, nullptr}}, true},                 // binary check (ligesis.rs:424-429)
        {eqa2, 5, {{o_alpha2, nullptr}}, false},                // hash channel (ligesis.rs:436-444)
        {eqr2, log_s, {{o_r2, nullptr}}, false},
        {big2, Lb, {{o_r1, B + o_gbi}, {o_pbh, B + o_gbi + 1}, {o_pbr, B + o_gbi + 2}}, true},  // 513-517
        {g_a, log_n, {{o_z2, B + o_ga}, {o_r6, B + o_ga + 1}}, false},                         // 518-522
    };
    for (int c = 0; c < 3; c++)
      if (R >= 1) want.push_back({eqr[c][0], R - 1, {{o_zp1[c], nullptr}}, false});  // Basefold rounds
    for (auto& w : want) {
      // B_I's two tables are evaluated from their factors in blocks of 256 consecutive entries,
      // which must share the high factor: 2^klo >= 256
      const uint32_t klo = w.factored ? bits_klo(w.k) : w.k / 2, khi = w.k - klo;
      EqTable tb{};
      tb.out = w.out;
      tb.n = uint64_t(1) << w.k;
      tb.klo = klo;
      tb.nterms = (uint32_t)w.terms.size();
      for (size_t q = 0; q < w.terms.size(); q++) {
        EqTerm t{};
        t.pt = B + w.terms[q].first;
        t.scale = w.terms[q].second;
        t.k = w.k;
        t.klo = klo;
        t.lo = (ef2*)(uintptr_t)eq_half_words;  // offsets into the half-table scratch for now
        eq_half_words += uint64_t(1) << klo;
        t.hi = (ef2*)(uintptr_t)eq_half_words;
        eq_half_words += uint64_t(1) << khi;
        eq_terms.push_back(t);
      }
      all_tabs.push_back(tb);
      factored.push_back(w.factored);
    }
  }
  // the folded high factors of B_I's eq tables for bits rounds 1..bits_rounds - 1: terms of
  // k = khi - tau variables (klo = 0) over pt[klo ..), scaled by the folds of the bound ones
  const uint32_t bits_rounds = std::min<uint32_t>(4, Lb > (uint32_t)kScTailLog ? Lb - kScTailLog : 0);
  const uint32_t b_klo = bits_klo(Lb), b_khi = Lb - b_klo;
  struct BitsTerm {
    int f;
    uint32_t tau, q;
  };
  std::vector<EqTerm> bits_terms;
  std::vector<BitsTerm> bits_meta;
  for (int f = 0; f < 2; f++)
    for (uint32_t tau = 1; tau < bits_rounds; tau++)
      for (uint32_t q = 0; q < bits_nterms[f]; q++) {
        EqTerm t{};
        t.pt = B + bits_offs[f][q] + b_klo;
        t.scale = B + o_bs[f][tau][q];
        t.k = b_khi - tau;
        t.klo = 0;
        t.lo = (ef2*)(uintptr_t)eq_half_words;
        eq_half_words += 1;
        t.hi = (ef2*)(uintptr_t)eq_half_words;
        eq_half_words += uint64_t(1) << t.k;
        bits_terms.push_back(t);
        bits_meta.push_back({f, tau, q});
      }
  ef2* eq_half = P.get<ef2>(eq_half_words);
  uint32_t eq_term_blocks = 0, eq_tab_blocks = 0;
  {
    size_t ti = 0;  // terms were appended table by table in `want` order
    for (size_t a = 0; a < all_tabs.size(); a++) {
      EqTable& tb = all_tabs[a];
      for (uint32_t q = 0; q < tb.nterms; q++, ti++) {
        EqTerm& t = eq_terms[ti];
        t.lo = eq_half + (uintptr_t)t.lo;
        t.hi = eq_half + (uintptr_t)t.hi;
        t.first_block = eq_term_blocks;
        eq_term_blocks += (uint32_t)(((uint64_t(1) << std::max(t.klo, t.k - t.klo)) + 255) / 256);
        tb.lo[q] = t.lo;
        tb.hi[q] = t.hi;
      }
      if (factored[a]) {  // big and big2, in that order
        eq_fact.push_back(tb);
        continue;
      }
      tb.first_block = eq_tab_blocks;
      eq_tab_blocks += (uint32_t)((tb.n + 255) / 256);
      eq_tabs.push_back(tb);
    }
  }
  // eq_fact: [0] big, [1] big2 (round 0), then for f = 0, 1 and tau = 1 .. bits_rounds - 1 the
  // round-tau factors: the same low halves, the folded high halves
  std::vector<size_t> bits_fact(8, 0);  // f * 4 + tau -> index into eq_fact
  if (eq_fact.size() != 2 || eq_fact[0].nterms != 1 || eq_fact[1].nterms != 3)
    throw Error(1, "prove: B_I's eq tables have one and three terms");
  for (size_t k = 0; k < bits_terms.size(); k++) {
    EqTerm& t = bits_terms[k];
    const BitsTerm& bm = bits_meta[k];
    t.lo = eq_half + (uintptr_t)t.lo;
    t.hi = eq_half + (uintptr_t)t.hi;
    t.first_block = eq_term_blocks;
    eq_term_blocks += (uint32_t)(((uint64_t(1) << t.k) + 255) / 256);
    eq_terms.push_back(t);
    if (bm.q == 0) {
      EqTable tb = eq_fact[bm.f];
      tb.out = nullptr;
      tb.n >>= bm.tau;
      bits_fact[bm.f * 4 + bm.tau] = eq_fact.size();
      eq_fact.push_back(tb);
    }
    eq_fact[bits_fact[bm.f * 4 + bm.tau]].hi[bm.q] = t.hi;
  }
  const size_t terms_bytes = eq_terms.size() * sizeof(EqTerm), tabs_bytes = eq_tabs.size() * sizeof(EqTable);
  const size_t fact_bytes = eq_fact.size() * sizeof(EqTable);
  uint8_t* eq_jobs = P.get<uint8_t>(terms_bytes + tabs_bytes + fact_bytes);
  {  // one upload: terms, tables, factored tables
    std::vector<uint8_t> h(terms_bytes + tabs_bytes + fact_bytes);
    std::memcpy(h.data(), eq_terms.data(), terms_bytes);
    std::memcpy(h.data() + terms_bytes, eq_tabs.data(), tabs_bytes);
    std::memcpy(h.data() + terms_bytes + tabs_bytes, eq_fact.data(), fact_bytes);
    P.arena.upload(eq_jobs, h.data(), h.size(), st);
  }
  const EqTable* d_eq_fact = (const EqTable*)(eq_jobs + terms_bytes + tabs_bytes);  // [0] big, [1] big2
  hip_check(launch_eq_group((const EqTerm*)eq_jobs, (int)eq_terms.size(), eq_term_blocks,
                            (const EqTable*)(eq_jobs + eq_terms.size() * sizeof(EqTerm)), (int)eq_tabs.size(),
                            eq_tab_blocks, st),
            "eq tables");

  // ---- a-bar = eq_{z1}^T F, y = a-bar(z2) (ligesis.rs:365-379); two copies for the sumchecks
  //      that fold a-bar in place ----
  ef2* a_vec = P.get<ef2>(n);
  ef2* a_copy = P.get<ef2>(n);
  ef2* a_copy2 = P.get<ef2>(n);
  // row chunks: about 2^19 threads in all (2048 workgroups at n = 2^16), few partials to add
  uint64_t chunks = std::max<uint64_t>(1, std::min<uint64_t>(m, (uint64_t(1) << 19) / n));
  while (m % chunks) chunks >>= 1;
  ef2* wpart = P.get<ef2>(chunks * n);
  ef2* wpart_g1 = P.get<ef2>(W);
  hip_check(launch_wrows(cm.f, (uint32_t)m, n, eqz1, wpart, (uint32_t)chunks, a_vec, st, a_copy,
                         a_copy2),
            "a-bar");
  hip_check(launch_dot(a_vec, eqz2, n, part, st), "y");

  hip_check(hipEventRecord(ctx.ev[9], st), "a-bar ready");  // a-bar's Basefold commitment waits for it


  // RS(a-bar) (ligesis.rs:387), both planes zero-padded to 2n, transformed here rather than
  // taken from a-bar's Basefold codeword (the same transform when log_interleave is 0), so that
  // the sumchecks need not wait for the other streams
  uint64_t* ra_planar = P.get<uint64_t>(2 * two_n);
  hip_check(launch_ntt((const uint64_t*)a_vec, 1, 2, (uint32_t)n, 2, logN, ctx.twiddle(logN), ra_planar, st),
            "RS(a-bar)");
  const uint64_t* ra_src = ra_planar;
  ef2* ra = P.get<ef2>(two_n);

  // ---- input tables of the seven Stage A sumchecks ----
  // hash / recon channels (ligesis.rs:432-485)
  ef2* g1 = P.get<ef2>(W);
  ef2* g2 = P.get<ef2>(W);
  ef2* g2b = P.get<ef2>(W);
  ef2* vv = P.get<ef2>(W);
  hip_check(launch_wrows(key.a.ptr, 32, W, eqa2, wpart_g1, 1, g1, st), "g1");
  ef2* sel = P.get<ef2>(two_n);
  // recon table and RS link (ligesis.rs:488-502); y5 = ra(r3) is sc_table's final evaluation of
  // ra, which is what the reference's verify compares (ligesis.rs:681)
  ef2* wv = P.get<ef2>(n);
  // g2 reads B_I from the second stream
  hip_check(hipStreamWaitEvent(st, ctx.ev[8], 0), "B_I ready");
  PrepJobs pj{};
  pj.bi = bi;
  pj.s = (uint32_t)s;
  pj.W = W;
  pj.eq_r2 = eqr2;
  pj.g2 = g2;
  pj.g2b = g2b;
  pj.sel = sel;
  pj.sel_n = two_n;
  pj.eqz1 = eqz1;
  pj.vv = vv;
  pj.r3 = B + o_r3;
  pj.log2n = logN;
  pj.n = n;
  pj.tw = ctx.twiddle(logN);
  pj.wv = wv;
  pj.c0 = ra_src;
  pj.c1 = ra_src + two_n;
  pj.pn = two_n;
  pj.ra = ra;
  hip_check(launch_prep_group(pj, st), "sumcheck inputs");
  hip_check(launch_sel_y4(didx, eqr2, (uint32_t)s, ra, sel, res + o_y4, st), "sel, y4");

  // ---- the Basefold commitments of B_I (ligesis.rs:411-416) and, once a-bar is there, of a-bar
  //      (ligesis.rs:384) on the second stream. Issued after every launch that the sumchecks
  //      wait for on st: the host issues a proof more slowly than the device runs it in its
  //      first part, and st's chain up to the sumchecks is the longer one. (A third stream for
  //      a-bar's gained nothing alone and costs hardware queues when several contexts share the
  //      device.) ----
  enqueue_basefold_commit(ctx, bi_state, bi, false, nv_bi, li_bi, 1, sb);
  mark(7, sb);
  hip_check(hipStreamWaitEvent(sb, ctx.ev[9], 0), "a-bar ready");
  enqueue_basefold_commit(ctx, av_state, (const uint64_t*)a_vec, true, log_n, li_a, 1, sb);
  mark(8, sb);
  const DevBasefold* states[3] = {&key.a_state, &bi_state, &av_state};
  const uint64_t* polys[3] = {key.a.ptr, bi, (const uint64_t*)a_vec};
  for (int c = 0; c < 3; c++)
    if (states[c]->D != D) throw Error(1, "prove: Basefold domains differ");

  // ---- the seven sumchecks, in lockstep ----
  // Every challenge is fixed before the first round (the oracle streams are inputs), so the
  // sumchecks only depend on their own tables: grouped launches cover every sumcheck with
  // rounds left, aligned so that all reach their last rounds (<= 2^kScTailLog entries)
  // together; one more launch runs those tails, one workgroup per sumcheck. Each round folds its
  // tables in place, so every sumcheck owns its tables (big2, a_copy2).
  struct ScRun {
    ef2* p[3];
    size_t ch_off;
  };
  const ScRun runs[7] = {{{bi_ef, nullptr, big}, o_chbin},   {{g1, g2, nullptr}, o_chhash},
                         {{vv, g2b, nullptr}, o_chrecon},     {{sel, ra, nullptr}, o_chtable},
                         {{a_copy, wv, nullptr}, o_chrs},     {{bi_ef2, big2, nullptr}, o_chbib},
                         {{a_copy2, g_a, nullptr}, o_chab}};
  // One grouped launch per round step; the first round of B_I's two sumchecks reads the 0/1
  // words alone (sc_bits*_body).
  uint32_t n_big[7], steps = 0;
  for (int i = 0; i < 7; i++) {
    n_big[i] = sc[i].nv > kScTailLog ? sc[i].nv - kScTailLog : 0;  // rounds before the tail
    steps = std::max(steps, n_big[i]);
  }
  if (n_big[0] == 0 || n_big[5] == 0)
    throw Error(2, "prove: B_I has fewer than 2^" + std::to_string(kScTailLog + 1) + " entries");
  // Partial sums alternate between two halves of scpart: the launch of step tau writes its half
  // while its extra blocks reduce the other half, written by step tau - 1 (the dot products'
  // partials for step 0); the tail launch reduces the last step's.
  // blocks of one round of sumcheck i: a bits round covers 256 kpt pairs per block, all inside
  // one run of 2^klo entries of its eq table's low factor (kpt <= 2^klo / 256)
  auto round_shape = [&](int i, uint32_t t, uint32_t& nblk, uint32_t& kpt) {
    const uint64_t half = uint64_t(1) << (sc[i].nv - 1 - t);
    if ((i == 0 || i == 5) && t < bits_rounds) {
      kpt = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(1, half >> 17), (uint64_t(1) << b_klo) >> 8);
      nblk = (uint32_t)(half / (256 * (uint64_t)kpt));
    } else {
      kpt = 0;
      nblk = sc_round_blocks(half);
    }
  };
  size_t part_half = 0;  // partials of one step, the largest step
  for (uint32_t tau = 0; tau < steps; tau++) {
    size_t w = 0;
    for (int i = 0; i < 7; i++) {
      if (tau + n_big[i] < steps) continue;
      uint32_t nb, kp;
      round_shape(i, tau + n_big[i] - steps, nb, kp);
      w += (size_t)nb * (sc[i].d + 1);
    }
    part_half = std::max(part_half, w);
  }
  ef2* scpart = P.get<ef2>(2 * part_half);
  std::vector<ScRoundJob> rjobs;
  std::vector<RedJob> reds = {RedJob{part, res + o_y, dot_blocks(n), 1}};
  std::vector<uint32_t> step_first(steps + 1), step_blocks(steps), red_first(steps + 2);
  for (uint32_t tau = 0; tau < steps; tau++) {
    step_first[tau] = (uint32_t)rjobs.size();
    red_first[tau + 1] = (uint32_t)reds.size();
    uint32_t blocks = 0;
    size_t poff = (tau & 1) * part_half;
    for (int i = 0; i < 7; i++) {
      if (tau + n_big[i] < steps) continue;  // starts later
      const uint32_t t = tau + n_big[i] - steps;
      const Sumcheck& S = sc[i];
      ScRoundJob j{};
      for (int q = 0; q < 3; q++) j.p[q] = runs[i].p[q];
      j.half = uint64_t(1) << (S.nv - 1 - t);
      j.r = B + runs[i].ch_off + t;
      if ((i == 0 || i == 5) && t < bits_rounds) {  // B_I's first rounds read the 0/1 bytes
        const int f = i == 0 ? 0 : 1;
        j.bits = bi8;
        j.tau = t;
        j.eqt = d_eq_fact + (t == 0 ? f : bits_fact[f * 4 + t]);
        if (t > 0) j.wts = B + o_bw[f][t];
        if (t + 1 < bits_rounds)  // only the last of them writes the folded tables
          for (int q = 0; q < 3; q++) j.p[q] = nullptr;
      }
      j.m1 = i == 0;  // the binary check's second table is B_I - 1
      round_shape(i, t, j.nblk, j.kpt);
      j.part = scpart + poff;
      poff += (size_t)j.nblk * (S.d + 1);
      reds.push_back(RedJob{j.part, res + S.rounds_at + (size_t)t * (S.d + 1), j.nblk, (uint32_t)S.d + 1});
      j.d = (uint32_t)S.d;
      j.first_block = blocks;
      blocks += j.nblk;
      rjobs.push_back(j);
    }
    step_blocks[tau] = blocks;
  }
  step_first[steps] = (uint32_t)rjobs.size();
  red_first[0] = 0;
  red_first[steps + 1] = (uint32_t)reds.size();
  // launch tau (tau = steps: the tail) runs reductions [red_first[tau], red_first[tau + 1])
  std::vector<ScTailJob> tjobs(7);
  for (int i = 0; i < 7; i++) {
    const Sumcheck& S = sc[i];
    const uint32_t t0 = n_big[i];
    for (int q = 0; q < 3; q++) tjobs[i].p[q] = runs[i].p[q];
    tjobs[i].len = 1u << (S.nv - t0);
    tjobs[i].chal = B + runs[i].ch_off + t0;
    tjobs[i].rounds = res + S.rounds_at + (size_t)t0 * (S.d + 1);
    tjobs[i].finals = res + S.finals_at;
    tjobs[i].d = (uint32_t)S.d;
    tjobs[i].m1 = i == 0;
  }
  // the step tables go to the launches as kernel arguments: no uploads on st before the rounds
  auto fill_reds = [&](RedJob* out, uint32_t& nr, uint32_t tau) {
    nr = red_first[tau + 1] - red_first[tau];
    if (nr > (uint32_t)kScMaxReds) throw Error(1, "prove: too many reductions in one sumcheck launch");
    for (uint32_t q = 0; q < nr; q++) out[q] = reds[red_first[tau] + q];
  };
  mark(1, st);
  for (uint32_t tau = 0; tau < steps; tau++) {
    // bits jobs (B_I's first rounds, all at the same round) go to a launch of their own
    ScStep bstep{}, step{};
    uint32_t btau = 0;
    for (uint32_t q = step_first[tau]; q < step_first[tau + 1]; q++) {
      ScRoundJob j = rjobs[q];
      ScStep& dst = j.bits ? bstep : step;
      if (j.bits) btau = j.tau;
      if (dst.nj == (uint32_t)kScMaxJobs) throw Error(1, "prove: too many sumchecks in one launch");
      j.first_block = dst.round_blocks;
      dst.round_blocks += j.nblk;
      dst.j[dst.nj++] = j;
    }
    ScStep& rs = step.nj ? step : bstep;  // the step's reductions ride on one of the launches
    fill_reds(rs.r, rs.nr, tau);
    if (bstep.nj) hip_check(launch_sc_bits_group(bstep, btau, st), "sumcheck bits round step");
    if (step.nj) hip_check(launch_sc_round_group(step, st), "sumcheck round step");
  }
  ScTail tail{};
  tail.nj = 7;
  for (int i = 0; i < 7; i++) tail.j[i] = tjobs[i];
  fill_reds(tail.r, tail.nr, steps);
  hip_check(launch_sc_tail_group(tail, st), "sumcheck tails");
  mark(2, st);

  // ---- Stage B: shared-domain batched Basefold open (basefold.rs:679-856), second stream ----
  // Its inputs come from sb itself and from st's work before the a-bar event (the eq tables, the
  // constants); its scratch (partial sums) is its own. Issued after Stage A: Stage A is the
  // longer chain, and issued first, Stage B (about a quarter of the proof's host issue time)
  // held back Stage A's start by more than it let Stage B start earlier.
  // l0 partial sums, alternating halves by round parity (each round reduces the previous one's)
  ef2* part_b = P.get<ef2>(2 * 3 * 1024);
  ef2* levels = P.get<ef2>(tot);
  P.levels = levels;
  BfChannels chans{};
  chans.count = 3;
  for (int c = 0; c < 3; c++) {
    chans.codes[c] = states[c]->codes.ptr;
    chans.wt[c] = B + o_wt[c];
    chans.chunks[c] = 1u << bli[c];
    chans.ext[c] = states[c]->ext ? 1 : 0;
  }
  hip_check(launch_bf_combine(chans, D, levels, sb), "batch combine");
  mark(9, sb);
  for (int c = 0; c < 3; c++)
    hip_check(launch_fold_low(polys[c], states[c]->ext ? 1 : 0, (int)bli[c], M, B + o_eq4[c], mbuf[c][0], sb),
              "batch fold");
  // rounds whose level has more than 2^kBfTailLog entries on the grid, the rest in one workgroup
  const uint32_t logD = (uint32_t)ilog2(D);
  const uint32_t t0 = logD > (uint32_t)kBfTailLog ? std::min<uint32_t>(R, logD - kBfTailLog) : 0;
  RedJobs last{};
  for (uint32_t t = 0; t < t0; t++) {
    const uint64_t half = uint64_t(1) << (R - 1 - t);
    BfRound br{};
    for (int c = 0; c < 3; c++) {
      br.m_in[c] = mbuf[c][t & 1];
      br.m_out[c] = mbuf[c][(t + 1) & 1];
      br.eq[c] = eqr[c][t & 1];
      br.eq_next[c] = eqr[c][(t + 1) & 1];
      br.prev[c] = last.j[c];  // zero (no reduction) in round 0
    }
    br.part = part_b + (size_t)(t & 1) * 3 * 1024;
    const uint64_t len = D >> t;
    br.cw_in = levels + lvl_off[t];
    br.cw_half = len / 2;
    br.cw_tw = ctx.twiddle(ilog2(len));
    br.cw_out = levels + lvl_off[t + 1];
    hip_check(launch_bf_round3(br, half, B + o_chbatch + t, sb), "batch round");
    const uint32_t nb = bf_round3_blocks(half, br.cw_half);
    for (int c = 0; c < 3; c++)
      last.j[c] = RedJob{br.part + (size_t)c * nb, res + o_l0 + (size_t)c * R + t, nb, 1};
    last.n = 3;
  }
  mark(10, sb);
  if (t0 < R) {
    BfTail bt{};
    for (int c = 0; c < 3; c++) {
      bt.m_in[c] = mbuf[c][t0 & 1];
      bt.eq[c] = eqr[c][t0 & 1];
      bt.prev[c] = last.j[c];
    }
    bt.cw_in = levels + lvl_off[t0];
    bt.lvl_out = levels + lvl_off[t0 + 1];
    bt.l0 = res + o_l0;
    bt.chal = B + o_chbatch;
    bt.tw = ctx.twiddle((int)logD);
    bt.R = R;
    bt.t0 = t0;
    bt.logD = logD;
    hip_check(launch_bf_tail(bt, sb), "batch tail rounds");
  } else {
    hip_check(launch_reduce_jobs(last, sb), "batch round sums");
  }
  // trees over the folded levels 1..R-1 (leaf = one F_p^2 value), built together
  std::vector<DevTree>& ftrees = P.ftrees;
  ftrees.resize(R > 1 ? R - 1 : 0);
  std::vector<MerkleSource> fsrc;
  for (uint32_t k = 1; k < R; k++) {
    DevTree& tr = ftrees[k - 1];
    tr.leaves = D >> k;
    tr.nodes = DevBuf<uint32_t>(pool, tr.node_count() * 8);
    fsrc.push_back(MerkleSource{(const uint64_t*)(levels + lvl_off[k]), 2, 1, 2, tr.leaves, tr.nodes.ptr});
  }
  std::vector<std::vector<MerkleJob>>& fjobs = P.fjobs;
  std::vector<DevBuf<uint64_t>>& ftables = P.ftables;
  enqueue_merkle_trees(ctx, fsrc, fjobs, ftables, sb, &P.arena);
  mark(3, sb);
  hip_check(hipEventRecord(ctx.ev[6], sb), "join");
  // the openings read Stage B's levels and trees
  hip_check(hipStreamWaitEvent(st, ctx.ev[6], 0), "join");
  };

  // With the option prove_graph (proofs on the commitment's own context), the plan stays with
  // the commitment, the second proof captures its launches into a graph, and later proofs
  // refill the blob and idx slots that the graph's upload nodes read and replay it. Otherwise
  // the steps are issued one by one into a plan that lives for this call.
  const bool own_ctx = cm.ctx == &ctx, graphs = own_ctx && ctx.prove_graph && !ctx.profile;
  std::shared_ptr<ProvePlan> plan = own_ctx ? cm.plan : nullptr;
  if (graphs && plan && plan->exec) {
    std::memcpy(plan->blob_host, blob.data(), blob.size() * sizeof(ef2));
    std::memcpy(plan->idx_host, idx.data(), s * 4);
    hip_check(hipGraphLaunch(plan->exec, st), "prove graph launch");
  } else {
    auto np = std::make_shared<ProvePlan>();
    np->pool = pool;
    // staging memory: the last kept plan's, else the context's spare (pinning new host memory
    // on every proof would cost more than the proof's uploads)
    np->arena.take(plan ? plan->arena : ctx.prove_arena);
    if (plan) {  // the last plan's buffers go back to the pool (its proof has completed)
      if (own_ctx) cm.plan.reset();
      plan.reset();
    }
    const bool capture = graphs && cm.prove_calls >= 1;
    if (capture) hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed), "prove capture");
    try {
      issue(*np);
    } catch (...) {
      if (capture) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(st, &g);
        if (g) (void)hipGraphDestroy(g);
      }
      throw;
    }
    if (capture) {
      hip_check(hipStreamEndCapture(st, &np->graph), "prove capture end");
      hip_check(hipGraphInstantiate(&np->exec, np->graph, nullptr, nullptr, 0), "prove graph instantiate");
      hip_check(hipGraphLaunch(np->exec, st), "prove graph launch");
    }
    plan = np;
    if (graphs) cm.plan = np;
  }
  if (own_ctx) cm.prove_calls++;
  ProvePlan& P = *plan;
  const DevBasefold* states[3] = {&key.a_state, &P.bi_state, &P.av_state};

  // Opening positions are fixed by the integer stream, so every index the proof opens is
  // computed here on the host while the device works through both stages.
  std::vector<uint64_t> base;
  for (uint64_t q = 0; q < 100; q++) base.push_back(ints[4 * s + q] % (D / 2));
  std::sort(base.begin(), base.end());
  base.erase(std::unique(base.begin(), base.end()), base.end());
  const std::vector<uint64_t> bpos = level_positions(base, D / 2);
  const std::vector<uint64_t> hnodes = multiproof_nodes(two_n, sorted);
  const size_t n_hopen = hnodes.size();
  const size_t nf = R > 1 ? R - 1 : 0;
  std::vector<std::vector<uint64_t>> fpos(nf), fnodes(nf);
  for (uint32_t k = 1; k < R; k++) {
    const uint64_t len = D >> k;
    fpos[k - 1] = level_positions(base, len / 2);
    fnodes[k - 1] = multiproof_nodes(len, fpos[k - 1]);
  }
  const std::vector<uint64_t> bnodes = multiproof_nodes(D, bpos);
  const size_t n_bproof = bnodes.size();
  // ---- openings: the results first, then every opened word, in one gather ----
  const uint64_t* lv = (const uint64_t*)P.levels;  // level k's entries from word 2 lvl_off[k]
  Gather g;
  const uint64_t g_res = g.add_range((const uint64_t*)P.res, 2 * res_n);
  const uint64_t g_hopen = g.add_nodes(cm.tree, hnodes);
  const uint64_t g_hcols = g.add(cm.h.ptr, sorted, 1, 1, 32, two_n);  // H[r][j], r inner
  const uint64_t g_com_av = g.add_root(P.av_state.tree), g_com_i = g.add_root(P.bi_state.tree);
  std::vector<uint64_t> g_froot(nf), g_fproof(nf), g_fval(nf);
  for (uint32_t k = 1; k < R; k++) g_froot[k - 1] = g.add_root(P.ftrees[k - 1]);
  const uint64_t g_final = g.add_range(lv + 2 * lvl_off[R], 2 * (D >> R));
  for (uint32_t k = 1; k < R; k++) {
    g_fproof[k - 1] = g.add_nodes(P.ftrees[k - 1], fnodes[k - 1]);
    g_fval[k - 1] = g.add(lv + 2 * lvl_off[k], fpos[k - 1], 2, 2);
  }
  uint64_t g_bproof[3], g_bval[3];
  for (int c = 0; c < 3; c++) {
    g_bproof[c] = g.add_nodes(states[c]->tree, bnodes);
    g_bval[c] = g.add(states[c]->codes.ptr, bpos, 1, 1, states[c]->code_rows(), D);  // rows inner
  }

  // one upload: the part table, then the positions
  const size_t stage_words = g.total;
  const size_t tab_words = g.parts.size() * (sizeof(GatherPart) / 8), pos_words = (g.pos.size() + 1) / 2;
  g.pos.resize(2 * pos_words);
  std::vector<uint64_t> gtab(tab_words + pos_words);
  std::memcpy(gtab.data(), g.parts.data(), tab_words * 8);
  std::memcpy(gtab.data() + tab_words, g.pos.data(), pos_words * 8);
  DevBuf<uint64_t> dstage(pool, stage_words);
  DevBuf<uint64_t> dgtab(pool, gtab.size());
  ctx.upload(dgtab.ptr, gtab.data(), gtab.size() * 8, st);
  hip_check(launch_gather_parts((const GatherPart*)dgtab.ptr, (uint32_t)g.parts.size(),
                                (const uint32_t*)(dgtab.ptr + tab_words), stage_words, dstage.ptr, st),
            "gather");
  mark(4, st);
  uint64_t* host = (uint64_t*)ctx.staging(stage_words * 8);
  hip_check(hipMemcpyAsync(host, dstage.ptr, stage_words * 8, hipMemcpyDeviceToHost, st), "download proof");
  mark(5, st);
  const auto t_enqueued = std::chrono::steady_clock::now();
  hip_check(hipStreamSynchronize(st), "prove sync");
  if (!cm.plan || cm.plan.get() != &P) ctx.prove_arena.take(P.arena);  // a plan of this call only
  if (ctx.profile) {  // host time spent issuing work, and the whole device-bound part
    ctx.stage_ms[5] = std::chrono::duration<double, std::milli>(t_enqueued - t_start).count();
    ctx.stage_ms[6] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    float ms = 0;
    for (int i = 1; i <= 10; i++) {
      hip_check(hipEventElapsedTime(&ms, ctx.timing_event(32), ctx.timing_event(32 + i)), "profile time");
      ctx.stage_ms[6 + i] = ms;
    }
  }
  const uint64_t* RS = host + g_res;  // results as words
  const uint64_t* GW = host;          // gathered words

  // ---- encode (field order of LigesisProof, ligesis.rs:99-119) ----
  if (out.vec && out.vec->size() < 8 * (stage_words + 4096)) out.vec->resize(8 * (stage_words + 4096));
  Writer w(out);
  w.ef(RS + 2 * o_y);
  w.raw(GW + g_com_av, 4);
  w.raw(GW + g_com_i, 4);
  w.ef(RS + 2 * o_y4);
  w.ef(RS + 2 * (sc[3].finals_at + 1));  // y5 = ra(r3)
  w.u64(s);
  for (uint64_t k = 0; k < s; k++) {
    w.u64(32);
    w.words(GW + g_hcols + 32 * k, 32);
  }
  w.u64(32 * n_hopen);
  w.raw(GW + g_hopen, 4 * n_hopen);
  for (int i = 0; i < 7; i++) {
    w.u64(sc[i].nv);
    for (uint32_t t = 0; t < sc[i].nv; t++) {
      w.u64(sc[i].d + 1);
      w.words(RS + 2 * (sc[i].rounds_at + (size_t)t * (sc[i].d + 1)), 2 * (sc[i].d + 1));
    }
    w.u64(sc[i].d);
    w.words(RS + 2 * sc[i].finals_at, 2 * sc[i].d);
  }
  // BatchProof (basefold.rs:151-162)
  w.u64(3);
  for (int c = 0; c < 3; c++) {
    w.u64(R);
    w.words(RS + 2 * (o_l0 + (size_t)c * R), 2 * R);
  }
  w.u64(R > 1 ? R - 1 : 0);
  for (uint32_t k = 1; k < R; k++) w.raw(GW + g_froot[k - 1], 4);
  w.u64(D >> R);
  w.words(GW + g_final, 2 * (D >> R));
  w.u64(R > 1 ? R - 1 : 0);
  for (uint32_t k = 1; k < R; k++) {
    w.u64(32 * fnodes[k - 1].size());
    w.raw(GW + g_fproof[k - 1], 4 * fnodes[k - 1].size());
    w.u64(fpos[k - 1].size());
    for (size_t q = 0; q < fpos[k - 1].size(); q++) {
      w.u64(fpos[k - 1][q]);
      w.u64(1);
      w.words(GW + g_fval[k - 1] + 2 * q, 2);
    }
  }
  for (int grp = 0; grp < 2; grp++) {  // base_bf (A-bar, B_I), then base_ef (a-bar)
    int c0 = grp == 0 ? 0 : 2, c1 = grp == 0 ? 2 : 3;
    w.u64(c1 - c0);
    for (int c = c0; c < c1; c++) {
      w.u64(32 * n_bproof);
      w.raw(GW + g_bproof[c], 4 * n_bproof);
      w.u64(bpos.size());
      const uint32_t rows = states[c]->code_rows();
      const uint32_t elems = states[c]->ext ? rows / 2 : rows;
      for (size_t q = 0; q < bpos.size(); q++) {
        w.u64(bpos[q]);
        w.u64(elems);
        w.words(GW + g_bval[c] + (size_t)rows * q, rows);
      }
    }
  }
  w.u64(1);  // code_rate
  w.u64(ilog2(D));
  if (out.vec) out.vec->resize(w.n);
  return (int64_t)w.n;
}

}  // namespace lgs
