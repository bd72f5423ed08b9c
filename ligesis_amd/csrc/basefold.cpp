// Standalone Basefold opening of one committed polynomial on the GPU (Basefold::prove,
// pcs/src/basefold.rs:384-516). The transcript is fixed before any launch: round t takes the
// t-th field of the randomness oracle and the queries take its first query_num(code_rate)
// integers (utils/src/oracle.rs:32-55), so the rounds, the folded levels' trees and the openings
// are enqueued back to back on the context stream and the proof is downloaded once.
#include <cstring>
#include <vector>

#include "prove_kernels.hpp"
#include "prove_util.hpp"
#include "runtime.hpp"

namespace lgs {

using namespace pu;

int64_t basefold_prove_impl(Context& ctx, const DevBasefold& bf, const uint64_t* point, uint64_t point_len,
                            const uint64_t* fields, uint64_t n_fields, const uint64_t* ints, uint64_t n_ints,
                            ProofOut out, uint64_t used[2]) {
  const uint32_t nv = bf.num_vars, li = bf.li, cr = bf.code_rate;
  if (point_len != nv) throw Error(1, "basefold prove: the point must have num_vars coordinates");
  const uint32_t R = nv - li;
  const uint64_t D = bf.D;  // 2^(R + code_rate)
  const uint32_t logD = (uint32_t)ilog2(D);
  const uint64_t M = uint64_t(1) << R;
  // transcript (basefold.rs:459, 487-494): one field per reduction round, then the queries
  if (n_fields < R) throw Error(5, "basefold prove: randomness oracle field stream exhausted");
  const uint64_t q = (100 + cr - 1) / cr;  // query_num, basefold.rs:107-109
  if (n_ints < q) throw Error(5, "basefold prove: randomness oracle integer stream exhausted");
  used[0] = R;
  used[1] = q;
  EfV pt(nv), chal(R);
  for (uint32_t i = 0; i < nv; i++) pt[i] = ef2{point[2 * i], point[2 * i + 1]};
  for (uint32_t t = 0; t < R; t++) chal[t] = ef2{fields[2 * t], fields[2 * t + 1]};
  const EfV eq4 = host_eq(EfV(pt.begin(), pt.begin() + li));
  const EfV zp(pt.begin() + li, pt.end());
  EfV blob;
  auto put = [&](const EfV& v) {
    const size_t at = blob.size();
    blob.insert(blob.end(), v.begin(), v.end());
    return at;
  };
  const size_t o_eq4 = put(eq4), o_chal = put(chal);
  const size_t o_zp1 = put(R ? EfV(zp.begin() + 1, zp.end()) : EfV{});
  if (blob.empty()) blob.push_back(ef_zero());

  std::vector<uint64_t> lvl_off(R + 1);
  uint64_t tot = 0;
  for (uint32_t k = 0; k <= R; k++) {
    lvl_off[k] = tot;
    tot += D >> k;
  }
  hipStream_t st = ctx.stream;
  DevicePool* pool = &ctx.pool;
  ctx.upload_reset();  // every earlier call on this context has been synchronized
  for (uint32_t t = 0; t <= R; t++) (void)ctx.twiddle(ilog2(D >> t));
  DevBuf<ef2> dblob(pool, blob.size());
  ctx.upload(dblob.ptr, blob.data(), blob.size() * sizeof(ef2), st);
  const ef2* B = dblob.ptr;

  DevBuf<ef2> levels(pool, tot), res(pool, 2 * (size_t)R + 1);
  // m and eq tables, double-buffered by round parity (one extra entry: the tail reads none past
  // the end, but the buffers stay valid for any launch shape)
  DevBuf<ef2> mb0(pool, M + 1), mb1(pool, M / 2 + 1), eb0(pool, M / 2 + 1), eb1(pool, M / 4 + 1);
  ef2* mbuf[2] = {mb0.ptr, mb1.ptr};
  ef2* ebuf[2] = {eb0.ptr, eb1.ptr};
  DevBuf<ef2> part(pool, 2 * 2 * 1024);  // l0 / l1 partials, alternating halves by round parity

  // eq table of the residual point's tail zp[1..R) for round 0 (MlPoly::new_eq, poly.rs:44-53)
  DevBuf<ef2> halves;
  DevBuf<uint8_t> eqjobs;
  if (R >= 1) {
    const uint32_t k = R - 1, klo = k / 2, khi = k - klo;
    halves = DevBuf<ef2>(pool, (size_t(1) << klo) + (size_t(1) << khi));
    EqTerm term{};
    term.pt = B + o_zp1;
    term.scale = nullptr;
    term.lo = halves.ptr;
    term.hi = halves.ptr + (size_t(1) << klo);
    term.k = k;
    term.klo = klo;
    term.first_block = 0;
    EqTable tab{};
    tab.out = ebuf[0];
    tab.n = uint64_t(1) << k;
    tab.lo[0] = term.lo;
    tab.hi[0] = term.hi;
    tab.klo = klo;
    tab.nterms = 1;
    tab.first_block = 0;
    std::vector<uint8_t> h(sizeof(EqTerm) + sizeof(EqTable));
    std::memcpy(h.data(), &term, sizeof(EqTerm));
    std::memcpy(h.data() + sizeof(EqTerm), &tab, sizeof(EqTable));
    eqjobs = DevBuf<uint8_t>(pool, h.size());
    ctx.upload(eqjobs.ptr, h.data(), h.size(), st);
    const uint32_t term_blocks = (uint32_t)(((uint64_t(1) << std::max(klo, khi)) + 255) / 256);
    const uint32_t tab_blocks = (uint32_t)((tab.n + 255) / 256);
    hip_check(launch_eq_group((const EqTerm*)eqjobs.ptr, 1, term_blocks,
                              (const EqTable*)(eqjobs.ptr + sizeof(EqTerm)), 1, tab_blocks, st),
              "basefold eq table");
  }
  // level 0: the eq-compressed codewords (basefold.rs:405-416); m = poly folded over its low li
  // variables (basefold.rs:418-421)
  BfChannels ch{};
  ch.count = 1;
  ch.codes[0] = bf.codes.ptr;
  ch.wt[0] = B + o_eq4;
  ch.chunks[0] = 1u << li;
  ch.ext[0] = bf.ext ? 1 : 0;
  hip_check(launch_bf_combine(ch, D, levels.ptr, st), "basefold combine");
  hip_check(launch_fold_low(bf.poly_ptr, bf.ext ? 1 : 0, (int)li, M, B + o_eq4, mbuf[0], st), "basefold fold");
  // reduction rounds (basefold.rs:423-467): on the grid while the level has more than
  // 2^kBfTailLog entries, then the rest in one workgroup
  const uint32_t t0 = logD > (uint32_t)kBfTailLog ? std::min<uint32_t>(R, logD - kBfTailLog) : 0;
  RedJob last{};
  for (uint32_t t = 0; t < t0; t++) {
    const uint64_t half = uint64_t(1) << (R - 1 - t), len = D >> t;
    Bf1Round br{};
    br.m_in = mbuf[t & 1];
    br.m_out = mbuf[(t + 1) & 1];
    br.eq = ebuf[t & 1];
    br.eq_next = ebuf[(t + 1) & 1];
    br.part = part.ptr + (size_t)(t & 1) * 2 * 1024;
    br.prev = last;
    br.cw_in = levels.ptr + lvl_off[t];
    br.cw_half = len / 2;
    br.cw_tw = ctx.twiddle(ilog2(len));
    br.cw_out = levels.ptr + lvl_off[t + 1];
    hip_check(launch_bf1_round(br, half, B + o_chal + t, st), "basefold round");
    last = RedJob{br.part, res.ptr + 2 * (size_t)t, bf1_round_blocks(half, br.cw_half), 2};
  }
  if (t0 < R) {
    Bf1Tail bt{};
    bt.m_in = mbuf[t0 & 1];
    bt.eq = ebuf[t0 & 1];
    bt.prev = last;
    bt.cw_in = levels.ptr + lvl_off[t0];
    bt.lvl_out = levels.ptr + lvl_off[t0 + 1];
    bt.lpair = res.ptr;
    bt.chal = B + o_chal;
    bt.tw = ctx.twiddle((int)logD);
    bt.R = R;
    bt.t0 = t0;
    bt.logD = logD;
    hip_check(launch_bf1_tail(bt, st), "basefold tail rounds");
  } else if (last.nblk) {
    RedJobs rj{};
    rj.j[0] = last;
    rj.n = 1;
    hip_check(launch_reduce_jobs(rj, st), "basefold round sums");
  }
  // trees over the folded levels 1..R-1 (leaf = one F_p^2 value, basefold.rs:471-481)
  std::vector<DevTree> ftrees(R > 1 ? R - 1 : 0);
  std::vector<MerkleSource> fsrc;
  for (uint32_t k = 1; k < R; k++) {
    DevTree& tr = ftrees[k - 1];
    tr.leaves = D >> k;
    tr.nodes = DevBuf<uint32_t>(pool, tr.node_count() * 8);
    fsrc.push_back(MerkleSource{(const uint64_t*)(levels.ptr + lvl_off[k]), 2, 1, 2, tr.leaves, tr.nodes.ptr});
  }
  std::vector<std::vector<MerkleJob>> fjobs;
  std::vector<DevBuf<uint64_t>> ftables;
  enqueue_merkle_trees(ctx, fsrc, fjobs, ftables, st);

  // queries (basefold.rs:483-494) and the positions every opening reveals
  std::vector<uint64_t> base;
  for (uint64_t k = 0; k < q; k++) base.push_back(ints[k] % (D / 2));
  std::sort(base.begin(), base.end());
  base.erase(std::unique(base.begin(), base.end()), base.end());
  const std::vector<uint64_t> bpos = level_positions(base, D / 2);
  const std::vector<uint64_t> bnodes = multiproof_nodes(D, bpos);
  const size_t nf = R > 1 ? R - 1 : 0;
  std::vector<std::vector<uint64_t>> fpos(nf), fnodes(nf);
  for (uint32_t k = 1; k < R; k++) {
    fpos[k - 1] = level_positions(base, (D >> k) / 2);
    fnodes[k - 1] = multiproof_nodes(D >> k, fpos[k - 1]);
  }
  // ---- openings: the round sums first, then every opened word, in one gather ----
  const uint64_t* lv = (const uint64_t*)levels.ptr;  // level k's entries from word 2 lvl_off[k]
  Gather g;
  // (no part may be empty: the gather finds a word's part by its first word)
  const uint64_t g_res = R ? g.add_range((const uint64_t*)res.ptr, 4 * (uint64_t)R) : 0;
  const uint64_t g_m0 = R ? 0 : g.add_range((const uint64_t*)mbuf[0], 2);  // the claim when R = 0
  std::vector<uint64_t> g_froot(nf), g_fproof(nf), g_fval(nf);
  for (uint32_t k = 1; k < R; k++) g_froot[k - 1] = g.add_root(ftrees[k - 1]);
  const uint64_t g_final = g.add_range(lv + 2 * lvl_off[R], 2 * (D >> R));
  for (uint32_t k = 1; k < R; k++) {
    g_fproof[k - 1] = g.add_nodes(ftrees[k - 1], fnodes[k - 1]);
    g_fval[k - 1] = g.add(lv + 2 * lvl_off[k], fpos[k - 1], 2, 2);
  }
  const uint64_t g_bproof = g.add_nodes(bf.tree, bnodes);
  const uint32_t rows = bf.code_rows();
  const uint64_t g_bval = g.add(bf.codes.ptr, bpos, 1, 1, rows, D);  // rows inner
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
            "basefold gather");
  uint64_t* host = (uint64_t*)ctx.staging(stage_words * 8);
  hip_check(hipMemcpyAsync(host, dstage.ptr, stage_words * 8, hipMemcpyDeviceToHost, st), "download proof");
  hip_check(hipStreamSynchronize(st), "basefold prove sync");
  const uint64_t* GW = host;

  // claimed_eval = f(point) = (1 - zp_0) l0 + zp_0 l1 of round 0 (the first fold of m at zp_0,
  // by linearity of the evaluation); with no rounds, m is the single value f(point)
  ef2 claimed;
  if (R == 0) {
    claimed = ef2{GW[g_m0], GW[g_m0 + 1]};
  } else {
    const ef2 l0{GW[g_res], GW[g_res + 1]}, l1{GW[g_res + 2], GW[g_res + 3]};
    claimed = ef_add(ef_mul(ef_sub(ef_one(), zp[0]), l0), ef_mul(zp[0], l1));
  }

  // ---- encode (field order of BasefoldProof, basefold.rs:84-103) ----
  Writer w(out);
  w.u64(R);
  w.words(GW + g_res, 4 * (size_t)R);  // (l0, l1) per round
  w.u64(nf);
  for (uint32_t k = 1; k < R; k++) w.raw(GW + g_froot[k - 1], 4);
  w.u64(D >> R);
  w.words(GW + g_final, 2 * (D >> R));
  w.u64(32 * bnodes.size());
  w.raw(GW + g_bproof, 4 * bnodes.size());
  w.u64(bpos.size());
  const uint32_t elems = bf.ext ? rows / 2 : rows;
  for (size_t k = 0; k < bpos.size(); k++) {
    w.u64(bpos[k]);
    w.u64(elems);
    w.words(GW + g_bval + (size_t)rows * k, rows);
  }
  w.u64(nf);
  for (uint32_t k = 1; k < R; k++) {
    w.u64(32 * fnodes[k - 1].size());
    w.raw(GW + g_fproof[k - 1], 4 * fnodes[k - 1].size());
    w.u64(fpos[k - 1].size());
    for (size_t i = 0; i < fpos[k - 1].size(); i++) {
      w.u64(fpos[k - 1][i]);
      w.u64(1);
      w.words(GW + g_fval[k - 1] + 2 * i, 2);
    }
  }
  const uint64_t cw[2] = {claimed.c0, claimed.c1};
  w.ef(cw);
  w.u64(nv);
  w.u64(cr);
  w.u64(li);
  if (out.vec) out.vec->resize(w.n);
  return (int64_t)w.n;
}

}  // namespace lgs
