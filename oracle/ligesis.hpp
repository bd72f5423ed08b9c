// Oracle (test infrastructure only): LigeSIS setup / commit / prove / verify.
//
// Restates pcs/src/ligesis.rs (constants :53-59, combine_eq :159-171, committed_vars /
// batch_rounds / shared_log_interleave :175-203, rs_encode :225-235, setup :239-273,
// commit :277-325, sample_columns :329-345, prove :347-549, rs_weights :553-567,
// verify :569-756, horner :760, combine_eq_eval :769) and pcs/src/subset_sum.rs:31-79.
// The loop structure of commit (per-row DFT, one byte per bit in column vectors, per-set-bit
// masked add over a[r][i]) follows the reference so that it doubles as the timed CPU baseline.
// Pinned by: prove/verify round trips (ligesis.rs:807-822), tamper rejection
// (ligesis.rs:824-868), y == f(z) by direct evaluation, and the reference's debug-build
// identities (ligesis.rs:461-471, 477, 492, 501), all checked in tests/test_oracle_ligesis.py.
//
// The Z_2^64 commitment (ring64 keys) is a paper-level variant the reference leaves out
// (subset_sum.rs:10-12): the same commit with H = A . x summed mod 2^64 over any 64-bit A (its
// words are serialized as they are); restated from that definition and pinned against big-integer
// sums mod 2^64 in tests/test_oracle_subset_sum.py.
//
// Proof bytes: the reference has no serializer (only size_bytes, ligesis.rs:124-138), so the
// canonical encoding is fixed here: struct fields in declaration order; u64 little-endian for
// integers and lengths; BF as canonical u64; EF as (c0, c1); [u8;32] raw; Vec<T> = u64 length
// + elements; Vec<u8> = u64 length + bytes; BTreeMap<usize, Vec<T>> = u64 count + (key, Vec<T>)
// in ascending key order.
#pragma once
#include "basefold.hpp"

namespace lgso {

constexpr size_t ETA = 64, C = 32, NUM_COLS_OPENED = 128, LOG_INTERLEAVE = 6, CODE_RATE = 1;

inline void committed_vars(size_t log_m, size_t log_n, size_t out[3]) {
  out[0] = 5 + 6 + log_m;  // A-bar : C * eta * m
  out[1] = 6 + log_m + 7;  // B_I   : eta * m * s
  out[2] = log_n;          // a-bar : n
}
inline size_t max_vars(size_t lm, size_t ln) {
  size_t v[3];
  committed_vars(lm, ln, v);
  return std::max(v[0], std::max(v[1], v[2]));
}
inline size_t min_vars(size_t lm, size_t ln) {
  size_t v[3];
  committed_vars(lm, ln, v);
  return std::min(v[0], std::min(v[1], v[2]));
}
inline bool shape_ok(size_t lm, size_t ln) { return max_vars(lm, ln) - min_vars(lm, ln) <= LOG_INTERLEAVE; }
inline size_t batch_rounds(size_t lm, size_t ln) { return max_vars(lm, ln) - LOG_INTERLEAVE; }
inline size_t shared_li(size_t nv, size_t lm, size_t ln) { return nv - batch_rounds(lm, ln); }

struct ProvingKey {
  std::vector<std::vector<uint64_t>> a;  // C rows of width eta*m
  size_t width = 0, log_m = 0, log_n = 0;
  BfState<uint64_t> a_state;
  H32 com_a{};
  bool has_state = false;
  bool ring64 = false;  // the Z_2^64 hash (commit only)
};

struct CommitData {
  std::vector<uint64_t> f;
  std::vector<std::vector<uint8_t>> b_cols;  // 2n columns of eta*m bits, one byte per bit
  std::vector<std::vector<uint64_t>> h;      // C x 2n
  MerkleTree mt;
  std::vector<uint64_t> enc;  // F' (m x 2n) when requested, for stage-by-stage comparison
};

inline std::vector<uint64_t> a_mle_flat(const ProvingKey& pk) {
  std::vector<uint64_t> flat;
  for (auto& row : pk.a) flat.insert(flat.end(), row.begin(), row.end());
  return flat;
}

// `setup` with the hash matrix A supplied (row-major C x eta*m, canonical) instead of sampled.
// `commit_a = false` builds only the hash (enough for commit; the reference commits A-bar too).
inline ProvingKey setup(size_t log_m, size_t log_n, const uint64_t* a, bool commit_a) {
  ProvingKey pk;
  pk.log_m = log_m;
  pk.log_n = log_n;
  pk.width = ETA << log_m;
  for (size_t r = 0; r < C; r++)
    pk.a.emplace_back(a + r * pk.width, a + (r + 1) * pk.width);
  if (commit_a) {
    if (!shape_ok(log_m, log_n)) throw std::runtime_error("setup: shape outside the shared-domain window");
    size_t a_vars = log2_exact(C * pk.width);
    pk.com_a = commit_on_domain(a_mle_flat(pk), CODE_RATE, shared_li(a_vars, log_m, log_n), pk.a_state);
    pk.has_state = true;
  }
  return pk;
}

// subset_sum.rs:43-54. Like p3's Goldilocks addition, the running sums stay in [0, 2^64)
// without canonicalisation (a carry out of 2^64 folds back as 2^32 - 1); they are made
// canonical once at the end, which gives the same residues.
inline uint64_t add_noncanon(uint64_t a, uint64_t b) {
  uint64_t s = a + b;
  uint64_t c = (uint64_t)(s < b) * EPS;  // branch-free: carries are data-dependent
  uint64_t t = s + c;
  if (t < c) t += EPS;  // second carry (rare)
  return t;
}

inline std::vector<uint64_t> hash_column(const ProvingKey& pk, const std::vector<uint8_t>& x) {
  uint64_t out[C] = {0};
  const uint64_t* rows[C];
  for (size_t r = 0; r < C; r++) rows[r] = pk.a[r].data();
  for (size_t i = 0; i < x.size(); i++)
    if (x[i]) {
      if (pk.ring64)
        for (size_t r = 0; r < C; r++) out[r] += rows[r][i];  // mod 2^64
      else
        for (size_t r = 0; r < C; r++) out[r] = add_noncanon(out[r], rows[r][i]);
    }
  std::vector<uint64_t> res(C);
  for (size_t r = 0; r < C; r++) res[r] = pk.ring64 ? out[r] : canon(out[r]);
  return res;
}

// A column of H as the leaf bytes: canonical field elements, or (Z_2^64 hash) the raw words.
inline std::vector<uint8_t> ser_h_col(const std::vector<std::vector<uint64_t>>& h, size_t j, bool ring64 = false) {
  std::vector<uint8_t> b;
  for (size_t r = 0; r < C; r++) {
    if (ring64)
      put_u64(b, h[r][j]);
    else
      ser_field(b, h[r][j]);
  }
  return b;
}

inline H32 commit(const ProvingKey& pk, std::vector<uint64_t> f, CommitData& cd, bool keep_enc) {
  size_t m = size_t(1) << pk.log_m, n = size_t(1) << pk.log_n, two_n = 2 * n;
  if (f.size() != m * n) throw std::runtime_error("commit: length");
  cd.b_cols.assign(two_n, std::vector<uint8_t>(ETA * m, 0));
  if (keep_enc) cd.enc.assign(m * two_n, 0);
  std::vector<uint64_t> row(two_n);
  for (size_t i = 0; i < m; i++) {
    std::fill(row.begin(), row.end(), 0);
    std::copy(f.begin() + i * n, f.begin() + (i + 1) * n, row.begin());
    dft_base(row);
    for (size_t j = 0; j < two_n; j++) {
      uint64_t v = row[j];
      if (keep_enc) cd.enc[i * two_n + j] = v;
      for (size_t k = 0; k < ETA; k++)
        if ((v >> k) & 1) cd.b_cols[j][ETA * i + k] = 1;
    }
  }
  cd.h.assign(C, std::vector<uint64_t>(two_n, 0));
  for (size_t j = 0; j < two_n; j++) {
    auto d = hash_column(pk, cd.b_cols[j]);
    for (size_t r = 0; r < C; r++) cd.h[r][j] = d[r];
  }
  std::vector<H32> leaves(two_n);
  for (size_t j = 0; j < two_n; j++) leaves[j] = blake3(ser_h_col(cd.h, j, pk.ring64));
  cd.mt = MerkleTree::from_leaf_hashes(std::move(leaves));
  cd.f = std::move(f);
  return cd.mt.root();
}

inline std::vector<size_t> sample_columns(OracleStream& o, size_t two_n, size_t s) {
  auto raw = o.next_n_ints(s * 4);
  std::vector<uint8_t> seen(two_n, 0);
  std::vector<size_t> idx;
  for (uint64_t v : raw) {
    size_t j = v % two_n;
    if (!seen[j]) {
      seen[j] = 1;
      idx.push_back(j);
      if (idx.size() == s) break;
    }
  }
  if (idx.size() != s) throw std::runtime_error("not enough distinct columns sampled from oracle");
  return idx;
}

inline EfVec combine_eq(const std::vector<EfVec>& pts, Ef gamma) {
  size_t len = size_t(1) << pts[0].size();
  EfVec out(len, ef_zero());
  Ef g = ef_one();
  for (auto& p : pts) {
    EfVec e = new_eq(p);
    for (size_t i = 0; i < len; i++) out[i] = ef_add(out[i], ef_mul(g, e[i]));
    g = ef_mul(g, gamma);
  }
  return out;
}

inline Ef horner(const EfVec& vals, Ef gamma) {
  Ef acc = ef_zero();
  for (size_t k = vals.size(); k-- > 0;) acc = ef_add(ef_mul(acc, gamma), vals[k]);
  return acc;
}

inline Ef combine_eq_eval(const std::vector<EfVec>& pts, Ef gamma, const EfVec& r) {
  Ef acc = ef_zero(), g = ef_one();
  for (auto& p : pts) {
    acc = ef_add(acc, ef_mul(g, eval_eq(p, r)));
    g = ef_mul(g, gamma);
  }
  return acc;
}

inline EfVec rs_encode_ext(const EfVec& c, size_t len) {
  EfVec x = c;
  x.resize(2 * len, ef_zero());
  dft_ext(x);
  return x;
}

inline EfVec rs_weights(const EfVec& r3, size_t n, size_t two_n) {
  int log_2n = log2_exact(two_n);
  uint64_t g = two_adic_generator(log_2n);
  EfVec w(n);
  for (size_t k = 0; k < n; k++) {
    Ef prod = ef_one();
    for (int t = 0; t < log_2n; t++) {
      uint64_t e = (uint64_t)(((u128)(uint64_t(1) << t) * k) % two_n);
      Ef gp = ef(fpow(g, e));
      prod = ef_mul(prod, ef_add(ef_sub(ef_one(), r3[t]), ef_mul(r3[t], gp)));
    }
    w[k] = prod;
  }
  return w;
}

struct LigesisProof {
  Ef y{};
  H32 com_a_vec{}, com_i{};
  Ef y4{}, y5{};
  std::vector<std::vector<uint64_t>> h_cols;
  std::vector<uint8_t> h_open;
  SumcheckProof sc_binary, sc_hash, sc_recon, sc_table, sc_rslink, sc_bi_batch, sc_a_batch;
  BatchProof batch;
};

inline std::vector<uint64_t> pow2_table() {
  std::vector<uint64_t> p(ETA);
  uint64_t c = 1;
  for (size_t k = 0; k < ETA; k++) {
    p[k] = c;
    c = fadd(c, c);
  }
  return p;
}

// Intermediates kept by prove() so stages can be compared one by one.
struct ProveTrace {
  EfVec a_vec, ra;
  std::vector<size_t> idx;
};

inline LigesisProof prove(const ProvingKey& pk, const CommitData& cd, const EfVec& z,
                          OracleStream& o, ProveTrace* trace = nullptr) {
  if (!pk.has_state) throw std::runtime_error("prove: key has no A-bar commitment");
  size_t log_m = pk.log_m, log_n = pk.log_n, m = size_t(1) << log_m, n = size_t(1) << log_n;
  size_t two_n = 2 * n, width = pk.width, s = NUM_COLS_OPENED, log_s = 7;
  size_t log_w = log2_exact(width);
  if (z.size() != log_m + log_n) throw std::runtime_error("prove: point length");
  auto pow2 = pow2_table();
  EfVec z2(z.begin(), z.begin() + log_n), z1(z.begin() + log_n, z.end());
  EfVec eqz1 = new_eq(z1);
  EfVec a_vec(n, ef_zero());
  for (size_t i = 0; i < m; i++)
    for (size_t j = 0; j < n; j++) a_vec[j] = ef_add(a_vec[j], ef_mul_base(eqz1[i], cd.f[i * n + j]));
  LigesisProof pf;
  pf.y = mle_eval(a_vec, z2);
  BfState<Ef> av_state;
  pf.com_a_vec = commit_on_domain(a_vec, CODE_RATE, shared_li(log_n, log_m, log_n), av_state);
  EfVec ra = rs_encode_ext(a_vec, n);
  auto idx = sample_columns(o, two_n, s);
  auto sorted = idx;
  std::sort(sorted.begin(), sorted.end());
  pf.h_open = cd.mt.open(sorted);
  for (size_t j : sorted) {
    std::vector<uint64_t> col;
    for (size_t r = 0; r < C; r++) col.push_back(cd.h[r][j]);
    pf.h_cols.push_back(col);
  }
  std::vector<uint64_t> bi_flat(width * s, 0);
  for (size_t i = 0; i < width; i++)
    for (size_t k = 0; k < s; k++)
      if (cd.b_cols[idx[k]][i]) bi_flat[i * s + k] = 1;
  BfState<uint64_t> bi_state;
  size_t bi_vars = log2_exact(width * s);
  pf.com_i = commit_on_domain(bi_flat, CODE_RATE, shared_li(bi_vars, log_m, log_n), bi_state);

  EfVec alpha1 = o.next_n_fields(log_w + log_s);
  EfVec alpha2 = o.next_n_fields(5);
  EfVec eq_a2 = new_eq(alpha2);
  EfVec bi_ef(bi_flat.size()), bi_m1(bi_flat.size());
  for (size_t i = 0; i < bi_flat.size(); i++) {
    bi_ef[i] = ef(bi_flat[i]);
    bi_m1[i] = ef_sub(bi_ef[i], ef_one());
  }
  EfVec ch;
  pf.sc_binary = sumcheck_prove({bi_ef, bi_m1, new_eq(alpha1)}, o, ch);
  EfVec r1 = rev(ch);
  EfVec r2 = o.next_n_fields(log_s);
  EfVec eq_r2 = new_eq(r2);
  EfVec g1(width), g2(width);
  for (size_t i = 0; i < width; i++) {
    Ef acc = ef_zero();
    for (size_t r = 0; r < C; r++) acc = ef_add(acc, ef_mul_base(eq_a2[r], pk.a[r][i]));
    g1[i] = acc;
    Ef acc2 = ef_zero();
    for (size_t k = 0; k < s; k++)
      if (cd.b_cols[idx[k]][i]) acc2 = ef_add(acc2, eq_r2[k]);
    g2[i] = acc2;
  }
  Ef y4 = ef_zero();
  for (size_t k = 0; k < s; k++) y4 = ef_add(y4, ef_mul(eq_r2[k], ra[idx[k]]));
  pf.y4 = y4;
  {  // ligesis.rs:461-471 hash channel == bilinear form
    Ef lhs = ef_zero(), rhs = ef_zero();
    for (size_t k = 0; k < s; k++) {
      Ef t = ef_zero();
      for (size_t r = 0; r < C; r++) t = ef_add(t, ef_mul_base(eq_a2[r], cd.h[r][idx[k]]));
      lhs = ef_add(lhs, ef_mul(eq_r2[k], t));
    }
    for (size_t i = 0; i < width; i++) rhs = ef_add(rhs, ef_mul(g1[i], g2[i]));
    if (lhs != rhs) throw std::runtime_error("prove: hash-channel identity failed");
  }
  EfVec v_vec(width);
  for (size_t i = 0; i < width; i++) v_vec[i] = ef_mul_base(eqz1[i / ETA], pow2[i % ETA]);
  {  // ligesis.rs:477
    Ef t = ef_zero();
    for (size_t i = 0; i < width; i++) t = ef_add(t, ef_mul(v_vec[i], g2[i]));
    if (t != y4) throw std::runtime_error("prove: recon identity failed");
  }
  pf.sc_hash = sumcheck_prove({g1, g2}, o, ch);
  EfVec r4 = rev(ch);
  pf.sc_recon = sumcheck_prove({v_vec, g2}, o, ch);
  EfVec r5 = rev(ch);
  EfVec sel(two_n, ef_zero());
  for (size_t k = 0; k < s; k++) sel[idx[k]] = eq_r2[k];
  pf.sc_table = sumcheck_prove({sel, ra}, o, ch);
  EfVec r3 = rev(ch);
  pf.y5 = mle_eval(ra, r3);
  EfVec w_vec = rs_weights(r3, n, two_n);
  {  // ligesis.rs:501
    Ef t = ef_zero();
    for (size_t k = 0; k < n; k++) t = ef_add(t, ef_mul(a_vec[k], w_vec[k]));
    if (t != pf.y5) throw std::runtime_error("prove: rs-link identity failed");
  }
  pf.sc_rslink = sumcheck_prove({a_vec, w_vec}, o, ch);
  EfVec r6 = rev(ch);
  EfVec p_a = cat(r4, alpha2), p_bi_hash = cat(r2, r4), p_bi_recon = cat(r2, r5);
  Ef gamma_bi = o.next_field();
  EfVec g_bi = combine_eq({r1, p_bi_hash, p_bi_recon}, gamma_bi);
  pf.sc_bi_batch = sumcheck_prove({bi_ef, g_bi}, o, ch);
  EfVec r_bi = rev(ch);
  Ef gamma_a = o.next_field();
  EfVec g_a = combine_eq({z2, r6}, gamma_a);
  pf.sc_a_batch = sumcheck_prove({a_vec, g_a}, o, ch);
  EfVec r_a = rev(ch);
  pf.batch = batch_prove({{&pk.a_state, p_a}, {&bi_state, r_bi}}, {{&av_state, r_a}}, o);
  if (trace) {
    trace->a_vec = a_vec;
    trace->ra = ra;
    trace->idx = idx;
  }
  return pf;
}

struct VerifyingKey {
  H32 com_a{};
  size_t log_m = 0, log_n = 0, width = 0;
};

inline bool verify(const VerifyingKey& vk, const H32& commitment, const EfVec& z,
                   const LigesisProof& pf, OracleStream& o) {
  size_t log_m = vk.log_m, log_n = vk.log_n, n = size_t(1) << log_n;
  size_t two_n = 2 * n, width = vk.width, s = NUM_COLS_OPENED, log_s = 7;
  if (z.size() != log_m + log_n) return false;
  auto pow2 = pow2_table();
  EfVec z2(z.begin(), z.begin() + log_n), z1(z.begin() + log_n, z.end());
  EfVec eqz1 = new_eq(z1);
  auto idx = sample_columns(o, two_n, s);
  auto sorted = idx;
  std::sort(sorted.begin(), sorted.end());
  if (pf.h_cols.size() != sorted.size()) return false;
  for (auto& c : pf.h_cols)
    if (c.size() != C) return false;
  std::vector<H32> hl;
  for (auto& c : pf.h_cols) hl.push_back(blake3(serialize_fields(c)));
  if (!merkle_verify(commitment, two_n, pf.h_open, sorted, hl)) return false;
  EfVec alpha1 = o.next_n_fields(log2_exact(width) + log_s);
  EfVec alpha2 = o.next_n_fields(5);
  EfVec ch;
  if (!sumcheck_verify(ef_zero(), pf.sc_binary, o, ch)) return false;
  EfVec r1 = rev(ch);
  {
    auto& fe = pf.sc_binary.final_evals;
    if (fe.size() != 3 || fe[1] != ef_sub(fe[0], ef_one()) || fe[2] != eval_eq(alpha1, r1)) return false;
  }
  EfVec r2 = o.next_n_fields(log_s);
  EfVec eq_r2 = new_eq(r2), eq_a2 = new_eq(alpha2);
  Ef y2 = ef_zero();
  for (size_t k = 0; k < s; k++) {
    size_t pos = std::lower_bound(sorted.begin(), sorted.end(), idx[k]) - sorted.begin();
    const auto& col = pf.h_cols[pos];
    Ef t = ef_zero();
    for (size_t r = 0; r < C; r++) t = ef_add(t, ef_mul_base(eq_a2[r], col[r]));
    y2 = ef_add(y2, ef_mul(eq_r2[k], t));
  }
  if (!sumcheck_verify(y2, pf.sc_hash, o, ch)) return false;
  EfVec r4 = rev(ch);
  if (pf.sc_hash.final_evals.size() != 2) return false;
  if (!sumcheck_verify(pf.y4, pf.sc_recon, o, ch)) return false;
  EfVec r5 = rev(ch);
  EfVec v_vec(width);
  for (size_t i = 0; i < width; i++) v_vec[i] = ef_mul_base(eqz1[i / ETA], pow2[i % ETA]);
  {
    auto& fe = pf.sc_recon.final_evals;
    if (fe.size() != 2 || fe[0] != mle_eval(v_vec, r5)) return false;
  }
  if (!sumcheck_verify(pf.y4, pf.sc_table, o, ch)) return false;
  EfVec r3 = rev(ch);
  EfVec sel(two_n, ef_zero());
  for (size_t k = 0; k < s; k++) sel[idx[k]] = eq_r2[k];
  {
    auto& fe = pf.sc_table.final_evals;
    if (fe.size() != 2 || fe[0] != mle_eval(sel, r3) || fe[1] != pf.y5) return false;
  }
  if (!sumcheck_verify(pf.y5, pf.sc_rslink, o, ch)) return false;
  EfVec r6 = rev(ch);
  EfVec w_vec = rs_weights(r3, n, two_n);
  {
    auto& fe = pf.sc_rslink.final_evals;
    if (fe.size() != 2 || fe[1] != mle_eval(w_vec, r6)) return false;
  }
  EfVec p_a = cat(r4, alpha2), p_bi_hash = cat(r2, r4), p_bi_recon = cat(r2, r5);
  Ef gamma_bi = o.next_field();
  std::vector<EfVec> bi_pts{r1, p_bi_hash, p_bi_recon};
  if (pf.sc_binary.final_evals.size() != 3 || pf.sc_recon.final_evals.size() != 2) return false;
  Ef claim_bi = horner({pf.sc_binary.final_evals[0], pf.sc_hash.final_evals[1],
                        pf.sc_recon.final_evals[1]}, gamma_bi);
  if (!sumcheck_verify(claim_bi, pf.sc_bi_batch, o, ch)) return false;
  EfVec r_bi = rev(ch);
  if (pf.sc_bi_batch.final_evals.size() != 2 ||
      pf.sc_bi_batch.final_evals[1] != combine_eq_eval(bi_pts, gamma_bi, r_bi))
    return false;
  Ef v_bi = pf.sc_bi_batch.final_evals[0];
  Ef gamma_a = o.next_field();
  std::vector<EfVec> a_pts{z2, r6};
  Ef claim_a = horner({pf.y, pf.sc_rslink.final_evals[0]}, gamma_a);
  if (!sumcheck_verify(claim_a, pf.sc_a_batch, o, ch)) return false;
  EfVec r_a = rev(ch);
  if (pf.sc_a_batch.final_evals.size() != 2 ||
      pf.sc_a_batch.final_evals[1] != combine_eq_eval(a_pts, gamma_a, r_a))
    return false;
  Ef v_a = pf.sc_a_batch.final_evals[0];
  size_t nv_a = log2_exact(C * width), nv_bi = log2_exact(width * s);
  std::vector<BfClaim> bfc{{vk.com_a, p_a, pf.sc_hash.final_evals[0], shared_li(nv_a, log_m, log_n)},
                           {pf.com_i, r_bi, v_bi, shared_li(nv_bi, log_m, log_n)}};
  std::vector<BfClaim> efc{{pf.com_a_vec, r_a, v_a, shared_li(log_n, log_m, log_n)}};
  return batch_verify(bfc, efc, pf.batch, o);
}

// ---- canonical proof encoding -------------------------------------------------------------

struct Writer {
  std::vector<uint8_t> b;
  void u64(uint64_t v) { put_u64(b, v); }
  void bf(uint64_t v) { ser_field(b, v); }
  void e(const Ef& v) { ser_field(b, v); }
  void h32(const H32& h) { b.insert(b.end(), h.begin(), h.end()); }
  void bytes(const std::vector<uint8_t>& v) {
    u64(v.size());
    b.insert(b.end(), v.begin(), v.end());
  }
  void efvec(const EfVec& v) {
    u64(v.size());
    for (auto& x : v) e(x);
  }
  void bfvec(const std::vector<uint64_t>& v) {
    u64(v.size());
    for (auto x : v) bf(x);
  }
  void sc(const SumcheckProof& s) {
    u64(s.round_polys.size());
    for (auto& r : s.round_polys) efvec(r);
    efvec(s.final_evals);
  }
};

struct Reader {
  const uint8_t* p;
  size_t n, c = 0;
  bool ok = true;
  uint64_t u64() {
    if (c + 8 > n) {
      ok = false;
      return 0;
    }
    uint64_t v = 0;
    for (int i = 0; i < 8; i++) v |= (uint64_t)p[c + i] << (8 * i);
    c += 8;
    return v;
  }
  size_t len(size_t elem) {  // a length prefix that must fit in what is left
    uint64_t v = u64();
    if (!ok || (elem && v > (n - c) / elem)) {
      ok = false;
      return 0;
    }
    return (size_t)v;
  }
  uint64_t bf() {
    uint64_t v = u64();
    if (v >= P) ok = false;
    return v;
  }
  Ef e() {
    uint64_t a = bf(), b2 = bf();
    return Ef{a, b2};
  }
  H32 h32() {
    H32 h{};
    if (c + 32 > n) {
      ok = false;
      return h;
    }
    std::memcpy(h.data(), p + c, 32);
    c += 32;
    return h;
  }
  std::vector<uint8_t> bytes() {
    size_t L = len(1);
    std::vector<uint8_t> v;
    if (ok) {
      v.assign(p + c, p + c + L);
      c += L;
    }
    return v;
  }
  EfVec efvec() {
    size_t L = len(16);
    EfVec v;
    for (size_t i = 0; ok && i < L; i++) v.push_back(e());
    return v;
  }
  std::vector<uint64_t> bfvec() {
    size_t L = len(8);
    std::vector<uint64_t> v;
    for (size_t i = 0; ok && i < L; i++) v.push_back(bf());
    return v;
  }
  SumcheckProof sc() {
    SumcheckProof s;
    size_t L = len(8);
    for (size_t i = 0; ok && i < L; i++) s.round_polys.push_back(efvec());
    s.final_evals = efvec();
    return s;
  }
};

inline std::vector<uint8_t> encode_proof(const LigesisProof& pf) {
  Writer w;
  w.e(pf.y);
  w.h32(pf.com_a_vec);
  w.h32(pf.com_i);
  w.e(pf.y4);
  w.e(pf.y5);
  w.u64(pf.h_cols.size());
  for (auto& c : pf.h_cols) w.bfvec(c);
  w.bytes(pf.h_open);
  for (const SumcheckProof* s : {&pf.sc_binary, &pf.sc_hash, &pf.sc_recon, &pf.sc_table,
                                 &pf.sc_rslink, &pf.sc_bi_batch, &pf.sc_a_batch})
    w.sc(*s);
  const BatchProof& b = pf.batch;
  w.u64(b.reduction_msgs.size());
  for (auto& r : b.reduction_msgs) w.efvec(r);
  w.u64(b.fold_roots.size());
  for (auto& h : b.fold_roots) w.h32(h);
  w.efvec(b.final_codeword);
  w.u64(b.fold_openings.size());
  for (auto& op : b.fold_openings) {
    w.bytes(op.proof_bytes);
    w.u64(op.values.size());
    for (auto& kv : op.values) {
      w.u64(kv.first);
      w.efvec(kv.second);
    }
  }
  w.u64(b.base_bf.size());
  for (auto& op : b.base_bf) {
    w.bytes(op.proof_bytes);
    w.u64(op.values.size());
    for (auto& kv : op.values) {
      w.u64(kv.first);
      w.bfvec(kv.second);
    }
  }
  w.u64(b.base_ef.size());
  for (auto& op : b.base_ef) {
    w.bytes(op.proof_bytes);
    w.u64(op.values.size());
    for (auto& kv : op.values) {
      w.u64(kv.first);
      w.efvec(kv.second);
    }
  }
  w.u64(b.code_rate);
  w.u64(b.log_len);
  return w.b;
}

inline bool decode_proof(const uint8_t* data, size_t len, LigesisProof& pf) {
  Reader r{data, len};
  pf.y = r.e();
  pf.com_a_vec = r.h32();
  pf.com_i = r.h32();
  pf.y4 = r.e();
  pf.y5 = r.e();
  size_t nh = r.len(8);
  for (size_t i = 0; r.ok && i < nh; i++) pf.h_cols.push_back(r.bfvec());
  pf.h_open = r.bytes();
  for (SumcheckProof* s : {&pf.sc_binary, &pf.sc_hash, &pf.sc_recon, &pf.sc_table, &pf.sc_rslink,
                           &pf.sc_bi_batch, &pf.sc_a_batch})
    *s = r.sc();
  BatchProof& b = pf.batch;
  size_t nr = r.len(8);
  for (size_t i = 0; r.ok && i < nr; i++) b.reduction_msgs.push_back(r.efvec());
  size_t nf = r.len(32);
  for (size_t i = 0; r.ok && i < nf; i++) b.fold_roots.push_back(r.h32());
  b.final_codeword = r.efvec();
  size_t no = r.len(16);
  for (size_t i = 0; r.ok && i < no; i++) {
    RoundOpening op;
    op.proof_bytes = r.bytes();
    size_t nv = r.len(16);
    for (size_t k = 0; r.ok && k < nv; k++) {
      uint64_t key = r.u64();
      op.values[key] = r.efvec();
    }
    b.fold_openings.push_back(op);
  }
  size_t nb = r.len(16);
  for (size_t i = 0; r.ok && i < nb; i++) {
    BaseOpeningBf op;
    op.proof_bytes = r.bytes();
    size_t nv = r.len(16);
    for (size_t k = 0; r.ok && k < nv; k++) {
      uint64_t key = r.u64();
      op.values[key] = r.bfvec();
    }
    b.base_bf.push_back(op);
  }
  size_t ne = r.len(16);
  for (size_t i = 0; r.ok && i < ne; i++) {
    BaseOpeningEf op;
    op.proof_bytes = r.bytes();
    size_t nv = r.len(16);
    for (size_t k = 0; r.ok && k < nv; k++) {
      uint64_t key = r.u64();
      op.values[key] = r.efvec();
    }
    b.base_ef.push_back(op);
  }
  b.code_rate = r.u64();
  b.log_len = r.u64();
  return r.ok && r.c == len;
}

}  // namespace lgso
