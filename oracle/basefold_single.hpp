// Oracle (test infrastructure only): the single-polynomial Basefold opening.
//
// Restates pcs/src/basefold.rs:
//   BasefoldProof :84-103 (field order of the canonical encoding below),
//   prove :384-516 (claimed evaluation, eq-compression of the 2^li codewords, the fold of the
//     polynomial over its low li variables, R = num_vars - li reduction rounds sending both
//     l_i(0) and l_i(1), codeword folds, Merkle trees of levels 1..R-1, query_num(code_rate)
//     queries mod L/2, sorted and de-duplicated, level-0 and folded-level openings),
//   verify :518-642 (claim chain, the same queries, Merkle checks of every opening, the fold
//     chain of every query, a constant final codeword equal to the reduced claim).
// Pinned by: the reference tests basefold.rs:1054-1185 restated in
// tests/test_oracle_basefold.py (base and extension round trips, claimed_eval equal to an
// independent multilinear evaluation, a tampered reduction message rejected), and the same
// commitment, fold and opening code as the batched open, pinned there.
#pragma once
#include <type_traits>

#include "ligesis.hpp"

namespace lgso {

struct SingleProof {
  std::vector<std::pair<Ef, Ef>> reduction_msgs;
  std::vector<H32> fold_roots;
  EfVec final_codeword;
  std::vector<uint8_t> base_proof;
  std::map<uint64_t, std::vector<uint64_t>> base_values;  // leaf-field words: 1 (base) or 2 (ext) per value
  std::vector<RoundOpening> fold_openings;
  Ef claimed_eval{};
  uint64_t num_vars = 0, code_rate = 0, log_interleave = 0;
  bool ext = false;  // the leaf field (BasefoldProof<LF, EF>'s LF): a type, not encoded
};

inline void leaf_words(const uint64_t& v, std::vector<uint64_t>& out) { out.push_back(v); }
inline void leaf_words(const Ef& v, std::vector<uint64_t>& out) {
  out.push_back(v.c0);
  out.push_back(v.c1);
}

template <class LF>
inline SingleProof single_prove(const BfState<LF>& st, const EfVec& point, OracleStream& o) {
  const size_t n = st.num_vars, li = st.log_interleave;
  if (point.size() != n || n < li) throw std::runtime_error("basefold prove: point length mismatch");
  const size_t rounds = n - li, l = st.codes[0].size();
  SingleProof pf;
  pf.claimed_eval = mle_eval(st.poly, point);
  EfVec eq4 = new_eq(EfVec(point.begin(), point.begin() + li));
  EfVec oracle0(l, ef_zero());
  for (size_t j = 0; j < st.codes.size(); j++)
    for (size_t i = 0; i < l; i++) oracle0[i] = ef_add(oracle0[i], ef_mul(eq4[j], lift(st.codes[j][i])));
  EfVec m = st.poly;
  mle_fold(m, EfVec(point.begin(), point.begin() + li));
  EfVec zp(point.begin() + li, point.end());
  std::vector<EfVec> levels{oracle0};
  for (size_t round = 0; round < rounds; round++) {
    size_t half = m.size() / 2;
    EfVec m0(half), m1(half);
    for (size_t i = 0; i < half; i++) {
      m0[i] = m[2 * i];
      m1[i] = m[2 * i + 1];
    }
    EfVec rest(zp.begin() + round + 1, zp.end());
    Ef l0 = mle_eval(m0, rest), l1 = mle_eval(m1, rest);
    pf.reduction_msgs.push_back({l0, l1});
    Ef r = o.next_field();
    mle_fold(m, EfVec{r});
    levels.push_back(eq_fold_codeword(levels.back(), r));
  }
  pf.final_codeword = levels[rounds];
  std::vector<MerkleTree> fmts;
  for (size_t k = 1; k < rounds; k++) {
    std::vector<H32> lh;
    for (auto& v : levels[k]) lh.push_back(blake3(serialize_fields(EfVec{v})));
    fmts.push_back(MerkleTree::from_leaf_hashes(std::move(lh)));
    pf.fold_roots.push_back(fmts.back().root());
  }
  const size_t q = query_num(st.code_rate);
  std::vector<size_t> base;
  for (uint64_t v : o.next_n_ints(q)) base.push_back(v % (l / 2));
  std::sort(base.begin(), base.end());
  base.erase(std::unique(base.begin(), base.end()), base.end());
  auto bpos = level_positions(base, l / 2);
  pf.base_proof = st.mt.open(bpos);
  for (size_t p : bpos) {
    std::vector<uint64_t> col;
    for (auto& c : st.codes) leaf_words(c[p], col);
    pf.base_values[p] = col;
  }
  for (size_t k = 1; k < rounds; k++) {
    size_t cur = l >> k;
    auto pos = level_positions(base, cur / 2);
    RoundOpening ro;
    ro.proof_bytes = fmts[k - 1].open(pos);
    for (size_t p : pos) ro.values[p] = EfVec{levels[k][p]};
    pf.fold_openings.push_back(ro);
  }
  pf.num_vars = n;
  pf.code_rate = st.code_rate;
  pf.log_interleave = li;
  pf.ext = std::is_same<LF, Ef>::value;
  return pf;
}

// `ext`: the commitment's leaves hold extension values (commit_ext_on_domain). A malformed
// proof (wrong lengths, missing openings) is rejected, never read out of bounds.
inline bool single_verify(const H32& commit, const EfVec& point, const SingleProof& pf, bool ext, OracleStream& o) {
  const size_t n = pf.num_vars, li = pf.log_interleave;
  if (point.size() != n || n < li || li > 30 || pf.code_rate > 30 || n - li + pf.code_rate > 40) return false;
  const size_t rounds = n - li, l = (size_t(1) << (n - li)) << pf.code_rate;
  if (pf.reduction_msgs.size() != rounds || pf.fold_roots.size() != (rounds ? rounds - 1 : 0) ||
      pf.fold_openings.size() != pf.fold_roots.size())
    return false;
  EfVec eq4 = new_eq(EfVec(point.begin(), point.begin() + li));
  EfVec zp(point.begin() + li, point.end());
  Ef y = pf.claimed_eval;
  EfVec ch;
  for (size_t round = 0; round < rounds; round++) {
    Ef l0 = pf.reduction_msgs[round].first, l1 = pf.reduction_msgs[round].second;
    Ef zi = zp[round];
    if (ef_add(ef_mul(ef_sub(ef_one(), zi), l0), ef_mul(zi, l1)) != y) return false;
    Ef r = o.next_field();
    ch.push_back(r);
    y = ef_add(ef_mul(ef_sub(ef_one(), r), l0), ef_mul(r, l1));
  }
  const size_t q = query_num(pf.code_rate);
  std::vector<size_t> base;
  for (uint64_t v : o.next_n_ints(q)) base.push_back(v % (l / 2));
  std::sort(base.begin(), base.end());
  base.erase(std::unique(base.begin(), base.end()), base.end());
  const size_t width = (ext ? 2 : 1) << li;  // words per level-0 leaf
  {
    std::vector<size_t> idx;
    std::vector<H32> lh;
    for (auto& kv : pf.base_values) {
      if (kv.second.size() != width) return false;
      idx.push_back(kv.first);
      std::vector<uint8_t> ser;
      for (uint64_t w : kv.second) ser_field(ser, w);
      lh.push_back(blake3(ser));
    }
    if (idx.empty() || !merkle_verify(commit, l, pf.base_proof, idx, lh)) return false;
  }
  for (size_t k = 1; k < rounds; k++) {
    size_t cur = l >> k;
    const RoundOpening& op = pf.fold_openings[k - 1];
    std::vector<size_t> idx;
    std::vector<H32> lh;
    for (auto& kv : op.values) {
      if (kv.second.size() != 1) return false;
      idx.push_back(kv.first);
      lh.push_back(blake3(serialize_fields(kv.second)));
    }
    if (idx.empty() || !merkle_verify(pf.fold_roots[k - 1], cur, op.proof_bytes, idx, lh)) return false;
  }
  auto val_at = [&](size_t k, size_t p, bool& ok) -> Ef {
    if (k == 0) {
      auto it = pf.base_values.find(p);
      if (it == pf.base_values.end()) {
        ok = false;
        return ef_zero();
      }
      Ef acc = ef_zero();
      for (size_t j = 0; j < eq4.size(); j++) {
        Ef v = ext ? Ef{it->second[2 * j], it->second[2 * j + 1]} : ef(it->second[j]);
        acc = ef_add(acc, ef_mul(eq4[j], v));
      }
      return acc;
    }
    auto it = pf.fold_openings[k - 1].values.find(p);
    if (it == pf.fold_openings[k - 1].values.end()) {
      ok = false;
      return ef_zero();
    }
    return it->second[0];
  };
  const size_t final_len = l >> rounds;
  if (pf.final_codeword.size() != final_len) return false;
  if (rounds == 0) return false;  // the reference unwraps the fold chain's value (basefold.rs:624)
  for (size_t idx : base) {
    bool have = false, ok = true;
    Ef derived = ef_zero();
    for (size_t k = 0; k < rounds; k++) {
      size_t cur = l >> k, half = cur / 2, pos = idx % half, sib = pos + half;
      Ef a = val_at(k, pos, ok), b = val_at(k, sib, ok);
      if (!ok) return false;
      if (have && derived != ((idx % cur) == pos ? a : b)) return false;
      derived = eq_fold_at(a, b, ch[k], cur, pos);
      have = true;
    }
    if (derived != pf.final_codeword[idx % final_len]) return false;
  }
  Ef c0 = pf.final_codeword[0];
  for (auto& v : pf.final_codeword)
    if (v != c0) return false;
  return y == c0;
}

inline std::vector<uint8_t> encode_single_proof(const SingleProof& pf) {
  Writer w;
  w.u64(pf.reduction_msgs.size());
  for (auto& r : pf.reduction_msgs) {
    w.e(r.first);
    w.e(r.second);
  }
  w.u64(pf.fold_roots.size());
  for (auto& h : pf.fold_roots) w.h32(h);
  w.efvec(pf.final_codeword);
  w.bytes(pf.base_proof);
  w.u64(pf.base_values.size());
  for (auto& kv : pf.base_values) {
    w.u64(kv.first);
    w.u64(kv.second.size() / (pf.ext ? 2 : 1));  // Vec<LF> length, then its words
    for (uint64_t x : kv.second) w.bf(x);
  }
  w.u64(pf.fold_openings.size());
  for (auto& op : pf.fold_openings) {
    w.bytes(op.proof_bytes);
    w.u64(op.values.size());
    for (auto& kv : op.values) {
      w.u64(kv.first);
      w.efvec(kv.second);
    }
  }
  w.e(pf.claimed_eval);
  w.u64(pf.num_vars);
  w.u64(pf.code_rate);
  w.u64(pf.log_interleave);
  return w.b;
}

inline bool decode_single_proof(const uint8_t* data, size_t len, bool ext, SingleProof& pf) {
  Reader r{data, len};
  pf.ext = ext;
  size_t nr = r.len(32);
  for (size_t i = 0; r.ok && i < nr; i++) {
    Ef a = r.e(), b = r.e();
    pf.reduction_msgs.push_back({a, b});
  }
  size_t nf = r.len(32);
  for (size_t i = 0; r.ok && i < nf; i++) pf.fold_roots.push_back(r.h32());
  pf.final_codeword = r.efvec();
  pf.base_proof = r.bytes();
  size_t nb = r.len(16);
  for (size_t i = 0; r.ok && i < nb; i++) {
    uint64_t key = r.u64();
    size_t cnt = r.len(ext ? 16 : 8);
    std::vector<uint64_t> v;
    for (size_t k = 0; r.ok && k < cnt * (ext ? 2 : 1); k++) v.push_back(r.bf());
    pf.base_values[key] = v;
  }
  size_t no = r.len(16);
  for (size_t i = 0; r.ok && i < no; i++) {
    RoundOpening op;
    op.proof_bytes = r.bytes();
    size_t nv = r.len(16);
    for (size_t k = 0; r.ok && k < nv; k++) {
      uint64_t key = r.u64();
      op.values[key] = r.efvec();
    }
    pf.fold_openings.push_back(op);
  }
  pf.claimed_eval = r.e();
  pf.num_vars = r.u64();
  pf.code_rate = r.u64();
  pf.log_interleave = r.u64();
  return r.ok && r.c == len;
}

}  // namespace lgso
