// Oracle (test infrastructure only): Basefold shared-domain commitment and batched open.
//
// Restates pcs/src/basefold.rs:
//   eq_fold_codeword :201-218, eq_fold_at :222-229, level_positions :234-244, query_num :107-109,
//   to_commit :257-307 (leaf p = column [codes[0][p] .. codes[2^li-1][p]], serialized),
//   commit_{base,ext}_on_domain :341-382 (chunk j = f[j :: 2^li], zero-pad x2^cr, DFT),
//   make_channel :648-671, batch_prove :679-856, batch_verify :860-1035.
// Pinned by: batch round trips with mixed sizes/fields (the reference test basefold.rs:1187-1211
// restated in tests/test_oracle_basefold.py), tamper rejection (basefold.rs:1213-1267), the
// reduction identity (1-z)l0 + z l1 = y (basefold.rs:757), and the final codeword being
// constant and equal to the combined claim.
#pragma once
#include <map>

#include "dft.hpp"
#include "merkle.hpp"
#include "poly.hpp"

namespace lgso {

inline size_t query_num(size_t code_rate) { return (100 + code_rate - 1) / code_rate; }

inline EfVec eq_fold_codeword(const EfVec& v, Ef r) {
  size_t len = v.size(), h = len / 2;
  uint64_t g = two_adic_generator(log2_exact(len));
  uint64_t ginv = finv(g), inv2 = finv(2);
  uint64_t tw = 1;
  EfVec out;
  out.reserve(h);
  Ef omr = ef_sub(ef_one(), r);
  for (size_t i = 0; i < h; i++) {
    Ef x = v[i], nx = v[i + h];
    Ef pe = ef_mul_base(ef_add(x, nx), inv2);
    Ef po = ef_mul_base(ef_mul_base(ef_sub(x, nx), tw), inv2);
    out.push_back(ef_add(ef_mul(omr, pe), ef_mul(r, po)));
    tw = fmul(tw, ginv);
  }
  return out;
}

inline Ef eq_fold_at(Ef a, Ef b, Ef r, size_t len, size_t pos) {
  uint64_t g = two_adic_generator(log2_exact(len));
  uint64_t tw = fpow(finv(g), pos), inv2 = finv(2);
  Ef pe = ef_mul_base(ef_add(a, b), inv2);
  Ef po = ef_mul_base(ef_mul_base(ef_sub(a, b), tw), inv2);
  return ef_add(ef_mul(ef_sub(ef_one(), r), pe), ef_mul(r, po));
}

inline std::vector<size_t> level_positions(const std::vector<size_t>& base, size_t half) {
  std::vector<size_t> p;
  for (size_t idx : base) {
    p.push_back(idx % half);
    p.push_back(idx % half + half);
  }
  std::sort(p.begin(), p.end());
  p.erase(std::unique(p.begin(), p.end()), p.end());
  return p;
}

template <class LF>
struct BfState {
  std::vector<std::vector<LF>> codes;  // 2^li codewords of length D
  MerkleTree mt;
  EfVec poly;  // full polynomial lifted to EF
  size_t num_vars = 0, code_rate = 0, log_interleave = 0;
};

template <class LF>
inline H32 commit_on_domain(const std::vector<LF>& poly, size_t code_rate, size_t li,
                            BfState<LF>& st) {
  auto chunks = mle_split(poly, poly.size() >> li);
  size_t code_len = chunks[0].size() << code_rate;
  st.codes.clear();
  for (auto& c : chunks) {
    std::vector<LF> x = c;
    x.resize(code_len, LF{});
    dft_any(x);
    st.codes.push_back(std::move(x));
  }
  std::vector<H32> leaves(code_len);
  std::vector<LF> col(st.codes.size());
  for (size_t p = 0; p < code_len; p++) {
    for (size_t j = 0; j < st.codes.size(); j++) col[j] = st.codes[j][p];
    leaves[p] = blake3(serialize_fields(col));
  }
  st.mt = MerkleTree::from_leaf_hashes(std::move(leaves));
  st.poly.clear();
  for (auto& v : poly) st.poly.push_back(lift(v));
  st.num_vars = log2_exact(poly.size());
  st.code_rate = code_rate;
  st.log_interleave = li;
  return st.mt.root();
}

struct RoundOpening {
  std::vector<uint8_t> proof_bytes;
  std::map<uint64_t, EfVec> values;
};
struct BaseOpeningBf {
  std::vector<uint8_t> proof_bytes;
  std::map<uint64_t, std::vector<uint64_t>> values;
};
struct BaseOpeningEf {
  std::vector<uint8_t> proof_bytes;
  std::map<uint64_t, EfVec> values;
};

struct BatchProof {
  std::vector<EfVec> reduction_msgs;
  std::vector<H32> fold_roots;
  EfVec final_codeword;
  std::vector<RoundOpening> fold_openings;
  std::vector<BaseOpeningBf> base_bf;
  std::vector<BaseOpeningEf> base_ef;
  uint64_t code_rate = 0, log_len = 0;
};

template <class LF>
inline void make_channel(const BfState<LF>& st, const EfVec& point, size_t d, EfVec& oracle0,
                         EfVec& m, EfVec& zp, Ef& y) {
  size_t li = st.log_interleave;
  EfVec eq4 = new_eq(EfVec(point.begin(), point.begin() + li));
  oracle0.assign(d, ef_zero());
  for (size_t j = 0; j < st.codes.size(); j++)
    for (size_t i = 0; i < d; i++) oracle0[i] = ef_add(oracle0[i], ef_mul(eq4[j], lift(st.codes[j][i])));
  m = st.poly;
  mle_fold(m, EfVec(point.begin(), point.begin() + li));
  y = mle_eval(st.poly, point);
  zp.assign(point.begin() + li, point.end());
}

template <class LF, class Open>
inline Open base_open(const BfState<LF>& st, const std::vector<size_t>& pos) {
  Open op;
  op.proof_bytes = st.mt.open(pos);
  for (size_t p : pos) {
    std::vector<LF> col;
    for (auto& c : st.codes) col.push_back(c[p]);
    op.values[p] = col;
  }
  return op;
}

inline BatchProof batch_prove(const std::vector<std::pair<const BfState<uint64_t>*, EfVec>>& bf,
                              const std::vector<std::pair<const BfState<Ef>*, EfVec>>& efs,
                              OracleStream& o) {
  size_t d = !bf.empty() ? bf[0].first->codes[0].size() : efs.at(0).first->codes[0].size();
  size_t cr = !bf.empty() ? bf[0].first->code_rate : efs[0].first->code_rate;
  for (auto& t : bf)
    if (t.first->codes[0].size() != d || t.first->code_rate != cr)
      throw std::runtime_error("batch_prove: domain mismatch");
  for (auto& t : efs)
    if (t.first->codes[0].size() != d || t.first->code_rate != cr)
      throw std::runtime_error("batch_prove: domain mismatch");
  size_t log_len = log2_exact(d), rounds = log_len - cr;
  size_t nch = bf.size() + efs.size();
  std::vector<EfVec> o0(nch), ms(nch), zps(nch);
  EfVec ys(nch);
  for (size_t c = 0; c < bf.size(); c++)
    make_channel(*bf[c].first, bf[c].second, d, o0[c], ms[c], zps[c], ys[c]);
  for (size_t c = 0; c < efs.size(); c++) {
    size_t k = bf.size() + c;
    make_channel(*efs[c].first, efs[c].second, d, o0[k], ms[k], zps[k], ys[k]);
  }
  EfVec coeffs = o.next_n_fields(nch);
  EfVec combined(d, ef_zero());
  for (size_t c = 0; c < nch; c++)
    for (size_t i = 0; i < d; i++) combined[i] = ef_add(combined[i], ef_mul(coeffs[c], o0[c][i]));

  BatchProof pf;
  pf.reduction_msgs.assign(nch, EfVec());
  std::vector<EfVec> levels{combined};
  for (size_t round = 0; round < rounds; round++) {
    std::vector<std::pair<Ef, Ef>> l01;
    for (size_t c = 0; c < nch; c++) {
      const EfVec& m = ms[c];
      size_t half = m.size() / 2;
      EfVec m0(half), m1(half);
      for (size_t i = 0; i < half; i++) {
        m0[i] = m[2 * i];
        m1[i] = m[2 * i + 1];
      }
      EfVec rest(zps[c].begin() + round + 1, zps[c].end());
      Ef l0 = mle_eval(m0, rest), l1 = mle_eval(m1, rest);
      pf.reduction_msgs[c].push_back(l0);
      l01.push_back({l0, l1});
    }
    Ef r = o.next_field();
    for (size_t c = 0; c < nch; c++) {
      ys[c] = ef_add(ef_mul(ef_sub(ef_one(), r), l01[c].first), ef_mul(r, l01[c].second));
      mle_fold(ms[c], EfVec{r});
    }
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
  size_t q = query_num(cr);
  std::vector<size_t> base;
  for (uint64_t v : o.next_n_ints(q)) base.push_back(v % (d / 2));
  std::sort(base.begin(), base.end());
  base.erase(std::unique(base.begin(), base.end()), base.end());
  auto bpos = level_positions(base, d / 2);
  for (auto& t : bf) pf.base_bf.push_back(base_open<uint64_t, BaseOpeningBf>(*t.first, bpos));
  for (auto& t : efs) pf.base_ef.push_back(base_open<Ef, BaseOpeningEf>(*t.first, bpos));
  for (size_t k = 1; k < rounds; k++) {
    size_t cur = d >> k;
    auto pos = level_positions(base, cur / 2);
    RoundOpening ro;
    ro.proof_bytes = fmts[k - 1].open(pos);
    for (size_t p : pos) ro.values[p] = EfVec{levels[k][p]};
    pf.fold_openings.push_back(ro);
  }
  pf.code_rate = cr;
  pf.log_len = log_len;
  return pf;
}

struct BfClaim {
  H32 commit;
  EfVec point;
  Ef claim;
  size_t li;
};

template <class Op>
inline bool check_base_opening(const Op& op, const H32& commit, size_t d,
                               const std::vector<size_t>& bpos) {
  std::vector<size_t> idx;
  for (auto& kv : op.values) idx.push_back(kv.first);
  if (idx != bpos) return false;
  std::vector<H32> lh;
  for (size_t p : idx) lh.push_back(blake3(serialize_fields(op.values.at(p))));
  return merkle_verify(commit, d, op.proof_bytes, idx, lh);
}

inline bool batch_verify(const std::vector<BfClaim>& bfc, const std::vector<BfClaim>& efc,
                         const BatchProof& pf, OracleStream& o) {
  if (pf.log_len < pf.code_rate || pf.log_len > 40) return false;
  size_t d = size_t(1) << pf.log_len, cr = pf.code_rate, rounds = pf.log_len - cr;
  size_t nbf = bfc.size(), nch = nbf + efc.size();
  size_t nfold = rounds >= 1 ? rounds - 1 : 0;
  if (pf.reduction_msgs.size() != nch || pf.base_bf.size() != nbf ||
      pf.base_ef.size() != efc.size() || pf.fold_roots.size() != nfold ||
      pf.fold_openings.size() != nfold)
    return false;
  std::vector<const BfClaim*> all;
  for (auto& c : bfc) all.push_back(&c);
  for (auto& c : efc) all.push_back(&c);
  std::vector<EfVec> eq4s;
  EfVec ys;
  for (size_t c = 0; c < nch; c++) {
    const BfClaim& cl = *all[c];
    if (cl.point.size() < cl.li || cl.point.size() - cl.li != rounds ||
        pf.reduction_msgs[c].size() != rounds)
      return false;
    eq4s.push_back(new_eq(EfVec(cl.point.begin(), cl.point.begin() + cl.li)));
    ys.push_back(cl.claim);
  }
  EfVec coeffs = o.next_n_fields(nch);
  EfVec ch;
  for (size_t round = 0; round < rounds; round++) {
    Ef r = o.next_field();
    ch.push_back(r);
    for (size_t c = 0; c < nch; c++) {
      Ef l0 = pf.reduction_msgs[c][round];
      Ef zi = all[c]->point[all[c]->li + round];
      if (zi == ef_zero()) return false;
      Ef l1 = ef_mul(ef_sub(ys[c], ef_mul(ef_sub(ef_one(), zi), l0)), ef_inv(zi));
      ys[c] = ef_add(ef_mul(ef_sub(ef_one(), r), l0), ef_mul(r, l1));
    }
  }
  size_t q = query_num(cr);
  std::vector<size_t> base;
  for (uint64_t v : o.next_n_ints(q)) base.push_back(v % (d / 2));
  std::sort(base.begin(), base.end());
  base.erase(std::unique(base.begin(), base.end()), base.end());
  auto bpos = level_positions(base, d / 2);
  for (size_t i = 0; i < nbf; i++)
    if (!check_base_opening(pf.base_bf[i], bfc[i].commit, d, bpos)) return false;
  for (size_t i = 0; i < efc.size(); i++)
    if (!check_base_opening(pf.base_ef[i], efc[i].commit, d, bpos)) return false;
  for (size_t k = 1; k < rounds; k++) {
    size_t cur = d >> k;
    const RoundOpening& op = pf.fold_openings[k - 1];
    std::vector<size_t> idx;
    std::vector<H32> lh;
    for (auto& kv : op.values) {
      idx.push_back(kv.first);
      lh.push_back(blake3(serialize_fields(kv.second)));
    }
    if (idx.empty() || !merkle_verify(pf.fold_roots[k - 1], cur, op.proof_bytes, idx, lh))
      return false;
  }
  auto val_at = [&](size_t k, size_t p, bool& ok) -> Ef {
    if (k == 0) {
      Ef acc = ef_zero();
      for (size_t c = 0; c < nch; c++) {
        Ef a = ef_zero();
        if (c < nbf) {
          auto it = pf.base_bf[c].values.find(p);
          if (it == pf.base_bf[c].values.end() || it->second.size() != eq4s[c].size()) {
            ok = false;
            return a;
          }
          for (size_t j = 0; j < it->second.size(); j++)
            a = ef_add(a, ef_mul(eq4s[c][j], ef(it->second[j])));
        } else {
          auto it = pf.base_ef[c - nbf].values.find(p);
          if (it == pf.base_ef[c - nbf].values.end() || it->second.size() != eq4s[c].size()) {
            ok = false;
            return a;
          }
          for (size_t j = 0; j < it->second.size(); j++)
            a = ef_add(a, ef_mul(eq4s[c][j], it->second[j]));
        }
        acc = ef_add(acc, ef_mul(coeffs[c], a));
      }
      return acc;
    }
    auto it = pf.fold_openings[k - 1].values.find(p);
    if (it == pf.fold_openings[k - 1].values.end() || it->second.empty()) {
      ok = false;
      return ef_zero();
    }
    return it->second[0];
  };
  size_t final_len = d >> rounds;
  if (pf.final_codeword.size() != final_len) return false;
  for (size_t idx : base) {
    bool have = false, ok = true;
    Ef derived = ef_zero();
    for (size_t k = 0; k < rounds; k++) {
      size_t cur = d >> k, half = cur / 2, pos = idx % half, sib = pos + half;
      Ef a = val_at(k, pos, ok), b = val_at(k, sib, ok);
      if (!ok) return false;
      if (have) {
        size_t prev = idx % cur;
        if (derived != (prev == pos ? a : b)) return false;
      }
      derived = eq_fold_at(a, b, ch[k], cur, pos);
      have = true;
    }
    if (derived != pf.final_codeword[idx % final_len]) return false;
  }
  Ef c0 = pf.final_codeword[0];
  for (auto& v : pf.final_codeword)
    if (v != c0) return false;
  Ef comb = ef_zero();
  for (size_t c = 0; c < nch; c++) comb = ef_add(comb, ef_mul(coeffs[c], ys[c]));
  return comb == c0;
}

}  // namespace lgso
