// Oracle (test infrastructure only): multilinear polynomials, the product sumcheck and the
// pre-sampled randomness oracle.
//
// Restates utils/src/poly.rs:7-61 (MlPoly eval/split/fold/new_eq/eval_eq, low variable first),
// utils/src/sumcheck.rs:37-130 (prove binds the HIGH variable first; round message = g(0..d);
// verify with Lagrange interpolation over 0..d) and utils/src/oracle.rs:10-64 (two independent
// cursors over pre-drawn field and integer streams).
// Pinned by: hypercube-sum closed forms and g(0)+g(1)=claim invariants, the reference tests
// sumcheck.rs:155-183 (round trip / wrong claim / tampered round), eq-table identities
// (sum eq = 1, MlPoly(v).eval(p) = <v, eq_p>) in tests/test_oracle_poly.py.
#pragma once
#include <stdexcept>
#include <vector>

#include "field.hpp"

namespace lgso {

typedef std::vector<Ef> EfVec;

// MlPoly::eval (poly.rs:7-18): bind point[0] to the lowest index bit first.
inline Ef mle_eval(EfVec s, const EfVec& point) {
  if ((size_t(1) << point.size()) != s.size()) throw std::runtime_error("mle_eval: size");
  size_t cur = s.size() >> 1;
  for (const Ef& r : point) {
    for (size_t i = 0; i < cur; i++)
      s[i] = ef_add(s[2 * i], ef_mul(ef_sub(s[2 * i + 1], s[2 * i]), r));
    cur >>= 1;
  }
  return s[0];
}

// MlPoly::fold (poly.rs:33-42)
inline void mle_fold(EfVec& s, const EfVec& point) {
  size_t cur = s.size();
  for (const Ef& r : point) {
    cur >>= 1;
    for (size_t i = 0; i < cur; i++)
      s[i] = ef_add(s[2 * i], ef_mul(ef_sub(s[2 * i + 1], s[2 * i]), r));
  }
  s.resize(cur);
}

// MlPoly::split (poly.rs:20-31): chunk j = v[j], v[j + nc], v[j + 2 nc], ... with nc = len / n.
template <class T>
inline std::vector<std::vector<T>> mle_split(const std::vector<T>& v, size_t n) {
  size_t len = v.size(), nc = len / n;
  std::vector<std::vector<T>> out(nc);
  for (size_t i = 0; i < len; i += nc)
    for (size_t j = 0; j < nc; j++) out[j].push_back(v[i + j]);
  return out;
}

// MlPoly::new_eq (poly.rs:44-53): eq[b] = prod_t (b_t ? p_t : 1 - p_t), bit t of b <-> p[t].
inline EfVec new_eq(const EfVec& point) {
  EfVec e{ef_one()};
  for (size_t k = point.size(); k-- > 0;) {
    const Ef& p = point[k];
    Ef q = ef_sub(ef_one(), p);
    EfVec n;
    n.reserve(e.size() * 2);
    for (const Ef& x : e) {
      n.push_back(ef_mul(q, x));
      n.push_back(ef_mul(p, x));
    }
    e.swap(n);
  }
  return e;
}

// MlPoly::eval_eq (poly.rs:55-61)
inline Ef eval_eq(const EfVec& a, const EfVec& b) {
  Ef r = ef_one();
  for (size_t i = 0; i < a.size() && i < b.size(); i++) {
    Ef t = ef_add(ef_mul(a[i], b[i]),
                  ef_mul(ef_sub(ef_one(), a[i]), ef_sub(ef_one(), b[i])));
    r = ef_mul(r, t);
  }
  return r;
}

struct OracleStream {
  std::vector<Ef> fields;
  std::vector<uint64_t> ints;
  size_t fi = 0, ii = 0;
  Ef next_field() {
    if (fi >= fields.size()) throw std::runtime_error("oracle field stream exhausted");
    return fields[fi++];
  }
  EfVec next_n_fields(size_t n) {
    EfVec v;
    for (size_t i = 0; i < n; i++) v.push_back(next_field());
    return v;
  }
  uint64_t next_int() {
    if (ii >= ints.size()) throw std::runtime_error("oracle int stream exhausted");
    return ints[ii++];
  }
  std::vector<uint64_t> next_n_ints(size_t n) {
    std::vector<uint64_t> v;
    for (size_t i = 0; i < n; i++) v.push_back(next_int());
    return v;
  }
  void restart() { fi = ii = 0; }
};

struct SumcheckProof {
  std::vector<EfVec> round_polys;
  EfVec final_evals;
};

// sumcheck::prove (sumcheck.rs:37-84)
inline SumcheckProof sumcheck_prove(std::vector<EfVec> polys, OracleStream& o, EfVec& challenges) {
  if (polys.empty()) throw std::runtime_error("sumcheck: no polys");
  size_t d = polys.size();
  int n = 0;
  while ((size_t(1) << n) < polys[0].size()) n++;
  SumcheckProof pf;
  challenges.clear();
  for (int round = 0; round < n; round++) {
    size_t half = polys[0].size() / 2;
    EfVec evals(d + 1, ef_zero());
    for (size_t i = 0; i < half; i++) {
      for (size_t k = 0; k <= d; k++) {
        Ef t = ef((uint64_t)k);
        Ef prod = ef_one();
        for (auto& p : polys) {
          Ef lo = p[i], hi = p[i + half];
          prod = ef_mul(prod, ef_add(lo, ef_mul(t, ef_sub(hi, lo))));
        }
        evals[k] = ef_add(evals[k], prod);
      }
    }
    pf.round_polys.push_back(evals);
    Ef r = o.next_field();
    challenges.push_back(r);
    for (auto& p : polys) {
      for (size_t i = 0; i < half; i++) p[i] = ef_add(p[i], ef_mul(r, ef_sub(p[i + half], p[i])));
      p.resize(half);
    }
  }
  for (auto& p : polys) pf.final_evals.push_back(p[0]);
  return pf;
}

// interpolate (sumcheck.rs:88-104)
inline Ef sc_interpolate(const EfVec& evals, Ef r) {
  Ef acc = ef_zero();
  size_t n = evals.size();
  for (size_t k = 0; k < n; k++) {
    Ef num = ef_one(), den = ef_one();
    for (size_t m = 0; m < n; m++) {
      if (m == k) continue;
      num = ef_mul(num, ef_sub(r, ef((uint64_t)m)));
      den = ef_mul(den, ef_sub(ef((uint64_t)k), ef((uint64_t)m)));
    }
    acc = ef_add(acc, ef_mul(evals[k], ef_mul(num, ef_inv(den))));
  }
  return acc;
}

// sumcheck::verify (sumcheck.rs:110-130)
inline bool sumcheck_verify(Ef claim, const SumcheckProof& pf, OracleStream& o, EfVec& challenges) {
  Ef cur = claim;
  challenges.clear();
  for (auto& ev : pf.round_polys) {
    if (ev.size() < 2 || ef_add(ev[0], ev[1]) != cur) return false;
    Ef r = o.next_field();
    cur = sc_interpolate(ev, r);
    challenges.push_back(r);
  }
  Ef prod = ef_one();
  for (auto& f : pf.final_evals) prod = ef_mul(prod, f);
  return prod == cur;
}

inline EfVec rev(const EfVec& v) { return EfVec(v.rbegin(), v.rend()); }
inline EfVec cat(const EfVec& lo, const EfVec& hi) {
  EfVec r = lo;
  r.insert(r.end(), hi.begin(), hi.end());
  return r;
}

}  // namespace lgso
