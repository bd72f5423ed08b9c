// Oracle (test infrastructure only): C entry points over the CPU restatement, loaded by
// tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg through ctypes. The product
// library (ligesis_amd/csrc) never links or calls anything here.
#include <cstring>
#include <exception>
#include <string>

#include "basefold_single.hpp"
#include "ligesis.hpp"

using namespace lgso;

namespace {
thread_local std::string g_err;

template <class F>
int64_t guard(F f) {
  try {
    return f();
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

OracleStream make_stream(const uint64_t* fields, uint64_t nf, const uint64_t* ints, uint64_t ni) {
  OracleStream o;
  for (uint64_t i = 0; i < nf; i++) o.fields.push_back(Ef{fields[2 * i], fields[2 * i + 1]});
  o.ints.assign(ints, ints + ni);
  return o;
}

EfVec load_ef(const uint64_t* p, size_t n) {
  EfVec v(n);
  for (size_t i = 0; i < n; i++) v[i] = Ef{p[2 * i], p[2 * i + 1]};
  return v;
}

void store_ef(const EfVec& v, uint64_t* out) {
  for (size_t i = 0; i < v.size(); i++) {
    out[2 * i] = v[i].c0;
    out[2 * i + 1] = v[i].c1;
  }
}
}  // namespace

extern "C" {

const char* lgso_last_error() { return g_err.c_str(); }

// ---- field ----
void lgso_field_ops(const uint64_t* a, const uint64_t* b, uint64_t n, uint64_t* add, uint64_t* sub,
                    uint64_t* mul) {
  for (uint64_t i = 0; i < n; i++) {
    add[i] = fadd(a[i], b[i]);
    sub[i] = fsub(a[i], b[i]);
    mul[i] = fmul(a[i], b[i]);
  }
}
void lgso_reduce128(const uint64_t* lo, const uint64_t* hi, uint64_t n, uint64_t* out) {
  for (uint64_t i = 0; i < n; i++) out[i] = reduce128(((u128)hi[i] << 64) | lo[i]);
}
void lgso_ef_ops(const uint64_t* a, const uint64_t* b, uint64_t n, uint64_t* mul, uint64_t* inv) {
  for (uint64_t i = 0; i < n; i++) {
    Ef x{a[2 * i], a[2 * i + 1]}, y{b[2 * i], b[2 * i + 1]};
    Ef m = ef_mul(x, y);
    mul[2 * i] = m.c0;
    mul[2 * i + 1] = m.c1;
    Ef v = (x == ef_zero()) ? ef_zero() : ef_inv(x);
    inv[2 * i] = v.c0;
    inv[2 * i + 1] = v.c1;
  }
}
uint64_t lgso_two_adic_generator(uint64_t bits) { return two_adic_generator((int)bits); }

// ---- dft ----
void lgso_dft(uint64_t* a, uint64_t n) {
  std::vector<uint64_t> v(a, a + n);
  dft_base(v);
  std::memcpy(a, v.data(), n * 8);
}
void lgso_dft_ext(uint64_t* a, uint64_t n) {
  EfVec v = load_ef(a, n);
  dft_ext(v);
  store_ef(v, a);
}
// Per-row rate-1/2 RS encoding (ligesis.rs:231-235): rows x n -> rows x 2n.
void lgso_rs_encode_rows(const uint64_t* f, uint64_t rows, uint64_t n, uint64_t* out) {
  std::vector<uint64_t> r(2 * n);
  for (uint64_t i = 0; i < rows; i++) {
    std::fill(r.begin(), r.end(), 0);
    std::copy(f + i * n, f + (i + 1) * n, r.begin());
    dft_base(r);
    std::memcpy(out + i * 2 * n, r.data(), 2 * n * 8);
  }
}

// ---- blake3 / merkle ----
void lgso_blake3(const uint8_t* data, uint64_t len, uint8_t* out) {
  H32 h = blake3(data, len);
  std::memcpy(out, h.data(), 32);
}
int64_t lgso_merkle_root(const uint8_t* leaves, uint64_t n, uint64_t leaf_len, uint8_t* root) {
  return guard([&] {
    std::vector<H32> lh(n);
    for (uint64_t i = 0; i < n; i++) lh[i] = blake3(leaves + i * leaf_len, leaf_len);
    H32 r = MerkleTree::from_leaf_hashes(lh).root();
    std::memcpy(root, r.data(), 32);
    return (int64_t)0;
  });
}
// Multiproof for sorted distinct indices; returns the byte length (writes if cap suffices).
int64_t lgso_merkle_open(const uint8_t* leaves, uint64_t n, uint64_t leaf_len, const uint64_t* idx,
                         uint64_t k, uint8_t* out, uint64_t cap) {
  return guard([&] {
    std::vector<H32> lh(n);
    for (uint64_t i = 0; i < n; i++) lh[i] = blake3(leaves + i * leaf_len, leaf_len);
    auto t = MerkleTree::from_leaf_hashes(lh);
    std::vector<size_t> ix(idx, idx + k);
    auto pf = t.open(ix);
    if (pf.size() <= cap) std::memcpy(out, pf.data(), pf.size());
    return (int64_t)pf.size();
  });
}
int64_t lgso_merkle_verify(const uint8_t* root, uint64_t n_leaves, const uint8_t* proof,
                           uint64_t plen, const uint64_t* idx, uint64_t k, const uint8_t* leaves,
                           uint64_t leaf_len) {
  return guard([&] {
    H32 r;
    std::memcpy(r.data(), root, 32);
    std::vector<size_t> ix(idx, idx + k);
    std::vector<H32> lh(k);
    for (uint64_t i = 0; i < k; i++) lh[i] = blake3(leaves + i * leaf_len, leaf_len);
    std::vector<uint8_t> pf(proof, proof + plen);
    return (int64_t)merkle_verify(r, n_leaves, pf, ix, lh);
  });
}

// ---- multilinear / sumcheck ----
void lgso_mle_eval(const uint64_t* v, uint64_t nv, const uint64_t* point, uint64_t* out) {
  Ef r = mle_eval(load_ef(v, size_t(1) << nv), load_ef(point, nv));
  out[0] = r.c0;
  out[1] = r.c1;
}
void lgso_new_eq(const uint64_t* point, uint64_t nv, uint64_t* out) {
  store_ef(new_eq(load_ef(point, nv)), out);
}
// polys: d blocks of 2^nv EF. rounds out: nv x (d+1) EF. finals: d EF.
int64_t lgso_sumcheck_prove(uint64_t d, uint64_t nv, const uint64_t* polys, const uint64_t* fields,
                            uint64_t nf, uint64_t* rounds, uint64_t* finals) {
  return guard([&] {
    std::vector<EfVec> ps;
    size_t len = size_t(1) << nv;
    for (uint64_t j = 0; j < d; j++) ps.push_back(load_ef(polys + 2 * j * len, len));
    OracleStream o = make_stream(fields, nf, nullptr, 0);
    EfVec ch;
    SumcheckProof pf = sumcheck_prove(ps, o, ch);
    for (size_t r = 0; r < pf.round_polys.size(); r++) store_ef(pf.round_polys[r], rounds + 2 * r * (d + 1));
    store_ef(pf.final_evals, finals);
    return (int64_t)o.fi;
  });
}
int64_t lgso_sumcheck_verify(uint64_t d, uint64_t nv, const uint64_t* claim, const uint64_t* rounds,
                             const uint64_t* finals, const uint64_t* fields, uint64_t nf) {
  return guard([&] {
    SumcheckProof pf;
    for (uint64_t r = 0; r < nv; r++) pf.round_polys.push_back(load_ef(rounds + 2 * r * (d + 1), d + 1));
    pf.final_evals = load_ef(finals, d);
    OracleStream o = make_stream(fields, nf, nullptr, 0);
    EfVec ch;
    return (int64_t)sumcheck_verify(Ef{claim[0], claim[1]}, pf, o, ch);
  });
}

// ---- basefold: one commitment root (commit_{base,ext}_on_domain, basefold.rs:341-382) ----
int64_t lgso_basefold_commit(const uint64_t* poly, uint64_t nv, int64_t ext, uint64_t li, uint64_t code_rate,
                             uint8_t* root) {
  return guard([&] {
    const size_t len = size_t(1) << nv;
    H32 c;
    if (ext) {
      BfState<Ef> st;
      c = commit_on_domain(load_ef(poly, len), code_rate, li, st);
    } else {
      BfState<uint64_t> st;
      c = commit_on_domain(std::vector<uint64_t>(poly, poly + len), code_rate, li, st);
    }
    std::copy(c.begin(), c.end(), root);
    return int64_t(1);
  });
}

// ---- basefold, one polynomial (basefold.rs:384-642) ----
// Commits poly (2^nv base values, or EF pairs with ext) with log_interleave li and code_rate,
// proves it at `point` (nv EF) from the oracle streams; returns the proof length and writes it
// when cap suffices; used[0..1] = fields / ints drawn; root receives the commitment.
int64_t lgso_basefold_prove(const uint64_t* poly, uint64_t nv, int64_t ext, uint64_t li, uint64_t code_rate,
                            const uint64_t* point, const uint64_t* fields, uint64_t nf, const uint64_t* ints,
                            uint64_t ni, uint8_t* root, uint8_t* out, uint64_t cap, uint64_t* used) {
  return guard([&] {
    const size_t len = size_t(1) << nv;
    OracleStream o = make_stream(fields, nf, ints, ni);
    SingleProof pf;
    H32 c;
    if (ext) {
      BfState<Ef> st;
      c = commit_on_domain(load_ef(poly, len), code_rate, li, st);
      pf = single_prove(st, load_ef(point, nv), o);
    } else {
      BfState<uint64_t> st;
      c = commit_on_domain(std::vector<uint64_t>(poly, poly + len), code_rate, li, st);
      pf = single_prove(st, load_ef(point, nv), o);
    }
    auto bytes = encode_single_proof(pf);
    if (root) std::copy(c.begin(), c.end(), root);
    if (used) {
      used[0] = o.fi;
      used[1] = o.ii;
    }
    if (out && cap >= bytes.size()) std::memcpy(out, bytes.data(), bytes.size());
    return (int64_t)bytes.size();
  });
}

// 1 when the proof decodes and verifies against root at `point` (np EF), else 0.
int64_t lgso_basefold_verify(const uint8_t* root, const uint64_t* point, uint64_t np, const uint8_t* proof,
                             uint64_t plen, int64_t ext, const uint64_t* fields, uint64_t nf, const uint64_t* ints,
                             uint64_t ni) {
  return guard([&] {
    SingleProof pf;
    if (!decode_single_proof(proof, plen, ext != 0, pf)) return int64_t(0);
    H32 c;
    std::copy(root, root + 32, c.begin());
    OracleStream o = make_stream(fields, nf, ints, ni);
    return int64_t(single_verify(c, load_ef(point, np), pf, ext != 0, o) ? 1 : 0);
  });
}

// ---- basefold (shared-domain batch, mixed base/extension polynomials) ----
// bf_polys: concatenated base polys of sizes 2^bf_nv[i]; ef_polys likewise (EF pairs).
// points: concatenated EF points, base polys first. Returns 1 when the oracle's own
// batch_verify accepts the batch_prove output (with `tamper` applied first), 0 otherwise.
int64_t lgso_basefold_batch_roundtrip(const uint64_t* bf_nv, uint64_t nbf, const uint64_t* bf_polys,
                                      const uint64_t* ef_nv, uint64_t nef, const uint64_t* ef_polys,
                                      const uint64_t* points, uint64_t rounds, uint64_t code_rate,
                                      const uint64_t* fields, uint64_t nf, const uint64_t* ints,
                                      uint64_t ni, int64_t tamper) {
  return guard([&] {
    std::vector<BfState<uint64_t>> bst(nbf);
    std::vector<BfState<Ef>> est(nef);
    std::vector<BfClaim> bfc, efc;
    std::vector<std::pair<const BfState<uint64_t>*, EfVec>> bt;
    std::vector<std::pair<const BfState<Ef>*, EfVec>> et;
    size_t off = 0, poff = 0;
    for (uint64_t i = 0; i < nbf; i++) {
      size_t len = size_t(1) << bf_nv[i];
      std::vector<uint64_t> p(bf_polys + off, bf_polys + off + len);
      off += len;
      EfVec pt = load_ef(points + 2 * poff, bf_nv[i]);
      poff += bf_nv[i];
      H32 c = commit_on_domain(p, code_rate, bf_nv[i] - rounds, bst[i]);
      bfc.push_back({c, pt, mle_eval(bst[i].poly, pt), bf_nv[i] - rounds});
      bt.push_back({&bst[i], pt});
    }
    off = 0;
    for (uint64_t i = 0; i < nef; i++) {
      size_t len = size_t(1) << ef_nv[i];
      EfVec p = load_ef(ef_polys + 2 * off, len);
      off += len;
      EfVec pt = load_ef(points + 2 * poff, ef_nv[i]);
      poff += ef_nv[i];
      H32 c = commit_on_domain(p, code_rate, ef_nv[i] - rounds, est[i]);
      efc.push_back({c, pt, mle_eval(p, pt), ef_nv[i] - rounds});
      et.push_back({&est[i], pt});
    }
    OracleStream o = make_stream(fields, nf, ints, ni);
    BatchProof pf = batch_prove(bt, et, o);
    if (tamper == 1) pf.reduction_msgs[0][0] = ef_add(pf.reduction_msgs[0][0], ef_one());
    if (tamper == 2) bfc[0].claim = ef_add(bfc[0].claim, ef_one());
    if (tamper == 3) {
      auto& v = pf.base_bf[0].values.begin()->second;
      v[0] = fadd(v[0], 1);
    }
    if (tamper == 4) pf.final_codeword[0] = ef_add(pf.final_codeword[0], ef_one());
    o.restart();
    return (int64_t)batch_verify(bfc, efc, pf, o);
  });
}

// ---- subset-sum hash ----
// bits: ncols x width bytes (column-major, one byte per bit); a: C x width; out: C x ncols.
void lgso_hash_columns(const uint64_t* a, uint64_t width, const uint8_t* bits, uint64_t ncols,
                       uint64_t* out, int ring64) {
  ProvingKey pk;
  pk.width = width;
  pk.ring64 = ring64 != 0;
  for (size_t r = 0; r < C; r++) pk.a.emplace_back(a + r * width, a + (r + 1) * width);
  for (uint64_t j = 0; j < ncols; j++) {
    std::vector<uint8_t> col(bits + j * width, bits + (j + 1) * width);
    auto d = hash_column(pk, col);
    for (size_t r = 0; r < C; r++) out[r * ncols + j] = d[r];
  }
}

// ---- LigeSIS ----
int64_t lgso_shape_ok(uint64_t log_m, uint64_t log_n) { return shape_ok(log_m, log_n); }

// mode: bit 0 commit the A-bar Basefold state, bit 1 the Z_2^64 hash (then no A-bar state)
void* lgso_setup(uint64_t log_m, uint64_t log_n, const uint64_t* a, int mode) {
  try {
    auto* pk = new ProvingKey(setup(log_m, log_n, a, (mode & 1) != 0 && (mode & 2) == 0));
    pk->ring64 = (mode & 2) != 0;
    return pk;
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}
void lgso_free_pk(void* pk) { delete (ProvingKey*)pk; }
int64_t lgso_pk_com_a(void* pk, uint8_t* out) {
  auto* k = (ProvingKey*)pk;
  if (!k->has_state) return -1;
  std::memcpy(out, k->com_a.data(), 32);
  return 0;
}

void* lgso_commit(void* pk, const uint64_t* f, uint8_t* root, int keep_enc) {
  try {
    auto* k = (ProvingKey*)pk;
    size_t len = size_t(1) << (k->log_m + k->log_n);
    auto* cd = new CommitData();
    H32 r = commit(*k, std::vector<uint64_t>(f, f + len), *cd, keep_enc != 0);
    std::memcpy(root, r.data(), 32);
    return cd;
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}
void lgso_free_cd(void* cd) { delete (CommitData*)cd; }
int64_t lgso_cd_enc(void* cd, uint64_t* out) {
  auto* c = (CommitData*)cd;
  if (c->enc.empty()) return -1;
  std::memcpy(out, c->enc.data(), c->enc.size() * 8);
  return 0;
}
int64_t lgso_cd_h(void* cd, uint64_t* out) {
  auto* c = (CommitData*)cd;
  size_t cols = c->h[0].size();
  for (size_t r = 0; r < C; r++) std::memcpy(out + r * cols, c->h[r].data(), cols * 8);
  return 0;
}

// Returns the proof length; writes it when cap suffices. used[0..1] = fields/ints drawn.
// trace (optional): a_vec (n EF) then ra (2n EF) then idx (128 u64).
int64_t lgso_prove(void* pk, void* cd, const uint64_t* z, const uint64_t* fields, uint64_t nf,
                   const uint64_t* ints, uint64_t ni, uint8_t* out, uint64_t cap, uint64_t* used,
                   uint64_t* trace) {
  return guard([&] {
    auto* k = (ProvingKey*)pk;
    OracleStream o = make_stream(fields, nf, ints, ni);
    ProveTrace tr;
    LigesisProof pf = prove(*k, *(CommitData*)cd, load_ef(z, k->log_m + k->log_n), o, &tr);
    auto bytes = encode_proof(pf);
    if (used) {
      used[0] = o.fi;
      used[1] = o.ii;
    }
    if (trace) {
      store_ef(tr.a_vec, trace);
      store_ef(tr.ra, trace + 2 * tr.a_vec.size());
      for (size_t i = 0; i < tr.idx.size(); i++) trace[2 * (tr.a_vec.size() + tr.ra.size()) + i] = tr.idx[i];
    }
    if (bytes.size() <= cap) std::memcpy(out, bytes.data(), bytes.size());
    return (int64_t)bytes.size();
  });
}

// 1 = accept, 0 = reject, -1 = error (e.g. the oracle stream ran out).
int64_t lgso_verify(uint64_t log_m, uint64_t log_n, const uint8_t* com_a, const uint8_t* root,
                    const uint64_t* z, const uint8_t* proof, uint64_t plen, const uint64_t* fields,
                    uint64_t nf, const uint64_t* ints, uint64_t ni) {
  return guard([&] {
    VerifyingKey vk;
    vk.log_m = log_m;
    vk.log_n = log_n;
    vk.width = ETA << log_m;
    std::memcpy(vk.com_a.data(), com_a, 32);
    H32 r;
    std::memcpy(r.data(), root, 32);
    LigesisProof pf;
    if (!decode_proof(proof, plen, pf)) return (int64_t)0;
    OracleStream o = make_stream(fields, nf, ints, ni);
    return (int64_t)verify(vk, r, load_ef(z, log_m + log_n), pf, o);
  });
}

// Re-encode a decoded proof (round-trip check of the canonical encoding).
int64_t lgso_proof_reencode(const uint8_t* proof, uint64_t plen, uint8_t* out) {
  LigesisProof pf;
  if (!decode_proof(proof, plen, pf)) return -1;
  auto b = encode_proof(pf);
  std::memcpy(out, b.data(), b.size());
  return (int64_t)b.size();
}

}  // extern "C"
