// Oracle (test infrastructure only): CPU restatement of the Goldilocks field and its degree-2
// extension used by the reference. Only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline may load this code; the product library never links it.
//
// Restates the third-party crates p3-goldilocks / p3-field 0.5.3 (reference Cargo.lock:333-367):
//   p  = 2^64 - 2^32 + 1                      (reference README.md:65)
//   EF = F_p[X]/(X^2 - 7)                     (BinomialExtensionField<Goldilocks,2>, README.md:66)
//   two_adic_generator(k) = g32^(2^(32-k)),  g32 = 1753635133440165772
// Pinned by: brute force against Python big integers (tests/test_oracle_field.py), the
// two-adic chain g_1 = p-1, g_2 = 2^48, g_3 = p-2^24 (closed forms), and X^2 = 7 being a
// non-residue (Euler's criterion).
#pragma once
#include <cstdint>
#include <cstddef>
#include <vector>

namespace lgso {

typedef unsigned __int128 u128;
constexpr uint64_t P = 0xffffffff00000001ULL;
constexpr uint64_t EPS = 0xffffffffULL;  // 2^64 mod p
constexpr uint64_t EXT_W = 7;            // X^2 = EXT_W
constexpr uint64_t G32 = 1753635133440165772ULL;

inline uint64_t canon(uint64_t x) { return x >= P ? x - P : x; }

inline uint64_t fadd(uint64_t a, uint64_t b) {
  uint64_t s = a + b;
  if (s < a) s += EPS;  // carry out of 2^64 is worth 2^32 - 1
  return canon(s);
}

inline uint64_t fsub(uint64_t a, uint64_t b) { return a >= b ? a - b : a + (P - b); }

inline uint64_t fneg(uint64_t a) { return a == 0 ? 0 : P - a; }

// x mod p for x < 2^128, using 2^64 = 2^32 - 1 and 2^96 = -1 (mod p).
inline uint64_t reduce128(u128 x) {
  uint64_t lo = (uint64_t)x;
  uint64_t hi = (uint64_t)(x >> 64);
  uint64_t hh = hi >> 32, hl = hi & EPS;
  uint64_t t0 = lo - hh;
  if (lo < hh) t0 -= EPS;
  uint64_t t1 = hl * EPS;
  uint64_t r = t0 + t1;
  if (r < t1) r += EPS;
  return canon(r);
}

inline uint64_t fmul(uint64_t a, uint64_t b) { return reduce128((u128)a * b); }

inline uint64_t fpow(uint64_t a, uint64_t e) {
  uint64_t r = 1;
  while (e) {
    if (e & 1) r = fmul(r, a);
    a = fmul(a, a);
    e >>= 1;
  }
  return r;
}

inline uint64_t finv(uint64_t a) { return fpow(a, P - 2); }

// p3 TwoAdicField::two_adic_generator(bits): generator of the order-2^bits subgroup.
inline uint64_t two_adic_generator(int bits) {
  uint64_t g = G32;
  for (int i = bits; i < 32; i++) g = fmul(g, g);
  return g;
}

struct Ef {
  uint64_t c0, c1;
};

inline bool operator==(const Ef& a, const Ef& b) { return a.c0 == b.c0 && a.c1 == b.c1; }
inline bool operator!=(const Ef& a, const Ef& b) { return !(a == b); }

inline Ef ef(uint64_t b) { return Ef{b, 0}; }
inline Ef ef_zero() { return Ef{0, 0}; }
inline Ef ef_one() { return Ef{1, 0}; }
inline Ef ef_add(Ef a, Ef b) { return Ef{fadd(a.c0, b.c0), fadd(a.c1, b.c1)}; }
inline Ef ef_sub(Ef a, Ef b) { return Ef{fsub(a.c0, b.c0), fsub(a.c1, b.c1)}; }
inline Ef ef_neg(Ef a) { return Ef{fneg(a.c0), fneg(a.c1)}; }
inline Ef ef_mul(Ef a, Ef b) {
  uint64_t c0 = fadd(fmul(a.c0, b.c0), fmul(EXT_W, fmul(a.c1, b.c1)));
  uint64_t c1 = fadd(fmul(a.c0, b.c1), fmul(a.c1, b.c0));
  return Ef{c0, c1};
}
inline Ef ef_mul_base(Ef a, uint64_t b) { return Ef{fmul(a.c0, b), fmul(a.c1, b)}; }
inline Ef ef_inv(Ef a) {
  // (a0 + a1 X)^-1 = (a0 - a1 X) / (a0^2 - 7 a1^2)
  uint64_t norm = fsub(fmul(a.c0, a.c0), fmul(EXT_W, fmul(a.c1, a.c1)));
  uint64_t ni = finv(norm);
  return Ef{fmul(a.c0, ni), fmul(fneg(a.c1), ni)};
}

inline Ef lift(uint64_t b) { return ef(b); }
inline Ef lift(Ef e) { return e; }

}  // namespace lgso
