// Goldilocks (p = 2^64 - 2^32 + 1) and F_{p^2} = F_p[X]/(X^2 - 7) arithmetic for host and
// device code. Values are kept canonical (< p) after every operation, which is what the
// reference serializes (utils/src/merkle.rs:24-29) and compares.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#define LGS_HD __host__ __device__ __forceinline__

namespace lgs {

constexpr uint64_t GL_P = 0xffffffff00000001ull;
constexpr uint64_t GL_EPS = 0xffffffffull;  // 2^64 mod p
constexpr uint64_t GL_G32 = 1753635133440165772ull;  // generator of the order-2^32 subgroup
constexpr uint64_t GL_INV2 = 0x7fffffff80000001ull;  // (p + 1) / 2

LGS_HD uint64_t gl_canon(uint64_t x) { return x >= GL_P ? x - GL_P : x; }

// Canonical inputs and outputs. Written as 32-bit carry chains (v_add_co / v_addc_co on the
// device: six full-rate instructions for an add, five for a subtract, instead of 64-bit adds
// followed by 64-bit compares to recover the carries).
LGS_HD uint64_t gl_add(uint64_t a, uint64_t b) {
  // s = a + b mod 2^64 and t = s + (2^32 - 1) mod 2^64 = a + b - p mod 2^64. The true sum is
  // >= p exactly when either addition carries: a carry out of a + b means a + b >= 2^64 > p,
  // and then t = a + b - p < p without a second wrap; else s + 2^32 - 1 >= 2^64 iff s >= p.
  unsigned c0, c1, c2, c3;
  const unsigned sl = __builtin_addc((unsigned)a, (unsigned)b, 0u, &c0);
  const unsigned sh = __builtin_addc((unsigned)(a >> 32), (unsigned)(b >> 32), c0, &c1);
  const unsigned tl = __builtin_addc(sl, 0xffffffffu, 0u, &c2);
  const unsigned th = __builtin_addc(sh, 0u, c2, &c3);
  return (c1 | c3) ? (((uint64_t)th << 32) | tl) : (((uint64_t)sh << 32) | sl);
}

// gl_add as one 64-bit add, two 64-bit compares and a 64-bit add of 0 or 2^32 - 1
// (v_lshl_add_u64 / v_cmp_lt_u64 on gfx950): s = a + b wraps exactly when the true sum reaches
// 2^64, and then, as when s >= p, s + 2^32 - 1 = s - p mod 2^64 is the reduced sum. One VALU
// instruction fewer than the carry chains where the zero high word stays in a register (the row
// transforms' butterflies).
LGS_HD uint64_t gl_add64(uint64_t a, uint64_t b) {
  const uint64_t s = a + b;
  const uint32_t m = ((s < a) | (s >= GL_P)) ? 0xffffffffu : 0u;
  return s + (uint64_t)m;
}

LGS_HD uint64_t gl_sub(uint64_t a, uint64_t b) {
  // a - b, and on a borrow a - b + 2^64 - (2^32 - 1) = a - b + p
  unsigned b0, b1, b2, b3;
  const unsigned dl = __builtin_subc((unsigned)a, (unsigned)b, 0u, &b0);
  const unsigned dh = __builtin_subc((unsigned)(a >> 32), (unsigned)(b >> 32), b0, &b1);
  const unsigned rl = __builtin_subc(dl, b1 ? 0xffffffffu : 0u, 0u, &b2);
  const unsigned rh = __builtin_subc(dh, 0u, b2, &b3);
  return ((uint64_t)rh << 32) | rl;
}

LGS_HD uint64_t gl_neg(uint64_t a) { return a ? GL_P - a : 0; }

// (hi * 2^64 + lo) mod p, using 2^64 = 2^32 - 1 and 2^96 = -1. (A 32-bit carry-chain form of
// this reduction measured slower on gfx950: its carry reads need wait states.)
LGS_HD uint64_t gl_reduce(uint64_t lo, uint64_t hi) {
  uint64_t hh = hi >> 32, hl = hi & GL_EPS;
  uint64_t t0 = lo - hh;
  t0 -= (lo < hh) ? GL_EPS : 0;
  uint64_t t1 = (hl << 32) - hl;
  uint64_t r = t0 + t1;
  r += (r < t1) ? GL_EPS : 0;
  return gl_canon(r);
}

// Full 64 x 64 -> 128-bit product from four 32 x 32 + 64 multiply-adds (v_mad_u64_u32); no
// partial sum can overflow: (2^32 - 1)^2 + 2 (2^32 - 1) = 2^64 - 1.
LGS_HD void mul_wide(uint64_t a, uint64_t b, uint64_t& lo, uint64_t& hi) {
  const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
  const uint64_t p00 = (uint64_t)a0 * b0;
  const uint64_t p01 = (uint64_t)a0 * b1 + (p00 >> 32);
  const uint64_t p10 = (uint64_t)a1 * b0 + (uint32_t)p01;
  hi = (uint64_t)a1 * b1 + (p01 >> 32) + (p10 >> 32);
  lo = (p10 << 32) | (uint32_t)p00;
}

LGS_HD uint64_t gl_mul(uint64_t a, uint64_t b) {
  uint64_t lo, hi;
  mul_wide(a, b, lo, hi);
  return gl_reduce(lo, hi);
}

// a * w * 2^-64 mod p, canonical, for a < 2^64 and w < p: the product of a and a factor kept in
// Montgomery form (w = c 2^64 mod p gives a c). Montgomery reduction of T = hi 2^64 + lo:
// p^-1 = 1 + 2^32 (mod 2^64), so m = lo (1 + 2^32) mod 2^64 makes T - m p divisible by 2^64, and
// (T - m p) / 2^64 = hi - k with k = floor(m p / 2^64) = m - (m >> 32) - e, e the carry out of
// lo + (lo << 32) (e = 1 exactly when the high half of m is below its low half). T < 2^64 p
// gives hi < p, and k < p, so hi - k lies in (-p, p): the result is hi - k, plus p on a borrow,
// and needs no final compare. On 32-bit halves: one add, two two-word subtractions and a
// conditional -(2^32 - 1) -- against the 128-bit reduction's two 64-bit corrections, 32 x 32
// product and final compare.
LGS_HD uint64_t gl_mul_mont(uint64_t a, uint64_t w) {
  uint64_t lo, hi;
  mul_wide(a, w, lo, hi);
  unsigned e, c0, c1, d0, d1, f0, f1;
  const unsigned ml = (unsigned)lo;
  const unsigned mh = __builtin_addc((unsigned)(lo >> 32), ml, 0u, &e);
  const unsigned kl = __builtin_subc(ml, mh, e, &c0);
  const unsigned kh = __builtin_subc(mh, 0u, c0, &c1);
  unsigned rl = __builtin_subc((unsigned)hi, kl, 0u, &d0);
  unsigned rh = __builtin_subc((unsigned)(hi >> 32), kh, d0, &d1);
  // a borrow: r + 2^64 - (2^32 - 1) = hi - k + p, in (0, p)
  rl = __builtin_subc(rl, d1 ? 0xffffffffu : 0u, 0u, &f0);
  rh = __builtin_subc(rh, 0u, f0, &f1);
  return ((uint64_t)rh << 32) | rl;
}

// x * 2^-K mod p for canonical x and 1 <= K <= 32 (K a constant where this is inlined). Since
// p = 1 (mod 2^32), t = -x mod 2^K makes x + t p divisible by 2^K, and
// (x + t p) / 2^K = (x + t) / 2^K + t 2^(64-K) - t 2^(32-K): the first two terms occupy disjoint
// bits, x + t < 2^64, and the subtraction cannot borrow (t 2^(64-K) >= t 2^(32-K)); one
// canonicalization remains. No 128-bit product or reduction.
LGS_HD uint64_t gl_mul_inv_pow2(uint64_t x, int K) {
  const uint64_t t = (0 - x) & ((uint64_t(1) << K) - 1);
  const uint64_t u = x + t;
  const uint64_t a = (u >> K) | (t << (64 - K));
  return gl_canon(a - (t << (32 - K)));
}

// x * 2^e mod p for canonical x and 32 < e < 64 (e a constant where this is inlined), without
// a 128-bit reduction. With x = x1 2^32 + x0 and 2^(32+e) = 2^(e-32) 2^64 = 2^e - 2^(e-32):
// x 2^e = s 2^e - x1 2^(e-32), s = x0 + x1 = s0 + c 2^32. Splitting s0 at bit 64 - e,
// s0 2^e = (s0 mod 2^(64-e)) 2^e + (s0 >> (64-e)) (2^32 - 1), and c 2^(32+e) = c (2^e - 2^(e-32)),
// so x 2^e = (rotl(s0, e - 32) + c 2^(e-32)) 2^32 - D with D = (s0 >> (64-e)) + (x1 + c) 2^(e-32)
// < 2^64 (x1 + c cannot wrap: x1 = 2^32 - 1 only with x0 = 0 for x < p). A carry k out of the
// high word is k 2^64 = k (2^32 - 1); a borrow of the final subtraction adds p. The high word is
// below 2^32, so the result is canonical without a compare.
LGS_HD uint64_t gl_mul_pow2_mid(uint64_t x, uint32_t e) {
  const unsigned x0 = (unsigned)x, x1 = (unsigned)(x >> 32);
  unsigned c, k, b0, b1, f0, f1;
  const unsigned s0 = __builtin_addc(x0, x1, 0u, &c);
  const unsigned xc = x1 + c;
  const unsigned rot = (s0 << (e - 32)) | (s0 >> (64 - e));
  const unsigned h = __builtin_addc(rot, c << (e - 32), 0u, &k);
  const unsigned dlo = (s0 >> (64 - e)) | (xc << (e - 32));
  const unsigned dhi = xc >> (64 - e);
  unsigned rl = __builtin_subc(k ? 0xffffffffu : 0u, dlo, 0u, &b0);
  unsigned rh = __builtin_subc(h, dhi, b0, &b1);
  // a borrow: r + 2^64 - (2^32 - 1) = r + p
  rl = __builtin_subc(rl, b1 ? 0xffffffffu : 0u, 0u, &f0);
  rh = __builtin_subc(rh, 0u, f0, &f1);
  return ((uint64_t)rh << 32) | rl;
}

LGS_HD uint64_t gl_pow(uint64_t a, uint64_t e) {
  uint64_t r = 1;
  while (e) {
    if (e & 1) r = gl_mul(r, a);
    a = gl_mul(a, a);
    e >>= 1;
  }
  return r;
}

LGS_HD uint64_t gl_inv(uint64_t a) { return gl_pow(a, GL_P - 2); }

// Generator of the order-2^bits subgroup (p3 TwoAdicField::two_adic_generator).
LGS_HD uint64_t gl_root(int bits) {
  uint64_t g = GL_G32;
  for (int i = bits; i < 32; i++) g = gl_mul(g, g);
  return g;
}

// Sum of up to 2^32 products x y of canonical values, kept as the integer lo + 2^64 hi +
// 2^128 top and reduced once at the end (2^128 = -2^32 mod p): a multiply-add costs one wide
// product and a few adds instead of a reduction.
struct WideAcc {
  uint64_t lo, hi;
  uint32_t top;
};
LGS_HD WideAcc wide_zero() { return WideAcc{0, 0, 0}; }
LGS_HD void wide_mul_add(WideAcc& a, uint64_t x, uint64_t y) {
  uint64_t pl, ph;
  mul_wide(x, y, pl, ph);
  a.lo += pl;
  // ph <= (p - 1)^2 / 2^64 < 2^64 - 2^33 + 2, so ph + carry does not wrap
  const uint64_t t = ph + (a.lo < pl ? 1 : 0);
  a.hi += t;
  a.top += a.hi < t ? 1u : 0u;
}
LGS_HD void wide_add(WideAcc& a, uint64_t x) {
  a.lo += x;
  if (a.lo < x) {
    a.hi += 1;
    a.top += a.hi == 0 ? 1u : 0u;
  }
}
LGS_HD uint64_t wide_reduce(const WideAcc& a) {
  return gl_sub(gl_reduce(a.lo, a.hi), gl_canon((uint64_t)a.top << 32));
}

struct ef2 {
  uint64_t c0, c1;
};

LGS_HD ef2 ef_make(uint64_t a, uint64_t b) { return ef2{a, b}; }
LGS_HD ef2 ef_base(uint64_t a) { return ef2{a, 0}; }
LGS_HD ef2 ef_add(ef2 a, ef2 b) { return ef2{gl_add(a.c0, b.c0), gl_add(a.c1, b.c1)}; }
LGS_HD ef2 ef_sub(ef2 a, ef2 b) { return ef2{gl_sub(a.c0, b.c0), gl_sub(a.c1, b.c1)}; }
LGS_HD ef2 ef_neg(ef2 a) { return ef2{gl_neg(a.c0), gl_neg(a.c1)}; }
LGS_HD bool ef_eq(ef2 a, ef2 b) { return a.c0 == b.c0 && a.c1 == b.c1; }
// (a0 + a1 X)(b0 + b1 X) = (a0 b0 + 7 a1 b1) + (a0 b1 + a1 b0) X with two reductions of 128-bit
// sums instead of four of single products:
//  c0: a0 b0 < p^2 < 2^128 - 2^97 and 7 (a1 b1 mod p) < 2^67, so their sum fits 128 bits;
//  c1: a0 b1 + a1 b0 < 2^129; a carry out of bit 127 is 2^128 = -2^32 (mod p).
LGS_HD ef2 ef_mul(ef2 a, ef2 b) {
  uint64_t l00, h00, l01, h01, l10, h10, l11, h11;
  mul_wide(a.c0, b.c0, l00, h00);
  mul_wide(a.c1, b.c1, l11, h11);
  mul_wide(a.c0, b.c1, l01, h01);
  mul_wide(a.c1, b.c0, l10, h10);
  const uint64_t v = gl_reduce(l11, h11);
  // 7 v = 8 v - v as (hi7, lo7), hi7 < 8
  const uint64_t v8 = v << 3;
  const uint64_t lo7 = v8 - v;
  const uint64_t hi7 = (v >> 61) - (v8 < v ? 1 : 0);
  const uint64_t lo0 = l00 + lo7;
  const uint64_t hi0 = h00 + hi7 + (lo0 < lo7 ? 1 : 0);
  const uint64_t lo1 = l01 + l10;
  const uint64_t c = lo1 < l10 ? 1 : 0;
  const uint64_t hs = h01 + c;  // h01 < 2^64 - 2^33, no wrap
  const uint64_t hi1 = hs + h10;
  const uint64_t r1 = gl_reduce(lo1, hi1);
  return ef2{gl_reduce(lo0, hi0), hi1 < hs ? gl_sub(r1, 0x100000000ull) : r1};
}
LGS_HD ef2 ef_mul_base(ef2 a, uint64_t b) { return ef2{gl_mul(a.c0, b), gl_mul(a.c1, b)}; }
LGS_HD ef2 ef_inv(ef2 a) {
  uint64_t norm = gl_sub(gl_mul(a.c0, a.c0), gl_mul(7, gl_mul(a.c1, a.c1)));
  uint64_t ni = gl_inv(norm);
  return ef2{gl_mul(a.c0, ni), gl_mul(gl_neg(a.c1), ni)};
}
LGS_HD ef2 ef_one() { return ef2{1, 0}; }
LGS_HD ef2 ef_zero() { return ef2{0, 0}; }

}  // namespace lgs
