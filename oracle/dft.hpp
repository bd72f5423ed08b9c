// Oracle (test infrastructure only): natural-order radix-2 DFT.
//
// Restates p3-dft 0.5.3 `Radix2Dit::dft` / `dft_algebra` as called by the reference
// (pcs/src/ligesis.rs:225-235, pcs/src/basefold.rs:351-358, 371-378):
//   out[j] = sum_k c_k * g^(j*k),  g = two_adic_generator(log2 len)
// `dft_algebra` over EF uses base-field twiddles, i.e. it is the component-wise base DFT.
// Pinned by the O(n^2) closed form (tests/test_oracle_dft.py) and by the reference's own
// rs_weights identity (pcs/src/ligesis.rs:501) exercised inside prove().
#pragma once
#include "field.hpp"

namespace lgso {

inline int log2_exact(size_t n) {
  int l = 0;
  while ((size_t(1) << l) < n) l++;
  return l;
}

inline void dft_base(std::vector<uint64_t>& a) {
  size_t n = a.size();
  if (n <= 1) return;
  int logn = log2_exact(n);
  for (size_t i = 0, j = 0; i < n; i++) {  // bit-reversal permutation
    if (i < j) std::swap(a[i], a[j]);
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j |= bit;
  }
  for (int s = 1; s <= logn; s++) {
    size_t len = size_t(1) << s, half = len >> 1;
    uint64_t wl = two_adic_generator(s);
    std::vector<uint64_t> tw(half);
    tw[0] = 1;
    for (size_t j = 1; j < half; j++) tw[j] = fmul(tw[j - 1], wl);
    for (size_t i = 0; i < n; i += len) {
      for (size_t j = 0; j < half; j++) {
        uint64_t u = a[i + j], v = fmul(a[i + j + half], tw[j]);
        a[i + j] = fadd(u, v);
        a[i + j + half] = fsub(u, v);
      }
    }
  }
}

inline void dft_ext(std::vector<Ef>& a) {
  std::vector<uint64_t> x(a.size()), y(a.size());
  for (size_t i = 0; i < a.size(); i++) {
    x[i] = a[i].c0;
    y[i] = a[i].c1;
  }
  dft_base(x);
  dft_base(y);
  for (size_t i = 0; i < a.size(); i++) a[i] = Ef{x[i], y[i]};
}

inline void dft_any(std::vector<uint64_t>& a) { dft_base(a); }
inline void dft_any(std::vector<Ef>& a) { dft_ext(a); }

}  // namespace lgso
