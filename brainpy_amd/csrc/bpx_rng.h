// bpx_rng.h -- LFSR88 (L'Ecuyer's taus88 as written in the reference,
// brainpy/_src/math/tifunc.py:47-180) and exact u32 division by an invariant
// divisor, usable on host and device.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define BPX_HD __host__ __device__ __forceinline__
#else
#define BPX_HD inline
#endif

namespace bpx {

struct Lfsr88 {
  uint32_t s0, s1, s2;

  // lfsr88_key (tifunc.py:67): (seed+1, seed+7, seed+15) with u32 wrap
  BPX_HD explicit Lfsr88(uint32_t seed) : s0(seed + 1u), s1(seed + 7u), s2(seed + 15u) {}

  // lfsr88_next_key (tifunc.py:80-86); returns k0^k1^k2 of the new state
  BPX_HD uint32_t next() {
#if defined(__HIP_DEVICE_COMPILE__)
    // ((s & ~(2^j - 1)) << k) ^ (t >> r) with j + k = 32 - r: the two parts
    // occupy disjoint bits, so the update is the funnel shift of
    // {s >> j, t} right by r (v_alignbit_b32), the same bits in 4
    // instructions per component instead of 6
    s0 = __builtin_amdgcn_alignbit(s0 >> 1, (s0 << 13) ^ s0, 19u);
    s1 = __builtin_amdgcn_alignbit(s1 >> 3, (s1 << 2) ^ s1, 25u);
    s2 = __builtin_amdgcn_alignbit(s2 >> 4, (s2 << 3) ^ s2, 11u);
#else
    uint32_t b = ((s0 << 13) ^ s0) >> 19;
    s0 = ((s0 & 4294967294u) << 12) ^ b;
    b = ((s1 << 2) ^ s1) >> 25;
    s1 = ((s1 & 4294967288u) << 4) ^ b;
    b = ((s2 << 3) ^ s2) >> 11;
    s2 = ((s2 & 4294967280u) << 17) ^ b;
#endif
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(s0, s1, s2, 0x96);   // s0 ^ s1 ^ s2 in one instruction
#else
    return s0 ^ s1 ^ s2;
#endif
  }
};

// Exact n / d and n % d for u32 n and an invariant u32 d >= 1, after
// Granlund & Montgomery, "Division by invariant integers using
// multiplication" (PLDI 1994), Fig. 4.1:
//   l = ceil(log2 d); m = floor(2^32 (2^l - d) / d) + 1;
//   t = mulhi(m, n); q = (t + ((n - t) >> min(l,1))) >> max(l-1,0).
struct DivU32 {
  uint32_t d, m, s1, s2;

  BPX_HD static DivU32 make(uint32_t d) {
    DivU32 r;
    r.d = d;
    uint32_t l = 0;
    while (l < 32 && (uint64_t(1) << l) < d) ++l;
    uint64_t two_l = uint64_t(1) << l;
    r.m = (uint32_t)(((uint64_t(1) << 32) * (two_l - d)) / d + 1);
    r.s1 = l < 1 ? l : 1;
    r.s2 = l > 0 ? l - 1 : 0;
    return r;
  }

  BPX_HD uint32_t div(uint32_t n) const {
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t t = __umulhi(m, n);
#else
    uint32_t t = (uint32_t)(((uint64_t)m * n) >> 32);
#endif
    return (t + ((n - t) >> s1)) >> s2;
  }
  BPX_HD uint32_t mod(uint32_t n) const { return n - div(n) * d; }
  // same result when 256 < d < 2^24 (then q = n / d < 2^24 and q * d fits the
  // full-rate 24-bit multiplier)
  BPX_HD uint32_t mod24(uint32_t n) const {
#if defined(__HIP_DEVICE_COMPILE__)
    return n - __umul24(div(n), d);
#else
    return n - div(n) * d;
#endif
  }
  BPX_HD bool small_quotient() const { return d > 256u && d < (1u << 24); }
};

// Short form of the same division, for divisors that admit it (about 3 in 4
// below 2^18, 19999 among them): with m = ceil(2^(32+s) / d) < 2^32 and
// m*d - 2^(32+s) <= 2^s, q = mulhi(n, m) >> s equals n / d for every u32 n
// (Granlund & Montgomery 1994, Theorem 4.2 with N = 32).  When also
// 512 < d < 2^23, the quotient is below 2^23 and the remainder is one signed
// 24-bit multiply-add, n + q * (-d).
struct DivShort {
  uint32_t d, m, s;

  // false when no such m exists
  static bool make(uint32_t d, DivShort* out) {
    if (d < 2) return false;
    for (uint32_t sh = 0; sh < 32; ++sh) {
      const unsigned __int128 two_p = (unsigned __int128)1 << (32 + sh);
      const unsigned __int128 m = (two_p + d - 1) / d;
      if (m >= ((unsigned __int128)1 << 32)) return false;
      if (m * d - two_p <= ((unsigned __int128)1 << sh)) {
        out->d = d;
        out->m = (uint32_t)m;
        out->s = sh;
        return true;
      }
    }
    return false;
  }
  static bool mad24_ok(uint32_t d) { return d > 512u && d < (1u << 23); }

  BPX_HD uint32_t div(uint32_t n) const {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umulhi(m, n) >> s;
#else
    return (uint32_t)(((uint64_t)m * n) >> 32) >> s;
#endif
  }
  // n % d, for divisors with mad24_ok(d)
  BPX_HD uint32_t mod24(uint32_t n) const {
#if defined(__HIP_DEVICE_COMPILE__)
    return n + (uint32_t)__mul24((int)div(n), -(int)d);
#else
    return n - div(n) * d;
#endif
  }
};

// lfsr88_random_integers(key, 0, clen-1) and (key, 1, clen) (tifunc.py:136-147):
// the divisor (high + 1 - low) equals clen at every call site of the jitconn
// kernels, so both reduce to bits % clen plus low.
BPX_HD int32_t rand_first(Lfsr88& k, const DivU32& dv) { return (int32_t)dv.mod(k.next()); }
BPX_HD int32_t rand_inc(Lfsr88& k, const DivU32& dv) { return (int32_t)(dv.mod(k.next()) + 1u); }
template <bool Q24>
BPX_HD uint32_t rand_inc_u(Lfsr88& k, const DivU32& dv) {
  return (Q24 ? dv.mod24(k.next()) : dv.mod(k.next())) + 1u;
}

}  // namespace bpx
