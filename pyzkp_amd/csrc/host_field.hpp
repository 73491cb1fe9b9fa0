// Host-side BN254 Fr/Fq Montgomery arithmetic (4 x 64-bit limbs, R = 2^256).
// Used by the library's host code for table generation and for the final
// window combination of an MSM; never on the per-element device path.
#pragma once
#include <stdint.h>
#include <string.h>

namespace zkh {

typedef uint64_t u64;
typedef unsigned __int128 u128;

struct Fe { u64 v[4]; };

struct Mod {
  u64 p[4];
  u64 pinv;   // -p^{-1} mod 2^64
  u64 one[4]; // R mod p
  u64 r2[4];  // R^2 mod p
};

static const Mod FR = {
    {0x43e1f593f0000001ull, 0x2833e84879b97091ull, 0xb85045b68181585dull, 0x30644e72e131a029ull},
    0xc2e1f593efffffffull,
    {0xac96341c4ffffffbull, 0x36fc76959f60cd29ull, 0x666ea36f7879462eull, 0x0e0a77c19a07df2full},
    {0x1bb8e645ae216da7ull, 0x53fe3ab1e35c59e3ull, 0x8c49833d53bb8085ull, 0x0216d0b17f4e44a5ull}};

static const Mod FQ = {
    {0x3c208c16d87cfd47ull, 0x97816a916871ca8dull, 0xb85045b68181585dull, 0x30644e72e131a029ull},
    0x87d20782e4866389ull,
    {0xd35d438dc58f0d9dull, 0x0a78eb28f5c70b3dull, 0x666ea36f7879462cull, 0x0e0a77c19a07df2full},
    {0xf32cfc5b538afa89ull, 0xb5e71911d44501fbull, 0x47ab1eff0a417ff6ull, 0x06d89f71cab8351full}};

inline bool geq(const u64* a, const u64* b) {
  for (int i = 3; i >= 0; i--) {
    if (a[i] != b[i]) return a[i] > b[i];
  }
  return true;
}

inline void sub_raw(u64* r, const u64* a, const u64* b) {
  u64 br = 0;
  for (int i = 0; i < 4; i++) {
    u128 d = (u128)a[i] - b[i] - br;
    r[i] = (u64)d;
    br = (u64)(d >> 127);
  }
}

inline bool is_zero(const Fe& a) { return (a.v[0] | a.v[1] | a.v[2] | a.v[3]) == 0; }
inline bool eq(const Fe& a, const Fe& b) { return memcmp(a.v, b.v, 32) == 0; }

// The reductions below select with masks, not branches: whether a sum or a
// Montgomery product needs the final subtraction is data-dependent and
// close to a coin flip, and a mispredicted branch there costs more than the
// whole product (measured 24.8 vs 3.6 ns per Fq product on 8 independent
// chains).
//
// r <- hi 2^256 + r - p if that is >= 0, else r unchanged (hi is 0 or 1)
inline void csub(u64* r, const u64* p, u64 hi = 0) {
  u64 d[4];
  u64 br = 0;
  for (int i = 0; i < 4; i++) {
    const u128 z = (u128)r[i] - p[i] - br;
    d[i] = (u64)z;
    br = (u64)(z >> 127);
  }
  const u64 keep = (u64)0 - (u64)((hi != 0) | (br == 0));
  for (int i = 0; i < 4; i++) r[i] = (d[i] & keep) | (r[i] & ~keep);
}

inline Fe add(const Mod& M, const Fe& a, const Fe& b) {
  Fe r;
  u64 c = 0;
  for (int i = 0; i < 4; i++) {
    u128 s = (u128)a.v[i] + b.v[i] + c;
    r.v[i] = (u64)s;
    c = (u64)(s >> 64);
  }
  csub(r.v, M.p, c);
  return r;
}

inline Fe sub(const Mod& M, const Fe& a, const Fe& b) {
  Fe r;
  u64 br = 0;
  for (int i = 0; i < 4; i++) {
    const u128 z = (u128)a.v[i] - b.v[i] - br;
    r.v[i] = (u64)z;
    br = (u64)(z >> 127);
  }
  // a < b: add p back (the sum wraps past 2^256 exactly once)
  const u64 m = (u64)0 - br;
  u64 c = 0;
  for (int i = 0; i < 4; i++) {
    const u128 s = (u128)r.v[i] + (M.p[i] & m) + c;
    r.v[i] = (u64)s;
    c = (u64)(s >> 64);
  }
  return r;
}

inline Fe neg(const Mod& M, const Fe& a) {
  Fe r;
  sub_raw(r.v, M.p, a.v);
  const u64 nz = (u64)0 - (u64)!is_zero(a);
  for (int i = 0; i < 4; i++) r.v[i] &= nz;
  return r;
}

// CIOS Montgomery product with an explicit carry word (any a, b below
// 2^256 with a b < 2^256 p; output below p for canonical inputs).  The
// modulus is a template argument, so its words are immediates and the
// limb loops unroll completely.
template <u64 P0, u64 P1, u64 P2, u64 P3, u64 PINV>
inline Fe mul_mod(const Fe& a, const Fe& b) {
  constexpr u64 p[4] = {P0, P1, P2, P3};
  u64 t0 = 0, t1 = 0, t2 = 0, t3 = 0, t4 = 0;
#pragma GCC unroll 4
  for (int i = 0; i < 4; i++) {
    const u64 bi = b.v[i];
    u128 s = (u128)a.v[0] * bi + t0;
    t0 = (u64)s;
    u64 c = (u64)(s >> 64);
    s = (u128)a.v[1] * bi + t1 + c;
    t1 = (u64)s;
    c = (u64)(s >> 64);
    s = (u128)a.v[2] * bi + t2 + c;
    t2 = (u64)s;
    c = (u64)(s >> 64);
    s = (u128)a.v[3] * bi + t3 + c;
    t3 = (u64)s;
    c = (u64)(s >> 64);
    s = (u128)t4 + c;
    t4 = (u64)s;
    const u64 t5 = (u64)(s >> 64);
    const u64 m = t0 * PINV;
    s = (u128)m * p[0] + t0;
    c = (u64)(s >> 64);
    s = (u128)m * p[1] + t1 + c;
    t0 = (u64)s;
    c = (u64)(s >> 64);
    s = (u128)m * p[2] + t2 + c;
    t1 = (u64)s;
    c = (u64)(s >> 64);
    s = (u128)m * p[3] + t3 + c;
    t2 = (u64)s;
    c = (u64)(s >> 64);
    s = (u128)t4 + c;
    t3 = (u64)s;
    t4 = t5 + (u64)(s >> 64);
  }
  Fe r = {{t0, t1, t2, t3}};
  csub(r.v, p, t4);
  return r;
}

// the same product for a modulus known only at run time
inline Fe mul_any(const Mod& M, const Fe& a, const Fe& b) {
  u64 t[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 4; i++) {
    u64 c = 0;
    for (int j = 0; j < 4; j++) {
      u128 s = (u128)a.v[j] * b.v[i] + t[j] + c;
      t[j] = (u64)s;
      c = (u64)(s >> 64);
    }
    u128 s = (u128)t[4] + c;
    t[4] = (u64)s;
    t[5] = (u64)(s >> 64);
    u64 m = t[0] * M.pinv;
    s = (u128)m * M.p[0] + t[0];
    c = (u64)(s >> 64);
    for (int j = 1; j < 4; j++) {
      s = (u128)m * M.p[j] + t[j] + c;
      t[j - 1] = (u64)s;
      c = (u64)(s >> 64);
    }
    s = (u128)t[4] + c;
    t[3] = (u64)s;
    t[4] = t[5] + (u64)(s >> 64);
  }
  Fe r;
  memcpy(r.v, t, 32);
  csub(r.v, M.p, t[4]);
  return r;
}

inline Fe mul(const Mod& M, const Fe& a, const Fe& b) {
  if (&M == &FQ)
    return mul_mod<0x3c208c16d87cfd47ull, 0x97816a916871ca8dull, 0xb85045b68181585dull, 0x30644e72e131a029ull,
                   0x87d20782e4866389ull>(a, b);
  if (&M == &FR)
    return mul_mod<0x43e1f593f0000001ull, 0x2833e84879b97091ull, 0xb85045b68181585dull, 0x30644e72e131a029ull,
                   0xc2e1f593efffffffull>(a, b);
  return mul_any(M, a, b);
}

inline Fe one(const Mod& M) { Fe r; memcpy(r.v, M.one, 32); return r; }
inline Fe zero() { Fe r = {{0, 0, 0, 0}}; return r; }

inline Fe to_mont(const Mod& M, const Fe& a) {
  Fe r2; memcpy(r2.v, M.r2, 32);
  return mul(M, a, r2);
}
inline Fe from_mont(const Mod& M, const Fe& a) {
  Fe o = zero(); o.v[0] = 1;
  return mul(M, a, o);
}
inline Fe from_u64(const Mod& M, u64 x) {
  Fe a = zero(); a.v[0] = x;
  return to_mont(M, a);
}

// a^e for a Montgomery-form base and a plain 256-bit exponent
inline Fe pow(const Mod& M, const Fe& a, const u64 e[4]) {
  Fe r = one(M);
  for (int i = 3; i >= 0; i--) {
    for (int b = 63; b >= 0; b--) {
      r = mul(M, r, r);
      if ((e[i] >> b) & 1) r = mul(M, r, a);
    }
  }
  return r;
}
inline Fe pow_u64(const Mod& M, const Fe& a, u64 e) {
  u64 ee[4] = {e, 0, 0, 0};
  return pow(M, a, ee);
}
// inverse by Fermat: a^(p-2); returns 0 for 0
inline Fe inv(const Mod& M, const Fe& a) {
  u64 e[4];
  memcpy(e, M.p, 32);
  e[0] -= 2;  // p is odd and > 2: no borrow out of limb 0
  return pow(M, a, e);
}

}  // namespace zkh
