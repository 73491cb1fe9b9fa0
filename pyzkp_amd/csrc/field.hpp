// BN254 prime-field arithmetic for gfx950: Fr (scalar field, modulus r) and
// Fq (base field, modulus q), both 254-bit, in Montgomery form with R = 2^256.
//
// Storage: 8 x 32-bit little-endian limbs (32 bytes), which is bit-identical
// to the 4 x u64 little-endian layout of the reference's device buffers
// (runtime/metal/runtime.py:23-28, runtime/kernels/cpu/kernels.py:135-140).
//
// Multiplication is word-serial Montgomery (CIOS) on 32-bit words with
// 32x32+64 -> 64 multiply-adds (v_mad_u64_u32).  Both moduli have a top word
// below 2^31 - 1, so the CIOS "no final carry word" variant applies: the
// running value t stays below 2p < 2^255 and fits in 8 words, and each round
// costs 8 + 8 + 1 = 17 multiplies (136 per product).
#pragma once
#include <stdint.h>

// Arithmetic is __host__ __device__ so the same code is unit-tested on the
// CPU (a plain C++ build of this header, see hostcheck.cpp).
#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define ZK_HD __host__ __device__ __forceinline__
#else
#define ZK_HD inline
#endif

namespace zk {

typedef uint32_t u32;
typedef uint64_t u64;

struct FrParams {
  static constexpr u32 P[8] = {0xf0000001u, 0x43e1f593u, 0x79b97091u, 0x2833e848u,
                               0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};
  static constexpr u32 PINV = 0xefffffffu;  // -p^{-1} mod 2^32
  static constexpr u32 ONE[8] = {0x4ffffffbu, 0xac96341cu, 0x9f60cd29u, 0x36fc7695u,
                                 0x7879462eu, 0x666ea36fu, 0x9a07df2fu, 0x0e0a77c1u};  // R mod p
  static constexpr u32 R2[8] = {0xae216da7u, 0x1bb8e645u, 0xe35c59e3u, 0x53fe3ab1u,
                                0x53bb8085u, 0x8c49833du, 0x7f4e44a5u, 0x0216d0b1u};  // R^2 mod p
};

struct FqParams {
  static constexpr u32 P[8] = {0xd87cfd47u, 0x3c208c16u, 0x6871ca8du, 0x97816a91u,
                               0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};
  static constexpr u32 PINV = 0xe4866389u;
  static constexpr u32 ONE[8] = {0xc58f0d9du, 0xd35d438du, 0xf5c70b3du, 0x0a78eb28u,
                                 0x7879462cu, 0x666ea36fu, 0x9a07df2fu, 0x0e0a77c1u};
  static constexpr u32 R2[8] = {0x538afa89u, 0xf32cfc5bu, 0xd44501fbu, 0xb5e71911u,
                                0x0a417ff6u, 0x47ab1effu, 0xcab8351fu, 0x06d89f71u};
};

// One field element.  Plain aggregate so that arrays of it are the 32-byte
// wire format and loads can be issued as two 16-byte vector loads.
struct alignas(16) fe {
  u32 w[8];
};

ZK_HD u32 addc(u32 a, u32 b, u32& carry) {
  u64 s = (u64)a + b + carry;
  carry = (u32)(s >> 32);
  return (u32)s;
}

ZK_HD u32 subb(u32 a, u32 b, u32& borrow) {
  u64 d = (u64)a - b - borrow;
  borrow = (u32)(d >> 63);
  return (u32)d;
}

template <class F>
struct Field {
  static ZK_HD fe zero() {
    fe r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.w[i] = 0;
    return r;
  }
  static ZK_HD fe one() {
    fe r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.w[i] = F::ONE[i];
    return r;
  }
  static ZK_HD bool is_zero(const fe& a) {
    u32 x = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) x |= a.w[i];
    return x == 0;
  }
  static ZK_HD bool eq(const fe& a, const fe& b) {
    u32 x = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) x |= a.w[i] ^ b.w[i];
    return x == 0;
  }

  // r = a - p if a >= p else a   (a < 2p)
  static ZK_HD fe reduce_once(const fe& a) {
    fe d;
    u32 br = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) d.w[i] = subb(a.w[i], F::P[i], br);
    // br == 1  <=>  a < p: keep a
    fe r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.w[i] = br ? a.w[i] : d.w[i];
    return r;
  }

  static ZK_HD fe add(const fe& a, const fe& b) {
    fe s;
    u32 c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) s.w[i] = addc(a.w[i], b.w[i], c);
    // a + b < 2p < 2^255: no carry out of the top word
    return reduce_once(s);
  }

  static ZK_HD fe sub(const fe& a, const fe& b) {
    fe d;
    u32 br = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) d.w[i] = subb(a.w[i], b.w[i], br);
    if (br) {
      u32 c = 0;
#pragma unroll
      for (int i = 0; i < 8; i++) d.w[i] = addc(d.w[i], F::P[i], c);
    }
    return d;
  }

  static ZK_HD fe neg(const fe& a) {
    if (is_zero(a)) return a;
    fe d;
    u32 br = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) d.w[i] = subb(F::P[i], a.w[i], br);
    return d;
  }

  static ZK_HD fe dbl(const fe& a) { return add(a, a); }

  // Montgomery product a*b*R^{-1} mod p, CIOS without the extra carry word.
  static ZK_HD fe mul(const fe& a, const fe& b) {
    u32 t[8];
#pragma unroll
    for (int i = 0; i < 8; i++) t[i] = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const u32 bi = b.w[i];
      u64 s = (u64)a.w[0] * bi + t[0];
      u32 A = (u32)(s >> 32);
      const u32 t0 = (u32)s;
      const u32 m = t0 * F::PINV;
      u64 s2 = (u64)m * F::P[0] + t0;
      u32 C = (u32)(s2 >> 32);
#pragma unroll
      for (int j = 1; j < 8; j++) {
        s = (u64)a.w[j] * bi + t[j] + A;
        A = (u32)(s >> 32);
        s2 = (u64)m * F::P[j] + (u32)s + C;
        C = (u32)(s2 >> 32);
        t[j - 1] = (u32)s2;
      }
      t[7] = C + A;
    }
    fe r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.w[i] = t[i];
    return reduce_once(r);
  }

  static ZK_HD fe sqr(const fe& a) { return mul(a, a); }

  // Montgomery -> canonical: a * 1 * R^{-1}
  static ZK_HD fe from_mont(const fe& a) {
    fe one_plain = zero();
    one_plain.w[0] = 1;
    return mul(a, one_plain);
  }
  // canonical -> Montgomery: a * R^2 * R^{-1}.  Input must be < p.
  static ZK_HD fe to_mont(const fe& a) {
    fe r2;
#pragma unroll
    for (int i = 0; i < 8; i++) r2.w[i] = F::R2[i];
    return mul(a, r2);
  }
  // canonical a (any 256-bit value) -> a mod p, for inputs below 2^256 < 6p
  static ZK_HD fe reduce_full(const fe& a) {
    fe x = a;
#pragma unroll
    for (int k = 0; k < 5; k++) {
      fe d;
      u32 br = 0;
#pragma unroll
      for (int i = 0; i < 8; i++) d.w[i] = subb(x.w[i], F::P[i], br);
#pragma unroll
      for (int i = 0; i < 8; i++) x.w[i] = br ? x.w[i] : d.w[i];
    }
    return x;
  }
};

typedef Field<FrParams> Fr;
typedef Field<FqParams> Fq;

#if defined(__HIP__)
__device__ __forceinline__ fe load_fe(const fe* p) {
  const uint4* q = reinterpret_cast<const uint4*>(p);
  uint4 a = q[0], b = q[1];
  fe r;
  r.w[0] = a.x; r.w[1] = a.y; r.w[2] = a.z; r.w[3] = a.w;
  r.w[4] = b.x; r.w[5] = b.y; r.w[6] = b.z; r.w[7] = b.w;
  return r;
}

__device__ __forceinline__ void store_fe(fe* p, const fe& v) {
  uint4* q = reinterpret_cast<uint4*>(p);
  q[0] = make_uint4(v.w[0], v.w[1], v.w[2], v.w[3]);
  q[1] = make_uint4(v.w[4], v.w[5], v.w[6], v.w[7]);
}
#endif

}  // namespace zk
