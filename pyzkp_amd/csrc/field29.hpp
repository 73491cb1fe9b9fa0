// BN254 Fq / Fr arithmetic on 9 x 29-bit limbs for the hot kernels
// (MSM bucket accumulation and reduction, NTT butterflies).
//
// Why 29 bits: with 29-bit limbs a product-scanning Montgomery multiply
// (FIPS: every column of a*b + m*p summed in one 64-bit accumulator) never
// overflows 64 bits (at most 18 products of < 2^58 plus a < 2^35 carry), so
// each partial product is exactly one v_mad_u64_u32 with the accumulator as
// its 64-bit addend and no carry handling.  The 8 x 32-bit CIOS form needs
// an add-with-carry and register moves around every multiply-add; on gfx950
// this form issues about 2.5x fewer VALU instructions per product.
//
// Representation: Montgomery with R = 2^261 (9 * 29), values kept lazily
// in [0, 2p).  R > 4p, so mul() takes inputs below 2p and returns below 2p
// without a final subtraction.  add / sub / neg return values in [0, 2p);
// canonical() maps to [0, p).  unpack / pack convert to and from the
// 32-byte little-endian wire format.  The carry-free forms (add_cf, sub_cf,
// neg2p_cf, neg_cf, sub_4p_a_2b_cf) add and subtract limb by limb, so limbs
// may pass 29 bits while the value stays exact; their callers state the limb
// bounds that keep every product column below 2^64.  normalize() and
// reduce_est() bring limbs back to 29 bits.
//
// Domain bookkeeping: a 32-byte buffer holding x * 2^256 (the reference's
// Montgomery form) read as an R = 2^261 value represents x * 2^-5.
// Linear maps (the NTT) can therefore run on such buffers unchanged, with
// twiddles in R = 2^261 form; the point arithmetic converts coordinates with
// to261() / to256() (one product each).
#pragma once
#include "field.hpp"

namespace zk {

struct f9 {
  u32 l[9];
};

static constexpr u32 M29 = (1u << 29) - 1;

struct Fq29Params {
  static constexpr u32 P[9] = {0x187cfd47u, 0x010460b6u, 0x1c72a34fu, 0x02d522d0u, 0x1585d978u, 0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu};
  static constexpr u32 P2[9] = {0x10f9fa8eu, 0x0208c16du, 0x18e5469eu, 0x05aa45a1u, 0x0b0bb2f0u, 0x05b68181u, 0x014dc282u, 0x1cb84c68u, 0x0060c89cu};
  static constexpr u32 P3[9] = {0x0976f7d5u, 0x030d2224u, 0x1557e9edu, 0x087f6872u, 0x00918c68u, 0x0891c242u, 0x01f4a3c3u, 0x0b14729cu, 0x00912cebu};
  static constexpr u32 P4[9] = {0x01f3f51cu, 0x041182dbu, 0x11ca8d3cu, 0x0b548b43u, 0x161765e0u, 0x0b6d0302u, 0x029b8504u, 0x197098d0u, 0x00c19139u};
  static constexpr u32 PINV = 0x04866389u;  // -p^-1 mod 2^29
  // 2p with limbs 0..7 raised by 2^29 (borrowed from the next limb): every
  // limb of a normalised value below 1.2p is at most the matching limb here
  static constexpr u32 K2P[9] = {0x30f9fa8eu, 0x2208c16cu, 0x38e5469du, 0x25aa45a0u, 0x2b0bb2efu, 0x25b68180u, 0x214dc281u, 0x3cb84c67u, 0x0060c89bu};
  // 4p in the same raised form
  static constexpr u32 K4P[9] = {0x21f3f51cu, 0x241182dau, 0x31ca8d3bu, 0x2b548b42u, 0x361765dfu, 0x2b6d0301u, 0x229b8503u, 0x397098cfu, 0x00c19138u};
  // 4p with limbs 0..7 raised by 2^31 (at least 3 * 2^29 each)
  static constexpr u32 K4P31[9] = {0x81f3f51cu, 0x841182d7u, 0x91ca8d38u, 0x8b548b3fu, 0x961765dcu, 0x8b6d02feu, 0x829b8500u, 0x997098ccu, 0x00c19135u};
  static constexpr u32 ONE[9] = {0x157ccc21u, 0x141c2758u, 0x185230d3u, 0x014c0419u, 0x0aa36fb9u, 0x1d4240ceu, 0x11d54c07u, 0x052ac7a8u, 0x000dc836u};                    // 2^261 mod p
  static constexpr u32 TO261[9] = {0x13349ca1u, 0x1a5d84a8u, 0x0a3e5cacu, 0x100249e0u, 0x12b951e8u, 0x0e92d304u, 0x14cb95b3u, 0x041b9d3du, 0x00058003u};         // 2^266 mod p: x 2^256 -> x 2^261
  static constexpr u32 TO256[9] = {0x058f0d9du, 0x1aea1c6eu, 0x11c2cf74u, 0x11d651ebu, 0x1462c0a7u, 0x11b7bc3cu, 0x1cbd99bau, 0x183340fbu, 0x000e0a77u};         // 2^256 mod p: x 2^261 -> x 2^256
  static constexpr u32 R2_522[9] = {0x059bac10u, 0x0d1503a3u, 0x018016b8u, 0x10ab0ca8u, 0x02632639u, 0x02c0169fu, 0x169bfd53u, 0x11869d4cu, 0x002a11a6u};        // 2^522 mod p: plain x -> x 2^261
};
struct Fr29Params {
  static constexpr u32 P[9] = {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u, 0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu};
  static constexpr u32 P2[9] = {0x00000002u, 0x1e1f593fu, 0x1cb848a1u, 0x0fa121e6u, 0x0b0ba506u, 0x05b68181u, 0x014dc282u, 0x1cb84c68u, 0x0060c89cu};
  static constexpr u32 P3[9] = {0x10000003u, 0x1d2f05deu, 0x0b146cf2u, 0x1771b2dau, 0x00917789u, 0x0891c242u, 0x01f4a3c3u, 0x0b14729cu, 0x00912cebu};
  static constexpr u32 P4[9] = {0x00000004u, 0x1c3eb27eu, 0x19709143u, 0x1f4243cdu, 0x16174a0cu, 0x0b6d0302u, 0x029b8504u, 0x197098d0u, 0x00c19139u};
  static constexpr u32 PINV = 0x0fffffffu;  // -p^-1 mod 2^29
  static constexpr u32 K2P[9] = {0x20000002u, 0x3e1f593eu, 0x3cb848a0u, 0x2fa121e5u, 0x2b0ba505u, 0x25b68180u, 0x214dc281u, 0x3cb84c67u, 0x0060c89bu};
  static constexpr u32 K4P[9] = {0x20000004u, 0x3c3eb27du, 0x39709142u, 0x3f4243ccu, 0x36174a0bu, 0x2b6d0301u, 0x229b8503u, 0x397098cfu, 0x00c19138u};
  static constexpr u32 K4P31[9] = {0x80000004u, 0x9c3eb27au, 0x9970913fu, 0x9f4243c9u, 0x96174a08u, 0x8b6d02feu, 0x829b8500u, 0x997098ccu, 0x00c19135u};
  static constexpr u32 ONE[9] = {0x0fffff57u, 0x1ea70ab4u, 0x052c068bu, 0x17504f49u, 0x0aa8075bu, 0x1d4240ceu, 0x11d54c07u, 0x052ac7a8u, 0x000dc836u};                    // 2^261 mod p
  static constexpr u32 TO261[9] = {0x0fffead7u, 0x1d5444f4u, 0x04438aa5u, 0x03b4d096u, 0x134c84dau, 0x0e92d304u, 0x14cb95b3u, 0x041b9d3du, 0x00058003u};         // 2^266 mod p: x 2^256 -> x 2^261
  static constexpr u32 TO256[9] = {0x0ffffffbu, 0x04b1a0e2u, 0x18334a6bu, 0x18ed2b3eu, 0x1462e36fu, 0x11b7bc3cu, 0x1cbd99bau, 0x183340fbu, 0x000e0a77u};         // 2^256 mod p: x 2^261 -> x 2^256
};

// Each multiply-add of a product column is chained through the accumulator:
// the empty asm makes the running sum opaque, so the compiler keeps one
// v_mad_u64_u32 per partial product instead of re-associating a column into
// separate sums that then need 64-bit additions (about 14 fewer
// instructions per product; the waves of a SIMD hide the chain's latency).
#if defined(__HIP_DEVICE_COMPILE__)
#define ZK_OPAQUE(x) asm("" : "+v"(x))
#else
#define ZK_OPAQUE(x) (void)0
#endif
#define ZK_MAD(acc, x, y)        \
  do {                           \
    (acc) += (u64)(x) * (y);     \
    ZK_OPAQUE(acc);              \
  } while (0)

template <class F>
struct Mont29 {
  static ZK_HD f9 zero() {
    f9 r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.l[i] = 0;
    return r;
  }
  static ZK_HD f9 one() {
    f9 r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.l[i] = F::ONE[i];
    return r;
  }
  static ZK_HD f9 constant(const u32 (&c)[9]) {
    f9 r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.l[i] = c[i];
    return r;
  }

  // Reduction part of FIPS column k: add the m*p terms, then either pick
  // m_k (k < 9) so the column's low 29 bits vanish, or emit limb k - 9.
  static ZK_HD void redc_column(u64& acc, u32 (&m)[9], f9& r, int k) {
#pragma unroll
    for (int i = 0; i < 9; i++) {
      const int j = k - i;
      if (i < k && j >= 0 && j < 9) ZK_MAD(acc, m[i], F::P[j]);
    }
    if (k < 9) {
      m[k] = ((u32)acc * F::PINV) & M29;
      ZK_MAD(acc, m[k], F::P[0]);
      acc >>= 29;
    } else {
      r.l[k - 9] = (u32)acc & M29;
      acc >>= 29;
    }
  }

  // FIPS Montgomery product: columns k = 0..16 of a*b + m*p in one u64.
  // Output < p + a*b/R: below 1.03p for inputs below 2p, below 2p for
  // inputs whose product is below R*p (e.g. 13p x 13p).
  static ZK_HD f9 mul(const f9& a, const f9& b) {
    u32 m[9];
    f9 r;
    u64 acc = 0;
#pragma unroll
    for (int k = 0; k < 17; k++) {
#pragma unroll
      for (int i = 0; i < 9; i++) {
        const int j = k - i;
        if (j >= 0 && j < 9) ZK_MAD(acc, a.l[i], b.l[j]);
      }
      redc_column(acc, m, r, k);
    }
    r.l[8] = (u32)acc;
    return r;
  }

  // two independent products a*b / R and c*d / R with their columns
  // interleaved: two accumulator chains, so twice the instruction-level
  // parallelism of two mul() calls for the same instructions (for callers
  // with few waves, where one chain's latency is exposed)
  static ZK_HD void mul_x2(const f9& a, const f9& b, const f9& c, const f9& d, f9& r, f9& s) {
    u32 m[9], n[9];
    u64 acc = 0, acc2 = 0;
#pragma unroll
    for (int k = 0; k < 17; k++) {
#pragma unroll
      for (int i = 0; i < 9; i++) {
        const int j = k - i;
        if (j >= 0 && j < 9) {
          ZK_MAD(acc, a.l[i], b.l[j]);
          ZK_MAD(acc2, c.l[i], d.l[j]);
        }
      }
      redc_column(acc, m, r, k);
      redc_column(acc2, n, s, k);
    }
    r.l[8] = (u32)acc;
    s.l[8] = (u32)acc2;
  }

  // (a*b + c*d) / R with one reduction: a column holds at most 27 products
  // of < 2^58 plus a < 2^35 carry, still below 2^64.
  static ZK_HD f9 mul2(const f9& a, const f9& b, const f9& c, const f9& d) {
    u32 m[9];
    f9 r;
    u64 acc = 0;
#pragma unroll
    for (int k = 0; k < 17; k++) {
#pragma unroll
      for (int i = 0; i < 9; i++) {
        const int j = k - i;
        if (j >= 0 && j < 9) {
          ZK_MAD(acc, a.l[i], b.l[j]);
          ZK_MAD(acc, c.l[i], d.l[j]);
        }
      }
      redc_column(acc, m, r, k);
    }
    r.l[8] = (u32)acc;
    return r;
  }

  // squaring: cross products once with a doubled operand; v (any 32-bit
  // limbs, the sum below 2^261) is added to the result through the high
  // columns, so the output is a^2 / R + v, normalised, with no separate
  // addition pass
  static ZK_HD f9 sqr_plus(const f9& a, const f9& v) {
    u32 a2[9];
#pragma unroll
    for (int i = 0; i < 9; i++) a2[i] = a.l[i] << 1;
    u32 m[9];
    f9 r;
    u64 acc = 0;
#pragma unroll
    for (int k = 0; k < 17; k++) {
#pragma unroll
      for (int i = 0; i < 9; i++) {
        const int j = k - i;
        if (j > i && j < 9) ZK_MAD(acc, a2[i], a.l[j]);
      }
      if ((k & 1) == 0 && k / 2 < 9) ZK_MAD(acc, a.l[k / 2], a.l[k / 2]);
      if (k >= 9) acc += v.l[k - 9];
      redc_column(acc, m, r, k);
    }
    r.l[8] = (u32)acc + v.l[8];
    return r;
  }
  static ZK_HD f9 sqr(const f9& a) { return sqr_plus(a, zero()); }
  // two independent squarings with their columns interleaved (as mul_x2)
  static ZK_HD void sqr_x2(const f9& a, const f9& c, f9& r, f9& s) {
    u32 a2[9], c2[9];
#pragma unroll
    for (int i = 0; i < 9; i++) {
      a2[i] = a.l[i] << 1;
      c2[i] = c.l[i] << 1;
    }
    u32 m[9], n[9];
    u64 acc = 0, acc2 = 0;
#pragma unroll
    for (int k = 0; k < 17; k++) {
#pragma unroll
      for (int i = 0; i < 9; i++) {
        const int j = k - i;
        if (j > i && j < 9) {
          ZK_MAD(acc, a2[i], a.l[j]);
          ZK_MAD(acc2, c2[i], c.l[j]);
        }
      }
      if ((k & 1) == 0 && k / 2 < 9) {
        ZK_MAD(acc, a.l[k / 2], a.l[k / 2]);
        ZK_MAD(acc2, c.l[k / 2], c.l[k / 2]);
      }
      redc_column(acc, m, r, k);
      redc_column(acc2, n, s, k);
    }
    r.l[8] = (u32)acc;
    s.l[8] = (u32)acc2;
  }

  // s - K if s >= K else s (K normalised limbs; s - K must fit below 2^261)
  static ZK_HD f9 sub_if_ge(const f9& s, const u32 (&K)[9]) {
    f9 d;
    int32_t br = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
      const int32_t t = (int32_t)s.l[i] - (int32_t)K[i] + br;
      d.l[i] = (u32)t & M29;
      br = t >> 29;
    }
    f9 r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.l[i] = br ? s.l[i] : d.l[i];
    return r;
  }

  // 2p - a for a <= 2p, no reduction: result in [0, 2p]
  static ZK_HD f9 neg_nr(const f9& a) {
    f9 s;
    int32_t c = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
      const int32_t t = (int32_t)F::P2[i] - (int32_t)a.l[i] + c;
      s.l[i] = (u32)t & M29;
      c = t >> 29;
    }
    return s;
  }

  // 4p - a - 2b, no reduction; needs a + 2b <= 4p
  static ZK_HD f9 sub_4p_a_2b(const f9& a, const f9& b) {
    f9 s;
    int32_t c = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
      const int32_t t = (int32_t)F::P4[i] - (int32_t)a.l[i] - (int32_t)(b.l[i] << 1) + c;
      s.l[i] = (u32)t & M29;
      c = t >> 29;
    }
    return s;
  }

  // x - q p with q = floor(x_8 / (p_8 + 1)), for any x below 2^261 whose
  // limbs 0..7 are below 2^32: q p <= x, and the result is normalised and
  // below p + p / 2^18 (limbs 0..7 past 29 bits add under 8 * 2^232 to x
  // beyond x_8 2^232, a few millionths of p; normalised limbs give p + p / 2^20)
  static ZK_HD f9 reduce_est(const f9& x) {
    const u32 q = x.l[8] / (F::P[8] + 1);
    f9 r;
    int64_t c = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
      const int64_t t = (int64_t)x.l[i] - (int64_t)((u64)q * F::P[i]) + c;
      r.l[i] = (u32)t & M29;
      c = t >> 29;
    }
    return r;
  }

  // a in [0, 4p): a == 0 mod p  <=>  a is 0, p, 2p or 3p
  static ZK_HD bool is_zero_lt4p(const f9& a) {
    u32 z = 0, e1 = 0, e2 = 0, e3 = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
      z |= a.l[i];
      e1 |= a.l[i] ^ F::P[i];
      e2 |= a.l[i] ^ F::P2[i];
      e3 |= a.l[i] ^ F::P3[i];
    }
    return z == 0 || e1 == 0 || e2 == 0 || e3 == 0;
  }

  // necessary condition for is_zero_lt4p: the top limb of 0, p, 2p or 3p
  static ZK_HD bool top_may_be_zero_lt4p(const f9& a) {
    const u32 t = a.l[8];
    return t == 0 || t == F::P[8] || t == F::P2[8] || t == F::P3[8];
  }

  // s -> s - 2p if s >= 2p  (s < 4p, limbs normalised)
  static ZK_HD f9 reduce2p(const f9& s) { return sub_if_ge(s, F::P2); }

  static ZK_HD f9 add(const f9& a, const f9& b) {
    f9 s;
    u32 c = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
      const u32 t = a.l[i] + b.l[i] + c;
      s.l[i] = t & M29;
      c = t >> 29;
    }
    return reduce2p(s);
  }

  // a - b + 2p in (0, 4p), then into [0, 2p)
  static ZK_HD f9 sub(const f9& a, const f9& b) {
    f9 s;
    int32_t c = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
      const int32_t t = (int32_t)a.l[i] - (int32_t)b.l[i] + (int32_t)F::P2[i] + c;
      s.l[i] = (u32)t & M29;
      c = t >> 29;
    }
    return reduce2p(s);
  }

  // a + b and a - b + 2p without the final reduction (normalised limbs):
  // for a, b < 2p the results are below 4p, which mul() accepts
  static ZK_HD f9 add_nr(const f9& a, const f9& b) {
    f9 s;
    u32 c = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
      const u32 t = a.l[i] + b.l[i] + c;
      s.l[i] = t & M29;
      c = t >> 29;
    }
    return s;
  }
  static ZK_HD f9 sub_nr(const f9& a, const f9& b) {
    f9 s;
    int32_t c = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
      const int32_t t = (int32_t)a.l[i] - (int32_t)b.l[i] + (int32_t)F::P2[i] + c;
      s.l[i] = (u32)t & M29;
      c = t >> 29;
    }
    return s;
  }

  // Carry-free forms (the NTT butterflies).  Limbs may exceed 29 bits; the
  // value sum_i l_i 2^(29 i) is exact as long as no limb leaves [0, 2^32).
  // a + b, limb by limb
  static ZK_HD f9 add_cf(const f9& a, const f9& b) {
    f9 s;
#pragma unroll
    for (int i = 0; i < 9; i++) s.l[i] = a.l[i] + b.l[i];
    return s;
  }
  // a - b + 2p, limb by limb, for b normalised with value below 1.99p (then
  // b's limbs are at most K2P's: its lower limbs are below 2^29, K2P's above,
  // and its top limb is at most K2P's, 2 p_8 - 1, so no limb goes negative);
  // limbs grow by less than 2^30
  static ZK_HD f9 sub_cf(const f9& a, const f9& b) {
    f9 s;
#pragma unroll
    for (int i = 0; i < 9; i++) s.l[i] = a.l[i] + (F::K2P[i] - b.l[i]);
    return s;
  }
  // 2p - b, limb by limb, for b normalised and below 1.2p; limbs below 2^30
  static ZK_HD f9 neg2p_cf(const f9& b) {
    f9 s;
#pragma unroll
    for (int i = 0; i < 9; i++) s.l[i] = F::K2P[i] - b.l[i];
    return s;
  }
  // 4p - a - 2b, limb by limb, for a and b normalised and below 1.02p (the
  // top limbs then sum below K4P31's); limbs below 2^32
  static ZK_HD f9 sub_4p_a_2b_cf(const f9& a, const f9& b) {
    f9 s;
#pragma unroll
    for (int i = 0; i < 9; i++) s.l[i] = F::K4P31[i] - (a.l[i] + (b.l[i] << 1));
    return s;
  }
  // 4p - b, limb by limb, for b normalised and below 2p (its top limb is
  // then below K4P's); limbs below 2^30
  static ZK_HD f9 neg_cf(const f9& b) {
    f9 s;
#pragma unroll
    for (int i = 0; i < 9; i++) s.l[i] = F::K4P[i] - b.l[i];
    return s;
  }
  // carry propagation back to 29-bit limbs (same value; it must be below 2^261)
  static ZK_HD f9 normalize(const f9& a) {
    f9 s;
    u32 c = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
      const u32 t = a.l[i] + c;
      s.l[i] = (i < 8) ? (t & M29) : t;
      c = t >> 29;
    }
    return s;
  }

  static ZK_HD f9 dbl(const f9& a) { return add(a, a); }
  static ZK_HD f9 neg(const f9& a) { return sub(zero(), a); }

  // a in [0, 2p): a == 0 mod p  <=>  a == 0 or a == p
  static ZK_HD bool is_zero(const f9& a) {
    u32 z = 0, e = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
      z |= a.l[i];
      e |= a.l[i] ^ F::P[i];
    }
    return z == 0 || e == 0;
  }

  // [0, 2p) -> [0, p)
  static ZK_HD f9 canonical(const f9& a) {
    f9 d;
    int32_t br = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
      const int32_t t = (int32_t)a.l[i] - (int32_t)F::P[i] + br;
      d.l[i] = (u32)t & M29;
      br = t >> 29;
    }
    f9 r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.l[i] = br ? a.l[i] : d.l[i];
    return r;
  }

  static ZK_HD f9 to261(const f9& x256) { return mul(x256, constant(F::TO261)); }
  static ZK_HD f9 to256(const f9& x261) { return mul(x261, constant(F::TO256)); }
};

typedef Mont29<Fq29Params> Fq9;
typedef Mont29<Fr29Params> Fr9;

// 8 x 32-bit words (value < 2^256) <-> 9 x 29-bit limbs
ZK_HD f9 unpack9(const fe& x) {
  f9 r;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const int bit = 29 * i, w = bit >> 5, sh = bit & 31;
    u32 v = x.w[w] >> sh;
    if (sh > 3 && w + 1 < 8) v |= x.w[w + 1] << (32 - sh);
    r.l[i] = v & M29;
  }
  return r;
}

// value must be below 2^256
ZK_HD fe pack9(const f9& a) {
  fe r;
#pragma unroll
  for (int w = 0; w < 8; w++) {
    const int bit = 32 * w, i = bit / 29, sh = bit % 29;
    u32 v = a.l[i] >> sh;
    if (i + 1 < 9) v |= a.l[i + 1] << (29 - sh);
    if (sh > 26 && i + 2 < 9) v |= a.l[i + 2] << (58 - sh);
    r.w[w] = v;
  }
  return r;
}

#if defined(__HIP__)
__device__ __forceinline__ f9 load_f9(const f9* p) {
  const u32* q = reinterpret_cast<const u32*>(p);
  f9 r;
#pragma unroll
  for (int i = 0; i < 9; i++) r.l[i] = q[i];
  return r;
}
__device__ __forceinline__ void store_f9(f9* p, const f9& v) {
  u32* q = reinterpret_cast<u32*>(p);
#pragma unroll
  for (int i = 0; i < 9; i++) q[i] = v.l[i];
}
#endif

}  // namespace zk
