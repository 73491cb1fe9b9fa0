// CPU build of the device field / curve arithmetic (field.hpp, curve.hpp are
// __host__ __device__).  Exposes a few entry points so the host-logic tests
// can check the exact code the kernels run against the oracle without a GPU.
#include <vector>
#include <string.h>
#include "field.hpp"
#include "curve.hpp"

using namespace zk;

static fe rd(const uint64_t* w) {
  fe r;
  for (int i = 0; i < 4; i++) {
    r.w[2 * i] = (uint32_t)w[i];
    r.w[2 * i + 1] = (uint32_t)(w[i] >> 32);
  }
  return r;
}
static void wr(uint64_t* w, const fe& r) {
  for (int i = 0; i < 4; i++) w[i] = (uint64_t)r.w[2 * i] | ((uint64_t)r.w[2 * i + 1] << 32);
}
static xyzz rdp(const uint64_t* w) {
  xyzz p;
  p.x = rd(w);
  p.y = rd(w + 4);
  p.zz = rd(w + 8);
  p.zzz = rd(w + 12);
  return p;
}
static void wrp(uint64_t* w, const xyzz& p) {
  wr(w, p.x);
  wr(w + 4, p.y);
  wr(w + 8, p.zz);
  wr(w + 12, p.zzz);
}

extern "C" {
// op: 0 mul, 1 add, 2 sub, 3 neg, 4 to_mont, 5 from_mont ; field: 0 Fr, 1 Fq
int zkhost_field(int field, int op, const uint64_t* a, const uint64_t* b, uint64_t* out, uint64_t n) {
  for (uint64_t i = 0; i < n; i++) {
    fe x = rd(a + 4 * i), y = b ? rd(b + 4 * i) : Fr::zero(), r;
    if (field == 0) {
      switch (op) {
        case 0: r = Fr::mul(x, y); break;
        case 1: r = Fr::add(x, y); break;
        case 2: r = Fr::sub(x, y); break;
        case 3: r = Fr::neg(x); break;
        case 4: r = Fr::to_mont(Fr::reduce_full(x)); break;
        case 5: r = Fr::from_mont(x); break;
        default: return 1;
      }
    } else {
      switch (op) {
        case 0: r = Fq::mul(x, y); break;
        case 1: r = Fq::add(x, y); break;
        case 2: r = Fq::sub(x, y); break;
        case 3: r = Fq::neg(x); break;
        case 4: r = Fq::to_mont(Fq::reduce_full(x)); break;
        case 5: r = Fq::from_mont(x); break;
        default: return 1;
      }
    }
    wr(out + 4 * i, r);
  }
  return 0;
}

// XYZZ ops on Montgomery Fq points (16 words each):
// op 0: add(p, q)  1: dbl(p)  2: madd(p, affine q = words 0..7 of q)  3: mul_small(p, k)
int zkhost_xyzz(int op, const uint64_t* p, const uint64_t* q, uint32_t k, uint64_t* out) {
  xyzz P = rdp(p), R;
  switch (op) {
    case 0: R = xyzz_add(P, rdp(q)); break;
    case 1: R = xyzz_dbl(P); break;
    case 2: {
      aff a;
      a.x = rd(q);
      a.y = rd(q + 4);
      R = P;
      if (!aff_is_inf(a)) xyzz_madd(R, a);
      break;
    }
    case 3: R = xyzz_mul_small(P, k); break;
    default: return 1;
  }
  wrp(out, R);
  return 0;
}
}

// ---- 29-bit limb arithmetic (field29.hpp / curve29.hpp), same word I/O:
// inputs / outputs are 4 x u64 little-endian values (< 2^256).
#include "curve29.hpp"

extern "C" {
// op: 0 mul, 1 add, 2 sub, 3 neg, 4 sqr, 5 canonical, 6 to261, 7 to256, 8 is_zero; field 0 Fr 1 Fq
int zkhost_field29(int field, int op, const uint64_t* a, const uint64_t* b, uint64_t* out, uint64_t n) {
  for (uint64_t i = 0; i < n; i++) {
    f9 x = unpack9(rd(a + 4 * i)), y = b ? unpack9(rd(b + 4 * i)) : Fr9::zero(), r;
    if (field == 0) {
      switch (op) {
        case 0: r = Fr9::mul(x, y); break;
        case 1: r = Fr9::add(x, y); break;
        case 2: r = Fr9::sub(x, y); break;
        case 3: r = Fr9::neg(x); break;
        case 4: r = Fr9::sqr(x); break;
        case 5: r = Fr9::canonical(x); break;
        case 6: r = Fr9::to261(x); break;
        case 7: r = Fr9::to256(x); break;
        case 8: r = Fr9::zero(); r.l[0] = Fr9::is_zero(x); break;
        default: return 1;
      }
    } else {
      switch (op) {
        case 0: r = Fq9::mul(x, y); break;
        case 1: r = Fq9::add(x, y); break;
        case 2: r = Fq9::sub(x, y); break;
        case 3: r = Fq9::neg(x); break;
        case 4: r = Fq9::sqr(x); break;
        case 5: r = Fq9::canonical(x); break;
        case 6: r = Fq9::to261(x); break;
        case 7: r = Fq9::to256(x); break;
        case 8: r = Fq9::zero(); r.l[0] = Fq9::is_zero(x); break;
        default: return 1;
      }
    }
    wr(out + 4 * i, pack9(r));
  }
  return 0;
}

// XYZZ on 29-bit limbs: points given as 16 words of R = 2^261 values
// op 0: add(p, q)  1: dbl(p)  2: madd(p, affine q)  3: mul_small(p, k)  4: k doublings of p,
// failing (2) if a coordinate leaves [0, 2p) on the way
int zkhost_xyzz29(int op, const uint64_t* p, const uint64_t* q, uint32_t k, uint64_t* out) {
  xyzz9 P, R;
  P.x = unpack9(rd(p)); P.y = unpack9(rd(p + 4)); P.zz = unpack9(rd(p + 8)); P.zzz = unpack9(rd(p + 12));
  switch (op) {
    case 0: {
      xyzz9 Q;
      Q.x = unpack9(rd(q)); Q.y = unpack9(rd(q + 4)); Q.zz = unpack9(rd(q + 8)); Q.zzz = unpack9(rd(q + 12));
      R = xyzz9_add(P, Q);
      break;
    }
    case 1: R = xyzz9_dbl(P); break;
    case 4: {
      // below 2p with 29-bit limbs: s - 2p would borrow, sub_if_ge keeps s
      auto ok = [](const f9& c) {
        const f9 d = Fq9::sub_if_ge(c, Fq29Params::P2);
        for (int j = 0; j < 9; j++)
          if (d.l[j] != c.l[j] || (j < 8 && (c.l[j] >> 29))) return false;
        return true;
      };
      R = P;
      for (uint32_t i = 0; i < k; i++) {
        R = xyzz9_dbl(R);
        if (!ok(R.x) || !ok(R.y) || !ok(R.zz) || !ok(R.zzz)) return 2;
      }
      break;
    }
    case 2: {
      aff9 a;
      a.x = unpack9(rd(q));
      a.y = unpack9(rd(q + 4));
      R = P;
      if (!aff9_is_inf(a)) xyzz9_madd(R, a);
      break;
    }
    case 3: R = xyzz9_mul_small(P, k); break;
    default: return 1;
  }
  wr(out, pack9(R.x)); wr(out + 4, pack9(R.y)); wr(out + 8, pack9(R.zz)); wr(out + 12, pack9(R.zzz));
  return 0;
}

// Lazy-reduction primitives (4 inputs of 4 words each; unused ones may be null):
// op 0 mul2(a, b, c, d)  1 sqr_plus(a, b)  2 sub_nr(a, b)  3 neg_nr(a)
//    4 sub_4p_a_2b(a, b)  5 sub_if_ge(a, 4p)  6 is_zero_lt4p(a)  7 reduce_est(a)
//    8 sub_4p_a_2b_cf(a, b)  9 sub_cf(a, b)  10 neg2p_cf(a)  11 neg_cf(a)  12 normalize(a)
// Inputs of ops 7 and 12 are 9 x u32 limbs (value below 2^261).  Outputs are 9 x u32 limbs per value (results may exceed 2^256).
int zkhost_field29_lazy(int field, int op, const uint64_t* a, const uint64_t* b, const uint64_t* c,
                        const uint64_t* d, uint32_t* out, uint64_t n) {
  for (uint64_t i = 0; i < n; i++) {
    f9 x;
    if (op == 7 || op == 12) {
      const uint32_t* l = reinterpret_cast<const uint32_t*>(a) + 9 * i;
      for (int j = 0; j < 9; j++) x.l[j] = l[j];
    } else {
      x = unpack9(rd(a + 4 * i));
    }
    const f9 y = b ? unpack9(rd(b + 4 * i)) : Fr9::zero();
    const f9 z = c ? unpack9(rd(c + 4 * i)) : Fr9::zero();
    const f9 w = d ? unpack9(rd(d + 4 * i)) : Fr9::zero();
    f9 r;
    if (field == 0) {
      switch (op) {
        case 0: r = Fr9::mul2(x, y, z, w); break;
        case 1: r = Fr9::sqr_plus(x, y); break;
        case 2: r = Fr9::sub_nr(x, y); break;
        case 3: r = Fr9::neg_nr(x); break;
        case 4: r = Fr9::sub_4p_a_2b(x, y); break;
        case 5: r = Fr9::sub_if_ge(x, Fr29Params::P4); break;
        case 6: r = Fr9::zero(); r.l[0] = Fr9::is_zero_lt4p(x); break;
        case 7: r = Fr9::reduce_est(x); break;
        case 8: r = Fr9::sub_4p_a_2b_cf(x, y); break;
        case 9: r = Fr9::sub_cf(x, y); break;
        case 10: r = Fr9::neg2p_cf(x); break;
        case 11: r = Fr9::neg_cf(x); break;
        case 12: r = Fr9::normalize(x); break;
        default: return 1;
      }
    } else {
      switch (op) {
        case 0: r = Fq9::mul2(x, y, z, w); break;
        case 1: r = Fq9::sqr_plus(x, y); break;
        case 2: r = Fq9::sub_nr(x, y); break;
        case 3: r = Fq9::neg_nr(x); break;
        case 4: r = Fq9::sub_4p_a_2b(x, y); break;
        case 5: r = Fq9::sub_if_ge(x, Fq29Params::P4); break;
        case 6: r = Fq9::zero(); r.l[0] = Fq9::is_zero_lt4p(x); break;
        case 7: r = Fq9::reduce_est(x); break;
        case 8: r = Fq9::sub_4p_a_2b_cf(x, y); break;
        case 9: r = Fq9::sub_cf(x, y); break;
        case 10: r = Fq9::neg2p_cf(x); break;
        case 11: r = Fq9::neg_cf(x); break;
        case 12: r = Fq9::normalize(x); break;
        default: return 1;
      }
    }
    for (int j = 0; j < 9; j++) out[9 * i + j] = r.l[j];
  }
  return 0;
}

// The X3 / Y3 step shared by the mixed and full additions: (R, PPP, Q, S1)
// as R = 2^261 words -> X3, Y3 as 9 x u32 limbs each (18 per value).
int zkhost_xyzz9_finish(const uint64_t* R, const uint64_t* ppp, const uint64_t* q, const uint64_t* s1,
                        uint32_t* out, uint64_t n) {
  for (uint64_t i = 0; i < n; i++) {
    f9 x3, y3;
    xyzz9_finish(unpack9(rd(R + 4 * i)), unpack9(rd(ppp + 4 * i)), unpack9(rd(q + 4 * i)),
                 unpack9(rd(s1 + 4 * i)), x3, y3);
    for (int j = 0; j < 9; j++) {
      out[18 * i + j] = x3.l[j];
      out[18 * i + 9 + j] = y3.l[j];
    }
  }
  return 0;
}

// The bucket-accumulation step of the MSM: p += each affine base in turn
// (R = 2^261 values, 8 words per base), the base negated as 2p - y where
// neg[i] is set; 16 words of XYZZ out after every step.
int zkhost_madd_chain(const uint64_t* p16, const uint64_t* affs, const uint8_t* neg, uint32_t k, uint64_t* out) {
  xyzz9 P;
  P.x = unpack9(rd(p16)); P.y = unpack9(rd(p16 + 4)); P.zz = unpack9(rd(p16 + 8)); P.zzz = unpack9(rd(p16 + 12));
  bool inf = xyzz9_is_inf(P);  // tracked as the accumulation kernel tracks it
  for (uint32_t i = 0; i < k; i++) {
    aff9 a;
    a.x = unpack9(rd(affs + 8 * i));
    a.y = unpack9(rd(affs + 8 * i + 4));
    if (!aff9_is_inf(a)) xyzz9_madd_t(P, inf, a, neg[i] != 0);
    if (inf != xyzz9_is_inf(P)) return 2;
    uint64_t* o = out + 16 * i;
    wr(o, pack9(P.x)); wr(o + 4, pack9(P.y)); wr(o + 8, pack9(P.zz)); wr(o + 12, pack9(P.zzz));
  }
  return 0;
}

// The arithmetic of one NTT pass holding the whole transform (n = 2^log_n
// <= 4096 elements, global stage 0 first), stage by stage (the kernel runs
// the same butterflies two stages at a time):
// caller words reduced below 4p on load, stage 0 with the carrying forms,
// later stages with the carry-free forms and u normalised at every even
// stage, then the quotient-estimate reduction (which takes the wide limbs)
// and canonical form.
// tw: omega^e (e < n/2), 4 words each, R = 2^261 canonical.  max_limb gets
// the largest limb any stage wrote.
int zkhost_ntt_tile(const uint64_t* in, const uint64_t* tw, uint32_t log_n, uint64_t* out, uint32_t* max_limb) {
  if (log_n == 0 || log_n > 12) return 1;
  const uint32_t n = 1u << log_n;
  std::vector<f9> a(n), b(n);
  for (uint32_t j = 0; j < n; j++) a[j] = Fr9::reduce2p(unpack9(rd(in + 4 * j)));
  uint32_t mx = 0;
  for (uint32_t t = 0; t < log_n; t++) {
    for (uint32_t q = 0; q < n / 2; q++) {
      const uint32_t d = q & ((1u << t) - 1);
      const f9 u = a[q], v = a[q + n / 2];
      const uint32_t o = ((q - d) << 1) + d;
      f9 lo, hi;
      if (t != 0) {
        const f9 vv = Fr9::mul(v, unpack9(rd(tw + 4 * ((uint64_t)d << (log_n - 1 - t)))));
        const f9 uu = ((t & 1) == 0) ? Fr9::normalize(u) : u;
        lo = Fr9::add_cf(uu, vv);
        hi = Fr9::sub_cf(uu, vv);
      } else {
        const f9 vv = Fr9::reduce2p(v);
        lo = Fr9::add_nr(u, vv);
        hi = Fr9::sub_nr(u, vv);
      }
      for (int i = 0; i < 9; i++) {
        mx = lo.l[i] > mx ? lo.l[i] : mx;
        mx = hi.l[i] > mx ? hi.l[i] : mx;
      }
      b[o] = lo;
      b[o + (1u << t)] = hi;
    }
    for (uint32_t j = 0; j < n; j++) a[j] = b[j];
  }
  for (uint32_t j = 0; j < n; j++) wr(out + 4 * j, pack9(Fr9::canonical(Fr9::reduce_est(a[j]))));
  *max_limb = mx;
  return 0;
}

// wire affine (R = 2^256) -> limb form -> XYZZ (R = 2^256, canonical) round trip
int zkhost_wire_roundtrip(const uint64_t* xy, uint64_t* out16) {
  aff9 a = aff9_from_wire(rd(xy), rd(xy + 4));
  xyzz r = xyzz_from9(xyzz9_from_aff(a));
  wr(out16, r.x); wr(out16 + 4, r.y); wr(out16 + 8, r.zz); wr(out16 + 12, r.zzz);
  return 0;
}
}
