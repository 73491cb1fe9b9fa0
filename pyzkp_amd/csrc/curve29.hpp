// BN254 G1 XYZZ arithmetic on the 29-bit-limb field (field29.hpp), used by
// the MSM accumulation, fix-up and reduction kernels.  Same formulas and
// exceptional-case handling as curve.hpp (add-2008-s, madd-2008-s,
// dbl-2008-s, mdbl-2008-s for a = 0); coordinates are R = 2^261 Montgomery
// values in [0, 2p); the identity is ZZ == 0 (mod p); an affine base with
// x = y = 0 is the point at infinity.
#pragma once
#include "curve.hpp"
#include "field29.hpp"

namespace zk {

struct aff9 {
  f9 x, y;
};

struct xyzz9 {
  f9 x, y, zz, zzz;
};

ZK_HD bool aff9_is_inf(const aff9& p) {
  u32 z = 0;
#pragma unroll
  for (int i = 0; i < 9; i++) z |= p.x.l[i] | p.y.l[i];
  return z == 0;
}

ZK_HD bool xyzz9_is_inf(const xyzz9& p) { return Fq9::is_zero(p.zz); }

ZK_HD xyzz9 xyzz9_inf() {
  xyzz9 r;
  r.x = Fq9::one();
  r.y = Fq9::one();
  r.zz = Fq9::zero();
  r.zzz = Fq9::zero();
  return r;
}

ZK_HD xyzz9 xyzz9_from_aff(const aff9& a) {
  if (aff9_is_inf(a)) return xyzz9_inf();
  xyzz9 r;
  r.x = a.x;
  r.y = a.y;
  r.zz = Fq9::one();
  r.zzz = Fq9::one();
  return r;
}

// mdbl-2008-s
ZK_HD xyzz9 xyzz9_dbl_aff(const aff9& a) {
  f9 u = Fq9::dbl(a.y);
  f9 v = Fq9::sqr(u);
  f9 w = Fq9::mul(u, v);
  f9 s = Fq9::mul(a.x, v);
  f9 x2 = Fq9::sqr(a.x);
  f9 m = Fq9::add(Fq9::dbl(x2), x2);
  xyzz9 r;
  r.x = Fq9::sub(Fq9::sqr(m), Fq9::dbl(s));
  r.y = Fq9::sub(Fq9::mul(m, Fq9::sub(s, r.x)), Fq9::mul(w, a.y));
  r.zz = v;
  r.zzz = w;
  return r;
}

// dbl-2008-s with lazy reductions.  Bounds, in units of p, for coordinates
// below 2p (normalised limbs):
//   U = 2Y < 4, limb by limb (limbs below 2^30: squaring columns stay
//     below 2^63.5, product columns below 2^63)
//   V = U^2 < 1.1, W = U V < 1.03, S = X V < 1.02, X^2 < 1.03
//   M = 3 X^2 < 3.1, normalised (it is squared)
//   X3 = M^2 + (4p - 2S) < 5.1 (bracket carry-free, in the high columns),
//        then below 1.0001p by one quotient estimate
//   Y3 = M (S - X3 + 2p) + W (4p - Y), one reduction: < 1.09 (brackets
//        carry-free, limbs below 1.5 * 2^30 and 2^30: columns under 27 * 2^59)
//   ZZ3 = V ZZ1, ZZZ3 = W ZZZ1 < 1.02
ZK_HD xyzz9 xyzz9_dbl(const xyzz9& p) {
  if (xyzz9_is_inf(p)) return p;
  const f9 u = Fq9::add_cf(p.y, p.y);
  f9 v, x2;
  Fq9::sqr_x2(u, p.x, v, x2);
  f9 w, s;
  Fq9::mul_x2(u, v, p.x, v, w, s);
  const f9 m = Fq9::normalize(Fq9::add_cf(Fq9::add_cf(x2, x2), x2));
  xyzz9 r;
  r.x = Fq9::reduce_est(Fq9::sqr_plus(m, Fq9::sub_4p_a_2b_cf(Fq9::zero(), s)));
  r.y = Fq9::mul2(m, Fq9::sub_cf(s, r.x), w, Fq9::neg_cf(p.y));
  Fq9::mul_x2(v, p.zz, w, p.zzz, r.zz, r.zzz);
  return r;
}

// Additions with lazy reductions.  Bounds, in units of p, for inputs below
// 2p (p / 2^261 < 0.0060, and a product is below p + a*b / 2^261):
//   U, S products < 1.03          P = U2 - U1 + 2p, R = S2 - S1 + 2p in (0, 3.03)
//   PP < 1.06, PPP and Q < 1.02   (P == 0 mod p iff P is p, 2p or 3p)
//   X3 = R^2 + (4p - PPP - 2Q) < 5.06, the bracket in carry-free limbs
//        (below 2^32, added into the squaring's 64-bit high columns), then
//        below 1.0001p by one quotient-estimate reduction
//   Y3 = R (Q - X3 + 2p) + S1 (2p - PPP), one reduction for both products,
//        < 1.08; both brackets in carry-free limbs below 1.5 * 2^30 and
//        2^30, so a column sums at most 9 products below 1.5 * 2^59, 9
//        below 2^59, 9 below 2^58 (m by p) and a carry: under 27 * 2^59
//   ZZ3, ZZZ3 < 1.02
// so every output coordinate is again below 2p, with normalised limbs.
ZK_HD void xyzz9_finish(const f9& R, const f9& ppp, const f9& q, const f9 s1, f9& x3, f9& y3) {
  x3 = Fq9::reduce_est(Fq9::sqr_plus(R, Fq9::sub_4p_a_2b_cf(ppp, q)));
  y3 = Fq9::mul2(R, Fq9::sub_cf(q, x3), s1, Fq9::neg2p_cf(ppp));
}

// madd-2008-s: p += a, or p -= a when neg is set, with the caller keeping
// p_inf == (p is the identity) up to date, so the bucket accumulation does
// not test ZZ at every step (a's coordinates below 2p).  For the negated
// base, S2 = (4p - y) ZZZ1 takes 4p - y in the carry-free form (limbs below
// 2^30, a valid product input; S2 stays below 1.03p); the identity and
// doubling cases, which use the coordinate further, take the normalised
// 2p - y.
ZK_HD void xyzz9_madd_t(xyzz9& p, bool& p_inf, const aff9& a, bool neg) {
  if (p_inf) {
    p_inf = aff9_is_inf(a);
    p = xyzz9_from_aff(a);
    if (neg && !p_inf) p.y = Fq9::neg_nr(a.y);
    return;
  }
  const f9 u2 = Fq9::mul(a.x, p.zz);
  const f9 s2 = Fq9::mul(neg ? Fq9::neg_cf(a.y) : a.y, p.zzz);
  const f9 P = Fq9::sub_nr(u2, p.x);
  const f9 R = Fq9::sub_nr(s2, p.y);
  // P is 0 mod p only as 0, p, 2p or 3p (normalised limbs below 4p): their
  // top limbs are the only candidates, so the full test runs only when a
  // lane's top limb matches one of them (almost never for distinct points)
  if (Fq9::top_may_be_zero_lt4p(P) && Fq9::is_zero_lt4p(P)) {
    if (Fq9::is_zero_lt4p(R)) {
      aff9 b = a;
      if (neg) b.y = Fq9::neg_nr(a.y);
      p = xyzz9_dbl_aff(b);
    } else {
      p = xyzz9_inf();
      p_inf = true;
    }
    return;
  }
  const f9 pp = Fq9::sqr(P);
  const f9 ppp = Fq9::mul(P, pp);
  const f9 q = Fq9::mul(p.x, pp);
  xyzz9_finish(R, ppp, q, p.y, p.x, p.y);
  p.zz = Fq9::mul(p.zz, pp);
  p.zzz = Fq9::mul(p.zzz, ppp);
}

// madd-2008-s: p += a (a affine, not the identity)
ZK_HD void xyzz9_madd(xyzz9& p, const aff9& a) {
  bool inf = xyzz9_is_inf(p);
  xyzz9_madd_t(p, inf, a, false);
}

// add-2008-s
ZK_HD xyzz9 xyzz9_add(const xyzz9& p, const xyzz9& q) {
  if (xyzz9_is_inf(p)) return q;
  if (xyzz9_is_inf(q)) return p;
  // independent products in pairs (mul_x2): the reduction's kernels run
  // few waves, so the latency of one product chain would be exposed
  f9 u1, u2, s1, s2;
  Fq9::mul_x2(p.x, q.zz, q.x, p.zz, u1, u2);
  Fq9::mul_x2(p.y, q.zzz, q.y, p.zzz, s1, s2);
  const f9 P = Fq9::sub_nr(u2, u1);
  const f9 R = Fq9::sub_nr(s2, s1);
  if (Fq9::top_may_be_zero_lt4p(P) && Fq9::is_zero_lt4p(P)) {
    if (Fq9::is_zero_lt4p(R)) return xyzz9_dbl(p);
    return xyzz9_inf();
  }
  f9 zz12, zzz12;
  Fq9::mul_x2(p.zz, q.zz, p.zzz, q.zzz, zz12, zzz12);
  const f9 pp = Fq9::sqr(P);
  f9 ppp, qq;
  Fq9::mul_x2(P, pp, u1, pp, ppp, qq);
  xyzz9 r;
  xyzz9_finish(R, ppp, qq, s1, r.x, r.y);
  Fq9::mul_x2(zz12, pp, zzz12, ppp, r.zz, r.zzz);
  return r;
}

ZK_HD xyzz9 xyzz9_mul_small(const xyzz9& p, uint32_t k) {
  xyzz9 acc = xyzz9_inf();
  if (k == 0 || xyzz9_is_inf(p)) return acc;
  int top = 31;
  while (top > 0 && !((k >> top) & 1)) top--;
  acc = p;
  for (int b = top - 1; b >= 0; b--) {
    acc = xyzz9_dbl(acc);
    if ((k >> b) & 1) acc = xyzz9_add(acc, p);
  }
  return acc;
}

// R = 2^256 32-bit-word point (curve.hpp) <-> R = 2^261 limb point
ZK_HD aff9 aff9_from_wire(const fe& x256, const fe& y256) {
  aff9 r;
  f9 x = unpack9(x256), y = unpack9(y256);
  r.x = Fq9::to261(x);  // 0 stays 0: the infinity marker survives
  r.y = Fq9::to261(y);
  return r;
}

ZK_HD xyzz xyzz_from9(const xyzz9& p) {
  xyzz r;
  r.x = pack9(Fq9::canonical(Fq9::to256(p.x)));
  r.y = pack9(Fq9::canonical(Fq9::to256(p.y)));
  r.zz = pack9(Fq9::canonical(Fq9::to256(p.zz)));
  r.zzz = pack9(Fq9::canonical(Fq9::to256(p.zzz)));
  return r;
}

}  // namespace zk
