// BN254 G1 (y^2 = x^3 + 3 over Fq) point arithmetic for the MSM kernels.
//
// Buckets are kept in XYZZ coordinates (x = X/ZZ, y = Y/ZZZ, ZZ^3 = ZZZ^2),
// which give the cheapest mixed addition with an affine base (8M + 2S) and
// need no inversion.  Formulas: add-2008-s, madd-2008-s, dbl-2008-s and
// mdbl-2008-s from the Explicit-Formulas Database for a = 0.  Every
// exceptional case is handled exactly (P = Q, P = -Q, either operand the
// identity), so repeated bases inside one bucket give the right result.
//
// The reference keeps points in Jacobian coordinates with Z = 0 for the
// identity (runtime/metal/ecc_source.py:142-305); here the identity is
// ZZ = 0, and an affine base with x = y = 0 (not on the curve) stands for
// the point at infinity in the packed base array.
#pragma once
#include "field.hpp"

namespace zk {

struct aff {
  fe x, y;
};

struct xyzz {
  fe x, y, zz, zzz;
};

ZK_HD bool aff_is_inf(const aff& p) { return Fq::is_zero(p.x) && Fq::is_zero(p.y); }

ZK_HD bool xyzz_is_inf(const xyzz& p) { return Fq::is_zero(p.zz); }

ZK_HD xyzz xyzz_inf() {
  xyzz r;
  r.x = Fq::one();
  r.y = Fq::one();
  r.zz = Fq::zero();
  r.zzz = Fq::zero();
  return r;
}

ZK_HD xyzz xyzz_from_aff(const aff& a) {
  if (aff_is_inf(a)) return xyzz_inf();
  xyzz r;
  r.x = a.x;
  r.y = a.y;
  r.zz = Fq::one();
  r.zzz = Fq::one();
  return r;
}

// mdbl-2008-s: 2 * (x, y) for an affine point
ZK_HD xyzz xyzz_dbl_aff(const aff& a) {
  fe u = Fq::dbl(a.y);
  fe v = Fq::sqr(u);
  fe w = Fq::mul(u, v);
  fe s = Fq::mul(a.x, v);
  fe x2 = Fq::sqr(a.x);
  fe m = Fq::add(Fq::dbl(x2), x2);
  xyzz r;
  r.x = Fq::sub(Fq::sqr(m), Fq::dbl(s));
  r.y = Fq::sub(Fq::mul(m, Fq::sub(s, r.x)), Fq::mul(w, a.y));
  r.zz = v;
  r.zzz = w;
  return r;
}

// dbl-2008-s
ZK_HD xyzz xyzz_dbl(const xyzz& p) {
  if (xyzz_is_inf(p)) return p;
  fe u = Fq::dbl(p.y);
  fe v = Fq::sqr(u);
  fe w = Fq::mul(u, v);
  fe s = Fq::mul(p.x, v);
  fe x2 = Fq::sqr(p.x);
  fe m = Fq::add(Fq::dbl(x2), x2);
  xyzz r;
  r.x = Fq::sub(Fq::sqr(m), Fq::dbl(s));
  r.y = Fq::sub(Fq::mul(m, Fq::sub(s, r.x)), Fq::mul(w, p.y));
  r.zz = Fq::mul(v, p.zz);
  r.zzz = Fq::mul(w, p.zzz);
  return r;
}

// madd-2008-s: p += a (a affine, not the identity)
ZK_HD void xyzz_madd(xyzz& p, const aff& a) {
  if (xyzz_is_inf(p)) {
    p = xyzz_from_aff(a);
    return;
  }
  fe u2 = Fq::mul(a.x, p.zz);
  fe s2 = Fq::mul(a.y, p.zzz);
  fe P = Fq::sub(u2, p.x);
  fe R = Fq::sub(s2, p.y);
  if (Fq::is_zero(P)) {
    if (Fq::is_zero(R)) {
      p = xyzz_dbl_aff(a);
    } else {
      p = xyzz_inf();
    }
    return;
  }
  fe pp = Fq::sqr(P);
  fe ppp = Fq::mul(P, pp);
  fe q = Fq::mul(p.x, pp);
  fe x3 = Fq::sub(Fq::sub(Fq::sqr(R), ppp), Fq::dbl(q));
  fe y3 = Fq::sub(Fq::mul(R, Fq::sub(q, x3)), Fq::mul(p.y, ppp));
  p.x = x3;
  p.y = y3;
  p.zz = Fq::mul(p.zz, pp);
  p.zzz = Fq::mul(p.zzz, ppp);
}

// add-2008-s: p + q
ZK_HD xyzz xyzz_add(const xyzz& p, const xyzz& q) {
  if (xyzz_is_inf(p)) return q;
  if (xyzz_is_inf(q)) return p;
  fe u1 = Fq::mul(p.x, q.zz);
  fe u2 = Fq::mul(q.x, p.zz);
  fe s1 = Fq::mul(p.y, q.zzz);
  fe s2 = Fq::mul(q.y, p.zzz);
  fe P = Fq::sub(u2, u1);
  fe R = Fq::sub(s2, s1);
  if (Fq::is_zero(P)) {
    if (Fq::is_zero(R)) return xyzz_dbl(p);
    return xyzz_inf();
  }
  fe pp = Fq::sqr(P);
  fe ppp = Fq::mul(P, pp);
  fe qq = Fq::mul(u1, pp);
  xyzz r;
  r.x = Fq::sub(Fq::sub(Fq::sqr(R), ppp), Fq::dbl(qq));
  r.y = Fq::sub(Fq::mul(R, Fq::sub(qq, r.x)), Fq::mul(s1, ppp));
  r.zz = Fq::mul(Fq::mul(p.zz, q.zz), pp);
  r.zzz = Fq::mul(Fq::mul(p.zzz, q.zzz), ppp);
  return r;
}

ZK_HD xyzz xyzz_neg(const xyzz& p) {
  xyzz r = p;
  r.y = Fq::neg(p.y);
  return r;
}

// k * p for a small non-negative integer k (double-and-add, MSB first)
ZK_HD xyzz xyzz_mul_small(const xyzz& p, uint32_t k) {
  xyzz acc = xyzz_inf();
  if (k == 0 || xyzz_is_inf(p)) return acc;
  int top = 31;
  while (top > 0 && !((k >> top) & 1)) top--;
  acc = p;
  for (int b = top - 1; b >= 0; b--) {
    acc = xyzz_dbl(acc);
    if ((k >> b) & 1) acc = xyzz_add(acc, p);
  }
  return acc;
}

}  // namespace zk
