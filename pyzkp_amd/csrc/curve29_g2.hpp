// BN254 G2 XYZZ arithmetic on the 29-bit-limb field, for the device G2 MSM.
// G2 lives on the twist y^2 = x^3 + 3 / (9 + u) over Fq2 = Fq[u] / (u^2 + 1);
// the XYZZ formulas (add-2008-s, madd-2008-s, dbl-2008-s, mdbl-2008-s for
// a = 0) do not involve the curve constant, so they are the G1 formulas over
// Fq2.  Every Fq2 coefficient is an R = 2^261 Montgomery value in [0, 2p)
// with normalised limbs (the fully reducing Fq9 forms throughout: add, sub,
// mul); the identity is ZZ == 0; an affine base with x = y = 0 is the point
// at infinity.
#pragma once
#include "curve29.hpp"

namespace zk {

struct f92 {
  f9 c0, c1;
};

ZK_HD f92 f2_zero9() { return {Fq9::zero(), Fq9::zero()}; }
ZK_HD f92 f2_one9() { return {Fq9::one(), Fq9::zero()}; }
ZK_HD f92 f2_add9(const f92& a, const f92& b) { return {Fq9::add(a.c0, b.c0), Fq9::add(a.c1, b.c1)}; }
ZK_HD f92 f2_sub9(const f92& a, const f92& b) { return {Fq9::sub(a.c0, b.c0), Fq9::sub(a.c1, b.c1)}; }
ZK_HD f92 f2_dbl9(const f92& a) { return f2_add9(a, a); }
ZK_HD f92 f2_neg9(const f92& a) { return f2_sub9(f2_zero9(), a); }
ZK_HD bool f2_is_zero9(const f92& a) { return Fq9::is_zero(a.c0) && Fq9::is_zero(a.c1); }
// (a0 + a1 u)(b0 + b1 u) = a0 b0 - a1 b1 + ((a0 + a1)(b0 + b1) - a0 b0 - a1 b1) u;
// the sums are below 4p (normalised), and a product of two such is below
// 16 p^2 < 2^261 p, so the third product also ends below 2p
ZK_HD f92 f2_mul9(const f92& a, const f92& b) {
  f9 t0, t1;
  Fq9::mul_x2(a.c0, b.c0, a.c1, b.c1, t0, t1);
  const f9 t2 = Fq9::mul(Fq9::add_nr(a.c0, a.c1), Fq9::add_nr(b.c0, b.c1));
  return {Fq9::sub(t0, t1), Fq9::sub(Fq9::sub(t2, t0), t1)};
}
// (a0 + a1 u)^2 = (a0 + a1)(a0 - a1) + 2 a0 a1 u (a0 - a1 + 2p below 4p)
ZK_HD f92 f2_sqr9(const f92& a) {
  f9 s, m;
  Fq9::mul_x2(Fq9::add_nr(a.c0, a.c1), Fq9::sub_nr(a.c0, a.c1), a.c0, a.c1, s, m);
  return {s, Fq9::dbl(m)};
}

struct aff9g2 {
  f92 x, y;
};

struct xyzz9g2 {
  f92 x, y, zz, zzz;
};

ZK_HD bool aff9g2_is_inf(const aff9g2& p) {
  u32 z = 0;
#pragma unroll
  for (int i = 0; i < 9; i++) z |= p.x.c0.l[i] | p.x.c1.l[i] | p.y.c0.l[i] | p.y.c1.l[i];
  return z == 0;
}

ZK_HD bool xyzz9g2_is_inf(const xyzz9g2& p) { return f2_is_zero9(p.zz); }

ZK_HD xyzz9g2 xyzz9g2_inf() {
  xyzz9g2 r;
  r.x = f2_one9();
  r.y = f2_one9();
  r.zz = f2_zero9();
  r.zzz = f2_zero9();
  return r;
}

// mdbl-2008-s (a affine, not the identity)
ZK_HD xyzz9g2 xyzz9g2_dbl_aff(const aff9g2& a) {
  const f92 u = f2_dbl9(a.y);
  const f92 v = f2_sqr9(u);
  const f92 w = f2_mul9(u, v);
  const f92 s = f2_mul9(a.x, v);
  const f92 x2 = f2_sqr9(a.x);
  const f92 m = f2_add9(f2_dbl9(x2), x2);
  xyzz9g2 r;
  r.x = f2_sub9(f2_sqr9(m), f2_dbl9(s));
  r.y = f2_sub9(f2_mul9(m, f2_sub9(s, r.x)), f2_mul9(w, a.y));
  r.zz = v;
  r.zzz = w;
  return r;
}

// dbl-2008-s
ZK_HD xyzz9g2 xyzz9g2_dbl(const xyzz9g2& p) {
  if (xyzz9g2_is_inf(p)) return p;
  const f92 u = f2_dbl9(p.y);
  const f92 v = f2_sqr9(u);
  const f92 w = f2_mul9(u, v);
  const f92 s = f2_mul9(p.x, v);
  const f92 x2 = f2_sqr9(p.x);
  const f92 m = f2_add9(f2_dbl9(x2), x2);
  xyzz9g2 r;
  r.x = f2_sub9(f2_sqr9(m), f2_dbl9(s));
  r.y = f2_sub9(f2_mul9(m, f2_sub9(s, r.x)), f2_mul9(w, p.y));
  r.zz = f2_mul9(v, p.zz);
  r.zzz = f2_mul9(w, p.zzz);
  return r;
}

// madd-2008-s: p += a, or p -= a when neg is set, with the caller keeping
// p_inf == (p is the identity), as xyzz9_madd_t
ZK_HD void xyzz9g2_madd_t(xyzz9g2& p, bool& p_inf, const aff9g2& a, bool neg) {
  aff9g2 b = a;
  if (neg) b.y = f2_neg9(a.y);
  if (p_inf) {
    p_inf = aff9g2_is_inf(a);
    if (p_inf) return;
    p.x = b.x;
    p.y = b.y;
    p.zz = f2_one9();
    p.zzz = f2_one9();
    return;
  }
  const f92 u2 = f2_mul9(b.x, p.zz);
  const f92 s2 = f2_mul9(b.y, p.zzz);
  const f92 P = f2_sub9(u2, p.x);
  const f92 R = f2_sub9(s2, p.y);
  if (f2_is_zero9(P)) {
    if (f2_is_zero9(R)) {
      p = xyzz9g2_dbl_aff(b);
    } else {
      p = xyzz9g2_inf();
      p_inf = true;
    }
    return;
  }
  const f92 pp = f2_sqr9(P);
  const f92 ppp = f2_mul9(P, pp);
  const f92 q = f2_mul9(p.x, pp);
  const f92 x3 = f2_sub9(f2_sub9(f2_sqr9(R), ppp), f2_dbl9(q));
  p.y = f2_sub9(f2_mul9(R, f2_sub9(q, x3)), f2_mul9(p.y, ppp));
  p.x = x3;
  p.zz = f2_mul9(p.zz, pp);
  p.zzz = f2_mul9(p.zzz, ppp);
}

// add-2008-s
ZK_HD xyzz9g2 xyzz9g2_add(const xyzz9g2& p, const xyzz9g2& q) {
  if (xyzz9g2_is_inf(p)) return q;
  if (xyzz9g2_is_inf(q)) return p;
  const f92 u1 = f2_mul9(p.x, q.zz), u2 = f2_mul9(q.x, p.zz);
  const f92 s1 = f2_mul9(p.y, q.zzz), s2 = f2_mul9(q.y, p.zzz);
  const f92 P = f2_sub9(u2, u1);
  const f92 R = f2_sub9(s2, s1);
  if (f2_is_zero9(P)) {
    if (f2_is_zero9(R)) return xyzz9g2_dbl(p);
    return xyzz9g2_inf();
  }
  const f92 pp = f2_sqr9(P);
  const f92 ppp = f2_mul9(P, pp);
  const f92 qq = f2_mul9(u1, pp);
  xyzz9g2 r;
  r.x = f2_sub9(f2_sub9(f2_sqr9(R), ppp), f2_dbl9(qq));
  r.y = f2_sub9(f2_mul9(R, f2_sub9(qq, r.x)), f2_mul9(s1, ppp));
  r.zz = f2_mul9(f2_mul9(p.zz, q.zz), pp);
  r.zzz = f2_mul9(f2_mul9(p.zzz, q.zzz), ppp);
  return r;
}

}  // namespace zk
