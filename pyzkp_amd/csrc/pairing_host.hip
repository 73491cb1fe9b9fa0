// Host-side BN254 optimal-ate pairing, for verifying Groth16 proofs
// (the reference verifies with py_ecc's pairing: crypto/pairing/bn254.py:10-11,
// protocols/groth16/verify.py:11-23).
//
// Tower: Fq2 = Fq[i]/(i^2 + 1), Fq6 = Fq2[v]/(v^3 - xi) with xi = 9 + i,
// Fq12 = Fq6[w]/(w^2 - v).  G2 lives on the D-type twist
// E'(Fq2): y^2 = x^3 + 3/xi, untwisted by psi(x, y) = (x w^2, y w^3).
// Miller loop over 6u + 2 (u = 4965661367192848881) with affine T on the
// twist; the line through T and Q (slope lambda) evaluated at P = (xP, yP)
// is yP - lambda xP w + (lambda xT - yT) w^3, and two final lines add
// pi(Q) and -pi^2(Q).  Final exponentiation (p^12 - 1)/r: easy part
// f^(p^6 - 1)(p^2 + 1), hard part by plain square-and-multiply with
// (p^4 - p^2 + 1)/r.
#include <string.h>
#include <vector>
#include "internal.hpp"

namespace zk {
namespace pr {

using zkh::Fe;
using zkh::FQ;

struct F2 {
  Fe a, b;
};
static inline F2 add2(const F2& x, const F2& y) { return {zkh::add(FQ, x.a, y.a), zkh::add(FQ, x.b, y.b)}; }
static inline F2 sub2(const F2& x, const F2& y) { return {zkh::sub(FQ, x.a, y.a), zkh::sub(FQ, x.b, y.b)}; }
static inline F2 neg2(const F2& x) { return {zkh::neg(FQ, x.a), zkh::neg(FQ, x.b)}; }
static inline F2 mul2(const F2& x, const F2& y) {
  Fe t0 = zkh::mul(FQ, x.a, y.a), t1 = zkh::mul(FQ, x.b, y.b);
  Fe t2 = zkh::mul(FQ, zkh::add(FQ, x.a, x.b), zkh::add(FQ, y.a, y.b));
  return {zkh::sub(FQ, t0, t1), zkh::sub(FQ, zkh::sub(FQ, t2, t0), t1)};
}
static inline F2 muls2(const F2& x, const Fe& s) { return {zkh::mul(FQ, x.a, s), zkh::mul(FQ, x.b, s)}; }
static inline F2 conj2(const F2& x) { return {x.a, zkh::neg(FQ, x.b)}; }
static inline bool zero2(const F2& x) { return zkh::is_zero(x.a) && zkh::is_zero(x.b); }
static inline F2 inv2(const F2& x) {
  Fe n = zkh::add(FQ, zkh::mul(FQ, x.a, x.a), zkh::mul(FQ, x.b, x.b));
  Fe ni = zkh::inv(FQ, n);
  return {zkh::mul(FQ, x.a, ni), zkh::neg(FQ, zkh::mul(FQ, x.b, ni))};
}
static F2 xi() { return {zkh::from_u64(FQ, 9), zkh::one(FQ)}; }
static inline F2 mulxi(const F2& x) {  // (9 + i)(a + b i) = (9a - b) + (a + 9b) i
  Fe nine = zkh::from_u64(FQ, 9);
  return {zkh::sub(FQ, zkh::mul(FQ, x.a, nine), x.b), zkh::add(FQ, x.a, zkh::mul(FQ, x.b, nine))};
}
static F2 zero2v() { return {zkh::zero(), zkh::zero()}; }
static F2 one2() { return {zkh::one(FQ), zkh::zero()}; }

struct F6 {
  F2 c0, c1, c2;
};
static inline F6 add6(const F6& x, const F6& y) { return {add2(x.c0, y.c0), add2(x.c1, y.c1), add2(x.c2, y.c2)}; }
static inline F6 sub6(const F6& x, const F6& y) { return {sub2(x.c0, y.c0), sub2(x.c1, y.c1), sub2(x.c2, y.c2)}; }
static inline F6 neg6(const F6& x) { return {neg2(x.c0), neg2(x.c1), neg2(x.c2)}; }
static F6 mul6(const F6& x, const F6& y) {
  // schoolbook with v^3 = xi
  F2 t00 = mul2(x.c0, y.c0), t11 = mul2(x.c1, y.c1), t22 = mul2(x.c2, y.c2);
  F2 t01 = add2(mul2(x.c0, y.c1), mul2(x.c1, y.c0));
  F2 t02 = add2(add2(mul2(x.c0, y.c2), mul2(x.c2, y.c0)), t11);
  F2 t12 = add2(mul2(x.c1, y.c2), mul2(x.c2, y.c1));
  return {add2(t00, mulxi(t12)), add2(t01, mulxi(t22)), t02};
}
static inline F6 mulv6(const F6& x) { return {mulxi(x.c2), x.c0, x.c1}; }  // x v
static F6 inv6(const F6& x) {
  F2 A = sub2(mul2(x.c0, x.c0), mulxi(mul2(x.c1, x.c2)));
  F2 B = sub2(mulxi(mul2(x.c2, x.c2)), mul2(x.c0, x.c1));
  F2 C = sub2(mul2(x.c1, x.c1), mul2(x.c0, x.c2));
  F2 F = add2(mul2(x.c0, A), mulxi(add2(mul2(x.c2, B), mul2(x.c1, C))));
  F2 Fi = inv2(F);
  return {mul2(A, Fi), mul2(B, Fi), mul2(C, Fi)};
}
static F6 zero6() { return {zero2v(), zero2v(), zero2v()}; }
static F6 one6() { return {one2(), zero2v(), zero2v()}; }

struct F12 {
  F6 c0, c1;
};
static F12 mul12(const F12& x, const F12& y) {
  F6 t0 = mul6(x.c0, y.c0), t1 = mul6(x.c1, y.c1);
  F6 m = mul6(add6(x.c0, x.c1), add6(y.c0, y.c1));
  return {add6(t0, mulv6(t1)), sub6(sub6(m, t0), t1)};
}
static F12 sqr12(const F12& x) { return mul12(x, x); }
static F12 conj12(const F12& x) { return {x.c0, neg6(x.c1)}; }
static F12 inv12(const F12& x) {
  F6 t = sub6(mul6(x.c0, x.c0), mulv6(mul6(x.c1, x.c1)));
  F6 ti = inv6(t);
  return {mul6(x.c0, ti), neg6(mul6(x.c1, ti))};
}
static F12 one12() { return {one6(), zero6()}; }
static bool is_one12(const F12& x) {
  F12 o = one12();
  return memcmp(&x, &o, sizeof(F12)) == 0;
}

// big exponent (little-endian 64-bit words), square-and-multiply
static F12 pow12(const F12& a, const std::vector<uint64_t>& e) {
  F12 r = one12();
  for (int i = (int)e.size() - 1; i >= 0; i--)
    for (int b = 63; b >= 0; b--) {
      r = sqr12(r);
      if ((e[i] >> b) & 1) r = mul12(r, a);
    }
  return r;
}
static F2 pow2(const F2& a, const std::vector<uint64_t>& e) {
  F2 r = one2();
  for (int i = (int)e.size() - 1; i >= 0; i--)
    for (int b = 63; b >= 0; b--) {
      r = mul2(r, r);
      if ((e[i] >> b) & 1) r = mul2(r, a);
    }
  return r;
}

// --- small multiprecision helpers for the exponents (little-endian u64)
typedef std::vector<uint64_t> Big;
static Big big_from(const uint64_t* w, int n) { return Big(w, w + n); }
static Big big_mul(const Big& a, const Big& b) {
  Big r(a.size() + b.size(), 0);
  for (size_t i = 0; i < a.size(); i++) {
    unsigned __int128 c = 0;
    for (size_t j = 0; j < b.size(); j++) {
      c += (unsigned __int128)a[i] * b[j] + r[i + j];
      r[i + j] = (uint64_t)c;
      c >>= 64;
    }
    r[i + b.size()] = (uint64_t)c;
  }
  while (r.size() > 1 && r.back() == 0) r.pop_back();
  return r;
}
static Big big_addsmall(Big a, int64_t s) {  // a + s (s may be negative, |s| small, result >= 0)
  if (s >= 0) {
    unsigned __int128 c = (uint64_t)s;
    for (size_t i = 0; i < a.size() && c; i++) {
      c += a[i];
      a[i] = (uint64_t)c;
      c >>= 64;
    }
    if (c) a.push_back((uint64_t)c);
  } else {
    uint64_t br = (uint64_t)(-s);
    for (size_t i = 0; i < a.size() && br; i++) {
      uint64_t old = a[i];
      a[i] = old - br;
      br = old < br ? 1 : 0;
    }
  }
  while (a.size() > 1 && a.back() == 0) a.pop_back();
  return a;
}
static Big big_sub(const Big& a, const Big& b) {  // a >= b
  Big r(a);
  uint64_t br = 0;
  for (size_t i = 0; i < r.size(); i++) {
    unsigned __int128 d = (unsigned __int128)r[i] - (i < b.size() ? b[i] : 0) - br;
    r[i] = (uint64_t)d;
    br = (uint64_t)(d >> 127);
  }
  while (r.size() > 1 && r.back() == 0) r.pop_back();
  return r;
}
static Big big_add(const Big& a, const Big& b) {
  Big r(std::max(a.size(), b.size()) + 1, 0);
  unsigned __int128 c = 0;
  for (size_t i = 0; i < r.size(); i++) {
    c += (unsigned __int128)(i < a.size() ? a[i] : 0) + (i < b.size() ? b[i] : 0);
    r[i] = (uint64_t)c;
    c >>= 64;
  }
  while (r.size() > 1 && r.back() == 0) r.pop_back();
  return r;
}
// a / d for a small-word divisor d (d fits the bigint routine via long division by a Big)
static Big big_divexact(const Big& a, const Big& d) {
  // binary long division (a and d are < 2^1024; called once at init)
  const int nb = (int)a.size() * 64;
  Big q(a.size(), 0), r(1, 0);
  for (int bit = nb - 1; bit >= 0; bit--) {
    // r = 2r + bit
    Big r2(r.size() + 1, 0);
    uint64_t c = (a[bit / 64] >> (bit % 64)) & 1;
    for (size_t i = 0; i < r.size(); i++) {
      r2[i] = (r[i] << 1) | c;
      c = r[i] >> 63;
    }
    r2[r.size()] = c;
    while (r2.size() > 1 && r2.back() == 0) r2.pop_back();
    r = r2;
    // if r >= d: r -= d, q |= bit
    bool ge;
    if (r.size() != d.size()) ge = r.size() > d.size();
    else {
      ge = true;
      for (int i = (int)r.size() - 1; i >= 0; i--)
        if (r[i] != d[i]) {
          ge = r[i] > d[i];
          break;
        }
    }
    if (ge) {
      r = big_sub(r, d);
      q[bit / 64] |= 1ull << (bit % 64);
    }
  }
  while (q.size() > 1 && q.back() == 0) q.pop_back();
  return q;
}

struct Consts {
  bool ready = false;
  F2 g12, g13;    // xi^((p-1)/3), xi^((p-1)/2): Frobenius on the twist
  Big p2;         // p^2
  Big hard;       // (p^4 - p^2 + 1) / r
  F2 b2;          // 3 / xi
};
static Consts K;

static void init() {
  if (K.ready) return;
  Big p = big_from(FQ.p, 4);
  Big r = big_from(zkh::FR.p, 4);
  Big pm1 = big_addsmall(p, -1);
  Big e3 = big_divexact(pm1, Big{3});
  Big e2 = big_divexact(pm1, Big{2});
  K.g12 = pow2(xi(), e3);
  K.g13 = pow2(xi(), e2);
  K.p2 = big_mul(p, p);
  Big p4 = big_mul(K.p2, K.p2);
  Big num = big_addsmall(big_sub(p4, K.p2), 1);
  K.hard = big_divexact(num, r);
  K.b2 = mul2({zkh::from_u64(FQ, 3), zkh::zero()}, inv2(xi()));
  K.ready = true;
}

struct A2 {
  F2 x, y;
  bool inf;
};

static A2 frob(const A2& q) {  // pi(x, y) = (conj(x) g12, conj(y) g13)
  return {mul2(conj2(q.x), K.g12), mul2(conj2(q.y), K.g13), q.inf};
}

// line through T and the next point (slope lam), evaluated at P
static F12 line(const F2& lam, const A2& T, const Fe& xP, const Fe& yP) {
  F12 l;
  l.c0 = {{yP, zkh::zero()}, zero2v(), zero2v()};
  l.c1 = {neg2(muls2(lam, xP)), sub2(mul2(lam, T.x), T.y), zero2v()};
  return l;
}

static F12 miller(const Fe& xP, const Fe& yP, const A2& Q) {
  const uint64_t loop = 0x9d797039be763ba8ull;  // low 64 bits of 6u + 2 = 29793968203157093288
  // 6u + 2 = 0x19d797039be763ba8 (65 bits): top bit at position 64
  F12 f = one12();
  A2 T = Q;
  auto dbl_step = [&]() {
    F2 lam = mul2(muls2(mul2(T.x, T.x), zkh::from_u64(FQ, 3)), inv2(add2(T.y, T.y)));
    F12 l = line(lam, T, xP, yP);
    F2 x3 = sub2(sub2(mul2(lam, lam), T.x), T.x);
    F2 y3 = sub2(mul2(lam, sub2(T.x, x3)), T.y);
    T.x = x3;
    T.y = y3;
    return l;
  };
  auto add_step = [&](const A2& S) {
    if (zero2(sub2(S.x, T.x))) {
      // T = +-S: only reached for degenerate inputs; a vertical line is 1 after
      // the final exponentiation, a tangent is the doubling line
      if (zero2(add2(S.y, T.y))) {
        T.inf = true;
        return one12();
      }
      return dbl_step();
    }
    F2 lam = mul2(sub2(S.y, T.y), inv2(sub2(S.x, T.x)));
    F12 l = line(lam, T, xP, yP);
    F2 x3 = sub2(sub2(mul2(lam, lam), T.x), S.x);
    F2 y3 = sub2(mul2(lam, sub2(T.x, x3)), T.y);
    T.x = x3;
    T.y = y3;
    return l;
  };
  // bit 64 is the leading 1; process bits 63..0
  for (int b = 63; b >= 0; b--) {
    f = mul12(sqr12(f), dbl_step());
    if ((loop >> b) & 1) f = mul12(f, add_step(Q));
  }
  A2 q1 = frob(Q);
  A2 q2 = frob(q1);
  q2.y = neg2(q2.y);
  f = mul12(f, add_step(q1));
  f = mul12(f, add_step(q2));
  return f;
}

static F12 final_exp(const F12& f) {
  F12 t = mul12(conj12(f), inv12(f));        // f^(p^6 - 1)
  t = mul12(pow12(t, K.p2), t);              // ^(p^2 + 1)
  return pow12(t, K.hard);                   // ^((p^4 - p^2 + 1) / r)
}

static Fe rd_fq(const uint64_t* w) {
  Fe a;
  memcpy(a.v, w, 32);
  return zkh::to_mont(FQ, a);
}
static void wr_fq(uint64_t* w, const Fe& a) {
  Fe c = zkh::from_mont(FQ, a);
  memcpy(w, c.v, 32);
}

}  // namespace pr
}  // namespace zk

using namespace zk;

extern "C" {

/* prod_i e(P_i, Q_i) for n pairs: P_i affine G1 canonical (8 words, zeros =
 * infinity), Q_i affine G2 canonical (16 words).  Pairs with an identity
 * contribute 1.  out: the Fq12 result, 48 canonical words in the order
 * c0.c0.a, c0.c0.b, c0.c1.a, ..., c1.c2.b. */
int zk_pairing_product(const uint64_t* g1, const uint64_t* g2, uint32_t n, uint64_t* out48) {
  if ((n && (!g1 || !g2)) || !out48) return ZK_ERR_INVALID;
  pr::init();
  pr::F12 acc = pr::one12();
  for (uint32_t i = 0; i < n; i++) {
    const uint64_t* p = g1 + 8ull * i;
    const uint64_t* q = g2 + 16ull * i;
    bool pinf = true, qinf = true;
    for (int k = 0; k < 8; k++) pinf &= p[k] == 0;
    for (int k = 0; k < 16; k++) qinf &= q[k] == 0;
    if (pinf || qinf) continue;
    pr::A2 Q{{pr::rd_fq(q), pr::rd_fq(q + 4)}, {pr::rd_fq(q + 8), pr::rd_fq(q + 12)}, false};
    acc = pr::mul12(acc, pr::miller(pr::rd_fq(p), pr::rd_fq(p + 4), Q));
  }
  pr::F12 r = pr::final_exp(acc);
  const pr::F2* parts[6] = {&r.c0.c0, &r.c0.c1, &r.c0.c2, &r.c1.c0, &r.c1.c1, &r.c1.c2};
  for (int k = 0; k < 6; k++) {
    pr::wr_fq(out48 + 8 * k, parts[k]->a);
    pr::wr_fq(out48 + 8 * k + 4, parts[k]->b);
  }
  return ZK_OK;
}

/* 1 in *ok iff prod_i e(P_i, Q_i) == 1 (one final exponentiation). */
int zk_pairing_check(const uint64_t* g1, const uint64_t* g2, uint32_t n, int* ok) {
  if (!ok) return ZK_ERR_INVALID;
  uint64_t out[48];
  int rc = zk_pairing_product(g1, g2, n, out);
  if (rc) return rc;
  bool one = out[0] == 1;
  for (int k = 1; k < 48; k++) one &= out[k] == 0;
  *ok = one ? 1 : 0;
  return ZK_OK;
}

}  // extern "C"
