// Host-side BN254 G1 scalar multiplication for proof assembly (the few
// [r] delta, [s] A, [r] B1, [r s] delta products a Groth16 proof needs after
// its MSMs: protocols/groth16/prove.py:119-126).  Jacobian coordinates over
// the 4 x 64-bit host field; points cross the C-ABI as canonical affine
// little-endian words (x, y; 8 x u64, all zero = infinity).
#include <string.h>
#include "internal.hpp"
#include "curve.hpp"

namespace zk {
namespace g1h {

using zkh::Fe;
using zkh::FQ;

struct J1 {
  Fe X, Y, Z;
};

static inline bool inf(const J1& p) { return zkh::is_zero(p.Z); }
static inline Fe dbl(const Fe& a) { return zkh::add(FQ, a, a); }

// dbl-2009-l (a = 0)
static J1 jdbl(const J1& p) {
  if (inf(p) || zkh::is_zero(p.Y)) return {zkh::one(FQ), zkh::one(FQ), zkh::zero()};
  Fe A = zkh::mul(FQ, p.X, p.X), B = zkh::mul(FQ, p.Y, p.Y), C = zkh::mul(FQ, B, B);
  Fe xb = zkh::add(FQ, p.X, B);
  Fe D = dbl(zkh::sub(FQ, zkh::sub(FQ, zkh::mul(FQ, xb, xb), A), C));
  Fe E = zkh::add(FQ, dbl(A), A);
  Fe F = zkh::mul(FQ, E, E);
  J1 r;
  r.X = zkh::sub(FQ, F, dbl(D));
  r.Y = zkh::sub(FQ, zkh::mul(FQ, E, zkh::sub(FQ, D, r.X)), dbl(dbl(dbl(C))));
  r.Z = dbl(zkh::mul(FQ, p.Y, p.Z));
  return r;
}

// add-2007-bl with the exceptional cases
static J1 jadd(const J1& p, const J1& q) {
  if (inf(p)) return q;
  if (inf(q)) return p;
  Fe z1z1 = zkh::mul(FQ, p.Z, p.Z), z2z2 = zkh::mul(FQ, q.Z, q.Z);
  Fe u1 = zkh::mul(FQ, p.X, z2z2), u2 = zkh::mul(FQ, q.X, z1z1);
  Fe s1 = zkh::mul(FQ, p.Y, zkh::mul(FQ, q.Z, z2z2)), s2 = zkh::mul(FQ, q.Y, zkh::mul(FQ, p.Z, z1z1));
  Fe h = zkh::sub(FQ, u2, u1), rr = zkh::sub(FQ, s2, s1);
  if (zkh::is_zero(h)) {
    if (zkh::is_zero(rr)) return jdbl(p);
    return {zkh::one(FQ), zkh::one(FQ), zkh::zero()};
  }
  Fe h2 = dbl(h);
  Fe i = zkh::mul(FQ, h2, h2);
  Fe j = zkh::mul(FQ, h, i);
  rr = dbl(rr);
  Fe v = zkh::mul(FQ, u1, i);
  J1 r;
  r.X = zkh::sub(FQ, zkh::sub(FQ, zkh::mul(FQ, rr, rr), j), dbl(v));
  r.Y = zkh::sub(FQ, zkh::mul(FQ, rr, zkh::sub(FQ, v, r.X)), dbl(zkh::mul(FQ, s1, j)));
  Fe zz = zkh::add(FQ, p.Z, q.Z);
  r.Z = zkh::mul(FQ, zkh::sub(FQ, zkh::sub(FQ, zkh::mul(FQ, zz, zz), z1z1), z2z2), h);
  return r;
}

static Fe words64(const fe& a) {
  Fe r;
  for (int i = 0; i < 4; i++) r.v[i] = (uint64_t)a.w[2 * i] | ((uint64_t)a.w[2 * i + 1] << 32);
  return r;
}

// XYZZ (x = X / ZZ, y = Y / ZZZ, ZZ^3 = ZZZ^2) -> Jacobian (X ZZ^2, Y ZZZ^2, ZZZ)
static J1 from_xyzz(const xyzz& p) {
  const Fe zz = words64(p.zz), zzz = words64(p.zzz);
  if (zkh::is_zero(zz)) return {zkh::one(FQ), zkh::one(FQ), zkh::zero()};
  return {zkh::mul(FQ, words64(p.x), zkh::mul(FQ, zz, zz)), zkh::mul(FQ, words64(p.y), zkh::mul(FQ, zzz, zzz)),
          zzz};
}

}  // namespace g1h

// The window combination of an MSM, sum_w 2^(c w) S_w by Horner's rule, on
// the host in Jacobian coordinates over the 4 x 64-bit field (a doubling is
// 7 products).  S_w: XYZZ, R = 2^256 Montgomery words.  out: Jacobian X, Y,
// Z (4 words each), all zero for the identity.
void g1_window_combine(const xyzz* s, uint32_t W, int c, uint64_t out[12]) {
  memset(out, 0, 12 * sizeof(uint64_t));
  if (W == 0) return;
  g1h::J1 acc = g1h::from_xyzz(s[W - 1]);
  for (int w = (int)W - 2; w >= 0; w--) {
    for (int d = 0; d < c; d++) acc = g1h::jdbl(acc);
    acc = g1h::jadd(acc, g1h::from_xyzz(s[w]));
  }
  if (g1h::inf(acc)) return;
  memcpy(out, acc.X.v, 32);
  memcpy(out + 4, acc.Y.v, 32);
  memcpy(out + 8, acc.Z.v, 32);
}

// sum_i 2^pos[i] t[i] by Horner's rule, the terms given in strictly
// decreasing bit position (XYZZ, R = 2^256 Montgomery words); out as
// g1_window_combine's
void g1_bits_combine(const xyzz* t, const uint32_t* pos, uint32_t count, uint64_t out[12]) {
  memset(out, 0, 12 * sizeof(uint64_t));
  if (count == 0) return;
  g1h::J1 acc = g1h::from_xyzz(t[0]);
  uint32_t at = pos[0];
  for (uint32_t i = 1; i < count; i++) {
    for (; at > pos[i]; at--) acc = g1h::jdbl(acc);
    acc = g1h::jadd(acc, g1h::from_xyzz(t[i]));
  }
  for (; at > 0; at--) acc = g1h::jdbl(acc);
  if (g1h::inf(acc)) return;
  memcpy(out, acc.X.v, 32);
  memcpy(out + 4, acc.Y.v, 32);
  memcpy(out + 8, acc.Z.v, 32);
}

}  // namespace zk

using namespace zk;

extern "C" {

/* out = [k] P: P affine canonical (8 words, zeros = infinity), k canonical
 * (reduced mod r first), out affine canonical. */
int zk_g1_mul_host(const uint64_t p[8], const uint64_t k[4], uint64_t out[8]) {
  if (!p || !k || !out) return ZK_ERR_INVALID;
  using zkh::Fe;
  Fe x, y, s;
  memcpy(x.v, p, 32);
  memcpy(y.v, p + 4, 32);
  memcpy(s.v, k, 32);
  while (zkh::geq(s.v, zkh::FR.p)) zkh::sub_raw(s.v, s.v, zkh::FR.p);
  g1h::J1 m[16];
  m[0] = {zkh::one(zkh::FQ), zkh::one(zkh::FQ), zkh::zero()};
  const bool pinf = zkh::is_zero(x) && zkh::is_zero(y);
  if (!pinf) m[1] = {zkh::to_mont(zkh::FQ, x), zkh::to_mont(zkh::FQ, y), zkh::one(zkh::FQ)};
  else m[1] = m[0];
  for (int d = 2; d < 16; d++) m[d] = g1h::jadd(m[d - 1], m[1]);
  g1h::J1 acc = m[0];
  for (int j = 63; j >= 0; j--) {
    for (int t = 0; t < 4; t++) acc = g1h::jdbl(acc);
    const uint32_t d = (uint32_t)((s.v[j >> 4] >> (4 * (j & 15))) & 0xf);
    if (d) acc = g1h::jadd(acc, m[d]);
  }
  if (g1h::inf(acc)) {
    memset(out, 0, 64);
    return ZK_OK;
  }
  Fe zi = zkh::inv(zkh::FQ, acc.Z), zi2 = zkh::mul(zkh::FQ, zi, zi);
  Fe ax = zkh::from_mont(zkh::FQ, zkh::mul(zkh::FQ, acc.X, zi2));
  Fe ay = zkh::from_mont(zkh::FQ, zkh::mul(zkh::FQ, acc.Y, zkh::mul(zkh::FQ, zi2, zi)));
  memcpy(out, ax.v, 32);
  memcpy(out + 4, ay.v, 32);
  return ZK_OK;
}

}  // extern "C"
