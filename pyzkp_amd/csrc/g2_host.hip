// Host-side BN254 G2 arithmetic: Fq2 = Fq[u]/(u^2 + 1), the sextic twist
// E'(Fq2): y^2 = x^3 + 3/(9 + u), Jacobian group law, a multi-threaded
// Pippenger MSM and a fixed-base multiplier.
//
// The served Groth16 prover keeps the G2 MSM of B (b_query_g2 . w) on the
// CPU, as the reference does (MSM_G2 has no GPU kernel there either:
// crypto/msm/pippenger.py:67-121 behind runtime/kernels/cpu/kernels.py:338-348),
// overlapped with the GPU work of other proofs.  Points cross the C-ABI as
// canonical (non-Montgomery) little-endian words: affine (x0, x1, y0, y1)
// with all-zero for infinity, 16 x u64 per point.
#include <string.h>
#include <atomic>
#include <thread>
#include <vector>
#include "internal.hpp"

namespace zk {
namespace g2 {

using zkh::Fe;
using zkh::FQ;

struct F2 {
  Fe c0, c1;
};

static inline F2 f2_add(const F2& a, const F2& b) { return {zkh::add(FQ, a.c0, b.c0), zkh::add(FQ, a.c1, b.c1)}; }
static inline F2 f2_sub(const F2& a, const F2& b) { return {zkh::sub(FQ, a.c0, b.c0), zkh::sub(FQ, a.c1, b.c1)}; }
static inline F2 f2_dbl(const F2& a) { return f2_add(a, a); }
static inline F2 f2_neg(const F2& a) { return {zkh::neg(FQ, a.c0), zkh::neg(FQ, a.c1)}; }
static inline bool f2_zero(const F2& a) { return zkh::is_zero(a.c0) && zkh::is_zero(a.c1); }
// (a0 + a1 u)(b0 + b1 u) = (a0 b0 - a1 b1) + ((a0 + a1)(b0 + b1) - a0 b0 - a1 b1) u
static inline F2 f2_mul(const F2& a, const F2& b) {
  Fe t0 = zkh::mul(FQ, a.c0, b.c0), t1 = zkh::mul(FQ, a.c1, b.c1);
  Fe t2 = zkh::mul(FQ, zkh::add(FQ, a.c0, a.c1), zkh::add(FQ, b.c0, b.c1));
  return {zkh::sub(FQ, t0, t1), zkh::sub(FQ, zkh::sub(FQ, t2, t0), t1)};
}
static inline F2 f2_sqr(const F2& a) {
  // (a0 + a1)(a0 - a1) + 2 a0 a1 u
  Fe t = zkh::mul(FQ, a.c0, a.c1);
  return {zkh::mul(FQ, zkh::add(FQ, a.c0, a.c1), zkh::sub(FQ, a.c0, a.c1)), zkh::add(FQ, t, t)};
}
static inline F2 f2_inv(const F2& a) {
  // (a0 - a1 u) / (a0^2 + a1^2)
  Fe n = zkh::add(FQ, zkh::mul(FQ, a.c0, a.c0), zkh::mul(FQ, a.c1, a.c1));
  Fe ni = zkh::inv(FQ, n);
  return {zkh::mul(FQ, a.c0, ni), zkh::neg(FQ, zkh::mul(FQ, a.c1, ni))};
}

struct J2 {
  F2 X, Y, Z;
};

static inline J2 j_inf() {
  J2 r;
  r.X = {zkh::one(FQ), zkh::zero()};
  r.Y = r.X;
  r.Z = {zkh::zero(), zkh::zero()};
  return r;
}
static inline bool j_is_inf(const J2& p) { return f2_zero(p.Z); }

// dbl-2009-l (a = 0)
static J2 j_dbl(const J2& p) {
  if (j_is_inf(p) || f2_zero(p.Y)) return j_inf();
  F2 A = f2_sqr(p.X), B = f2_sqr(p.Y), C = f2_sqr(B);
  F2 xb = f2_add(p.X, B);
  F2 D = f2_dbl(f2_sub(f2_sub(f2_sqr(xb), A), C));
  F2 E = f2_add(f2_dbl(A), A);
  F2 F = f2_sqr(E);
  J2 r;
  r.X = f2_sub(F, f2_dbl(D));
  F2 c8 = f2_dbl(f2_dbl(f2_dbl(C)));
  r.Y = f2_sub(f2_mul(E, f2_sub(D, r.X)), c8);
  r.Z = f2_dbl(f2_mul(p.Y, p.Z));
  return r;
}

// add-2007-bl
static J2 j_add(const J2& p, const J2& q) {
  if (j_is_inf(p)) return q;
  if (j_is_inf(q)) return p;
  F2 z1z1 = f2_sqr(p.Z), z2z2 = f2_sqr(q.Z);
  F2 u1 = f2_mul(p.X, z2z2), u2 = f2_mul(q.X, z1z1);
  F2 s1 = f2_mul(p.Y, f2_mul(q.Z, z2z2)), s2 = f2_mul(q.Y, f2_mul(p.Z, z1z1));
  F2 h = f2_sub(u2, u1), rr = f2_sub(s2, s1);
  if (f2_zero(h)) {
    if (f2_zero(rr)) return j_dbl(p);
    return j_inf();
  }
  F2 i = f2_sqr(f2_dbl(h));
  F2 j = f2_mul(h, i);
  rr = f2_dbl(rr);
  F2 v = f2_mul(u1, i);
  J2 r;
  r.X = f2_sub(f2_sub(f2_sqr(rr), j), f2_dbl(v));
  r.Y = f2_sub(f2_mul(rr, f2_sub(v, r.X)), f2_dbl(f2_mul(s1, j)));
  F2 zz = f2_add(p.Z, q.Z);
  r.Z = f2_mul(f2_sub(f2_sub(f2_sqr(zz), z1z1), z2z2), h);
  return r;
}

// madd-2007-bl: p + (x, y) affine (not infinity)
static J2 j_madd(const J2& p, const F2& x, const F2& y) {
  if (j_is_inf(p)) {
    J2 r;
    r.X = x;
    r.Y = y;
    r.Z = {zkh::one(FQ), zkh::zero()};
    return r;
  }
  F2 z1z1 = f2_sqr(p.Z);
  F2 u2 = f2_mul(x, z1z1), s2 = f2_mul(y, f2_mul(p.Z, z1z1));
  F2 h = f2_sub(u2, p.X), rr = f2_sub(s2, p.Y);
  if (f2_zero(h)) {
    if (f2_zero(rr)) return j_dbl(p);
    return j_inf();
  }
  F2 hh = f2_sqr(h);
  F2 i = f2_dbl(f2_dbl(hh));
  F2 j = f2_mul(h, i);
  rr = f2_dbl(rr);
  F2 v = f2_mul(p.X, i);
  J2 r;
  r.X = f2_sub(f2_sub(f2_sqr(rr), j), f2_dbl(v));
  r.Y = f2_sub(f2_mul(rr, f2_sub(v, r.X)), f2_dbl(f2_mul(p.Y, j)));
  r.Z = f2_sub(f2_sub(f2_sqr(f2_add(p.Z, h)), z1z1), hh);
  return r;
}

static inline Fe rd_fq(const uint64_t* w) {
  Fe a;
  memcpy(a.v, w, 32);
  return zkh::to_mont(FQ, a);
}
static inline void wr_fq(uint64_t* w, const Fe& a) {
  Fe c = zkh::from_mont(FQ, a);
  memcpy(w, c.v, 32);
}

struct A2 {
  F2 x, y;
  bool inf;
};

static A2 rd_aff(const uint64_t* w) {
  A2 p;
  p.x = {rd_fq(w), rd_fq(w + 4)};
  p.y = {rd_fq(w + 8), rd_fq(w + 12)};
  p.inf = f2_zero(p.x) && f2_zero(p.y);
  return p;
}

static void wr_jac_affine(uint64_t* out, const J2& p) {
  if (j_is_inf(p)) {
    memset(out, 0, 128);
    return;
  }
  F2 zi = f2_inv(p.Z), zi2 = f2_sqr(zi), zi3 = f2_mul(zi2, zi);
  F2 x = f2_mul(p.X, zi2), y = f2_mul(p.Y, zi3);
  wr_fq(out, x.c0);
  wr_fq(out + 4, x.c1);
  wr_fq(out + 8, y.c0);
  wr_fq(out + 12, y.c1);
}

template <class F>
static void parallel_for(uint64_t n, int threads, F fn) {
  if (threads <= 1 || n <= 1) {
    for (uint64_t i = 0; i < n; i++) fn(i);
    return;
  }
  std::vector<std::thread> ts;
  std::atomic<uint64_t> next(0);
  for (int t = 0; t < threads; t++)
    ts.emplace_back([&]() {
      for (;;) {
        uint64_t i = next.fetch_add(1);
        if (i >= n) break;
        fn(i);
      }
    });
  for (auto& t : ts) t.join();
}

static uint32_t bits_at(const uint64_t* s, uint32_t off, uint32_t c) {
  const uint32_t wi = off >> 6, sh = off & 63;
  unsigned __int128 v = (unsigned __int128)s[wi] >> sh;
  if (wi + 1 < 4) v |= (unsigned __int128)s[wi + 1] << (64 - sh);
  return (uint32_t)(v & ((1ull << c) - 1));
}

// sum_i s_i Q_i: unsigned c-bit windows; each (window, bucket slice) task
// runs on one thread; tasks combine in order.
static void msm(const uint64_t* bases, const uint64_t* scalars, uint64_t n, int c, int threads,
                uint64_t out[16]) {
  std::vector<A2> pts(n);
  std::vector<uint64_t> ss(4 * n);
  parallel_for(n, threads, [&](uint64_t i) {
    pts[i] = rd_aff(bases + 16 * i);
    Fe s;
    memcpy(s.v, scalars + 4 * i, 32);
    while (zkh::geq(s.v, zkh::FR.p)) zkh::sub_raw(s.v, s.v, zkh::FR.p);
    memcpy(&ss[4 * i], s.v, 32);
  });
  const uint32_t W = (254 + c - 1) / c;
  const uint32_t nb = 1u << c;
  std::vector<J2> wsum(W);
  parallel_for(W, threads, [&](uint64_t w) {
    std::vector<J2> buckets(nb, j_inf());
    for (uint64_t i = 0; i < n; i++) {
      if (pts[i].inf) continue;
      uint32_t d = bits_at(&ss[4 * i], (uint32_t)w * c, c);
      if (d) buckets[d] = j_madd(buckets[d], pts[i].x, pts[i].y);
    }
    J2 run = j_inf(), sum = j_inf();
    for (uint32_t d = nb - 1; d > 0; d--) {
      run = j_add(run, buckets[d]);
      sum = j_add(sum, run);
    }
    wsum[w] = sum;
  });
  J2 acc = j_inf();
  for (int w = (int)W - 1; w >= 0; w--) {
    for (int k = 0; k < c; k++) acc = j_dbl(acc);
    acc = j_add(acc, wsum[w]);
  }
  wr_jac_affine(out, acc);
}

}  // namespace g2
}  // namespace zk

using namespace zk;

extern "C" {

/* G2 MSM on the host: bases affine canonical (16 words each, zeros =
 * infinity), scalars canonical (4 words each), window c (0: default rule),
 * threads (0: hardware concurrency).  out: affine canonical, 16 words. */
int zk_g2_msm_host(const uint64_t* bases, const uint64_t* scalars, uint64_t n, int window_bits, int threads,
                   uint64_t out[16]) {
  if (n && (!bases || !scalars)) return ZK_ERR_INVALID;
  if (!out) return ZK_ERR_INVALID;
  if (n == 0) {
    memset(out, 0, 128);
    return ZK_OK;
  }
  int c = window_bits;
  if (c <= 0) {
    int bl = 0;
    while ((n >> bl) != 0) bl++;
    c = bl - 2;
    if (c > 16) c = 16;
    if (c < 4) c = 4;
  }
  if (c > 20) return ZK_ERR_INVALID;
  if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
  if (threads < 1) threads = 1;
  g2::msm(bases, scalars, n, c, threads, out);
  return ZK_OK;
}

/* out[i] = [k_i] Q for one point Q (affine canonical, 16 words) and n
 * canonical scalars, affine canonical outputs (16 words each). */
int zk_g2_mul_fixed_host(const uint64_t* q, const uint64_t* scalars, uint64_t n, int threads, uint64_t* out) {
  if (!q || (n && (!scalars || !out))) return ZK_ERR_INVALID;
  if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
  if (threads < 1) threads = 1;
  g2::A2 base = g2::rd_aff(q);
  if (n < 64) {
    // few scalars: 4-bit fixed windows over 16 multiples, no table set-up
    g2::J2 m[16];
    m[0] = g2::j_inf();
    for (int d = 1; d < 16; d++) m[d] = base.inf ? g2::j_inf() : g2::j_madd(m[d - 1], base.x, base.y);
    for (uint64_t i = 0; i < n; i++) {
      zkh::Fe k;
      memcpy(k.v, scalars + 4 * i, 32);
      while (zkh::geq(k.v, zkh::FR.p)) zkh::sub_raw(k.v, k.v, zkh::FR.p);
      g2::J2 acc = g2::j_inf();
      for (int j = 63; j >= 0; j--) {
        for (int t = 0; t < 4; t++) acc = g2::j_dbl(acc);
        uint32_t d = (uint32_t)((k.v[j >> 4] >> (4 * (j & 15))) & 0xf);
        if (d) acc = g2::j_add(acc, m[d]);
      }
      g2::wr_jac_affine(out + 16 * i, acc);
    }
    return ZK_OK;
  }
  // table[j][d] = d 2^(8 j) Q as Jacobian, j < 32, d < 256
  std::vector<g2::J2> tab(32 * 256, g2::j_inf());
  if (!base.inf) {
    g2::J2 b;
    b.X = base.x;
    b.Y = base.y;
    b.Z = {zkh::one(zkh::FQ), zkh::zero()};
    for (int j = 0; j < 32; j++) {
      for (int d = 1; d < 256; d++) tab[j * 256 + d] = g2::j_add(tab[j * 256 + d - 1], b);
      b = g2::j_add(tab[j * 256 + 255], b);
    }
  }
  // normalise the table so the per-scalar loop uses mixed additions
  std::vector<g2::A2> at(32 * 256);
  g2::parallel_for(32 * 256, threads, [&](uint64_t t) {
    if (g2::j_is_inf(tab[t])) {
      at[t].inf = true;
      return;
    }
    g2::F2 zi = g2::f2_inv(tab[t].Z), zi2 = g2::f2_sqr(zi);
    at[t].x = g2::f2_mul(tab[t].X, zi2);
    at[t].y = g2::f2_mul(tab[t].Y, g2::f2_mul(zi2, zi));
    at[t].inf = false;
  });
  g2::parallel_for(n, threads, [&](uint64_t i) {
    zkh::Fe k;
    memcpy(k.v, scalars + 4 * i, 32);
    while (zkh::geq(k.v, zkh::FR.p)) zkh::sub_raw(k.v, k.v, zkh::FR.p);
    g2::J2 acc = g2::j_inf();
    for (int j = 0; j < 32; j++) {
      uint32_t d = (uint32_t)((k.v[j >> 3] >> (8 * (j & 7))) & 0xff);
      if (d && !at[j * 256 + d].inf) acc = g2::j_madd(acc, at[j * 256 + d].x, at[j * 256 + d].y);
    }
    g2::wr_jac_affine(out + 16 * i, acc);
  });
  return ZK_OK;
}

/* p + q on G2 (affine canonical in / out, 16 words). */
int zk_g2_add_host(const uint64_t* p, const uint64_t* q, uint64_t out[16]) {
  if (!p || !q || !out) return ZK_ERR_INVALID;
  g2::A2 a = g2::rd_aff(p), b = g2::rd_aff(q);
  g2::J2 r = g2::j_inf();
  if (!a.inf) r = g2::j_madd(r, a.x, a.y);
  if (!b.inf) r = g2::j_madd(r, b.x, b.y);
  g2::wr_jac_affine(out, r);
  return ZK_OK;
}

}  // extern "C"
