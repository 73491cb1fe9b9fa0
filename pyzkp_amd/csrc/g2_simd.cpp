// Host G2 MSM bucket accumulation and bucket sums on AVX-512 IFMA.
//
// Eight Fq values per vector, each as five 52-bit limbs in Montgomery form
// with R = 2^260.  A product is operand scanning over b's limbs: per limb,
// a * b_i and m * q go into 64-bit column accumulators through
// vpmadd52luq / vpmadd52huq (the low and high 52 bits of a 52 x 52-bit
// product), m = t_0 (-q^-1) mod 2^52 clears the low column, and its carry
// moves up.  For inputs a, b with a b < q 2^260 (both below 9q, say) the
// result (a b + m q) / 2^260 is below 2q, so values stay lazily reduced and
// only stored coordinates are made canonical.  Differences a - b are taken
// as a + k q - b with a k q above any value b can have, and the limb
// carries of a sum or difference are propagated with arithmetic shifts.
//
// The bucket additions are affine and batched (the algorithm of the scalar
// path in g2_host.hip): a batch of additions to distinct buckets shares one
// inversion.  Eight additions run side by side, one per lane; every lane
// keeps its own product chain of the denominators' norms (1 / d = conj(d) /
// N(d), N(d) = d0^2 + d1^2 in Fq, nonzero for d != 0 because -1 is not a
// square mod q), and one vector exponentiation by q - 2 inverts all eight
// chains.  A point whose bucket already has an addition queued goes to a
// replica slot of that bucket, and the replicas are folded back pairwise
// through the same batched additions, so any digit distribution stays on
// the batched path.
//
// The weighted bucket sum sum_b (b + 1) B_b then runs on the lanes too:
// each lane keeps XYZZ running sums over one eighth of the buckets.
//
// This file is the only one built with AVX-512 code generation and it is
// entered only after the caller checked the CPU.  It shares no inline
// function or template instantiation with the other translation units
// (everything here is static, it includes no C++ library templates), so the
// linker cannot pick an AVX-512 copy of a function that code on the scalar
// path also calls.
#include <immintrin.h>
#include <stdlib.h>
#include <string.h>

#include "g2_simd.hpp"

namespace zk {
namespace g2simd {
namespace {

typedef __m512i V;
static constexpr uint64_t M52 = (1ull << 52) - 1;

// ---- limb conversions (values below 2^256) ----
static inline void to52(const uint64_t* w, uint64_t* l) {
  l[0] = w[0] & M52;
  l[1] = ((w[0] >> 52) | (w[1] << 12)) & M52;
  l[2] = ((w[1] >> 40) | (w[2] << 24)) & M52;
  l[3] = ((w[2] >> 28) | (w[3] << 36)) & M52;
  l[4] = w[3] >> 16;
}

// ---- constants ----
struct Consts {
  uint64_t q[5];
  uint64_t qinv;       // -q^-1 mod 2^52
  uint64_t kq[9][5];   // k q, limbs normalised (k <= 8: below 2^257)
  uint64_t r2[5];      // 2^520 mod q: mont(x, r2) = x R
  uint64_t one[5];     // R mod q = 2^260 mod q
  uint64_t to256[5];   // 2^256 mod q: mont(x R, to256) = x 2^256 (the 4 x 64-bit form)
  uint64_t e[4];       // q - 2
};

static const uint64_t QW64[4] = {0x3c208c16d87cfd47ull, 0x97816a916871ca8dull, 0xb85045b68181585dull,
                                  0x30644e72e131a029ull};

// t <- 2 t mod q for t < q (four words)
static void dbl_mod(uint64_t* t) {
  uint64_t s[4], d[4], c = 0, br = 0;
  for (int j = 0; j < 4; j++) {
    const unsigned __int128 v = ((unsigned __int128)t[j] << 1) + c;
    s[j] = (uint64_t)v;
    c = (uint64_t)(v >> 64);
  }
  for (int j = 0; j < 4; j++) {
    const unsigned __int128 v = (unsigned __int128)s[j] - QW64[j] - br;
    d[j] = (uint64_t)v;
    br = (uint64_t)(v >> 127);
  }
  const bool ge = c || !br;  // 2t >= q
  for (int j = 0; j < 4; j++) t[j] = ge ? d[j] : s[j];
}

static Consts make_consts() {
  Consts c;
  const uint64_t* p = QW64;
  to52(p, c.q);
  uint64_t x = p[0];  // Newton iteration for p^-1 mod 2^64 (p odd: x = p is right mod 8)
  for (int i = 0; i < 6; i++) x *= 2 - p[0] * x;
  c.qinv = (0 - x) & M52;
  for (int k = 0; k <= 8; k++) {
    uint64_t carry = 0;
    for (int j = 0; j < 5; j++) {
      const uint64_t v = (uint64_t)k * c.q[j] + carry;
      c.kq[k][j] = j < 4 ? (v & M52) : v;
      carry = v >> 52;
    }
  }
  uint64_t t[4] = {1, 0, 0, 0};
  for (int i = 1; i <= 520; i++) {
    dbl_mod(t);  // t = 2^i mod q
    if (i == 256) to52(t, c.to256);
    if (i == 260) to52(t, c.one);
  }
  to52(t, c.r2);
  memcpy(c.e, p, 32);
  c.e[0] -= 2;  // p odd and > 2
  return c;
}
static const Consts& consts() {
  static const Consts c = make_consts();
  return c;
}

// ---- eight-lane field arithmetic ----
struct F {
  V l[5];
};
struct F2 {
  F c0, c1;
};

struct VC {  // broadcast constants
  V zero, m52, qinv;
  V q[5];
  V kq[9][5];
  F one;
  F to256;
};
static void make_vc(VC& v) {
  const Consts& c = consts();
  v.zero = _mm512_setzero_si512();
  v.m52 = _mm512_set1_epi64((long long)M52);
  v.qinv = _mm512_set1_epi64((long long)c.qinv);
  for (int j = 0; j < 5; j++) {
    v.q[j] = _mm512_set1_epi64((long long)c.q[j]);
    v.one.l[j] = _mm512_set1_epi64((long long)c.one[j]);
    v.to256.l[j] = _mm512_set1_epi64((long long)c.to256[j]);
    for (int k = 0; k <= 8; k++) v.kq[k][j] = _mm512_set1_epi64((long long)c.kq[k][j]);
  }
}

static inline V mlo(V acc, V a, V b) { return _mm512_madd52lo_epu64(acc, a, b); }
static inline V mhi(V acc, V a, V b) { return _mm512_madd52hi_epu64(acc, a, b); }

// r = a b R^-1, below 2q when a b < q 2^260; limbs of a, b below 2^52
static inline void mul(F& r, const F& a, const F& b, const VC& c) {
  V t0 = c.zero, t1 = c.zero, t2 = c.zero, t3 = c.zero, t4 = c.zero;
#pragma GCC unroll 5
  for (int i = 0; i < 5; i++) {
    const V bi = b.l[i];
    V t5 = c.zero;
    t0 = mlo(t0, a.l[0], bi);
    t1 = mhi(t1, a.l[0], bi);
    t1 = mlo(t1, a.l[1], bi);
    t2 = mhi(t2, a.l[1], bi);
    t2 = mlo(t2, a.l[2], bi);
    t3 = mhi(t3, a.l[2], bi);
    t3 = mlo(t3, a.l[3], bi);
    t4 = mhi(t4, a.l[3], bi);
    t4 = mlo(t4, a.l[4], bi);
    t5 = mhi(t5, a.l[4], bi);
    const V m = mlo(c.zero, t0, c.qinv);
    t0 = mlo(t0, m, c.q[0]);
    t1 = mhi(t1, m, c.q[0]);
    t1 = mlo(t1, m, c.q[1]);
    t2 = mhi(t2, m, c.q[1]);
    t2 = mlo(t2, m, c.q[2]);
    t3 = mhi(t3, m, c.q[2]);
    t3 = mlo(t3, m, c.q[3]);
    t4 = mhi(t4, m, c.q[3]);
    t4 = mlo(t4, m, c.q[4]);
    t5 = mhi(t5, m, c.q[4]);
    // t0 is now a multiple of 2^52: drop the column, keep its carry
    t0 = _mm512_add_epi64(t1, _mm512_srli_epi64(t0, 52));
    t1 = t2;
    t2 = t3;
    t3 = t4;
    t4 = t5;
  }
  t1 = _mm512_add_epi64(t1, _mm512_srli_epi64(t0, 52));
  r.l[0] = _mm512_and_si512(t0, c.m52);
  t2 = _mm512_add_epi64(t2, _mm512_srli_epi64(t1, 52));
  r.l[1] = _mm512_and_si512(t1, c.m52);
  t3 = _mm512_add_epi64(t3, _mm512_srli_epi64(t2, 52));
  r.l[2] = _mm512_and_si512(t2, c.m52);
  t4 = _mm512_add_epi64(t4, _mm512_srli_epi64(t3, 52));
  r.l[3] = _mm512_and_si512(t3, c.m52);
  r.l[4] = t4;
}

// limb carries of a value that may have negative limbs but is >= 0
static inline void snorm(F& r, V t0, V t1, V t2, V t3, V t4, const VC& c) {
  t1 = _mm512_add_epi64(t1, _mm512_srai_epi64(t0, 52));
  r.l[0] = _mm512_and_si512(t0, c.m52);
  t2 = _mm512_add_epi64(t2, _mm512_srai_epi64(t1, 52));
  r.l[1] = _mm512_and_si512(t1, c.m52);
  t3 = _mm512_add_epi64(t3, _mm512_srai_epi64(t2, 52));
  r.l[2] = _mm512_and_si512(t2, c.m52);
  t4 = _mm512_add_epi64(t4, _mm512_srai_epi64(t3, 52));
  r.l[3] = _mm512_and_si512(t3, c.m52);
  r.l[4] = t4;
}
static inline void add(F& r, const F& a, const F& b, const VC& c) {
  snorm(r, _mm512_add_epi64(a.l[0], b.l[0]), _mm512_add_epi64(a.l[1], b.l[1]), _mm512_add_epi64(a.l[2], b.l[2]),
        _mm512_add_epi64(a.l[3], b.l[3]), _mm512_add_epi64(a.l[4], b.l[4]), c);
}
// a + b limb by limb, carries left in the limbs (below 2^53): only for a
// value that feeds a difference (subk normalises), never a product
static inline void add_raw(F& r, const F& a, const F& b) {
  for (int j = 0; j < 5; j++) r.l[j] = _mm512_add_epi64(a.l[j], b.l[j]);
}
// r = a + k q - b, for b < k q
template <int K>
static inline void subk(F& r, const F& a, const F& b, const VC& c) {
  V t[5];
  for (int j = 0; j < 5; j++) t[j] = _mm512_sub_epi64(_mm512_add_epi64(a.l[j], c.kq[K][j]), b.l[j]);
  snorm(r, t[0], t[1], t[2], t[3], t[4], c);
}
// a <- a - k q where that is >= 0
template <int K>
static inline void csubk(F& a, const VC& c) {
  F d;
  snorm(d, _mm512_sub_epi64(a.l[0], c.kq[K][0]), _mm512_sub_epi64(a.l[1], c.kq[K][1]),
        _mm512_sub_epi64(a.l[2], c.kq[K][2]), _mm512_sub_epi64(a.l[3], c.kq[K][3]),
        _mm512_sub_epi64(a.l[4], c.kq[K][4]), c);
  const __mmask8 ge = _mm512_cmpge_epi64_mask(d.l[4], c.zero);
  for (int j = 0; j < 5; j++) a.l[j] = _mm512_mask_blend_epi64(ge, a.l[j], d.l[j]);
}
// below 8q -> canonical
static inline void canon8(F& a, const VC& c) {
  csubk<4>(a, c);
  csubk<2>(a, c);
  csubk<1>(a, c);
}
static inline void blend(F& a, const F& b, __mmask8 m) {
  for (int j = 0; j < 5; j++) a.l[j] = _mm512_mask_blend_epi64(m, a.l[j], b.l[j]);
}

// a^e for the exponent q - 2 (4-bit fixed windows), lane by lane
static void inv(F& r, const F& a, const VC& c) {
  const Consts& k = consts();
  F tab[16];
  tab[0] = c.one;
  tab[1] = a;
  for (int i = 2; i < 16; i++) mul(tab[i], tab[i - 1], a, c);
  bool started = false;
  F acc = c.one;
  for (int nib = 63; nib >= 0; nib--) {
    const uint32_t d = (uint32_t)((k.e[nib >> 4] >> (4 * (nib & 15))) & 0xf);
    if (started) {
      for (int s = 0; s < 4; s++) mul(acc, acc, acc, c);
      if (d) mul(acc, acc, tab[d], c);
    } else if (d) {
      acc = tab[d];
      started = true;
    }
  }
  r = acc;
}

}  // namespace

// ---- point preparation ----
void prepare_points(const uint64_t* bases16, uint64_t n, uint64_t* out20, uint8_t* inf) {
  VC c;
  make_vc(c);
  const Consts& k = consts();
  F r2;
  for (int j = 0; j < 5; j++) r2.l[j] = _mm512_set1_epi64((long long)k.r2[j]);
  alignas(64) uint64_t s[PT_WORDS][8];
  for (uint64_t i0 = 0; i0 < n; i0 += 8) {
    const int m = (int)(n - i0 < 8 ? n - i0 : 8);
    memset(s, 0, sizeof(s));
    for (int lane = 0; lane < m; lane++) {
      const uint64_t* w = bases16 + 16 * (i0 + lane);
      uint64_t any = 0;
      for (int t = 0; t < 16; t++) any |= w[t];
      inf[i0 + lane] = any == 0;
      for (int co = 0; co < 4; co++) {
        uint64_t l[5];
        to52(w + 4 * co, l);
        for (int j = 0; j < 5; j++) s[5 * co + j][lane] = l[j];
      }
    }
    for (int co = 0; co < 4; co++) {
      F x;
      for (int j = 0; j < 5; j++) x.l[j] = _mm512_load_si512(s[5 * co + j]);
      mul(x, x, r2, c);  // x 2^520 / 2^260 = x R, below 2q
      csubk<1>(x, c);
      for (int j = 0; j < 5; j++) _mm512_store_si512(s[5 * co + j], x.l[j]);
    }
    for (int lane = 0; lane < m; lane++)
      for (int t = 0; t < PT_WORDS; t++) out20[PT_WORDS * (i0 + lane) + t] = s[t][lane];
  }
}

namespace {

// ---- batched affine bucket additions ----
static constexpr int BATCH = 512;
static constexpr int GROUPS = BATCH / 8;
// Queue rows are BATCH + 8 words apart: at exactly 4 KiB apart the 20 words
// of one entry would share an L1 set (measured: 563 -> 200 cycles per entry
// of queueing work).
static constexpr int ROW = BATCH + 8;

struct Work {
  alignas(64) uint64_t qp[PT_WORDS][ROW];   // the point added (before its sign is applied)
  alignas(64) uint64_t qb[PT_WORDS][ROW];   // the slot's value when the addition was queued
  alignas(64) uint64_t res[PT_WORDS][ROW];  // the sums, canonical
  F pref[GROUPS], nrm[GROUPS];
  uint32_t qslot[BATCH];
  uint8_t qneg[BATCH], qdbl[BATCH];
};

static inline void ld(F& f, const uint64_t (*a)[ROW], int base, int g) {
  for (int j = 0; j < 5; j++) f.l[j] = _mm512_load_si512(a[base + j] + 8 * g);
}
static inline void st(uint64_t (*a)[ROW], int base, int g, const F& f) {
  for (int j = 0; j < 5; j++) _mm512_store_si512(a[base + j] + 8 * g, f.l[j]);
}
static inline __mmask8 lane_mask(const uint8_t* flags, int g) {
  uint32_t m = 0;
  for (int l = 0; l < 8; l++) m |= (uint32_t)(flags[8 * g + l] != 0) << l;
  return (__mmask8)m;
}

// lam's denominator for group g: x_P - x_B, or 2 y_B where the lane doubles
static inline void denom(F2& den, const Work& w, int g, __mmask8 dm, const VC& c) {
  F px, bx;
  ld(px, w.qp, 0, g);
  ld(bx, w.qb, 0, g);
  subk<2>(den.c0, px, bx, c);
  ld(px, w.qp, 5, g);
  ld(bx, w.qb, 5, g);
  subk<2>(den.c1, px, bx, c);
  if (dm) {
    F by, d2;
    ld(by, w.qb, 10, g);
    add(d2, by, by, c);
    blend(den.c0, d2, dm);
    ld(by, w.qb, 15, g);
    add(d2, by, by, c);
    blend(den.c1, d2, dm);
  }
}

static void flush(Work& w, int cnt, const VC& c) {
  const int G = (cnt + 7) / 8;
  for (int e = cnt; e < 8 * G; e++) {  // pad with copies of entry 0 (nonzero denominator)
    for (int t = 0; t < PT_WORDS; t++) {
      w.qp[t][e] = w.qp[t][0];
      w.qb[t][e] = w.qb[t][0];
    }
    w.qneg[e] = w.qneg[0];
    w.qdbl[e] = w.qdbl[0];
  }
  // forward: norms of the denominators and their running products
  F acc = c.one;
  for (int g = 0; g < G; g++) {
    F2 den;
    denom(den, w, g, lane_mask(w.qdbl, g), c);
    F s0, s1;
    mul(s0, den.c0, den.c0, c);
    mul(s1, den.c1, den.c1, c);
    add(w.nrm[g], s0, s1, c);  // < 4q
    w.pref[g] = acc;
    mul(acc, acc, w.nrm[g], c);
  }
  F iv;
  inv(iv, acc, c);
  // backward: each norm's inverse, then the chord (or tangent) step
  const F zero5 = {{c.zero, c.zero, c.zero, c.zero, c.zero}};
  for (int g = G - 1; g >= 0; g--) {
    const __mmask8 dm = lane_mask(w.qdbl, g), nm = lane_mask(w.qneg, g);
    F ni;
    mul(ni, iv, w.pref[g], c);
    mul(iv, iv, w.nrm[g], c);
    F2 den;
    denom(den, w, g, dm, c);
    F2 ie;  // 1 / den = conj(den) / N(den)
    F t;
    mul(ie.c0, den.c0, ni, c);
    mul(t, den.c1, ni, c);
    subk<2>(ie.c1, zero5, t, c);
    F2 px, py, bx, by;
    ld(px.c0, w.qp, 0, g);
    ld(px.c1, w.qp, 5, g);
    ld(py.c0, w.qp, 10, g);
    ld(py.c1, w.qp, 15, g);
    ld(bx.c0, w.qb, 0, g);
    ld(bx.c1, w.qb, 5, g);
    ld(by.c0, w.qb, 10, g);
    ld(by.c1, w.qb, 15, g);
    if (nm) {
      F n0, n1;
      subk<1>(n0, zero5, py.c0, c);
      subk<1>(n1, zero5, py.c1, c);
      blend(py.c0, n0, nm);
      blend(py.c1, n1, nm);
    }
    F2 num;  // y_P - y_B (< 4q), or 3 x_B^2 (< 6q) where the lane doubles
    subk<2>(num.c0, py.c0, by.c0, c);
    subk<2>(num.c1, py.c1, by.c1, c);
    if (dm) {
      F a, b, s0, s1, x0, x1;
      add(a, bx.c0, bx.c1, c);
      subk<1>(b, bx.c0, bx.c1, c);
      mul(x0, a, b, c);  // x0^2 - x1^2
      add(a, bx.c0, bx.c0, c);
      mul(x1, a, bx.c1, c);  // 2 x0 x1
      add(s0, x0, x0, c);
      add(s0, s0, x0, c);
      add(s1, x1, x1, c);
      add(s1, s1, x1, c);
      blend(num.c0, s0, dm);
      blend(num.c1, s1, dm);
    }
    // lam = num / den (Karatsuba): lam0 < 4q, and lam1 < 6q brought below 4q
    F2 lam;
    {
      F m0, m1, m2, sa, sb, m01;
      mul(m0, num.c0, ie.c0, c);
      mul(m1, num.c1, ie.c1, c);
      add(sa, num.c0, num.c1, c);
      add(sb, ie.c0, ie.c1, c);
      mul(m2, sa, sb, c);
      subk<2>(lam.c0, m0, m1, c);
      add_raw(m01, m0, m1);
      subk<4>(lam.c1, m2, m01, c);
      csubk<2>(lam.c1, c);
    }
    // x3 = lam^2 - x_B - x_P: (lam0 + lam1)(lam0 - lam1 + 4q) and 2 lam0 lam1
    // (products below 64q^2 and 32q^2), so c0, c1 < 2q and x3 < 4q
    F2 x3;
    {
      F s, d, t, c0, c1, bp;
      add(s, lam.c0, lam.c1, c);
      subk<4>(d, lam.c0, lam.c1, c);
      mul(c0, s, d, c);
      add(t, lam.c0, lam.c0, c);
      mul(c1, t, lam.c1, c);
      add_raw(bp, bx.c0, px.c0);
      subk<2>(x3.c0, c0, bp, c);
      add_raw(bp, bx.c1, px.c1);
      subk<2>(x3.c1, c1, bp, c);
      csubk<2>(x3.c0, c);
      csubk<1>(x3.c0, c);
      csubk<2>(x3.c1, c);
      csubk<1>(x3.c1, c);
    }
    // y3 = lam (x_B - x3) - y_B
    F2 y3;
    {
      F d0, d1, m0, m1, m2, sa, sb, m01, t0, t1;
      subk<1>(d0, bx.c0, x3.c0, c);
      subk<1>(d1, bx.c1, x3.c1, c);
      mul(m0, lam.c0, d0, c);
      mul(m1, lam.c1, d1, c);
      add(sa, lam.c0, lam.c1, c);
      add(sb, d0, d1, c);
      mul(m2, sa, sb, c);
      subk<2>(t0, m0, m1, c);
      add_raw(m01, m0, m1);
      subk<4>(t1, m2, m01, c);
      subk<1>(y3.c0, t0, by.c0, c);
      subk<1>(y3.c1, t1, by.c1, c);
      canon8(y3.c0, c);
      canon8(y3.c1, c);
    }
    st(w.res, 0, g, x3.c0);
    st(w.res, 5, g, x3.c1);
    st(w.res, 10, g, y3.c0);
    st(w.res, 15, g, y3.c1);
  }
}

// q - y for canonical y (0 stays 0), five 52-bit limbs
static inline void neg52(const uint64_t* y, uint64_t* r) {
  const uint64_t* q = consts().q;
  uint64_t any = 0;
  for (int j = 0; j < 5; j++) any |= y[j];
  int64_t br = 0;
  for (int j = 0; j < 5; j++) {
    int64_t v = (int64_t)q[j] - (int64_t)y[j] + br;
    br = v >> 52;  // arithmetic: -1 on borrow
    r[j] = (uint64_t)v & M52;
  }
  if (!any)
    for (int j = 0; j < 5; j++) r[j] = 0;
}

// Bucket slots: the nb buckets, then replicas.  A point whose bucket has an
// addition pending goes to a replica of that bucket (a copy when the
// replica is empty, else an addition to it), so no point ever waits; the
// replicas of a bucket form a chain from the bucket, and merge_all() adds
// them back pairwise (each round halves every chain, and the additions of a
// round go to distinct slots, so they batch like any others).
static constexpr uint32_t REPLICAS = 2 * BATCH;

struct Slots {
  uint64_t* bk;       // nb bucket values (the caller's)
  uint64_t* rp;       // REPLICAS replica values
  uint8_t* empty;     // nb + REPLICAS
  uint8_t* busy;      // nb + REPLICAS
  int32_t* next;      // chain links, -1 at the end
  uint32_t* tail;     // nb: last slot of the bucket's chain
  uint32_t* cursor;   // nb: where the bucket's next point starts looking
  uint32_t* chained;  // buckets with replicas, in order of first use
  uint8_t* is_chained;
  uint32_t nb, n_chained, n_rep;
  uint64_t* at(uint32_t s) const { return s < nb ? bk + (uint64_t)PT_WORDS * s : rp + (uint64_t)PT_WORDS * (s - nb); }
};

// Fr modulus: scalars at or above it are reduced first
static const uint64_t RW64[4] = {0x43e1f593f0000001ull, 0x2833e84879b97091ull, 0xb85045b68181585dull,
                                 0x30644e72e131a029ull};

// signed digit of window w (bit offset off = w c): bits [off, off + c) of
// (s mod r) + K, minus 2^(c-1)
static inline int32_t digit(const uint64_t* sp, const uint64_t* K, uint32_t off, uint32_t c) {
  uint64_t s[4];
  memcpy(s, sp, 32);
  for (;;) {  // s >= r happens only for non-canonical input
    bool ge = true;
    for (int j = 3; j >= 0; j--)
      if (s[j] != RW64[j]) {
        ge = s[j] > RW64[j];
        break;
      }
    if (!ge) break;
    uint64_t br = 0;
    for (int j = 0; j < 4; j++) {
      const unsigned __int128 v = (unsigned __int128)s[j] - RW64[j] - br;
      s[j] = (uint64_t)v;
      br = (uint64_t)(v >> 127);
    }
  }
  uint64_t t[5], cy = 0;
  for (int j = 0; j < 4; j++) {
    const unsigned __int128 v = (unsigned __int128)s[j] + K[j] + cy;
    t[j] = (uint64_t)v;
    cy = (uint64_t)(v >> 64);
  }
  t[4] = K[4] + cy;
  const uint32_t wi = off >> 6, sh = off & 63;
  uint64_t v = t[wi] >> sh;
  if (sh && wi < 4) v |= t[wi + 1] << (64 - sh);
  return (int32_t)(v & ((1ull << c) - 1)) - (int32_t)(1u << (c - 1));
}

}  // namespace

uint64_t accumulate(const uint64_t* pts20, const uint8_t* inf, const uint64_t* scalars, const uint64_t* K,
                    uint32_t win, uint32_t cbits, uint64_t i0, uint64_t i1, uint32_t nb, uint64_t* bk20,
                    uint8_t* empty) {
  VC c;
  make_vc(c);
  const uint32_t ns = nb + REPLICAS;
  Work* wp = (Work*)aligned_alloc(64, (sizeof(Work) + 63) / 64 * 64);
  Slots S;
  S.bk = bk20;
  S.rp = (uint64_t*)malloc((size_t)REPLICAS * PT_WORDS * sizeof(uint64_t));
  S.empty = (uint8_t*)malloc(ns);
  S.busy = (uint8_t*)calloc(ns, 1);
  S.next = (int32_t*)malloc((size_t)ns * sizeof(int32_t));
  S.tail = (uint32_t*)malloc((size_t)nb * sizeof(uint32_t));
  S.cursor = (uint32_t*)malloc((size_t)nb * sizeof(uint32_t));
  S.chained = (uint32_t*)malloc((size_t)nb * sizeof(uint32_t));
  S.is_chained = (uint8_t*)calloc(nb, 1);
  uint32_t* touched = (uint32_t*)malloc((size_t)nb * sizeof(uint32_t));
  uint8_t* is_touched = (uint8_t*)calloc(nb, 1);
  auto release = [&]() {
    free(wp);
    free(S.rp);
    free(S.empty);
    free(S.busy);
    free(S.next);
    free(S.tail);
    free(S.cursor);
    free(S.chained);
    free(S.is_chained);
    free(touched);
    free(is_touched);
  };
  if (!wp || !S.rp || !S.empty || !S.busy || !S.next || !S.tail || !S.cursor || !S.chained || !S.is_chained ||
      !touched || !is_touched) {
    release();
    return UINT64_MAX;
  }
  S.nb = nb;
  S.n_chained = S.n_rep = 0;
  memcpy(S.empty, empty, nb);
  memset(S.empty + nb, 1, REPLICAS);
  for (uint32_t s = 0; s < ns; s++) S.next[s] = -1;
  for (uint32_t b = 0; b < nb; b++) S.tail[b] = S.cursor[b] = b;
  uint32_t n_touched = 0;
  Work& w = *wp;
  int cnt = 0;

  auto do_flush = [&]() {
    if (cnt == 0) return;
    flush(w, cnt, c);
    for (int e = 0; e < cnt; e++) {
      uint64_t* B = S.at(w.qslot[e]);
      for (int t = 0; t < PT_WORDS; t++) B[t] = w.res[t][e];
      S.busy[w.qslot[e]] = 0;
    }
    cnt = 0;
    for (uint32_t k = 0; k < n_touched; k++) {
      S.cursor[touched[k]] = touched[k];
      is_touched[touched[k]] = 0;
    }
    n_touched = 0;
  };
  auto enqueue = [&](const uint64_t* P, uint32_t s, uint8_t neg, uint8_t dbl) {
    const uint64_t* B = S.at(s);
    for (int t = 0; t < PT_WORDS; t++) {
      w.qp[t][cnt] = P[t];
      w.qb[t][cnt] = B[t];
    }
    w.qslot[cnt] = s;
    w.qneg[cnt] = neg;
    w.qdbl[cnt] = dbl;
    S.busy[s] = 1;
    if (++cnt == BATCH) do_flush();
  };
  // slot s (not busy) += (neg ? -P : P); P's y words are P + 10
  auto place = [&](const uint64_t* P, uint32_t s, uint32_t neg) {
    uint64_t* B = S.at(s);
    if (S.empty[s]) {
      memcpy(B, P, 10 * sizeof(uint64_t));
      if (neg) {
        neg52(P + 10, B + 10);
        neg52(P + 15, B + 15);
      } else {
        memcpy(B + 10, P + 10, 10 * sizeof(uint64_t));
      }
      S.empty[s] = 0;
      return;
    }
    if (memcmp(B, P, 10 * sizeof(uint64_t)) == 0) {  // same x: P = B or P = -B
      uint64_t y[10];
      if (neg) {
        neg52(P + 10, y);
        neg52(P + 15, y + 5);
      } else {
        memcpy(y, P + 10, sizeof(y));
      }
      if (memcmp(B + 10, y, sizeof(y)) != 0) {
        S.empty[s] = 1;  // P = -B
      } else {
        uint64_t any = 0;
        for (int t = 0; t < 10; t++) any |= y[t];
        if (!any) S.empty[s] = 1;  // y = 0: 2B is the point at infinity
        else enqueue(B, s, 0, 1);  // tangent step on B itself
      }
      return;
    }
    enqueue(P, s, (uint8_t)neg, 0);
  };
  // adds every replica back into its bucket; the queue is empty on entry
  // and on exit, and the replica pool is free again
  auto merge_all = [&]() {
    bool more = S.n_chained > 0;
    while (more) {
      more = false;
      for (uint32_t k = 0; k < S.n_chained; k++) {
        const uint32_t b = S.chained[k];
        for (int32_t s = (int32_t)b; s >= 0 && S.next[s] >= 0; s = S.next[s]) {
          const uint32_t t = (uint32_t)S.next[s];
          if (!S.empty[t]) place(S.at(t), (uint32_t)s, 0);  // P is copied when queued
          S.next[s] = S.next[t];
          if (S.next[s] >= 0) more = true;
        }
      }
      do_flush();
    }
    for (uint32_t k = 0; k < S.n_chained; k++) {
      const uint32_t b = S.chained[k];
      S.next[b] = -1;
      S.tail[b] = S.cursor[b] = b;
      S.is_chained[b] = 0;
    }
    S.n_chained = 0;
    memset(S.empty + nb, 1, REPLICAS);
    for (uint32_t r = 0; r < S.n_rep; r++) S.next[nb + r] = -1;
    S.n_rep = 0;
  };
  auto push = [&](const uint64_t* P, uint32_t b, uint32_t neg) {
    int32_t s = (int32_t)S.cursor[b];
    while (s >= 0 && S.busy[s]) s = S.next[s];
    if (s < 0) {
      if (S.n_rep == REPLICAS) {  // pool used up: fold every replica back, then the bucket is free
        do_flush();
        merge_all();
        s = (int32_t)b;
      } else {
        s = (int32_t)(nb + S.n_rep++);
        S.next[S.tail[b]] = s;
        S.tail[b] = (uint32_t)s;
        if (!S.is_chained[b]) {
          S.is_chained[b] = 1;
          S.chained[S.n_chained++] = b;
        }
      }
    }
    if ((uint32_t)s != b && !is_touched[b]) {
      is_touched[b] = 1;
      touched[n_touched++] = b;
    }
    S.cursor[b] = (uint32_t)s;
    place(P, (uint32_t)s, neg);
  };

  const uint32_t off = win * cbits;
  for (uint64_t i = i0; i < i1; i++) {
    if (inf[i]) continue;
    const int32_t d = digit(scalars + 4 * i, K, off, cbits);
    if (d == 0) continue;
    push(pts20 + (uint64_t)PT_WORDS * i, (uint32_t)(d < 0 ? -d : d) - 1, d < 0);
  }
  do_flush();
  merge_all();
  memcpy(empty, S.empty, nb);
  release();
  return 0;
}

namespace {

// ---- the weighted bucket sum: Fq2 and XYZZ points on eight lanes ----
// Fq2 components stay below 2q in and out of every operation.
struct V2 {
  F c0, c1;
};
struct P4 {  // XYZZ: x = X / ZZ, y = Y / ZZZ; infinity is tracked in lane masks
  V2 X, Y, ZZ, ZZZ;
};

static inline void add2(V2& r, const V2& a, const V2& b, const VC& c) {
  F r0, r1;
  add(r0, a.c0, b.c0, c);
  csubk<2>(r0, c);
  add(r1, a.c1, b.c1, c);
  csubk<2>(r1, c);
  r.c0 = r0;
  r.c1 = r1;
}
static inline void sub2(V2& r, const V2& a, const V2& b, const VC& c) {
  F r0, r1;
  subk<2>(r0, a.c0, b.c0, c);
  csubk<2>(r0, c);
  subk<2>(r1, a.c1, b.c1, c);
  csubk<2>(r1, c);
  r.c0 = r0;
  r.c1 = r1;
}
// Karatsuba: (a0 + a1)(b0 + b1) < 16q^2, so every product ends below 2q
static inline void mul2(V2& r, const V2& a, const V2& b, const VC& c) {
  F m0, m1, m2, sa, sb, t, r0, r1;
  mul(m0, a.c0, b.c0, c);
  mul(m1, a.c1, b.c1, c);
  add(sa, a.c0, a.c1, c);
  add(sb, b.c0, b.c1, c);
  mul(m2, sa, sb, c);
  subk<2>(r0, m0, m1, c);
  csubk<2>(r0, c);
  add(t, m0, m1, c);
  subk<4>(r1, m2, t, c);
  csubk<4>(r1, c);
  csubk<2>(r1, c);
  r.c0 = r0;
  r.c1 = r1;
}
// (a0 + a1)(a0 - a1 + 2q) and 2 a0 a1
static inline void sqr2(V2& r, const V2& a, const VC& c) {
  F s, d, t, r0, r1;
  add(s, a.c0, a.c1, c);
  subk<2>(d, a.c0, a.c1, c);
  mul(r0, s, d, c);
  add(t, a.c0, a.c0, c);
  mul(r1, t, a.c1, c);
  r.c0 = r0;
  r.c1 = r1;
}
static inline __mmask8 zero_lt2q(const F& a, const VC& c) {
  F t = a;
  csubk<1>(t, c);
  const V o = _mm512_or_si512(_mm512_or_si512(_mm512_or_si512(t.l[0], t.l[1]), _mm512_or_si512(t.l[2], t.l[3])),
                              t.l[4]);
  return _mm512_cmpeq_epi64_mask(o, c.zero);
}
static inline __mmask8 zero2(const V2& a, const VC& c) { return zero_lt2q(a.c0, c) & zero_lt2q(a.c1, c); }
static inline void blend2(V2& a, const V2& b, __mmask8 m) {
  blend(a.c0, b.c0, m);
  blend(a.c1, b.c1, m);
}
static inline void blendP(P4& a, const P4& b, __mmask8 m) {
  blend2(a.X, b.X, m);
  blend2(a.Y, b.Y, m);
  blend2(a.ZZ, b.ZZ, m);
  blend2(a.ZZZ, b.ZZZ, m);
}

// r = 2 (x, y) for an affine point (mdbl-2008-s-1, a = 0); lanes with y = 0
// are returned in yz (their double is infinity)
static void aff_dbl(P4& r, const V2& x, const V2& y, __mmask8& yz, const VC& c) {
  V2 U, V, W, S, x2, M, t, X3, Y3;
  yz = zero2(y, c);
  add2(U, y, y, c);
  sqr2(V, U, c);
  mul2(W, U, V, c);
  mul2(S, x, V, c);
  sqr2(x2, x, c);
  add2(M, x2, x2, c);
  add2(M, M, x2, c);
  sqr2(X3, M, c);
  add2(t, S, S, c);
  sub2(X3, X3, t, c);
  sub2(t, S, X3, c);
  mul2(Y3, M, t, c);
  mul2(t, W, y, c);
  sub2(Y3, Y3, t, c);
  r.X = X3;
  r.Y = Y3;
  r.ZZ = V;
  r.ZZZ = W;
}
// r = 2p (dbl-2008-s-1, a = 0); lanes with y = 0 in yz
static void xyzz_dbl(P4& r, const P4& p, __mmask8& yz, const VC& c) {
  V2 U, V, W, S, x2, M, t, X3, Y3, ZZ3, ZZZ3;
  yz = zero2(p.Y, c);
  add2(U, p.Y, p.Y, c);
  sqr2(V, U, c);
  mul2(W, U, V, c);
  mul2(S, p.X, V, c);
  sqr2(x2, p.X, c);
  add2(M, x2, x2, c);
  add2(M, M, x2, c);
  sqr2(X3, M, c);
  add2(t, S, S, c);
  sub2(X3, X3, t, c);
  sub2(t, S, X3, c);
  mul2(Y3, M, t, c);
  mul2(t, W, p.Y, c);
  sub2(Y3, Y3, t, c);
  mul2(ZZ3, V, p.ZZ, c);
  mul2(ZZZ3, W, p.ZZZ, c);
  r.X = X3;
  r.Y = Y3;
  r.ZZ = ZZ3;
  r.ZZZ = ZZZ3;
}

// p += (x, y) on the lanes of m (affine points, not infinity)
// (madd-2008-s); pinf: p's infinity lanes, updated
static void xyzz_madd(P4& p, __mmask8& pinf, const V2& x, const V2& y, __mmask8 m, const VC& c) {
  const __mmask8 was_inf = pinf;
  const __mmask8 fin = m & (__mmask8)~was_inf;
  if (fin) {
    V2 U2, S2, Pd, R, PP, PPP, Q, t, X3, Y3;
    P4 r;
    mul2(U2, x, p.ZZ, c);
    mul2(S2, y, p.ZZZ, c);
    sub2(Pd, U2, p.X, c);
    sub2(R, S2, p.Y, c);
    sqr2(PP, Pd, c);
    mul2(PPP, Pd, PP, c);
    mul2(Q, p.X, PP, c);
    sqr2(X3, R, c);
    sub2(X3, X3, PPP, c);
    add2(t, Q, Q, c);
    sub2(X3, X3, t, c);
    sub2(t, Q, X3, c);
    mul2(Y3, R, t, c);
    mul2(t, p.Y, PPP, c);
    sub2(Y3, Y3, t, c);
    mul2(r.ZZ, p.ZZ, PP, c);
    mul2(r.ZZZ, p.ZZZ, PPP, c);
    r.X = X3;
    r.Y = Y3;
    __mmask8 ok = fin;
    const __mmask8 pz = fin & zero2(Pd, c);
    if (pz) {  // same x: p = (x, y) or p = -(x, y)
      const __mmask8 dm = pz & zero2(R, c);
      __mmask8 cancel = pz & (__mmask8)~dm;
      if (dm) {
        P4 d;
        __mmask8 yz;
        aff_dbl(d, x, y, yz, c);
        blendP(r, d, dm);
        cancel |= dm & yz;
      }
      pinf |= cancel;
      ok &= (__mmask8)~cancel;
    }
    blendP(p, r, ok);
  }
  const __mmask8 fi = m & was_inf;
  if (fi) {
    P4 a;
    a.X = x;
    a.Y = y;
    a.ZZ.c0 = c.one;
    a.ZZ.c1 = F{{c.zero, c.zero, c.zero, c.zero, c.zero}};
    a.ZZZ = a.ZZ;
    blendP(p, a, fi);
    pinf &= (__mmask8)~fi;
  }
}

// p += q (add-2008-s); pinf, qinf: infinity lanes
static void xyzz_add(P4& p, __mmask8& pinf, const P4& q, __mmask8 qinf, const VC& c) {
  const __mmask8 both = (__mmask8)(~pinf & ~qinf);
  const __mmask8 take_q = pinf & (__mmask8)~qinf;
  if (both) {
    V2 U1, U2, S1, S2, Pd, R, PP, PPP, Q, t, X3, Y3;
    P4 r;
    mul2(U1, p.X, q.ZZ, c);
    mul2(U2, q.X, p.ZZ, c);
    mul2(S1, p.Y, q.ZZZ, c);
    mul2(S2, q.Y, p.ZZZ, c);
    sub2(Pd, U2, U1, c);
    sub2(R, S2, S1, c);
    sqr2(PP, Pd, c);
    mul2(PPP, Pd, PP, c);
    mul2(Q, U1, PP, c);
    sqr2(X3, R, c);
    sub2(X3, X3, PPP, c);
    add2(t, Q, Q, c);
    sub2(X3, X3, t, c);
    sub2(t, Q, X3, c);
    mul2(Y3, R, t, c);
    mul2(t, S1, PPP, c);
    sub2(Y3, Y3, t, c);
    mul2(t, p.ZZ, q.ZZ, c);
    mul2(r.ZZ, t, PP, c);
    mul2(t, p.ZZZ, q.ZZZ, c);
    mul2(r.ZZZ, t, PPP, c);
    r.X = X3;
    r.Y = Y3;
    __mmask8 ok = both;
    const __mmask8 pz = both & zero2(Pd, c);
    if (pz) {  // same x: p = q or p = -q
      const __mmask8 dm = pz & zero2(R, c);
      __mmask8 cancel = pz & (__mmask8)~dm;
      if (dm) {
        P4 d;
        __mmask8 yz;
        xyzz_dbl(d, q, yz, c);
        blendP(r, d, dm);
        cancel |= dm & yz;
      }
      pinf |= cancel;
      ok &= (__mmask8)~cancel;
    }
    blendP(p, r, ok);
  }
  if (take_q) {
    blendP(p, q, take_q);
    pinf &= (__mmask8)~take_q;
  }
}

// lane values of one Fq component as four canonical words in R = 2^256 form
static inline void export_f(const F& a, uint64_t* base, int stride, const VC& c) {
  F t;
  mul(t, a, c.to256, c);
  csubk<1>(t, c);
  alignas(64) uint64_t l5[5][8];
  for (int j = 0; j < 5; j++) _mm512_store_si512(l5[j], t.l[j]);
  for (int lane = 0; lane < 8; lane++) {
    uint64_t* w = base + (size_t)stride * lane;
    w[0] = l5[0][lane] | (l5[1][lane] << 52);
    w[1] = (l5[1][lane] >> 12) | (l5[2][lane] << 40);
    w[2] = (l5[2][lane] >> 24) | (l5[3][lane] << 28);
    w[3] = (l5[3][lane] >> 36) | (l5[4][lane] << 16);
  }
}
// Jacobian (X ZZ^2, Y ZZZ^2, ZZZ) of each lane: 24 words per lane
static void export_jac(const P4& p, uint64_t* out, const VC& c) {
  V2 t, Xj, Yj;
  sqr2(t, p.ZZ, c);
  mul2(Xj, p.X, t, c);
  sqr2(t, p.ZZZ, c);
  mul2(Yj, p.Y, t, c);
  export_f(Xj.c0, out + 0, 24, c);
  export_f(Xj.c1, out + 4, 24, c);
  export_f(Yj.c0, out + 8, 24, c);
  export_f(Yj.c1, out + 12, 24, c);
  export_f(p.ZZZ.c0, out + 16, 24, c);
  export_f(p.ZZZ.c1, out + 20, 24, c);
}

}  // namespace

uint32_t reduce_segments(const uint64_t* bk20, const uint8_t* empty, uint32_t nb, uint64_t* out_sum,
                         uint64_t* out_run, uint8_t* sum_inf, uint8_t* run_inf) {
  VC c;
  make_vc(c);
  const uint32_t L = (nb + 7) / 8;
  P4 run, sum;
  const F z5 = {{c.zero, c.zero, c.zero, c.zero, c.zero}};
  run.X.c0 = run.X.c1 = z5;
  run.Y = run.ZZ = run.ZZZ = run.X;
  sum = run;
  __mmask8 rinf = 0xff, sinf = 0xff;
  alignas(64) uint64_t tb[PT_WORDS][8];
  for (int64_t j = (int64_t)L - 1; j >= 0; j--) {
    __mmask8 present = 0;
    for (int l = 0; l < 8; l++) {
      const uint64_t b = (uint64_t)l * L + (uint64_t)j;
      if (b < nb && !empty[b]) {
        present |= (__mmask8)(1u << l);
        const uint64_t* B = bk20 + (uint64_t)PT_WORDS * b;
        for (int t = 0; t < PT_WORDS; t++) tb[t][l] = B[t];
      } else {
        for (int t = 0; t < PT_WORDS; t++) tb[t][l] = 0;
      }
    }
    if (present) {
      V2 x, y;
      for (int k = 0; k < 5; k++) {
        x.c0.l[k] = _mm512_load_si512(tb[k]);
        x.c1.l[k] = _mm512_load_si512(tb[5 + k]);
        y.c0.l[k] = _mm512_load_si512(tb[10 + k]);
        y.c1.l[k] = _mm512_load_si512(tb[15 + k]);
      }
      xyzz_madd(run, rinf, x, y, present, c);
    }
    xyzz_add(sum, sinf, run, rinf, c);
  }
  export_jac(sum, out_sum, c);
  export_jac(run, out_run, c);
  for (int l = 0; l < 8; l++) {
    sum_inf[l] = (sinf >> l) & 1;
    run_inf[l] = (rinf >> l) & 1;
  }
  return L;
}

}  // namespace g2simd
}  // namespace zk
