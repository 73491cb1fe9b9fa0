/*
 * C restatement of the reference's CPU NTT and CPU Pippenger MSM — TEST
 * INFRASTRUCTURE (large-size parity and the timed CPU baseline).  Not part
 * of the product library.
 *
 *   oracle_ntt          crypto/poly/ntt.py:42-146: bit-reversal permutation
 *                       then iterative radix-2 DIT with one twiddle per
 *                       stage (wlen = omega^(n/len)); the inverse runs the
 *                       forward transform with omega^-1 and scales by n^-1;
 *                       coset forms scale by shift^i before / shift^-i after
 *                       (shift 0: only index 0 survives, crypto/poly/ntt.py:139-145).
 *   oracle_msm_pippenger crypto/msm/pippenger.py:10-64: unsigned window_bits
 *                       windows over bitlen(r) = 254 bits, buckets 1..2^w-1,
 *                       running-sum window reduction, Horner combination.
 *   oracle_g1_mul        crypto/ecc/bn254.py:23-24 (double-and-add).
 *
 * Arithmetic: Fr / Fq in Montgomery form (R = 2^256) on 4 x 64-bit limbs,
 * product by full 8-limb schoolbook multiplication followed by a separate
 * word-by-word Montgomery reduction (SOS).  I/O is canonical little-endian
 * 4 x u64 words.  Points: affine (x, y) with x = y = 0 for the identity,
 * Jacobian internally.  OpenMP parallelises NTT stages and MSM windows when
 * threads > 1.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;
typedef struct { uint64_t v[4]; } fe;
typedef struct { uint64_t p[4]; uint64_t ninv; fe r2; fe one; } mod_t;

static mod_t MR, MQ;

static int geq(const uint64_t* a, const uint64_t* b) {
  for (int i = 3; i >= 0; i--) {
    if (a[i] > b[i]) return 1;
    if (a[i] < b[i]) return 0;
  }
  return 1;
}
static void sub4(uint64_t* r, const uint64_t* a, const uint64_t* b) {
  uint64_t br = 0;
  for (int i = 0; i < 4; i++) {
    uint64_t t = a[i] - b[i];
    uint64_t b1 = a[i] < b[i];
    uint64_t t2 = t - br;
    uint64_t b2 = t < br;
    r[i] = t2;
    br = b1 | b2;
  }
}
static fe fadd(const mod_t* m, fe a, fe b) {
  fe r;
  uint64_t c = 0;
  for (int i = 0; i < 4; i++) {
    u128 s = (u128)a.v[i] + b.v[i] + c;
    r.v[i] = (uint64_t)s;
    c = (uint64_t)(s >> 64);
  }
  if (c || geq(r.v, m->p)) sub4(r.v, r.v, m->p);
  return r;
}
static fe fsub(const mod_t* m, fe a, fe b) {
  fe r;
  if (geq(a.v, b.v)) {
    sub4(r.v, a.v, b.v);
  } else {
    fe t;
    sub4(t.v, m->p, b.v);
    r = fadd(m, a, t);
  }
  return r;
}
static int fzero(fe a) { return (a.v[0] | a.v[1] | a.v[2] | a.v[3]) == 0; }
static int feq(fe a, fe b) { return memcmp(&a, &b, sizeof(fe)) == 0; }

/* SOS: t = a * b (8 words), then 4 reduction rounds */
static fe fmul(const mod_t* m, fe a, fe b) {
  uint64_t t[9] = {0};
  for (int i = 0; i < 4; i++) {
    uint64_t c = 0;
    for (int j = 0; j < 4; j++) {
      u128 s = (u128)a.v[i] * b.v[j] + t[i + j] + c;
      t[i + j] = (uint64_t)s;
      c = (uint64_t)(s >> 64);
    }
    t[i + 4] = c;
  }
  for (int i = 0; i < 4; i++) {
    uint64_t u = t[i] * m->ninv;
    uint64_t c = 0;
    for (int j = 0; j < 4; j++) {
      u128 s = (u128)u * m->p[j] + t[i + j] + c;
      t[i + j] = (uint64_t)s;
      c = (uint64_t)(s >> 64);
    }
    for (int k = i + 4; k < 9 && c; k++) {
      u128 s = (u128)t[k] + c;
      t[k] = (uint64_t)s;
      c = (uint64_t)(s >> 64);
    }
  }
  fe r;
  memcpy(r.v, t + 4, 32);
  if (t[8] || geq(r.v, m->p)) sub4(r.v, r.v, m->p);
  return r;
}
static fe to_m(const mod_t* m, fe a) {
  while (geq(a.v, m->p)) sub4(a.v, a.v, m->p);
  return fmul(m, a, m->r2);
}
static fe from_m(const mod_t* m, fe a) {
  fe one = {{1, 0, 0, 0}};
  return fmul(m, a, one);
}
static fe fpow(const mod_t* m, fe a, const uint64_t* e) {
  fe r = m->one;
  for (int i = 3; i >= 0; i--)
    for (int b = 63; b >= 0; b--) {
      r = fmul(m, r, r);
      if ((e[i] >> b) & 1) r = fmul(m, r, a);
    }
  return r;
}
static fe finv(const mod_t* m, fe a) {
  uint64_t e[4];
  memcpy(e, m->p, 32);
  e[0] -= 2;
  return fpow(m, a, e);
}

/* R^2 mod p by doubling R mod p 256 more times; -p^-1 mod 2^64 by Newton */
static void init_mod(mod_t* m, const uint64_t p[4]) {
  memcpy(m->p, p, 32);
  uint64_t inv = 1;
  for (int i = 0; i < 7; i++) inv *= 2 - p[0] * inv;
  m->ninv = (uint64_t)0 - inv;
  fe x = {{1, 0, 0, 0}};
  for (int i = 0; i < 512; i++) {
    /* x = 2x mod p */
    uint64_t c = 0;
    for (int k = 0; k < 4; k++) {
      uint64_t nc = x.v[k] >> 63;
      x.v[k] = (x.v[k] << 1) | c;
      c = nc;
    }
    if (c || geq(x.v, p)) sub4(x.v, x.v, p);
    if (i == 255) m->one = x;
  }
  m->r2 = x;
}

static int g_init = 0;
static void init_all(void) {
  if (g_init) return;
  const uint64_t r[4] = {0x43e1f593f0000001ull, 0x2833e84879b97091ull, 0xb85045b68181585dull,
                         0x30644e72e131a029ull};
  const uint64_t q[4] = {0x3c208c16d87cfd47ull, 0x97816a916871ca8dull, 0xb85045b68181585dull,
                         0x30644e72e131a029ull};
  init_mod(&MR, r);
  init_mod(&MQ, q);
  g_init = 1;
}

static fe ld(const uint64_t* w) { fe r; memcpy(r.v, w, 32); return r; }
static void st(uint64_t* w, fe a) { memcpy(w, a.v, 32); }

/* ------------------------------------------------------------------ NTT */
int oracle_ntt(const uint64_t* in, uint64_t in_len, uint64_t* out, uint32_t log_n,
               const uint64_t omega[4], int inverse, const uint64_t* shift, int threads) {
  init_all();
  const uint64_t n = 1ull << log_n;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  fe* a = (fe*)malloc(sizeof(fe) * n);
  if (!a) return 1;
  fe w = to_m(&MR, ld(omega));
  fe s = shift ? to_m(&MR, ld(shift)) : MR.one;
  if (inverse) w = finv(&MR, w);
  /* load (zero-pad / truncate for the forward form) */
  uint64_t take = inverse ? n : (in_len < n ? in_len : n);
  for (uint64_t i = 0; i < n; i++) a[i] = (i < take) ? to_m(&MR, ld(in + 4 * i)) : (fe){{0, 0, 0, 0}};
  if (!inverse && shift) {
    fe pw = MR.one;
    for (uint64_t i = 0; i < take; i++) {
      a[i] = fmul(&MR, a[i], pw);
      pw = fmul(&MR, pw, s);
    }
  }
  /* bit reversal */
  for (uint64_t i = 1, j = 0; i < n; i++) {
    uint64_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) { fe t = a[i]; a[i] = a[j]; a[j] = t; }
  }
  /* powers of w for all stages: stage len uses wlen = w^(n/len), i.e.
     tw[k * (n/len)] for k < len/2 */
  fe* tw = (fe*)malloc(sizeof(fe) * (n / 2 + 1));
  tw[0] = MR.one;
  for (uint64_t k = 1; k < n / 2; k++) tw[k] = fmul(&MR, tw[k - 1], w);
  for (uint64_t len = 2; len <= n; len <<= 1) {
    const uint64_t half = len >> 1, step = n / len;
#pragma omp parallel for schedule(static) if (threads != 1)
    for (int64_t bi = 0; bi < (int64_t)(n / 2); bi++) {
      uint64_t blk = (uint64_t)bi / half, k = (uint64_t)bi % half;
      uint64_t i = blk * len;
      fe u = a[i + k];
      fe v = fmul(&MR, a[i + k + half], tw[k * step]);
      a[i + k] = fadd(&MR, u, v);
      a[i + k + half] = fsub(&MR, u, v);
    }
  }
  free(tw);
  if (inverse) {
    fe ninv = finv(&MR, to_m(&MR, (fe){{n, 0, 0, 0}}));
    fe si = (shift && !fzero(s)) ? finv(&MR, s) : (fe){{0, 0, 0, 0}};
    fe pw = MR.one;
    for (uint64_t i = 0; i < n; i++) {
      a[i] = fmul(&MR, a[i], ninv);
      if (shift) {
        a[i] = fmul(&MR, a[i], pw);
        pw = fmul(&MR, pw, si);
      }
    }
  }
  for (uint64_t i = 0; i < n; i++) st(out + 4 * i, from_m(&MR, a[i]));
  free(a);
  return 0;
}

/* ------------------------------------------------------------------ G1 */
typedef struct { fe X, Y, Z; } jac;

static jac jinf(void) { jac r; r.X = MQ.one; r.Y = MQ.one; memset(&r.Z, 0, sizeof(fe)); return r; }
static int jis_inf(const jac* p) { return fzero(p->Z); }

static jac jdbl(jac p) {
  if (jis_inf(&p) || fzero(p.Y)) return jinf();
  fe A = fmul(&MQ, p.X, p.X), B = fmul(&MQ, p.Y, p.Y), C = fmul(&MQ, B, B);
  fe xb = fadd(&MQ, p.X, B);
  fe D = fsub(&MQ, fsub(&MQ, fmul(&MQ, xb, xb), A), C);
  D = fadd(&MQ, D, D);
  fe E = fadd(&MQ, fadd(&MQ, A, A), A);
  fe F = fmul(&MQ, E, E);
  jac r;
  r.X = fsub(&MQ, F, fadd(&MQ, D, D));
  fe c8 = fadd(&MQ, C, C); c8 = fadd(&MQ, c8, c8); c8 = fadd(&MQ, c8, c8);
  r.Y = fsub(&MQ, fmul(&MQ, E, fsub(&MQ, D, r.X)), c8);
  fe yz = fmul(&MQ, p.Y, p.Z);
  r.Z = fadd(&MQ, yz, yz);
  return r;
}

static jac jadd(jac p, jac q) {
  if (jis_inf(&p)) return q;
  if (jis_inf(&q)) return p;
  fe z1z1 = fmul(&MQ, p.Z, p.Z), z2z2 = fmul(&MQ, q.Z, q.Z);
  fe u1 = fmul(&MQ, p.X, z2z2), u2 = fmul(&MQ, q.X, z1z1);
  fe s1 = fmul(&MQ, p.Y, fmul(&MQ, q.Z, z2z2)), s2 = fmul(&MQ, q.Y, fmul(&MQ, p.Z, z1z1));
  fe h = fsub(&MQ, u2, u1), rr = fsub(&MQ, s2, s1);
  if (fzero(h)) {
    if (fzero(rr)) return jdbl(p);
    return jinf();
  }
  fe hh = fmul(&MQ, h, h), hhh = fmul(&MQ, h, hh), v = fmul(&MQ, u1, hh);
  jac r;
  r.X = fsub(&MQ, fsub(&MQ, fmul(&MQ, rr, rr), hhh), fadd(&MQ, v, v));
  r.Y = fsub(&MQ, fmul(&MQ, rr, fsub(&MQ, v, r.X)), fmul(&MQ, s1, hhh));
  r.Z = fmul(&MQ, fmul(&MQ, p.Z, q.Z), h);
  return r;
}

static jac jfrom_aff(const uint64_t* xy) {
  fe x = ld(xy), y = ld(xy + 4);
  if (fzero(x) && fzero(y)) return jinf();
  jac r;
  r.X = to_m(&MQ, x);
  r.Y = to_m(&MQ, y);
  r.Z = MQ.one;
  return r;
}

static void jto_aff(jac p, uint64_t* out_xy, int* is_inf) {
  if (jis_inf(&p)) {
    memset(out_xy, 0, 64);
    *is_inf = 1;
    return;
  }
  fe zi = finv(&MQ, p.Z), zi2 = fmul(&MQ, zi, zi), zi3 = fmul(&MQ, zi2, zi);
  st(out_xy, from_m(&MQ, fmul(&MQ, p.X, zi2)));
  st(out_xy + 4, from_m(&MQ, fmul(&MQ, p.Y, zi3)));
  *is_inf = 0;
}

static jac jmul_u64(jac p, uint64_t k) {
  jac acc = jinf();
  for (int b = 63; b >= 0; b--) {
    acc = jdbl(acc);
    if ((k >> b) & 1) acc = jadd(acc, p);
  }
  return acc;
}

int oracle_g1_mul(const uint64_t* xy, const uint64_t* k, uint64_t* out_xy, int* is_inf) {
  init_all();
  jac p = jfrom_aff(xy), acc = jinf();
  for (int i = 3; i >= 0; i--)
    for (int b = 63; b >= 0; b--) {
      acc = jdbl(acc);
      if ((k[i] >> b) & 1) acc = jadd(acc, p);
    }
  jto_aff(acc, out_xy, is_inf);
  return 0;
}

static uint32_t digit_at(const uint64_t* s, int k, int w) {
  /* (s >> k) & (2^w - 1) for a 256-bit s */
  int wi = k >> 6, sh = k & 63;
  u128 v = s[wi] >> sh;
  if (sh && wi + 1 < 4) v |= (u128)s[wi + 1] << (64 - sh);
  else if (!sh && wi + 1 < 4) v |= (u128)s[wi + 1] << 64;
  return (uint32_t)(v & ((1u << w) - 1));
}

int oracle_msm_pippenger(const uint64_t* bases, const uint64_t* scalars, uint64_t n, int window_bits,
                         uint64_t* out_xy, int* is_inf, int threads) {
  init_all();
  if (window_bits <= 0 || window_bits > 20) return 2;
  const int w = window_bits, max_bits = 254;
  const int nwin = (max_bits + w - 1) / w;
  const uint32_t nb = 1u << w;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  jac* pts = (jac*)malloc(sizeof(jac) * (n ? n : 1));
  uint64_t* ss = (uint64_t*)malloc(32 * (n ? n : 1));
  for (uint64_t i = 0; i < n; i++) {
    pts[i] = jfrom_aff(bases + 8 * i);
    /* scalars reduced mod r (pippenger.py:28) */
    fe s = ld(scalars + 4 * i);
    while (geq(s.v, MR.p)) sub4(s.v, s.v, MR.p);
    memcpy(ss + 4 * i, s.v, 32);
  }
  jac* wsum = (jac*)malloc(sizeof(jac) * nwin);
#pragma omp parallel for schedule(dynamic, 1) if (threads != 1)
  for (int wi = 0; wi < nwin; wi++) {
    const int k = wi * w;
    jac* buckets = (jac*)malloc(sizeof(jac) * nb);
    for (uint32_t d = 0; d < nb; d++) buckets[d] = jinf();
    for (uint64_t i = 0; i < n; i++) {
      uint32_t d = digit_at(ss + 4 * i, k, w);
      if (d) buckets[d] = jadd(buckets[d], pts[i]);
    }
    jac running = jinf(), sum = jinf();
    for (uint32_t d = nb - 1; d > 0; d--) {
      running = jadd(running, buckets[d]);
      sum = jadd(sum, running);
    }
    wsum[wi] = sum;
    free(buckets);
  }
  jac acc = jinf();
  for (int wi = nwin - 1; wi >= 0; wi--) {
    for (int d = 0; d < w; d++) acc = jdbl(acc);
    acc = jadd(acc, wsum[wi]);
  }
  jto_aff(acc, out_xy, is_inf);
  free(wsum);
  free(pts);
  free(ss);
  return 0;
}

int oracle_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* f(x) = sum_j coeffs[j] x^j by Horner (canonical words in / out): the
 * definition an NTT output row must equal, evals[i] = f(omega^i). */
int oracle_poly_eval(const uint64_t* coeffs, uint64_t n, const uint64_t* x, uint64_t* out) {
  init_all();
  fe xm = to_m(&MR, ld(x));
  fe acc = {{0, 0, 0, 0}};
  for (uint64_t j = n; j-- > 0;) acc = fadd(&MR, fmul(&MR, acc, xm), to_m(&MR, ld(coeffs + 4 * j)));
  st(out, from_m(&MR, acc));
  return 0;
}
