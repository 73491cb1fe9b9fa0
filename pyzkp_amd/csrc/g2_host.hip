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
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <exception>
#include <memory>
#include <mutex>
#include <new>
#include <utility>
#include <thread>
#include <vector>
#include "g2_simd.hpp"
#include "internal.hpp"
#include "../../include/zkhip.h"

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
// Lazy-reduction Fq2 products.  All F2 coefficients are canonical (< q).
// The three 256 x 256-bit products of the Karatsuba form are kept as 512-bit
// integers and combined exactly:
//   c1 = (a0 + a1)(b0 + b1) - a0 b0 - a1 b1 = a0 b1 + a1 b0   in [0, 2 q^2)
//   c0 = a0 b0 + q^2 - a1 b1                                  in (0, 2 q^2)
// and only these two are Montgomery-reduced.  Both are below q 2^256
// (2 q < 2^256), so a reduction ends below 2 q and one conditional
// subtraction makes it canonical: two reductions and no modular additions
// instead of three reductions and five modular additions.
using zkh::u128;
using zkh::u64;
static constexpr u64 QW[4] = {0x3c208c16d87cfd47ull, 0x97816a916871ca8dull, 0xb85045b68181585dull,
                              0x30644e72e131a029ull};
static constexpr u64 QINV = 0x87d20782e4866389ull;  // -q^-1 mod 2^64
static constexpr u64 QSQ[8] = {0x3b5458a2275d69b1ull, 0xa602072d09eac101ull, 0x4a50189c6d96cadcull,
                               0x04689e957a1242c8ull, 0x26edfa5c34c6b38dull, 0xb00b855116375606ull,
                               0x599a6f7c0348d21cull, 0x0925c4b8763cbf9cull};  // q^2

// r = a b, 8 words
static inline void mul_wide(const u64* a, const u64* b, u64* r) {
  u64 t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma GCC unroll 4
  for (int i = 0; i < 4; i++) {
    u64 c = 0;
#pragma GCC unroll 4
    for (int j = 0; j < 4; j++) {
      const u128 s = (u128)a[j] * b[i] + t[i + j] + c;
      t[i + j] = (u64)s;
      c = (u64)(s >> 64);
    }
    t[i + 4] = c;
  }
  memcpy(r, t, 64);
}
// a += b (8 words; no carry out by the callers' bounds)
static inline void add8(u64* a, const u64* b) {
  u64 c = 0;
  for (int i = 0; i < 8; i++) {
    const u128 s = (u128)a[i] + b[i] + c;
    a[i] = (u64)s;
    c = (u64)(s >> 64);
  }
}
// a -= b (8 words; a >= b by the callers' bounds)
static inline void sub8(u64* a, const u64* b) {
  u64 br = 0;
  for (int i = 0; i < 8; i++) {
    const u128 d = (u128)a[i] - b[i] - br;
    a[i] = (u64)d;
    br = (u64)(d >> 127);
  }
}
// T 2^-256 mod q for T < q 2^256 (8 words, clobbered), canonical
static inline Fe redc(u64* t) {
  u64 hi = 0;  // carry out of word i + 3 of the previous step, owed to word i + 4
#pragma GCC unroll 4
  for (int i = 0; i < 4; i++) {
    const u64 m = t[i] * QINV;
    u64 c = 0;
#pragma GCC unroll 4
    for (int j = 0; j < 4; j++) {
      const u128 s = (u128)m * QW[j] + t[i + j] + c;
      t[i + j] = (u64)s;
      c = (u64)(s >> 64);
    }
    const u128 s = (u128)t[i + 4] + c + hi;
    t[i + 4] = (u64)s;
    hi = (u64)(s >> 64);
  }
  // (T + M q) / 2^256 < 2 q < 2^256, so hi is 0 here
  Fe r = {{t[4], t[5], t[6], t[7]}};
  zkh::csub(r.v, QW);
  return r;
}
// a + b without reduction (< 2 q for canonical inputs)
static inline Fe add_nr(const Fe& a, const Fe& b) {
  Fe r;
  u64 c = 0;
  for (int i = 0; i < 4; i++) {
    const u128 s = (u128)a.v[i] + b.v[i] + c;
    r.v[i] = (u64)s;
    c = (u64)(s >> 64);
  }
  return r;
}
// (a0 + a1 u)(b0 + b1 u) = (a0 b0 - a1 b1) + (a0 b1 + a1 b0) u
static inline F2 f2_mul(const F2& a, const F2& b) {
  u64 t0[8], t1[8], t2[8];
  const Fe sa = add_nr(a.c0, a.c1), sb = add_nr(b.c0, b.c1);
  mul_wide(a.c0.v, b.c0.v, t0);
  mul_wide(a.c1.v, b.c1.v, t1);
  mul_wide(sa.v, sb.v, t2);
  sub8(t2, t0);
  sub8(t2, t1);
  add8(t0, QSQ);
  sub8(t0, t1);
  return {redc(t0), redc(t2)};
}
// a0^2 + a1^2 (below 2 q^2 before the one reduction)
static inline Fe f2_norm(const F2& a) {
  u64 t0[8], t1[8];
  mul_wide(a.c0.v, a.c0.v, t0);
  mul_wide(a.c1.v, a.c1.v, t1);
  add8(t0, t1);
  return redc(t0);
}
static inline F2 f2_sqr(const F2& a) {
  // (a0 + a1)(a0 - a1) + 2 a0 a1 u; the unreduced factors are below 2 q,
  // their products below 4 q^2 < q 2^256, which the Montgomery product
  // takes to a canonical result
  Fe dq;
  zkh::sub_raw(dq.v, QW, a.c1.v);
  return {zkh::mul(FQ, add_nr(a.c0, a.c1), add_nr(a.c0, dq)), zkh::mul(FQ, a.c0, add_nr(a.c1, a.c1))};
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

// Host worker threads shared by every parallel loop of this file, started
// once.  A loop posts a job; idle workers join the oldest job that still has
// items (up to the job's thread cap) and the calling thread works on its own
// job too, so a job always progresses even when every worker is busy
// elsewhere.  Concurrent MSMs (the prove server runs two) thus share the
// cores instead of starting 16 threads each, and a small MSM pays no thread
// start-up.
class HostPool {
 public:
  static HostPool& get() {
    static HostPool* p = new HostPool();  // never destroyed: workers may be parked at exit
    return *p;
  }
  // fn(0 .. n-1) on at most `threads` threads (the caller is one of them);
  // an exception thrown by fn skips the items not yet started and is
  // rethrown here
  void run(uint64_t n, int threads, const std::function<void(uint64_t)>& fn) {
    if (threads <= 1 || n <= 1 || workers_.empty()) {
      for (uint64_t i = 0; i < n; i++) fn(i);
      return;
    }
    Job j;
    j.fn = &fn;
    j.n = n;
    j.max_helpers = std::min<int>(threads - 1, (int)workers_.size());
    {
      std::lock_guard<std::mutex> g(mu_);
      jobs_.push_back(&j);
    }
    cv_.notify_all();
    work(&j);
    {
      std::unique_lock<std::mutex> lk(mu_);
      done_cv_.wait(lk, [&] { return j.done.load() == j.n && j.helpers == 0; });
      jobs_.erase(std::find(jobs_.begin(), jobs_.end(), &j));
    }
    if (j.err) std::rethrow_exception(j.err);
  }

 private:
  struct Job {
    const std::function<void(uint64_t)>* fn = nullptr;
    uint64_t n = 0;
    std::atomic<uint64_t> next{0}, done{0};
    std::atomic<bool> stop{false};
    int max_helpers = 0, helpers = 0;  // helpers: under mu_
    std::exception_ptr err;
    std::mutex err_mu;
  };

  HostPool() {
    int hw = (int)std::thread::hardware_concurrency();
    if (hw < 1) hw = 1;
    for (int t = 0; t < hw; t++) workers_.emplace_back([this] { loop(); });
  }
  static void work(Job* j) {
    for (;;) {
      const uint64_t i = j->next.fetch_add(1);
      if (i >= j->n) break;
      if (!j->stop.load()) {
        try {
          (*j->fn)(i);
        } catch (...) {
          std::lock_guard<std::mutex> g(j->err_mu);
          if (!j->err) j->err = std::current_exception();
          j->stop.store(true);
        }
      }
      j->done.fetch_add(1);
    }
  }
  void loop() {
    for (;;) {
      Job* j = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] {
          for (Job* c : jobs_)
            if (c->next.load() < c->n && c->helpers < c->max_helpers) {
              j = c;
              return true;
            }
          return false;
        });
        j->helpers++;
      }
      work(j);
      {
        std::lock_guard<std::mutex> g(mu_);
        j->helpers--;
      }
      done_cv_.notify_all();
    }
  }

  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  std::vector<Job*> jobs_;
  std::vector<std::thread> workers_;
};

template <class F>
static void parallel_for(uint64_t n, int threads, F fn) {
  const std::function<void(uint64_t)> f(std::ref(fn));
  HostPool::get().run(n, threads, f);
}

static uint32_t bits_at(const uint64_t* s, uint32_t off, uint32_t c) {
  const uint32_t wi = off >> 6, sh = off & 63;
  unsigned __int128 v = (unsigned __int128)s[wi] >> sh;
  if (wi + 1 < 4) v |= (unsigned __int128)s[wi + 1] << (64 - sh);
  return (uint32_t)(v & ((1ull << c) - 1));
}

// sum_i s_i Q_i, signed-digit Pippenger with batch-affine bucket sums.
//
// Digits: W = ceil(255 / c) signed c-bit digits per scalar (carry into the
// next window), so a window has 2^(c-1) buckets.  Work is split into tasks
// (window, range of points); every task owns its buckets, so tasks run on
// any thread without locks, and a task's weighted bucket sum is its
// contribution to the window (the weighted sum is linear in the buckets).
//
// Bucket sums stay affine: a point is added to its bucket with the affine
// chord rule, and up to BA_BATCH such additions (to distinct buckets) share
// one Fq2 inversion (Montgomery's simultaneous inversion): about 5 Fq2
// products + 1 squaring per addition instead of 7 + 4 for a Jacobian mixed
// addition.  A point whose bucket already has an addition pending in the
// current batch goes into a Jacobian overflow bucket instead (so heavy
// collisions cost Jacobian speed, never more); P + P and P + (-P) are
// handled when they are scheduled.
static constexpr int BA_BATCH = 256;

static inline bool f2_eq(const F2& a, const F2& b) { return zkh::eq(a.c0, b.c0) && zkh::eq(a.c1, b.c1); }
static inline F2 f2_one() { return {zkh::one(FQ), zkh::zero()}; }

struct BucketTask {
  uint32_t w;
  uint64_t i0, i1;
};

// affine doubling (y != 0 on the prime-order subgroup)
static A2 a_dbl(const A2& p) {
  F2 x2 = f2_sqr(p.x);
  F2 lam = f2_mul(f2_add(f2_dbl(x2), x2), f2_inv(f2_dbl(p.y)));
  A2 r;
  r.x = f2_sub(f2_sqr(lam), f2_dbl(p.x));
  r.y = f2_sub(f2_mul(lam, f2_sub(p.x, r.x)), p.y);
  r.inf = false;
  return r;
}

static J2 bucket_task(const A2* pts, const int32_t* digits, uint64_t n_all, const BucketTask& t, int c) {
  const uint32_t nb = 1u << (c - 1);
  std::vector<A2> bk(nb);
  for (auto& b : bk) b.inf = true;
  std::vector<J2> over;  // Jacobian overflow buckets, for points whose bucket is busy
  std::vector<uint8_t> busy(nb, 0);
  uint32_t bidx[BA_BATCH];
  A2 pp[BA_BATCH];
  F2 den[BA_BATCH];
  Fe nrm[BA_BATCH], pref[BA_BATCH];
  int cnt = 0;

  // 1 / d = conj(d) / N(d) with the norm N(d) = d0^2 + d1^2 in Fq (nonzero
  // for d != 0: -1 is not a square mod q), so the simultaneous inversion
  // runs its prefix chain over the norms in Fq, not over the d in Fq2
  auto flush = [&]() {
    if (cnt == 0) return;
    Fe acc = zkh::one(FQ);
    for (int e = 0; e < cnt; e++) {
      nrm[e] = f2_norm(den[e]);
      pref[e] = acc;
      acc = zkh::mul(FQ, acc, nrm[e]);
    }
    Fe inv = zkh::inv(FQ, acc);
    for (int e = cnt - 1; e >= 0; e--) {
      const Fe ni = zkh::mul(FQ, inv, pref[e]);
      inv = zkh::mul(FQ, inv, nrm[e]);
      const F2 ie = {zkh::mul(FQ, den[e].c0, ni), zkh::neg(FQ, zkh::mul(FQ, den[e].c1, ni))};
      A2& B = bk[bidx[e]];
      const F2 lam = f2_mul(f2_sub(pp[e].y, B.y), ie);
      const F2 x3 = f2_sub(f2_sub(f2_sqr(lam), B.x), pp[e].x);
      B.y = f2_sub(f2_mul(lam, f2_sub(B.x, x3)), B.y);
      B.x = x3;
      busy[bidx[e]] = 0;
    }
    cnt = 0;
  };

  for (uint64_t i = t.i0; i < t.i1; i++) {
    const int32_t d = digits[(uint64_t)t.w * n_all + i];
    if (d == 0 || pts[i].inf) continue;
    A2 P = pts[i];
    if (d < 0) P.y = f2_neg(P.y);
    const uint32_t b = (uint32_t)(d < 0 ? -d : d) - 1;
    A2& B = bk[b];
    if (B.inf) {  // an empty bucket never has an addition pending
      B = P;
      continue;
    }
    if (busy[b]) {  // an addition to this bucket is pending in the batch
      if (over.empty()) over.assign(nb, j_inf());
      over[b] = j_madd(over[b], P.x, P.y);
      continue;
    }
    if (f2_eq(B.x, P.x)) {
      if (f2_eq(B.y, P.y)) B = a_dbl(B);
      else B.inf = true;  // P = -B
      continue;
    }
    busy[b] = 1;
    bidx[cnt] = b;
    pp[cnt] = P;
    den[cnt] = f2_sub(P.x, B.x);
    if (++cnt == BA_BATCH) flush();
  }
  flush();
  // sum_b (b + 1) B_b by running sums from the top bucket
  J2 run = j_inf(), sum = j_inf();
  for (int b = (int)nb - 1; b >= 0; b--) {
    if (!bk[b].inf) run = j_madd(run, bk[b].x, bk[b].y);
    if (!over.empty()) run = j_add(run, over[b]);
    sum = j_add(sum, run);
  }
  return sum;
}

// The same bucket sums on the AVX-512 IFMA path: the batched affine
// additions run eight at a time (g2_simd.cpp), and the weighted bucket sum
// runs as above.
static J2 bucket_task_simd(const uint64_t* pts20, const uint8_t* pinf, const uint64_t* scalars, const uint64_t* K,
                           const BucketTask& t, int c) {
  const uint32_t nb = 1u << (c - 1);
  std::vector<uint64_t> bk((size_t)nb * g2simd::PT_WORDS);
  std::vector<uint8_t> empty(nb, 1);
  if (g2simd::accumulate(pts20, pinf, scalars, K, t.w, (uint32_t)c, t.i0, t.i1, nb, bk.data(), empty.data()))
    throw std::bad_alloc();
  // sum_b (b + 1) B_b: eight running-sum segments on the vector lanes, then
  // sum_l sum_l + L sum_l l run_l here
  uint64_t sw[8 * 24], rw[8 * 24];
  uint8_t si[8], ri[8];
  const uint32_t L = g2simd::reduce_segments(bk.data(), empty.data(), nb, sw, rw, si, ri);
  auto jac = [](const uint64_t* w) {
    J2 p;
    memcpy(p.X.c0.v, w, 32);
    memcpy(p.X.c1.v, w + 4, 32);
    memcpy(p.Y.c0.v, w + 8, 32);
    memcpy(p.Y.c1.v, w + 12, 32);
    memcpy(p.Z.c0.v, w + 16, 32);
    memcpy(p.Z.c1.v, w + 20, 32);
    return p;
  };
  J2 sum = j_inf(), run = j_inf(), u = j_inf();
  for (int l = 7; l >= 0; l--) {
    if (!si[l]) sum = j_add(sum, jac(sw + 24 * l));
    if (l == 0) break;
    if (!ri[l]) run = j_add(run, jac(rw + 24 * l));
    u = j_add(u, run);  // u = sum_(l >= 1) l run_l
  }
  J2 m = j_inf();
  for (int bit = 31; bit >= 0; bit--) {
    m = j_dbl(m);
    if ((L >> bit) & 1) m = j_add(m, u);
  }
  return j_add(sum, m);
}

// 1: the AVX-512 IFMA path where the CPU has it; 0: the scalar path
static std::atomic<int> g_simd_mode(1);
static bool cpu_ifma() {
  static const bool has = g2simd::cpu_has_ifma();
  return has;
}
static bool simd_on() { return cpu_ifma() && g_simd_mode.load() != 0; }

// bases in the IFMA path's form, converted in parallel (each worker first
// touches the pages it writes)
static void prepare(const uint64_t* bases, uint64_t n, int threads, uint64_t* pts20, uint8_t* inf) {
  parallel_for((n + 1023) / 1024, threads, [&](uint64_t blk) {
    const uint64_t i0 = blk * 1024, e = i0 + 1024 < n ? i0 + 1024 : n;
    g2simd::prepare_points(bases + 16 * i0, e - i0, pts20 + (uint64_t)g2simd::PT_WORDS * i0, inf + i0);
  });
}

// sum_i s_i Q_i.  pts20 / pinf: the bases in the IFMA form (then that path
// runs; its digits come from the scalars on the fly), or null.
static void msm(const uint64_t* bases, const uint64_t* pts20, const uint8_t* pinf, const uint64_t* scalars,
                uint64_t n, int c, int threads, uint64_t out[16]) {
  const uint32_t W = (255 + c - 1) / c;
  const bool simd = pts20 != nullptr;
  std::vector<A2> pts(simd ? 0 : n);
  std::vector<int32_t> digits(simd ? 0 : (size_t)n * W);
  // K = sum_w 2^(w c + c - 1): window w of (s mod r) + K, minus 2^(c-1), is
  // the signed digit that the carry recoding below gives (the carry out of
  // the lower windows is the carry of the addition); W c >= 255 keeps
  // r + K below 2^(W c)
  uint64_t K[5] = {0, 0, 0, 0, 0};
  for (uint32_t w = 0; w < W; w++) {
    const uint32_t bit = w * (uint32_t)c + (uint32_t)c - 1;
    K[bit >> 6] |= 1ull << (bit & 63);
  }
  if (!simd) {
    parallel_for((n + 1023) / 1024, threads, [&](uint64_t blk) {
      const uint64_t e = (blk + 1) * 1024 < n ? (blk + 1) * 1024 : n;
      for (uint64_t i = blk * 1024; i < e; i++) {
        pts[i] = rd_aff(bases + 16 * i);
        Fe s;
        memcpy(s.v, scalars + 4 * i, 32);
        while (zkh::geq(s.v, zkh::FR.p)) zkh::sub_raw(s.v, s.v, zkh::FR.p);
        // signed digits: v = window + carry; v >= 2^(c-1) -> v - 2^c, carry 1
        const int64_t half = 1ll << (c - 1), full = 1ll << c;
        int64_t carry = 0;
        for (uint32_t w = 0; w < W; w++) {
          const uint32_t off = w * (uint32_t)c, wi = off >> 6, sh = off & 63;
          unsigned __int128 v = wi < 4 ? ((unsigned __int128)s.v[wi] >> sh) : 0;
          if (wi + 1 < 4 && sh) v |= (unsigned __int128)s.v[wi + 1] << (64 - sh);
          int64_t dv = (int64_t)(uint64_t)(v & (uint64_t)(full - 1)) + carry;
          if (dv >= half) {
            dv -= full;
            carry = 1;
          } else {
            carry = 0;
          }
          digits[(uint64_t)w * n + i] = (int32_t)dv;
        }
      }
    });
  }
  // tasks: pieces (window, point range), each with its own buckets
  std::vector<BucketTask> tasks;
  std::vector<std::vector<uint32_t>> chunks;  // the pieces a worker takes together (IFMA path)
  if (simd) {
    // The W n (window, point) pairs are cut into one equal chunk per
    // thread; a chunk may hold the end of one window and the start of the
    // next.  Cuts fall on a grid of n / 32 points, so no piece is a sliver
    // that pays a whole bucket sum for a few points.  (Equal pieces per
    // window balance only when their count fills whole rounds of threads.)
    const uint64_t g = n >= 32 ? n / 32 : 1;
    auto cut = [&](uint64_t k) {  // chunk boundary k as (window, point)
      const unsigned __int128 u = (unsigned __int128)W * n * k / (uint64_t)threads;
      uint64_t w = (uint64_t)(u / n), i = (uint64_t)(u % n);
      i = (i + g / 2) / g * g;
      if (i >= n) {
        i = 0;
        w++;
      }
      return std::make_pair(w, i);
    };
    for (int k = 0; k < threads; k++) {
      auto a = cut((uint64_t)k), b = cut((uint64_t)k + 1);
      std::vector<uint32_t> ids;
      for (uint64_t w = a.first; w <= b.first && w < W; w++) {
        const uint64_t i0 = (w == a.first) ? a.second : 0;
        const uint64_t i1 = (w == b.first) ? b.second : n;
        if (i1 > i0) {
          ids.push_back((uint32_t)tasks.size());
          tasks.push_back({(uint32_t)w, i0, i1});
        }
      }
      if (!ids.empty()) chunks.push_back(ids);
    }
  } else {
    const uint64_t min_pts = 1ull << 12;
    uint64_t S = ((uint64_t)threads * 2 + W - 1) / W;
    if (S > (n + min_pts - 1) / min_pts) S = (n + min_pts - 1) / min_pts;
    if (S < 1) S = 1;
    for (uint32_t w = 0; w < W; w++)
      for (uint64_t k = 0; k < S; k++) tasks.push_back({w, n * k / S, n * (k + 1) / S});
  }
  std::vector<J2> part(tasks.size());
  if (simd) {
    parallel_for(chunks.size(), threads, [&](uint64_t k) {
      for (uint32_t t : chunks[k]) part[t] = bucket_task_simd(pts20, pinf, scalars, K, tasks[t], c);
    });
  } else {
    parallel_for(tasks.size(), threads,
                 [&](uint64_t t) { part[t] = bucket_task(pts.data(), digits.data(), n, tasks[t], c); });
  }
  std::vector<J2> win(W, j_inf());
  for (size_t t = 0; t < tasks.size(); t++) win[tasks[t].w] = j_add(win[tasks[t].w], part[t]);
  J2 acc = j_inf();
  for (int w = (int)W - 1; w >= 0; w--) {
    for (int k = 0; k < c; k++) acc = j_dbl(acc);
    acc = j_add(acc, win[w]);
  }
  wr_jac_affine(out, acc);
}

// window for n points on `threads` threads
static int default_window(uint64_t n, int threads, bool simd) {
  // the points one task sees: a window is split over about 2 threads / W
  // tasks (W ~ 18..22)
  const uint64_t split = ((uint64_t)threads * 2 + 17) / 18;
  const uint64_t per = n / (split ? split : 1);
  int bl = 0;
  while ((per >> bl) != 0) bl++;
  int c;
  if (simd) {
    // additions and bucket sums both run on the vector lanes and the work
    // is cut into equal chunks per thread, so the window follows the whole
    // MSM: c = log2(n) - 5 for large n (about 2^6 points per bucket), up to
    // log2(n) - 3 (at most 13) for small n, where fewer windows matter more
    // than bucket sums (measured on 16 cores: best c 8-9 at 2^12, 11 at
    // 2^14, 12-13 at 2^16, 13 at 2^18, 15 at 2^20)
    int ln = 0;
    while ((n >> ln) > 1) ln++;
    c = ln - 5;
    if (c < ln - 3 && c < 13) c = ln - 3 < 13 ? ln - 3 : 13;
    if (c > 16) c = 16;
  } else {
    // about 2^5 points per bucket; at most 2^13 buckets (1.1 MB of affine
    // points) per task, so a task's buckets stay in a core's L2 (c = 15, 16
    // measured 3-10% slower at 2^19..2^20 points)
    c = bl - 5;
    if (c > 14) c = 14;
  }
  return c < 4 ? 4 : c;
}

// A G2 base set in both forms: the canonical words (scalar path) and, where
// the CPU has IFMA, that path's form, converted once.
struct Bases {
  uint64_t n = 0;
  std::unique_ptr<uint64_t[]> words;
  std::unique_ptr<uint64_t[]> pts20;
  std::unique_ptr<uint8_t[]> inf;
};

}  // namespace g2

// Horner over the device G2 MSM's window terms (XYZZ -> Jacobian as
// (X ZZ^2, Y ZZZ^2, ZZZ), which holds because ZZ^3 = ZZZ^2)
void g2_bits_combine(const uint64_t* t, const uint32_t* pos, uint32_t count, uint64_t out[16]) {
  using namespace g2;
  memset(out, 0, 16 * sizeof(uint64_t));
  auto term = [&](uint32_t i) {
    const uint64_t* w = t + (size_t)32 * i;
    F2 c[4];
    for (int k = 0; k < 4; k++) {
      memcpy(c[k].c0.v, w + 8 * k, 32);
      memcpy(c[k].c1.v, w + 8 * k + 4, 32);
    }
    if (f2_zero(c[2])) return j_inf();
    J2 r;
    r.X = f2_mul(c[0], f2_sqr(c[2]));
    r.Y = f2_mul(c[1], f2_sqr(c[3]));
    r.Z = c[3];
    return r;
  };
  if (count == 0) return;
  J2 acc = term(0);
  uint32_t at = pos[0];
  for (uint32_t i = 1; i < count; i++) {
    for (; at > pos[i]; at--) acc = j_dbl(acc);
    acc = j_add(acc, term(i));
  }
  for (; at > 0; at--) acc = j_dbl(acc);
  if (j_is_inf(acc)) return;
  wr_jac_affine(out, acc);
}
}  // namespace zk

using namespace zk;

extern "C" {

static int norm_threads(int threads) {
  if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
  return threads < 1 ? 1 : threads;
}

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
  threads = norm_threads(threads);
  const bool simd = g2::simd_on();
  const int c = window_bits > 0 ? window_bits : g2::default_window(n, threads, simd);
  if (c < 2 || c > 20) return ZK_ERR_INVALID;
  try {
    std::unique_ptr<uint64_t[]> pts20;
    std::unique_ptr<uint8_t[]> inf;
    if (simd) {  // uninitialised: the conversion's workers write every word
      pts20.reset(new uint64_t[(size_t)n * g2simd::PT_WORDS]);
      inf.reset(new uint8_t[n]);
      g2::prepare(bases, n, threads, pts20.get(), inf.get());
    }
    g2::msm(bases, pts20.get(), inf.get(), scalars, n, c, threads, out);
  } catch (const std::bad_alloc&) {
    return ZK_ERR_RUNTIME;
  }
  return ZK_OK;
}

/* A G2 base set prepared once for many MSMs (the proving key's
 * b_query_g2): the points are copied, and converted to the IFMA path's
 * form where the CPU has it.  Bases as for zk_g2_msm_host. */
int zk_g2_bases_create(const uint64_t* bases, uint64_t n, int threads, zk_g2_bases** out) {
  if (!out || (n && !bases)) return ZK_ERR_INVALID;
  *out = nullptr;
  threads = norm_threads(threads);
  try {
    std::unique_ptr<g2::Bases> h(new g2::Bases);
    h->n = n;
    h->words.reset(new uint64_t[(size_t)(n ? n : 1) * 16]);
    if (n) memcpy(h->words.get(), bases, (size_t)n * 128);
    if (g2::cpu_ifma() && n) {
      h->pts20.reset(new uint64_t[(size_t)n * g2simd::PT_WORDS]);
      h->inf.reset(new uint8_t[n]);
      g2::prepare(bases, n, threads, h->pts20.get(), h->inf.get());
    }
    *out = reinterpret_cast<zk_g2_bases*>(h.release());
  } catch (const std::bad_alloc&) {
    return ZK_ERR_RUNTIME;
  }
  return ZK_OK;
}

int zk_g2_bases_destroy(zk_g2_bases* h) {
  delete reinterpret_cast<g2::Bases*>(h);
  return ZK_OK;
}

/* sum_i s_i Q_i over the first n points of a prepared set (n <= its size),
 * otherwise as zk_g2_msm_host. */
int zk_g2_msm_prepared(const zk_g2_bases* h, const uint64_t* scalars, uint64_t n, int window_bits, int threads,
                       uint64_t out[16]) {
  const g2::Bases* B = reinterpret_cast<const g2::Bases*>(h);
  if (!B || !out || (n && !scalars) || n > B->n) return ZK_ERR_INVALID;
  if (n == 0) {
    memset(out, 0, 128);
    return ZK_OK;
  }
  threads = norm_threads(threads);
  const bool simd = g2::simd_on() && B->pts20;
  const int c = window_bits > 0 ? window_bits : g2::default_window(n, threads, simd);
  if (c < 2 || c > 20) return ZK_ERR_INVALID;
  try {
    g2::msm(B->words.get(), simd ? B->pts20.get() : nullptr, simd ? B->inf.get() : nullptr, scalars, n, c, threads,
            out);
  } catch (const std::bad_alloc&) {
    return ZK_ERR_RUNTIME;
  }
  return ZK_OK;
}

/* out[i] = [k_i] Q for one point Q (affine canonical, 16 words) and n
 * canonical scalars, affine canonical outputs (16 words each). */
int zk_g2_mul_fixed_host(const uint64_t* q, const uint64_t* scalars, uint64_t n, int threads, uint64_t* out) {
  if (!q || (n && (!scalars || !out))) return ZK_ERR_INVALID;
  threads = norm_threads(threads);
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

/* The host G2 MSM's AVX-512 IFMA path: mode 1 uses it where the CPU has
 * AVX-512F and IFMA (the default), 0 uses the scalar path, any other value
 * only asks.  Returns 1 if the IFMA path is in use after the call.  Both
 * paths give the same points. */
int zk_host_simd(int mode) {
  if (mode == 0 || mode == 1) g2::g_simd_mode.store(mode);
  return g2::simd_on() ? 1 : 0;
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
