// Fr NTT / iNTT / coset NTT on gfx950: self-sorting (Stockham) radix-2
// stages, grouped into passes of k stages that run on an LDS tile.
//
// Semantics follow the reference's CPU NTT (crypto/poly/ntt.py:42-146) and
// its Metal Stockham v2 (runtime/kernels/metal/kernels.py:383-498, 601-702,
// runtime/metal/runtime.py:336-362): natural-order input and output,
// evals[i] = sum_j a_j w^(ij); the inverse uses w^-1 and scales by n^-1;
// the coset forms scale coefficient j by g^j (forward, before) or g^-j
// (inverse, after).  Forward input shorter than n is zero-padded and longer
// input is truncated (crypto/poly/ntt.py:103-110).
//
// Pass structure.  Global radix-2 Stockham stage s (sub-transform size 2^s)
// maps input index [b | mid | low_s] to output [mid | b' | low_s]: it takes
// the top index bit, butterflies it with twiddle w_(2^(s+1))^(low_s) and
// parks it at bit s.  k consecutive stages starting at s0 therefore touch
// disjoint groups of R = 2^k elements spaced n/R apart, i.e. one radix-R
// Stockham pass.  A workgroup takes T adjacent groups (an R x T tile, T
// consecutive "columns" j so global reads and writes are T*32-byte runs),
// runs the k stages in LDS with the exact global twiddles (one Fr product
// per butterfly, none extra for the pass split), and writes the tile back.
// Coset / n^-1 scaling is fused into the first pass's loads and the last
// pass's stores.
//
// Arithmetic runs on 9 x 29-bit limbs (field29.hpp, R = 2^261).  The data
// buffers hold R = 2^256 Montgomery values; since the transform is linear,
// reading them as R = 2^261 values scales every element by the same 2^-5,
// which the output carries back out.  Twiddles and scaling tables are kept
// in R = 2^261 form (packed 32 bytes per entry).  Inside a pass values are
// not reduced at all (a stage adds at most 2p to their bound, and 9 x 29
// bits leave room for 24p); a pass ends with one quotient-estimate
// reduction per element (below 1.0001p), and the last pass stores canonical
// values.
#include <algorithm>
#include "internal.hpp"
#include "field29.hpp"

namespace zk {

// Tiles: 2^10 elements (36 KiB as limbs, 256 threads: four workgroups per
// CU) or 2^12 (144 KiB of the 160 KiB LDS, 1024 threads: one workgroup per
// CU), so that a pass can hold up to 12 stages.  A thread runs one radix-4
// group per step, so a tile has four elements per thread.
static constexpr int NTT_TILE_BITS = 10;
static constexpr int NTT_BIG_TILE_BITS = 12;
template <int TB>
struct NttTile {
  static constexpr int ELEMS = 1 << TB;
  static constexpr int THREADS = ELEMS / 4;
  static constexpr int BFLY_PER_THREAD = ELEMS / 2 / THREADS;
};
static constexpr int POW_LO_BITS = POW_TABLE_LO_BITS;  // two-level power tables: lo has 2^10 entries

// tables: R = 2^261 values packed in 32 bytes
__device__ __forceinline__ f9 tab9(const fe* __restrict__ t, uint64_t i) { return unpack9(load_fe(t + i)); }
// twiddles: 9 limbs per entry
__device__ __forceinline__ f9 tw9(const uint32_t* __restrict__ t, uint32_t i) {
  const uint32_t* e = t + (uint64_t)i * 9;
  f9 r;
#pragma unroll
  for (int k = 0; k < 9; k++) r.l[k] = e[k];
  return r;
}

__device__ __forceinline__ f9 pow_lookup(const fe* __restrict__ hi, const fe* __restrict__ lo, uint64_t idx) {
  return Fr9::mul(tab9(hi, idx >> POW_LO_BITS), tab9(lo, idx & ((1u << POW_LO_BITS) - 1)));
}

// scale modes
// SCALE_TABLE: two-level power tables (one product per lookup); SCALE_FULL:
// one 9-limb entry per element (scale * base^j, built once per size and
// shift), read at the element's own index
enum { SCALE_NONE = 0, SCALE_CONST = 1, SCALE_TABLE = 2, SCALE_FULL = 3 };

// xcd_group > 1: tiles are numbered so that the workgroups one XCD receives
// (the hardware deals blockIdx.x round robin over the 8 XCDs) take
// consecutive tiles; with one-column tiles (T = 1) the 32-byte runs of
// neighbouring columns then share a cache line inside one XCD's L2.
// INV: the transform with the inverse root, read from the forward table.
// w^-e = w^(n - e) = -w^(n/2 - e), and the table holds w^e for 0 <= e <=
// n/2 (w^(n/2) = -1), so the twiddle is the entry n/2 - e negated; negating
// the twiddle of a butterfly (u + w v, u - w v) only exchanges its two
// outputs, so those stages write the sum where the difference goes and back.
// IP: in-place stages.  The tile is loaded with its rows in bit-reversed
// order (row r of the LDS tile holds input row bitrev_k(r)); a radix-4 step
// then reads its four values from the very rows it writes (o2, o2 + 2^(t+1),
// o2 + 2^t, o2 + 3 2^t), the in-place form of the same butterflies with the
// same twiddles, and the last stage leaves the rows in natural order.  A
// thread only overwrites what it read, so a step needs one barrier (before
// the next step's reads) instead of two.
template <int TB, bool INV, bool IP>
__global__ __launch_bounds__(NttTile<TB>::THREADS) void ntt_pass_kernel(
    const fe* __restrict__ in, fe* __restrict__ out, const uint32_t* __restrict__ tw, uint32_t log_n,
    uint32_t s0, uint32_t k, uint32_t log_t, uint64_t in_len, int pre_mode,
    const fe* __restrict__ pre_hi, const fe* __restrict__ pre_lo, int post_mode,
    const fe* __restrict__ post_hi, const fe* __restrict__ post_lo, int last, uint32_t xcd_group,
    int tw_stage, const uint32_t* __restrict__ pre_full, const uint32_t* __restrict__ post_full) {
  constexpr int NTT_THREADS = NttTile<TB>::THREADS;
  constexpr int NTT_BFLY_PER_THREAD = NttTile<TB>::BFLY_PER_THREAD;
  __shared__ f9 lds[NttTile<TB>::ELEMS];
  const uint64_t n = 1ull << log_n;
  const uint32_t R = 1u << k;
  const uint32_t T = 1u << log_t;
  const uint32_t RT = R << log_t;
  const uint64_t cols = n >> k;  // n / R
  const uint64_t boff = (uint64_t)blockIdx.y * n;
  in += boff;
  out += boff;
  uint64_t tile = blockIdx.x;
  if (xcd_group > 1) {
    const uint64_t per = gridDim.x / xcd_group;  // the host passes a grid divisible by xcd_group
    tile = (tile % xcd_group) * per + tile / xcd_group;
  }
  const uint64_t j0 = tile << log_t;
  const uint64_t Ns = 1ull << s0;
  const uint32_t tid = threadIdx.x;

  // ---- load the R x T tile: element (r, c) <- in[j0 + c + r * n/R]
  for (uint32_t e = tid; e < RT; e += NTT_THREADS) {
    const uint32_t r = e >> log_t, c = e & (T - 1);
    const uint32_t rr = IP ? (__brev(r) >> (32 - k)) : r;  // k >= 1
    const uint64_t idx = j0 + c + (uint64_t)rr * cols;
    f9 v;
    if (idx < in_len) {
      v = unpack9(load_fe(in + idx));
      if (s0 == 0) v = Fr9::reduce2p(v);  // caller data: any value < 2^256 (< 5.3 p) -> < 4p
      if (pre_mode == SCALE_TABLE) v = Fr9::mul(v, pow_lookup(pre_hi, pre_lo, idx));
      else if (pre_mode == SCALE_FULL) v = Fr9::mul(v, tw9(pre_full, (uint32_t)idx));
    } else {
      v = Fr9::zero();
    }
    lds[e] = v;
  }
  __syncthreads();

  // ---- k radix-2 Stockham stages in LDS.  Values grow lazily (bounds in
  //      units of p): a stage forms u + w v and u - w v + 2p with w v < 1.2p,
  //      so it adds at most 2 to the bound, from < 4 after the load to < 24
  //      after 10 stages, inside the 2^261 the limbs hold; the product takes
  //      such inputs (its output is below p + v w / 2^261).  Global stage 0
  //      has twiddle 1 and reduces v to [0, 2p) instead.
  //      The additions carry nothing between limbs: a limb of u + w v or
  //      u - w v + 2p (the latter through 2p with limbs near 2^29, K2P) is
  //      below u's limb + 2^30.  From 29-bit limbs one stage gives limbs
  //      below 1.5 * 2^30 and a second below 2.5 * 2^30, so u is normalised
  //      at every even pass stage after the first and no limb exceeds
  //      2.5 * 2^30.  v enters the product with such limbs: a column then
  //      sums at most 9 products below 2.5 * 2^59 (v by the twiddle), 9
  //      below 2^58 (m by p) and a carry below 2^35, under 27 * 2^59 < 2^64.
  //      Stages run in pairs (t, t + 1), t even, as one radix-4 step: a
  //      thread takes the rows q, q + R/4, q + R/2, q + 3R/4 of one column
  //      (q < R/4) through both stages in registers.  Stage t pairs rows
  //      (q, q + R/2) and (q + R/4, q + 3R/4), both with d = q mod 2^t and
  //      so one twiddle; its outputs land on rows o, o + 2^t, o + R/2,
  //      o + R/2 + 2^t (o = 2(q - d) + d), which stage t + 1 pairs as
  //      (o, o + R/2) with d and (o + 2^t, o + 2^t + R/2) with d + 2^t,
  //      writing rows o2, o2 + 2^(t+1), o2 + 2^t, o2 + 2^t + 2^(t+1) with
  //      o2 = 4(q - d) + d.  An odd stage count ends with one radix-2 stage.
  const uint32_t halfRT = RT >> 1, quarterRT = RT >> 2;
  // twiddle of global stage s = s0 + t: w_(2^(s+1))^x, x = low + d Ns.  One
  // table of w^e (e <= n/2) reads entry x 2^(log_n - 1 - s); per-stage
  // tables (tw_stage) hold w_(2^(s+1))^x for 0 <= x <= 2^s at 2^s - 1 + s,
  // so the adjacent columns of a tile read adjacent entries at every stage.
  // The inverse reads the reflected entry (n/2 - e, or 2^s - x).
  //  Indices stay in 32 bits (n <= 2^28: x < 2^27, entries below 2^28 + 28).
  const uint32_t jc0 = (uint32_t)j0, nmask = (uint32_t)Ns - 1u;
  auto tw_cd = [&](uint32_t c, uint32_t d, uint32_t t) -> f9 {
    const uint32_t x = ((jc0 + c) & nmask) + (d << s0);
    const uint32_t s = s0 + t;
    if (tw_stage) return tw9(tw, ((1u << s) - 1u + s) + (INV ? (1u << s) - x : x));
    const uint32_t e = x << (log_n - 1 - s);
    return tw9(tw, INV ? (uint32_t)(n >> 1) - e : e);
  };
  // the two outputs of a table-twiddle butterfly: (u + w v, u - w v + 2p),
  // exchanged in the inverse (the table gives -w there)
  auto bfly = [&](const f9& u, const f9& wv, f9& lo, f9& hi) {
    if (INV) {
      lo = Fr9::sub_cf(u, wv);
      hi = Fr9::add_cf(u, wv);
    } else {
      lo = Fr9::add_cf(u, wv);
      hi = Fr9::sub_cf(u, wv);
    }
  };
  uint32_t t = 0;
  for (; t + 1 < k; t += 2) {
    const bool act = tid < quarterRT;
    const uint32_t c = tid & (T - 1), q = tid >> log_t;
    const uint32_t d = q & ((1u << t) - 1);
    const uint32_t R4 = R >> 2;
    f9 x0, x1, x2, x3, w1, w2, w3;
    if (act) {
      if (IP) {
        const uint32_t o2 = 4 * (q - d) + d;
        x0 = lds[(o2 << log_t) + c];
        x1 = lds[((o2 + (2u << t)) << log_t) + c];
        x2 = lds[((o2 + (1u << t)) << log_t) + c];
        x3 = lds[((o2 + (3u << t)) << log_t) + c];
      } else {
        x0 = lds[(q << log_t) + c];
        x1 = lds[((q + R4) << log_t) + c];
        x2 = lds[((q + 2 * R4) << log_t) + c];
        x3 = lds[((q + 3 * R4) << log_t) + c];
      }
      if (s0 + t != 0) {
        w1 = tw_cd(c, d, t);
        w2 = tw_cd(c, d, t + 1);
      }
      w3 = tw_cd(c, d + (1u << t), t + 1);
    }
    if (!IP) __syncthreads();
    if (act) {
      f9 A, B, C, D;
      f9 r0, r1, r2, r3;
      if (s0 + t != 0) {
        const f9 u0 = (t != 0) ? Fr9::normalize(x0) : x0;
        const f9 u1 = (t != 0) ? Fr9::normalize(x1) : x1;
        f9 v2, v3;
        Fr9::mul_x2(x2, w1, x3, w1, v2, v3);  // two chains interleaved
        bfly(u0, v2, A, B);
        bfly(u1, v3, C, D);
        const f9 vc = Fr9::mul(C, w2), vd = Fr9::mul(D, w3);
        bfly(A, vc, r0, r1);
        bfly(B, vd, r2, r3);
      } else {
        // global stage 0: v below 2p, not 1.2p: the carrying forms
        const f9 v2 = Fr9::reduce2p(x2), v3 = Fr9::reduce2p(x3);
        A = Fr9::add_nr(x0, v2);
        B = Fr9::sub_nr(x0, v2);
        C = Fr9::add_nr(x1, v3);
        D = Fr9::sub_nr(x1, v3);
        // global stage 1: the pair (A, C) has twiddle w^0 = 1 (the inverse's
        // table gives -1 there with the outputs exchanged: the same pair), so
        // C (< 6p) is only reduced, not multiplied
        const f9 vc = Fr9::reduce_est(C), vd = Fr9::mul(D, w3);
        r0 = Fr9::add_cf(A, vc);
        r1 = Fr9::sub_cf(A, vc);
        bfly(B, vd, r2, r3);
      }
      const uint32_t o2 = 4 * (q - d) + d;
      lds[(o2 << log_t) + c] = r0;
      lds[((o2 + (2u << t)) << log_t) + c] = r1;
      lds[((o2 + (1u << t)) << log_t) + c] = r2;
      lds[((o2 + (3u << t)) << log_t) + c] = r3;
    }
    __syncthreads();
  }
  if (t < k) {
    const uint32_t s = s0 + t;
    f9 u[NTT_BFLY_PER_THREAD], v[NTT_BFLY_PER_THREAD], w[NTT_BFLY_PER_THREAD];
#pragma unroll
    for (int i = 0; i < NTT_BFLY_PER_THREAD; i++) {
      const uint32_t b = tid + i * NTT_THREADS;
      if (b < halfRT) {
        const uint32_t c = b & (T - 1), q = b >> log_t;
        const uint32_t d = q & ((1u << t) - 1);
        const uint32_t pu = IP ? ((q - d) << 1) + d : q;
        const uint32_t pv = IP ? pu + (1u << t) : q + (R >> 1);
        u[i] = lds[(pu << log_t) + c];
        v[i] = lds[(pv << log_t) + c];
        if (s != 0) w[i] = tw_cd(c, d, t);
      }
    }
    if (!IP) __syncthreads();
#pragma unroll
    for (int i = 0; i < NTT_BFLY_PER_THREAD; i++) {
      const uint32_t b = tid + i * NTT_THREADS;
      if (b < halfRT) {
        const uint32_t c = b & (T - 1), q = b >> log_t;
        const uint32_t d = q & ((1u << t) - 1);
        const uint32_t o = ((q - d) << 1) + d;
        f9 lo, hi;
        if (s != 0) {
          const f9 vv = Fr9::mul(v[i], w[i]);
          const f9 uu = (t != 0) ? Fr9::normalize(u[i]) : u[i];  // t is even here
          bfly(uu, vv, lo, hi);
        } else {
          const f9 vv = Fr9::reduce2p(v[i]);
          lo = Fr9::add_nr(u[i], vv);
          hi = Fr9::sub_nr(u[i], vv);
        }
        lds[(o << log_t) + c] = lo;
        lds[((o + (1u << t)) << log_t) + c] = hi;
      }
    }
    __syncthreads();
  }

  // ---- store: element (sp, c) -> out[(j / Ns) * Ns * R + (j % Ns) + sp * Ns], j = j0 + c.
  //      Values leave below 2p: through the scaling product (< 1.3p) or the
  //      quotient-estimate reduction (< 1.0001p); the last pass makes them
  //      canonical.
  const f9 post_c = (post_mode == SCALE_CONST) ? tab9(post_hi, 0) : Fr9::zero();
  for (uint32_t e = tid; e < RT; e += NTT_THREADS) {
    uint32_t sp, c;
    if (s0 == 0) {  // Ns = 1: a column's R outputs are contiguous
      c = e >> k;
      sp = e & (R - 1);
    } else {
      sp = e >> log_t;
      c = e & (T - 1);
    }
    const uint64_t j = j0 + c;
    const uint64_t pos = ((j >> s0) << (s0 + k)) + (j & (Ns - 1)) + ((uint64_t)sp << s0);
    f9 x = lds[(sp << log_t) + c];
    if (post_mode == SCALE_CONST) {
      x = Fr9::mul(x, post_c);
    } else if (post_mode == SCALE_TABLE) {
      x = Fr9::mul(x, pow_lookup(post_hi, post_lo, pos));
    } else if (post_mode == SCALE_FULL) {
      x = Fr9::mul(x, tw9(post_full, (uint32_t)pos));
    } else {
      x = Fr9::reduce_est(x);  // limbs up to 2.5 * 2^30 in, normalised out
    }
    if (last) x = Fr9::canonical(x);
    store_fe(out + pos, pack9(x));
  }
}

// twiddle table tw[e] = hi[e >> 10] * lo[e & 1023], stored as 9 limbs per
// entry so the butterflies load it without unpacking
__global__ __launch_bounds__(256) void pow_expand_kernel(const fe* __restrict__ hi,
                                                         const fe* __restrict__ lo,
                                                         uint32_t* __restrict__ out, uint64_t count) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const f9 t = Fr9::canonical(pow_lookup(hi, lo, i));
#pragma unroll
  for (int k = 0; k < 9; k++) out[9 * i + k] = t.l[k];
}

// per-stage twiddle tables: stage s (0 <= s < log_n) holds w_(2^(s+1))^x =
// w^(x 2^(log_n - 1 - s)) for 0 <= x <= 2^s at offset 2^s - 1 + s
__global__ __launch_bounds__(256) void stage_expand_kernel(const fe* __restrict__ hi,
                                                           const fe* __restrict__ lo, uint32_t log_n,
                                                           uint32_t* __restrict__ out, uint64_t count) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= count) return;
  uint32_t s = 0;
  while (s + 1 < log_n && g >= (1ull << (s + 1)) - 1 + (s + 1)) s++;
  const uint64_t x = g - ((1ull << s) - 1 + s);
  const f9 t = Fr9::canonical(pow_lookup(hi, lo, x << (log_n - 1 - s)));
#pragma unroll
  for (int k = 0; k < 9; k++) out[9 * g + k] = t.l[k];
}

// host-built R = 2^256 tables -> R = 2^261 packed form, in place
__global__ __launch_bounds__(256) void table_to261_kernel(fe* __restrict__ t, uint64_t count) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) store_fe(t + i, pack9(Fr9::canonical(Fr9::to261(unpack9(load_fe(t + i))))));
}

// n == 1: out[0] = (in_len ? in[0] : 0) * scale
__global__ void ntt_size1_kernel(const fe* __restrict__ in, fe* __restrict__ out, uint64_t in_len,
                                 uint32_t batch) {
  uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < batch) store_fe(out + b, in_len ? load_fe(in + b) : Fr::zero());
}

// ---------------------------------------------------------------- host side

// hi[h] = scale * base^(1024 h), lo[i] = base^i   (all Montgomery)
static void build_pow_tables(const zkh::Fe& base, const zkh::Fe& scale, uint64_t count,
                             std::vector<zkh::Fe>& hi, std::vector<zkh::Fe>& lo) {
  const uint64_t L = 1ull << POW_LO_BITS;
  lo.resize(L);
  zkh::Fe x = zkh::one(zkh::FR);
  for (uint64_t i = 0; i < L; i++) {
    lo[i] = x;
    x = zkh::mul(zkh::FR, x, base);
  }
  const zkh::Fe step = x;  // base^1024
  uint64_t nh = (count + L - 1) / L;
  if (nh == 0) nh = 1;
  hi.resize(nh);
  zkh::Fe y = scale;
  for (uint64_t h = 0; h < nh; h++) {
    hi[h] = y;
    y = zkh::mul(zkh::FR, y, step);
  }
}

// Frees least recently used, unpinned tables that the current table call
// has not touched until `need` more bytes fit under the cap (or nothing is
// left to free).  Launches still in flight may read a table, so the device
// is drained before the first free.
static int make_table_room(Ctx* c, size_t need) {
  if (c->table_cap == 0) return ZK_OK;
  bool drained = false;
  while (c->table_bytes + need > c->table_cap) {
    uint64_t oldest = UINT64_MAX;
    int which = -1;
    std::array<uint64_t, 5> kt{};
    std::array<uint64_t, 9> kp{};
    for (auto& kv : c->twiddles)
      if (!kv.second.pinned && kv.second.last_call != c->table_call && kv.second.last_call < oldest) {
        oldest = kv.second.last_call;
        which = 0;
        kt = kv.first;
      }
    for (auto& kv : c->pow_tables)
      if (!kv.second.pinned && kv.second.last_call != c->table_call && kv.second.last_call < oldest) {
        oldest = kv.second.last_call;
        which = 1;
        kp = kv.first;
      }
    if (which < 0) break;  // everything left is pinned or in use: go over the cap
    if (!drained) {
      ZK_HIP(c, hipDeviceSynchronize());
      drained = true;
    }
    CachedTable t = which == 0 ? c->twiddles[kt] : c->pow_tables[kp];
    if (t.a.ptr) ZK_HIP(c, hipFree(t.a.ptr));
    if (t.b.ptr) ZK_HIP(c, hipFree(t.b.ptr));
    c->table_bytes -= t.bytes();
    if (which == 0) c->twiddles.erase(kt);
    else c->pow_tables.erase(kp);
  }
  return ZK_OK;
}

static void touch(Ctx* c, CachedTable& t) {
  t.last_call = c->table_call;
  if (c->table_pin) t.pinned = true;
}

static int get_pow_tables(Ctx* c, uint32_t log_n, const zkh::Fe& base, const zkh::Fe& scale,
                          const fe** hi, const fe** lo, hipStream_t st) {
  std::array<uint64_t, 9> key;
  key[0] = log_n;
  for (int i = 0; i < 4; i++) key[1 + i] = base.v[i];
  for (int i = 0; i < 4; i++) key[5 + i] = scale.v[i];
  auto it = c->pow_tables.find(key);
  if (it == c->pow_tables.end()) {
    std::vector<zkh::Fe> h, l;
    build_pow_tables(base, scale, 1ull << log_n, h, l);
    int rc = make_table_room(c, (h.size() + l.size()) * 32);
    if (rc) return rc;
    CachedTable t;
    rc = upload_table(c, h, &t.a, st);
    if (rc) return rc;
    rc = upload_table(c, l, &t.b, st);
    if (rc) {
      hipFree(t.a.ptr);
      return rc;
    }
    table_to261_kernel<<<grid_for(h.size(), 256), 256, 0, st>>>((fe*)t.a.ptr, h.size());
    table_to261_kernel<<<grid_for(l.size(), 256), 256, 0, st>>>((fe*)t.b.ptr, l.size());
    ZK_HIP(c, hipGetLastError());
    c->table_bytes += t.bytes();
    it = c->pow_tables.emplace(key, t).first;
  }
  touch(c, it->second);
  *hi = (const fe*)it->second.a.ptr;
  *lo = (const fe*)it->second.b.ptr;
  return ZK_OK;
}

// A full scaling table: scale * base^j for 0 <= j < 2^log_n, 9 limbs per
// entry, expanded once from the two-level tables and cached with them (its
// key marks log_n with 0x100).  Used for sizes up to 2^FULL_SCALE_MAX_BITS.
static constexpr uint32_t FULL_SCALE_MAX_BITS = 26;
static int get_scale_full(Ctx* c, uint32_t log_n, const zkh::Fe& base, const zkh::Fe& scale, const uint32_t** out,
                          hipStream_t st) {
  std::array<uint64_t, 9> key;
  key[0] = log_n | 0x100u;
  for (int i = 0; i < 4; i++) key[1 + i] = base.v[i];
  for (int i = 0; i < 4; i++) key[5 + i] = scale.v[i];
  auto it = c->pow_tables.find(key);
  if (it == c->pow_tables.end()) {
    const fe *hi, *lo;
    int rc = get_pow_tables(c, log_n, base, scale, &hi, &lo, st);
    if (rc) return rc;
    const uint64_t count = 1ull << log_n;
    rc = make_table_room(c, count * 36);
    if (rc) return rc;
    CachedTable t;
    t.a.bytes = count * 36;
    ZK_HIP(c, hipMalloc(&t.a.ptr, t.a.bytes));
    pow_expand_kernel<<<grid_for(count, 256), 256, 0, st>>>(hi, lo, (uint32_t*)t.a.ptr, count);
    ZK_HIP(c, hipGetLastError());
    c->table_bytes += t.bytes();
    it = c->pow_tables.emplace(key, t).first;
  }
  touch(c, it->second);
  *out = (const uint32_t*)it->second.a.ptr;
  return ZK_OK;
}

int ntt_pow_tables(Ctx* c, uint32_t log_n, const zkh::Fe& base, const zkh::Fe& scale, const fe** hi,
                   const fe** lo, hipStream_t st) {
  c->table_call++;
  return get_pow_tables(c, log_n, base, scale, hi, lo, st);
}

// stage: per-stage tables (2^s + 1 entries for each stage s) instead of
// one table of w^e, 0 <= e <= n/2; either serves both directions
static int get_twiddles(Ctx* c, uint32_t log_n, const zkh::Fe& w, const uint32_t** out, hipStream_t st,
                        bool stage) {
  std::array<uint64_t, 5> key = {log_n | (stage ? 0x100u : 0u), w.v[0], w.v[1], w.v[2], w.v[3]};
  auto it = c->twiddles.find(key);
  if (it == c->twiddles.end()) {
    const uint64_t count = stage ? (1ull << log_n) - 1 + log_n : (1ull << log_n) / 2 + 1;
    const fe *hi, *lo;
    int rc = get_pow_tables(c, log_n, w, zkh::one(zkh::FR), &hi, &lo, st);
    if (rc) return rc;
    rc = make_table_room(c, count * 36);
    if (rc) return rc;
    CachedTable t;
    t.a.bytes = count * 36;
    ZK_HIP(c, hipMalloc(&t.a.ptr, t.a.bytes));
    if (stage)
      stage_expand_kernel<<<grid_for(count, 256), 256, 0, st>>>(hi, lo, log_n, (uint32_t*)t.a.ptr, count);
    else
      pow_expand_kernel<<<grid_for(count, 256), 256, 0, st>>>(hi, lo, (uint32_t*)t.a.ptr, count);
    ZK_HIP(c, hipGetLastError());
    c->table_bytes += t.bytes();
    it = c->twiddles.emplace(key, t).first;
  }
  touch(c, it->second);
  *out = (const uint32_t*)it->second.a.ptr;
  return ZK_OK;
}

// omega_for_size (crypto/poly/ntt.py:13-24): ROOT_2_28^(2^(28 - log_n))
static zkh::Fe standard_root(uint32_t log_n) {
  // ROOT_2_28 = 5^((r-1)/2^28) (crypto/poly/ntt.py:8), canonical little-endian words
  const zkh::Fe root_plain = {{0x9bd61b6e725b19f0ull, 0x402d111e41112ed4ull,
                               0x00e0a7eb8ef62abcull, 0x2a3c09f0a58a7e85ull}};
  zkh::Fe w = zkh::to_mont(zkh::FR, root_plain);
  for (uint32_t i = log_n; i < 28; i++) w = zkh::mul(zkh::FR, w, w);
  return w;
}

// choose stages per pass: as few passes of at most 10 stages as the size
// needs, with near-equal pass sizes.  A 10-stage pass on the 2^10-element
// tile moves one-column (32-byte) runs on its strided side; at 2^20 two
// such passes (tiles grouped per XCD) measured 4% faster than three
// passes with 128-byte runs.  Passes of 11 or 12 stages (max_pass_bits)
// use the 2^12-element tile; at 2^24 two 12-stage passes are bit-exact but
// measured 37% slower than three 8-stage passes: one workgroup per CU
// leaves each tile's load and store phases exposed.
static void plan_passes(uint32_t log_n, int max_pass_bits, std::vector<uint32_t>& ks) {
  ks.clear();
  if (log_n == 0) return;
  // passes of more than 10 stages (asked for through max_pass_bits) run on
  // the 2^12-element tiles
  const uint32_t tile_bits = max_pass_bits > NTT_TILE_BITS ? NTT_BIG_TILE_BITS : NTT_TILE_BITS;
  uint32_t kmax = (max_pass_bits > 0) ? (uint32_t)max_pass_bits : 10;
  if (kmax > tile_bits) kmax = tile_bits;
  if (log_n <= tile_bits && (max_pass_bits <= 0 || log_n <= kmax)) {
    ks.push_back(log_n);
    return;
  }
  uint32_t p = (log_n + kmax - 1) / kmax;
  uint32_t base = log_n / p, extra = log_n % p;
  for (uint32_t i = 0; i < p; i++) ks.push_back(base + (i < extra ? 1 : 0));
  // Passes run their stages two at a time; an odd pass ends with a lone
  // radix-2 stage that costs about as much as a pair.  With the same pass
  // count, prefer even sizes (largest first) when they fit under kmax:
  // 20 = 8 + 6 + 6 rather than 7 + 7 + 6.
  if (log_n % 2 == 0) {
    std::vector<uint32_t> ev(p, 0);
    uint32_t left = log_n / 2;  // pairs to place
    for (uint32_t i = 0; i < p; i++) {
      const uint32_t rest = p - 1 - i;
      uint32_t take = (left + rest) / (rest + 1);   // near-equal split of the pairs
      if (2 * take > kmax) take = kmax / 2;
      ev[i] = 2 * take;
      left -= take;
    }
    bool ok = left == 0;
    for (uint32_t v : ev) ok = ok && v >= 1 && v <= kmax;
    if (ok) {
      std::sort(ev.begin(), ev.end(), [](uint32_t a, uint32_t b) { return a > b; });
      ks = ev;
    }
  }
}

int ntt_run(Ctx* c, const fe* in, uint64_t in_len, fe* out, uint32_t log_n, uint32_t batch,
            bool inverse, const zkh::Fe* omega, const zkh::Fe* shift, hipStream_t st,
            int max_pass_bits, fe* tmp_buf) {
  if (log_n > 28) return set_err(c, ZK_ERR_INVALID, "domain too large for BN254 root of unity");
  if (batch == 0) return ZK_OK;
  c->table_call++;
  const uint64_t n = 1ull << log_n;
  if (batch > 1 && in_len != n)
    return set_err(c, ZK_ERR_INVALID, "batched NTT needs in_len == n");
  if (log_n == 0) {
    // n^-1 = 1 and g^0 = 1: the transform of one element is the element
    ntt_size1_kernel<<<grid_for(batch, 64), 64, 0, st>>>(in, out, in_len, batch);
    ZK_HIP(c, hipGetLastError());
    return ZK_OK;
  }

  zkh::Fe w = omega ? *omega : standard_root(log_n);
  // w must be a primitive n-th root: w^(n/2) == -1
  {
    zkh::Fe h = zkh::pow_u64(zkh::FR, w, n / 2);
    zkh::Fe m1 = zkh::neg(zkh::FR, zkh::one(zkh::FR));
    if (!zkh::eq(h, m1)) return set_err(c, ZK_ERR_INVALID, "omega is not a primitive n-th root of unity");
  }
  // one table per (n, w) serves both directions
  const uint32_t* tw;
  const bool tw_stage = c->ntt_stage_twiddles != 0;
  int rc = get_twiddles(c, log_n, w, &tw, st, tw_stage);
  if (rc) return rc;

  int pre_mode = 0, post_mode = 0;
  const fe *pre_hi = nullptr, *pre_lo = nullptr, *post_hi = nullptr, *post_lo = nullptr;
  const uint32_t *pre_full = nullptr, *post_full = nullptr;
  const bool full = c->ntt_full_scale && log_n <= FULL_SCALE_MAX_BITS;
  if (!inverse) {
    if (shift) {
      if (full) {
        pre_mode = SCALE_FULL;
        rc = get_scale_full(c, log_n, *shift, zkh::one(zkh::FR), &pre_full, st);
      } else {
        pre_mode = SCALE_TABLE;
        rc = get_pow_tables(c, log_n, *shift, zkh::one(zkh::FR), &pre_hi, &pre_lo, st);
      }
      if (rc) return rc;
    }
  } else {
    zkh::Fe ninv = zkh::inv(zkh::FR, zkh::from_u64(zkh::FR, n));
    if (shift) {
      zkh::Fe ginv = zkh::is_zero(*shift) ? zkh::zero() : zkh::inv(zkh::FR, *shift);
      if (full) {
        post_mode = SCALE_FULL;
        rc = get_scale_full(c, log_n, ginv, ninv, &post_full, st);
      } else {
        post_mode = SCALE_TABLE;
        rc = get_pow_tables(c, log_n, ginv, ninv, &post_hi, &post_lo, st);
      }
    } else {
      post_mode = 1;
      // constant n^-1: a one-entry "hi" table
      rc = get_pow_tables(c, 0, zkh::zero(), ninv, &post_hi, &post_lo, st);
    }
    if (rc) return rc;
  }

  std::vector<uint32_t> ks;
  plan_passes(log_n, max_pass_bits, ks);
  const size_t P = ks.size();

  // buffer sequence: in -> ... -> out, intermediates never alias their source
  fe* tmp = tmp_buf;
  if (P > 1 && !tmp) {
    void* p;
    rc = ensure_scratch(c, 0, (size_t)n * batch * 32, &p);
    if (rc) return rc;
    tmp = (fe*)p;
  }
  std::vector<const fe*> src(P);
  std::vector<fe*> dst(P);
  for (size_t i = 0; i < P; i++) {
    // pass i writes out when (P - 1 - i) is even, tmp otherwise
    dst[i] = (((P - 1 - i) & 1) == 0) ? out : tmp;
  }
  if (P > 1 && (const void*)in == (const void*)out && dst[0] == out) {
    if (tmp_buf) return set_err(c, ZK_ERR_INVALID, "in-place NTT with a caller scratch buffer is not supported");
    // in-place with an odd pass count: copy the input aside first
    void* p2;
    rc = ensure_scratch(c, 1, (size_t)n * batch * 32, &p2);
    if (rc) return rc;
    ZK_HIP(c, hipMemcpyAsync(p2, in, (size_t)(batch > 1 ? n * batch : (in_len < n ? in_len : n)) * 32,
                             hipMemcpyDeviceToDevice, st));
    in = (const fe*)p2;
  }
  uint32_t s0 = 0;
  hipEvent_t ev;
  phase_begin(c, "ntt", st, &ev);
  for (size_t i = 0; i < P; i++) {
    const uint32_t k = ks[i];
    const uint64_t cols = n >> k;
    const bool big = k > NTT_TILE_BITS || c->ntt_tile_bits == NTT_BIG_TILE_BITS;
    const uint32_t tb = big ? NTT_BIG_TILE_BITS : NTT_TILE_BITS;
    uint32_t log_t = tb - k;
    while ((1ull << log_t) > cols) log_t--;
    const uint64_t tiles = cols >> log_t;
    uint32_t xg = c->ntt_xcd_group ? (uint32_t)c->ntt_xcd_group : (log_t <= 1 ? 8u : 1u);
    if (xg > 1 && tiles % xg != 0) xg = 1;
    const fe* s = (i == 0) ? in : dst[i - 1];
    const uint64_t len = (i == 0) ? in_len : n;
    dim3 grid((unsigned)tiles, batch);
#define ZK_NTT_PASS(TBV, INVV, IPV)                                                                     \
  ntt_pass_kernel<TBV, INVV, IPV><<<grid, NttTile<TBV>::THREADS, 0, st>>>(                              \
      s, dst[i], tw, log_n, s0, k, log_t, len, i == 0 ? pre_mode : 0, pre_hi, pre_lo,                   \
      i == P - 1 ? post_mode : 0, post_hi, post_lo, i == P - 1 ? 1 : 0, xg, tw_stage ? 1 : 0, pre_full,     \
      post_full)
#define ZK_NTT_PASS_IP(TBV, INVV)            \
  do {                                       \
    if (c->ntt_inplace) ZK_NTT_PASS(TBV, INVV, true); \
    else ZK_NTT_PASS(TBV, INVV, false);      \
  } while (0)
    if (big) {
      if (inverse) ZK_NTT_PASS_IP(NTT_BIG_TILE_BITS, true);
      else ZK_NTT_PASS_IP(NTT_BIG_TILE_BITS, false);
    } else {
      if (inverse) ZK_NTT_PASS_IP(NTT_TILE_BITS, true);
      else ZK_NTT_PASS_IP(NTT_TILE_BITS, false);
    }
#undef ZK_NTT_PASS_IP
#undef ZK_NTT_PASS
    ZK_HIP(c, hipGetLastError());
    s0 += k;
  }
  phase_end(c, "ntt", st, ev);
  return ZK_OK;
}

}  // namespace zk

namespace zk {
// Build the twiddle / scaling tables an NTT of this shape needs (so a later
// launch sequence can run without allocating, e.g. under graph capture).
int ntt_prepare(Ctx* c, uint32_t log_n, bool inverse, const zkh::Fe* omega, const zkh::Fe* shift,
                hipStream_t st) {
  if (log_n == 0) return ZK_OK;
  if (log_n > 28) return set_err(c, ZK_ERR_INVALID, "domain too large for BN254 root of unity");
  const uint64_t n = 1ull << log_n;
  zkh::Fe w = omega ? *omega : standard_root(log_n);
  // a plan's recorded launches keep these tables' addresses: pin them
  c->table_call++;
  c->table_pin = true;
  const uint32_t* tw;
  int rc = get_twiddles(c, log_n, w, &tw, st, c->ntt_stage_twiddles != 0);
  const fe *h, *l;
  const uint32_t* f;
  const bool full = c->ntt_full_scale && log_n <= FULL_SCALE_MAX_BITS;
  if (rc == ZK_OK && !inverse) {
    if (shift) {
      rc = full ? get_scale_full(c, log_n, *shift, zkh::one(zkh::FR), &f, st)
                : get_pow_tables(c, log_n, *shift, zkh::one(zkh::FR), &h, &l, st);
    }
  } else if (rc == ZK_OK) {
    zkh::Fe ninv = zkh::inv(zkh::FR, zkh::from_u64(zkh::FR, n));
    if (shift) {
      zkh::Fe ginv = zkh::is_zero(*shift) ? zkh::zero() : zkh::inv(zkh::FR, *shift);
      rc = full ? get_scale_full(c, log_n, ginv, ninv, &f, st) : get_pow_tables(c, log_n, ginv, ninv, &h, &l, st);
    } else {
      rc = get_pow_tables(c, 0, zkh::zero(), ninv, &h, &l, st);
    }
  }
  c->table_pin = false;
  return rc;
}
}  // namespace zk
