// BN254 G1 multi-scalar multiplication on gfx950: signed-digit Pippenger
// with CSR buckets built by a device counting sort, load-balanced bucket
// accumulation in XYZZ coordinates, and a parallel bucket reduction.
//
// Follows the reference's Metal MSM v2 (runtime/metal/msm.py:145-332,
// runtime/metal/ecc_source.py:437-566) in what is computed:
//   * window count W = ceil(255 / c), signed digits in [-2^(c-1), 2^(c-1)-1]
//     with a carry into the next window (runtime/metal/msm.py:145-177),
//   * 2^(c-1) + 1 buckets per window, bucket |d| takes +-P (sign from d),
//   * window sums Sum_d d * B_d, combined as Sum_w 2^(cw) S_w on the host
//     (runtime/metal/msm.py:322-332).
// It differs in how: digits come from the device-resident Montgomery
// scalars on the device (no host loop), the CSR build is a parallel
// histogram / scan / scatter, each thread owns a fixed-length segment of the
// sorted (bucket, point) list instead of a whole bucket (so skewed buckets
// do not serialise a thread), and the bucket reduction is split into chunks
// whose weighted partial sums are recombined in parallel.
#include <stdio.h>
#include <mutex>
#include "internal.hpp"
#include "curve.hpp"
#include "curve29.hpp"
#include "curve29_g2.hpp"
#include <type_traits>

namespace zk {

static constexpr uint32_t NO_KEY = 0xffffffffu;
static constexpr uint32_t SIGN = 0x80000000u;

__device__ __forceinline__ xyzz load_xyzz(const xyzz* p) {
  const fe* f = reinterpret_cast<const fe*>(p);
  xyzz r;
  r.x = load_fe(f);
  r.y = load_fe(f + 1);
  r.zz = load_fe(f + 2);
  r.zzz = load_fe(f + 3);
  return r;
}

__device__ __forceinline__ void store_xyzz(xyzz* p, const xyzz& v) {
  fe* f = reinterpret_cast<fe*>(p);
  store_fe(f, v.x);
  store_fe(f + 1, v.y);
  store_fe(f + 2, v.zz);
  store_fe(f + 3, v.zzz);
}

// Bases: coordinates are R = 2^261 limb values (field29.hpp) packed back
// into 32 bytes each, so a base is 64 bytes (half a 128-byte line per
// gather).
__device__ __forceinline__ aff9 load_aff9(const fe* __restrict__ bases, uint32_t idx) {
  aff9 a;
  a.x = unpack9(load_fe(bases + 2 * (uint64_t)idx));
  a.y = unpack9(load_fe(bases + 2 * (uint64_t)idx + 1));
  return a;
}

// Buckets, head partials and reduction partials stay in limb form: 4 x 9
// limbs (144 bytes, nine 16-byte words), so the hot loop's bucket stores
// and the fix-up / reduction loads need no packing.
struct __align__(16) pt9 {
  uint4 q[9];
};

__device__ __forceinline__ xyzz9 load_xyzz9(const pt9* p) {
  uint32_t w[36];
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const uint4 v = p->q[i];
    w[4 * i] = v.x;
    w[4 * i + 1] = v.y;
    w[4 * i + 2] = v.z;
    w[4 * i + 3] = v.w;
  }
  xyzz9 r;
#pragma unroll
  for (int j = 0; j < 9; j++) {
    r.x.l[j] = w[j];
    r.y.l[j] = w[9 + j];
    r.zz.l[j] = w[18 + j];
    r.zzz.l[j] = w[27 + j];
  }
  return r;
}

__device__ __forceinline__ void store_xyzz9(pt9* p, const xyzz9& v) {
  uint32_t w[36];
#pragma unroll
  for (int j = 0; j < 9; j++) {
    w[j] = v.x.l[j];
    w[9 + j] = v.y.l[j];
    w[18 + j] = v.zz.l[j];
    w[27 + j] = v.zzz.l[j];
  }
#pragma unroll
  for (int i = 0; i < 9; i++) p->q[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
}

// G2 buckets, head partials and sums: 8 x 9 limbs (288 bytes)
struct __align__(16) pt18 {
  uint4 q[18];
};

// coordinate k of a G2 XYZZ point: x.c0, x.c1, y.c0, y.c1, zz.c0, zz.c1,
// zzz.c0, zzz.c1 (k a compile-time constant after unrolling)
__device__ __forceinline__ f9& g2_coord(xyzz9g2& p, int k) {
  f92& e = k < 2 ? p.x : k < 4 ? p.y : k < 6 ? p.zz : p.zzz;
  return (k & 1) ? e.c1 : e.c0;
}
__device__ __forceinline__ const f9& g2_coord(const xyzz9g2& p, int k) {
  const f92& e = k < 2 ? p.x : k < 4 ? p.y : k < 6 ? p.zz : p.zzz;
  return (k & 1) ? e.c1 : e.c0;
}

__device__ __forceinline__ xyzz9g2 load_xyzz9g2(const pt18* p) {
  uint32_t w[72];
#pragma unroll
  for (int i = 0; i < 18; i++) {
    const uint4 v = p->q[i];
    w[4 * i] = v.x;
    w[4 * i + 1] = v.y;
    w[4 * i + 2] = v.z;
    w[4 * i + 3] = v.w;
  }
  xyzz9g2 r;
#pragma unroll
  for (int k = 0; k < 8; k++)
#pragma unroll
    for (int j = 0; j < 9; j++) g2_coord(r, k).l[j] = w[9 * k + j];
  return r;
}

__device__ __forceinline__ void store_xyzz9g2(pt18* p, const xyzz9g2& v) {
  uint32_t w[72];
#pragma unroll
  for (int k = 0; k < 8; k++)
#pragma unroll
    for (int j = 0; j < 9; j++) w[9 * k + j] = g2_coord(v, k).l[j];
#pragma unroll
  for (int i = 0; i < 18; i++) p->q[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
}

// G2 bases: x.c0, x.c1, y.c0, y.c1 as R = 2^261 limb values packed in 32
// bytes each (128 bytes a base)
__device__ __forceinline__ aff9g2 load_aff9g2(const fe* __restrict__ bases, uint32_t idx) {
  const fe* b = bases + 4 * (uint64_t)idx;
  aff9g2 a;
  a.x.c0 = unpack9(load_fe(b));
  a.x.c1 = unpack9(load_fe(b + 1));
  a.y.c0 = unpack9(load_fe(b + 2));
  a.y.c1 = unpack9(load_fe(b + 3));
  return a;
}

// The point kinds of the accumulation, fix-up and digit-split reduction:
// the bases' affine form, the XYZZ accumulator and its stored form, the
// group law, and the export for the host combination (R = 2^256 canonical
// XYZZ words; the identity as ZZ = ZZZ = 0).
struct PtG1 {
  using Aff = aff9;
  using Acc = xyzz9;
  using Sto = pt9;
  static constexpr uint32_t EXP_FE = 4;  // exported fe per point
  __device__ __forceinline__ static Aff load_base(const fe* b, uint32_t i) { return load_aff9(b, i); }
  __device__ __forceinline__ static bool base_inf(const Aff& a) { return aff9_is_inf(a); }
  __device__ __forceinline__ static void madd(Acc& p, bool& inf, const Aff& a, bool neg) { xyzz9_madd_t(p, inf, a, neg); }
  __device__ __forceinline__ static Acc inf() { return xyzz9_inf(); }
  __device__ __forceinline__ static bool is_inf(const Acc& p) { return xyzz9_is_inf(p); }
  __device__ __forceinline__ static Acc add(const Acc& p, const Acc& q) { return xyzz9_add(p, q); }
  __device__ __forceinline__ static Acc dbl(const Acc& p) { return xyzz9_dbl(p); }
  __device__ __forceinline__ static Acc load(const Sto* s) { return load_xyzz9(s); }
  __device__ __forceinline__ static void store(Sto* s, const Acc& p) { store_xyzz9(s, p); }
  __device__ __forceinline__ static void export_to(fe* out, const Acc& p) {
    xyzz r;
    if (xyzz9_is_inf(p)) {
      r.x = Fq::one();
      r.y = Fq::one();
      r.zz = Fq::zero();
      r.zzz = Fq::zero();
    } else {
      r = xyzz_from9(p);
    }
    store_xyzz(reinterpret_cast<xyzz*>(out), r);
  }
};

struct PtG2 {
  using Aff = aff9g2;
  using Acc = xyzz9g2;
  using Sto = pt18;
  static constexpr uint32_t EXP_FE = 8;
  __device__ __forceinline__ static Aff load_base(const fe* b, uint32_t i) { return load_aff9g2(b, i); }
  __device__ __forceinline__ static bool base_inf(const Aff& a) { return aff9g2_is_inf(a); }
  __device__ __forceinline__ static void madd(Acc& p, bool& inf, const Aff& a, bool neg) { xyzz9g2_madd_t(p, inf, a, neg); }
  __device__ __forceinline__ static Acc inf() { return xyzz9g2_inf(); }
  __device__ __forceinline__ static bool is_inf(const Acc& p) { return xyzz9g2_is_inf(p); }
  __device__ __forceinline__ static Acc add(const Acc& p, const Acc& q) { return xyzz9g2_add(p, q); }
  __device__ __forceinline__ static Acc dbl(const Acc& p) { return xyzz9g2_dbl(p); }
  __device__ __forceinline__ static Acc load(const Sto* s) { return load_xyzz9g2(s); }
  __device__ __forceinline__ static void store(Sto* s, const Acc& p) { store_xyzz9g2(s, p); }
  __device__ __forceinline__ static void export_to(fe* out, const Acc& p) {
    const bool inf = xyzz9g2_is_inf(p);
#pragma unroll
    for (int k = 0; k < 8; k++)
      store_fe(out + k, inf ? Fq::zero() : pack9(Fq9::canonical(Fq9::to256(g2_coord(p, k)))));
  }
};

// wire bases (affine, R = 2^256, x = y = 0 for infinity) -> R = 2^261 form
__global__ __launch_bounds__(256) void msm_bases_convert_kernel(const fe* __restrict__ in, uint64_t n,
                                                                fe* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const aff9 a = aff9_from_wire(load_fe(in + 2 * i), load_fe(in + 2 * i + 1));
  store_fe(out + 2 * i, pack9(Fq9::canonical(a.x)));
  store_fe(out + 2 * i + 1, pack9(Fq9::canonical(a.y)));
}

// window sums: limb form -> R = 2^256 canonical XYZZ for the host combination
template <class P>
__global__ void msm_wsum_export_kernel(const typename P::Sto* __restrict__ in, uint32_t W, fe* __restrict__ out) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  P::export_to(out + (uint64_t)w * P::EXP_FE, P::load(in + w));
}


// 1. scalar (Montgomery) -> canonical -> W signed c-bit digits
//    digits[w * n + i] = |d| | (d < 0 ? SIGN : 0)
template <class Sink>
__device__ __forceinline__ void signed_digits(const fe& s, uint32_t c, uint32_t W, Sink&& sink) {
  uint64_t buf = (uint64_t)s.w[0] | ((uint64_t)s.w[1] << 32);
  int have = 64;
  uint32_t nx[6];
#pragma unroll
  for (int k = 0; k < 6; k++) nx[k] = s.w[k + 2];
  const uint32_t mask = (1u << c) - 1;
  const uint32_t half = 1u << (c - 1);
  uint32_t carry = 0;
  for (uint32_t w = 0; w < W; w++) {
    if (have < (int)c) {
      buf |= (uint64_t)nx[0] << have;
      have += 32;
#pragma unroll
      for (int k = 0; k < 5; k++) nx[k] = nx[k + 1];
      nx[5] = 0;
    }
    const uint32_t val = (uint32_t)(buf & mask) + carry;
    buf >>= c;
    have -= (int)c;
    uint32_t out;
    if (val >= half) {
      const uint32_t mag = (1u << c) - val;  // digit = val - 2^c
      carry = 1;
      out = mag ? (mag | SIGN) : 0u;
    } else {
      carry = 0;
      out = val;
    }
    sink(w, out);
  }
}

__global__ __launch_bounds__(256) void msm_digits_kernel(const fe* __restrict__ scalars, uint64_t n,
                                                         uint32_t c, uint32_t W,
                                                         uint32_t* __restrict__ digits) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  signed_digits(Fr::from_mont(load_fe(scalars + i)), c, W,
                [&](uint32_t w, uint32_t out) { digits[(uint64_t)w * n + i] = out; });
}

// 2. bucket sort of the (row, |d|) pairs in two levels.  A row is one window
//    of a plain MSM, or one whole MSM of a fixed-base batch (all its windows
//    share one bucket set).  MsmRow describes where a row's digits live and
//    how a position p inside the row maps to a base index:
//      base = (p / inner) * stride + boff + p % inner.
//    Bins are |d| - 1 (zero digits are dropped).  A bin splits into a coarse
//    part k = bin >> S (at most SORT_MAX_COARSE values) and a fine part
//    f = bin & (2^S - 1), so that every level works with few counters:
//      a. count: each tile of SORT_TILE digits of a row histograms its coarse
//         keys in LDS into hist[col(r, k) + tile], col(r, k) =
//         tile0(r) NC + k ntiles(r), i.e. columns in (row, coarse key) order
//         with the row's tiles contiguous;
//      b. one exclusive scan of the whole histogram gives every (row, coarse
//         key, tile) its output offset (the sorted order is (row, k, f));
//      c. scatter: a tile ranks its digits per coarse key in LDS, stages them
//         in coarse-key order and writes runs of consecutive slots, so global
//         writes are contiguous per (tile, k) instead of one 4-byte store per
//         bucket;
//      d. fine: each (row, k) segment is cut into chunks; the chunks count
//         their fine keys (count / scan / scatter again, so one segment that
//         holds a huge bucket is spread over many workgroups), write the CSR
//         row pointers of the segment's 2^S buckets and place each entry at
//         its bucket position.
static constexpr uint32_t SORT_TILE = 1u << 13;
static constexpr uint32_t SORT_THREADS = 512;
static constexpr uint32_t SORT_ITEMS = SORT_TILE / SORT_THREADS;
static constexpr uint32_t SORT_MAX_COARSE = 2 * SORT_THREADS;
static constexpr uint32_t SORT_MAX_FINE_BITS = 13;   // 2^13 u32 LDS counters = 32 KiB
static constexpr uint32_t FINE_THREADS = 256;
static constexpr uint32_t FINE_CHUNK = 1u << 14;   // entries per fine-level workgroup (at least)

struct MsmRow {
  uint64_t off;     // first digit of the row
  uint32_t len;     // digits in the row
  uint32_t inner;   // digits per window inside the row
  uint32_t stride;  // base-index stride between windows
  uint32_t ntiles;  // ceil(len / SORT_TILE)
  uint32_t tile0;   // first tile slot of the row
  uint32_t boff;    // base-index offset of the row (fixed-base ranges)
};

__global__ __launch_bounds__(SORT_THREADS) void msm_sort_count_kernel(const uint32_t* __restrict__ digits,
                                                                      const MsmRow* __restrict__ rows,
                                                                      uint32_t S, uint32_t NC,
                                                                      uint32_t* __restrict__ hist) {
  __shared__ uint32_t cnt[SORT_MAX_COARSE];
  const MsmRow row = rows[blockIdx.y];
  const uint32_t tile = blockIdx.x;
  if (tile >= row.ntiles) return;
  for (uint32_t k = threadIdx.x; k < NC; k += SORT_THREADS) cnt[k] = 0;
  __syncthreads();
  const uint32_t p0 = tile * SORT_TILE;
  const uint32_t p1 = (p0 + SORT_TILE < row.len) ? p0 + SORT_TILE : row.len;
  for (uint32_t p = p0 + threadIdx.x; p < p1; p += SORT_THREADS) {
    const uint32_t mag = digits[row.off + p] & ~SIGN;
    if (mag) atomicAdd(&cnt[(mag - 1) >> S], 1u);
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < NC; k += SORT_THREADS)
    hist[(uint64_t)row.tile0 * NC + (uint64_t)k * row.ntiles + tile] = cnt[k];
}

// the same digits, one workgroup per sort tile (SORT_TILE scalars), which
// also counts the tile's coarse keys of every window in LDS and writes the
// coarse histogram the sort scans (window w is row w, tile0 = w * nt), so
// the sort does not read the digits once more just to count them
static constexpr uint32_t DIGITS_COUNT_THREADS = 512;
__global__ __launch_bounds__(DIGITS_COUNT_THREADS) void msm_digits_count_kernel(
    const fe* __restrict__ scalars, uint64_t n, uint32_t c, uint32_t W, uint32_t S, uint32_t NC, uint32_t nt,
    uint32_t* __restrict__ digits, uint32_t* __restrict__ hist) {
  extern __shared__ uint32_t wcnt[];  // W x NC
  const uint32_t tid = threadIdx.x;
  for (uint32_t k = tid; k < W * NC; k += DIGITS_COUNT_THREADS) wcnt[k] = 0;
  __syncthreads();
  const uint64_t p0 = (uint64_t)blockIdx.x * SORT_TILE;
  for (uint32_t j = tid; j < SORT_TILE; j += DIGITS_COUNT_THREADS) {
    const uint64_t i = p0 + j;
    if (i >= n) break;
    signed_digits(Fr::from_mont(load_fe(scalars + i)), c, W, [&](uint32_t w, uint32_t out) {
      digits[(uint64_t)w * n + i] = out;
      const uint32_t mag = out & ~SIGN;
      if (mag) atomicAdd(&wcnt[w * NC + ((mag - 1) >> S)], 1u);
    });
  }
  __syncthreads();
  for (uint32_t idx = tid; idx < W * NC; idx += DIGITS_COUNT_THREADS) {
    const uint32_t w = idx / NC, k = idx - w * NC;
    hist[((uint64_t)w * nt) * NC + (uint64_t)k * nt + blockIdx.x] = wcnt[idx];
  }
}

// exclusive scan building blocks (a hierarchical scan over any length)
static constexpr int SCAN_THREADS = 256;
static constexpr int SCAN_PER_THREAD = 8;
static constexpr int SCAN_CHUNK = SCAN_THREADS * SCAN_PER_THREAD;

// (A wave-shuffle form with two barriers measured slower in the bucket sort:
// its six dependent lane permutes cost more than the barriers they save.)
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh, uint32_t& total) {
  const uint32_t tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (uint32_t off = 1; off < blockDim.x; off <<= 1) {
    uint32_t x = (tid >= off) ? sh[tid - off] : 0;
    __syncthreads();
    sh[tid] += x;
    __syncthreads();
  }
  total = sh[blockDim.x - 1];
  const uint32_t incl = sh[tid];
  __syncthreads();
  return incl - v;
}

// out[i] = exclusive prefix of in[] within its chunk of SCAN_CHUNK; the
// chunk total goes to sums[chunk]
// (in and out may be the same buffer: a block loads its chunk before it stores)
__global__ __launch_bounds__(SCAN_THREADS) void scan_chunks_kernel(const uint32_t* in, uint32_t* out,
                                                                   uint32_t* sums, uint64_t len) {
  __shared__ uint32_t sh[SCAN_THREADS];
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_CHUNK + threadIdx.x * SCAN_PER_THREAD;
  uint32_t v[SCAN_PER_THREAD];
  uint32_t local = 0;
#pragma unroll
  for (int k = 0; k < SCAN_PER_THREAD; k++) {
    v[k] = (base + k < len) ? in[base + k] : 0;
    local += v[k];
  }
  uint32_t total;
  uint32_t run = block_exclusive_scan(local, sh, total);
#pragma unroll
  for (int k = 0; k < SCAN_PER_THREAD; k++) {
    if (base + k < len) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// out[i] += sums[chunk of i]; out[len] = sums[number of chunks] (the total)
__global__ __launch_bounds__(SCAN_THREADS) void scan_add_kernel(uint32_t* out, const uint32_t* sums,
                                                                uint64_t len) {
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_CHUNK;
  const uint32_t add = sums[blockIdx.x];
  if (blockIdx.x == 0 && threadIdx.x == 0) out[len] = sums[gridDim.x];
  for (uint32_t k = threadIdx.x; k < SCAN_CHUNK; k += SCAN_THREADS) {
    const uint64_t idx = base + k;
    if (idx < len) out[idx] += add;
  }
}

// c. scatter of one tile into coarse-key order.  Packed form (PK): the
//    fine key rides in bits [ps, 31) of the entry (ps = 31 - S, used when
//    every base index of the launch is below 2^ps), so no separate u16
//    array is written, staged or read by the fine level.  KT / NCX: the
//    staged coarse key's type and the key arrays' size; with at most 256
//    coarse keys (every window up to 22 bits) u8 keys and 256-entry arrays
//    keep the kernel at 45 KiB of LDS, three workgroups per CU (1 024
//    threads of 8 digits, two workgroups and 32 waves per CU, measured no
//    faster).
template <bool PK, typename KT, uint32_t NCX>
__global__ __launch_bounds__(SORT_THREADS) void msm_sort_scatter_kernel(
    const uint32_t* __restrict__ digits, const MsmRow* __restrict__ rows, uint32_t S, uint32_t NC,
    const uint32_t* __restrict__ hoff, uint32_t* __restrict__ items_v, uint16_t* __restrict__ items_f) {
  __shared__ uint32_t cnt[NCX];
  __shared__ uint32_t cstart[NCX];
  __shared__ uint32_t gbase[NCX];
  __shared__ uint32_t stage_v[SORT_TILE];
  __shared__ uint16_t stage_f[PK ? 1 : SORT_TILE];
  __shared__ KT stage_k[SORT_TILE];
  __shared__ uint32_t scan_sh[SORT_THREADS];
  const MsmRow row = rows[blockIdx.y];
  const uint32_t tile = blockIdx.x;
  if (tile >= row.ntiles) return;
  const uint32_t tid = threadIdx.x;
  for (uint32_t k = tid; k < NC; k += SORT_THREADS) cnt[k] = 0;
  __syncthreads();
  const uint32_t p0 = tile * SORT_TILE;
  const uint32_t fmask = (1u << S) - 1;
  const uint32_t ps = 31 - S;
  uint32_t d[SORT_ITEMS], rank[SORT_ITEMS];
#pragma unroll
  for (uint32_t i = 0; i < SORT_ITEMS; i++) {
    const uint32_t p = p0 + i * SORT_THREADS + tid;
    d[i] = (p < row.len) ? digits[row.off + p] : 0u;
    const uint32_t mag = d[i] & ~SIGN;
    rank[i] = mag ? atomicAdd(&cnt[(mag - 1) >> S], 1u) : 0u;
  }
  __syncthreads();
  // exclusive scan of the NC coarse counts, two per thread (NC <= 2 SORT_THREADS)
  uint32_t total;
  const uint32_t k0 = 2 * tid, k1 = 2 * tid + 1;
  const uint32_t a = (k0 < NC) ? cnt[k0] : 0u, b = (k1 < NC) ? cnt[k1] : 0u;
  const uint32_t ex = block_exclusive_scan(a + b, scan_sh, total);
  const uint64_t hrow = (uint64_t)row.tile0 * NC + tile;
  if (k0 < NC) {
    cstart[k0] = ex;
    gbase[k0] = hoff[hrow + (uint64_t)k0 * row.ntiles] - ex;
  }
  if (k1 < NC) {
    cstart[k1] = ex + a;
    gbase[k1] = hoff[hrow + (uint64_t)k1 * row.ntiles] - (ex + a);
  }
  __syncthreads();
#pragma unroll
  for (uint32_t i = 0; i < SORT_ITEMS; i++) {
    const uint32_t mag = d[i] & ~SIGN;
    if (!mag) continue;
    const uint32_t p = p0 + i * SORT_THREADS + tid;
    const uint32_t k = (mag - 1) >> S;
    const uint32_t slot = cstart[k] + rank[i];
    const uint32_t wi = p / row.inner;
    const uint32_t idx = (wi * row.stride + row.boff + (p - wi * row.inner)) | (d[i] & SIGN);
    if (PK) {
      stage_v[slot] = idx | (((mag - 1) & fmask) << ps);
    } else {
      stage_v[slot] = idx;
      stage_f[slot] = (uint16_t)((mag - 1) & fmask);
    }
    stage_k[slot] = (KT)k;
  }
  __syncthreads();
  for (uint32_t slot = tid; slot < total; slot += SORT_THREADS) {
    const uint32_t g = gbase[stage_k[slot]] + slot;
    items_v[g] = stage_v[slot];
    if (!PK) items_f[g] = stage_f[slot];
  }
}

// d. fine level.  Segment g = (row r, coarse key k) holds the entries
//    [s0, s1) of the coarse order, and the same range of the sorted order.
//    A segment one workgroup can hold (packed entries) is placed whole by
//    msm_fine_fused_kernel.  Any other segment is cut into chunks of
//    FINE_CHUNK entries (at least one chunk, so empty segments still write
//    their row pointers); chunk j of g is one workgroup, which counts the
//    fine keys in LDS into fhist[fc0(g) NF + f nch(g) + j] (columns in
//    (segment, f) order, chunks contiguous).  One exclusive scan of fhist
//    then gives the offset of every (segment, f, chunk) among the chunked
//    segments' entries; rebased on s0 it is the output position, and the
//    bucket start of (r, bin = k 2^S + f) is the value at chunk 0.  A
//    segment holding one huge bucket is thus spread over many workgroups.
__device__ __forceinline__ void fine_segment(const uint32_t* __restrict__ hoff, const MsmRow* __restrict__ rows,
                                             uint32_t NC, uint32_t g, uint32_t& r, uint32_t& k, uint32_t& s0,
                                             uint32_t& s1) {
  r = g / NC;
  k = g - r * NC;
  const MsmRow row = rows[r];
  const uint64_t col = (uint64_t)row.tile0 * NC + (uint64_t)k * row.ntiles;
  s0 = hoff[col];
  s1 = hoff[col + row.ntiles];  // the next column's start (the grand total after the last)
}

// Segments of at most FUSE_MAX entries are placed whole by one workgroup of
// msm_fine_fused_kernel (fused_max = FUSE_MAX) and get no chunks here; with
// fused_max = 0 every segment is chunked.
__global__ __launch_bounds__(256) void msm_fine_chunks_kernel(const uint32_t* __restrict__ hoff,
                                                              const MsmRow* __restrict__ rows, uint32_t NC,
                                                              uint32_t nseg, uint32_t fch, uint32_t fused_max,
                                                              uint32_t* __restrict__ nch,
                                                              uint32_t* __restrict__ maxrun) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g == 0) *maxrun = 0;  // longest head run of this launch sequence, raised by the accumulation
  if (g >= nseg) return;
  uint32_t r, k, s0, s1;
  fine_segment(hoff, rows, NC, g, r, k, s0, s1);
  if (fused_max && s1 - s0 <= fused_max) {
    nch[g] = 0;
    return;
  }
  const uint32_t m = (s1 - s0 + fch - 1) / fch;
  nch[g] = m ? m : 1u;
}

// workgroup b -> (segment g, chunk j) with fc0[g] <= b < fc0[g + 1]; false
// for the spare workgroups of the upper-bound grid
__device__ __forceinline__ bool fine_chunk(const uint32_t* __restrict__ fc0, uint32_t nseg, uint32_t b,
                                           uint32_t& g, uint32_t& j, uint32_t& nc) {
  if (b >= fc0[nseg]) return false;
  uint32_t lo = 0, hi = nseg;  // fc0[lo] <= b < fc0[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (fc0[mid] <= b) lo = mid; else hi = mid;
  }
  g = lo;
  j = b - fc0[g];
  nc = fc0[g + 1] - fc0[g];
  return true;
}

// fine key of coarse-ordered entry p (packed form: bits [31 - S, 31) of the entry)
template <bool PK>
__device__ __forceinline__ uint32_t fine_key(const uint32_t* __restrict__ items_v,
                                             const uint16_t* __restrict__ items_f, uint32_t p, uint32_t S) {
  if (PK) return (items_v[p] >> (31 - S)) & ((1u << S) - 1);
  return items_f[p];
}
// the entry as the accumulation reads it (base index | sign)
template <bool PK>
__device__ __forceinline__ uint32_t entry_value(uint32_t v, uint32_t S) {
  return PK ? (v & ~(((1u << S) - 1) << (31 - S))) : v;
}

// Fine level of one whole segment in one workgroup (packed entries): the
// segment's entries are read once into registers, ranked per fine key in
// LDS, staged in bucket order and written out as one contiguous run, and the
// workgroup writes the CSR row pointers of the segment's buckets itself, so
// no chunk histogram has to be written, scanned and read back.  At 2^22
// points every coarse segment (about n / 256 entries) takes this path.
static constexpr uint32_t FUSE_THREADS = 1024;
static constexpr uint32_t FUSE_ITEMS = 17;
static constexpr uint32_t FUSE_MAX = FUSE_THREADS * FUSE_ITEMS;  // 17 408 entries, 68 KiB staged
static constexpr uint32_t FUSE_KEYS = FUSE_THREADS;             // one fine key per thread

__global__ __launch_bounds__(FUSE_THREADS) void msm_fine_fused_kernel(
    const uint32_t* __restrict__ items_v, const uint32_t* __restrict__ hoff, const MsmRow* __restrict__ rows,
    uint32_t S, uint32_t NC, uint32_t nseg, uint32_t nbins, uint32_t B, uint32_t* __restrict__ row_ptr,
    uint32_t* __restrict__ sorted) {
  __shared__ uint32_t lcnt[FUSE_KEYS];
  __shared__ uint32_t stage_v[FUSE_MAX];
  __shared__ uint32_t scan_sh[FUSE_THREADS];
  const uint32_t g = blockIdx.x;
  uint32_t r, k, s0, s1;
  fine_segment(hoff, rows, NC, g, r, k, s0, s1);
  const uint32_t cnt = s1 - s0;
  if (cnt > FUSE_MAX) return;  // placed by the chunked kernels
  const uint32_t NF = 1u << S;
  const uint32_t tid = threadIdx.x;
  if (tid < NF) lcnt[tid] = 0;
  __syncthreads();
  uint32_t v[FUSE_ITEMS], rank[FUSE_ITEMS];
#pragma unroll
  for (uint32_t i = 0; i < FUSE_ITEMS; i++) {
    const uint32_t q = i * FUSE_THREADS + tid;
    if (q < cnt) {
      v[i] = items_v[s0 + q];
      rank[i] = atomicAdd(&lcnt[(v[i] >> (31 - S)) & (NF - 1)], 1u);
    }
  }
  __syncthreads();
  // bucket starts inside the segment: exclusive scan of the counts
  uint32_t total;
  const uint32_t c0 = (tid < NF) ? lcnt[tid] : 0u;
  const uint32_t ex = block_exclusive_scan(c0, scan_sh, total);
  if (tid < NF) {
    lcnt[tid] = ex;
    if ((k << S) + tid < nbins) row_ptr[(uint64_t)r * B + 1 + ((uint64_t)k << S) + tid] = s0 + ex;
  }
  if (k == 0 && tid == 0) row_ptr[(uint64_t)r * B] = s0;
  if (g == nseg - 1 && tid == 0) row_ptr[(uint64_t)(nseg / NC) * B] = s1;  // the total
  __syncthreads();
#pragma unroll
  for (uint32_t i = 0; i < FUSE_ITEMS; i++) {
    const uint32_t q = i * FUSE_THREADS + tid;
    if (q < cnt) stage_v[lcnt[(v[i] >> (31 - S)) & (NF - 1)] + rank[i]] = v[i];
  }
  __syncthreads();
  for (uint32_t slot = tid; slot < cnt; slot += FUSE_THREADS) sorted[s0 + slot] = entry_value<true>(stage_v[slot], S);
}

template <bool PK>
__global__ __launch_bounds__(FINE_THREADS) void msm_fine_count_kernel(
    const uint32_t* __restrict__ items_v, const uint16_t* __restrict__ items_f,
    const uint32_t* __restrict__ hoff, const MsmRow* __restrict__ rows,
    uint32_t S, uint32_t NC, uint32_t nseg, uint32_t fch, const uint32_t* __restrict__ fc0,
    uint32_t* __restrict__ fhist) {
  extern __shared__ uint32_t fcnt[];
  uint32_t g, j, nc;
  if (!fine_chunk(fc0, nseg, blockIdx.x, g, j, nc)) return;
  uint32_t r, k, s0, s1;
  fine_segment(hoff, rows, NC, g, r, k, s0, s1);
  const uint32_t NF = 1u << S;
  for (uint32_t f = threadIdx.x; f < NF; f += FINE_THREADS) fcnt[f] = 0;
  __syncthreads();
  const uint32_t a = s0 + j * fch;
  const uint32_t e = (s1 - a > fch) ? a + fch : s1;
  for (uint32_t p = a + threadIdx.x; p < e; p += FINE_THREADS) atomicAdd(&fcnt[fine_key<PK>(items_v, items_f, p, S)], 1u);
  __syncthreads();
  uint32_t* out = fhist + (uint64_t)fc0[g] * NF + j;
  for (uint32_t f = threadIdx.x; f < NF; f += FINE_THREADS) out[(uint64_t)f * nc] = fcnt[f];
}

// places the chunk's entries at their bucket positions; chunk 0 of a
// segment writes the CSR row pointers of its buckets (bucket |d| = bin + 1,
// key = row B + |d|; bucket 0 of each row is empty)
template <bool PK>
__global__ __launch_bounds__(FINE_THREADS) void msm_fine_scatter_kernel(
    const uint32_t* __restrict__ items_v, const uint16_t* __restrict__ items_f,
    const uint32_t* __restrict__ hoff, const MsmRow* __restrict__ rows, uint32_t S, uint32_t NC,
    uint32_t nseg, uint32_t fch, const uint32_t* __restrict__ fc0, const uint32_t* __restrict__ foff,
    uint32_t nbins, uint32_t B, uint32_t* __restrict__ row_ptr, uint32_t* __restrict__ sorted) {
  extern __shared__ uint32_t fpos[];
  uint32_t g, j, nc;
  if (!fine_chunk(fc0, nseg, blockIdx.x, g, j, nc)) return;
  uint32_t r, k, s0, s1;
  fine_segment(hoff, rows, NC, g, r, k, s0, s1);
  const uint32_t NF = 1u << S;
  const uint32_t* in = foff + (uint64_t)fc0[g] * NF + j;
  const uint32_t rebase = s0 - foff[(uint64_t)fc0[g] * NF];  // the scan counts chunked segments only
  for (uint32_t f = threadIdx.x; f < NF; f += FINE_THREADS) fpos[f] = in[(uint64_t)f * nc] + rebase;
  __syncthreads();
  if (j == 0) {
    const uint64_t key0 = (uint64_t)r * B + 1 + ((uint64_t)k << S);
    for (uint32_t f = threadIdx.x; f < NF; f += FINE_THREADS)
      if (((k << S) + f) < nbins) row_ptr[key0 + f] = fpos[f];
    if (k == 0 && threadIdx.x == 0) row_ptr[(uint64_t)r * B] = s0;
    if (g == nseg - 1 && threadIdx.x == 0) row_ptr[(uint64_t)(nseg / NC) * B] = s1;  // the total
  }
  __syncthreads();  // the bucket starts are read before the placement below advances them
  const uint32_t a = s0 + j * fch;
  const uint32_t e = (s1 - a > fch) ? a + fch : s1;
  for (uint32_t p = a + threadIdx.x; p < e; p += FINE_THREADS) {
    const uint32_t pos = atomicAdd(&fpos[fine_key<PK>(items_v, items_f, p, S)], 1u);
    sorted[pos] = entry_value<PK>(items_v[p], S);
  }
}

// Staged form of the placement for up to FINE_STAGE_KEYS fine keys: the
// chunk (at most FINE_STAGE entries) is ranked in LDS, staged in fine-key
// order and written out by consecutive threads, so the global writes are
// contiguous runs per fine key instead of one scattered 4-byte store per
// entry.
static constexpr uint32_t FINE_STAGE = 8192;
static constexpr uint32_t FINE_STAGE_THREADS = 512;
static constexpr uint32_t FINE_STAGE_ITEMS = FINE_STAGE / FINE_STAGE_THREADS;
static constexpr uint32_t FINE_STAGE_KEYS = 2 * FINE_STAGE_THREADS;

template <bool PK>
__global__ __launch_bounds__(FINE_STAGE_THREADS) void msm_fine_scatter_staged_kernel(
    const uint32_t* __restrict__ items_v, const uint16_t* __restrict__ items_f,
    const uint32_t* __restrict__ hoff, const MsmRow* __restrict__ rows, uint32_t S, uint32_t NC,
    uint32_t nseg, uint32_t fch, const uint32_t* __restrict__ fc0, const uint32_t* __restrict__ foff,
    uint32_t nbins, uint32_t B, uint32_t* __restrict__ row_ptr, uint32_t* __restrict__ sorted) {
  __shared__ uint32_t gpos[FINE_STAGE_KEYS];
  __shared__ uint32_t lcnt[FINE_STAGE_KEYS];
  __shared__ uint32_t stage_v[FINE_STAGE];
  __shared__ uint16_t stage_f[PK ? 1 : FINE_STAGE];
  __shared__ uint32_t scan_sh[FINE_STAGE_THREADS];
  uint32_t g, j, nc;
  if (!fine_chunk(fc0, nseg, blockIdx.x, g, j, nc)) return;
  uint32_t r, k, s0, s1;
  fine_segment(hoff, rows, NC, g, r, k, s0, s1);
  const uint32_t NF = 1u << S;
  const uint32_t tid = threadIdx.x;
  const uint32_t* in = foff + (uint64_t)fc0[g] * NF + j;
  const uint32_t rebase = s0 - foff[(uint64_t)fc0[g] * NF];  // the scan counts chunked segments only
  for (uint32_t f = tid; f < NF; f += FINE_STAGE_THREADS) {
    gpos[f] = in[(uint64_t)f * nc] + rebase;
    lcnt[f] = 0;
  }
  __syncthreads();
  if (j == 0) {
    const uint64_t key0 = (uint64_t)r * B + 1 + ((uint64_t)k << S);
    for (uint32_t f = tid; f < NF; f += FINE_STAGE_THREADS)
      if (((k << S) + f) < nbins) row_ptr[key0 + f] = gpos[f];
    if (k == 0 && tid == 0) row_ptr[(uint64_t)r * B] = s0;
    if (g == nseg - 1 && tid == 0) row_ptr[(uint64_t)(nseg / NC) * B] = s1;  // the total
  }
  const uint32_t a = s0 + j * fch;
  const uint32_t cnt = (s1 - a > fch) ? fch : s1 - a;
  uint32_t v[FINE_STAGE_ITEMS], f[FINE_STAGE_ITEMS], rank[FINE_STAGE_ITEMS];
#pragma unroll
  for (uint32_t i = 0; i < FINE_STAGE_ITEMS; i++) {
    const uint32_t q = i * FINE_STAGE_THREADS + tid;
    if (q < cnt) {
      v[i] = items_v[a + q];
      f[i] = PK ? (v[i] >> (31 - S)) & (NF - 1) : items_f[a + q];
      rank[i] = atomicAdd(&lcnt[f[i]], 1u);
    }
  }
  __syncthreads();
  // exclusive scan of the local counts, two keys per thread
  uint32_t total;
  const uint32_t k0 = 2 * tid, k1 = 2 * tid + 1;
  const uint32_t c0 = (k0 < NF) ? lcnt[k0] : 0u, c1 = (k1 < NF) ? lcnt[k1] : 0u;
  const uint32_t ex = block_exclusive_scan(c0 + c1, scan_sh, total);
  __syncthreads();
  if (k0 < NF) lcnt[k0] = ex;
  if (k1 < NF) lcnt[k1] = ex + c0;
  __syncthreads();
#pragma unroll
  for (uint32_t i = 0; i < FINE_STAGE_ITEMS; i++) {
    const uint32_t q = i * FINE_STAGE_THREADS + tid;
    if (q < cnt) {
      const uint32_t slot = lcnt[f[i]] + rank[i];
      stage_v[slot] = v[i];
      if (!PK) stage_f[slot] = (uint16_t)f[i];
    }
  }
  __syncthreads();
  for (uint32_t slot = tid; slot < cnt; slot += FINE_STAGE_THREADS) {
    const uint32_t sv = stage_v[slot];
    const uint32_t ff = PK ? (sv >> (31 - S)) & (NF - 1) : stage_f[slot];
    sorted[gpos[ff] + (slot - lcnt[ff])] = entry_value<PK>(sv, S);
  }
}

// 4. bucket accumulation over fixed-length segments of the sorted list.
//    A run (bucket) that starts inside the segment is owned by it and its
//    partial sum goes straight to buckets[key]; the run the segment starts
//    in the middle of is its "head" partial, added in by the fix-up pass.
template <class P>
__global__ __launch_bounds__(256) void msm_accumulate_kernel(
    const fe* __restrict__ bases, const uint32_t* __restrict__ sorted,
    const uint32_t* __restrict__ row_ptr, uint32_t nkeys, uint32_t L, typename P::Sto* __restrict__ buckets,
    uint32_t* __restrict__ head_key, typename P::Sto* __restrict__ head, uint32_t* __restrict__ maxrun) {
  const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t M = row_ptr[nkeys];
  const uint64_t start = s * L;
  if (start >= M) return;
  const uint64_t end = (start + L < M) ? start + L : M;
  // largest k with row_ptr[k] <= start  (row_ptr[nkeys] = M > start)
  uint32_t lo = 0, hi = nkeys;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (row_ptr[mid] <= start) lo = mid; else hi = mid;
  }
  uint32_t k = lo;
  uint32_t rbeg = row_ptr[k], rend = row_ptr[k + 1];
  uint32_t hk = NO_KEY;
  typename P::Acc acc = P::inf();
  bool acc_inf = true;
  // a head partial of bucket k: the run of heads it belongs to spans `len`
  // segments; the fix-up rounds only have work for runs longer than their stride
  auto note_run = [&](uint32_t a, uint32_t b) {
    const uint32_t len = (b - 1) / L - a / L;
    if (len > 1) atomicMax(maxrun, len);
  };
  for (uint64_t p = start; p < end; p++) {
    while (p >= rend) {
      if (rend > rbeg) {
        if (rbeg < start) {
          hk = k;
          P::store(head + s, acc);
          note_run(rbeg, rend);
        } else {
          P::store(buckets + k, acc);
        }
      }
      acc = P::inf();
      acc_inf = true;
      k++;
      rbeg = rend;
      rend = row_ptr[k + 1];
    }
    const uint32_t v = sorted[p];
    const typename P::Aff a = P::load_base(bases, v & ~SIGN);
    if (!P::base_inf(a)) P::madd(acc, acc_inf, a, (v & SIGN) != 0);
  }
  if (rbeg < start) {
    hk = k;
    P::store(head + s, acc);
    note_run(rbeg, rend);
  } else {
    P::store(buckets + k, acc);
  }
  head_key[s] = hk;
}

// 5. fix-up of buckets that span several segments.  The head partials of
//    bucket k sit in consecutive segments first..first+len-1 with
//    first = row_ptr[k] / L + 1.  Round r folds F partials spaced F^r apart
//    into the one at offset 0 mod F^(r+1); after the launched rounds the
//    run's sum is at `first` and is added to buckets[k] (which the owning
//    segment wrote).  F is chosen on the device from the longest run the
//    accumulation recorded (fixup_fan), so a chain of rounds (F - 1)
//    dependent additions instead of len: a few huge buckets (the top window
//    when c does not divide the scalar bits, equal scalars) fold in a fraction
//    of the time a fixed fan-in of FIXUP_FAN took.
static constexpr uint32_t FIXUP_FAN = 32;

__device__ __forceinline__ bool head_run(const uint32_t* __restrict__ row_ptr, uint32_t k, uint32_t L,
                                         uint64_t s, uint64_t& first, uint64_t& len) {
  const uint64_t a = row_ptr[k], b = row_ptr[k + 1];
  first = a / L + 1;
  len = (b - 1) / L - a / L;
  return s >= first && s < first + len;
}

// The fan-in of `rounds` launched rounds: the smallest F >= 2 with F^rounds
// >= mr (the longest run), so the rounds fold every run and the chain of
// dependent additions, rounds (F - 1), is as short as they allow.  The host
// launches ceil(log_FIXUP_FAN(segments)) rounds, so F <= FIXUP_FAN.
__device__ __forceinline__ uint64_t fixup_fan(uint32_t mr, uint32_t rounds) {
  if (rounds >= 32 || mr <= (1u << rounds)) return 2;  // uniform scalars: runs of two heads at most
  auto covers = [&](uint64_t f) {
    uint64_t p = 1;
    for (uint32_t i = 0; i < rounds; i++) {
      p *= f;
      if (p >= mr) return true;
    }
    return false;
  };
  uint64_t lo = 2, hi = FIXUP_FAN;  // covers(FIXUP_FAN) holds: mr < segments <= FIXUP_FAN^rounds
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (covers(mid)) hi = mid; else lo = mid + 1;
  }
  return lo;
}

template <class P>
__global__ __launch_bounds__(256) void msm_fixup_round_kernel(const uint32_t* __restrict__ row_ptr,
                                                              uint32_t nkeys, uint32_t L, uint32_t round,
                                                              uint32_t rounds,
                                                              const uint32_t* __restrict__ head_key,
                                                              typename P::Sto* __restrict__ head,
                                                              const uint32_t* __restrict__ maxrun) {
  const uint32_t mr = *maxrun;
  // F >= 2, so the stride F^round is at least 2^round: no run longer than
  // that means nothing to fold (the common case, checked before anything else)
  if (mr <= (1u << round)) return;
  const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t M = row_ptr[nkeys];
  const uint64_t S = (M + L - 1) / L;
  if (s >= S) return;
  const uint32_t k = head_key[s];
  if (k == NO_KEY) return;
  uint64_t first, len;
  if (!head_run(row_ptr, k, L, s, first, len)) return;
  const uint64_t off = s - first;
  const uint64_t F = fixup_fan(mr, rounds);
  uint64_t stride = 1;
  for (uint32_t i = 0; i < round; i++) stride *= F;
  if (off % (stride * F) != 0 || off + stride >= len) return;
  typename P::Acc acc = P::load(head + s);
  for (uint32_t j = 1; j < F; j++) {
    const uint64_t t = off + (uint64_t)j * stride;
    if (t >= len) break;
    acc = P::add(acc, P::load(head + first + t));
  }
  P::store(head + s, acc);
}

template <class P>
__global__ __launch_bounds__(256) void msm_fixup_final_kernel(const uint32_t* __restrict__ row_ptr,
                                                              uint32_t nkeys, uint32_t L,
                                                              typename P::Sto* __restrict__ buckets,
                                                              const uint32_t* __restrict__ head_key,
                                                              const typename P::Sto* __restrict__ head) {
  const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t M = row_ptr[nkeys];
  const uint64_t S = (M + L - 1) / L;
  if (s >= S) return;
  const uint32_t k = head_key[s];
  if (k == NO_KEY) return;
  uint64_t first, len;
  if (!head_run(row_ptr, k, L, s, first, len) || s != first) return;
  P::store(buckets + k, P::add(P::load(buckets + k), P::load(head + s)));
}

// 6. bucket reduction, chunked: chunk t of window w covers buckets
//    [lo, lo + K) with lo = 1 + t K; its weighted sum
//      Sum_b b B_b = Sum_b (b - lo + 1) B_b + (lo - 1) Sum_b B_b
//    is a running sum from the top plus one small scalar multiple.  A
//    workgroup (RED_CHUNK chunks of one window: grid (ceil(cpw / RED_CHUNK),
//    W)) then sums its chunks in an LDS tree and writes one partial, so the
//    first level of the per-window tree needs no launch of its own.
static constexpr uint32_t RED_CHUNK = 128;
__global__ __launch_bounds__(RED_CHUNK) void msm_reduce_chunks_kernel(const pt9* __restrict__ buckets,
                                                                      const uint32_t* __restrict__ row_ptr,
                                                                      uint32_t B, uint32_t K, uint32_t cpw,
                                                                      pt9* __restrict__ partial) {
  __shared__ xyzz9 sh[RED_CHUNK];
  const uint32_t w = blockIdx.y, tid = threadIdx.x;
  const uint32_t t = blockIdx.x * RED_CHUNK + tid;
  xyzz9 sum = xyzz9_inf();
  if (t < cpw) {
    const uint32_t lo = 1 + t * K;
    uint32_t hi = lo + K - 1;
    if (hi > B - 1) hi = B - 1;
    xyzz9 run = xyzz9_inf();
    for (uint32_t b = hi; b >= lo; b--) {
      const uint32_t key = w * B + b;
      if (row_ptr[key + 1] > row_ptr[key]) run = xyzz9_add(run, load_xyzz9(buckets + key));
      sum = xyzz9_add(sum, run);
    }
    if (lo > 1) sum = xyzz9_add(sum, xyzz9_mul_small(run, lo - 1));
  }
  sh[tid] = sum;
  __syncthreads();
  for (uint32_t step = RED_CHUNK >> 1; step > 0; step >>= 1) {
    if (tid < step) sh[tid] = xyzz9_add(sh[tid], sh[tid + step]);
    __syncthreads();
  }
  if (tid == 0) store_xyzz9(partial + (uint64_t)w * gridDim.x + blockIdx.x, sh[0]);
}

// 7. per-window tree sum of the chunk partials, in levels: group g of window
//    w sums in[w m + 64 g .. + 64) into out[w G + g], G = ceil(m / 64);
//    the host repeats until one value per window is left
static constexpr uint32_t RED_GROUP = 64;

__global__ __launch_bounds__(RED_GROUP) void msm_reduce_groups_kernel(const pt9* __restrict__ in, uint32_t m,
                                                                     uint32_t G, pt9* __restrict__ out) {
  __shared__ xyzz9 sh[RED_GROUP];
  const uint32_t w = blockIdx.y, g = blockIdx.x, tid = threadIdx.x;
  const uint32_t t = g * RED_GROUP + tid;
  sh[tid] = (t < m) ? load_xyzz9(in + (uint64_t)w * m + t) : xyzz9_inf();
  __syncthreads();
  for (uint32_t step = RED_GROUP >> 1; step > 0; step >>= 1) {
    if (tid < step) sh[tid] = xyzz9_add(sh[tid], sh[tid + step]);
    __syncthreads();
  }
  if (tid == 0) store_xyzz9(out + (uint64_t)w * G + g, sh[0]);
}

// 6b. bucket reduction by digit split, for rows that are the windows of one
//     MSM (the host already doubles between windows, so it can take terms at
//     any bit position).  Row r holds buckets b = 1..M, M = 2^m.  Writing
//     b = h 2^ml + l (ml = ceil(m / 2), mh = m - ml, l < 2^ml):
//       S_r = sum_b b B_b = sum_l l L_l + 2^ml sum_h h H_h,
//       H_h = sum_l B_(h 2^ml + l) for h < 2^mh,   H_(2^mh) = B_M,
//       L_l = sum_(h < 2^mh) B_(h 2^ml + l),
//     and by the bits of l and h
//       S_r = sum_(j = 0..m) 2^j E_j,
//       E_j = sum of L_l over the l with bit j set (j < ml), else of H_h over
//       the h with bit j - ml set.
//     Every bucket enters one H and one L sum: two additions per bucket, as
//     in a running sum, but all of them in plain sums (trees), with no long
//     dependent chains and no per-chunk scalar multiple.  The planes are
//     folded SPLIT_FOLD at a time, G_g = sum_(i < SPLIT_FOLD) 2^i E_(g SPLIT_FOLD + i),
//     and the host adds G_g at bit c r + SPLIT_FOLD g of its Horner combination.
static constexpr uint32_t SPLIT_THREADS = 256;
static constexpr uint32_t SPLIT_WG_PER_CU = 3;  // sums workgroups resident per CU (164 VGPRs: 3 waves per SIMD)
static constexpr uint32_t SPLIT_FOLD = 4;                           // planes per folded term
static constexpr uint32_t SPLIT_PLANE_LANES = 32;                   // threads per plane sum

// blocks [0, nblk_h): H sums, from nblk_h on: L sums.  The LANES partials
// of one sum sit PER = SPLIT_THREADS / LANES threads apart, so the tree below
// retires whole waves (step s keeps the threads below s PER busy) instead of
// running every level in every wave.  LANES = 32 halves each thread's chain
// of dependent additions and doubles the workgroups; the launch takes it
// when the doubled grid still fits one round of resident workgroups.
template <class P, uint32_t LANES>
__global__ __launch_bounds__(SPLIT_THREADS) void msm_split_sums_kernel(const typename P::Sto* __restrict__ buckets,
                                                                       const uint32_t* __restrict__ row_ptr,
                                                                       uint32_t B, uint32_t ml, uint32_t mh,
                                                                       uint32_t nblk_h, typename P::Sto* __restrict__ hsum,
                                                                       typename P::Sto* __restrict__ lsum) {
  constexpr uint32_t PER = SPLIT_THREADS / LANES;
  __shared__ typename P::Acc sh[SPLIT_THREADS];
  const uint32_t r = blockIdx.y, tid = threadIdx.x;
  const uint32_t M = 1u << (ml + mh);
  const bool hrow = blockIdx.x < nblk_h;
  const uint32_t o = (hrow ? blockIdx.x : blockIdx.x - nblk_h) * PER + tid % PER;
  const uint32_t p = tid / PER;
  const uint32_t nout = hrow ? (1u << mh) + 1 : (1u << ml);
  typename P::Acc acc = P::inf();
  if (o < nout) {
    const uint32_t cnt = hrow ? (o < (1u << mh) ? (1u << ml) : 1u) : (1u << mh);
    for (uint32_t i = p; i < cnt; i += LANES) {
      const uint32_t b = hrow ? (o << ml) + i : (i << ml) + o;  // H_(2^mh): o << ml = M
      if (b == 0 || b > M) continue;
      const uint32_t key = r * B + b;
      if (row_ptr[key + 1] > row_ptr[key]) acc = P::add(acc, P::load(buckets + key));
    }
  }
  sh[tid] = acc;
  __syncthreads();
  for (uint32_t step = LANES >> 1; step > 0; step >>= 1) {
    if (p < step) sh[tid] = P::add(sh[tid], sh[tid + step * PER]);
    __syncthreads();
  }
  if (p == 0 && o < nout) P::store((hrow ? hsum : lsum) + (uint64_t)r * nout + o, sh[tid]);
}

// G_g of row r: SPLIT_FOLD plane sums side by side (SPLIT_PLANE_LANES threads
// each), then 2^i-weighted by Horner's rule in one thread
template <class P>
__global__ __launch_bounds__(SPLIT_FOLD* SPLIT_PLANE_LANES) void msm_split_planes_kernel(
    const typename P::Sto* __restrict__ hsum, const typename P::Sto* __restrict__ lsum, uint32_t ml, uint32_t mh, uint32_t ng,
    typename P::Sto* __restrict__ gsum) {
  __shared__ typename P::Acc sh[SPLIT_FOLD * SPLIT_PLANE_LANES];
  const uint32_t r = blockIdx.y, g = blockIdx.x, tid = threadIdx.x;
  const uint32_t q = tid / SPLIT_PLANE_LANES, p = tid % SPLIT_PLANE_LANES;
  const uint32_t j = g * SPLIT_FOLD + q, m = ml + mh;
  typename P::Acc acc = P::inf();
  if (j <= m) {
    const bool low = j < ml;
    const uint32_t jb = low ? j : j - ml;  // the index bit this plane selects
    const uint32_t span = low ? ml : mh;   // index bits of the L / H sums below the top
    const typename P::Sto* src = low ? lsum + (uint64_t)r * (1u << ml) : hsum + (uint64_t)r * ((1u << mh) + 1);
    if (jb < span) {
      // the indices below 2^span with bit jb set: k with a one inserted at bit jb
      const uint32_t cnt = 1u << (span - 1);
      for (uint32_t k = p; k < cnt; k += SPLIT_PLANE_LANES) {
        const uint32_t x = ((k >> jb) << (jb + 1)) | (1u << jb) | (k & ((1u << jb) - 1));
        acc = P::add(acc, P::load(src + x));
      }
    } else if (p == 0) {
      acc = P::load(src + (1u << mh));  // j = m: only H_(2^mh) = B_M has bit mh
    }
  }
  sh[tid] = acc;
  __syncthreads();
  for (uint32_t step = SPLIT_PLANE_LANES >> 1; step > 0; step >>= 1) {
    if (p < step) sh[tid] = P::add(sh[tid], sh[tid + step]);
    __syncthreads();
  }
  if (tid == 0) {
    typename P::Acc G = sh[(SPLIT_FOLD - 1) * SPLIT_PLANE_LANES];
    for (int i = (int)SPLIT_FOLD - 2; i >= 0; i--) G = P::add(P::dbl(G), sh[i * SPLIT_PLANE_LANES]);
    P::store(gsum + (uint64_t)r * ng + g, G);
  }
}

// 8. fixed-base batches: the tables and the digit split.
//    Table row w holds 2^(c w) * B_i (affine, R = 2^261 packed), so every
//    signed digit d_(i,w) of any scalar maps to the same bucket set:
//      sum_i s_i B_i = sum_i sum_w d_(i,w) (2^(c w) B_i).
__device__ __forceinline__ f9 fq9_inv(const f9& a) {
  // a^(q-2), q - 2 as little-endian 32-bit words
  const uint32_t e[8] = {0xd87cfd45u, 0x3c208c16u, 0x6871ca8du, 0x97816a91u,
                         0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};
  f9 r = Fq9::one();
  for (int wi = 7; wi >= 0; wi--) {
    const uint32_t word = e[wi];
    for (int b = 31; b >= 0; b--) {
      r = Fq9::sqr(r);
      if ((word >> b) & 1u) r = Fq9::mul(r, a);
    }
  }
  return r;
}

__global__ __launch_bounds__(128) void msm_fixed_table_kernel(const fe* __restrict__ bases9, uint64_t n,
                                                              uint32_t c, uint32_t W,
                                                              fe* __restrict__ table) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  xyzz9 acc = xyzz9_from_aff(load_aff9(bases9, (uint32_t)i));
  for (uint32_t w = 0; w < W; w++) {
    f9 x = Fq9::zero(), y = Fq9::zero();
    if (!xyzz9_is_inf(acc)) {
      const f9 inv = fq9_inv(Fq9::mul(acc.zz, acc.zzz));
      x = Fq9::canonical(Fq9::mul(acc.x, Fq9::mul(inv, acc.zzz)));
      y = Fq9::canonical(Fq9::mul(acc.y, Fq9::mul(inv, acc.zz)));
    }
    fe* t = table + 2 * ((uint64_t)w * n + i);
    store_fe(t, pack9(x));
    store_fe(t + 1, pack9(y));
    if (w + 1 < W)
      for (uint32_t d = 0; d < c; d++) acc = xyzz9_dbl(acc);
  }
}

// digits of MSM j (scalars at soff[j], count sizes[j]) into row j:
// digits[row_off[j] + w * sizes[j] + i]
__global__ __launch_bounds__(256) void msm_digits_fixed_kernel(const fe* __restrict__ scalars,
                                                               const MsmRow* __restrict__ rows,
                                                               const uint64_t* __restrict__ soff,
                                                               uint32_t c, uint32_t W,
                                                               uint32_t* __restrict__ digits) {
  const MsmRow row = rows[blockIdx.y];
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= row.inner) return;
  const fe s = Fr::from_mont(load_fe(scalars + soff[blockIdx.y] + i));
  uint64_t buf = (uint64_t)s.w[0] | ((uint64_t)s.w[1] << 32);
  int have = 64;
  uint32_t nx[6];
#pragma unroll
  for (int k = 0; k < 6; k++) nx[k] = s.w[k + 2];
  const uint32_t mask = (1u << c) - 1;
  const uint32_t half = 1u << (c - 1);
  uint32_t carry = 0;
  for (uint32_t w = 0; w < W; w++) {
    if (have < (int)c) {
      buf |= (uint64_t)nx[0] << have;
      have += 32;
#pragma unroll
      for (int k = 0; k < 5; k++) nx[k] = nx[k + 1];
      nx[5] = 0;
    }
    const uint32_t val = (uint32_t)(buf & mask) + carry;
    buf >>= c;
    have -= (int)c;
    uint32_t out;
    if (val >= half) {
      const uint32_t mag = (1u << c) - val;
      carry = 1;
      out = mag ? (mag | SIGN) : 0u;
    } else {
      carry = 0;
      out = val;
    }
    digits[row.off + (uint64_t)w * row.inner + i] = out;
  }
}

// ---------------------------------------------------------------- host side

// c = max(4, min(16, bitlen(n) - 2))   (runtime/metal/msm.py:191,
// runtime/kernels/cpu/kernels.py:557-560)
static int default_window_bits(uint64_t n) {
  int bl = 0;
  while ((n >> bl) != 0) bl++;
  int c = bl - 2;
  if (c > 16) c = 16;
  if (c < 4) c = 4;
  return c;
}

// The window of a variable-base MSM called with window_bits 0: the fastest
// c on MI355X among 5, 6, 8, 10, 12, 13, 15, 16, 17, 18 for n = 2^8 .. 2^23
// (uniform scalars, one process).  Windows whose count times width comes to
// 255 or 256 bits (c = 5, 8, 15, 17) spend no pass on a top window of one or
// two bits, whose few buckets hold nearly every point.  Against the rule
// above: 2^10 0.60 -> 0.45 ms, 2^13 0.78 -> 0.60, 2^15 1.05 -> 0.69, 2^22
// 7.24 -> 5.74 ms (c = 16 -> 17).  Fixed-base tables (whose size grows with
// the window count) and the no-allocation entry point (whose arena limit is
// stated per shape) keep the rule above.
static int fast_window_bits(uint64_t n) {
  if (n < (1ull << 12)) return 5;
  if (n < (1ull << 15)) return 8;
  if (n < (1ull << 20)) return 15;
  return 17;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Shape of one pipeline launch sequence over `nrows` rows of digits.
struct MsmShape {
  bool packed = false;  // fine key carried in the entry's bits [31 - S, 31)
  int c = 0;
  uint32_t W = 0;       // windows per scalar
  uint32_t B = 0;       // buckets per row (incl. the empty bucket 0)
  uint32_t nrows = 0;
  uint32_t nkeys = 0;   // nrows * B
  uint32_t nbins = 0;   // B - 1
  uint32_t S = 0;       // fine-key bits
  uint32_t NC = 0;      // coarse keys per row
  uint32_t max_tiles = 0;
  uint32_t tile_slots = 0;
  uint64_t hist_len = 0;  // tile_slots * NC
  uint32_t fch = 0;       // entries per fine-level chunk
  uint32_t fchunks = 0;   // upper bound of the fine-level chunks
  uint64_t pairs = 0;   // digits over all rows
  uint64_t Smax = 0;    // accumulation segments
  uint32_t L = 0, K = 0, cpw = 0;
  uint32_t nchunks = 0;
  uint32_t ml = 0, mh = 0, ng = 0;  // digit-split reduction: bucket index bits below / above, folded terms per row
  uint32_t pt_bytes = sizeof(pt9);    // a stored point (bucket, head partial, sum): pt9, or pt18 for G2
  uint32_t exp_bytes = sizeof(xyzz);  // an exported point: 4 fe, or 8 for G2
};

struct MsmBuffers {
  uint32_t *digits, *hist, *items_v, *row_ptr, *sums, *sorted, *hkey, *fc0, *fhist, *maxrun;
  uint16_t* items_f;
  pt9 *buckets, *head, *partial, *partial2, *rsum;
  pt9 *hsum, *lsum, *gsum;  // digit-split reduction
  xyzz *rexp, *gexp;
  MsmRow* rows;
};

// Accumulation segment length when the msm_segment knob is 0: the launch's
// waves run in rounds of num_cus x 4 SIMDs x ACC_WAVES_PER_SIMD (the
// accumulation's 145-147 VGPRs allow three per SIMD), each round about as
// long as its segments, so L in [SEG_MIN, SEG_MAX] minimises rounds x L
// (the longer L on ties: fewer head partials to fold).  Measured against a
// fixed 80 (one process): 2^16 0.79 vs 1.11 ms, 2^18 0.89 vs 1.11, 2^20
// 1.86 vs 2.11, 2^21 3.27 vs 3.38; 2^22 keeps 80.
static constexpr uint32_t SEG_MIN = 24, SEG_MAX = 80, ACC_WAVES_PER_SIMD = 3;
static uint32_t auto_segment(const Ctx* c, uint64_t pairs) {
  const uint64_t cap = (uint64_t)(c->num_cus ? c->num_cus : 1) * 4 * ACC_WAVES_PER_SIMD;
  uint32_t best = SEG_MAX;
  uint64_t best_cost = UINT64_MAX;
  for (uint32_t L = SEG_MAX; L >= SEG_MIN; L--) {
    const uint64_t waves = (pairs + 64ull * L - 1) / (64ull * L);
    const uint64_t cost = ((waves + cap - 1) / cap) * L;
    if (cost < best_cost) {
      best_cost = cost;
      best = L;
    }
  }
  return best;
}

// fixed_segment: SEG_MAX when the knob is 0 (a caller's fixed-size arena,
// whose capacity must not depend on the launch size)
static int make_shape(Ctx* c, int cb, uint32_t nrows, const std::vector<MsmRow>& rows, MsmShape& sh,
                      bool fixed_segment = false) {
  if (cb < 2 || cb > 24) return set_err(c, ZK_ERR_INVALID, "window_bits must be in [2, 24]");
  sh.c = cb;
  sh.W = (255 + cb - 1) / cb;
  sh.B = (1u << (cb - 1)) + 1;
  sh.nrows = nrows;
  sh.nkeys = nrows * sh.B;
  sh.nbins = sh.B - 1;
  {
    uint32_t lb = 0;
    while ((1u << lb) < sh.nbins) lb++;
    sh.S = lb > 8 ? lb - 8 : 0;
    if (sh.S > SORT_MAX_FINE_BITS) sh.S = SORT_MAX_FINE_BITS;
    sh.NC = (sh.nbins + (1u << sh.S) - 1) >> sh.S;
    if (sh.NC > SORT_MAX_COARSE) return set_err(c, ZK_ERR_INVALID, "window too large for the bucket sort");
  }
  sh.max_tiles = 0;
  sh.tile_slots = 0;
  sh.pairs = 0;
  for (const MsmRow& r : rows) {
    sh.max_tiles = r.ntiles > sh.max_tiles ? r.ntiles : sh.max_tiles;
    sh.tile_slots += r.ntiles;
    sh.pairs += r.len;
  }
  if (sh.pairs >= (1ull << 32)) return set_err(c, ZK_ERR_INVALID, "total digit count must be below 2^32");
  {
    // packed sort entries need every base index below 2^(31 - S)
    uint64_t bound = 0;
    for (const MsmRow& r : rows) {
      if (r.len == 0) continue;
      const uint64_t nw = (r.len + r.inner - 1) / r.inner;
      const uint64_t e = (nw - 1) * r.stride + r.boff + (r.inner < r.len ? r.inner : r.len);
      bound = e > bound ? e : bound;
    }
    sh.packed = bound <= (1ull << (31 - sh.S));
  }
  if (nrows > 65535) return set_err(c, ZK_ERR_INVALID, "too many MSM rows in one launch");
  sh.hist_len = (uint64_t)sh.tile_slots * sh.NC;
  if ((1u << sh.S) <= FINE_STAGE_KEYS) sh.fch = FINE_STAGE;  // staged placement
  else sh.fch = FINE_CHUNK > 16u * (1u << sh.S) ? FINE_CHUNK : 16u * (1u << sh.S);
  {
    const uint64_t fc = sh.pairs / sh.fch + (uint64_t)nrows * sh.NC;
    if (fc * (1ull << sh.S) >= (1ull << 32)) return set_err(c, ZK_ERR_INVALID, "MSM too large for the bucket sort");
    sh.fchunks = (uint32_t)fc;
  }
  sh.L = c->msm_segment ? c->msm_segment : (fixed_segment ? SEG_MAX : auto_segment(c, sh.pairs));
  sh.Smax = (sh.pairs + sh.L - 1) / sh.L;
  if (sh.Smax == 0) sh.Smax = 1;
  uint32_t K = c->msm_reduce_chunk;
  if (K == 0) {
    // ~64 K reduction threads, at most 8 buckets each (measured best at 2^22, c = 17)
    const uint64_t per = ((uint64_t)nrows * sh.nbins) / 65536;
    K = per < 2 ? 2 : (per > 8 ? 8 : (uint32_t)per);
  }
  if (K > sh.nbins) K = sh.nbins;
  sh.K = K;
  sh.cpw = (sh.nbins + K - 1) / K;
  {
    uint32_t m = 0;  // nbins = 2^m
    while ((1u << m) < sh.nbins) m++;
    sh.ml = (m + 1) / 2;
    sh.mh = m - sh.ml;
    sh.ng = (m + 1 + SPLIT_FOLD - 1) / SPLIT_FOLD;
  }
  {
    uint64_t longest = sh.hist_len;  // scan partial sums, all levels, of the longest scan
    if ((uint64_t)sh.fchunks << sh.S > longest) longest = (uint64_t)sh.fchunks << sh.S;
    sh.nchunks = (uint32_t)(longest / (SCAN_CHUNK - 1) + 16);
  }
  return ZK_OK;
}

// digits_ext: the rows' digits live elsewhere (then the sorted entries reuse
// that region, which must hold sh.pairs words and belong to these rows only)
static size_t layout(const MsmShape& sh, char* base, MsmBuffers* b, uint32_t* digits_ext = nullptr) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off = align256(off + bytes);
    return base ? (void*)(base + o) : nullptr;
  };
  MsmBuffers t;
  t.digits = digits_ext ? digits_ext : (uint32_t*)take(sh.pairs * 4);
  t.sorted = t.digits;  // the digits are dead once the scatter pass has run
  t.hist = (uint32_t*)take((size_t)(sh.hist_len + 1) * 4);
  t.items_v = (uint32_t*)take(sh.pairs * 4);
  t.items_f = (uint16_t*)take(sh.pairs * 2);
  t.row_ptr = (uint32_t*)take((size_t)(sh.nkeys + 1) * 4);
  t.sums = (uint32_t*)take((size_t)sh.nchunks * 4);
  t.fc0 = (uint32_t*)take(((size_t)sh.nrows * sh.NC + 1) * 4);
  t.fhist = (uint32_t*)take(((size_t)sh.fchunks << sh.S) * 4 + 4);
  t.buckets = (pt9*)take((size_t)sh.nkeys * sh.pt_bytes);
  t.hkey = (uint32_t*)take(sh.Smax * 4);
  t.head = (pt9*)take(sh.Smax * sh.pt_bytes);
  t.partial = (pt9*)take((size_t)sh.nrows * sh.cpw * sizeof(pt9));
  t.partial2 = (pt9*)take((size_t)sh.nrows * ((sh.cpw + RED_GROUP - 1) / RED_GROUP) * sizeof(pt9));
  t.rsum = (pt9*)take((size_t)sh.nrows * sizeof(pt9));
  t.rexp = (xyzz*)take((size_t)sh.nrows * sizeof(xyzz));
  t.hsum = (pt9*)take((size_t)sh.nrows * ((1u << sh.mh) + 1) * sh.pt_bytes);
  t.lsum = (pt9*)take((size_t)sh.nrows * (1u << sh.ml) * sh.pt_bytes);
  t.gsum = (pt9*)take((size_t)sh.nrows * sh.ng * sh.pt_bytes);
  t.gexp = (xyzz*)take((size_t)sh.nrows * sh.ng * sh.exp_bytes);
  t.rows = (MsmRow*)take((size_t)sh.nrows * sizeof(MsmRow));
  t.maxrun = (uint32_t*)take(4);
  if (b) *b = t;
  return off;
}

// exclusive scan of in[0, len) into out[0, len) (in place allowed) with the
// total in out[len]; tmp holds the partial sums of every level
// (len / (SCAN_CHUNK - 1) + 16 words suffice)
static void scan_exclusive(const uint32_t* in, uint32_t* out, uint64_t len, uint32_t* tmp, hipStream_t st) {
  if (len <= (uint64_t)SCAN_CHUNK) {
    scan_chunks_kernel<<<1, SCAN_THREADS, 0, st>>>(in, out, out + len, len);
    return;
  }
  const uint64_t nch = (len + SCAN_CHUNK - 1) / SCAN_CHUNK;
  scan_chunks_kernel<<<(unsigned)nch, SCAN_THREADS, 0, st>>>(in, out, tmp, len);
  scan_exclusive(tmp, tmp, nch, tmp + nch + 1, st);  // tmp[nch] = total
  scan_add_kernel<<<(unsigned)nch, SCAN_THREADS, 0, st>>>(out, tmp, len);
}

// sort -> accumulate -> fix-up -> per-row reduction -> export (R = 2^256 XYZZ
// per row in b.rexp).  No allocation, no host synchronisation.
// coarse scatter and fine placement (after the coarse histogram scan);
// fuse: segments of at most FUSE_MAX entries by msm_fine_fused_kernel
template <bool PK>
static void sort_placement(const MsmShape& sh, const MsmBuffers& b, hipStream_t st, bool fuse) {
  const dim3 g(sh.max_tiles, sh.nrows);
  if (sh.max_tiles > 0) {
    if (sh.NC <= 256)
      msm_sort_scatter_kernel<PK, uint8_t, 256><<<g, SORT_THREADS, 0, st>>>(b.digits, b.rows, sh.S, sh.NC, b.hist,
                                                                           b.items_v, b.items_f);
    else
      msm_sort_scatter_kernel<PK, uint16_t, SORT_MAX_COARSE><<<g, SORT_THREADS, 0, st>>>(
          b.digits, b.rows, sh.S, sh.NC, b.hist, b.items_v, b.items_f);
  }
  const uint32_t nseg = sh.nrows * sh.NC, NF = 1u << sh.S;
  fuse = fuse && PK && NF <= FUSE_KEYS;
  if (fuse)
    msm_fine_fused_kernel<<<nseg, FUSE_THREADS, 0, st>>>(b.items_v, b.hist, b.rows, sh.S, sh.NC, nseg, sh.nbins,
                                                         sh.B, b.row_ptr, b.sorted);
  msm_fine_chunks_kernel<<<grid_for(nseg, 256), 256, 0, st>>>(b.hist, b.rows, sh.NC, nseg, sh.fch,
                                                              fuse ? FUSE_MAX : 0u, b.fc0, b.maxrun);
  scan_exclusive(b.fc0, b.fc0, nseg, b.sums, st);  // fc0[nseg] = chunks in use (<= sh.fchunks)
  msm_fine_count_kernel<PK><<<sh.fchunks, FINE_THREADS, NF * 4, st>>>(b.items_v, b.items_f, b.hist, b.rows, sh.S,
                                                                      sh.NC, nseg, sh.fch, b.fc0, b.fhist);
  // entries past the chunks in use are never read; scanning them is harmless
  scan_exclusive(b.fhist, b.fhist, (uint64_t)sh.fchunks * NF, b.sums, st);
  if (NF <= FINE_STAGE_KEYS)
    msm_fine_scatter_staged_kernel<PK><<<sh.fchunks, FINE_STAGE_THREADS, 0, st>>>(
        b.items_v, b.items_f, b.hist, b.rows, sh.S, sh.NC, nseg, sh.fch, b.fc0, b.fhist, sh.nbins, sh.B,
        b.row_ptr, b.sorted);
  else
    msm_fine_scatter_kernel<PK><<<sh.fchunks, FINE_THREADS, NF * 4, st>>>(b.items_v, b.items_f, b.hist, b.rows,
                                                                          sh.S, sh.NC, nseg, sh.fch, b.fc0,
                                                                          b.fhist, sh.nbins, sh.B, b.row_ptr,
                                                                          b.sorted);
}

// counted: the coarse histogram is already written (msm_digits_count_kernel)
// split: rows are the windows of one MSM, reduced by digit split (their
// folded terms in b.gexp, bit positions c r + SPLIT_FOLD g); otherwise one
// sum per row in b.rexp
template <class P = PtG1>
static int run_rows(Ctx* c, const MsmShape& sh, const MsmBuffers& b, const fe* bases9, hipStream_t st,
                    hipEvent_t sorted_ev = nullptr, bool counted = false, bool split = false) {
  hipEvent_t ev;
  phase_begin(c, "msm_sort", st, &ev);
  const dim3 g(sh.max_tiles, sh.nrows);
  if (sh.max_tiles > 0 && !counted)
    msm_sort_count_kernel<<<g, SORT_THREADS, 0, st>>>(b.digits, b.rows, sh.S, sh.NC, b.hist);
  scan_exclusive(b.hist, b.hist, sh.hist_len, b.sums, st);  // in place; hist[hist_len] = total
  if (sh.packed)
    sort_placement<true>(sh, b, st, c->msm_sort_fused != 0);
  else
    sort_placement<false>(sh, b, st, false);
  ZK_HIP(c, hipGetLastError());
  phase_end(c, "msm_sort", st, ev);
  if (sorted_ev) ZK_HIP(c, hipEventRecord(sorted_ev, st));

  using Sto = typename P::Sto;
  Sto* buckets = reinterpret_cast<Sto*>(b.buckets);
  Sto* head = reinterpret_cast<Sto*>(b.head);
  phase_begin(c, "msm_accumulate", st, &ev);
  msm_accumulate_kernel<P><<<grid_for(sh.Smax, 256), 256, 0, st>>>(bases9, b.sorted, b.row_ptr, sh.nkeys, sh.L,
                                                                  buckets, b.hkey, head, b.maxrun);
  ZK_HIP(c, hipGetLastError());
  phase_end(c, "msm_accumulate", st, ev);

  phase_begin(c, "msm_fixup", st, &ev);
  uint32_t rounds = 0;  // ceil(log_FIXUP_FAN(Smax)): enough for a run of every segment
  for (uint64_t span = 1; span < sh.Smax; span *= FIXUP_FAN) rounds++;
  for (uint32_t r = 0; r < rounds; r++)
    msm_fixup_round_kernel<P><<<grid_for(sh.Smax, 256), 256, 0, st>>>(b.row_ptr, sh.nkeys, sh.L, r, rounds,
                                                                     b.hkey, head, b.maxrun);
  msm_fixup_final_kernel<P><<<grid_for(sh.Smax, 256), 256, 0, st>>>(b.row_ptr, sh.nkeys, sh.L, buckets, b.hkey,
                                                                   head);
  ZK_HIP(c, hipGetLastError());
  phase_end(c, "msm_fixup", st, ev);

  phase_begin(c, "msm_reduce", st, &ev);
  if (split) {
    // 32 partials per sum when that grid fits one round of resident workgroups
    // (with 5% slack), else 16
    const uint32_t bh32 = ((1u << sh.mh) + 1 + 7) / 8, bl32 = ((1u << sh.ml) + 7) / 8;
    if ((uint64_t)sh.nrows * (bh32 + bl32) * 20 <= (uint64_t)c->num_cus * SPLIT_WG_PER_CU * 21) {
      msm_split_sums_kernel<P, 32><<<dim3(bh32 + bl32, sh.nrows), SPLIT_THREADS, 0, st>>>(
          buckets, b.row_ptr, sh.B, sh.ml, sh.mh, bh32, reinterpret_cast<Sto*>(b.hsum), reinterpret_cast<Sto*>(b.lsum));
    } else {
      const uint32_t bh = ((1u << sh.mh) + 1 + 15) / 16, bl = ((1u << sh.ml) + 15) / 16;
      msm_split_sums_kernel<P, 16><<<dim3(bh + bl, sh.nrows), SPLIT_THREADS, 0, st>>>(
          buckets, b.row_ptr, sh.B, sh.ml, sh.mh, bh, reinterpret_cast<Sto*>(b.hsum), reinterpret_cast<Sto*>(b.lsum));
    }
    msm_split_planes_kernel<P><<<dim3(sh.ng, sh.nrows), SPLIT_FOLD * SPLIT_PLANE_LANES, 0, st>>>(
        reinterpret_cast<Sto*>(b.hsum), reinterpret_cast<Sto*>(b.lsum), sh.ml, sh.mh, sh.ng,
        reinterpret_cast<Sto*>(b.gsum));
    msm_wsum_export_kernel<P><<<grid_for((uint64_t)sh.nrows * sh.ng, 64), 64, 0, st>>>(
        reinterpret_cast<Sto*>(b.gsum), sh.nrows * sh.ng, reinterpret_cast<fe*>(b.gexp));
    ZK_HIP(c, hipGetLastError());
    phase_end(c, "msm_reduce", st, ev);
    return ZK_OK;
  }
  if constexpr (!std::is_same<P, PtG1>::value) {
    return set_err(c, ZK_ERR_INVALID, "the chunked bucket reduction is G1 only");
  } else {
  const uint32_t g1 = (sh.cpw + RED_CHUNK - 1) / RED_CHUNK;  // partials per window after the chunk kernel
  msm_reduce_chunks_kernel<<<dim3(g1, sh.nrows), RED_CHUNK, 0, st>>>(b.buckets, b.row_ptr, sh.B, sh.K, sh.cpw,
                                                                     b.partial);
  {
    // workgroup partials -> one sum per window, 64 to 1 per level
    pt9* src = b.partial;
    pt9* spare[2] = {b.partial2, b.partial};
    uint32_t m = g1;
    for (int lvl = 0;; lvl++) {
      const uint32_t G = (m + RED_GROUP - 1) / RED_GROUP;
      pt9* dst = (G == 1) ? b.rsum : spare[lvl & 1];
      msm_reduce_groups_kernel<<<dim3(G, sh.nrows), RED_GROUP, 0, st>>>(src, m, G, dst);
      if (G == 1) break;
      src = dst;
      m = G;
    }
  }
  msm_wsum_export_kernel<P><<<grid_for(sh.nrows, 64), 64, 0, st>>>(b.rsum, sh.nrows, reinterpret_cast<fe*>(b.rexp));
  ZK_HIP(c, hipGetLastError());
  phase_end(c, "msm_reduce", st, ev);
  return ZK_OK;
  }
}

// XYZZ (R = 2^256) -> Jacobian (X ZZ^2, Y ZZZ^2, ZZZ) words; zeros for infinity
static void export_jacobian(const xyzz& acc, uint64_t out[12]) {
  for (int i = 0; i < 12; i++) out[i] = 0;
  if (xyzz_is_inf(acc)) return;
  fe zz2 = Fq::sqr(acc.zz), zzz2 = Fq::sqr(acc.zzz);
  fe X = Fq::mul(acc.x, zz2), Y = Fq::mul(acc.y, zzz2), Z = acc.zzz;
  for (int i = 0; i < 4; i++) {
    out[i] = (uint64_t)X.w[2 * i] | ((uint64_t)X.w[2 * i + 1] << 32);
    out[4 + i] = (uint64_t)Y.w[2 * i] | ((uint64_t)Y.w[2 * i + 1] << 32);
    out[8 + i] = (uint64_t)Z.w[2 * i] | ((uint64_t)Z.w[2 * i + 1] << 32);
  }
}

// With the msm_groups tuning knob above 1, the windows are split into G
// groups, each its own sort -> accumulate -> fix-up -> reduce sequence;
// groups alternate between the context stream and a second stream, and
// group g + 1 starts sorting when group g's sort is done, so that the sort
// and reduction of one group can run beside the accumulation of its
// neighbour.  On MI355X at 2^22 this measured no faster than one group
// (prepared bases, digit-split reduction: 6.01 ms for one or two groups,
// 6.09 for three, 6.53 for five).  Registers are not the reason: with the
// accumulation held to 144 VGPRs, so that a sort wave fits beside its three
// waves per SIMD, two groups were still no faster than one (6.08 against
// 6.06 ms) and the cap cost the accumulation 1%.  The default is one group.
// One launch sequence over all W windows (the default shape).  `arena`
// null: the context's growable scratch; else the caller's arena of
// `arena_bytes`, which must hold the layout (nothing is allocated then).
// `bases9` non-null: bases already in the R = 2^261 form (a prepared set),
// so the per-call conversion is skipped.
static int msm_g1_single(Ctx* c, const fe* bases, const fe* bases9_ready, const fe* scalars, uint64_t n, int cb,
                         uint64_t out_jac[12], hipStream_t st, char* arena, size_t arena_bytes) {
  const uint32_t W = (255 + cb - 1) / cb;
  const uint32_t nt = (uint32_t)((n + SORT_TILE - 1) / SORT_TILE);
  std::vector<MsmRow> rows(W);
  for (uint32_t w = 0; w < W; w++) {
    MsmRow& r = rows[w];
    r.off = (uint64_t)w * n;
    r.len = (uint32_t)n;
    r.inner = (uint32_t)n;
    r.stride = 0;
    r.ntiles = nt;
    r.tile0 = w * nt;
    r.boff = 0;
  }
  MsmShape sh;
  int rc = make_shape(c, cb, W, rows, sh, arena != nullptr);
  if (rc) return rc;
  const size_t o_bases = align256((size_t)W * n * 4);
  const size_t o_rest = bases9_ready ? o_bases : align256(o_bases + (size_t)n * 64);
  uint32_t dummy = 0;
  const size_t need = align256(o_rest + layout(sh, nullptr, nullptr, &dummy));
  char* base = arena;
  if (!base) {
    void* p;
    rc = ensure_scratch(c, 2, need, &p);
    if (rc) return rc;
    base = (char*)p;
  } else if (need > arena_bytes) {
    char msg[160];
    snprintf(msg, sizeof(msg), "MSM of n = %llu at window %d needs %zu scratch bytes, above the %zu available",
             (unsigned long long)n, cb, need, arena_bytes);
    return set_err(c, ZK_ERR_INVALID, msg);
  }
  uint32_t* digits = (uint32_t*)base;
  const fe* bases9 = bases9_ready ? bases9_ready : (const fe*)(base + o_bases);
  MsmBuffers b;
  layout(sh, base + o_rest, &b, digits);
  ZK_HIP(c, hipMemcpyAsync(b.rows, rows.data(), rows.size() * sizeof(MsmRow), hipMemcpyHostToDevice, st));

  hipEvent_t ev;
  phase_begin(c, "msm_digits", st, &ev);
  if (!bases9_ready) msm_bases_convert_kernel<<<grid_for(n, 256), 256, 0, st>>>(bases, n, (fe*)bases9);
  msm_digits_count_kernel<<<nt, DIGITS_COUNT_THREADS, (size_t)W * sh.NC * 4, st>>>(
      scalars, n, (uint32_t)cb, W, sh.S, sh.NC, nt, digits, b.hist);
  ZK_HIP(c, hipGetLastError());
  phase_end(c, "msm_digits", st, ev);
  const bool split = c->msm_reduce_split != 0;
  rc = run_rows(c, sh, b, bases9, st, nullptr, true, split);
  if (rc) return rc;
  if (split) {
    // term g of window w sits at bit c w + SPLIT_FOLD g; hand them over top first.
    // Read back through the context's pinned buffer (a direct copy), except
    // on a caller's arena, where nothing may be allocated.
    const uint32_t T = W * sh.ng;
    std::vector<xyzz> tv;
    xyzz* ts;
    if (arena) {
      tv.resize(T);
      ts = tv.data();
    } else {
      void* hp;
      rc = ensure_host_out(c, (size_t)T * sizeof(xyzz), &hp);
      if (rc) return rc;
      ts = (xyzz*)hp;
    }
    ZK_HIP(c, hipMemcpyAsync(ts, b.gexp, (size_t)T * sizeof(xyzz), hipMemcpyDeviceToHost, st));
    ZK_HIP(c, hipStreamSynchronize(st));
    std::vector<xyzz> desc(T);
    std::vector<uint32_t> pos(T);
    for (uint32_t i = 0; i < T; i++) {
      const uint32_t k = T - 1 - i, w = k / sh.ng, g = k % sh.ng;
      desc[i] = ts[k];
      pos[i] = (uint32_t)cb * w + SPLIT_FOLD * g;
    }
    g1_bits_combine(desc.data(), pos.data(), T, out_jac);
    return ZK_OK;
  }
  std::vector<xyzz> hs(W);
  ZK_HIP(c, hipMemcpyAsync(hs.data(), b.rexp, (size_t)W * sizeof(xyzz), hipMemcpyDeviceToHost, st));
  ZK_HIP(c, hipStreamSynchronize(st));
  g1_window_combine(hs.data(), W, cb, out_jac);
  return ZK_OK;
}

// arena: a caller's fixed arena (the no-allocation entry point): the
// reference's default window rule, else the measured one
static int msm_check_args(Ctx* c, uint64_t n, int window_bits, int* cb, bool arena = false) {
  if (n >= (1ull << 31)) return set_err(c, ZK_ERR_INVALID, "MSM size must be below 2^31");
  *cb = window_bits > 0 ? window_bits : (arena ? default_window_bits(n) : fast_window_bits(n));
  if (*cb < 2 || *cb > 24) return set_err(c, ZK_ERR_INVALID, "window_bits must be in [2, 24]");
  return ZK_OK;
}

// ------------------------------------------------------------ prepared base sets
// A base set used by many MSMs (an SRS, a proving key's query vector) is
// converted once to the R = 2^261 form the accumulation gathers, and stays
// resident; MSMs over a prefix of it skip the per-call conversion.
int msm_bases_create(Ctx* c, const fe* bases, uint64_t n, MsmBases** out, hipStream_t st) {
  *out = nullptr;
  if (n >= (1ull << 31)) return set_err(c, ZK_ERR_INVALID, "MSM size must be below 2^31");
  MsmBases* b = new MsmBases();
  b->n = n;
  b->conv.bytes = (n ? n : 1) * 64;
  hipError_t e = hipMalloc(&b->conv.ptr, b->conv.bytes);
  if (e != hipSuccess) {
    delete b;
    return hip_check(c, e, "hipMalloc(prepared bases)");
  }
  if (n) msm_bases_convert_kernel<<<grid_for(n, 256), 256, 0, st>>>(bases, n, (fe*)b->conv.ptr);
  e = hipGetLastError();
  if (e != hipSuccess) {
    hipFree(b->conv.ptr);
    delete b;
    return hip_check(c, e, "msm_bases_convert_kernel");
  }
  *out = b;
  return ZK_OK;
}

void msm_bases_destroy(MsmBases* b) {
  if (!b) return;
  if (b->conv.ptr) hipFree(b->conv.ptr);
  delete b;
}

static int msm_g1_grouped(Ctx* c, const fe* bases, const fe* bases9_ready, const fe* scalars, uint64_t n, int cb,
                          uint32_t G, uint64_t out_jac[12], hipStream_t st);
static uint32_t msm_group_count(Ctx* c, int cb);

int msm_g1_prepared_run(Ctx* c, const MsmBases* b, const fe* scalars, uint64_t n, int window_bits,
                        uint64_t out_jac[12], hipStream_t st) {
  for (int i = 0; i < 12; i++) out_jac[i] = 0;
  if (n > b->n) return set_err(c, ZK_ERR_INVALID, "more scalars than prepared bases");
  if (n == 0) return ZK_OK;
  int cb;
  int rc = msm_check_args(c, n, window_bits, &cb);
  if (rc) return rc;
  const uint32_t G = msm_group_count(c, cb);
  if (G > 1) return msm_g1_grouped(c, nullptr, (const fe*)b->conv.ptr, scalars, n, cb, G, out_jac, st);
  return msm_g1_single(c, nullptr, (const fe*)b->conv.ptr, scalars, n, cb, out_jac, st, nullptr, 0);
}

// ------------------------------------------------------------ no-allocation MSM
// `scratch` is the MSM arena of zk_msm_g1_noalloc, for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).  Each GPU the library runs on holds its
// own copy for the life of the process, at its own address, looked up once
// per device by msm_noalloc_init.  Calls on one GPU share it and keep the
// last call's contents; nothing in it needs clearing, because every word a
// launch sequence reads was written earlier in the same sequence.  Calls
// that overlap on one GPU race.
static constexpr size_t MSM_NOALLOC_BYTES = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[MSM_NOALLOC_BYTES];
static constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_noalloc_mu;
static char* g_noalloc_ptr[NOALLOC_MAX_DEVICES];  // device address of `scratch`, per device

int msm_noalloc_init(Ctx* c) {
  if (c->device < 0 || c->device >= NOALLOC_MAX_DEVICES)
    return set_err(c, ZK_ERR_INVALID, "device ordinal beyond the no-allocation table");
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (g_noalloc_ptr[c->device]) return ZK_OK;
  void* p = nullptr;
  ZK_HIP(c, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) return set_err(c, ZK_ERR_RUNTIME, "no device address for the MSM scratch array");
  g_noalloc_ptr[c->device] = (char*)p;
  return ZK_OK;
}

int msm_g1_noalloc_run(Ctx* c, const fe* bases, const fe* scalars, uint64_t n, int window_bits,
                       uint64_t out_jac[12], hipStream_t st) {
  for (int i = 0; i < 12; i++) out_jac[i] = 0;
  if (n == 0) return ZK_OK;
  int cb;
  int rc = msm_check_args(c, n, window_bits, &cb, true);
  if (rc) return rc;
  char* arena;
  {
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    arena = (c->device >= 0 && c->device < NOALLOC_MAX_DEVICES) ? g_noalloc_ptr[c->device] : nullptr;
  }
  if (!arena) return set_err(c, ZK_ERR_RUNTIME, "MSM scratch array not set up for this device");
  return msm_g1_single(c, bases, nullptr, scalars, n, cb, out_jac, st, arena, MSM_NOALLOC_BYTES);
}

// The windows split into G groups (the msm_groups tuning knob), each its own
// sort -> accumulate -> fix-up -> reduce sequence, on two streams (see above).
// `bases9_ready` as in msm_g1_single.
static int msm_g1_grouped(Ctx* c, const fe* bases, const fe* bases9_ready, const fe* scalars, uint64_t n, int cb,
                          uint32_t G, uint64_t out_jac[12], hipStream_t st) {
  const uint32_t W = (255 + cb - 1) / cb;
  std::vector<uint32_t> wb(G + 1);
  for (uint32_t g = 0; g <= G; g++) wb[g] = (uint32_t)((uint64_t)g * W / G);
  // rows = windows: row w holds digits[w n .. w n + n), base index = position
  const uint32_t nt = (uint32_t)((n + SORT_TILE - 1) / SORT_TILE);
  std::vector<std::vector<MsmRow>> rows(G);
  std::vector<MsmShape> sh(G);
  std::vector<size_t> off(G + 1);
  const size_t o_bases = align256((size_t)W * n * 4);
  off[0] = bases9_ready ? o_bases : align256(o_bases + (size_t)n * 64);
  for (uint32_t g = 0; g < G; g++) {
    for (uint32_t w = wb[g]; w < wb[g + 1]; w++) {
      MsmRow r;
      r.off = (uint64_t)(w - wb[g]) * n;
      r.len = (uint32_t)n;
      r.inner = (uint32_t)n;
      r.stride = 0;
      r.ntiles = nt;
      r.tile0 = (w - wb[g]) * nt;
      r.boff = 0;
      rows[g].push_back(r);
    }
    int rc = make_shape(c, cb, wb[g + 1] - wb[g], rows[g], sh[g]);
    if (rc) return rc;
    uint32_t dummy = 0;
    off[g + 1] = align256(off[g] + layout(sh[g], nullptr, nullptr, &dummy));
  }
  void* base;
  int rc = ensure_scratch(c, 2, off[G], &base);
  if (rc) return rc;
  rc = ensure_aux(c, G + 2);
  if (rc) return rc;
  uint32_t* digits = (uint32_t*)base;
  const fe* bases9 = bases9_ready ? bases9_ready : (const fe*)((char*)base + o_bases);
  std::vector<MsmBuffers> b(G);
  for (uint32_t g = 0; g < G; g++) {
    layout(sh[g], (char*)base + off[g], &b[g], digits + (uint64_t)wb[g] * n);
    ZK_HIP(c, hipMemcpyAsync(b[g].rows, rows[g].data(), rows[g].size() * sizeof(MsmRow), hipMemcpyHostToDevice, st));
  }

  hipEvent_t ev;
  phase_begin(c, "msm_digits", st, &ev);
  if (!bases9_ready) msm_bases_convert_kernel<<<grid_for(n, 256), 256, 0, st>>>(bases, n, (fe*)bases9);
  msm_digits_kernel<<<grid_for(n, 256), 256, 0, st>>>(scalars, n, (uint32_t)cb, W, digits);
  ZK_HIP(c, hipGetLastError());
  phase_end(c, "msm_digits", st, ev);

  const bool split = c->msm_reduce_split != 0;
  hipStream_t S[2] = {st, c->aux_stream};
  hipEvent_t* E = c->events.data();  // E[g]: group g sorted; E[G]: digits ready; E[G + 1]: aux done
  ZK_HIP(c, hipEventRecord(E[G], st));
  ZK_HIP(c, hipStreamWaitEvent(S[1], E[G], 0));
  for (uint32_t g = 0; g < G; g++) {
    hipStream_t s = S[g & 1];
    if (g > 0) ZK_HIP(c, hipStreamWaitEvent(s, E[g - 1], 0));
    rc = run_rows(c, sh[g], b[g], bases9, s, E[g], false, split);
    if (rc) return rc;
  }
  ZK_HIP(c, hipEventRecord(E[G + 1], S[1]));
  ZK_HIP(c, hipStreamWaitEvent(st, E[G + 1], 0));

  // window combination on the host: acc = Sum_w 2^(c w) S_w (Horner)
  if (split) {
    // group g's term (row r, fold j) sits at bit c (wb[g] + r) + SPLIT_FOLD j
    const uint32_t ng = sh[0].ng;
    void* hp;
    rc = ensure_host_out(c, (size_t)W * ng * sizeof(xyzz), &hp);
    if (rc) return rc;
    xyzz* ts = (xyzz*)hp;  // the context's pinned read-back buffer
    for (uint32_t g = 0; g < G; g++)
      ZK_HIP(c, hipMemcpyAsync(ts + (size_t)wb[g] * ng, b[g].gexp,
                               (size_t)(wb[g + 1] - wb[g]) * ng * sizeof(xyzz), hipMemcpyDeviceToHost, st));
    ZK_HIP(c, hipStreamSynchronize(st));
    const uint32_t T = W * ng;
    std::vector<xyzz> desc(T);
    std::vector<uint32_t> pos(T);
    for (uint32_t i = 0; i < T; i++) {
      const uint32_t k = T - 1 - i, w = k / ng, j = k % ng;
      desc[i] = ts[k];
      pos[i] = (uint32_t)cb * w + SPLIT_FOLD * j;
    }
    g1_bits_combine(desc.data(), pos.data(), T, out_jac);
    return ZK_OK;
  }
  std::vector<xyzz> hs(W);
  for (uint32_t g = 0; g < G; g++)
    ZK_HIP(c, hipMemcpyAsync(hs.data() + wb[g], b[g].rexp, (size_t)(wb[g + 1] - wb[g]) * sizeof(xyzz),
                             hipMemcpyDeviceToHost, st));
  ZK_HIP(c, hipStreamSynchronize(st));
  g1_window_combine(hs.data(), W, cb, out_jac);
  return ZK_OK;
}

static uint32_t msm_group_count(Ctx* c, int cb) {
  const uint32_t W = (255 + cb - 1) / cb;
  uint32_t G = c->msm_groups ? c->msm_groups : 1u;
  return G > W ? W : G;
}

int msm_g1_run(Ctx* c, const fe* bases, const fe* scalars, uint64_t n, int window_bits,
               uint64_t out_jac[12], hipStream_t st) {
  for (int i = 0; i < 12; i++) out_jac[i] = 0;
  if (n == 0) return ZK_OK;
  int cb;
  int rc0 = msm_check_args(c, n, window_bits, &cb);
  if (rc0) return rc0;
  const uint32_t G = msm_group_count(c, cb);
  if (G == 1) return msm_g1_single(c, bases, nullptr, scalars, n, cb, out_jac, st, nullptr, 0);
  return msm_g1_grouped(c, bases, nullptr, scalars, n, cb, G, out_jac, st);
}

// ------------------------------------------------------------ G2 MSM
// The G2 sum over a device base set, through the same digits, bucket sort,
// accumulation, fix-up and digit-split reduction as G1 (the PtG2 point kind:
// Fq2 coordinates), with the window terms combined on the host.  A base set
// holds the host's plain affine words (x0, x1, y0, y1; zeros for the
// identity) converted once to the R = 2^261 limb form.
struct MsmG2Bases {
  uint64_t n = 0;
  DevBuf conv;
};

__global__ __launch_bounds__(256) void msm_g2_bases_convert_kernel(const fe* __restrict__ in, uint64_t n,
                                                                   fe* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const f9 r2 = Fq9::constant(Fq29Params::R2_522);  // plain x -> x 2^261 (zero stays zero)
#pragma unroll
  for (int k = 0; k < 4; k++)
    store_fe(out + 4 * i + k, pack9(Fq9::canonical(Fq9::mul(unpack9(load_fe(in + 4 * i + k)), r2))));
}

int msm_g2_bases_create(Ctx* c, const uint64_t* host_words, uint64_t n, MsmG2Bases** out, hipStream_t st) {
  *out = nullptr;
  if (n >= (1ull << 31)) return set_err(c, ZK_ERR_INVALID, "MSM size must be below 2^31");
  MsmG2Bases* b = new MsmG2Bases();
  b->n = n;
  b->conv.bytes = (n ? n : 1) * 128;
  void* stage = nullptr;
  hipError_t e = hipMalloc(&b->conv.ptr, b->conv.bytes);
  if (e == hipSuccess && n) e = hipMalloc(&stage, n * 128);
  if (e == hipSuccess && n) e = hipMemcpyAsync(stage, host_words, n * 128, hipMemcpyHostToDevice, st);
  if (e == hipSuccess && n) {
    msm_g2_bases_convert_kernel<<<grid_for(n, 256), 256, 0, st>>>((const fe*)stage, n, (fe*)b->conv.ptr);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (stage) (void)hipFree(stage);
  if (e != hipSuccess) {
    if (b->conv.ptr) (void)hipFree(b->conv.ptr);
    delete b;
    return hip_check(c, e, "G2 base set");
  }
  *out = b;
  return ZK_OK;
}

void msm_g2_bases_destroy(MsmG2Bases* b) {
  if (!b) return;
  if (b->conv.ptr) (void)hipFree(b->conv.ptr);
  delete b;
}

int msm_g2_run(Ctx* c, const MsmG2Bases* gb, const fe* scalars, uint64_t n, int window_bits, uint64_t out[16],
               hipStream_t st) {
  for (int i = 0; i < 16; i++) out[i] = 0;
  if (n > gb->n) return set_err(c, ZK_ERR_INVALID, "more scalars than G2 bases");
  if (n == 0) return ZK_OK;
  int cb;
  int rc = msm_check_args(c, n, window_bits, &cb);
  if (rc) return rc;
  const uint32_t W = (255 + cb - 1) / cb;
  const uint32_t nt = (uint32_t)((n + SORT_TILE - 1) / SORT_TILE);
  std::vector<MsmRow> rows(W);
  for (uint32_t w = 0; w < W; w++) {
    MsmRow& r = rows[w];
    r.off = (uint64_t)w * n;
    r.len = (uint32_t)n;
    r.inner = (uint32_t)n;
    r.stride = 0;
    r.ntiles = nt;
    r.tile0 = w * nt;
    r.boff = 0;
  }
  MsmShape sh;
  rc = make_shape(c, cb, W, rows, sh);
  if (rc) return rc;
  sh.pt_bytes = sizeof(pt18);
  sh.exp_bytes = PtG2::EXP_FE * sizeof(fe);
  const size_t o_rest = align256((size_t)W * n * 4);
  uint32_t dummy = 0;
  const size_t need = align256(o_rest + layout(sh, nullptr, nullptr, &dummy));
  void* p;
  rc = ensure_scratch(c, 3, need, &p);
  if (rc) return rc;
  char* base = (char*)p;
  uint32_t* digits = (uint32_t*)base;
  MsmBuffers b;
  layout(sh, base + o_rest, &b, digits);
  ZK_HIP(c, hipMemcpyAsync(b.rows, rows.data(), rows.size() * sizeof(MsmRow), hipMemcpyHostToDevice, st));
  hipEvent_t ev;
  phase_begin(c, "msm_digits", st, &ev);
  msm_digits_count_kernel<<<nt, DIGITS_COUNT_THREADS, (size_t)W * sh.NC * 4, st>>>(
      scalars, n, (uint32_t)cb, W, sh.S, sh.NC, nt, digits, b.hist);
  ZK_HIP(c, hipGetLastError());
  phase_end(c, "msm_digits", st, ev);
  rc = run_rows<PtG2>(c, sh, b, (const fe*)gb->conv.ptr, st, nullptr, true, true);
  if (rc) return rc;
  // term g of window w (8 fe) sits at bit c w + SPLIT_FOLD g; hand them over top first
  const uint32_t T = W * sh.ng;
  void* hp;
  rc = ensure_host_out(c, (size_t)T * sh.exp_bytes, &hp);
  if (rc) return rc;
  ZK_HIP(c, hipMemcpyAsync(hp, b.gexp, (size_t)T * sh.exp_bytes, hipMemcpyDeviceToHost, st));
  ZK_HIP(c, hipStreamSynchronize(st));
  const uint64_t* ts = (const uint64_t*)hp;
  std::vector<uint64_t> desc((size_t)T * 32);
  std::vector<uint32_t> pos(T);
  for (uint32_t i = 0; i < T; i++) {
    const uint32_t k = T - 1 - i, w = k / sh.ng, g = k % sh.ng;
    memcpy(desc.data() + (size_t)32 * i, ts + (size_t)32 * k, 256);
    pos[i] = (uint32_t)cb * w + SPLIT_FOLD * g;
  }
  g2_bits_combine(desc.data(), pos.data(), T, out);
  return ZK_OK;
}

// ------------------------------------------------------------ fixed-base sets

int fixed_bases_create(Ctx* c, const fe* bases, uint64_t n, int window_bits, FixedBases** out,
                       hipStream_t st) {
  *out = nullptr;
  if (n == 0 || n >= (1ull << 31)) return set_err(c, ZK_ERR_INVALID, "fixed-base set size must be in [1, 2^31)");
  const int cb = window_bits > 0 ? window_bits : default_window_bits(n);
  if (cb < 2 || cb > 24) return set_err(c, ZK_ERR_INVALID, "window_bits must be in [2, 24]");
  const uint32_t W = (255 + cb - 1) / cb;
  if ((uint64_t)W * n >= (1ull << 31)) return set_err(c, ZK_ERR_INVALID, "fixed-base table too large");
  FixedBases* fb = new FixedBases();
  fb->n = n;
  fb->c = cb;
  fb->W = W;
  fb->table.bytes = (size_t)W * n * 64;
  hipError_t e = hipMalloc(&fb->table.ptr, fb->table.bytes);
  if (e != hipSuccess) {
    delete fb;
    return hip_check(c, e, "hipMalloc(fixed-base table)");
  }
  void* tmp;
  int rc = ensure_scratch(c, 3, (size_t)n * 64, &tmp);
  if (rc) {
    hipFree(fb->table.ptr);
    delete fb;
    return rc;
  }
  msm_bases_convert_kernel<<<grid_for(n, 256), 256, 0, st>>>(bases, n, (fe*)tmp);
  msm_fixed_table_kernel<<<grid_for(n, 128), 128, 0, st>>>((const fe*)tmp, n, (uint32_t)cb, W,
                                                          (fe*)fb->table.ptr);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    hipFree(fb->table.ptr);
    delete fb;
    return hip_check(c, e, "fixed-base table build");
  }
  *out = fb;
  return ZK_OK;
}

void fixed_bases_destroy(FixedBases* fb) {
  if (!fb) return;
  if (fb->table.ptr) hipFree(fb->table.ptr);
  delete fb;
}

// MSM j = sum_{i < sizes[j]} scalars[offsets[j] + i] * B_i over one fixed set.
int msm_fixed_batch_run(Ctx* c, const FixedBases* fb, const fe* scalars, const uint64_t* offsets,
                        const uint64_t* sizes, const uint64_t* base_offsets, uint32_t count, uint64_t* out_jac,
                        hipStream_t st) {
  for (uint64_t i = 0; i < 12ull * count; i++) out_jac[i] = 0;
  if (count == 0) return ZK_OK;
  std::vector<MsmRow> rows(count);
  std::vector<uint64_t> soff(count);
  uint64_t off = 0;
  uint32_t tiles = 0, max_inner = 0;
  for (uint32_t j = 0; j < count; j++) {
    const uint64_t bo = base_offsets ? base_offsets[j] : 0;
    if (bo + sizes[j] > fb->n) return set_err(c, ZK_ERR_INVALID, "scalars length exceeds points length");
    const uint64_t len = sizes[j] * fb->W;
    rows[j].off = off;
    rows[j].len = (uint32_t)len;
    rows[j].inner = (uint32_t)sizes[j];
    rows[j].stride = (uint32_t)fb->n;
    rows[j].ntiles = (uint32_t)((len + SORT_TILE - 1) / SORT_TILE);
    rows[j].tile0 = tiles;
    rows[j].boff = (uint32_t)bo;
    tiles += rows[j].ntiles;
    off += len;
    soff[j] = offsets[j];
    if (sizes[j] > max_inner) max_inner = (uint32_t)sizes[j];
  }
  if (off >= (1ull << 32)) return set_err(c, ZK_ERR_INVALID, "total digit count must be below 2^32");
  MsmShape sh;
  int rc = make_shape(c, fb->c, count, rows, sh);
  if (rc) return rc;
  const size_t bytes = layout(sh, nullptr, nullptr);
  const size_t o_soff = align256(bytes);
  void* base;
  rc = ensure_scratch(c, 2, o_soff + (size_t)count * 8, &base);
  if (rc) return rc;
  MsmBuffers b;
  layout(sh, (char*)base, &b);
  uint64_t* d_soff = (uint64_t*)((char*)base + o_soff);
  ZK_HIP(c, hipMemcpyAsync(b.rows, rows.data(), rows.size() * sizeof(MsmRow), hipMemcpyHostToDevice, st));
  ZK_HIP(c, hipMemcpyAsync(d_soff, soff.data(), soff.size() * 8, hipMemcpyHostToDevice, st));

  if (max_inner > 0) {
    hipEvent_t ev;
    phase_begin(c, "msm_digits", st, &ev);
    dim3 g(grid_for(max_inner, 256), count);
    msm_digits_fixed_kernel<<<g, 256, 0, st>>>(scalars, b.rows, d_soff, (uint32_t)fb->c, fb->W, b.digits);
    ZK_HIP(c, hipGetLastError());
    phase_end(c, "msm_digits", st, ev);
    rc = run_rows(c, sh, b, (const fe*)fb->table.ptr, st);
    if (rc) return rc;
  }
  std::vector<xyzz> hs(count);
  if (max_inner > 0)
    ZK_HIP(c, hipMemcpyAsync(hs.data(), b.rexp, (size_t)count * sizeof(xyzz), hipMemcpyDeviceToHost, st));
  ZK_HIP(c, hipStreamSynchronize(st));
  for (uint32_t j = 0; j < count; j++) {
    if (sizes[j] == 0) continue;
    export_jacobian(hs[j], out_jac + 12ull * j);
  }
  return ZK_OK;
}

// ------------------------------------------------------------------ batches
// A batch plan: many fixed-base MSMs (uneven sizes, one base set) plus many
// NTT jobs, with every buffer and table allocated up front so the whole
// launch sequence can be captured once into a hipGraph and replayed.  The
// NTT jobs run on a forked stream, concurrent with the MSM pipeline.

struct BatchPlan {
  const FixedBases* fb = nullptr;
  uint32_t n_msm = 0;
  std::vector<uint64_t> sizes;
  const fe* scalars = nullptr;
  uint32_t max_inner = 0;
  MsmShape sh;
  MsmBuffers b;
  uint64_t* d_soff = nullptr;
  DevBuf msm_scratch;
  std::vector<NttJob> ntts;
  std::vector<DevBuf> ntt_tmp;
  hipStream_t aux = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  xyzz* host_res = nullptr;  // pinned
};

static int batch_enqueue(Ctx* c, BatchPlan* p, hipStream_t st) {
  ZK_HIP(c, hipEventRecord(p->fork, st));
  ZK_HIP(c, hipStreamWaitEvent(p->aux, p->fork, 0));
  for (size_t k = 0; k < p->ntts.size(); k++) {
    const NttJob& j = p->ntts[k];
    int rc = ntt_run(c, j.in, 1ull << j.log_n, j.out, j.log_n, j.batch, j.inverse,
                     j.has_omega ? &j.omega : nullptr, j.has_shift ? &j.shift : nullptr, p->aux,
                     c->ntt_max_pass_bits, (fe*)p->ntt_tmp[k].ptr);
    if (rc) return rc;
  }
  ZK_HIP(c, hipEventRecord(p->join, p->aux));
  if (p->n_msm && p->max_inner) {
    dim3 g(grid_for(p->max_inner, 256), p->n_msm);
    msm_digits_fixed_kernel<<<g, 256, 0, st>>>(p->scalars, p->b.rows, p->d_soff, (uint32_t)p->fb->c, p->fb->W,
                                               p->b.digits);
    ZK_HIP(c, hipGetLastError());
    int rc = run_rows(c, p->sh, p->b, (const fe*)p->fb->table.ptr, st);
    if (rc) return rc;
    ZK_HIP(c, hipMemcpyAsync(p->host_res, p->b.rexp, (size_t)p->n_msm * sizeof(xyzz), hipMemcpyDeviceToHost, st));
  }
  ZK_HIP(c, hipStreamWaitEvent(st, p->join, 0));
  return ZK_OK;
}

void batch_destroy(BatchPlan* p) {
  if (!p) return;
  hipDeviceSynchronize();
  if (p->exec) hipGraphExecDestroy(p->exec);
  if (p->graph) hipGraphDestroy(p->graph);
  if (p->msm_scratch.ptr) hipFree(p->msm_scratch.ptr);
  for (auto& t : p->ntt_tmp)
    if (t.ptr) hipFree(t.ptr);
  if (p->host_res) hipHostFree(p->host_res);
  if (p->fork) hipEventDestroy(p->fork);
  if (p->join) hipEventDestroy(p->join);
  if (p->aux) hipStreamDestroy(p->aux);
  delete p;
}

int batch_create(Ctx* c, const FixedBases* fb, const fe* scalars, const uint64_t* offsets,
                 const uint64_t* sizes, const uint64_t* base_offsets, uint32_t n_msm, const NttJob* jobs,
                 uint32_t n_ntt, bool use_graph, BatchPlan** out, hipStream_t st) {
  *out = nullptr;
  if (n_msm && !fb) return set_err(c, ZK_ERR_INVALID, "MSMs need a fixed-base set");
  BatchPlan* p = new BatchPlan();
  auto fail = [&](int rc) {
    batch_destroy(p);
    return rc;
  };
  hipError_t e = hipStreamCreateWithFlags(&p->aux, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&p->fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&p->join, hipEventDisableTiming);
  if (e != hipSuccess) return fail(hip_check(c, e, "batch streams"));
  p->fb = fb;
  p->n_msm = n_msm;
  p->scalars = scalars;
  p->use_graph = use_graph;
  if (n_msm) {
    std::vector<MsmRow> rows(n_msm);
    std::vector<uint64_t> soff(n_msm);
    uint64_t off = 0;
    uint32_t tiles = 0;
    for (uint32_t j = 0; j < n_msm; j++) {
      const uint64_t bo = base_offsets ? base_offsets[j] : 0;
      if (bo + sizes[j] > fb->n) return fail(set_err(c, ZK_ERR_INVALID, "scalars length exceeds points length"));
      const uint64_t len = sizes[j] * fb->W;
      rows[j] = MsmRow{off, (uint32_t)len, (uint32_t)sizes[j], (uint32_t)fb->n,
                       (uint32_t)((len + SORT_TILE - 1) / SORT_TILE), tiles, (uint32_t)bo};
      tiles += rows[j].ntiles;
      off += len;
      soff[j] = offsets[j];
      p->sizes.push_back(sizes[j]);
      if (sizes[j] > p->max_inner) p->max_inner = (uint32_t)sizes[j];
    }
    if (off >= (1ull << 32)) return fail(set_err(c, ZK_ERR_INVALID, "total digit count must be below 2^32"));
    int rc = make_shape(c, fb->c, n_msm, rows, p->sh);
    if (rc) return fail(rc);
    const size_t bytes = layout(p->sh, nullptr, nullptr);
    const size_t o_soff = align256(bytes);
    p->msm_scratch.bytes = o_soff + (size_t)n_msm * 8;
    e = hipMalloc(&p->msm_scratch.ptr, p->msm_scratch.bytes);
    if (e != hipSuccess) return fail(hip_check(c, e, "hipMalloc(batch MSM scratch)"));
    layout(p->sh, (char*)p->msm_scratch.ptr, &p->b);
    p->d_soff = (uint64_t*)((char*)p->msm_scratch.ptr + o_soff);
    e = hipMemcpy(p->b.rows, rows.data(), rows.size() * sizeof(MsmRow), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_soff, soff.data(), soff.size() * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipHostMalloc((void**)&p->host_res, (size_t)n_msm * sizeof(xyzz), hipHostMallocDefault);
    if (e != hipSuccess) return fail(hip_check(c, e, "batch MSM setup"));
  }
  for (uint32_t k = 0; k < n_ntt; k++) {
    const NttJob& j = jobs[k];
    if (j.log_n > 28 || j.batch == 0) return fail(set_err(c, ZK_ERR_INVALID, "bad NTT job"));
    if ((const void*)j.in == (const void*)j.out)
      return fail(set_err(c, ZK_ERR_INVALID, "batched NTT jobs must be out of place"));
    int rc = ntt_prepare(c, j.log_n, j.inverse, j.has_omega ? &j.omega : nullptr, j.has_shift ? &j.shift : nullptr,
                         st);
    if (rc) return fail(rc);
    DevBuf t;
    t.bytes = ((size_t)j.batch << j.log_n) * 32;
    e = hipMalloc(&t.ptr, t.bytes);
    if (e != hipSuccess) return fail(hip_check(c, e, "hipMalloc(batch NTT scratch)"));
    p->ntt_tmp.push_back(t);
    p->ntts.push_back(j);
  }
  e = hipStreamSynchronize(st);
  if (e != hipSuccess) return fail(hip_check(c, e, "batch setup"));
  if (use_graph) {
    const bool prof = c->profiling;
    c->profiling = false;
    e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
      c->profiling = prof;
      return fail(hip_check(c, e, "hipStreamBeginCapture"));
    }
    int rc = batch_enqueue(c, p, st);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(st, &g);
    c->profiling = prof;
    if (rc) {
      if (g) hipGraphDestroy(g);
      return fail(rc);
    }
    if (e != hipSuccess) return fail(hip_check(c, e, "hipStreamEndCapture"));
    p->graph = g;
    e = hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) return fail(hip_check(c, e, "hipGraphInstantiate"));
  }
  *out = p;
  return ZK_OK;
}

int batch_run(Ctx* c, BatchPlan* p, uint64_t* out_jac, hipStream_t st) {
  if (p->use_graph) {
    ZK_HIP(c, hipGraphLaunch(p->exec, st));
  } else {
    int rc = batch_enqueue(c, p, st);
    if (rc) return rc;
  }
  ZK_HIP(c, hipStreamSynchronize(st));
  for (uint32_t j = 0; j < p->n_msm; j++) {
    if (p->sizes[j] == 0) {
      for (int i = 0; i < 12; i++) out_jac[12ull * j + i] = 0;
      continue;
    }
    export_jacobian(p->host_res[j], out_jac + 12ull * j);
  }
  return ZK_OK;
}

}  // namespace zk
