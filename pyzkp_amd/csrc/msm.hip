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
#include "internal.hpp"
#include "curve.hpp"
#include "curve29.hpp"

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

// Hot-loop storage: coordinates are R = 2^261 limb values (field29.hpp)
// packed back into 32 bytes each, so a base is 64 bytes and a bucket 128.
__device__ __forceinline__ aff9 load_aff9(const fe* __restrict__ bases, uint32_t idx) {
  aff9 a;
  a.x = unpack9(load_fe(bases + 2 * (uint64_t)idx));
  a.y = unpack9(load_fe(bases + 2 * (uint64_t)idx + 1));
  return a;
}

__device__ __forceinline__ xyzz9 load_xyzz9(const xyzz* p) {
  const fe* f = reinterpret_cast<const fe*>(p);
  xyzz9 r;
  r.x = unpack9(load_fe(f));
  r.y = unpack9(load_fe(f + 1));
  r.zz = unpack9(load_fe(f + 2));
  r.zzz = unpack9(load_fe(f + 3));
  return r;
}

__device__ __forceinline__ void store_xyzz9(xyzz* p, const xyzz9& v) {
  fe* f = reinterpret_cast<fe*>(p);
  store_fe(f, pack9(v.x));
  store_fe(f + 1, pack9(v.y));
  store_fe(f + 2, pack9(v.zz));
  store_fe(f + 3, pack9(v.zzz));
}

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
__global__ void msm_wsum_export_kernel(const xyzz* __restrict__ in, uint32_t W, xyzz* __restrict__ out) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  const xyzz9 p = load_xyzz9(in + w);
  xyzz r;
  if (xyzz9_is_inf(p)) {
    r.x = Fq::one();
    r.y = Fq::one();
    r.zz = Fq::zero();
    r.zzz = Fq::zero();
  } else {
    r = xyzz_from9(p);
  }
  store_xyzz(out + w, r);
}


// 1. scalar (Montgomery) -> canonical -> W signed c-bit digits
//    digits[w * n + i] = |d| | (d < 0 ? SIGN : 0)
__global__ __launch_bounds__(256) void msm_digits_kernel(const fe* __restrict__ scalars, uint64_t n,
                                                         uint32_t c, uint32_t W,
                                                         uint32_t* __restrict__ digits) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const fe s = Fr::from_mont(load_fe(scalars + i));
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
    digits[(uint64_t)w * n + i] = out;
  }
}

// 2. counting sort of the (window, |d|) pairs, one window per grid row.
//    A tile of SORT_TILE digits of one window is histogrammed in LDS with
//    16-bit counters (two per word; a tile has at most 2^15 entries so a
//    counter cannot overflow).  Bins are |d| - 1 (zero digits are dropped);
//    a window with more than SORT_RANGE bins is covered by several bin
//    ranges (grid z), each re-reading the tile.
//    tile_hist[(w * ntiles + tile) * nbins + bin] = count
static constexpr uint32_t SORT_TILE = 1u << 15;
static constexpr uint32_t SORT_RANGE = 1u << 15;  // bins per LDS pass (64 KiB)
static constexpr uint32_t SORT_THREADS = 1024;

__global__ __launch_bounds__(SORT_THREADS) void msm_tile_count_kernel(
    const uint32_t* __restrict__ digits, uint64_t n, uint32_t nbins, uint32_t ntiles,
    uint32_t* __restrict__ tile_hist) {
  __shared__ uint32_t cnt[SORT_RANGE / 2];
  const uint32_t tile = blockIdx.x, w = blockIdx.y;
  const uint32_t lo = blockIdx.z * SORT_RANGE;
  const uint32_t hi = (lo + SORT_RANGE < nbins) ? lo + SORT_RANGE : nbins;
  for (uint32_t k = threadIdx.x; k < SORT_RANGE / 2; k += SORT_THREADS) cnt[k] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)w * n + (uint64_t)tile * SORT_TILE;
  const uint64_t end = (uint64_t)w * n + (((uint64_t)tile + 1) * SORT_TILE < n ? ((uint64_t)tile + 1) * SORT_TILE : n);
  for (uint64_t p = base + threadIdx.x; p < end; p += SORT_THREADS) {
    const uint32_t mag = digits[p] & ~SIGN;
    if (mag == 0) continue;
    const uint32_t bin = mag - 1;
    if (bin < lo || bin >= hi) continue;
    const uint32_t b = bin - lo;
    atomicAdd(&cnt[b >> 1], 1u << ((b & 1) * 16));
  }
  __syncthreads();
  uint32_t* out = tile_hist + ((uint64_t)w * ntiles + tile) * nbins;
  for (uint32_t b = lo + threadIdx.x; b < hi; b += SORT_THREADS) {
    const uint32_t k = b - lo;
    out[b] = (cnt[k >> 1] >> ((k & 1) * 16)) & 0xffffu;
  }
}

// exclusive prefix of each (window, bin) column over the tiles, in place;
// the column total becomes counts[w * B + bin + 1] (bucket |d| = bin + 1)
__global__ __launch_bounds__(256) void msm_tile_prefix_kernel(uint32_t* __restrict__ tile_hist,
                                                              uint32_t W, uint32_t nbins,
                                                              uint32_t ntiles, uint32_t B,
                                                              uint32_t* __restrict__ counts) {
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (uint64_t)W * nbins) return;
  const uint32_t w = (uint32_t)(gid / nbins), bin = (uint32_t)(gid - (uint64_t)w * nbins);
  uint32_t run = 0;
  uint32_t* col = tile_hist + (uint64_t)w * ntiles * nbins + bin;
  for (uint32_t t = 0; t < ntiles; t++) {
    const uint32_t v = col[(uint64_t)t * nbins];
    col[(uint64_t)t * nbins] = run;
    run += v;
  }
  counts[w * B + bin + 1] = run;
  if (bin == 0) counts[w * B] = 0;
}

// 2. exclusive scan of the bucket counts -> row pointers (3 small kernels)
static constexpr int SCAN_THREADS = 256;
static constexpr int SCAN_PER_THREAD = 8;
static constexpr int SCAN_CHUNK = SCAN_THREADS * SCAN_PER_THREAD;

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

__global__ __launch_bounds__(SCAN_THREADS) void scan_chunks_kernel(const uint32_t* __restrict__ in,
                                                                   uint32_t* __restrict__ out,
                                                                   uint32_t* __restrict__ sums,
                                                                   uint32_t len) {
  __shared__ uint32_t sh[SCAN_THREADS];
  const uint32_t base = blockIdx.x * SCAN_CHUNK + threadIdx.x * SCAN_PER_THREAD;
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

// one block scans all chunk sums (at most SCAN_CHUNK of them) in place and
// writes the grand total to out_total
__global__ __launch_bounds__(SCAN_THREADS) void scan_sums_kernel(uint32_t* __restrict__ sums,
                                                                 uint32_t nsums,
                                                                 uint32_t* __restrict__ out_total) {
  __shared__ uint32_t sh[SCAN_THREADS];
  const uint32_t base = threadIdx.x * SCAN_PER_THREAD;
  uint32_t v[SCAN_PER_THREAD];
  uint32_t local = 0;
#pragma unroll
  for (int k = 0; k < SCAN_PER_THREAD; k++) {
    v[k] = (base + k < nsums) ? sums[base + k] : 0;
    local += v[k];
  }
  uint32_t total;
  uint32_t run = block_exclusive_scan(local, sh, total);
#pragma unroll
  for (int k = 0; k < SCAN_PER_THREAD; k++) {
    if (base + k < nsums) sums[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 0) *out_total = total;
}

__global__ __launch_bounds__(SCAN_THREADS) void scan_add_kernel(uint32_t* __restrict__ out,
                                                                const uint32_t* __restrict__ sums,
                                                                uint32_t len) {
  const uint32_t base = blockIdx.x * SCAN_CHUNK;
  const uint32_t add = sums[blockIdx.x];
  for (uint32_t k = threadIdx.x; k < SCAN_CHUNK; k += SCAN_THREADS) {
    const uint32_t idx = base + k;
    if (idx < len) out[idx] += add;
  }
}

// 3. scatter: sorted[row_ptr[w B + |d|] + tile prefix + rank in tile] =
//    point index | sign; ranks come from 16-bit LDS counters as in the count pass
__global__ __launch_bounds__(SORT_THREADS) void msm_tile_scatter_kernel(
    const uint32_t* __restrict__ digits, uint64_t n, uint32_t nbins, uint32_t ntiles, uint32_t B,
    const uint32_t* __restrict__ tile_hist, const uint32_t* __restrict__ row_ptr,
    uint32_t* __restrict__ sorted) {
  __shared__ uint32_t cnt[SORT_RANGE / 2];
  const uint32_t tile = blockIdx.x, w = blockIdx.y;
  const uint32_t lo = blockIdx.z * SORT_RANGE;
  const uint32_t hi = (lo + SORT_RANGE < nbins) ? lo + SORT_RANGE : nbins;
  for (uint32_t k = threadIdx.x; k < SORT_RANGE / 2; k += SORT_THREADS) cnt[k] = 0;
  __syncthreads();
  const uint32_t* pre = tile_hist + ((uint64_t)w * ntiles + tile) * nbins;
  const uint32_t* rp = row_ptr + (uint64_t)w * B + 1;
  const uint64_t i0 = (uint64_t)tile * SORT_TILE;
  const uint64_t i1 = (i0 + SORT_TILE < n) ? i0 + SORT_TILE : n;
  for (uint64_t i = i0 + threadIdx.x; i < i1; i += SORT_THREADS) {
    const uint32_t d = digits[(uint64_t)w * n + i];
    const uint32_t mag = d & ~SIGN;
    if (mag == 0) continue;
    const uint32_t bin = mag - 1;
    if (bin < lo || bin >= hi) continue;
    const uint32_t b = bin - lo;
    const uint32_t sh = (b & 1) * 16;
    const uint32_t rank = (atomicAdd(&cnt[b >> 1], 1u << sh) >> sh) & 0xffffu;
    sorted[rp[bin] + pre[bin] + rank] = (uint32_t)i | (d & SIGN);
  }
}

// 4. bucket accumulation over fixed-length segments of the sorted list.
//    A run (bucket) that starts inside the segment is owned by it and its
//    partial sum goes straight to buckets[key]; the run the segment starts
//    in the middle of is its "head" partial, added in by the fix-up pass.
__global__ __launch_bounds__(256) void msm_accumulate_kernel(
    const fe* __restrict__ bases, const uint32_t* __restrict__ sorted,
    const uint32_t* __restrict__ row_ptr, uint32_t nkeys, uint32_t L, xyzz* __restrict__ buckets,
    uint32_t* __restrict__ head_key, xyzz* __restrict__ head) {
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
  xyzz9 acc = xyzz9_inf();
  for (uint64_t p = start; p < end; p++) {
    while (p >= rend) {
      if (rend > rbeg) {
        if (rbeg < start) {
          hk = k;
          store_xyzz9(head + s, acc);
        } else {
          store_xyzz9(buckets + k, acc);
        }
      }
      acc = xyzz9_inf();
      k++;
      rbeg = rend;
      rend = row_ptr[k + 1];
    }
    const uint32_t v = sorted[p];
    aff9 a = load_aff9(bases, v & ~SIGN);
    if (!aff9_is_inf(a)) {
      if (v & SIGN) a.y = Fq9::neg(a.y);
      xyzz9_madd(acc, a);
    }
  }
  if (rbeg < start) {
    hk = k;
    store_xyzz9(head + s, acc);
  } else {
    store_xyzz9(buckets + k, acc);
  }
  head_key[s] = hk;
}

// 5. fix-up of buckets that span several segments.  The head partials of
//    bucket k sit in consecutive segments first..first+len-1 with
//    first = row_ptr[k] / L + 1.  Round r folds FIXUP_FAN partials spaced
//    FIXUP_FAN^r apart into the one at offset 0 mod FIXUP_FAN^(r+1); after
//    enough rounds the run's sum is at `first` and is added to buckets[k]
//    (which the owning segment wrote).  Depth log_FAN(len) instead of len.
static constexpr uint32_t FIXUP_FAN = 32;

__device__ __forceinline__ bool head_run(const uint32_t* __restrict__ row_ptr, uint32_t k, uint32_t L,
                                         uint64_t s, uint64_t& first, uint64_t& len) {
  const uint64_t a = row_ptr[k], b = row_ptr[k + 1];
  first = a / L + 1;
  len = (b - 1) / L - a / L;
  return s >= first && s < first + len;
}

__global__ __launch_bounds__(256) void msm_fixup_round_kernel(const uint32_t* __restrict__ row_ptr,
                                                              uint32_t nkeys, uint32_t L,
                                                              uint64_t stride,
                                                              const uint32_t* __restrict__ head_key,
                                                              xyzz* __restrict__ head) {
  const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t M = row_ptr[nkeys];
  const uint64_t S = (M + L - 1) / L;
  if (s >= S) return;
  const uint32_t k = head_key[s];
  if (k == NO_KEY) return;
  uint64_t first, len;
  if (!head_run(row_ptr, k, L, s, first, len)) return;
  const uint64_t off = s - first;
  if (off % (stride * FIXUP_FAN) != 0 || off + stride >= len) return;
  xyzz9 acc = load_xyzz9(head + s);
  for (uint32_t j = 1; j < FIXUP_FAN; j++) {
    const uint64_t t = off + (uint64_t)j * stride;
    if (t >= len) break;
    acc = xyzz9_add(acc, load_xyzz9(head + first + t));
  }
  store_xyzz9(head + s, acc);
}

__global__ __launch_bounds__(256) void msm_fixup_final_kernel(const uint32_t* __restrict__ row_ptr,
                                                              uint32_t nkeys, uint32_t L,
                                                              xyzz* __restrict__ buckets,
                                                              const uint32_t* __restrict__ head_key,
                                                              const xyzz* __restrict__ head) {
  const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t M = row_ptr[nkeys];
  const uint64_t S = (M + L - 1) / L;
  if (s >= S) return;
  const uint32_t k = head_key[s];
  if (k == NO_KEY) return;
  uint64_t first, len;
  if (!head_run(row_ptr, k, L, s, first, len) || s != first) return;
  store_xyzz9(buckets + k, xyzz9_add(load_xyzz9(buckets + k), load_xyzz9(head + s)));
}

// 6. bucket reduction, chunked: chunk t of window w covers buckets
//    [lo, lo + K) with lo = 1 + t K; its weighted sum
//      Sum_b b B_b = Sum_b (b - lo + 1) B_b + (lo - 1) Sum_b B_b
//    is a running sum from the top plus one small scalar multiple.
__global__ __launch_bounds__(128) void msm_reduce_chunks_kernel(const xyzz* __restrict__ buckets,
                                                                const uint32_t* __restrict__ row_ptr,
                                                                uint32_t W, uint32_t B, uint32_t K,
                                                                uint32_t cpw,
                                                                xyzz* __restrict__ partial) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= W * cpw) return;
  const uint32_t w = gid / cpw, t = gid - w * cpw;
  const uint32_t lo = 1 + t * K;
  uint32_t hi = lo + K - 1;
  if (hi > B - 1) hi = B - 1;
  xyzz9 run = xyzz9_inf(), sum = xyzz9_inf();
  for (uint32_t b = hi; b >= lo; b--) {
    const uint32_t key = w * B + b;
    if (row_ptr[key + 1] > row_ptr[key]) run = xyzz9_add(run, load_xyzz9(buckets + key));
    sum = xyzz9_add(sum, run);
  }
  if (lo > 1) sum = xyzz9_add(sum, xyzz9_mul_small(run, lo - 1));
  store_xyzz9(partial + gid, sum);
}

// 7. per-window tree sum of the chunk partials
__global__ __launch_bounds__(128) void msm_reduce_windows_kernel(const xyzz* __restrict__ partial,
                                                                 uint32_t cpw,
                                                                 xyzz* __restrict__ window_sums) {
  __shared__ xyzz9 sh[128];
  const uint32_t w = blockIdx.x, tid = threadIdx.x;
  xyzz9 acc = xyzz9_inf();
  for (uint32_t t = tid; t < cpw; t += blockDim.x) acc = xyzz9_add(acc, load_xyzz9(partial + (uint64_t)w * cpw + t));
  sh[tid] = acc;
  __syncthreads();
  for (uint32_t step = blockDim.x >> 1; step > 0; step >>= 1) {
    if (tid < step) sh[tid] = xyzz9_add(sh[tid], sh[tid + step]);
    __syncthreads();
  }
  if (tid == 0) store_xyzz9(window_sums + w, sh[0]);
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

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

int msm_g1_run(Ctx* c, const fe* bases, const fe* scalars, uint64_t n, int window_bits,
               uint64_t out_jac[12], hipStream_t st) {
  for (int i = 0; i < 12; i++) out_jac[i] = 0;
  if (n == 0) return ZK_OK;
  if (n >= (1ull << 31)) return set_err(c, ZK_ERR_INVALID, "MSM size must be below 2^31");
  const int cb = window_bits > 0 ? window_bits : default_window_bits(n);
  if (cb < 2 || cb > 24) return set_err(c, ZK_ERR_INVALID, "window_bits must be in [2, 24]");
  const uint32_t W = (255 + cb - 1) / cb;
  const uint32_t B = (1u << (cb - 1)) + 1;
  const uint32_t nkeys = W * B;
  const uint32_t L = c->msm_segment;
  uint32_t K = c->msm_reduce_chunk;
  if (K == 0) {
    // enough reduction threads to fill the device (~64K), at least 2 buckets each
    const uint64_t per = ((uint64_t)W * (B - 1)) / 65536;
    K = per < 2 ? 2 : (uint32_t)per;
  }
  if (K > B - 1) K = B - 1;
  const uint32_t cpw = (B - 1 + K - 1) / K;
  const uint64_t pairs = n * W;
  if (pairs >= (1ull << 32)) return set_err(c, ZK_ERR_INVALID, "n * window count must be below 2^32");
  const uint64_t Smax = (pairs + L - 1) / L;
  const uint32_t nchunks = (nkeys + SCAN_CHUNK - 1) / SCAN_CHUNK;
  if (nchunks > SCAN_CHUNK) return set_err(c, ZK_ERR_INVALID, "too many buckets for the scan");
  const uint32_t nbins = B - 1;  // |d| in 1..2^(c-1)
  const uint32_t ntiles = (uint32_t)((n + SORT_TILE - 1) / SORT_TILE);
  const uint32_t nranges = (nbins + SORT_RANGE - 1) / SORT_RANGE;

  // scratch layout
  size_t off = 0;
  const size_t o_digits = off; off = align256(off + pairs * 4);
  const size_t o_hist = off; off = align256(off + (size_t)W * ntiles * nbins * 4);
  const size_t o_counts = off; off = align256(off + (size_t)(nkeys + 1) * 4);
  const size_t o_rowptr = off; off = align256(off + (size_t)(nkeys + 1) * 4);
  const size_t o_sums = off; off = align256(off + (size_t)nchunks * 4);
  const size_t o_sorted = off; off = align256(off + pairs * 4);
  const size_t o_buckets = off; off = align256(off + (size_t)nkeys * sizeof(xyzz));
  const size_t o_hkey = off; off = align256(off + Smax * 4);
  const size_t o_head = off; off = align256(off + Smax * sizeof(xyzz));
  const size_t o_partial = off; off = align256(off + (size_t)W * cpw * sizeof(xyzz));
  const size_t o_wsum = off; off = align256(off + (size_t)W * sizeof(xyzz));
  const size_t o_wexp = off; off = align256(off + (size_t)W * sizeof(xyzz));
  const size_t o_bases = off; off = align256(off + (size_t)n * 64);
  void* base;
  int rc = ensure_scratch(c, 2, off, &base);
  if (rc) return rc;
  char* b8 = (char*)base;
  uint32_t* digits = (uint32_t*)(b8 + o_digits);
  uint32_t* hist = (uint32_t*)(b8 + o_hist);
  uint32_t* counts = (uint32_t*)(b8 + o_counts);
  uint32_t* row_ptr = (uint32_t*)(b8 + o_rowptr);
  uint32_t* sums = (uint32_t*)(b8 + o_sums);
  uint32_t* sorted = (uint32_t*)(b8 + o_sorted);
  xyzz* buckets = (xyzz*)(b8 + o_buckets);
  uint32_t* hkey = (uint32_t*)(b8 + o_hkey);
  xyzz* head = (xyzz*)(b8 + o_head);
  xyzz* partial = (xyzz*)(b8 + o_partial);
  xyzz* wsum = (xyzz*)(b8 + o_wsum);
  xyzz* wexp = (xyzz*)(b8 + o_wexp);
  fe* bases9 = (fe*)(b8 + o_bases);

  hipEvent_t ev;
  phase_begin(c, "msm_digits", st, &ev);
  msm_bases_convert_kernel<<<grid_for(n, 256), 256, 0, st>>>(bases, n, bases9);
  msm_digits_kernel<<<grid_for(n, 256), 256, 0, st>>>(scalars, n, (uint32_t)cb, W, digits);
  ZK_HIP(c, hipGetLastError());
  phase_end(c, "msm_digits", st, ev);

  phase_begin(c, "msm_sort", st, &ev);
  {
    dim3 g(ntiles, W, nranges);
    msm_tile_count_kernel<<<g, SORT_THREADS, 0, st>>>(digits, n, nbins, ntiles, hist);
    msm_tile_prefix_kernel<<<grid_for((uint64_t)W * nbins, 256), 256, 0, st>>>(hist, W, nbins, ntiles, B,
                                                                               counts);
    scan_chunks_kernel<<<nchunks, SCAN_THREADS, 0, st>>>(counts, row_ptr, sums, nkeys);
    scan_sums_kernel<<<1, SCAN_THREADS, 0, st>>>(sums, nchunks, row_ptr + nkeys);
    scan_add_kernel<<<nchunks, SCAN_THREADS, 0, st>>>(row_ptr, sums, nkeys);
    msm_tile_scatter_kernel<<<g, SORT_THREADS, 0, st>>>(digits, n, nbins, ntiles, B, hist, row_ptr, sorted);
  }
  ZK_HIP(c, hipGetLastError());
  phase_end(c, "msm_sort", st, ev);

  phase_begin(c, "msm_accumulate", st, &ev);
  msm_accumulate_kernel<<<grid_for(Smax, 256), 256, 0, st>>>(bases9, sorted, row_ptr, nkeys, L, buckets,
                                                            hkey, head);
  ZK_HIP(c, hipGetLastError());
  phase_end(c, "msm_accumulate", st, ev);

  phase_begin(c, "msm_fixup", st, &ev);
  for (uint64_t stride = 1; stride < Smax; stride *= FIXUP_FAN)
    msm_fixup_round_kernel<<<grid_for(Smax, 256), 256, 0, st>>>(row_ptr, nkeys, L, stride, hkey, head);
  msm_fixup_final_kernel<<<grid_for(Smax, 256), 256, 0, st>>>(row_ptr, nkeys, L, buckets, hkey, head);
  ZK_HIP(c, hipGetLastError());
  phase_end(c, "msm_fixup", st, ev);

  phase_begin(c, "msm_reduce", st, &ev);
  msm_reduce_chunks_kernel<<<grid_for((uint64_t)W * cpw, 128), 128, 0, st>>>(buckets, row_ptr, W, B, K,
                                                                           cpw, partial);
  msm_reduce_windows_kernel<<<W, 128, 0, st>>>(partial, cpw, wsum);
  msm_wsum_export_kernel<<<grid_for(W, 64), 64, 0, st>>>(wsum, W, wexp);
  ZK_HIP(c, hipGetLastError());
  phase_end(c, "msm_reduce", st, ev);

  // window combination on the host: acc = Sum_w 2^(c w) S_w (Horner)
  std::vector<xyzz> hs(W);
  ZK_HIP(c, hipMemcpyAsync(hs.data(), wexp, (size_t)W * sizeof(xyzz), hipMemcpyDeviceToHost, st));
  ZK_HIP(c, hipStreamSynchronize(st));
  xyzz acc = hs[W - 1];
  for (int w = (int)W - 2; w >= 0; w--) {
    for (int d = 0; d < cb; d++) acc = xyzz_dbl(acc);
    acc = xyzz_add(acc, hs[w]);
  }
  if (!xyzz_is_inf(acc)) {
    // Jacobian (X ZZ^2, Y ZZZ^2, ZZZ) represents (X/ZZ, Y/ZZZ)
    fe zz2 = Fq::sqr(acc.zz), zzz2 = Fq::sqr(acc.zzz);
    fe X = Fq::mul(acc.x, zz2), Y = Fq::mul(acc.y, zzz2), Z = acc.zzz;
    for (int i = 0; i < 4; i++) {
      out_jac[i] = (uint64_t)X.w[2 * i] | ((uint64_t)X.w[2 * i + 1] << 32);
      out_jac[4 + i] = (uint64_t)Y.w[2 * i] | ((uint64_t)Y.w[2 * i + 1] << 32);
      out_jac[8 + i] = (uint64_t)Z.w[2 * i] | ((uint64_t)Z.w[2 * i + 1] << 32);
    }
  }
  return ZK_OK;
}

}  // namespace zk
