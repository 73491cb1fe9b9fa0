// Element-wise Fr/Fq kernels: Montgomery conversion, pointwise product,
// sum, difference, scaling by a constant.  Counterparts of the reference's
// pointwise_mul_fr_mont / poly_sub_fr_mont (runtime/metal/runtime.py:204-224)
// and of its host-side Montgomery packing (runtime/kernels/cpu/kernels.py:98-172),
// which this library does on the device instead of in a host loop.
#include <string.h>
#include "internal.hpp"

namespace zk {

template <int OP>
__global__ __launch_bounds__(256) void elt_kernel(const fe* __restrict__ a, const fe* __restrict__ b,
                                                  fe* __restrict__ out, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    fe x = load_fe(a + i);
    fe r;
    if constexpr (OP == ELT_FR_TO_MONT) {
      r = Fr::to_mont(Fr::reduce_full(x));
    } else if constexpr (OP == ELT_FR_FROM_MONT) {
      r = Fr::from_mont(x);
    } else if constexpr (OP == ELT_FQ_TO_MONT) {
      r = Fq::to_mont(Fq::reduce_full(x));
    } else if constexpr (OP == ELT_FQ_FROM_MONT) {
      r = Fq::from_mont(x);
    } else if constexpr (OP == ELT_FR_MUL) {
      r = Fr::mul(x, load_fe(b + i));
    } else if constexpr (OP == ELT_FR_ADD) {
      r = Fr::add(x, load_fe(b + i));
    } else if constexpr (OP == ELT_FR_SUB) {
      r = Fr::sub(x, load_fe(b + i));
    } else {
      r = x;
    }
    store_fe(out + i, r);
  }
}

int elementwise(Ctx* c, int op, const fe* a, const fe* b, fe* out, uint64_t n, hipStream_t st) {
  if (n == 0) return ZK_OK;
  const unsigned block = 256;
  uint64_t want = (n + block - 1) / block;
  unsigned grid = (unsigned)(want < 65536 ? want : 65536);
  switch (op) {
    case ELT_FR_TO_MONT: elt_kernel<ELT_FR_TO_MONT><<<grid, block, 0, st>>>(a, b, out, n); break;
    case ELT_FR_FROM_MONT: elt_kernel<ELT_FR_FROM_MONT><<<grid, block, 0, st>>>(a, b, out, n); break;
    case ELT_FQ_TO_MONT: elt_kernel<ELT_FQ_TO_MONT><<<grid, block, 0, st>>>(a, b, out, n); break;
    case ELT_FQ_FROM_MONT: elt_kernel<ELT_FQ_FROM_MONT><<<grid, block, 0, st>>>(a, b, out, n); break;
    case ELT_FR_MUL: elt_kernel<ELT_FR_MUL><<<grid, block, 0, st>>>(a, b, out, n); break;
    case ELT_FR_ADD: elt_kernel<ELT_FR_ADD><<<grid, block, 0, st>>>(a, b, out, n); break;
    case ELT_FR_SUB: elt_kernel<ELT_FR_SUB><<<grid, block, 0, st>>>(a, b, out, n); break;
    default: return set_err(c, ZK_ERR_INVALID, "unknown element-wise op");
  }
  ZK_HIP(c, hipGetLastError());
  return ZK_OK;
}

__global__ __launch_bounds__(256) void scale_kernel(const fe* __restrict__ a, fe cst, fe* __restrict__ out,
                                                    uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) store_fe(out + i, Fr::mul(load_fe(a + i), cst));
}

int elementwise_scale(Ctx* c, const fe* a, const uint64_t c_mont[4], fe* out, uint64_t n, hipStream_t st) {
  if (n == 0) return ZK_OK;
  fe cst;
  for (int i = 0; i < 4; i++) {
    cst.w[2 * i] = (uint32_t)c_mont[i];
    cst.w[2 * i + 1] = (uint32_t)(c_mont[i] >> 32);
  }
  uint64_t want = (n + 255) / 256;
  unsigned grid = (unsigned)(want < 65536 ? want : 65536);
  scale_kernel<<<grid, 256, 0, st>>>(a, cst, out, n);
  ZK_HIP(c, hipGetLastError());
  return ZK_OK;
}

}  // namespace zk

namespace zk {

// out[i] = scale * base^i (Montgomery Fr), two-level tables built on the host
__global__ __launch_bounds__(256) void fr_powers_kernel(const fe* __restrict__ hi, const fe* __restrict__ lo,
                                                        uint64_t n, fe* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) store_fe(out + i, Fr::mul(load_fe(hi + (i >> 10)), load_fe(lo + (i & 1023))));
}

int fr_powers_run(Ctx* c, const uint64_t base_mont[4], const uint64_t scale_mont[4], uint64_t n, fe* out,
                  hipStream_t st) {
  if (n == 0) return ZK_OK;
  zkh::Fe b, s;
  memcpy(b.v, base_mont, 32);
  memcpy(s.v, scale_mont, 32);
  std::vector<zkh::Fe> lo(1024), hi((n + 1023) / 1024);
  zkh::Fe x = zkh::one(zkh::FR);
  for (int i = 0; i < 1024; i++) {
    lo[i] = x;
    x = zkh::mul(zkh::FR, x, b);
  }
  zkh::Fe y = s;
  for (size_t h = 0; h < hi.size(); h++) {
    hi[h] = y;
    y = zkh::mul(zkh::FR, y, x);
  }
  void* p;
  int rc = ensure_scratch(c, 5, (hi.size() + 1024) * 32, &p);
  if (rc) return rc;
  ZK_HIP(c, hipMemcpyAsync(p, lo.data(), 1024 * 32, hipMemcpyHostToDevice, st));
  ZK_HIP(c, hipMemcpyAsync((char*)p + 1024 * 32, hi.data(), hi.size() * 32, hipMemcpyHostToDevice, st));
  fr_powers_kernel<<<grid_for(n, 256), 256, 0, st>>>((const fe*)p + 1024, (const fe*)p, n, out);
  ZK_HIP(c, hipGetLastError());
  ZK_HIP(c, hipStreamSynchronize(st));  // host tables die here
  return ZK_OK;
}

}  // namespace zk
