// Device helpers for the Groth16 prover around the MSM / NTT path:
//   * R1CS row evaluation  out[j] = sum_k coeff[j,k] * w[col[j,k]]  (CSR),
//     the device form of eval_r1cs_vectors (protocols/groth16/r1cs.py:77-92);
//   * gather out[k] = w[idx[k]] (the aux-witness slice fed to the l-query MSM,
//     protocols/groth16/prove.py:105);
//   * element-wise inversion with 0 -> 0 (the BATCH_INV op,
//     runtime/kernels/cpu/kernels.py:253-269, crypto/field/batch.py:8-38);
//   * an all-zero check (used to reject witnesses that do not satisfy the
//     R1CS, the reference's non-zero-remainder error, protocols/groth16/qap.py:98-99).
#include "internal.hpp"

namespace zk {

__global__ __launch_bounds__(256) void r1cs_eval_kernel(const uint32_t* __restrict__ row_ptr,
                                                        const uint32_t* __restrict__ cols,
                                                        const fe* __restrict__ coeffs,
                                                        const fe* __restrict__ w, uint64_t rows,
                                                        fe* __restrict__ out) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= rows) return;
  fe acc = Fr::zero();
  for (uint32_t k = row_ptr[j]; k < row_ptr[j + 1]; k++)
    acc = Fr::add(acc, Fr::mul(load_fe(coeffs + k), load_fe(w + cols[k])));
  store_fe(out + j, acc);
}

__global__ __launch_bounds__(256) void gather_kernel(const fe* __restrict__ src, const uint32_t* __restrict__ idx,
                                                     uint64_t n, fe* __restrict__ out) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) store_fe(out + k, load_fe(src + idx[k]));
}

// Montgomery inverse: a R -> a^-1 R, by a^(r-2) (0 -> 0)
// Element-wise inverse with 0 -> 0 (crypto/field/batch.py:8-38), by
// simultaneous inversion: thread t takes the elements t, t + T, t + 2T, ...
// (coalesced across threads), keeps their running product in out[] on the
// way up, inverts the product once (a^(r-2)) and walks back down:
// about 3 products per element plus one exponentiation per thread.
__device__ fe fr_pow_r_minus_2(const fe& x) {
  const uint32_t e[8] = {0xefffffffu, 0x43e1f593u, 0x79b97091u, 0x2833e848u,
                         0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};  // r - 2
  fe r = Fr::one();
  for (int wi = 7; wi >= 0; wi--) {
    const uint32_t word = e[wi];
    for (int b = 31; b >= 0; b--) {
      r = Fr::sqr(r);
      if ((word >> b) & 1u) r = Fr::mul(r, x);
    }
  }
  return r;
}

// pre holds the running products (it is `out` itself unless a and out alias)
__global__ __launch_bounds__(256) void fr_inv_kernel(const fe* a, uint64_t n, uint64_t T, fe* pre, fe* out) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T || t >= n) return;
  fe acc = Fr::one();
  for (uint64_t i = t; i < n; i += T) {
    const fe x = load_fe(a + i);
    store_fe(pre + i, acc);  // product of the earlier non-zero elements
    if (!Fr::is_zero(x)) acc = Fr::mul(acc, x);
  }
  fe inv = fr_pow_r_minus_2(acc);
  const uint64_t last = t + ((n - 1 - t) / T) * T;
  for (uint64_t i = last + T; i >= t + T; i -= T) {
    const uint64_t k = i - T;
    const fe x = load_fe(a + k);
    const fe p = load_fe(pre + k);
    if (Fr::is_zero(x)) {
      store_fe(out + k, Fr::zero());
    } else {
      store_fe(out + k, Fr::mul(p, inv));
      inv = Fr::mul(inv, x);
    }
  }
}

__global__ __launch_bounds__(256) void count_nonzero_kernel(const fe* __restrict__ a, uint64_t n,
                                                            uint32_t* __restrict__ count) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && !Fr::is_zero(load_fe(a + i))) atomicAdd(count, 1u);
}

int r1cs_eval(Ctx* c, const uint32_t* row_ptr, const uint32_t* cols, const fe* coeffs, const fe* w, uint64_t rows,
              fe* out, hipStream_t st) {
  if (rows == 0) return ZK_OK;
  r1cs_eval_kernel<<<grid_for(rows, 256), 256, 0, st>>>(row_ptr, cols, coeffs, w, rows, out);
  ZK_HIP(c, hipGetLastError());
  return ZK_OK;
}

int fr_gather(Ctx* c, const fe* src, const uint32_t* idx, uint64_t n, fe* out, hipStream_t st) {
  if (n == 0) return ZK_OK;
  gather_kernel<<<grid_for(n, 256), 256, 0, st>>>(src, idx, n, out);
  ZK_HIP(c, hipGetLastError());
  return ZK_OK;
}

int fr_inv(Ctx* c, const fe* a, uint64_t n, fe* out, hipStream_t st) {
  if (n == 0) return ZK_OK;
  // about 64 elements per thread once there are enough of them for the device
  uint64_t T = n / 64;
  if (T < 65536) T = n < 65536 ? n : 65536;
  fe* pre = out;
  if ((const void*)a == (const void*)out) {
    void* p;
    int rc = ensure_scratch(c, 6, (size_t)n * 32, &p);
    if (rc) return rc;
    pre = (fe*)p;
  }
  fr_inv_kernel<<<grid_for(T, 256), 256, 0, st>>>(a, n, T, pre, out);
  ZK_HIP(c, hipGetLastError());
  return ZK_OK;
}

int fr_count_nonzero(Ctx* c, const fe* a, uint64_t n, uint64_t* result, hipStream_t st) {
  *result = 0;
  if (n == 0) return ZK_OK;
  void* p;
  int rc = ensure_scratch(c, 4, 256, &p);
  if (rc) return rc;
  ZK_HIP(c, hipMemsetAsync(p, 0, 4, st));
  count_nonzero_kernel<<<grid_for(n, 256), 256, 0, st>>>(a, n, (uint32_t*)p);
  ZK_HIP(c, hipGetLastError());
  uint32_t h = 0;
  ZK_HIP(c, hipMemcpyAsync(&h, p, 4, hipMemcpyDeviceToHost, st));
  ZK_HIP(c, hipStreamSynchronize(st));
  *result = h;
  return ZK_OK;
}

}  // namespace zk
