// Fixed-base scalar multiplication of one point by many scalars, with
// affine output: out_i = [k_i] P.  The host supplies the table
// T[j][d] = [d * 2^(8j)] P (j < 32, d < 256, affine Montgomery), so each
// product is at most 32 mixed additions followed by one Fermat inversion.
//
// This is the device counterpart of the setup-time loops that build G1
// query vectors and SRS powers one g1_mul at a time
// (crypto/kzg/cpu_ref.py:28-39 setup_srs, protocols/groth16/setup.py:63-96),
// and of the fixed-base tables of crypto/msm/pippenger.py:252-272.  The
// benchmarks use it to build bases B_i = [k_i] G1 on the device.
#include "internal.hpp"
#include "curve.hpp"

namespace zk {

// a^(q-2) mod q (Fermat inverse, 0 -> 0), square-and-multiply over the
// fixed exponent words
__device__ __forceinline__ fe fq_inv(const fe& a) {
  // q - 2, little-endian 32-bit words
  const uint32_t e[8] = {0xd87cfd45u, 0x3c208c16u, 0x6871ca8du, 0x97816a91u,
                         0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};
  fe r = Fq::one();
  for (int wi = 7; wi >= 0; wi--) {
    const uint32_t word = e[wi];
    for (int b = 31; b >= 0; b--) {
      r = Fq::sqr(r);
      if ((word >> b) & 1u) r = Fq::mul(r, a);
    }
  }
  return r;
}

__global__ __launch_bounds__(128) void g1_fixed_base_kernel(const fe* __restrict__ table,
                                                            const fe* __restrict__ scalars,
                                                            uint64_t n, fe* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const fe k = Fr::from_mont(load_fe(scalars + i));
  xyzz acc = xyzz_inf();
#pragma unroll 1
  for (int j = 0; j < 32; j++) {
    const uint32_t d = (k.w[j >> 2] >> (8 * (j & 3))) & 0xffu;
    if (d) {
      const fe* t = table + 2 * ((uint64_t)j * 256 + d);
      aff a;
      a.x = load_fe(t);
      a.y = load_fe(t + 1);
      if (!aff_is_inf(a)) xyzz_madd(acc, a);
    }
  }
  fe x = Fq::zero(), y = Fq::zero();
  if (!xyzz_is_inf(acc)) {
    const fe inv = fq_inv(Fq::mul(acc.zz, acc.zzz));  // 1 / (ZZ ZZZ)
    x = Fq::mul(acc.x, Fq::mul(inv, acc.zzz));       // X / ZZ
    y = Fq::mul(acc.y, Fq::mul(inv, acc.zz));        // Y / ZZZ
  }
  store_fe(out + 2 * i, x);
  store_fe(out + 2 * i + 1, y);
}

int g1_fixed_base_run(Ctx* c, const fe* table, const fe* scalars, uint64_t n, fe* out, hipStream_t st) {
  if (n == 0) return ZK_OK;
  g1_fixed_base_kernel<<<grid_for(n, 128), 128, 0, st>>>(table, scalars, n, out);
  ZK_HIP(c, hipGetLastError());
  return ZK_OK;
}

}  // namespace zk
