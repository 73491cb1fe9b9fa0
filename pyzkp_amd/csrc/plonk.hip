// PLONK quotient numerator on the extended coset, one fused element-wise
// pass: the device form of PLONK_T_QUOTIENT_EVALS
// (runtime/kernels/cpu/kernels.py:607-694).  For every point
// x_i = shift omega_m^i of the size-m coset it computes, from the 15 coset
// evaluation vectors (a, b, c, z, pi, ql, qr, qm, qo, qc, s1, s2, s3,
// z(omega X), L1):
//   gate  = ql a + qr b + qm a b + qo c + qc + pi
//   perm  = (a + beta x + gamma)(b + beta k1 x + gamma)(c + beta k2 x + gamma) z
//         - (a + beta s1 + gamma)(b + beta s2 + gamma)(c + beta s3 + gamma) z(omega x)
//   num_i = gate + alpha perm + alpha^2 (z - 1) L1
//   zh_i  = shift_n u^i - 1,  u = omega_m^n
// All vectors are Fr Montgomery (R = 2^256); x_i comes from the two-level
// power table of (omega_m, shift) that the coset NTT of the same shape uses.
#include "internal.hpp"
#include "field29.hpp"

namespace zk {

static constexpr uint32_t PLONK_MAX_RATIO = 16;  // m / n: u^i takes at most this many values

struct PlonkQArgs {
  const fe* ext[15];
  fe beta, gamma, alpha, alpha2, bk1, bk2;
  fe zh[PLONK_MAX_RATIO];
  uint32_t ratio;
};

__global__ __launch_bounds__(256) void plonk_quotient_kernel(const PlonkQArgs A, uint64_t m,
                                                             const fe* __restrict__ hi,
                                                             const fe* __restrict__ lo, fe* __restrict__ num,
                                                             fe* __restrict__ zh) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const f9 x9 = Fr9::mul(unpack9(load_fe(hi + (i >> POW_TABLE_LO_BITS))),
                         unpack9(load_fe(lo + (i & ((1u << POW_TABLE_LO_BITS) - 1)))));
  const fe x = pack9(Fr9::canonical(Fr9::to256(x9)));
  const fe a = load_fe(A.ext[0] + i), b = load_fe(A.ext[1] + i), c = load_fe(A.ext[2] + i);
  const fe z = load_fe(A.ext[3] + i), pi = load_fe(A.ext[4] + i);
  fe gate = Fr::add(Fr::mul(load_fe(A.ext[5] + i), a), Fr::mul(load_fe(A.ext[6] + i), b));
  gate = Fr::add(gate, Fr::mul(Fr::mul(load_fe(A.ext[7] + i), a), b));
  gate = Fr::add(gate, Fr::add(Fr::mul(load_fe(A.ext[8] + i), c), Fr::add(load_fe(A.ext[9] + i), pi)));
  const fe t1 = Fr::add(Fr::add(a, Fr::mul(A.beta, x)), A.gamma);
  const fe t2 = Fr::add(Fr::add(b, Fr::mul(A.bk1, x)), A.gamma);
  const fe t3 = Fr::add(Fr::add(c, Fr::mul(A.bk2, x)), A.gamma);
  const fe left = Fr::mul(Fr::mul(Fr::mul(t1, t2), t3), z);
  const fe u1 = Fr::add(Fr::add(a, Fr::mul(A.beta, load_fe(A.ext[10] + i))), A.gamma);
  const fe u2 = Fr::add(Fr::add(b, Fr::mul(A.beta, load_fe(A.ext[11] + i))), A.gamma);
  const fe u3 = Fr::add(Fr::add(c, Fr::mul(A.beta, load_fe(A.ext[12] + i))), A.gamma);
  const fe right = Fr::mul(Fr::mul(Fr::mul(u1, u2), u3), load_fe(A.ext[13] + i));
  const fe perm = Fr::sub(left, right);
  const fe boundary = Fr::mul(Fr::sub(z, Fr::one()), load_fe(A.ext[14] + i));
  const fe v = Fr::add(gate, Fr::add(Fr::mul(A.alpha, perm), Fr::mul(A.alpha2, boundary)));
  store_fe(num + i, v);
  store_fe(zh + i, A.zh[i % A.ratio]);
}

static fe to_fe(const zkh::Fe& x) {
  fe r;
  memcpy(r.w, x.v, 32);
  return r;
}

int plonk_quotient_evals(Ctx* c, const fe* const ext[15], uint64_t n, uint64_t m, const zkh::Fe& omega_m,
                         const zkh::Fe& shift, const zkh::Fe& shift_n, const zkh::Fe& alpha,
                         const zkh::Fe& beta, const zkh::Fe& gamma, const zkh::Fe& k1, const zkh::Fe& k2,
                         fe* num, fe* zh, hipStream_t st) {
  if (m == 0 || (m & (m - 1))) return set_err(c, ZK_ERR_INVALID, "m must be a power of two");
  if (n == 0 || m % n != 0 || m / n > PLONK_MAX_RATIO)
    return set_err(c, ZK_ERR_INVALID, "m must be a multiple of n by at most 16");
  uint32_t log_m = 0;
  while ((1ull << log_m) < m) log_m++;
  if (log_m > 28) return set_err(c, ZK_ERR_INVALID, "domain too large for BN254 root of unity");
  for (int k = 0; k < 15; k++)
    if (!ext[k]) return set_err(c, ZK_ERR_INVALID, "null evaluation vector");
  PlonkQArgs A;
  for (int k = 0; k < 15; k++) A.ext[k] = ext[k];
  A.beta = to_fe(beta);
  A.gamma = to_fe(gamma);
  A.alpha = to_fe(alpha);
  A.alpha2 = to_fe(zkh::mul(zkh::FR, alpha, alpha));
  A.bk1 = to_fe(zkh::mul(zkh::FR, beta, k1));
  A.bk2 = to_fe(zkh::mul(zkh::FR, beta, k2));
  A.ratio = (uint32_t)(m / n);
  // zh_j = shift_n u^j - 1, u = omega_m^n (u^(m / n) = 1)
  const zkh::Fe u = zkh::pow_u64(zkh::FR, omega_m, n);
  zkh::Fe up = zkh::one(zkh::FR);
  for (uint32_t j = 0; j < A.ratio; j++) {
    A.zh[j] = to_fe(zkh::sub(zkh::FR, zkh::mul(zkh::FR, shift_n, up), zkh::one(zkh::FR)));
    up = zkh::mul(zkh::FR, up, u);
  }
  const fe *hi, *lo;
  int rc = ntt_pow_tables(c, log_m, omega_m, shift, &hi, &lo, st);
  if (rc) return rc;
  plonk_quotient_kernel<<<grid_for(m, 256), 256, 0, st>>>(A, m, hi, lo, num, zh);
  ZK_HIP(c, hipGetLastError());
  return ZK_OK;
}

}  // namespace zk
