// Interface between the host G2 MSM (g2_host.hip) and its AVX-512 IFMA
// bucket accumulation (g2_simd.cpp, the only translation unit built with
// AVX-512 code generation).  Nothing here names a vector type, so the HIP
// translation units include it without x86 vector headers; g2_host.hip calls
// into g2_simd.cpp only after checking that the CPU has AVX-512F and IFMA.
#pragma once
#include <stdint.h>

namespace zk {
namespace g2simd {

// Points in the accumulation's form: affine (x0, x1, y0, y1), each an Fq
// value in Montgomery form with R = 2^260, canonical (< q), as five 52-bit
// limbs: 20 words per point.
constexpr int PT_WORDS = 20;

// The CPU and OS support AVX-512F and AVX-512 IFMA (host_cpu.cpp).
bool cpu_has_ifma();

// out[20 i ..] <- point i of bases (16 canonical words each, all-zero =
// infinity), inf[i] <- 1 for the point at infinity (its 20 words are zero).
void prepare_points(const uint64_t* bases16, uint64_t n, uint64_t* out20, uint8_t* inf);

// Bucket sums of window w (c bits) over points [i0, i1): the signed digit d
// of scalar i (4 words; bits [w c, w c + c) of (s mod r) + K, minus
// 2^(c-1), K = sum over the windows of 2^(w c + c - 1), 5 words) adds
// sign(d) P_i to bucket |d| - 1 of nb.  bk20 (nb x 20 words) and empty (nb)
// are the buckets, in and out.  Additions are affine, batched so that one
// inversion serves a whole batch; any digit distribution is handled (a point
// whose bucket is busy goes to a replica of it, and replicas are added back
// at the end).  Returns 0, or UINT64_MAX when out of memory.
uint64_t accumulate(const uint64_t* pts20, const uint8_t* inf, const uint64_t* scalars, const uint64_t* K,
                    uint32_t w, uint32_t c, uint64_t i0, uint64_t i1, uint32_t nb, uint64_t* bk20, uint8_t* empty);

// The weighted sum sum_b (b + 1) B_b of nb affine buckets (bk20 / empty as
// accumulate leaves them), eight segments at a time: lane l runs XYZZ running
// sums over buckets [l L, (l + 1) L), L = ceil(nb / 8), and returns
//   sum_l = sum_j (j + 1) B_(l L + j)  and  run_l = sum_j B_(l L + j)
// as Jacobian points (X ZZ^2, Y ZZZ^2, ZZZ): 24 words per lane (X0 X1 Y0 Y1
// Z0 Z1, canonical, Montgomery with R = 2^256), with their infinity flags.
// The full sum is sum_l sum_l + L sum_l l run_l.  Returns L.
uint32_t reduce_segments(const uint64_t* bk20, const uint8_t* empty, uint32_t nb, uint64_t* out_sum,
                         uint64_t* out_run, uint8_t* sum_inf, uint8_t* run_inf);

}  // namespace g2simd
}  // namespace zk
