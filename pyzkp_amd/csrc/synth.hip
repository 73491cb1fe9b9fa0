// Synthetic Groth16 workload: the witness of the benchmark circuit (no
// counterpart in the reference, whose circuits come from its frontend).
//
// Circuit: `chains` independent chains of `rounds` rounds over a 3-lane
// state.  A round applies the S-box x -> x^5 to every lane (x2 = x x,
// x4 = x2 x2, x5 = x4 x: 9 constraints) and a 3-term linear mix
// y_k = sum_j M[k][j] x5_j (3 constraints, each (sum) * 1 = y_k).
// Variable ids: lane k of chain c's initial state is 3 c + k; round r of
// chain c uses base = 3 chains + 12 (c rounds + r): x2_k = base + k,
// x4_k = base + 3 + k, x5_k = base + 6 + k, y_k = base + 9 + k.  The
// constant-one variable follows the last id (one_id = m, as in
// protocols/groth16/r1cs.py:40-41).
#include <string.h>
#include <atomic>
#include <thread>
#include <vector>
#include "internal.hpp"

using namespace zk;

extern "C" {

/* w[0..m] (canonical little-endian words, 4 per element) for the circuit
 * above: mix = 9 canonical elements (row-major M), init = 3 chains
 * canonical elements; w[m] = 1. */
int zk_synth_witness(const uint64_t* mix, const uint64_t* init, uint32_t chains, uint32_t rounds, int threads,
                     uint64_t* w) {
  if (!mix || !init || !w) return ZK_ERR_INVALID;
  const uint64_t m = 3ull * chains + 12ull * chains * rounds;
  zkh::Fe M[9];
  for (int k = 0; k < 9; k++) {
    memcpy(M[k].v, mix + 4 * k, 32);
    M[k] = zkh::to_mont(zkh::FR, M[k]);
  }
  auto put = [&](uint64_t id, const zkh::Fe& v) {
    zkh::Fe cnv = zkh::from_mont(zkh::FR, v);
    memcpy(w + 4 * id, cnv.v, 32);
  };
  auto chain = [&](uint32_t c) {
    zkh::Fe x[3];
    for (int k = 0; k < 3; k++) {
      memcpy(x[k].v, init + 4 * (3ull * c + k), 32);
      x[k] = zkh::to_mont(zkh::FR, x[k]);
      put(3ull * c + k, x[k]);
    }
    for (uint32_t r = 0; r < rounds; r++) {
      const uint64_t base = 3ull * chains + 12ull * ((uint64_t)c * rounds + r);
      zkh::Fe x5[3];
      for (int k = 0; k < 3; k++) {
        zkh::Fe x2 = zkh::mul(zkh::FR, x[k], x[k]);
        zkh::Fe x4 = zkh::mul(zkh::FR, x2, x2);
        x5[k] = zkh::mul(zkh::FR, x4, x[k]);
        put(base + k, x2);
        put(base + 3 + k, x4);
        put(base + 6 + k, x5[k]);
      }
      for (int k = 0; k < 3; k++) {
        zkh::Fe y = zkh::zero();
        for (int j = 0; j < 3; j++) y = zkh::add(zkh::FR, y, zkh::mul(zkh::FR, M[3 * k + j], x5[j]));
        put(base + 9 + k, y);
        x[k] = y;
      }
    }
  };
  if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
  if (threads < 1) threads = 1;
  std::vector<std::thread> ts;
  std::atomic<uint32_t> next(0);
  for (int t = 0; t < threads && t < (int)chains; t++)
    ts.emplace_back([&]() {
      for (;;) {
        uint32_t c = next.fetch_add(1);
        if (c >= chains) break;
        chain(c);
      }
    });
  for (auto& t : ts) t.join();
  memset(w + 4 * m, 0, 32);
  w[4 * m] = 1;
  return ZK_OK;
}

}  // extern "C"
