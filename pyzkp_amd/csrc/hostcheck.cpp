// CPU build of the device field / curve arithmetic (field.hpp, curve.hpp are
// __host__ __device__).  Exposes a few entry points so the host-logic tests
// can check the exact code the kernels run against the oracle without a GPU.
#include <string.h>
#include "field.hpp"
#include "curve.hpp"

using namespace zk;

static fe rd(const uint64_t* w) {
  fe r;
  for (int i = 0; i < 4; i++) {
    r.w[2 * i] = (uint32_t)w[i];
    r.w[2 * i + 1] = (uint32_t)(w[i] >> 32);
  }
  return r;
}
static void wr(uint64_t* w, const fe& r) {
  for (int i = 0; i < 4; i++) w[i] = (uint64_t)r.w[2 * i] | ((uint64_t)r.w[2 * i + 1] << 32);
}
static xyzz rdp(const uint64_t* w) {
  xyzz p;
  p.x = rd(w);
  p.y = rd(w + 4);
  p.zz = rd(w + 8);
  p.zzz = rd(w + 12);
  return p;
}
static void wrp(uint64_t* w, const xyzz& p) {
  wr(w, p.x);
  wr(w + 4, p.y);
  wr(w + 8, p.zz);
  wr(w + 12, p.zzz);
}

extern "C" {
// op: 0 mul, 1 add, 2 sub, 3 neg, 4 to_mont, 5 from_mont ; field: 0 Fr, 1 Fq
int zkhost_field(int field, int op, const uint64_t* a, const uint64_t* b, uint64_t* out, uint64_t n) {
  for (uint64_t i = 0; i < n; i++) {
    fe x = rd(a + 4 * i), y = b ? rd(b + 4 * i) : Fr::zero(), r;
    if (field == 0) {
      switch (op) {
        case 0: r = Fr::mul(x, y); break;
        case 1: r = Fr::add(x, y); break;
        case 2: r = Fr::sub(x, y); break;
        case 3: r = Fr::neg(x); break;
        case 4: r = Fr::to_mont(Fr::reduce_full(x)); break;
        case 5: r = Fr::from_mont(x); break;
        default: return 1;
      }
    } else {
      switch (op) {
        case 0: r = Fq::mul(x, y); break;
        case 1: r = Fq::add(x, y); break;
        case 2: r = Fq::sub(x, y); break;
        case 3: r = Fq::neg(x); break;
        case 4: r = Fq::to_mont(Fq::reduce_full(x)); break;
        case 5: r = Fq::from_mont(x); break;
        default: return 1;
      }
    }
    wr(out + 4 * i, r);
  }
  return 0;
}

// XYZZ ops on Montgomery Fq points (16 words each):
// op 0: add(p, q)  1: dbl(p)  2: madd(p, affine q = words 0..7 of q)  3: mul_small(p, k)
int zkhost_xyzz(int op, const uint64_t* p, const uint64_t* q, uint32_t k, uint64_t* out) {
  xyzz P = rdp(p), R;
  switch (op) {
    case 0: R = xyzz_add(P, rdp(q)); break;
    case 1: R = xyzz_dbl(P); break;
    case 2: {
      aff a;
      a.x = rd(q);
      a.y = rd(q + 4);
      R = P;
      if (!aff_is_inf(a)) xyzz_madd(R, a);
      break;
    }
    case 3: R = xyzz_mul_small(P, k); break;
    default: return 1;
  }
  wrp(out, R);
  return 0;
}
}
