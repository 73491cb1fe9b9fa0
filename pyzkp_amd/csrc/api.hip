// C-ABI entry points (declared in include/zkhip.h): context lifetime, error
// reporting, memory and copies, and argument checking in front of the
// NTT / MSM / element-wise drivers.
#include <string.h>
#include <stdio.h>
#include "internal.hpp"
#include "../../include/zkhip.h"

struct zk_ctx {
  zk::Ctx c;
};

namespace zk {

static thread_local std::string g_init_err;

int set_err(Ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int hip_check(Ctx* c, hipError_t e, const char* what) {
  char buf[512];
  snprintf(buf, sizeof(buf), "HIP error %d (%s) in %s", (int)e, hipGetErrorString(e), what);
  return set_err(c, ZK_ERR_RUNTIME, buf);
}

int ensure_aux(Ctx* c, size_t count) {
  if (!c->aux_stream) ZK_HIP(c, hipStreamCreateWithFlags(&c->aux_stream, hipStreamNonBlocking));
  while (c->events.size() < count) {
    hipEvent_t e;
    ZK_HIP(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    c->events.push_back(e);
  }
  return ZK_OK;
}

int ensure_scratch(Ctx* c, int slot, size_t bytes, void** out) {
  DevBuf& b = c->scratch[slot];
  if (b.bytes < bytes) {
    if (b.ptr) {
      // earlier launches on any stream may still read the old arena
      ZK_HIP(c, hipDeviceSynchronize());
      ZK_HIP(c, hipFree(b.ptr));
      b.ptr = nullptr;
      b.bytes = 0;
    }
    size_t want = bytes + bytes / 8;
    ZK_HIP(c, hipMalloc(&b.ptr, want));
    b.bytes = want;
  }
  *out = b.ptr;
  return ZK_OK;
}

int ensure_host_out(Ctx* c, size_t bytes, void** out) {
  if (c->host_out_bytes < bytes) {
    if (c->host_out) {
      ZK_HIP(c, hipHostFree(c->host_out));
      c->host_out = nullptr;
      c->host_out_bytes = 0;
    }
    const size_t want = bytes < 65536 ? 65536 : bytes;
    ZK_HIP(c, hipHostMalloc(&c->host_out, want, hipHostMallocDefault));
    c->host_out_bytes = want;
  }
  *out = c->host_out;
  return ZK_OK;
}

int upload_table(Ctx* c, const std::vector<zkh::Fe>& host, DevBuf* out, hipStream_t st) {
  out->bytes = host.size() * 32;
  ZK_HIP(c, hipMalloc(&out->ptr, out->bytes));
  // synchronous copy: the host vector dies when the caller returns
  ZK_HIP(c, hipStreamSynchronize(st));
  ZK_HIP(c, hipMemcpy(out->ptr, host.data(), out->bytes, hipMemcpyHostToDevice));
  return ZK_OK;
}

void phase_begin(Ctx* c, const char* name, hipStream_t st, hipEvent_t* ev) {
  (void)name;
  *ev = nullptr;
  if (!c->profiling) return;
  if (hipEventCreate(ev) != hipSuccess) {
    *ev = nullptr;
    return;
  }
  hipEventRecord(*ev, st);
}

void phase_end(Ctx* c, const char* name, hipStream_t st, hipEvent_t ev) {
  if (!c->profiling || ev == nullptr) return;
  hipEvent_t stop;
  if (hipEventCreate(&stop) != hipSuccess) return;
  hipEventRecord(stop, st);
  c->phases[name].pending.emplace_back(ev, stop);
}

static void collect(Ctx* c) {
  for (auto& kv : c->phases) {
    PhaseTimer& t = kv.second;
    for (auto& pr : t.pending) {
      hipEventSynchronize(pr.second);
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
        t.total_ms += ms;
        t.launches += 1;
      }
      hipEventDestroy(pr.first);
      hipEventDestroy(pr.second);
    }
    t.pending.clear();
  }
}

static zkh::Fe words_to_fe(const uint64_t w[4]) {
  zkh::Fe f;
  memcpy(f.v, w, 32);
  return f;
}

}  // namespace zk

using namespace zk;

#define CTX_CHECK(ctx)                       \
  do {                                       \
    if (!(ctx)) return ZK_ERR_INVALID;       \
    hipSetDevice((ctx)->c.device);           \
  } while (0)

static inline hipStream_t pick(zk_ctx* ctx, void* stream) {
  return stream ? (hipStream_t)stream : ctx->c.stream;
}

extern "C" {

int zk_version(void) { return 100; }

int zk_device_count(int* count) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  *count = n;
  return ZK_OK;
}

int zk_init(int ordinal, zk_ctx** out) {
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    g_init_err = "no HIP device available";
    return ZK_ERR_NO_DEVICE;
  }
  if (ordinal < 0 || ordinal >= n) {
    g_init_err = "device ordinal out of range";
    return ZK_ERR_INVALID;
  }
  hipError_t e = hipSetDevice(ordinal);
  if (e != hipSuccess) {
    g_init_err = std::string("hipSetDevice failed: ") + hipGetErrorString(e);
    return ZK_ERR_RUNTIME;
  }
  int cus = 0;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ordinal);
  if (e != hipSuccess || cus <= 0) {
    g_init_err = std::string("hipDeviceGetAttribute(compute units) failed: ") + hipGetErrorString(e);
    return ZK_ERR_RUNTIME;
  }
  zk_ctx* ctx = new zk_ctx();
  ctx->c.device = ordinal;
  ctx->c.num_cus = (uint32_t)cus;
  e = hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    g_init_err = std::string("hipStreamCreate failed: ") + hipGetErrorString(e);
    delete ctx;
    return ZK_ERR_RUNTIME;
  }
  // device address of the no-allocation MSM arena, looked up once per device
  if (int rc = msm_noalloc_init(&ctx->c)) {
    g_init_err = ctx->c.err;
    hipStreamDestroy(ctx->c.stream);
    delete ctx;
    return rc;
  }
  *out = ctx;
  return ZK_OK;
}

int zk_destroy(zk_ctx* ctx) {
  if (!ctx) return ZK_OK;
  hipSetDevice(ctx->c.device);
  hipDeviceSynchronize();
  collect(&ctx->c);
  for (auto& kv : ctx->c.twiddles) hipFree(kv.second.a.ptr);
  for (auto& kv : ctx->c.pow_tables) {
    hipFree(kv.second.a.ptr);
    hipFree(kv.second.b.ptr);
  }
  for (auto& b : ctx->c.scratch)
    if (b.ptr) hipFree(b.ptr);
  for (hipEvent_t e : ctx->c.events) hipEventDestroy(e);
  if (ctx->c.host_out) hipHostFree(ctx->c.host_out);
  if (ctx->c.aux_stream) hipStreamDestroy(ctx->c.aux_stream);
  hipStreamDestroy(ctx->c.stream);
  delete ctx;
  return ZK_OK;
}

const char* zk_last_error(const zk_ctx* ctx) {
  if (!ctx) return g_init_err.c_str();
  return ctx->c.err.c_str();
}

int zk_alloc(zk_ctx* ctx, size_t bytes, void** out) {
  CTX_CHECK(ctx);
  *out = nullptr;
  if (bytes == 0) bytes = 256;
  ZK_HIP(&ctx->c, hipMalloc(out, bytes));
  return ZK_OK;
}

int zk_free(zk_ctx* ctx, void* ptr) {
  CTX_CHECK(ctx);
  if (ptr) ZK_HIP(&ctx->c, hipFree(ptr));
  return ZK_OK;
}

int zk_pinned_alloc(zk_ctx* ctx, size_t bytes, void** out) {
  CTX_CHECK(ctx);
  *out = nullptr;
  if (bytes == 0) bytes = 256;
  ZK_HIP(&ctx->c, hipHostMalloc(out, bytes, hipHostMallocDefault));
  return ZK_OK;
}

int zk_pinned_free(zk_ctx* ctx, void* ptr) {
  CTX_CHECK(ctx);
  if (ptr) ZK_HIP(&ctx->c, hipHostFree(ptr));
  return ZK_OK;
}

int zk_h2d(zk_ctx* ctx, void* dst, const void* src, size_t bytes, void* stream) {
  CTX_CHECK(ctx);
  if (bytes == 0) return ZK_OK;
  ZK_HIP(&ctx->c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, pick(ctx, stream)));
  return ZK_OK;
}

int zk_d2h(zk_ctx* ctx, void* dst, const void* src, size_t bytes, void* stream) {
  CTX_CHECK(ctx);
  if (bytes == 0) return ZK_OK;
  ZK_HIP(&ctx->c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, pick(ctx, stream)));
  return ZK_OK;
}

int zk_d2d(zk_ctx* ctx, void* dst, const void* src, size_t bytes, void* stream) {
  CTX_CHECK(ctx);
  if (bytes == 0) return ZK_OK;
  ZK_HIP(&ctx->c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, pick(ctx, stream)));
  return ZK_OK;
}

int zk_memset_zero(zk_ctx* ctx, void* dst, size_t bytes, void* stream) {
  CTX_CHECK(ctx);
  if (bytes == 0) return ZK_OK;
  ZK_HIP(&ctx->c, hipMemsetAsync(dst, 0, bytes, pick(ctx, stream)));
  return ZK_OK;
}

int zk_sync(zk_ctx* ctx, void* stream) {
  CTX_CHECK(ctx);
  ZK_HIP(&ctx->c, hipStreamSynchronize(pick(ctx, stream)));
  return ZK_OK;
}

static int elt(zk_ctx* ctx, int op, const void* a, const void* b, void* out, uint64_t n, void* stream) {
  CTX_CHECK(ctx);
  if (n && (!a || !out)) return set_err(&ctx->c, ZK_ERR_INVALID, "null buffer");
  return elementwise(&ctx->c, op, (const fe*)a, (const fe*)b, (fe*)out, n, pick(ctx, stream));
}

int zk_fr_to_mont(zk_ctx* ctx, const void* in, void* out, uint64_t n, void* stream) {
  return elt(ctx, ELT_FR_TO_MONT, in, nullptr, out, n, stream);
}
int zk_fr_from_mont(zk_ctx* ctx, const void* in, void* out, uint64_t n, void* stream) {
  return elt(ctx, ELT_FR_FROM_MONT, in, nullptr, out, n, stream);
}
int zk_fq_to_mont(zk_ctx* ctx, const void* in, void* out, uint64_t n, void* stream) {
  return elt(ctx, ELT_FQ_TO_MONT, in, nullptr, out, n, stream);
}
int zk_fq_from_mont(zk_ctx* ctx, const void* in, void* out, uint64_t n, void* stream) {
  return elt(ctx, ELT_FQ_FROM_MONT, in, nullptr, out, n, stream);
}
int zk_fr_mul(zk_ctx* ctx, const void* a, const void* b, void* out, uint64_t n, void* stream) {
  if (n && !b) return ctx ? set_err(&ctx->c, ZK_ERR_INVALID, "null buffer") : ZK_ERR_INVALID;
  return elt(ctx, ELT_FR_MUL, a, b, out, n, stream);
}
int zk_fr_add(zk_ctx* ctx, const void* a, const void* b, void* out, uint64_t n, void* stream) {
  if (n && !b) return ctx ? set_err(&ctx->c, ZK_ERR_INVALID, "null buffer") : ZK_ERR_INVALID;
  return elt(ctx, ELT_FR_ADD, a, b, out, n, stream);
}
int zk_fr_sub(zk_ctx* ctx, const void* a, const void* b, void* out, uint64_t n, void* stream) {
  if (n && !b) return ctx ? set_err(&ctx->c, ZK_ERR_INVALID, "null buffer") : ZK_ERR_INVALID;
  return elt(ctx, ELT_FR_SUB, a, b, out, n, stream);
}

int zk_fr_scale(zk_ctx* ctx, const void* a, const uint64_t c_mont[4], void* out, uint64_t n,
                void* stream) {
  CTX_CHECK(ctx);
  if (!c_mont) return set_err(&ctx->c, ZK_ERR_INVALID, "null constant");
  return elementwise_scale(&ctx->c, (const fe*)a, c_mont, (fe*)out, n, pick(ctx, stream));
}

int zk_ntt(zk_ctx* ctx, const void* in, uint64_t in_len, void* out, uint32_t log_n, int inverse,
           const uint64_t omega_mont[4], const uint64_t shift_mont[4], void* stream) {
  CTX_CHECK(ctx);
  if (!out || (in_len && !in)) return set_err(&ctx->c, ZK_ERR_INVALID, "null buffer");
  if (inverse && in_len != (1ull << (log_n > 63 ? 63 : log_n)))
    return set_err(&ctx->c, ZK_ERR_INVALID, "inverse NTT needs in_len == n");
  zkh::Fe w, s;
  if (omega_mont) w = words_to_fe(omega_mont);
  if (shift_mont) s = words_to_fe(shift_mont);
  return ntt_run(&ctx->c, (const fe*)in, in_len, (fe*)out, log_n, 1, inverse != 0,
                 omega_mont ? &w : nullptr, shift_mont ? &s : nullptr, pick(ctx, stream),
                 ctx->c.ntt_max_pass_bits);
}

int zk_ntt_batch(zk_ctx* ctx, const void* in, void* out, uint32_t log_n, uint32_t batch,
                 int inverse, const uint64_t omega_mont[4], const uint64_t shift_mont[4],
                 void* stream) {
  CTX_CHECK(ctx);
  if (batch && (!in || !out)) return set_err(&ctx->c, ZK_ERR_INVALID, "null buffer");
  if (log_n > 28) return set_err(&ctx->c, ZK_ERR_INVALID, "domain too large for BN254 root of unity");
  zkh::Fe w, s;
  if (omega_mont) w = words_to_fe(omega_mont);
  if (shift_mont) s = words_to_fe(shift_mont);
  return ntt_run(&ctx->c, (const fe*)in, 1ull << log_n, (fe*)out, log_n, batch, inverse != 0,
                 omega_mont ? &w : nullptr, shift_mont ? &s : nullptr, pick(ctx, stream),
                 ctx->c.ntt_max_pass_bits);
}

int zk_msm_g1(zk_ctx* ctx, const void* bases, const void* scalars, uint64_t n, int window_bits,
              uint64_t out_jac_mont[12], void* stream) {
  CTX_CHECK(ctx);
  if (!out_jac_mont) return set_err(&ctx->c, ZK_ERR_INVALID, "null output");
  if (n && (!bases || !scalars)) return set_err(&ctx->c, ZK_ERR_INVALID, "null buffer");
  return msm_g1_run(&ctx->c, (const fe*)bases, (const fe*)scalars, n, window_bits, out_jac_mont,
                    pick(ctx, stream));
}

int zk_msm_bases_create(zk_ctx* ctx, const void* bases, uint64_t n, zk_msm_bases** out, void* stream) {
  CTX_CHECK(ctx);
  if (!out) return set_err(&ctx->c, ZK_ERR_INVALID, "null output");
  *out = nullptr;
  if (n && !bases) return set_err(&ctx->c, ZK_ERR_INVALID, "null buffer");
  MsmBases* b = nullptr;
  int rc = msm_bases_create(&ctx->c, (const fe*)bases, n, &b, pick(ctx, stream));
  *out = (zk_msm_bases*)b;
  return rc;
}

int zk_msm_bases_destroy(zk_ctx* ctx, zk_msm_bases* b) {
  CTX_CHECK(ctx);
  hipDeviceSynchronize();
  msm_bases_destroy((MsmBases*)b);
  return ZK_OK;
}

int zk_msm_g1_prepared(zk_ctx* ctx, const zk_msm_bases* b, const void* scalars, uint64_t n, int window_bits,
                       uint64_t out_jac_mont[12], void* stream) {
  CTX_CHECK(ctx);
  if (!out_jac_mont || !b) return set_err(&ctx->c, ZK_ERR_INVALID, "null argument");
  if (n && !scalars) return set_err(&ctx->c, ZK_ERR_INVALID, "null buffer");
  return msm_g1_prepared_run(&ctx->c, (const MsmBases*)b, (const fe*)scalars, n, window_bits, out_jac_mont,
                             pick(ctx, stream));
}

int zk_msm_g2_bases_create(zk_ctx* ctx, const uint64_t* bases, uint64_t n, zk_msm_g2_bases** out, void* stream) {
  CTX_CHECK(ctx);
  if (!out) return set_err(&ctx->c, ZK_ERR_INVALID, "null output");
  *out = nullptr;
  if (n && !bases) return set_err(&ctx->c, ZK_ERR_INVALID, "null buffer");
  MsmG2Bases* b = nullptr;
  int rc = msm_g2_bases_create(&ctx->c, bases, n, &b, pick(ctx, stream));
  *out = (zk_msm_g2_bases*)b;
  return rc;
}

int zk_msm_g2_bases_destroy(zk_ctx* ctx, zk_msm_g2_bases* b) {
  CTX_CHECK(ctx);
  hipDeviceSynchronize();
  msm_g2_bases_destroy((MsmG2Bases*)b);
  return ZK_OK;
}

int zk_msm_g2(zk_ctx* ctx, const zk_msm_g2_bases* b, const void* scalars, uint64_t n, int window_bits,
              uint64_t out[16], void* stream) {
  CTX_CHECK(ctx);
  if (!out || !b) return set_err(&ctx->c, ZK_ERR_INVALID, "null argument");
  if (n && !scalars) return set_err(&ctx->c, ZK_ERR_INVALID, "null buffer");
  return msm_g2_run(&ctx->c, (const MsmG2Bases*)b, (const fe*)scalars, n, window_bits, out, pick(ctx, stream));
}

int zk_msm_g1_noalloc(zk_ctx* ctx, const void* bases, const void* scalars, uint64_t n, int window_bits,
                      uint64_t out_jac_mont[12], void* stream) {
  CTX_CHECK(ctx);
  if (!out_jac_mont) return set_err(&ctx->c, ZK_ERR_INVALID, "null output");
  if (n && (!bases || !scalars)) return set_err(&ctx->c, ZK_ERR_INVALID, "null buffer");
  return msm_g1_noalloc_run(&ctx->c, (const fe*)bases, (const fe*)scalars, n, window_bits, out_jac_mont,
                            pick(ctx, stream));
}

int zk_g1_fixed_base(zk_ctx* ctx, const void* table, const void* scalars, uint64_t n, void* out,
                     void* stream) {
  CTX_CHECK(ctx);
  if (n && (!table || !scalars || !out)) return set_err(&ctx->c, ZK_ERR_INVALID, "null buffer");
  return g1_fixed_base_run(&ctx->c, (const fe*)table, (const fe*)scalars, n, (fe*)out, pick(ctx, stream));
}

int zk_fixed_bases_create(zk_ctx* ctx, const void* bases, uint64_t n, int window_bits,
                          zk_fixed_bases** out, void* stream) {
  CTX_CHECK(ctx);
  if (!out) return set_err(&ctx->c, ZK_ERR_INVALID, "null output");
  *out = nullptr;
  if (!bases) return set_err(&ctx->c, ZK_ERR_INVALID, "null buffer");
  FixedBases* fb = nullptr;
  int rc = fixed_bases_create(&ctx->c, (const fe*)bases, n, window_bits, &fb, pick(ctx, stream));
  *out = (zk_fixed_bases*)fb;
  return rc;
}

int zk_fixed_bases_destroy(zk_ctx* ctx, zk_fixed_bases* fb) {
  CTX_CHECK(ctx);
  hipDeviceSynchronize();
  fixed_bases_destroy((FixedBases*)fb);
  return ZK_OK;
}

int zk_fixed_bases_info(const zk_fixed_bases* fb, uint64_t* n, int* window_bits) {
  if (!fb) return ZK_ERR_INVALID;
  const FixedBases* f = (const FixedBases*)fb;
  if (n) *n = f->n;
  if (window_bits) *window_bits = f->c;
  return ZK_OK;
}

int zk_msm_g1_fixed_batch(zk_ctx* ctx, const zk_fixed_bases* fb, const void* scalars,
                          const uint64_t* offsets, const uint64_t* sizes, const uint64_t* base_offsets,
                          uint32_t count, uint64_t* out_jac_mont, void* stream) {
  CTX_CHECK(ctx);
  if (!fb) return set_err(&ctx->c, ZK_ERR_INVALID, "null fixed-base set");
  if (count && (!scalars || !offsets || !sizes || !out_jac_mont))
    return set_err(&ctx->c, ZK_ERR_INVALID, "null buffer");
  return msm_fixed_batch_run(&ctx->c, (const FixedBases*)fb, (const fe*)scalars, offsets, sizes, base_offsets,
                             count, out_jac_mont, pick(ctx, stream));
}

int zk_batch_create(zk_ctx* ctx, const zk_fixed_bases* fb, const void* msm_scalars,
                    const uint64_t* msm_offsets, const uint64_t* msm_sizes, const uint64_t* msm_base_offsets,
                    uint32_t n_msm, const zk_ntt_job* ntts, uint32_t n_ntt, int use_graph, zk_batch** out,
                    void* stream) {
  CTX_CHECK(ctx);
  if (!out) return set_err(&ctx->c, ZK_ERR_INVALID, "null output");
  *out = nullptr;
  if (n_msm && (!msm_scalars || !msm_offsets || !msm_sizes)) return set_err(&ctx->c, ZK_ERR_INVALID, "null buffer");
  if (n_ntt && !ntts) return set_err(&ctx->c, ZK_ERR_INVALID, "null NTT job list");
  std::vector<NttJob> jobs(n_ntt);
  for (uint32_t k = 0; k < n_ntt; k++) {
    if (!ntts[k].in || !ntts[k].out) return set_err(&ctx->c, ZK_ERR_INVALID, "null NTT buffer");
    jobs[k].in = (const fe*)ntts[k].in;
    jobs[k].out = (fe*)ntts[k].out;
    jobs[k].log_n = ntts[k].log_n;
    jobs[k].batch = ntts[k].batch;
    jobs[k].inverse = ntts[k].inverse != 0;
    jobs[k].has_omega = ntts[k].has_omega != 0;
    jobs[k].has_shift = ntts[k].has_shift != 0;
    jobs[k].omega = words_to_fe(ntts[k].omega_mont);
    jobs[k].shift = words_to_fe(ntts[k].shift_mont);
  }
  BatchPlan* p = nullptr;
  int rc = batch_create(&ctx->c, (const FixedBases*)fb, (const fe*)msm_scalars, msm_offsets, msm_sizes,
                        msm_base_offsets, n_msm, jobs.data(), n_ntt, use_graph != 0, &p, pick(ctx, stream));
  *out = (zk_batch*)p;
  return rc;
}

int zk_batch_run(zk_ctx* ctx, zk_batch* b, uint64_t* msm_out_jac_mont, void* stream) {
  CTX_CHECK(ctx);
  if (!b) return set_err(&ctx->c, ZK_ERR_INVALID, "null batch");
  return batch_run(&ctx->c, (BatchPlan*)b, msm_out_jac_mont, pick(ctx, stream));
}

int zk_batch_destroy(zk_ctx* ctx, zk_batch* b) {
  CTX_CHECK(ctx);
  batch_destroy((BatchPlan*)b);
  return ZK_OK;
}

int zk_fr_powers(zk_ctx* ctx, const uint64_t base_mont[4], const uint64_t scale_mont[4], uint64_t n, void* out,
                 void* stream) {
  CTX_CHECK(ctx);
  if (!base_mont || !scale_mont || (n && !out)) return set_err(&ctx->c, ZK_ERR_INVALID, "null buffer");
  return fr_powers_run(&ctx->c, base_mont, scale_mont, n, (fe*)out, pick(ctx, stream));
}

int zk_r1cs_eval(zk_ctx* ctx, const uint32_t* row_ptr, const uint32_t* cols, const void* coeffs, const void* w,
                 uint64_t rows, void* out, void* stream) {
  CTX_CHECK(ctx);
  if (rows && (!row_ptr || !w || !out)) return set_err(&ctx->c, ZK_ERR_INVALID, "null buffer");
  return r1cs_eval(&ctx->c, row_ptr, cols, (const fe*)coeffs, (const fe*)w, rows, (fe*)out, pick(ctx, stream));
}

int zk_fr_gather(zk_ctx* ctx, const void* src, const uint32_t* idx, uint64_t n, void* out, void* stream) {
  CTX_CHECK(ctx);
  if (n && (!src || !idx || !out)) return set_err(&ctx->c, ZK_ERR_INVALID, "null buffer");
  return fr_gather(&ctx->c, (const fe*)src, idx, n, (fe*)out, pick(ctx, stream));
}

int zk_fr_inv(zk_ctx* ctx, const void* a, void* out, uint64_t n, void* stream) {
  CTX_CHECK(ctx);
  if (n && (!a || !out)) return set_err(&ctx->c, ZK_ERR_INVALID, "null buffer");
  return fr_inv(&ctx->c, (const fe*)a, n, (fe*)out, pick(ctx, stream));
}

int zk_plonk_quotient_evals(zk_ctx* ctx, const void* const* ext, uint64_t n, uint64_t m,
                            const uint64_t omega_m[4], const uint64_t shift[4], const uint64_t shift_n[4],
                            const uint64_t alpha[4], const uint64_t beta[4], const uint64_t gamma[4],
                            const uint64_t k1[4], const uint64_t k2[4], void* num_out, void* zh_out,
                            void* stream) {
  CTX_CHECK(ctx);
  if (!ext || !num_out || !zh_out || !omega_m || !shift || !shift_n || !alpha || !beta || !gamma || !k1 || !k2)
    return set_err(&ctx->c, ZK_ERR_INVALID, "null argument");
  const fe* e[15];
  for (int k = 0; k < 15; k++) e[k] = (const fe*)ext[k];
  return plonk_quotient_evals(&ctx->c, e, n, m, words_to_fe(omega_m), words_to_fe(shift), words_to_fe(shift_n),
                              words_to_fe(alpha), words_to_fe(beta), words_to_fe(gamma), words_to_fe(k1),
                              words_to_fe(k2), (fe*)num_out, (fe*)zh_out, pick(ctx, stream));
}

int zk_fr_count_nonzero(zk_ctx* ctx, const void* a, uint64_t n, uint64_t* result, void* stream) {
  CTX_CHECK(ctx);
  if (!result || (n && !a)) return set_err(&ctx->c, ZK_ERR_INVALID, "null buffer");
  return fr_count_nonzero(&ctx->c, (const fe*)a, n, result, pick(ctx, stream));
}

int zk_set_tuning(zk_ctx* ctx, const char* key, int64_t value) {
  CTX_CHECK(ctx);
  if (!key) return set_err(&ctx->c, ZK_ERR_INVALID, "null key");
  std::string k(key);
  if (k == "ntt_max_pass_bits") {
    if (value < 0 || value > 12) return set_err(&ctx->c, ZK_ERR_INVALID, "ntt_max_pass_bits must be in [0, 12]");
    ctx->c.ntt_max_pass_bits = (int)value;
  } else if (k == "ntt_table_cap_mb") {
    if (value < 0) return set_err(&ctx->c, ZK_ERR_INVALID, "ntt_table_cap_mb must be >= 0");
    ctx->c.table_cap = (size_t)value << 20;
  } else if (k == "ntt_tile_bits") {
    if (value != 0 && value != 10 && value != 12)
      return set_err(&ctx->c, ZK_ERR_INVALID, "ntt_tile_bits must be 0, 10 or 12");
    ctx->c.ntt_tile_bits = (int)value;
  } else if (k == "ntt_xcd_group") {
    if (value < 0 || value > 64) return set_err(&ctx->c, ZK_ERR_INVALID, "ntt_xcd_group must be in [0, 64]");
    ctx->c.ntt_xcd_group = (int)value;
  } else if (k == "ntt_inplace") {
    if (value != 0 && value != 1) return set_err(&ctx->c, ZK_ERR_INVALID, "ntt_inplace must be 0 or 1");
    ctx->c.ntt_inplace = (int)value;
  } else if (k == "ntt_stage_twiddles") {
    if (value != 0 && value != 1) return set_err(&ctx->c, ZK_ERR_INVALID, "ntt_stage_twiddles must be 0 or 1");
    ctx->c.ntt_stage_twiddles = (int)value;
  } else if (k == "ntt_full_scale") {
    if (value != 0 && value != 1) return set_err(&ctx->c, ZK_ERR_INVALID, "ntt_full_scale must be 0 or 1");
    ctx->c.ntt_full_scale = (int)value;
  } else if (k == "msm_segment") {
    if (value < 0 || value > 4096) return set_err(&ctx->c, ZK_ERR_INVALID, "msm_segment must be in [0, 4096]");
    ctx->c.msm_segment = (uint32_t)value;
  } else if (k == "msm_reduce_chunk") {
    if (value < 0 || value > 65536) return set_err(&ctx->c, ZK_ERR_INVALID, "msm_reduce_chunk must be in [0, 65536]");
    ctx->c.msm_reduce_chunk = (uint32_t)value;
  } else if (k == "msm_reduce_split") {
    if (value != 0 && value != 1) return set_err(&ctx->c, ZK_ERR_INVALID, "msm_reduce_split must be 0 or 1");
    ctx->c.msm_reduce_split = (uint32_t)value;
  } else if (k == "msm_sort_fused") {
    if (value != 0 && value != 1) return set_err(&ctx->c, ZK_ERR_INVALID, "msm_sort_fused must be 0 or 1");
    ctx->c.msm_sort_fused = (uint32_t)value;
  } else if (k == "msm_groups") {
    if (value < 0 || value > 16) return set_err(&ctx->c, ZK_ERR_INVALID, "msm_groups must be in [0, 16]");
    ctx->c.msm_groups = (uint32_t)value;
  } else {
    return set_err(&ctx->c, ZK_ERR_INVALID, "unknown tuning key: " + k);
  }
  return ZK_OK;
}

int zk_probe_mad_rate(zk_ctx* ctx, double* mads_per_s) {
  CTX_CHECK(ctx);
  if (!mads_per_s) return set_err(&ctx->c, ZK_ERR_INVALID, "null output");
  return probe_mad_rate(&ctx->c, mads_per_s);
}

int zk_profile_enable(zk_ctx* ctx, int on) {
  CTX_CHECK(ctx);
  ctx->c.profiling = on != 0;
  return ZK_OK;
}

int zk_profile_get(zk_ctx* ctx, const char* phase, double* total_ms, uint64_t* launches) {
  CTX_CHECK(ctx);
  collect(&ctx->c);
  auto it = ctx->c.phases.find(phase ? phase : "");
  *total_ms = it == ctx->c.phases.end() ? 0.0 : it->second.total_ms;
  *launches = it == ctx->c.phases.end() ? 0 : it->second.launches;
  return ZK_OK;
}

int zk_profile_reset(zk_ctx* ctx) {
  CTX_CHECK(ctx);
  collect(&ctx->c);
  ctx->c.phases.clear();
  return ZK_OK;
}

}  // extern "C"
