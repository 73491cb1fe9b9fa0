// Library-internal declarations shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <map>
#include <string>
#include <vector>
#include <array>

#include "field.hpp"
#include "host_field.hpp"

// Status codes of the C-ABI (mirrored in include/zkhip.h).
#define ZK_OK 0
#define ZK_ERR_INVALID 1
#define ZK_ERR_RUNTIME 2
#define ZK_ERR_NO_DEVICE 3

namespace zk {

// A device allocation owned by the context (tables, scratch).
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
};

// Per-phase device timing, recorded with HIP events on the launch stream
// when profiling is switched on for a context.
struct PhaseTimer {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;  // (start, stop)
  double total_ms = 0.0;
  uint64_t launches = 0;
};

// A cached NTT table (twiddles use `a` only; power tables are the pair
// (hi, lo)).  `last_call` is the table call that used it last; `pinned`
// tables (built for a batch plan, whose recorded launches keep their
// addresses) are never evicted.
struct CachedTable {
  DevBuf a, b;
  uint64_t last_call = 0;
  bool pinned = false;
  size_t bytes() const { return a.bytes + b.bytes; }
};

struct Ctx {
  int device = 0;
  uint32_t num_cus = 0;  // compute units of the device
  hipStream_t stream = nullptr;
  std::string err;
  // NTT twiddle tables omega^e, e <= n/2 (both directions), keyed by (log_n, omega Montgomery words)
  std::map<std::array<uint64_t, 5>, CachedTable> twiddles;
  // two-level power tables (hi, lo) for coset / inverse scaling,
  // keyed by (log_n, base words, scale words)
  std::map<std::array<uint64_t, 9>, CachedTable> pow_tables;
  // table cache accounting: bytes held, the byte cap (0: none; least
  // recently used unpinned tables are freed to stay under it) and the
  // number of the current table call (tables it uses are never freed)
  size_t table_bytes = 0;
  size_t table_cap = 16ull << 30;
  uint64_t table_call = 0;
  bool table_pin = false;  // tables fetched now are pinned (batch plans)
  // growable scratch arenas
  DevBuf scratch[8];  // 0-1 NTT, 2-3 MSM, 4 counters, 5 powers, 6 inversion, 7 PLONK quotient
  // tuning knobs (set through zk_set_tuning)
  int ntt_max_pass_bits = 0;      // 0: planner default
  int ntt_tile_bits = 0;          // 0: 2^10-element tiles unless a pass has more than 10 stages
  int ntt_xcd_group = 0;          // 0: group tiles per XCD only for one- or two-column tiles
  int ntt_inplace = 1;            // in-place pass stages (bit-reversed tile load, one barrier per step); 0: Stockham order
  int ntt_stage_twiddles = 1;     // per-stage twiddle tables (adjacent columns read adjacent entries); 0: one table of w^e
  int ntt_full_scale = 1;         // coset scaling from a full table per (size, shift) up to 2^26; 0: two-level tables
  uint32_t msm_segment = 0;       // sorted (bucket, point) pairs per accumulation thread (0: by launch size)
  uint32_t msm_reduce_chunk = 0;  // buckets per reduction thread (0: sized to fill the device)
  uint32_t msm_groups = 0;        // window groups pipelined over two streams (0: chosen per shape)
  uint32_t msm_reduce_split = 1;  // single-launch MSMs: digit-split bucket reduction (0: chunked running sums)
  uint32_t msm_sort_fused = 1;    // bucket sort: segments one workgroup holds placed whole (0: all chunked)
  // second stream (default priority) and events for pipelined launches, created on first use
  hipStream_t aux_stream = nullptr;
  // pinned host staging for small results read back within one call (MSM window terms)
  void* host_out = nullptr;
  size_t host_out_bytes = 0;
  std::vector<hipEvent_t> events;
  bool profiling = false;
  std::map<std::string, PhaseTimer> phases;
};

// error helpers --------------------------------------------------------------
int set_err(Ctx* c, int code, const std::string& msg);
int hip_check(Ctx* c, hipError_t e, const char* what);
#define ZK_HIP(ctx, call)                                          \
  do {                                                             \
    hipError_t _e = (call);                                        \
    if (_e != hipSuccess) return ::zk::hip_check((ctx), _e, #call); \
  } while (0)

// device memory helpers (context-owned) -------------------------------------
int ensure_scratch(Ctx* c, int slot, size_t bytes, void** out);
// c->host_out grown to at least `bytes` (pinned; only used synchronously inside one call)
int ensure_host_out(Ctx* c, size_t bytes, void** out);
// the context's second stream and at least `count` timing-free events
int ensure_aux(Ctx* c, size_t count);
int upload_table(Ctx* c, const std::vector<zkh::Fe>& host, DevBuf* out, hipStream_t st);

// profiling helpers: bracket a launch with events when profiling is on
void phase_begin(Ctx* c, const char* name, hipStream_t st, hipEvent_t* ev);
void phase_end(Ctx* c, const char* name, hipStream_t st, hipEvent_t ev);

// NTT ----------------------------------------------------------------------
int ntt_run(Ctx* c, const fe* in, uint64_t in_len, fe* out, uint32_t log_n, uint32_t batch,
            bool inverse, const zkh::Fe* omega, const zkh::Fe* shift, hipStream_t st,
            int max_pass_bits, fe* tmp_buf = nullptr);
// two-level power tables hi[h] = scale base^(2^POW_TABLE_LO_BITS h), lo[i] = base^i
// (R = 2^261 packed), cached per context
static constexpr uint32_t POW_TABLE_LO_BITS = 10;
int ntt_pow_tables(Ctx* c, uint32_t log_n, const zkh::Fe& base, const zkh::Fe& scale, const fe** hi,
                   const fe** lo, hipStream_t st);
// PLONK quotient numerator and Z_H on the extended coset
int plonk_quotient_evals(Ctx* c, const fe* const ext[15], uint64_t n, uint64_t m, const zkh::Fe& omega_m,
                         const zkh::Fe& shift, const zkh::Fe& shift_n, const zkh::Fe& alpha,
                         const zkh::Fe& beta, const zkh::Fe& gamma, const zkh::Fe& k1, const zkh::Fe& k2,
                         fe* num, fe* zh, hipStream_t st);
int ntt_prepare(Ctx* c, uint32_t log_n, bool inverse, const zkh::Fe* omega, const zkh::Fe* shift,
                hipStream_t st);

// one NTT job of a batch plan (out of place, `batch` transforms back to back)
struct NttJob {
  const fe* in = nullptr;
  fe* out = nullptr;
  uint32_t log_n = 0;
  uint32_t batch = 1;
  bool inverse = false;
  bool has_omega = false;
  bool has_shift = false;
  zkh::Fe omega, shift;
};

// element-wise -----------------------------------------------------------------
enum EltOp { ELT_FR_TO_MONT = 0, ELT_FR_FROM_MONT, ELT_FQ_TO_MONT, ELT_FQ_FROM_MONT,
             ELT_FR_MUL, ELT_FR_ADD, ELT_FR_SUB };
int elementwise(Ctx* c, int op, const fe* a, const fe* b, fe* out, uint64_t n, hipStream_t st);
int elementwise_scale(Ctx* c, const fe* a, const uint64_t c_mont[4], fe* out, uint64_t n, hipStream_t st);
int fr_powers_run(Ctx* c, const uint64_t base_mont[4], const uint64_t scale_mont[4], uint64_t n, fe* out,
                  hipStream_t st);

// MSM ------------------------------------------------------------------------
// A base set prepared for fixed-base MSMs: table[w][i] = 2^(c w) B_i.
struct FixedBases {
  uint64_t n = 0;
  int c = 0;
  uint32_t W = 0;
  DevBuf table;
};
int fixed_bases_create(Ctx* c, const fe* bases, uint64_t n, int window_bits, FixedBases** out,
                       hipStream_t st);
// sum_w 2^(c w) S_w over an MSM's XYZZ window sums, on the host (g1_host.hip)
struct xyzz;
void g1_window_combine(const xyzz* s, uint32_t W, int c, uint64_t out[12]);
void g1_bits_combine(const xyzz* t, const uint32_t* pos, uint32_t count, uint64_t out[12]);
// G2: terms as 32 words each (x, y, zz, zzz over Fq2, R = 2^256 Montgomery,
// canonical; zz = 0 is the identity) at strictly decreasing bit positions;
// out: affine canonical plain words (x0, x1, y0, y1), zeros for the identity
void g2_bits_combine(const uint64_t* t, const uint32_t* pos, uint32_t count, uint64_t out[16]);
struct MsmG2Bases;
int msm_g2_bases_create(Ctx* c, const uint64_t* host_words, uint64_t n, MsmG2Bases** out, hipStream_t st);
void msm_g2_bases_destroy(MsmG2Bases* b);
int msm_g2_run(Ctx* c, const MsmG2Bases* b, const fe* scalars, uint64_t n, int window_bits, uint64_t out[16],
               hipStream_t st);
void fixed_bases_destroy(FixedBases* fb);
int msm_fixed_batch_run(Ctx* c, const FixedBases* fb, const fe* scalars, const uint64_t* offsets,
                        const uint64_t* sizes, const uint64_t* base_offsets, uint32_t count, uint64_t* out_jac,
                        hipStream_t st);
struct BatchPlan;
int batch_create(Ctx* c, const FixedBases* fb, const fe* scalars, const uint64_t* offsets,
                 const uint64_t* sizes, const uint64_t* base_offsets, uint32_t n_msm, const NttJob* jobs,
                 uint32_t n_ntt, bool use_graph, BatchPlan** out, hipStream_t st);
int batch_run(Ctx* c, BatchPlan* p, uint64_t* out_jac, hipStream_t st);
void batch_destroy(BatchPlan* p);
int msm_g1_run(Ctx* c, const fe* bases, const fe* scalars, uint64_t n, int window_bits,
               uint64_t out_jac[12], hipStream_t st);
// base sets converted once for repeated MSMs
struct MsmBases {
  uint64_t n = 0;
  DevBuf conv;  // R = 2^261 affine coordinates, 64 bytes per point
};
int msm_bases_create(Ctx* c, const fe* bases, uint64_t n, MsmBases** out, hipStream_t st);
void msm_bases_destroy(MsmBases* b);
int msm_g1_prepared_run(Ctx* c, const MsmBases* b, const fe* scalars, uint64_t n, int window_bits,
                        uint64_t out_jac[12], hipStream_t st);
// the same MSM on a file-scope device arena: no device allocation
int msm_noalloc_init(Ctx* c);
int msm_g1_noalloc_run(Ctx* c, const fe* bases, const fe* scalars, uint64_t n, int window_bits,
                       uint64_t out_jac[12], hipStream_t st);

// fixed-base scalar multiplication --------------------------------------------
int g1_fixed_base_run(Ctx* c, const fe* table, const fe* scalars, uint64_t n, fe* out, hipStream_t st);

// Groth16 helpers -------------------------------------------------------------
int r1cs_eval(Ctx* c, const uint32_t* row_ptr, const uint32_t* cols, const fe* coeffs, const fe* w, uint64_t rows,
              fe* out, hipStream_t st);
int fr_gather(Ctx* c, const fe* src, const uint32_t* idx, uint64_t n, fe* out, hipStream_t st);
int fr_inv(Ctx* c, const fe* a, uint64_t n, fe* out, hipStream_t st);
int fr_count_nonzero(Ctx* c, const fe* a, uint64_t n, uint64_t* result, hipStream_t st);

// measurement probe ---------------------------------------------------------
int probe_mad_rate(Ctx* c, double* mads_per_s);

inline unsigned grid_for(uint64_t n, unsigned block) {
  uint64_t g = (n + block - 1) / block;
  return (unsigned)(g == 0 ? 1 : g);
}

}  // namespace zk
