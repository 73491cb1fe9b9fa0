/*
 * zkhip — C ABI of the MI355X (gfx950) prover hot path: BN254 G1 MSM and
 * Fr NTT / iNTT / coset NTT, plus the device-buffer and conversion calls the
 * operator layer needs.
 *
 * Conventions
 *   - Every call returns ZK_OK (0) or an error code; zk_last_error(ctx)
 *     gives the message.  ZK_ERR_INVALID maps to ValueError and the other
 *     codes to RuntimeError in the Python operator layer.
 *   - Field elements are 32 bytes: 4 x u64 (equivalently 8 x u32) little-
 *     endian words.  Device buffers hold Montgomery form (x * 2^256 mod p),
 *     the same layout as the reference's device buffers
 *     ("fr_mont_u64x4", runtime/metal/buffer.py:8-15,
 *     runtime/kernels/cpu/kernels.py:98-144).
 *   - G1 bases are affine Montgomery (x, y), 64 bytes each; x = y = 0 marks
 *     the point at infinity.  MSM results are Jacobian Montgomery (X, Y, Z),
 *     12 x u64, with Z = 0 for infinity — the layout the reference unpacks
 *     in runtime/metal/msm.py:36-63.
 *   - `stream` is a hipStream_t (NULL = the context's own stream).  Calls are
 *     stream-ordered; the caller owns every pointer it passes in.  A context
 *     is used from one host thread at a time.
 */
#ifndef ZKHIP_H
#define ZKHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ZK_OK 0
#define ZK_ERR_INVALID 1
#define ZK_ERR_RUNTIME 2
#define ZK_ERR_NO_DEVICE 3

typedef struct zk_ctx zk_ctx;
typedef struct zk_fixed_bases zk_fixed_bases;
typedef struct zk_batch zk_batch;
typedef struct zk_msm_bases zk_msm_bases;
typedef struct zk_msm_g2_bases zk_msm_g2_bases;
typedef struct zk_g2_bases zk_g2_bases;

/* One NTT job of a batch: `batch` transforms of size 2^log_n laid out back
 * to back, out of place (in != out), with the optional root / coset shift
 * of zk_ntt. */
typedef struct zk_ntt_job {
  const void* in;
  void* out;
  uint32_t log_n;
  uint32_t batch;
  int inverse;
  int has_omega;
  int has_shift;
  uint64_t omega_mont[4];
  uint64_t shift_mont[4];
} zk_ntt_job;

/* Library version (major * 10000 + minor * 100 + patch). */
int zk_version(void);

/* Number of visible HIP devices (0 on a machine without a GPU). */
int zk_device_count(int* count);

/* Create / destroy a context on device `ordinal`.  Replaces the reference's
 * MetalContext.create_default (runtime/context.py:31-36) and
 * MetalRuntime.create_default (runtime/metal/runtime.py:494-629). */
int zk_init(int ordinal, zk_ctx** out);
int zk_destroy(zk_ctx* ctx);
/* Message of the last failing call on ctx (or of the last failing zk_init
 * when ctx is NULL). */
const char* zk_last_error(const zk_ctx* ctx);

/* Device and pinned-host memory; replace MemoryPool.alloc_metal /
 * MTLBuffer storage (runtime/memory.py:75-110). */
int zk_alloc(zk_ctx* ctx, size_t bytes, void** out);
int zk_free(zk_ctx* ctx, void* ptr);
int zk_pinned_alloc(zk_ctx* ctx, size_t bytes, void** out);
int zk_pinned_free(zk_ctx* ctx, void* ptr);
int zk_h2d(zk_ctx* ctx, void* dst, const void* src, size_t bytes, void* stream);
int zk_d2h(zk_ctx* ctx, void* dst, const void* src, size_t bytes, void* stream);
int zk_d2d(zk_ctx* ctx, void* dst, const void* src, size_t bytes, void* stream);
int zk_memset_zero(zk_ctx* ctx, void* dst, size_t bytes, void* stream);
int zk_sync(zk_ctx* ctx, void* stream);

/* Canonical <-> Montgomery conversion of n elements on the device; replace
 * the host loops of _to_device / _from_device
 * (runtime/kernels/cpu/kernels.py:135-140, 168-170) and of
 * _pack_g1_point (runtime/metal/msm.py:13-34). Inputs >= p are reduced. */
int zk_fr_to_mont(zk_ctx* ctx, const void* in, void* out, uint64_t n, void* stream);
int zk_fr_from_mont(zk_ctx* ctx, const void* in, void* out, uint64_t n, void* stream);
int zk_fq_to_mont(zk_ctx* ctx, const void* in, void* out, uint64_t n, void* stream);
int zk_fq_from_mont(zk_ctx* ctx, const void* in, void* out, uint64_t n, void* stream);

/* Element-wise Fr (Montgomery) ops over n elements.  zk_fr_mul replaces
 * pointwise_mul_fr_mont (runtime/metal/runtime.py:204-213, dispatched at
 * runtime/kernels/metal/kernels.py:175-222); zk_fr_sub replaces
 * poly_sub_fr_mont (runtime/metal/runtime.py:215-224,
 * runtime/kernels/metal/kernels.py:122-167). */
int zk_fr_mul(zk_ctx* ctx, const void* a, const void* b, void* out, uint64_t n, void* stream);
int zk_fr_add(zk_ctx* ctx, const void* a, const void* b, void* out, uint64_t n, void* stream);
int zk_fr_sub(zk_ctx* ctx, const void* a, const void* b, void* out, uint64_t n, void* stream);
/* out[i] = a[i] * c for one Montgomery constant c (host words). */
int zk_fr_scale(zk_ctx* ctx, const void* a, const uint64_t c_mont[4], void* out, uint64_t n,
                void* stream);

/* NTT of size n = 2^log_n (log_n <= 28) over Montgomery Fr.
 *   inverse = 0: out[i] = sum_{j < min(n, in_len)} in[j] * s^j * w^(ij)
 *   inverse = 1: out[j] = n^-1 * s^-j * sum_i in[i] * w^(-ij)   (in_len must be n)
 * w = omega_mont (NULL: the standard root omega_for_size(n) =
 * ROOT_2_28^(2^(28-log_n))), s = shift_mont (NULL: no coset, s = 1).
 * Replaces ROOTS_EVALS_FROM_COEFFS / ROOTS_COEFFS_FROM_EVALS /
 * COSET_EVALS_FROM_COEFFS / COSET_COEFFS_FROM_EVALS on the Metal backend
 * (runtime/kernels/metal/kernels.py:225-381, 418-827) and the CPU ones
 * (runtime/kernels/cpu/kernels.py:175-250, crypto/poly/ntt.py:42-146).
 * in == out is allowed. */
int zk_ntt(zk_ctx* ctx, const void* in, uint64_t in_len, void* out, uint32_t log_n, int inverse,
           const uint64_t omega_mont[4], const uint64_t shift_mont[4], void* stream);
/* `batch` independent transforms of size 2^log_n laid out back to back. */
int zk_ntt_batch(zk_ctx* ctx, const void* in, void* out, uint32_t log_n, uint32_t batch,
                 int inverse, const uint64_t omega_mont[4], const uint64_t shift_mont[4],
                 void* stream);

/* G1 MSM: sum_i scalars[i] * bases[i] over n points, signed-digit
 * Pippenger with window_bits c (0: the fastest measured window for n on MI355X:
 * 5 below 2^12, 8 below 2^15, 15 below 2^20, 17 from there on; the
 * no-allocation entry point and fixed-base sets keep the reference's rule
 * c = max(4, min(16, bitlen(n) - 2))).
 * Replaces metal_msm_g1 / metal_msm_g1_v2 (runtime/metal/msm.py:65-143,
 * 179-332) behind the MSM_G1 op (runtime/kernels/metal/kernels.py:84-120)
 * and the CPU msm_pippenger (crypto/msm/pippenger.py:10-64).  Synchronises
 * `stream` before returning the result in out_jac_mont. */
int zk_msm_g1(zk_ctx* ctx, const void* bases, const void* scalars, uint64_t n, int window_bits,
              uint64_t out_jac_mont[12], void* stream);

/* Prepared base sets: `n` bases (the zk_msm_g1 layout) converted once into
 * the form the MSM kernels gather and kept resident, so repeated MSMs over
 * the set (or a prefix of it) skip the per-call conversion.  The device
 * form of the reference's per-points-identity caches (cached_points_tuple,
 * runtime/warmup.py:17-26; crypto/msm/pippenger.py:225-248), where the
 * Metal path re-packs the bases on every call (runtime/metal/msm.py:13-34).
 * zk_msm_g1_prepared computes the same sum as zk_msm_g1 over the first n
 * bases of the set (n <= the set's size). */
int zk_msm_bases_create(zk_ctx* ctx, const void* bases, uint64_t n, zk_msm_bases** out, void* stream);
int zk_msm_bases_destroy(zk_ctx* ctx, zk_msm_bases* b);

/* The G2 MSM on the device (replaces msm_pippenger_g2,
 * crypto/msm/pippenger.py:67-121, the reference's CPU MSM_G2 kernel at
 * runtime/kernels/cpu/kernels.py:338-348): zk_msm_g2_bases_create uploads a
 * G2 base set once (host words as zk_g2_msm_host takes them: affine
 * canonical plain (x0, x1, y0, y1), 16 words a point, zeros = infinity) and
 * converts it on the device; zk_msm_g2 sums scalars[i] * base[i] over the
 * first n bases with device Fr scalars in Montgomery form (as zk_msm_g1)
 * and writes the affine canonical result as zk_g2_msm_host does (16 words,
 * zeros for the identity).  window_bits 0 picks the window as for G1. */
int zk_msm_g2_bases_create(zk_ctx* ctx, const uint64_t* bases, uint64_t n, zk_msm_g2_bases** out, void* stream);
int zk_msm_g2_bases_destroy(zk_ctx* ctx, zk_msm_g2_bases* b);
int zk_msm_g2(zk_ctx* ctx, const zk_msm_g2_bases* b, const void* scalars, uint64_t n, int window_bits,
              uint64_t out[16], void* stream);
int zk_msm_g1_prepared(zk_ctx* ctx, const zk_msm_bases* b, const void* scalars, uint64_t n, int window_bits,
                       uint64_t out_jac_mont[12], void* stream);

/* zk_msm_g1 without any device allocation: the same result, on a 64 MiB
 * arena that is a file-scope device array (one per GPU, held until the
 * process ends) instead of the context's growable scratch.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Shapes whose arena exceeds 64 MiB are refused with
 * ZK_ERR_INVALID (n up to 2^15 at the default window, 2^16 at windows up
 * to 14).  Calls on one
 * GPU share the arena, so calls that overlap on one GPU race; the caller
 * serialises them. */
int zk_msm_g1_noalloc(zk_ctx* ctx, const void* bases, const void* scalars, uint64_t n, int window_bits,
                      uint64_t out_jac_mont[12], void* stream);

/* Fixed-base scalar multiplication of one G1 point P by n scalars:
 * out[i] = [scalars[i]] P as 64-byte affine Montgomery (x = y = 0 for the
 * identity).  `table` holds T[j][d] = [d * 2^(8j)] P for j < 32, d < 256
 * (32 * 256 affine Montgomery points, T[j][0] = identity).  Replaces the
 * one-g1_mul-per-element loops of SRS / query-vector generation
 * (crypto/kzg/cpu_ref.py:28-39, protocols/groth16/setup.py:63-96) and is
 * the device form of crypto/msm/pippenger.py:252-272's fixed-base table. */
int zk_g1_fixed_base(zk_ctx* ctx, const void* table, const void* scalars, uint64_t n, void* out,
                     void* stream);

/* Fixed-base sets: a base set used by many MSMs (an SRS, a proving key's
 * query vector) is expanded once into table[w][i] = 2^(c w) B_i (W windows,
 * 64 bytes per entry, resident in HBM).  Every MSM over a prefix of the set
 * then needs no per-window reduction or window combination: all signed
 * digits of all windows land in one bucket set.  This is the device form of
 * the reference's shared fixed-base precomputation (FixedBasePrecomp /
 * fixed_base_precompute / msm_fixed_base(_batch),
 * crypto/msm/pippenger.py:217-342, runtime/warmup.py:84-150).
 * window_bits 0: the default rule for n. */
int zk_fixed_bases_create(zk_ctx* ctx, const void* bases, uint64_t n, int window_bits,
                          zk_fixed_bases** out, void* stream);
int zk_fixed_bases_destroy(zk_ctx* ctx, zk_fixed_bases* fb);
int zk_fixed_bases_info(const zk_fixed_bases* fb, uint64_t* n, int* window_bits);

/* count MSMs over one fixed-base set: MSM j sums scalars[offsets[j] + i] *
 * B_(base_offsets[j] + i) for i < sizes[j] (base_offsets NULL: all zero;
 * base_offsets[j] + sizes[j] <= n; Montgomery scalars, offsets in
 * elements), so one set can hold several query vectors back to back.  Results: out_jac_mont[12 j .. 12 j + 12), Jacobian Montgomery.
 * One launch sequence for the whole batch; counterpart of
 * msm_fixed_base_batch / msm_pippenger_batch (crypto/msm/pippenger.py:124-214,
 * 312-342) behind MSM_G1_BATCH and KZG_BATCH_COMMIT
 * (runtime/kernels/cpu/kernels.py:441-496). */
int zk_msm_g1_fixed_batch(zk_ctx* ctx, const zk_fixed_bases* fb, const void* scalars,
                          const uint64_t* offsets, const uint64_t* sizes, const uint64_t* base_offsets,
                          uint32_t count, uint64_t* out_jac_mont, void* stream);

/* Batch plan: n_msm fixed-base MSMs over `fb` (as zk_msm_g1_fixed_batch)
 * plus n_ntt NTT jobs, every buffer and table allocated at creation.  The
 * NTT jobs run on a forked stream concurrently with the MSM pipeline.  With
 * use_graph the whole launch sequence is captured once into a hipGraph and
 * zk_batch_run replays it (inputs are read from the pointers given at
 * creation, so callers refill those buffers between runs).  zk_batch_run
 * synchronises `stream` and writes the MSM results (12 words each).
 * Counterpart of the reference's sequential re-use of one graph per proof
 * (Executor.run_repeated, runtime/executor.py:59-81; prove_batch,
 * protocols/groth16/prove.py:133-248). */
int zk_batch_create(zk_ctx* ctx, const zk_fixed_bases* fb, const void* msm_scalars,
                    const uint64_t* msm_offsets, const uint64_t* msm_sizes, const uint64_t* msm_base_offsets,
                    uint32_t n_msm, const zk_ntt_job* ntts, uint32_t n_ntt, int use_graph, zk_batch** out,
                    void* stream);
int zk_batch_run(zk_ctx* ctx, zk_batch* b, uint64_t* msm_out_jac_mont, void* stream);
int zk_batch_destroy(zk_ctx* ctx, zk_batch* b);

/* out[i] = scale * base^i for i < n (Montgomery Fr, host words in): the
 * geometric vectors of setup (tau^k for h_query, protocols/groth16/setup.py:85-89;
 * the domain roots, crypto/poly/ntt.py:27-38).  Synchronises `stream`. */
int zk_fr_powers(zk_ctx* ctx, const uint64_t base_mont[4], const uint64_t scale_mont[4], uint64_t n, void* out,
                 void* stream);

/* Synthetic benchmark witness (no reference counterpart): `chains` chains of
 * `rounds` rounds of x^5 S-boxes on 3 lanes plus a 3-term mix y = M x5;
 * layout documented in csrc/synth.hip.  mix: 9 canonical elements, init:
 * 3 chains canonical elements; w: 3 chains + 12 chains rounds + 1 elements
 * out (canonical, last = 1).  Host only. */
int zk_synth_witness(const uint64_t* mix, const uint64_t* init, uint32_t chains, uint32_t rounds, int threads,
                     uint64_t* w);

/* Groth16 helpers on the device (Montgomery Fr):
 * zk_r1cs_eval: out[j] = sum_{k in [row_ptr[j], row_ptr[j+1])} coeffs[k] * w[cols[k]]
 *   for j < rows (CSR; replaces eval_r1cs_vectors, protocols/groth16/r1cs.py:77-92);
 * zk_fr_gather: out[k] = src[idx[k]] (aux slice, protocols/groth16/prove.py:105);
 * zk_fr_inv: out[i] = a[i]^-1, 0 -> 0 (BATCH_INV, runtime/kernels/cpu/kernels.py:253-269);
 * zk_fr_count_nonzero: number of non-zero elements (synchronises `stream`;
 *   the R1CS-satisfaction check of protocols/groth16/qap.py:98-99). */
int zk_r1cs_eval(zk_ctx* ctx, const uint32_t* row_ptr, const uint32_t* cols, const void* coeffs, const void* w,
                 uint64_t rows, void* out, void* stream);
int zk_fr_gather(zk_ctx* ctx, const void* src, const uint32_t* idx, uint64_t n, void* out, void* stream);
int zk_fr_inv(zk_ctx* ctx, const void* a, void* out, uint64_t n, void* stream);
int zk_fr_count_nonzero(zk_ctx* ctx, const void* a, uint64_t n, uint64_t* result, void* stream);

/* PLONK quotient numerator and Z_H on the extended coset of size m (a
 * multiple of n by at most 16), fused into one device pass: replaces the
 * PLONK_T_QUOTIENT_EVALS operator (runtime/kernels/cpu/kernels.py:607-694,
 * used by protocols/plonk/prove.py:363-424).  ext holds 15 device vectors of
 * m Fr Montgomery elements in the order a, b, c, z, pi, ql, qr, qm, qo, qc,
 * s1, s2, s3, z(omega X), L1 (their coset evaluations); all scalars are
 * Montgomery words.  num_out[i] = gate + alpha perm + alpha^2 (z - 1) L1 at
 * x_i = shift omega_m^i, zh_out[i] = shift_n (omega_m^n)^i - 1. */
int zk_plonk_quotient_evals(zk_ctx* ctx, const void* const* ext, uint64_t n, uint64_t m,
                            const uint64_t omega_m[4], const uint64_t shift[4], const uint64_t shift_n[4],
                            const uint64_t alpha[4], const uint64_t beta[4], const uint64_t gamma[4],
                            const uint64_t k1[4], const uint64_t k2[4], void* num_out, void* zh_out,
                            void* stream);

/* G2 on the host (CPU threads).  Points are affine over Fq2 = Fq[u]/(u^2+1)
 * on the twist y^2 = x^3 + 3/(9+u), CANONICAL little-endian words
 * (x0, x1, y0, y1; 16 x u64, all zero = infinity); scalars canonical.
 * zk_g2_msm_host replaces msm_pippenger_g2 (crypto/msm/pippenger.py:67-121,
 * the CPU MSM_G2 op runtime/kernels/cpu/kernels.py:338-348) with signed
 * c-bit windows (window_bits in [2, 20]; 0 = sized from n and threads) and
 * batch-affine bucket sums; threads 0 = all hardware threads.
 * zk_g2_mul_fixed_host computes out[i] = [k_i] Q (setup-time g2_mul loops,
 * protocols/groth16/setup.py:63-83); zk_g2_add_host adds two points. */
int zk_g2_msm_host(const uint64_t* bases, const uint64_t* scalars, uint64_t n, int window_bits, int threads,
                   uint64_t out[16]);
int zk_g2_mul_fixed_host(const uint64_t* q, const uint64_t* scalars, uint64_t n, int threads, uint64_t* out);
int zk_g2_add_host(const uint64_t* p, const uint64_t* q, uint64_t out[16]);
/* A G2 base set prepared once for many host MSMs (the proving key's
 * b_query_g2, reused by every proof): zk_g2_bases_create copies the points
 * (affine canonical, 16 words, zeros = infinity) and converts them for the
 * IFMA path where the CPU has it; zk_g2_msm_prepared is zk_g2_msm_host over
 * its first n <= size points.  The host counterpart of zk_msm_bases_create
 * for the G2 MSM of the reference's prover (protocols/groth16/prove.py:
 * 76-90 runs an MSM_G2 node over b_query_g2 in every proof). */
int zk_g2_bases_create(const uint64_t* bases, uint64_t n, int threads, zk_g2_bases** out);
int zk_g2_bases_destroy(zk_g2_bases* b);
int zk_g2_msm_prepared(const zk_g2_bases* b, const uint64_t* scalars, uint64_t n, int window_bits, int threads,
                       uint64_t out[16]);
/* Selects the host G2 MSM's arithmetic: mode 1 (the default) runs its
 * batched affine bucket additions eight at a time on AVX-512 IFMA where the
 * CPU has AVX-512F and IFMA, 0 runs them one at a time on 64-bit scalar
 * words, any other mode only asks.  Returns 1 if the IFMA path is in use
 * after the call.  Both give the same points.  No counterpart in the
 * reference (its G2 MSM is Python, crypto/msm/pippenger.py:67-121). */
int zk_host_simd(int mode);

/* out = [k] P on the host for one G1 point (affine canonical, 8 words, zeros
 * = infinity; k canonical).  Replaces the g1_mul calls of proof assembly
 * (protocols/groth16/prove.py:119-126, crypto/ecc/bn254.py:15-16). */
int zk_g1_mul_host(const uint64_t p[8], const uint64_t k[4], uint64_t out[8]);

/* BN254 optimal-ate pairing on the host.  G1 points affine canonical (8
 * words, zeros = infinity), G2 points affine canonical (16 words).
 * zk_pairing_product writes prod_i e(P_i, Q_i) as 48 canonical words (Fq12
 * = Fq6[w]/(w^2 - v), Fq6 = Fq2[v]/(v^3 - (9 + i))); zk_pairing_check sets
 * *ok = 1 iff the product is 1.  Replaces pairing_g1_g2
 * (crypto/pairing/bn254.py:10-11) in the Groth16 verifier
 * (protocols/groth16/verify.py:11-23). */
int zk_pairing_product(const uint64_t* g1, const uint64_t* g2, uint32_t n, uint64_t* out48);
int zk_pairing_check(const uint64_t* g1, const uint64_t* g2, uint32_t n, int* ok);

/* Tuning knobs: "ntt_max_pass_bits" (stages per NTT pass, 1..12, passes of
 * 11 or 12 stages on 2^12-element tiles; 0 = auto), "ntt_tile_bits" (0, 10
 * or 12), "ntt_xcd_group" (tile numbering per XCD; 0 = auto),
 * "ntt_inplace" (1, the default: in-place pass stages on a bit-reversed
 * tile load; 0: Stockham order), "ntt_stage_twiddles" (1, the default:
 * per-stage twiddle tables; 0: one table of w^e), "ntt_full_scale" (1,
 * the default: coset factors from a full table per size and shift up to
 * 2^26; 0: two-level power tables),
 * "ntt_table_cap_mb" (bytes of cached NTT tables a context keeps before
 * freeing the least recently used unpinned ones; 0 = no cap, default 16384),
 * "msm_segment" (pairs per accumulation thread; 0, the default: chosen
 * per launch from its size, between 24 and 80), "msm_reduce_chunk"
 * (buckets per reduction thread of the chunked reduction),
 * "msm_reduce_split" (1, the default: single-launch MSMs reduce their
 * buckets by digit split; 0: chunked running sums), "msm_sort_fused" (1,
 * the default: bucket-sort segments that one workgroup holds are placed
 * whole in one pass; 0: every segment through the chunked count / scan /
 * place kernels).  No counterpart in the reference. */
int zk_set_tuning(zk_ctx* ctx, const char* key, int64_t value);

/* Per-phase device time from HIP events recorded on the launch stream
 * ("ntt", "msm_digits", "msm_sort", "msm_accumulate", "msm_reduce").
 * The counterpart of the reference's per-op trace
 * (runtime/executor.py:196-218), at kernel-phase granularity. */
int zk_profile_enable(zk_ctx* ctx, int on);
int zk_profile_get(zk_ctx* ctx, const char* phase, double* total_ms, uint64_t* launches);
int zk_profile_reset(zk_ctx* ctx);

/* Measured issue rate of v_mad_u64_u32 on this device (lane multiply-adds
 * per second): the VALU roofline the MSM / NTT kernels are reported
 * against.  No counterpart in the reference. */
int zk_probe_mad_rate(zk_ctx* ctx, double* mads_per_s);

#ifdef __cplusplus
}
#endif

#endif /* ZKHIP_H */
