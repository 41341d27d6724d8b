/* zorch_amd — C ABI of the MI355X dense sumcheck prover (libzorch_amd.so).
 *
 * Plain pointers and sizes only. Every field element crosses the boundary in
 * its storage encoding, little-endian: BabyBear = one u32 Montgomery word
 * (a * 2^32 mod p), BabyBear^4 = four such words c0..c3 (X^4 - 11),
 * Goldilocks = one canonical u64. Host buffers are borrowed for the duration
 * of a call; device pointers (tables_on_device = 1) must stay valid until the
 * call returns. A zs_ctx owns one HIP stream and is single-threaded; distinct
 * contexts are independent.
 *
 * Return codes: ZS_OK, or an error whose message zs_last_error() returns
 * (thread-local). The Python layer maps ZS_EINVAL to ValueError,
 * ZS_ENOMEM / ZS_EBUDGET to MemoryError and ZS_EHIP / ZS_ENODEV to
 * RuntimeError — the reference's conventions (structural errors raise
 * ValueError; soundness failures return ok = 0 and never raise).
 */
#ifndef ZORCH_AMD_H_
#define ZORCH_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ZS_OK 0
#define ZS_EINVAL 1
#define ZS_ENOMEM 2
#define ZS_EHIP 3
#define ZS_EBUDGET 4
#define ZS_ENODEV 5

/* Field ids. */
#define ZS_F_BABYBEAR 1
#define ZS_F_BABYBEAR4 2
#define ZS_F_GOLDILOCKS 3

/* Transcript kinds. */
#define ZS_TR_SHA256 1 /* Sha256FieldTranscript: counter squeeze SHA256(S || le64(ctr)) */
#define ZS_TR_BLAKE3 2 /* Blake3FieldTranscript: squeeze = BLAKE3 extendable output of S */

/* Largest number of rounds (variables) a stage may have: every table of a
 * round holds at most 2^31 entries. */
#define ZS_MAX_ROUNDS 31

/* Prover modes. */
#define ZS_MODE_PRODUCT 1 /* StandardRound(ProductSummand(degree)), natural domain */
#define ZS_MODE_EQ 2      /* EqPolyProver(ProductSummand(degree)), natural domain */
#define ZS_MODE_LOGUP 3   /* LogupSumcheckRound (zorch/logup_gkr/prover.py:232-300): five
                             stacked tables [eq, n0, d1, n1, d0] in one field, the summand
                             eq (lam (n0 d1 + n1 d0) + d0 d1), degree 3, natural domain */

typedef struct zs_ctx zs_ctx;
typedef struct zs_state zs_state;

/* Streaming SHA-256 state: midstate, pending tail, byte counts (112 bytes). */
typedef struct zs_sha256_state {
  uint32_t h[8];
  uint8_t pending[64];
  uint64_t pending_len;
  uint64_t total_len;
} zs_sha256_state;

/* Streaming BLAKE3 state (hash mode, 632 bytes): the current chunk's
 * chaining value, the stack of completed subtree chaining values, the current
 * block (not yet compressed: the last block of a chunk carries CHUNK_END),
 * the chunk index and counts. Replaces hash-frx's Blake3Stream
 * (zorch/blake3_field_transcript.py:206-215). */
typedef struct zs_blake3_state {
  uint32_t cv[8];
  uint32_t stack[16][8];
  uint8_t block[64];
  uint64_t chunk_counter;
  uint32_t block_len;
  uint32_t blocks_done;
  uint32_t stack_len;
  uint32_t reserved;
} zs_blake3_state;

/* A field transcript value (648 bytes). `kind` selects the hash (ZS_TR_*),
 * `field` is the transcript word field and `item_bytes` its storage width.
 * Plain data: copying it forks the transcript, as the reference's immutable
 * transcripts do. */
typedef struct zs_transcript {
  uint32_t kind;
  uint32_t field;
  uint32_t item_bytes;
  uint32_t reserved;
  union {
    zs_sha256_state sha;
    zs_blake3_state blake3;
  };
} zs_transcript;

/* Round-message wire forms. The round polynomial s has degree
 * D = degree (product) or degree + 1 (eq); npts = D + 1. */
#define ZS_WIRE_VALUES 0     /* s(0..D), the natural domain (StandardRound's and
                                EqPolyProver's default; reduce_evals) */
#define ZS_WIRE_DOMAIN 1     /* s sampled at an explicit evaluation domain: the
                                finite nodes in order, s(inf) (the leading
                                coefficient) inserted at inf_pos when has_inf
                                (EvalDomain.sample, zorch/sumcheck/domain.py:95-106;
                                the compressed U-hat domain, :140-146;
                                reduce_domain, zorch/sumcheck/reduce.py:43-52) */
#define ZS_WIRE_COEFFS 2     /* ascending coefficients of s (product_round_coeffs,
                                zorch/sumcheck/domain.py:216-223; reduce_coeffs,
                                zorch/sumcheck/reduce.py:35-40) */
#define ZS_WIRE_COMPRESSED 3 /* [c0, c2] of a two-factor product round
                                (CompressedProductRound, zorch/sumcheck/prover.py:177-212;
                                reduce_compressed, zorch/sumcheck/reduce.py:55-66) */

typedef struct zs_wire {
  uint32_t kind;      /* ZS_WIRE_* */
  uint32_t num_nodes; /* ZS_WIRE_DOMAIN: finite sample nodes (distinct) */
  uint32_t has_inf;   /* ZS_WIRE_DOMAIN: 1 if s(inf) is among the samples */
  uint32_t inf_pos;   /* ZS_WIRE_DOMAIN: its position, 0..num_nodes */
  const void* nodes;  /* ZS_WIRE_DOMAIN: num_nodes challenge-field elements (host) */
} zs_wire;

typedef struct zs_prove_args {
  uint32_t table_field;      /* storage field of `tables` */
  uint32_t chal_field;       /* challenge / message / claim field */
  uint32_t mode;             /* ZS_MODE_* */
  uint32_t degree;           /* ProductSummand degree; must equal num_tables */
  uint32_t num_tables;       /* k stacked factor tables */
  uint32_t rounds;           /* n: every table has 2^n entries */
  const void* tables;        /* k x 2^n elements, table-major */
  uint32_t tables_on_device; /* 1: `tables` is a device pointer */
  uint32_t refold_rounds;    /* ZS_MODE_PRODUCT, zs_prove and zs_state_* only: the first
                                refold_rounds rounds (at most rounds / 2) never fold the
                                input; each refolds it over the bound prefix with the
                                running eq(r[<i], .) table (the sqrt-space engine,
                                zorch/sumcheck/sqrt_space.py:40-132, which binds
                                rounds / 2 that way). The same messages as folding;
                                0: fold every round */
  const void* eq_point;      /* ZS_MODE_EQ: n challenge-field elements (host) */
  const void* claim;         /* one challenge-field element (host) */
  const zs_wire* wire;       /* message wire form; null: ZS_WIRE_VALUES (copied) */
  const void* eq_mass;       /* ZS_MODE_EQ: the eq mass already bound, one challenge-field
                                element (EqPolyState's second part); null: one */
  const void* lambda;        /* ZS_MODE_LOGUP: the layer's batching challenge lam, one
                                challenge-field element (host); null otherwise */
} zs_prove_args;

/* Caller-allocated host outputs. mlen = the message width of the wire form
 * (zs_prove_check): npts for values and coefficients, the sample count of a
 * domain, 2 for the compressed form. */
typedef struct zs_proof {
  void* msgs;         /* rounds x mlen challenge-field elements */
  void* point;        /* rounds x challenge-field raw squeezed words */
  void* value;        /* one challenge-field element */
  void* final_tables; /* optional: k folded values (+ eq mass for ZS_MODE_EQ) */
} zs_proof;

const char* zs_last_error(void);
const char* zs_version(void);
int zs_device_count(int* count);

/* ---- context ---------------------------------------------------------- */
/* Replaces the reference's implicit FRX device (frx.device_put / PJRT client,
 * zorch/transcript.py:949-962): owns the stream, work buffers and budget. */
int zs_ctx_create(int device, uint64_t hbm_budget_bytes, zs_ctx** out);
void zs_ctx_destroy(zs_ctx* ctx);
/* Bytes of device memory the context currently holds. */
int zs_ctx_reserved_bytes(zs_ctx* ctx, uint64_t* out);

/* Device buffers from the context's budgeted pool, for inputs that should be
 * resident in HBM before proving (tables_on_device = 1). The reference keeps
 * arrays on its FRX device the same way (frx.device_put,
 * zorch/transcript.py:949-962). Copies are synchronous on the context's
 * stream. */
int zs_device_alloc(zs_ctx* ctx, size_t bytes, void** dptr);
int zs_device_free(zs_ctx* ctx, void* dptr);
int zs_memcpy_htod(zs_ctx* ctx, void* dst, const void* src, size_t bytes);
int zs_memcpy_dtoh(zs_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Device-to-device copy on the context's stream (synchronous); also the plain
 * copy of known size that HBM byte counters are calibrated against. */
int zs_memcpy_dtod(zs_ctx* ctx, void* dst, const void* src, size_t bytes);

/* ---- HBM bandwidth (measurement aid) ---------------------------------------
 * No counterpart in the reference. The attainable HBM bandwidth of this
 * device, the peak the streaming rounds are judged against: a streaming-read
 * kernel (16-byte loads, every byte read once) over `bytes` of device memory
 * from the context's pool, and a streaming copy of the same size (read and
 * write, counted twice), each the best of `reps` runs timed with HIP events
 * on the context's stream; GB/s = 1e9 bytes per second. */
int zs_hbm_bandwidth(zs_ctx* ctx, size_t bytes, uint32_t reps, double* read_gbs,
                     double* copy_gbs);

/* ---- kernel timing -------------------------------------------------------
 * Profiling aid with no counterpart in the reference (whose only tracing is
 * XLA's dump flags): while enabled, every pass, finisher and tail kernel the
 * context launches is bracketed by HIP events on the context's stream. A read
 * returns one record per kernel with its algorithmic HBM bytes (fused-pass
 * model) and its measured duration, then clears the list. kind: 1 = pass,
 * 2 = finisher, 3 = tail (all remaining rounds in one block), 4 = batch step,
 * 5 = chain, 6-8 = the look-ahead pass, chain and tail, 9 = rounds 0-1 from
 * the input, 10 = round 2 from the input, 11 = refold round, 12 = engine
 * pass, 13 = LogUp round, 14 = jagged row round, 15 = jagged batch rounds. */
typedef struct zs_kernel_record {
  uint32_t round;
  uint32_t kind;
  uint32_t grid;
  uint32_t reserved;
  uint64_t bytes;
  double ms;
} zs_kernel_record;
int zs_ctx_profile(zs_ctx* ctx, int enable);
int zs_ctx_profile_read(zs_ctx* ctx, zs_kernel_record* out, uint32_t cap, uint32_t* count);
/* In-kernel phase trace (profiling aid): while enabled, kernels record
 * (round << 8 | phase, device wall-clock ticks) pairs; read copies up to `cap`
 * pairs, reports the wall-clock rate in kHz and clears the trace. Phases:
 * 1 pass start, 2 partial sums done, 3 block reduction done, 4 message ready,
 * 5 Fiat-Shamir hop done, 6 round bound; a chained launch adds 7 kernel entry
 * and 8 proof state loaded (with its first round's index). */
int zs_ctx_trace(zs_ctx* ctx, int enable);
int zs_ctx_trace_read(zs_ctx* ctx, uint64_t* pairs, size_t cap, size_t* count,
                      uint64_t* clock_khz);

/* ---- launch shapes -------------------------------------------------------
 * Tuning knobs with no counterpart in the reference (XLA picks its own
 * launch shapes): rounds with at most tail_pairs pairs run in one single-block
 * kernel; a multi-block pass uses a power-of-two grid of at most max_grid
 * blocks (<= 4096) sized for pairs_per_thread pairs per thread; in zs_prove
 * and the per-round calls, rounds from the third on with at most chain_pairs
 * pairs (and more than tail_pairs) run in one persistent launch of at most
 * chain_grid blocks (power of two; chain_pairs 0: one launch per round).
 * In batched and served proving a proof's tail runs at most step_rounds rounds
 * per step (0: all of its remaining rounds in one step). lookahead (0/1): in
 * zs_prove, rounds 0-2 come from the input in two launches (round 1's tables
 * are never written) when the stage starts with three pass rounds, and the
 * tail, the chained launch and the pass round before it fold and evaluate
 * each round while the previous round's message is hashed.
 * Results do not depend on them. Set applies to proofs started afterwards. */
typedef struct zs_tune {
  uint32_t tail_pairs;
  uint32_t max_grid;
  uint32_t pairs_per_thread;
  uint32_t chain_pairs;
  uint32_t chain_grid;
  uint32_t step_rounds;
  uint32_t lookahead;
  uint32_t reserved[1];
} zs_tune;
int zs_ctx_get_tune(zs_ctx* ctx, zs_tune* out);
int zs_ctx_set_tune(zs_ctx* ctx, const zs_tune* tune);

/* ---- transcript (host) ------------------------------------------------
 * Replaces hash-frx's streams under Sha256FieldTranscript
 * (zorch/sha256_field_transcript.py:45-50,100-114,178-181) and
 * Blake3FieldTranscript (zorch/blake3_field_transcript.py:206-215,293-330).
 * Framing is composed by the caller exactly as the reference's Python does. */
int zs_transcript_new(uint32_t kind, uint32_t field, const uint8_t* domain, size_t len,
                      zs_transcript* out);
/* Stream absorb (zorch/sha256_field_transcript.py:180,196;
 * zorch/blake3_field_transcript.py:300-301). */
int zs_transcript_absorb(zs_transcript* t, const uint8_t* data, size_t len);
/* A squeeze minus its framing absorb, re-absorbing what it read: SHA-256,
 * SHA256(S||le64(ctr)) blocks truncated to nbytes
 * (zorch/sha256_field_transcript.py:100-114); BLAKE3, the first nbytes of the
 * extendable output of S (zorch/blake3_field_transcript.py:206-215). */
int zs_transcript_squeeze(zs_transcript* t, size_t nbytes, uint8_t* out);
/* The 32-byte hash of the stream so far (the proof-of-work state digest,
 * zorch/sha256_field_transcript.py:277-284, blake3_field_transcript.py:409). */
int zs_transcript_digest(const zs_transcript* t, uint8_t out[32]);

/* ---- argument check (host) ----------------------------------------------
 * The structural checks every prove entry point makes, with no device work:
 * ZS_EINVAL (message in zs_last_error) for a malformed request, else ZS_OK
 * and the message width on the wire (elements per round). `transcript` may
 * be null. */
int zs_prove_check(const zs_prove_args* args, const zs_transcript* transcript,
                   uint32_t* msg_len);

/* ---- prover ------------------------------------------------------------
 * Whole stage on the device: SumcheckProver.prove / EqPolyProver.prove
 * (zorch/sumcheck/stage.py:52-69, zorch/sumcheck/eq/stage.py:62-92) driving
 * fold_rounds (zorch/prove.py:26-37). One fused fold+evaluate pass per round,
 * the Fiat-Shamir hop on the device; `transcript` is read and advanced. */
int zs_prove(zs_ctx* ctx, const zs_prove_args* args, zs_transcript* transcript,
             zs_proof* out);
/* No counterpart in the reference. For callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): zs_prove with the proof's device working set (control block,
 * messages, slot rows, eq families, folded tables and, for host tables, the
 * staged input) in a library-owned scratch array of ZS_NOALLOC_SCRATCH_BYTES
 * per device, held until the process ends; the call allocates no device or
 * pinned host memory. ZS_EINVAL when the stage does not fit (for instance a
 * two-table BabyBear^4 zerocheck with device-resident tables fits up to
 * n = 21). Calls that overlap on one device race on its array: serialise
 * them. Results are those of zs_prove. */
#define ZS_NOALLOC_SCRATCH_BYTES (64u << 20)
int zs_prove_noalloc(zs_ctx* ctx, const zs_prove_args* args, zs_transcript* transcript,
                     zs_proof* out);

/* ---- round-cost engines ---------------------------------------------------
 * The sqrt-space engine is zs_prove with refold_rounds = rounds / 2
 * (prove_sqrt_space, zorch/sumcheck/sqrt_space.py:99-132).
 *
 * Univariate skip: prove_univariate_skip (zorch/sumcheck/univariate_skip.py:
 * 136-190) for skip_rounds in [1, rounds] (skip_rounds 0 is zs_prove). The
 * first skip_rounds variables index the order-2^skip_rounds two-adic subgroup;
 * round 0 sends msg0, the ascending coefficients of
 * s0(Z) = sum_x prod_k f_k(Z, x) (degree * (2^skip_rounds - 1) + 1 table-field
 * elements, written to msg0; round0_message, :65-101), observed in the table
 * field, and samples r0; the prism is bound at r0 and the remaining
 * rounds - skip_rounds rounds are product rounds over challenge-field tables
 * (messages on args->wire, natural values when null). out->msgs holds those
 * rounds' messages, out->point r0 then their challenges (1 + rounds -
 * skip_rounds raw words), out->value and out->final_tables as zs_prove.
 * ZS_MODE_PRODUCT; tables BabyBear or Goldilocks, the transcript over the
 * tables' field. The subgroup's generator is the field's multiplicative
 * generator (BabyBear 31, Goldilocks 7) to the power (p - 1) / 2^k. */
int zs_prove_univariate_skip(zs_ctx* ctx, const zs_prove_args* args, uint32_t skip_rounds,
                             zs_transcript* transcript, void* msg0, zs_proof* out);
/* Round 0 of the skip alone (skip_round0, zorch/sumcheck/univariate_skip.py:
 * 104-133): msg0, the raw challenge r0 (chal), s0(r0) (reduced) and the prism
 * bound at r0 in `bound_tables`, a device buffer of num_tables x
 * 2^(rounds - skip_rounds) challenge-field elements (table-major), from which
 * any product tail can continue (zs_prove / zs_state_* with tables_on_device). */
int zs_univariate_skip_round0(zs_ctx* ctx, const zs_prove_args* args, uint32_t skip_rounds,
                              zs_transcript* transcript, void* msg0, void* chal, void* reduced,
                              void* bound_tables);

/* Small-value eq-poly sumcheck: prove_eq_poly_small_value
 * (zorch/sumcheck/eq/small_value.py:148-192) with l0 in [1, rounds - rounds/2]:
 * one pass over the tables precomputes the accumulators of rounds 1..l0
 * (zorch/sumcheck/eq/accumulators.py:41-91), whose rounds are then
 * contractions; the factors are refolded once over the bound prefix and the
 * remaining rounds (the transition round first) are eq-poly rounds from the
 * bound eq mass. Every message is on the compressed U-hat_degree domain (the
 * reference's only wire for this engine: args->wire must be null); the
 * messages equal those of EqPolyRound's (a v or y index of the accumulators is
 * read most-significant digit first: the reference writes it
 * least-significant first, which agrees up to l0 = 2). ZS_MODE_EQ,
 * degree 2..4, one transcript word a challenge (the transcript over the
 * challenge field, as SmallValueRound's observe_and_sample(msg, 1)). Outputs
 * as zs_prove. */
int zs_prove_small_value(zs_ctx* ctx, const zs_prove_args* args, uint32_t l0,
                         zs_transcript* transcript, zs_proof* out);

/* ---- dense LogUp-GKR circuit on the device --------------------------------
 * Replace layer_transition (zorch/logup_gkr/circuit.py:100-140) and
 * expand_eq_to_hypercube (zorch/poly/eq.py) for device-resident layers. A
 * layer is a device buffer of five rows of `width` elements each, laid out
 * as a ZS_MODE_LOGUP stage's tables [eq, n0, d1, n1, d0], so that it is
 * proved in place (tables_on_device = 1) once row 0 holds its eq table.
 * zs_logup_transition writes rows 1..4 of `next` (five rows of width / 2)
 * from rows 1..4 of `layer`; zs_logup_eq writes eq(point, x), x in {0,1}^n,
 * point[0] on the index MSB, to the 2^n elements at `out` (device); `point`
 * is n host elements. Both synchronous on the context's stream. */
int zs_logup_transition(zs_ctx* ctx, uint32_t field, const void* layer, size_t width,
                        void* next);
int zs_logup_eq(zs_ctx* ctx, uint32_t field, const void* point, uint32_t n, void* out);

/* ---- jagged LogUp-GKR layer --------------------------------------------
 * Replaces prove_jagged_layer / the sumcheck of JaggedGkrLayerRound
 * (zorch/logup_gkr/jagged_prover.py:124-448; round bodies
 * zorch/logup_gkr/_jagged_rounds.py, zorch/sumcheck/jagged/rounds.py:36-85,
 * fs.py:37-58). A jagged layer materialises row_counts[s] rows of segment s
 * (batch-major, planes n0 | n1 | d0 | d1 of width = sum(row_counts) each)
 * out of 2^row_variables virtual rows; the others are the neutral fraction
 * (n = 0, d = 1). The sumcheck of sum_x eq(eval_point, x) (lam (n0 d1 +
 * n1 d0) + d0 d1) over the virtual hypercube binds the row variables LSB
 * first, then the batch variables; a round sends the four ascending
 * coefficients of its cubic (interpolated through 0, 1, 1/2 and the root of
 * the round's eq factor, the virtual rows' mass added in closed form).
 * Planes, challenges and messages share `field`; the transcript is over
 * `field` or its base field. */
/* One transition of the jagged circuit on the device (replaces
 * jagged_layer_transition, zorch/logup_gkr/circuit.py:296-470): `planes` are
 * the four planes n0, n1, d0, d1 of a layer (device, row stride `width`,
 * segment s holding row_counts[s] rows back to back); `next` receives the
 * four planes of the next layer (row stride `out_width`) with
 * out_row_counts[s] >= ceil(row_counts[s] / 2) rows per segment: folded
 * pairs, then neutral rows, then zeros past the live rows. Counts are host
 * arrays of num_segments entries. Synchronous on the context's stream. */
int zs_jagged_transition(zs_ctx* ctx, uint32_t field, uint32_t num_segments,
                         const uint32_t* row_counts, const uint32_t* out_row_counts,
                         const void* planes, size_t width, void* next, size_t out_width);

typedef struct zs_jagged_args {
  uint32_t field;              /* ZS_F_* of planes, challenges and messages */
  uint32_t num_segments;       /* 2^niv segments (a power of two) */
  uint32_t row_variables;      /* nrv >= 1; niv + nrv <= ZS_MAX_ROUNDS */
  uint32_t planes_on_device;   /* 1: `planes` is a device pointer */
  const uint32_t* row_counts;  /* num_segments counts, 1 <= count <= 2^nrv (host) */
  const void* planes;          /* 4 x sum(row_counts) elements: n0, n1, d0, d1 */
  const void* eval_point;      /* niv + nrv elements, most significant first (host) */
  const void* lambda;          /* the batching challenge, one element (host) */
  const void* claim;           /* lam num_eval + den_eval, one element (host) */
} zs_jagged_args;

typedef struct zs_jagged_proof {
  void* round_polys;  /* (niv + nrv) x 4 ascending coefficients */
  void* challenges;   /* niv + nrv raw squeezed words, in sampling order (LSB first) */
  void* openings;     /* 4 elements: n0, n1, d0, d1 at the bound point */
  void* value;        /* the final reduced claim, one element */
} zs_jagged_proof;

int zs_prove_jagged_layer(zs_ctx* ctx, const zs_jagged_args* args, zs_transcript* transcript,
                          zs_jagged_proof* out);

/* Many independent instances (uneven sizes, mixed kinds) in one submission;
 * instance i uses args[i], transcripts[i], outs[i], with the same results as
 * zs_prove on each. Instances of one kind (field pair, factor count, mode)
 * advance together: step s is one launch that runs round s of every instance
 * still running, so the per-round latency is paid once per step for the whole
 * batch. One upload of all control blocks, one download of all outputs.
 * The reference proves instances one at a time (no batched entry point). */
int zs_prove_batch(zs_ctx* ctx, size_t count, const zs_prove_args* args,
                   zs_transcript* transcripts, zs_proof* outs);

/* ---- per-round state (drop-in under fold_rounds) -----------------------
 * initial_claim (zorch/sumcheck/prover.py:86-91) on device-resident tables,
 * then one ProverRound step per call: StandardRound.__call__
 * (zorch/sumcheck/prover.py:157-174) or EqPolyRound.__call__
 * (zorch/sumcheck/eq/eq_poly.py:214-221). */
int zs_state_create(zs_ctx* ctx, const zs_prove_args* args, zs_state** out);
/* Advance one round: writes the message (npts elements), the raw challenge
 * and the reduced claim; advances `transcript`. */
int zs_state_round(zs_state* st, zs_transcript* transcript, void* msg_out, void* chal_out,
                   void* claim_out);
/* Rounds already bound, and a copy of the current folded tables
 * (k x 2^(n-bound) challenge-field elements; before round 1 the tables are
 * returned lifted into the challenge field). */
int zs_state_bound(const zs_state* st, uint32_t* bound);
int zs_state_read_tables(zs_state* st, void* out, size_t out_bytes);
void zs_state_destroy(zs_state* st);

/* ---- verifier (host) ---------------------------------------------------
 * One SumcheckRound step (zorch/sumcheck/verifier.py:62-71): observe(msg),
 * sample the challenge, reduce_evals (zorch/sumcheck/reduce.py:27-32).
 * Writes the raw challenge words, the reduced claim and ok (1/0). Width
 * mismatch -> ZS_EINVAL. */
int zs_verify_round(uint32_t chal_field, zs_transcript* transcript, const void* claim,
                    const void* msg, uint32_t npts, uint32_t degree, void* chal_out,
                    void* reduced_out, int* ok);

/* ---- device field arithmetic (self-test) ---------------------------------
 * No counterpart in the reference, whose field arithmetic lives in its
 * third-party dtypes (zk_dtypes): runs `count` independent operations of the
 * device field library on the GPU (one kernel on the context's stream), the
 * same functions the round kernels call, for known-answer tests of edge
 * operands. a, b, out: host arrays of `count` storage elements (operands are
 * reduced first, as the kernels reduce squeezed words). */
#define ZS_OP_ADD 1
#define ZS_OP_SUB 2
#define ZS_OP_MUL 3
int zs_field_op(zs_ctx* ctx, uint32_t field, uint32_t op, const void* a, const void* b,
                void* out, size_t count);

/* ---- proof server (continuous batching) ----------------------------------
 * No counterpart in the reference, which proves one stage per call; this is
 * the serving loop for many independent requests on one GPU. A server takes
 * over a context: its scheduler thread owns the context's stream and launch
 * shapes, so do not prove on that context elsewhere until zs_server_destroy
 * (its device and pinned pools are thread-safe, so device buffers from it
 * may still be allocated and freed; inputs may also live on another context
 * of the same device).
 * Its scheduler thread admits submitted requests in arrival order against an
 * HBM budget (resident state per request: control block, outputs, eq
 * families, folded tables, and the input tables when they are on the host),
 * then repeatedly runs one step: the next round of every admitted request,
 * one batched launch per request kind, so that requests which arrive while
 * others are mid-proof join at the next step. A request whose remaining
 * rounds are small finishes them within one step (at most step_rounds of
 * them per step when that launch shape is set). Requests of 20 rounds or
 * more are instead proved whole, as by zs_prove, on one of three streams of
 * the server's own, concurrently with the steps and with one another. Results are those of
 * zs_prove for the same arguments and transcript. Submit, poll, wait and
 * completed may be called from any thread. */
typedef struct zs_server zs_server;
typedef struct zs_server_stats {
  uint64_t submitted, admitted, completed, failed, steps;
  uint64_t queued, active, resident_bytes, peak_resident_bytes, peak_active;
  uint64_t budget_bytes;
  double busy_s; /* wall time with at least one step queued on the device */
} zs_server_stats;
/* hbm_budget_bytes = 0: the context's budget. */
int zs_server_create(zs_ctx* ctx, uint64_t hbm_budget_bytes, zs_server** out);
/* args->tables must stay valid until the request completes; the claim, the
 * eq point and the wire form (with its nodes) are copied; the transcript is read. A request larger than the
 * whole budget completes with ZS_EBUDGET. */
int zs_server_submit(zs_server* srv, const zs_prove_args* args, const zs_transcript* transcript,
                     uint64_t* id);
/* done = 0: still running. done = 1: `out` filled as by zs_prove, the
 * transcript state advanced, latency_s = submit to completion on the
 * server's clock; the request is forgotten (its error, if any, is returned). */
int zs_server_poll(zs_server* srv, uint64_t id, zs_proof* out, zs_transcript* transcript,
                   double* latency_s, int* done);
/* As poll, blocking up to timeout_s seconds. */
int zs_server_wait(zs_server* srv, uint64_t id, zs_proof* out, zs_transcript* transcript,
                   double* latency_s, double timeout_s, int* done);
/* Ids of requests finished since the last call (at most cap). */
int zs_server_completed(zs_server* srv, uint64_t* ids, size_t cap, size_t* count);
int zs_server_stats_read(zs_server* srv, zs_server_stats* out);
/* Stops the scheduler after the step in flight; unfinished requests are
 * dropped. The context is free for other use afterwards. */
void zs_server_destroy(zs_server* srv);

#ifdef __cplusplus
}
#endif
#endif /* ZORCH_AMD_H_ */
