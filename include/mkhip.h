/*
 * mkhip.h -- C ABI of the MI355X (gfx950) multi-key mul+relin library.
 *
 * Drop-in boundary for the reference's evaluator calls (ZhangChuanxin1/MKF):
 *   KKLSS MK-CKKS  mkckks.Evaluator.MulRelinNew   software/kklss/mkckks/evaluator.go:419
 *   KKLSS MK-BFV   mkbfv.Evaluator.MulRelinNew    software/kklss/mkbfv/evaluator.go:85
 *   MKF-CKKS       (*CKKSCase).Run                software/mkf/mkf_cpu/ckks.go:102
 *   MKF-BFV        (*BFVCase).Run                 software/mkf/mkf_cpu/bfv.go:212
 * A Go host binds these through cgo; a C++ host through
 * mkf_amd/host/mkhip.hpp.  Plain pointers and sizes only.
 *
 * Data layout (the reference's fixture/dump layout: little-endian uint64,
 * modulus-major then coefficient-major, mkckks/fpga_fixture_test.go:125-148):
 *   poly        [limbs][N]
 *   ciphertext  [1 + k][level+1][N]: component "0" first, then one component per
 *               party in the order of the accompanying id array
 *   switching key (KKLSS) [beta][L+K][N]: beta PolyQP, Q limbs then P limbs,
 *               NTT + Montgomery form (x * 2^64 mod q)
 *   switching key (MKF)   [beta][2][L+K][N]: output_poly[0..1] per digit
 * All residues are canonical ([0, q)).  KKLSS ciphertexts are in the coefficient
 * domain, MKF-CKKS ciphertexts in the NTT domain (as in the reference).
 *
 * Ownership: the caller owns every host buffer; the library owns device memory
 * (keys stay resident in HBM until unregistered or the context is destroyed).
 * Threading: a context is not re-entrant (the reference's evaluators share
 * pools, mkrlwe/keyswitch.go:9-18); use one context per host thread.
 */
#ifndef MKHIP_H
#define MKHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (0 = OK).  The Go binding turns them into panics for the
 * KKLSS evaluators and into `error` values for MKF, as the reference does. ---- */
#define MKHIP_OK 0
#define MKHIP_ERR_ARG (-1)           /* null pointer / bad size / bad count          */
#define MKHIP_ERR_PARAMS (-2)        /* moduli not NTT friendly, unsupported N        */
#define MKHIP_ERR_GADGET (-3)        /* alpha = K/gamma is 0 (e.g. K=1, gamma=2)      */
#define MKHIP_ERR_UNKNOWN_PARTY (-4) /* keys.go:190-198 "there is no ... key"         */
#define MKHIP_ERR_LEVEL (-5)         /* keyswitch_hoisted.go:53-59 level mismatch     */
#define MKHIP_ERR_NO_KEY (-6)        /* CRS u / MKF keys not uploaded                 */
#define MKHIP_ERR_OOM (-7)           /* device memory exhausted (admission refused)   */
#define MKHIP_ERR_DEVICE (-8)        /* HIP runtime error                             */
#define MKHIP_ERR_SCALE (-9)         /* Rescale: minScale <= 0, scale 0, level 0      */
#define MKHIP_ERR_SCHEME (-10)       /* call not valid for this context's scheme      */
#define MKHIP_ERR_TOO_MANY (-11)     /* more parties than MKHIP_MAX_PARTIES           */
#define MKHIP_ERR_SCRATCH (-12)      /* work larger than the allocation-free scratch  */

#define MKHIP_MAX_PARTIES 32

/* context flags */
#define MKHIP_ROOT_RULE_SMALLEST2 1  /* primitive root search from 2 instead of 3  */

typedef struct mkhip_ctx mkhip_ctx;

const char *mkhip_status_string(int status);

/*
 * Parameter set and device tables.  Replaces the parameter/ring construction of
 * mkrlwe.NewParameters (mkrlwe/params.go:16-61) and mkckks.NewParameters
 * (mkckks/params.go:17-25): N = 2^logN, Q[0..L), P[0..K), gamma = 2 for the KKLSS
 * evaluators (alpha = K/gamma) and 1 for MKF (Lattigo rlwe, alpha = K).
 * logN in {13, 14, 15}.  t and qmul are for BFV contexts (0 / NULL otherwise): a
 * non-null qmul (L primes, the QMul chain of mkbfv.ParametersLiteral,
 * mkbfv/params.go:12-81) makes a KKLSS MK-BFV context over R = Q u QMul with
 * plaintext modulus t; it needs gamma = 2 and alpha = K/2 = 1 (K = 2 or 3).
 */
int mkhip_ctx_create(mkhip_ctx **out, int logN, const uint64_t *q, int L, const uint64_t *p, int K,
                     int gamma, uint64_t t, const uint64_t *qmul, int flags);
void mkhip_ctx_destroy(mkhip_ctx *ctx);
/* alpha, beta (at the top level), and the word count of one switching key */
int mkhip_ctx_info(const mkhip_ctx *ctx, int *alpha, int *beta, size_t *swk_words);

/*
 * Resident evaluation keys (HBM).  CRS index -1 is the KKLSS `u` (params.go:38),
 * index 0 its `a` (needed only by the key generator below);
 * the rotation / conjugation CRS indices are described with the Galois key switches.
 * Plans refer to the resident keys: destroy the plans that use a key before
 * removing or replacing it (the host-pointer entry points do this for their own
 * cached plan).
 * A party's relinearization key (b, d, v) replaces mkrlwe.RelinearizationKey
 * (keys.go:133-136) and RelinearizationKeySet.AddRelinearizationKey (keys.go:277).
 */
int mkhip_set_crs(mkhip_ctx *ctx, int index, const uint64_t *swk);
int mkhip_register_party(mkhip_ctx *ctx, int party, const uint64_t *b, const uint64_t *d, const uint64_t *v);
int mkhip_unregister_party(mkhip_ctx *ctx, int party);

/*
 * KKLSS MK-CKKS mul + relin + rescale, host buffers
 * (mkckks.Evaluator.MulRelinNew, evaluator.go:419-446 and :561-586).
 *   op0: [1+k0][level0+1][N] coefficient domain, parties ids0[k0]; op1 likewise.
 *   out: [1+ku][level_out+1][N]; ids_out receives the union of ids in ascending
 *        order (ku entries, ku <= k0+k1); level_out = min(level0, level1) - nres
 *        where nres comes from the reference's float64 scale loop with
 *        scale = scale0 * scale1 and minScale = min_scale (evaluator.go:383-388).
 * out must hold (1 + ku) * (min(level0,level1)+1) * N words.
 */
int mkhip_kklss_ckks_mulrelin(mkhip_ctx *ctx, const uint64_t *op0, int level0, double scale0, const int *ids0,
                              int k0, const uint64_t *op1, int level1, double scale1, const int *ids1, int k1,
                              double min_scale, uint64_t *out, int *ids_out, int *ku_out, int *level_out,
                              double *scale_out);

/* ---- device buffers (for callers that keep operands resident in HBM) ---- */
int mkhip_dev_alloc(mkhip_ctx *ctx, size_t bytes, void **dptr);
int mkhip_dev_free(mkhip_ctx *ctx, void *dptr);
int mkhip_h2d(mkhip_ctx *ctx, void *dst, const void *src, size_t bytes);
int mkhip_d2h(mkhip_ctx *ctx, void *dst, const void *src, size_t bytes);
/* device-to-device copy on the context's stream (asynchronous; see mkhip_sync) */
int mkhip_d2d(mkhip_ctx *ctx, void *dst, const void *src, size_t bytes);
int mkhip_sync(mkhip_ctx *ctx);

/*
 * Peak-rate probes of this GPU (no counterpart in the reference: measurement).
 * Runs `iters` timed launches (HIP events on the context's stream) of a probe and
 * writes its rate: MKHIP_PROBE_COPY bytes read + written per second by the
 * library's device copy over two buffers of `size` bytes (allocated and freed by
 * the call, size >= 1 MiB), MKHIP_PROBE_COPY_STREAM the same with a streaming
 * copy (several non-temporal 16-byte loads in flight a lane); the others count operations per second of work held
 * in registers, `size` steps per lane (0: 4096) -- v_mul_lo_u32, v_mad_u64_u32,
 * forward NTT butterflies (unreduced form of moduli below 2^57 / reduced form of
 * moduli up to 2^61), Montgomery products with a modular add, products summed in
 * carry-save form (MKHIP_PROBE_ACC3, the external products' MAC) and unreduced
 * inverse butterflies (MKHIP_PROBE_IBFLY_LAZY).
 */
#define MKHIP_PROBE_COPY 0
#define MKHIP_PROBE_MUL32 1
#define MKHIP_PROBE_MAD64 2
#define MKHIP_PROBE_BFLY_LAZY 3
#define MKHIP_PROBE_BFLY_RED 4
#define MKHIP_PROBE_MRED 5
#define MKHIP_PROBE_COPY_STREAM 6
#define MKHIP_PROBE_ACC3 7
#define MKHIP_PROBE_IBFLY_LAZY 8
int mkhip_probe_peak(mkhip_ctx *ctx, int kind, size_t size, int iters, double *rate);

/*
 * Batched KKLSS MK-CKKS mul + relin over device-resident operands.  A plan fixes
 * the batch shape (ops, parties, levels) and owns the descriptors and the
 * workspace; launching it replays one hipGraph of the whole batch.
 */
typedef struct {
    const uint64_t *op0, *op1; /* device pointers, layout as above           */
    uint64_t *out;             /* device pointer                              */
    int ids0[MKHIP_MAX_PARTIES], ids1[MKHIP_MAX_PARTIES];
    int k0, k1, level;         /* level of the product                        */
    int nres;                  /* rescales to apply (from mkhip_rescale_count) */
    int level0, level1;        /* KKLSS MK-CKKS: levels of op0 / op1 when above
                                  `level` (components of level0 + 1 / level1 + 1
                                  limbs, each operand decomposed at its own level,
                                  mkckks/evaluator.go:433-439, the product at
                                  `level`, :577-581); a value below `level` (e.g. 0
                                  from zero-initialisation) means `level`.  The
                                  KKLSS MK-BFV plans ignore them. */
} mkhip_kklss_op;

typedef struct mkhip_plan mkhip_plan;

int mkhip_plan_kklss_ckks(mkhip_ctx *ctx, int nops, const mkhip_kklss_op *ops, int use_graph, mkhip_plan **out);

/*
 * Allocation-free form of mkhip_plan_kklss_ckks + mkhip_plan_launch (eager) for
 * the same ops, returning when the outputs are written.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc).  The call allocates nothing on the
 * device: the plan's workspace and descriptor tables live in one fixed array of
 * MKHIP_NOALLOC_SCRATCH_BYTES per GPU (held by each GPU the library runs on until
 * the process ends); operands and outputs are the caller's device buffers.  Ops
 * whose workspace + descriptors exceed it fail with MKHIP_ERR_SCRATCH (one op of
 * the N = 2^14, 8 Q + 2 P, k = 4 shape takes about 180 MB).  Every call on a GPU
 * shares that GPU's array and leaves its contents behind (nothing needs
 * clearing: no kernel reads workspace it has not written in the same call);
 * calls that overlap on one GPU -- from two threads, or from contexts whose
 * streams run at once -- race.
 */
#define MKHIP_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int mkhip_kklss_ckks_mulrelin_noalloc(mkhip_ctx *ctx, int nops, const mkhip_kklss_op *ops);
int mkhip_plan_launch(mkhip_plan *plan);
/* average device time of the plan's kernels over `iters` launches (HIP events) */
int mkhip_plan_time(mkhip_plan *plan, int iters, double *ms_per_launch);
void mkhip_plan_destroy(mkhip_plan *plan);
size_t mkhip_plan_workspace_bytes(const mkhip_plan *plan);
/* A plan is a fixed sequence of kernel launches ("stages").  stage_times gives
 * the device time of each (ms, averaged over `iters` eager launches, HIP events
 * on the context stream); ms must hold mkhip_plan_stage_count(plan) doubles. */
int mkhip_plan_stage_count(const mkhip_plan *plan);
const char *mkhip_plan_stage_name(const mkhip_plan *plan, int stage);
int mkhip_plan_stage_times(mkhip_plan *plan, int iters, double *ms);

/*
 * MKF MK-CKKS (software/mkf/mkf_cpu/ckks.go:102-151, CKKSCase.Run): fold both
 * operands to the aggregate secret (common.go:79-126), Karatsuba tensor, one
 * relinearization, rescale, k-fold materialization of Gamma.  Needs a context
 * with gamma = 1 (alpha = K, the Lattigo rlwe decomposition).
 *
 * A group is one party set with its fold keys s_i -> s_aggregate and the
 * relinearization key of s_aggregate (common.go:32-51), each a Lattigo
 * SwitchingKey laid out [beta][2][L+K][N] (mkf_cpu/fpga_fixture_test.go:56-69).
 */
int mkhip_register_mkf_group(mkhip_ctx *ctx, int group, int k, const uint64_t *const *fold_keys,
                             const uint64_t *relin_key);
int mkhip_unregister_mkf_group(mkhip_ctx *ctx, int group);
/* host buffers: left/right [1+k][level+1][N] NTT domain; out [1+k][level_out+1][N]
 * NTT domain = c0 followed by k copies of Gamma.  Rescale errors are returned. */
int mkhip_mkf_ckks_mulrelin(mkhip_ctx *ctx, int group, const uint64_t *left, const uint64_t *right, int level,
                            double scale_l, double scale_r, double min_scale, uint64_t *out, int *level_out,
                            double *scale_out);

typedef struct {
    const uint64_t *left, *right; /* device, [1+k][level+1][N], NTT domain   */
    uint64_t *out;                /* device, [1+k][level+1-nres][N]          */
    int group, level, nres;       /* nres in {0, 1}                           */
} mkhip_mkf_op;

int mkhip_plan_mkf_ckks(mkhip_ctx *ctx, int nops, const mkhip_mkf_op *ops, int use_graph, mkhip_plan **out);

/*
 * MKF MK-BFV (software/mkf/mkf_cpu/bfv.go:212-252, BFVCase.Run) on a BFV context
 * with gamma = 1 (Lattigo rlwe decomposition; QMul = the KKLSS chain the MKF code
 * reuses, mkf_cpu/params.go:69-73): fold both sides, lift the left operand to R
 * and rescale the right one into R, Karatsuba tensor, quantize, relinearize,
 * k-fold materialization of Gamma.  Keys: an MKF group as above.
 * Host buffers left/right [1+k][L][N] coefficient domain (C0 then the parties'
 * components); out [1+k][L][N] coefficient domain: c0 then k copies of Gamma.
 */
int mkhip_mkf_bfv_mulrelin(mkhip_ctx *ctx, int group, const uint64_t *left, const uint64_t *right,
                           uint64_t *out);
/* batched over device operands: every op has level = L-1 and nres = 0 */
int mkhip_plan_mkf_bfv(mkhip_ctx *ctx, int nops, const mkhip_mkf_op *ops, int use_graph, mkhip_plan **out);

/*
 * KKLSS MK-BFV (software/kklss/mkbfv).  A party's key is mkbfv.RelinearizationKey
 * (mkbfv/keys.go:6-9): two mkrlwe relinearization keys, one per gadget branch;
 * the evaluator reads b and d of both branches and v of branch 0
 * (keyswitch_hoisted.go:105-229), each [beta][L+K][N] as above.  Replaces
 * RelinearizationKeySet.AddRelinearizationKey (mkbfv/keys.go:57-67).
 */
int mkhip_register_bfv_party(mkhip_ctx *ctx, int party, const uint64_t *b0, const uint64_t *d0,
                             const uint64_t *v0, const uint64_t *b1, const uint64_t *d1);
/* mkbfv.Evaluator.MulRelinNew (evaluator.go:85-91): host buffers op0 [1+k0][L][N],
 * op1 [1+k1][L][N] coefficient domain; out [1+ku][L][N] with the union of ids in
 * ascending order written to ids_out (ku entries). */
int mkhip_kklss_bfv_mulrelin(mkhip_ctx *ctx, const uint64_t *op0, const int *ids0, int k0, const uint64_t *op1,
                             const int *ids1, int k1, uint64_t *out, int *ids_out, int *ku_out);
/* batched over device operands: every op has level = L-1 and nres = 0 */
int mkhip_plan_kklss_bfv(mkhip_ctx *ctx, int nops, const mkhip_kklss_op *ops, int use_graph, mkhip_plan **out);
/* R-basis primitives of mkbfv.FastBasisExtender (mkbfv/basis_extension.go), host buffers:
 *   mkhip_bfv_modup_q_to_r   ModUpQtoR (:50-66): [L][N] -> [2L][N] (QMul limbs unreduced,
 *                            congruent, as multSum returns them)
 *   mkhip_bfv_rescale_q_to_r Rescale (:88-104): floor(QMul x / Q) into R, [L][N] -> [2L][N]
 *   mkhip_bfv_quantize       Quantize (:69-85): NTT-domain [2L][N] -> [L][N] coefficient
 *   mkhip_bfv_decompose      KeySwitcher.DecomposeBFV (mkbfv/keyswitch.go:58-83):
 *                            [2L][N] -> [2L][L+K][N] NTT domain */
int mkhip_bfv_modup_q_to_r(mkhip_ctx *ctx, const uint64_t *in, uint64_t *out);
int mkhip_bfv_rescale_q_to_r(mkhip_ctx *ctx, const uint64_t *in, uint64_t *out);
int mkhip_bfv_quantize(mkhip_ctx *ctx, const uint64_t *in_ntt, uint64_t *out);
int mkhip_bfv_decompose(mkhip_ctx *ctx, const uint64_t *aR, uint64_t *out);

/*
 * Galois key switches (rotations and conjugation).
 *
 * KKLSS (mkrlwe, both the MK-CKKS and the MK-BFV evaluators).  Keys are mkrlwe
 * switching keys [beta][L+K][N] (MForm, NTT) as the relinearization keys:
 *   mkhip_set_crs(ctx, r, a)            CRS[r] of rotation r (1 <= r < N/2) or of
 *                                       conjugation (r = -2) (params.go:35-58, AddCRS)
 *   mkhip_add_rotation_key(ctx, id, r)  RotationKeySet.AddRotationKey (keys.go:133-143),
 *                                       a key made by GenRotationKey (keygen.go:190-230)
 *   mkhip_del_rotation_key              RotationKeySet.DelRotationKey (keys.go:146-148)
 *   mkhip_add_conjugation_key           ConjugationKeySet.AddConjugationKey (keys.go:208-211),
 *                                       GenConjugationKey (keygen.go:240-268)
 *   mkhip_del_conjugation_key           ConjugationKeySet.DelConjugationKey (keys.go:213-216)
 *
 * mkhip_kklss_rotate: Evaluator.RotateNew (hoisted = 0; mkckks/evaluator.go:488-529,
 * mkbfv/evaluator.go:156-194) or RotateHoistedNew (hoisted = 1; mkckks :590-624).
 * rotidx is normalised into [0, N/2); 0 copies the input; a rotation with a CRS
 * entry is one key switch (keyswitch.go:243-307); otherwise RotateNew applies the
 * rotations 1, 2, 4, ... of rotidx's binary expansion one after the other and
 * RotateHoistedNew fails with MKHIP_ERR_NO_KEY (the reference panics).
 * mkhip_kklss_conjugate: Evaluator.ConjugateNew (keyswitch.go:309-344).
 * in / out: [1+k][level+1][N] coefficient domain, parties ids[k]; out keeps the
 * party order.  Every key switch uses the hoisted external product
 * (keyswitch_hoisted.go:11-45); the reference's non-hoisted ExternalProduct
 * (keyswitch.go:82-130) multiplies every key digit by the last digit's
 * decomposition, so for more than one digit its Rotate / Conjugate outputs differ
 * from these (and from its own RotateHoisted).
 */
#define MKHIP_MAX_ROTATIONS 16
#define MKHIP_CONJUGATE (-2) /* rot[] entry of a conjugation (its CRS index) */
int mkhip_add_rotation_key(mkhip_ctx *ctx, int party, int rotidx, const uint64_t *swk);
int mkhip_del_rotation_key(mkhip_ctx *ctx, int party, int rotidx);
int mkhip_add_conjugation_key(mkhip_ctx *ctx, int party, const uint64_t *swk);
int mkhip_del_conjugation_key(mkhip_ctx *ctx, int party);
int mkhip_kklss_rotate(mkhip_ctx *ctx, const uint64_t *in, int level, const int *ids, int k, int rotidx,
                       int hoisted, uint64_t *out);
int mkhip_kklss_conjugate(mkhip_ctx *ctx, const uint64_t *in, int level, const int *ids, int k, uint64_t *out);

/* Batched over device operands.  One decomposition of the input serves every
 * rotation of the op (the hoisting of RotateHoisted); a conjugation decomposes
 * the conjugated input.  rot[r] is a normalised rotidx with a CRS entry, or
 * MKHIP_CONJUGATE; out holds nrot ciphertexts back to back, each
 * [1+k][level+1][N].  Every op of a plan has the same level. */
typedef struct {
    const uint64_t *in; /* device, [1+k][level+1][N] coefficient domain */
    uint64_t *out;      /* device, nrot x [1+k][level+1][N]             */
    int ids[MKHIP_MAX_PARTIES];
    int k, level, nrot;
    int rot[MKHIP_MAX_ROTATIONS];
} mkhip_rotate_op;
int mkhip_plan_kklss_rotate(mkhip_ctx *ctx, int nops, const mkhip_rotate_op *ops, int use_graph,
                            mkhip_plan **out);

/*
 * MKF party-specific rotation (paperMKFCKKSEvaluator.RotateNew,
 * software/mkf/mkf_cpu/true_mkf_ckks.go:323-353) on a gamma = 1 CKKS context.
 * mkhip_register_mkf_rotation stores, for the Galois element of `rotation`
 * (5^rotation mod 2N), the group's k party keys sigma(s_i) -> s_aggregate
 * (Lattigo switching keys [beta][2][L+K][N], true_mkf_ckks.go:379-390).
 * mkhip_mkf_ckks_rotate: in [1+k][level+1][N] NTT domain; out [1+k][level+1][N]
 * NTT domain: sigma(C0) + sum_i u0_i, then k copies of sum_i u1_i.  A missing key
 * set is MKHIP_ERR_NO_KEY (the reference's error at :328-330).
 */
int mkhip_register_mkf_rotation(mkhip_ctx *ctx, int group, int rotation, const uint64_t *const *keys);
int mkhip_mkf_ckks_rotate(mkhip_ctx *ctx, int group, const uint64_t *in, int level, int rotation, uint64_t *out);
typedef struct {
    const uint64_t *in; /* device, [1+k][level+1][N] NTT domain */
    uint64_t *out;      /* device, [1+k][level+1][N] NTT domain */
    int group, level, rotation;
} mkhip_mkf_rotate_op;
int mkhip_plan_mkf_rotate(mkhip_ctx *ctx, int nops, const mkhip_mkf_rotate_op *ops, int use_graph,
                          mkhip_plan **out);

/*
 * MKF CKKS polynomial evaluation (paperMKFCKKSEvaluator.EvaluatePoly,
 * software/mkf/mkf_cpu/true_mkf_ckks_polynomial.go:79-273) and the evaluator's
 * public linear operations (true_mkf_ckks.go:187-303), on a gamma = 1 CKKS
 * context.  Ciphertexts are NTT-domain [1+k][level+1][N] (c0, then the k party
 * components), as paperMKFCKKSCiphertext holds them; scales are float64.
 *
 * The polynomial mirrors ckks.Polynomial: coefficients (imaginary parts may be
 * NULL), MaxDeg, Lead and the basis; A and B (Chebyshev interval) are the
 * caller's change of variable, applied before the call as in the reference.
 *   mkhip_plan_mkf_poly            nct ciphertexts (device) through the same
 *                                  polynomial, as one plan: its schedule of
 *                                  MulRelin (each operand folded at its own level),
 *                                  Rescale, Add/Sub, AddConst and
 *                                  MultByGaussianIntegerAndAdd steps; out[i] receives
 *                                  [1+k][level_out+1][N].  Errors: empty polynomial
 *                                  MKHIP_ERR_ARG; depth above the level (the
 *                                  reference's "levels < polynomial depth")
 *                                  MKHIP_ERR_LEVEL; rescale errors MKHIP_ERR_SCALE.
 *   mkhip_mkf_poly_query           output level, scale and MulRelin count only
 *   mkhip_mkf_ckks_evaluate_poly   one ciphertext, host buffers
 *   mkhip_mkf_ckks_add             AddNew / SubNew (sub = 1): the operand of smaller
 *                                  scale first multiplied by floor(ratio) when that is
 *                                  above 1; output at the smaller level, larger scale
 *   mkhip_mkf_ckks_add_const       AddConstNew(re + im i): c0 only, scaled by `scale`
 *   mkhip_mkf_ckks_mult_by_const   MultByConstNew: a constant with a fractional part
 *                                  is scaled by Q[level] and so is the output scale
 *   mkhip_mkf_ckks_mult_by_gaussian_integer_and_add
 *                                  acc (at level_acc, in place) += in (re + im i)
 *   mkhip_mkf_ckks_rescale         RescaleNew(min_scale): as many primes as the scale
 *                                  loop divides (mkckks/evaluator.go:360-407)
 *   mkhip_mkf_ckks_mulrelin_levels MulRelinNew: operands at their own levels, output at
 *                                  the smaller one, c0 then k copies of Gamma, no
 *                                  rescale (output scale = product of the scales)
 * DropLevelNew is a prefix of the limbs of each component: no call.
 */
#define MKHIP_BASIS_STANDARD 0
#define MKHIP_BASIS_CHEBYSHEV 1
typedef struct {
    const double *coeffs_re, *coeffs_im; /* ncoeffs each (coeffs_im may be NULL) */
    int ncoeffs;                         /* degree + 1                           */
    int max_deg, lead, basis;            /* Polynomial.MaxDeg, .Lead, .Basis      */
} mkhip_polynomial;
int mkhip_plan_mkf_poly(mkhip_ctx *ctx, int group, int nct, const uint64_t *const *in, int level, double scale,
                        const mkhip_polynomial *poly, double target_scale, uint64_t *const *out, int use_graph,
                        mkhip_plan **plan, int *level_out, double *scale_out);
int mkhip_mkf_poly_query(mkhip_ctx *ctx, int group, int level, double scale, const mkhip_polynomial *poly,
                         double target_scale, int *level_out, double *scale_out, int *mulrelins);
int mkhip_mkf_ckks_evaluate_poly(mkhip_ctx *ctx, int group, const uint64_t *in, int level, double scale,
                                 const mkhip_polynomial *poly, double target_scale, uint64_t *out, int *level_out,
                                 double *scale_out);
int mkhip_mkf_ckks_add(mkhip_ctx *ctx, int k, const uint64_t *a, int level_a, double scale_a, const uint64_t *b,
                       int level_b, double scale_b, int sub, uint64_t *out, int *level_out, double *scale_out);
int mkhip_mkf_ckks_add_const(mkhip_ctx *ctx, int k, const uint64_t *in, int level, double scale, double re,
                             double im, uint64_t *out);
int mkhip_mkf_ckks_mult_by_const(mkhip_ctx *ctx, int k, const uint64_t *in, int level, double scale, double re,
                                 double im, uint64_t *out, double *scale_out);
int mkhip_mkf_ckks_mult_by_gaussian_integer_and_add(mkhip_ctx *ctx, int k, const uint64_t *in, int level_in,
                                                    int64_t re, int64_t im, uint64_t *acc, int level_acc);
int mkhip_mkf_ckks_rescale(mkhip_ctx *ctx, int k, const uint64_t *in, int level, double scale, double min_scale,
                           uint64_t *out, int *level_out, double *scale_out);
int mkhip_mkf_ckks_mulrelin_levels(mkhip_ctx *ctx, int group, const uint64_t *left, int level_l,
                                   const uint64_t *right, int level_r, uint64_t *out);

/*
 * Key generation and public-key encryption on the device (mkrlwe/keygen.go,
 * encryptor.go; the step before the mul+relin path, feeding party registration).
 * The reference samples secrets, errors and the encryption's v inside these calls
 * from its crypto PRNG; here they are inputs: small signed polynomials, int64 [N]
 * (ternary or Gaussian coefficients; the caller samples them).  Secrets are
 * MForm(NTT(s)) over Q u P, [L+K][N]; keys [beta][L+K][N] as elsewhere.
 *   mkhip_gen_secret           genSecretKeyFromSampler (keygen.go:44-56)
 *   mkhip_gen_public_key       GenPublicKey (:88-110): pk0 = NTT(e) - MRed(sk, a0),
 *                              pk1 = a0 (digit 0 of CRS[0]), both [L+K][N]
 *   mkhip_gen_relin_key        GenRelinearizationKey(sk, r) (:137-187): errs [3][beta][N]
 *                              are the errors of b, d, v; uses CRS[0] and CRS[-1]
 *   mkhip_register_party_keygen the same, registered for `party` without leaving HBM
 *   mkhip_gen_rotation_key     GenRotationKey (:190-230), errs [beta][N], CRS[rotidx]
 *   mkhip_gen_conjugation_key  GenConjugationKey (:240-268), errs [beta][N], CRS[-2]
 *   mkhip_encrypt              Encryptor.Encrypt (encryptor.go:55-119) at `level`:
 *                              v ternary, e0 / e1 errors, pt [level+1][N] (NTT domain
 *                              when pt_ntt); ciphertext (c0, c1) [level+1][N], NTT
 *                              domain when ntt_out
 * CRS[0] is uploaded with mkhip_set_crs(ctx, 0, a).
 */
int mkhip_gen_secret(mkhip_ctx *ctx, const int64_t *s, uint64_t *sk);
int mkhip_gen_public_key(mkhip_ctx *ctx, const uint64_t *sk, const int64_t *e, const uint64_t *a0, uint64_t *pk0,
                         uint64_t *pk1);
int mkhip_gen_relin_key(mkhip_ctx *ctx, const uint64_t *sk, const uint64_t *r, const int64_t *errs, uint64_t *b,
                        uint64_t *d, uint64_t *v);
int mkhip_register_party_keygen(mkhip_ctx *ctx, int party, const uint64_t *sk, const uint64_t *r,
                                const int64_t *errs);
int mkhip_gen_rotation_key(mkhip_ctx *ctx, int rotidx, const uint64_t *sk, const int64_t *errs, uint64_t *rk);
int mkhip_gen_conjugation_key(mkhip_ctx *ctx, const uint64_t *sk, const int64_t *errs, uint64_t *ck);
int mkhip_encrypt(mkhip_ctx *ctx, const uint64_t *pk0, const uint64_t *pk1, int level, const int64_t *v,
                  const int64_t *e0, const int64_t *e1, const uint64_t *pt, int pt_ntt, int ntt_out, uint64_t *c0,
                  uint64_t *c1);

/*
 * KKLSS MK-BFV relinearization keys (mkbfv/keygen.go:24-162), a BFV context with
 * gamma 2 (one prime per digit): errs [5][beta][N] are the errors of b_1, b_2,
 * d_1, d_2 and v; CRS[0] (a_1), CRS[-3] (a_2, mkhip_set_crs(ctx, -3, a)) and u
 * must be uploaded.
 *   b_br = MForm(e) - MRed(a_br, s)                                  (:46-63)
 *   d_br = MForm(G_br,i s + e) - MRed(a_br, r), GenBFVSwitchingKey's gadget
 *          G_1,i = (Q QMul / q_i) t [(Q QMul / q_i)^-1]_q_i P / QMul over the Q
 *          primes and G_2,i likewise over the QMul primes        (:65-73, 91-162)
 *   v    = -(MForm(e + P r g_i) + MRed(u, s))                         (:75-85)
 * b and d hold both branches (2 beta digits: branch 1 then 2) as
 * mkhip_register_bfv_party takes them; mkhip_register_bfv_party_keygen generates
 * and registers them for `party` without leaving HBM.
 */
int mkhip_gen_bfv_relin_key(mkhip_ctx *ctx, const uint64_t *sk, const uint64_t *r, const int64_t *errs, uint64_t *b,
                            uint64_t *d, uint64_t *v);
int mkhip_register_bfv_party_keygen(mkhip_ctx *ctx, int party, const uint64_t *sk, const uint64_t *r,
                                    const int64_t *errs);

/*
 * MKF keys (mkf_cpu/common.go:32-51, Lattigo rlwe.KeyGenerator), a gamma-1 context:
 *   mkhip_gen_mkf_switching_key  GenSwitchingKey(s_in, s_out): [beta][2][L+K][N] =
 *                                (MForm(e + P s_in g_i) - MRed(a_i, s_out), a_i); a
 *                                [beta][L+K][N] uniform (the reference samples it
 *                                directly in NTT and Montgomery form), errs [beta][N]
 *   mkhip_gen_mkf_relin_key      GenRelinearizationKey(s, 1): the switching key s^2 -> s
 *   mkhip_register_mkf_group_keygen generateOnlineKeys for the k secrets sks
 *                                [k][L+K][N]: fold keys s_i -> s = sum s_i (aggregateSecret)
 *                                and the relinearization key of s, registered as `group`
 *                                without leaving HBM; a [k+1][beta][L+K][N] and errs
 *                                [k+1][beta][N] (fold keys in party order, then relin)
 */
int mkhip_gen_mkf_switching_key(mkhip_ctx *ctx, const uint64_t *s_in, const uint64_t *s_out, const uint64_t *a,
                                const int64_t *errs, uint64_t *swk);
int mkhip_gen_mkf_relin_key(mkhip_ctx *ctx, const uint64_t *s, const uint64_t *a, const int64_t *errs, uint64_t *swk);
int mkhip_register_mkf_group_keygen(mkhip_ctx *ctx, int group, int k, const uint64_t *sks, const uint64_t *a,
                                    const int64_t *errs);

/*
 * Tuning knobs (no counterpart in the reference: they choose between equivalent
 * kernel schedules; every setting gives the same bits).
 *   MKHIP_TUNE_FUSED_DECOMP  alpha = 1 KKLSS plans: the gadget decomposition fused
 *                            with the x / y key MAC (1) or as two launches (0)
 *   MKHIP_TUNE_XCD_REMAP     workgroups ordered XCD-contiguously (1) or as issued (0)
 *                            (every limb kernel but the external products)
 *   MKHIP_TUNE_DECOMP_ORDER  fused decomposition grid order: ops sharing a key row
 *                            adjacent (0), rows sharing an input limb adjacent (1), or
 *                            XCD-tiled (2: each XCD walks every op of one (side,
 *                            digit, block of up to 5 rows), so input limbs and key
 *                            rows are fetched from HBM about once each)
 *   MKHIP_TUNE_FP64          transforms of limbs whose modulus is below 2^50 in 64-bit
 *                            integer arithmetic (0, default) or in exact FP64
 *                            two-product arithmetic (1, opt-in; the same words)
 *   MKHIP_TUNE_SPLIT_LIMB    N = 2^14 one-prime digit decompositions (fused and
 *                            plain): each limb over two workgroups of half the LDS (1)
 *                            or one workgroup (0)
 *   MKHIP_TUNE_SPLIT_EXT     N = 2^14 external products: each limb over two workgroups
 *                            of half the LDS, the last inverse stage done by the ModDown
 *                            that reads the rows (1), or one workgroup (0); at N = 2^15
 *                            the external products always run so
 *   MKHIP_TUNE_DIGIT_PRE     MKF CKKS plans with CRT digits (alpha >= 2): the digit-only
 *                            part of the ModUp (y_l, the overflow estimate) computed once
 *                            per digit by k_digit_pre (1) or by every target row (0)
 *   MKHIP_TUNE_XCD_EXT       external-product kernels (k_ext, k_mac2) in XCD-contiguous
 *                            order (1, default) or as issued (0)
 *   MKHIP_TUNE_EXT_CS        external-product MACs (k_ext, k_mac2): each sum of key x digit
 *                            products kept as an exact carry-save 128-bit sum and reduced
 *                            once (1, default, for rows whose sums fit) or every product
 *                            Montgomery-reduced (0); the same words
 *   MKHIP_TUNE_INV_LAZY      inverse transforms of limbs with 2^32 <= q < 2^57: sums of
 *                            the Gentleman-Sande butterflies left unreduced, the values
 *                            brought below q only when the next pass could overflow
 *                            (1, default) or reduced every stage (0); the same words
 * A knob applies to plans built after it is set: a plan keeps the settings it was
 * built with for every later launch, graph capture and stage timing.  mkhip_autotune plans and times
 * every applicable combination of the decomposition knobs on the caller's batch, then
 * the best with MKHIP_TUNE_XCD_EXT flipped (ops: mkhip_kklss_op[] for the
 * KKLSS schemes, mkhip_mkf_op[] for MKF; their outputs are overwritten with the
 * results), keeps the fastest, and writes up to 16 candidate times (ms per launch).
 */
#define MKHIP_TUNE_FUSED_DECOMP 1
#define MKHIP_TUNE_XCD_REMAP 2
#define MKHIP_TUNE_DECOMP_ORDER 3
#define MKHIP_TUNE_FP64 4
#define MKHIP_TUNE_SPLIT_LIMB 5
#define MKHIP_TUNE_SPLIT_EXT 6
#define MKHIP_TUNE_DIGIT_PRE 7
#define MKHIP_TUNE_XCD_EXT 8
#define MKHIP_TUNE_EXT_CS 9
#define MKHIP_TUNE_INV_LAZY 10
#define MKHIP_SCHEME_KKLSS_CKKS 0
#define MKHIP_SCHEME_KKLSS_BFV 1
#define MKHIP_SCHEME_MKF_CKKS 2
#define MKHIP_SCHEME_MKF_BFV 3
#define MKHIP_SCHEME_KKLSS_ROTATE 4 /* ops: mkhip_rotate_op[]     */
#define MKHIP_SCHEME_MKF_ROTATE 5   /* ops: mkhip_mkf_rotate_op[] */
int mkhip_set_tuning(mkhip_ctx *ctx, int knob, int value);
int mkhip_get_tuning(const mkhip_ctx *ctx, int knob, int *value);
int mkhip_autotune(mkhip_ctx *ctx, int scheme, int nops, const void *ops, int iters, int *ncand, double *ms);

/*
 * Uneven batches in one hipGraph launch (no counterpart in the reference, which
 * runs one MulRelinNew at a time; this is the batch_submit / batch_wait surface a
 * Go caller uses to hand over many ops at once).  Each mkhip_batch_add gives the
 * ops of one context -- any mix of party counts, device-resident operands, one
 * scheme (MKHIP_SCHEME_*, ops of the matching struct type) -- as one branch.  Its
 * ops are cut in order into chunks whose plan workspace fits ws_budget bytes
 * (0: no limit; a chunk holds at least one op); the chunks of a branch run in
 * sequence in one shared workspace, the branches run concurrently.
 * mkhip_batch_finalize runs the batch once eagerly and captures it as one graph;
 * mkhip_batch_submit launches that graph (asynchronous), mkhip_batch_wait blocks
 * until it completed.  Contexts used by a batch must outlive it.
 */
typedef struct mkhip_batch mkhip_batch;
int mkhip_batch_create(mkhip_batch **out);
int mkhip_batch_add(mkhip_batch *batch, mkhip_ctx *ctx, int scheme, int nops, const void *ops, size_t ws_budget,
                    int *nchunks);
int mkhip_batch_finalize(mkhip_batch *batch);
int mkhip_batch_submit(mkhip_batch *batch);
int mkhip_batch_wait(mkhip_batch *batch);
/* average device time of `iters` back-to-back graph launches (HIP events), ms */
int mkhip_batch_time(mkhip_batch *batch, int iters, double *ms_per_launch);
/* device bytes held: workspace arenas + descriptor tables */
size_t mkhip_batch_workspace_bytes(const mkhip_batch *batch);
int mkhip_batch_branch_info(const mkhip_batch *batch, int branch, int *nchunks, size_t *arena_bytes);
void mkhip_batch_destroy(mkhip_batch *batch);

/*
 * Serving (no counterpart in the reference: the enqueue / poll surface of a
 * request loop).  A server owns `streams` HIP streams and a device pool of
 * mem_bytes; each class is one context (its parties' keys resident in HBM) and
 * one scheme.  A scheduler thread batches waiting requests continuously: when a
 * stream is free, the class with the oldest waiting request gets a batch of its
 * waiting requests (oldest first, at most max_batch) whose device block --
 * operands, outputs, workspace, descriptors -- fits the largest free extent of
 * the pool; the request waits otherwise (admission by free device memory).
 * Batches on different streams overlap on the device.  A request that cannot
 * fit an empty pool fails with MKHIP_ERR_OOM.
 *
 * Requests carry host operands at the top level (L limbs per component) in the
 * layouts of the single-op entry points; out receives the result ((1 + ku) x
 * (L - nres) x N for KKLSS, ku = |ids0 u ids1| in ascending id order; (1 + k) x
 * (L - nres) x N for MKF).  in0, in1 and out must stay valid until the request
 * completed; pinned buffers (mkhip_host_alloc) make the copies asynchronous.
 * With on_device = 1, in0, in1 and out are device pointers instead (ciphertexts
 * that stay in HBM between operations, such as the results of earlier ones): the
 * batch reads and writes them in place, with no copies, and takes pool memory
 * for workspace and descriptors only; latency then ends when the result is in
 * device memory.
 * Contexts attached to a server must not be used by other threads while it runs.
 */
typedef struct mkhip_server mkhip_server;
typedef struct {
    int streams;      /* batches in flight at once (>= 1)                        */
    int max_batch;    /* ops per batch at most (>= 1)                            */
    size_t mem_bytes; /* device pool for operands, outputs, workspace, descriptors */
} mkhip_server_config;
typedef struct {
    int cls;                                         /* from mkhip_server_add_class */
    const uint64_t *in0, *in1;                       /* host operands               */
    uint64_t *out;                                   /* host result                 */
    int ids0[MKHIP_MAX_PARTIES], ids1[MKHIP_MAX_PARTIES];
    int k0, k1;                                      /* KKLSS party ids per operand */
    int group;                                       /* MKF party group             */
    int nres;                                        /* CKKS rescales; 0 for BFV    */
    int on_device;                                   /* 1: in0, in1, out in HBM     */
} mkhip_request;
typedef struct {
    uint64_t completed, batches, max_batch;
    double mean_batch;
    int max_inflight;
    size_t pool_bytes, pool_free_largest;
} mkhip_server_stats_t;
int mkhip_server_create(const mkhip_server_config *cfg, mkhip_server **out);
int mkhip_server_add_class(mkhip_server *server, mkhip_ctx *ctx, int scheme, int *cls);
int mkhip_server_start(mkhip_server *server);
int mkhip_server_enqueue(mkhip_server *server, const mkhip_request *req, uint64_t *ticket);
/* done = 1 once the request finished; status, latency (enqueue -> results in
 * host memory) and queueing time (enqueue -> batch launch) are then reported
 * and the ticket is released */
int mkhip_server_poll(mkhip_server *server, uint64_t ticket, int *done, int *status, double *latency_ms,
                      double *queue_ms);
int mkhip_server_wait(mkhip_server *server, uint64_t ticket, int *status, double *latency_ms, double *queue_ms);
int mkhip_server_stats(mkhip_server *server, mkhip_server_stats_t *stats);
/* stops the scheduler after the batches in flight completed */
void mkhip_server_destroy(mkhip_server *server);
/* pinned host memory */
int mkhip_host_alloc(size_t bytes, void **ptr);
int mkhip_host_free(void *ptr);

/* reference scale loop (mkckks/evaluator.go:383-388): number of rescales */
int mkhip_rescale_count(const mkhip_ctx *ctx, int level, double scale, double min_scale, int *nres,
                        double *new_scale);

/*
 * RNS primitives on host buffers -- the per-primitive operations the reference
 * times in software/mkf/cmd/rnsbench/main.go:179-201.
 *   mkhip_ntt_limbs        ring NTTLvl / InvNTTLvl of nlimbs rows; ring_idx[i] selects
 *                          the modulus (0..L-1 = Q, L..L+K-1 = P, then BFV QMul) of row i
 *   mkhip_decompose        mkrlwe KeySwitcher.Decompose (mkrlwe/keyswitch.go:50-76):
 *                          a [level+1][N] coefficient -> [beta][level+1+K][N] NTT
 *   mkhip_moddown          FastBasisExtender.ModDownQPtoQ (basis_extension.go:192-232):
 *                          x [level+1+K][N] coefficient -> [level+1][N]
 *   mkhip_external_product KeySwitcher.ExternalProductHoisted (keyswitch_hoisted.go:11-45)
 */
int mkhip_ntt_limbs(mkhip_ctx *ctx, const uint64_t *in, uint64_t *out, const int *ring_idx, int nlimbs,
                    int inverse);
int mkhip_decompose(mkhip_ctx *ctx, const uint64_t *a, int level, uint64_t *out);
int mkhip_moddown(mkhip_ctx *ctx, const uint64_t *x, int level, uint64_t *out);
int mkhip_external_product(mkhip_ctx *ctx, const uint64_t *a_hoisted, const uint64_t *key, int level,
                           uint64_t *out);

#ifdef __cplusplus
}
#endif
#endif /* MKHIP_H */
