/* fls_wizard.h -- encoding selection on the GPU: exhaustive trial sizing and a learned selector.
 *
 * The reference picks an expression for a column by encoding it with every expression of its pool and keeping the
 * smallest (ExpColEncoder<PT>::encode_all + AllExpEncodedCol::GetBest, reference
 * src/encoder/exp_col_encoder.cpp:19-28, 117-129; pools in src/cor/exp_pool.cpp:23-144). Here the unit is a row-group
 * of 65536 int64 values (64 vectors, one page per candidate). The exhaustive form computes every candidate's exact
 * encoded size in closed form on the GPU (no bytes are produced); the learned form reads a sample of each row-group,
 * computes 64 statistics and runs a 64-256-256-K MLP (BF16 MFMA, FP32 accumulate) that was trained on the
 * exhaustive form's choices. Either way the chosen expression is then encoded by the bit-exact codec, so the choice
 * never changes decoded bytes -- only the size.
 *
 * All pointers named dev_* are device pointers; calls are asynchronous on `stream` (a hipStream_t, NULL = default).
 * Return codes are the FLS_E_* values of fls_gpu.h; fls_gpu_error_string() describes the last failure. */
#ifndef FLS_WIZARD_H
#define FLS_WIZARD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FLS_WIZ_RG_VALUES 65536u  /* values per row-group */
#define FLS_WIZ_N_FEATURES 64u
#define FLS_WIZ_SAMPLE_VECTORS 4u /* vectors of a row-group the features read: 0, 21, 42, 63 */
#define FLS_WIZ_MAX_CLASSES 32u

/* Candidate pool (int64 expression ids of include/fastlanes.h / the reference's i64 pool):
 * 0 UNCOMPRESSED, 1 FFOR, 11 DELTA, 6 RLE (FFOR-delta positions), 7 RLE (plain positions),
 * 16 DICT (plain codes), 17 DICT (FFOR codes), 22 DICT (delta codes). */
uint32_t fls_wizard_n_candidates(void);
/* Expression id of candidate k (k < fls_wizard_n_candidates()), or 255. */
uint8_t  fls_wizard_candidate_expr(uint32_t k);

/* Exhaustive trial sizing: dev_sizes[r * K + k] = exact encoded bytes of row-group r under candidate k (the size
 * encode_from_memory would return for those 65536 values). Allocates its hash-table workspace on first use. */
int fls_wizard_sizes(const int64_t* dev_values, uint64_t n_rowgroups, uint64_t* dev_sizes, void* stream);
/* argmin over candidates of dev_sizes (ties -> lowest k). */
int fls_wizard_best(const uint64_t* dev_sizes, uint64_t n_rowgroups, int32_t* dev_choice, void* stream);
/* The 64 sample statistics of each row-group (float32, row-major n_rowgroups x 64). */
int fls_wizard_features(const int64_t* dev_values, uint64_t n_rowgroups, float* dev_features, void* stream);

/* ---- MLP 64-256-256-K (K <= 32 classes): BF16 MFMA GEMMs, FP32 master weights, Adam -------------------------- */
typedef struct fls_mlp fls_mlp;

int  fls_mlp_create(uint32_t n_classes, uint32_t max_batch, uint64_t seed, fls_mlp** out);
void fls_mlp_destroy(fls_mlp* m);
/* Parameter block (float32): W1t[256][64] b1[256] W2t[256][256] b2[256] W3t[32][256] b3[32] (Wt[out][in]). */
uint64_t fls_mlp_n_params(void);
int fls_mlp_get_params(fls_mlp* m, float* host_params);
int fls_mlp_set_params(fls_mlp* m, const float* host_params);
/* Forward + backward on one batch (batch <= max_batch, multiple of 32): gradients of the mean softmax
 * cross-entropy w.r.t. every parameter into the gradient block (same layout), loss into *host_loss. */
int fls_mlp_forward_backward(fls_mlp* m, const float* dev_features, const int32_t* dev_labels, uint32_t batch,
                             float* host_loss, void* stream);
int fls_mlp_get_grads(fls_mlp* m, float* host_grads);
/* One Adam step with the current gradients (beta1 0.9, beta2 0.999, eps 1e-8). */
int fls_mlp_adam_step(fls_mlp* m, float lr, void* stream);
/* Batched inference: dev_choice[i] = argmax over the K logits of row i (any n; processed in max_batch chunks). */
int fls_mlp_predict(fls_mlp* m, const float* dev_features, uint64_t n, int32_t* dev_choice, void* stream);
/* Logits of the last forward pass (host copy, batch x 32 float32). */
int fls_mlp_get_logits(fls_mlp* m, float* host_logits, uint32_t batch);

/* ---- Structural wizard (reference Wizard::Spell, src/wizard/wizard.cpp:19-44) -------------------------------- */
/* Plan of each column of a row-group, in the reference's order: CONSTANT (all values equal), then EQUAL (same type
 * and values as an earlier column), then one-to-one maps (a first column with at most 25% distinct values whose
 * every value always meets the same value of a later column: the first becomes DICT + SUB_EXPRESSION, the second
 * FIXME; a first may take several seconds), and EXPRESSION (ZERO_COPY) for the rest. Values compare as the
 * reference's types do: integers bit for bit, doubles with == (-0.0 equals 0.0, a NaN equals nothing), strings byte
 * for byte. */
enum fls_spell_kind {
    FLS_SPELL_EXPRESSION = 0, /* ZERO_COPY: encoded by an expression of its own */
    FLS_SPELL_CONSTANT   = 1, /* CONSTANT */
    FLS_SPELL_EQUAL      = 2, /* EQUAL; ref = the first earlier column with the same values */
    FLS_SPELL_MAP_SOURCE = 3, /* DICT + SUB_EXPRESSION (uint16), once per target; ref = its first target */
    FLS_SPELL_MAP_TARGET = 4  /* FIXME; ref = its source */
};
struct fls_spell_column {
    uint8_t         type;    /* enum data_t of fastlanes.h (DOUBLE, INT8..UINT64, STR) */
    const void*     values;  /* device: n_rows values, or the string bytes */
    const uint64_t* offsets; /* device, strings only: n_rows + 1 byte offsets into values */
};
/* Plans n_columns device-resident columns of n_rows rows each (n_rows <= 2^30). out_kind / out_ref are host
 * arrays of n_columns (ref -1 where the kind has none). Runs on `stream` and returns when the plan is complete;
 * allocates its hash tables stream-ordered and frees them before returning. */
int fls_wizard_spell(uint32_t n_columns, const struct fls_spell_column* columns, uint64_t n_rows, int32_t* out_kind,
                     int32_t* out_ref, void* stream);

#ifdef __cplusplus
}
#endif

#endif
