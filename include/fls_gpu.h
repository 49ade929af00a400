/*
 * fls_gpu.h -- device-resident entry points of the FastLanes decoder for MI355X (gfx950).
 *
 * Plain C ABI: pointers, sizes and an opaque plan handle. `stream` is a hipStream_t passed as void*
 * (NULL = the default stream). Device pointers are HIP device memory (e.g. hipMalloc or a torch CUDA tensor).
 * Decoded output buffers must be 16-byte aligned. All functions return 0 on success, a negative FLS_E_* code
 * otherwise; fls_gpu_error_string() describes the last failure of the calling thread.
 *
 * The reference has no device boundary; each function states which reference interface it stands in for.
 */
#ifndef FLS_GPU_H
#define FLS_GPU_H

#include <stdint.h>

#include "fastlanes.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
    FLS_OK            = 0,
    FLS_E_ARG         = -1, /* bad argument (null pointer, misaligned output, unknown type) */
    FLS_E_FORMAT      = -2, /* encoded bytes fail validation */
    FLS_E_UNSUPPORTED = -3, /* expression or type not supported */
    FLS_E_CAPACITY    = -4, /* output or encode capacity too small */
    FLS_E_HIP         = -5, /* HIP runtime error */
    FLS_E_DATA        = -6  /* decoded data referenced an out-of-range run or dictionary entry */
};

const char* fls_gpu_error_string(void);

/* ---- column plans ------------------------------------------------------------------------------------------------
 * A plan is the host-side parse of one encoded column (all page entries): it stands in for the per-call parsing of
 * Decompressor::Init / InitDict (reference src/cor/eng/decompressor.cpp:21-94) and ExpColDecoder::full_decode
 * (src/encoder/exp_col_decoder.cpp:13-26), done once. host_encoded must hold the whole column; it is only read
 * during fls_gpu_plan_create. Decoding then reads the same bytes from device memory. */
typedef struct fls_gpu_plan fls_gpu_plan;

int      fls_gpu_plan_create(const uint8_t* host_encoded, uint64_t encoded_size, enum data_t type,
                             fls_gpu_plan** out_plan);
void     fls_gpu_plan_destroy(fls_gpu_plan* plan);
uint64_t fls_gpu_plan_n_values(const fls_gpu_plan* plan);  /* values (or strings) in the column */
uint64_t fls_gpu_plan_encoded_size(const fls_gpu_plan* plan);
uint32_t fls_gpu_plan_n_pages(const fls_gpu_plan* plan);

/* Decode the whole integer column to dev_out (n_values * sizeof(T) bytes, original value order).
 * Replaces decode_to_memory (reference include/fastlanes.h:84-86) for device-resident data. Asynchronous. */
int fls_gpu_plan_decode(fls_gpu_plan* plan, const uint8_t* dev_encoded, void* dev_out, uint64_t out_capacity,
                        void* stream);
/* Synchronizes the stream and reports FLS_E_DATA if a decode on this plan met an out-of-range position/code. */
int fls_gpu_plan_check(fls_gpu_plan* plan, void* stream);

/* Strings (FSST pages, reference include/fls/cor/exp/fsst/fsst_exp.hpp:35-52 decoded as
 * include/fls/encoder/assert_eq.hpp:166-210 does). Step 1 writes n+1 uint64 output offsets (Arrow layout);
 * dev_offsets[n] is the total decoded size. Step 2 writes the bytes. Both asynchronous. */
int fls_gpu_plan_string_offsets(fls_gpu_plan* plan, const uint8_t* dev_encoded, uint64_t* dev_offsets,
                                void* stream);
int fls_gpu_plan_decode_strings(fls_gpu_plan* plan, const uint8_t* dev_encoded, const uint64_t* dev_offsets,
                                uint8_t* dev_bytes, uint64_t bytes_capacity, void* stream);
/* Steps 1 and 2 in one pass over the compressed bytes: writes the n+1 offsets and the bytes (bytes_capacity bounds
 * the bytes; a column that needs more reports FLS_E_DATA at fls_gpu_plan_check). No reference counterpart: the
 * reference decodes string by string on the CPU. Calls on one plan must not overlap (they share its segment status
 * words). Kernel launches only: the call may be captured into a HIP graph; each replay takes its ticket base and
 * epoch on the device. */
int fls_gpu_plan_decode_strings_fused(fls_gpu_plan* plan, const uint8_t* dev_encoded, uint64_t* dev_offsets,
                                      uint8_t* dev_bytes, uint64_t bytes_capacity, void* stream);

/* ---- predicate pushdown (reference src/primitive/predicate/equal.cpp, include/fls/expression/predicate_operator.hpp,
 * include/fls/expression/selection_ds.hpp) ----------------------------------------------------------------------
 * fls_gpu_plan_select decodes every vector and compares its values with `constant` in the same launch, writing no
 * decoded values: per vector, 16 uint64 bitmap words (bit i of word j set when value 64j + i satisfies
 * `value OP constant`; the SelectionDS bitmap) and the selected count (n_selected_values). Values compare in the
 * column's type: signed for INT*, unsigned for UINT*, IEEE for DOUBLE; `constant` holds the value's bit pattern in
 * its low W bits. dev_bitmap: 16 * n_vec words; dev_counts: n_vec. Not for string columns. */
enum fls_predicate_op { FLS_PRED_EQ = 0, FLS_PRED_GT = 1, FLS_PRED_LT = 2, FLS_PRED_GE = 3, FLS_PRED_LE = 4, FLS_PRED_NE = 5 };
int fls_gpu_plan_select(fls_gpu_plan* plan, const uint8_t* dev_encoded, int op, uint64_t constant,
                        uint64_t* dev_bitmap, uint32_t* dev_counts, void* stream);
/* Index-array form of a selection (SelectionDS idx_arr): vector v's selected positions, ascending, at
 * dev_indices[v * 1024 ...]; entries past the count are unspecified. dev_counts may be NULL. */
int fls_gpu_selection_indices(const uint64_t* dev_bitmap, uint64_t n_vec, uint16_t* dev_indices, uint32_t* dev_counts,
                              void* stream);
/* AND of two selections (and_selection_ds_func, src/primitive/predicate/equal.cpp:96-118); dev_counts may be NULL. */
int fls_gpu_selection_and(const uint64_t* dev_a, const uint64_t* dev_b, uint64_t n_vec, uint64_t* dev_out,
                          uint32_t* dev_counts, void* stream);

/* GPU encoder (reference encode RPNs of include/fls/cor/exp/{uncompressed,ffor,delta}/ and the page writer
 * of src/cor/eng/compressor.cpp:41-86): encodes n device-resident values (n a multiple of 1024) of `type` with
 * pool expression 0 (UNCOMPRESSED), 1 (FFOR), 11 (DELTA), 6 or 7 (RLE) into dev_out, the same bytes encode_from_memory writes
 * for those values. *out_size receives the encoded size (also when it exceeds `capacity`, which returns
 * FLS_E_CAPACITY and writes nothing). Reads the per-vector bit widths back to lay out the pages, so it returns
 * after `stream` has completed. */
int fls_gpu_encode(const void* dev_values, uint64_t n, enum data_t type, uint8_t expr_id, uint8_t* dev_out,
                   uint64_t capacity, uint64_t* out_size, void* stream);
/* fls_gpu_encode without device allocation. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The encoder's scratch (per-vector bit widths, bases, buffer offsets and the page table) is a fixed device array
 * of about 1.1 MiB per GPU, kept until the process ends, instead of stream allocations; n is at most 2^24 values
 * (FLS_E_CAPACITY beyond). Every call on one GPU shares that GPU's array: calls that overlap on one GPU race.
 * Same bytes, errors and synchronization as fls_gpu_encode. */
int fls_gpu_encode_noalloc(const void* dev_values, uint64_t n, enum data_t type, uint8_t expr_id, uint8_t* dev_out,
                           uint64_t capacity, uint64_t* out_size, void* stream);

/* Encode a string column with FSST (1024-string vectors; n must be a multiple of 1024): strings i are
 * bytes[offsets[i], offsets[i+1]). Writes *out_size bytes. No counterpart in the reference's C API, which accepts
 * INT64 and DOUBLE only (reference src/encoder/single_col_encoder.cpp:13-33); the layout is the reference's FSST
 * expression (include/fls/cor/exp/fsst/fsst_exp.hpp:35-52). Host-side, no GPU involved. */
int fls_encode_strings(const uint32_t* offsets, const uint8_t* bytes, uint64_t n, uint8_t* out, uint64_t capacity,
                       uint64_t* out_size);

/* ---- raw primitive streams -----------------------------------------------------------------------------------------
 * n_vec packed vectors back to back (128*bw bytes each), decoded into n_vec*1024 values of `width` bits in the
 * lane-major order of the reference's generated unffor/unpack (primitives/fls_generated/fallback/scalar_aav_uf1/
 * fallback_scalar_aav_1024_uf1_unffor_src.cpp:22879-23077, ..._unpack_src.cpp:17191-17380). dev_bases holds one
 * base per vector (FFOR) or is NULL (plain unpack). */
int fls_gpu_unffor_stream(unsigned width, unsigned bw, const void* dev_packed, const void* dev_bases, void* dev_out,
                          uint64_t n_vec, void* stream);

/* ---- launch-shape tuning ------------------------------------------------------------------------------------------
 * The decoder reads a table of launch shapes (threads per block, non-temporal stores, grid cap) from tuning.txt
 * next to the library (or the file named by FLS_TUNING) on first use; fastlanes_amd/autotune.py measures and writes
 * it. family: 0 raw streams, 1 UNCOMPRESSED, 2 FFOR, 3 DELTA, 4 RLE, 5 dictionary pages; bw 255 = all bit widths.
 * nontemporal: bit 0 = non-temporal stores; bit 1 = direct stores of each lane's segment instead of staging the
 * output through LDS (families 1-3 with 32/64-bit values, families 4-5 with 64-bit values); bit 2 (family 4, 8/16-bit
 * values) = decode position lanes 2 per thread with staging instead of one per thread with direct stores.
 * family 6 is the row-group launch: block = KiB of decoded values per tile, max_blocks = blocks per compute unit
 * (0: as many as stay resident at once, each wave then owning one long run of the work list).
 * fls_gpu_tune_get reads the shape in effect for (family, width, bw); any output pointer may be NULL. */
int fls_gpu_tune_set(int family, unsigned width, unsigned bw, unsigned block, unsigned nontemporal,
                     unsigned max_blocks);
int fls_gpu_tune_get(int family, unsigned width, unsigned bw, unsigned* block, unsigned* nontemporal,
                     unsigned* max_blocks);
int fls_gpu_tune_load(const char* path);

/* ---- row groups --------------------------------------------------------------------------------------------------
 * Many columns (each an encoded column resident in device memory at its own offset of one buffer) decoded by one
 * persistent launch over a device-side work list of (column, vector range) tiles. Stands in for the per-column
 * decode loop behind Reader::get_chunk / Interpreter::Decoding (reference src/reader/reader.cpp:19-61,
 * src/expression/interpreter.cpp:105-137). */
typedef struct fls_gpu_rowgroup fls_gpu_rowgroup;

/* types[c] is the column's enum data_t value stored in one byte (the enum's size differs between C and C++). */
int  fls_gpu_rowgroup_create(uint32_t n_columns, const uint8_t* const* host_columns, const uint64_t* column_sizes,
                             const uint64_t* device_offsets, const uint8_t* types, fls_gpu_rowgroup** out);
void fls_gpu_rowgroup_destroy(fls_gpu_rowgroup* rg);
/* Output offsets (bytes, 256-aligned) of each column's values in the row-group's output buffer, and its size.
 * Each column occupies n_values rounded up to whole 1024-value vectors. */
int  fls_gpu_rowgroup_layout(const fls_gpu_rowgroup* rg, uint64_t* out_offsets, uint64_t* out_total_bytes);
/* mode 0: one persistent launch over the work list (dictionaries of the fused pages are decoded first; pages
 * outside the persistent kernel's set -- patched FFOR, RLE with plain positions, DICT_RLE, S_PATCH_REC -- decode
 * with their per-page kernels before it); mode 1: one launch per column page, in column order on `stream` (the
 * per-column form); mode 2: the same launches spread over parallel branches (4 streams forked from and joined back
 * into `stream`). dev_out must be 256-byte aligned and hold out_total_bytes. */
int  fls_gpu_rowgroup_decode(fls_gpu_rowgroup* rg, const uint8_t* dev_encoded, uint8_t* dev_out, int mode,
                             void* stream);
/* Captures mode 0 once in a HIP graph and replays it (one graph launch per call; recaptured when the buffers
 * change). */
int  fls_gpu_rowgroup_decode_graph(fls_gpu_rowgroup* rg, const uint8_t* dev_encoded, uint8_t* dev_out, void* stream);
/* The persistent launch's work list: tile count, then up to `capacity` tiles of 128 bytes (the layout of RgTile in
 * fastlanes_amd/csrc/rowgroup.hpp) in launch order, the launch shape (blocks, waves per block, LDS bytes of staged
 * ranges per wave) and each wave's run of the list (blocks * waves + 1 bounds: wave w decodes tiles
 * [wave_begin[w], wave_begin[w + 1])). Any pointer may be NULL. */
int  fls_gpu_rowgroup_work_list(const fls_gpu_rowgroup* rg, uint32_t* n_tiles, void* tiles, uint32_t capacity,
                                uint32_t* blocks, uint32_t* waves, uint32_t* wave_lds, uint32_t* wave_begin);
/* The same work list planned on the host alone (no device memory is touched; dictionary tiles carry a nonzero
 * placeholder instead of a dictionary address), with the same validation, for inspecting a row-group's tiles.
 * waves > 0 also assigns the tiles to that many waves (tiles in launch order, waves + 1 bounds in wave_begin). */
int  fls_rowgroup_plan_work_list(uint32_t n_columns, const uint8_t* const* host_columns, const uint64_t* column_sizes,
                                 const uint64_t* device_offsets, const uint8_t* types, uint32_t waves,
                                 uint32_t* n_tiles, void* tiles, uint32_t capacity, uint32_t* wave_begin,
                                 uint32_t* wave_lds);
int  fls_gpu_rowgroup_check(fls_gpu_rowgroup* rg, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* FLS_GPU_H */
