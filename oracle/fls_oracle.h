/*
 * fls_oracle.h -- CPU restatement of FastLanes' decode path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker (or as the
 * timed CPU baseline). The product decode path never calls into it.
 *
 * Every function restates the reference's algorithm from its closed form and
 * cites the reference file:line it follows. Pinning of each function
 * (golden vectors, compiled reference, invariants) is listed in the header
 * comment of oracle/fls_oracle.c.
 */
#ifndef FLS_ORACLE_H
#define FLS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- 1024-value vector primitives; W (element bits) in {8,16,32,64} ------------------------------------------ */
void    orc_transpose(unsigned W, const void* in, void* out);   /* transpose_i   */
void    orc_untranspose(unsigned W, const void* in, void* out); /* untranspose_i */
void    orc_pack(unsigned W, const void* in, void* out, unsigned bw);
void    orc_unpack(unsigned W, const void* in, void* out, unsigned bw);
void    orc_ffor(unsigned W, const void* in, void* out, unsigned bw, const void* base);
void    orc_unffor(unsigned W, const void* in, void* out, unsigned bw, const void* base);
void    orc_rsum(unsigned W, const void* in, void* out, const void* base);
void    orc_unrsum(unsigned W, const void* in, void* out);
uint8_t orc_ffor_analyze(unsigned W, const void* in, void* base_out);
uint8_t orc_ffor_bw_reference_formula(uint64_t delta);

/* Stream forms over n_vec back-to-back vectors (packed vector v starts at v*128*bw bytes). */
void orc_ffor_stream(unsigned W, const void* in, void* out, uint64_t n_vec, unsigned bw, const void* bases);
void orc_unffor_stream(unsigned W, const void* in, void* out, uint64_t n_vec, unsigned bw, const void* bases);

/* ---- page level ------------------------------------------------------------------------------------------------ */
/* data_t values follow include/fastlanes.h. Expression ids follow the reference's i64 pool
 * (src/cor/exp_pool.cpp:23-65): 0 UNCOMPRESSED, 1 FFOR, 6 RLE+DELTA, 7 RLE, 11 DELTA, 16 DICT, 17 DICT+FFOR,
 * 22 DICT+DELTA, 28 DICT_FRQ, 29 DICT_FRQ+FFOR; strings: 13 FSST. */
int64_t orc_encode_column(uint8_t dtype, uint8_t expr_id, const void* in, uint64_t n_vals, uint8_t* out,
                          uint64_t capacity);
/* Strings (expr 13 FSST, or 2/3 DICT over strings are not supported): values are given as
 * Arrow-style offsets (n+1 entries) + bytes. */
int64_t orc_encode_str_column(uint8_t expr_id, const uint32_t* offsets, const uint8_t* bytes, uint64_t n_vals,
                              uint8_t* out, uint64_t capacity);
/* Decode a column chunk produced by orc_encode_column. Returns number of values or <0 on error. */
int64_t orc_decode_column(uint8_t dtype, const uint8_t* in, uint64_t in_size, void* out, uint64_t out_capacity);
/* Decode an FSST string column: writes offsets (n+1) and bytes; returns n or <0. */
int64_t orc_decode_str_column(const uint8_t* in, uint64_t in_size, uint32_t* out_offsets, uint8_t* out_bytes,
                              uint64_t out_bytes_capacity, uint64_t* out_n_bytes);
/* Closed-form encoded size of a column chunk (bytes), without writing. */
int64_t orc_encoded_size(uint8_t dtype, uint8_t expr_id, const void* in, uint64_t n_vals);

/* ---- HZL (dictionary payload) ------------------------------------------------------------------------------------- */
int64_t orc_hzl_encode(unsigned W, const void* in, uint64_t n, uint8_t* out);
int64_t orc_hzl_decode(unsigned W, const uint8_t* in, void* out, uint64_t* consumed);

/* ---- FSST -------------------------------------------------------------------------------------------------------- */
/* Symbol table in the reference's decoder form (include/fls/cor/prm/fsst/fsst.h:92-98). */
typedef struct orc_fsst_decoder {
    unsigned long long version;
    unsigned char      zero_terminated;
    unsigned char      len[255];
    unsigned long long symbol[255];
} orc_fsst_decoder;
unsigned orc_fsst_import(orc_fsst_decoder* d, const uint8_t* buf);
size_t   orc_fsst_decompress(const orc_fsst_decoder* d, size_t len_in, const uint8_t* in, size_t size, uint8_t* out);
/* Builds a symbol table over a sample and writes the serialized table (fsst_export format) to buf.
 * Returns serialized length. */
unsigned orc_fsst_build(const uint32_t* offsets, const uint8_t* bytes, uint64_t n, uint8_t* buf);
/* Compress with a serialized table; returns compressed length of one string. */
size_t orc_fsst_compress_one(const uint8_t* table, const uint8_t* in, size_t len, uint8_t* out);

/* ---- timing helper for the CPU baseline: decode a chunk n_iter times with n_threads threads ----------------------- */
double orc_time_decode(uint8_t dtype, const uint8_t* in, uint64_t in_size, void* out, uint64_t out_capacity,
                       int n_iter, int n_threads);
double orc_time_unffor_stream(unsigned W, const void* in, void* out, uint64_t n_vec, unsigned bw,
                              const void* bases, int n_iter, int n_threads);

#ifdef __cplusplus
}
#endif
#endif
