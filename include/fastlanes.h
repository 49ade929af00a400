/*
 * fastlanes.h -- C API of the FastLanes decoder for MI355X (gfx950).
 *
 * Drop-in for the reference's include/fastlanes.h: same enums, same functions, same argument meaning.
 * Every entry point below replaces the reference declaration cited next to it. Differences in behaviour,
 * all in the direction of reporting errors the reference drops silently:
 *   - encode_from_memory returns false (instead of throwing or silently truncating) when the encoded column
 *     does not fit the capacity, when in_n_input is not a multiple of 1024, or for an expression/type the
 *     library does not encode; out_metadadata_size is written (the reference leaves it untouched).
 *   - decode_to_memory decodes on the GPU. Pointers may be host or device memory; host buffers are staged
 *     through device memory. Columns longer than 65535 vectors span several pages and decode in one call.
 *   - init_random_access / random_access are implemented (the reference declares them only).
 */
#ifndef FASTLANES_H
#define FASTLANES_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* data type; values as reference include/fastlanes.h:14-34 (mirrors fastlanes::DataType) */
enum data_t
#ifdef __cplusplus
    : uint8_t
#endif
{
    INVALID    = 0,
    DOUBLE     = 1,
    INT8       = 2,
    INT16      = 3,
    INT32      = 4,
    INT64      = 5,
    UINT8      = 6,
    UINT16     = 7,
    UINT32     = 8,
    UINT64     = 9,
    STR        = 10,
    BOOLEAN    = 11,
    DATE       = 12,
    FLOAT      = 13,
    BYTE_ARRAY = 14,
    LIST       = 15,
    STRUCT     = 16,
    MAP        = 17,
    FALLBACK   = 18,
};

/* i64 expression ids; reference include/fastlanes.h:37-41 (ids index src/cor/exp_pool.cpp:23-65). The same ids
 * select the same expressions for every integer type here, and the library accepts every id 0-38 of that pool:
 * FFOR patches 2-5, RLE 6-10, DELTA 11-15, DICT 16-26, DICT_RLE 27, DICT_FRQ 28-38. Patched FFOR and DELTA (2-5,
 * 12-15) need 32/64-bit values; the forms that store a buffer as one HZL chain of u64 values per page (3, 8-10,
 * 15) need 64-bit values; dictionary forms take every integer width. init_random_access reports no buffers for a
 * page with such a chain (3, 8-10, 15, 21, 26, 27, 33, 38): no vector owns a slice of it. */
enum i64_expr_encoding_t
#ifdef __cplusplus
    : uint8_t
#endif
{
    I64_UNCOMPRESSED   = 0,
    I64_FFOR_NO_PATCH  = 1,
    I64_DELTA_NO_PATCH = 11,
};

/* double expression ids; reference include/fastlanes.h:44-47. DBL_ALP is not supported (ALP lives in an external,
 * unpinned dependency of the reference). encode_from_memory also takes the double pool's RLE ids 6-10
 * (src/cor/exp_pool.cpp:105-112): runs split where double != holds, run values kept as 8-byte patterns. */
enum dbl_expr_encoding_t
#ifdef __cplusplus
    : uint8_t
#endif
{
    DBL_UNCOMPRESSED = 0,
    DBL_ALP          = 1,
};

typedef struct Connection connection_cpi;

/* reference include/fastlanes.h:53 (src/c_api/c_api_connector.cpp:6-9) */
connection_cpi* fls_cpi_connection_open(void);
/* reference include/fastlanes.h:54 (src/c_api/c_api_connector.cpp:11-13) */
void fls_cpi_connection_close(connection_cpi* fls_connection);

/* reference include/fastlanes.h:72-79 (src/c_api/c_api_connector.cpp:15-33) */
bool encode_from_memory(void* in_data, uint64_t in_n_input, uint64_t in_capacity, uint8_t* out_encoded_data,
                        uint64_t* out_encoded_size, uint64_t* out_metadadata_size, enum data_t in_datatype,
                        uint8_t in_expression_type);

/* reference include/fastlanes.h:84-86 (src/c_api/c_api_connector.cpp:35-37) */
void decode_to_memory(uint8_t* in_encoded_data, void* out_data, enum data_t in_datatype);

/* reference include/fastlanes.h:92-96 (declared only). For vector in_vector_id of the first page, writes
 * *output_n_buffer = number of page buffers and output_buffer[2*i], [2*i+1] = (offset from in_encoded_data, size)
 * of the vector's slice of buffer i. */
void init_random_access(uint8_t* in_encoded_data, uint64_t in_vector_id, uint64_t* output_n_buffer,
                        uint64_t* output_buffer);

/* reference include/fastlanes.h:101-107 (declared only). Decodes vector in_vector_id from the buffers the caller
 * loaded (output_buffers[i] holds the bytes named by pair i of init_random_access) into out_deencoded_vector
 * (1024 values, original order). in_encoded_data must hold at least the metadata (header, offsets, entry points
 * and, for dictionary expressions, the dictionary page). */
void random_access(uint8_t* in_encoded_data, uint64_t in_vector_id, uint64_t* output_n_buffer,
                   uint64_t* output_pair_offset_and_size, uint8_t** output_buffers, uint8_t* out_deencoded_vector);

#ifdef __cplusplus
}
#endif

#endif /* FASTLANES_H */
