// fls_reader.hpp -- the C++ boundary of the decode path: fastlanes::Connection, Reader and Chunk over a row-group
// directory (data.fls + footer.json), decoded on the GPU by one row-group launch.
//
// Signatures follow the reference: Reader(dir_path, Connection&), get_chunk(vec_idx), reset(), to_csv(dir_path)
// (reference include/fls/reader/reader.hpp:17-35, src/reader/reader.cpp:19-61); Connection::read_fls, reset,
// encode_from_memory, decode_to_memory (reference include/fls/connection.hpp:33-92). The footer is the
// reference's JSON (reference src/json/fls_json.cpp:30-61, include/fls/footer/rowgroup_footer.hpp:23-75):
//   {"1: N VEC": n_vec, "2: Rowgroup size": bytes of data.fls,
//    "3: Column Metadatas": [{"logical_type": DataType, "logical_type_str": name, "col_ofs": byte offset,
//                             "Encoding RPN": {"operators": [...], "operands": [...]}, "is_determined": true,
//                             "size": encoded bytes (optional), "pool id": expression id (optional)}, ...]}
// Each column's bytes at col_ofs are its encoded pages (the format encode_from_memory writes); the pages carry
// their own expression, so "Encoding RPN" and "pool id" are descriptive. A column's size, when absent, runs to
// the next column's offset (or the end of data.fls).
//
// Differences from the reference: the whole row-group is decoded once, on the GPU, when the Reader is built, and
// a Chunk points at the decoded vector in device memory (copy_to_host gives the values); to_csv writes every
// vector of the row-group, not only the last.
//
// Errors are reported as std::runtime_error, as the reference does.
#pragma once

#include <stdint.h>

#include <filesystem>
#include <memory>
#include <string>
#include <vector>

namespace fastlanes {

using path = std::filesystem::path;
using n_t  = uint64_t;

// reference include/fls/expression/data_type.hpp:9-29 (the same numbering as include/fastlanes.h data_t)
enum class DataType : uint8_t {
    INVALID = 0,
    DOUBLE  = 1,
    INT8    = 2,
    INT16   = 3,
    INT32   = 4,
    INT64   = 5,
    UINT8   = 6,
    UINT16  = 7,
    UINT32  = 8,
    UINT64  = 9,
    STR     = 10,
};
// bytes of one value (0 for types the reader does not decode)
uint64_t SizeOf(DataType type);

constexpr auto const* FOOTER_FILE_NAME    = "footer.json";
constexpr auto const* FASTLANES_FILE_NAME = "data.fls";
constexpr auto const* CSV_FILE_NAME       = "extracted_from_fls.csv";

struct ColumnMetadata {
    DataType    data_type = DataType::INVALID;
    uint64_t    offset    = 0;  // "col_ofs"
    uint64_t    size      = 0;  // "size", or up to the next column
    int         pool_id   = -1; // "pool id" when present
    std::string rpn;            // "Encoding RPN" as written (JSON text)
};

struct Footer {
    n_t                         n_vec         = 0;  // "1: N VEC"
    uint64_t                    rowgroup_size = 0;  // "2: Rowgroup size"
    std::vector<ColumnMetadata> column_metadatas;   // "3: Column Metadatas"
    static Footer               read(const path& file);
};

class Connection;
struct ReaderState;

// One vector of every column: column c's 1024 decoded values in device memory.
class Chunk {
public:
    [[nodiscard]] size_t      size() const { return vecs.size(); }
    [[nodiscard]] n_t         vec_idx() const { return m_vec_idx; }
    [[nodiscard]] DataType    type(size_t c) const { return types.at(c); }
    [[nodiscard]] const void* device_data(size_t c) const { return vecs.at(c); }
    // copies column c's 1024 values (1024 * SizeOf(type(c)) bytes) to host memory at dst
    void copy_to_host(size_t c, void* dst) const;

private:
    friend class Reader;
    n_t                      m_vec_idx = 0;
    std::vector<const void*> vecs;
    std::vector<DataType>    types;
    void*                    stream = nullptr;
};

class Reader {
public:
    explicit Reader(const path& dir_path, Connection& fls);
    ~Reader();
    Reader(const Reader&)            = delete;
    Reader& operator=(const Reader&) = delete;

    ///! the chunk of vector vec_idx (decodes the row-group first if reset() was called)
    Chunk& get_chunk(n_t vec_idx);
    ///! drops the decoded row-group; the next get_chunk decodes it again
    void reset();
    ///! writes every vector, one row per value and '|' between columns, to dir_path/extracted_from_fls.csv
    void to_csv(const path& dir_path);

    [[nodiscard]] const Footer& footer() const;

private:
    std::unique_ptr<ReaderState> m;
    void                         decode();
};

class Connection {
public:
    Connection();
    ~Connection();
    ///! read a row-group directory; the Reader lives until the next read_fls or reset
    Reader& read_fls(const path& dir_path);
    ///!
    Connection& reset();

    static void encode_from_memory(void* input_p, n_t n_tup, n_t capacity, void* encoded_p, uint64_t* encoded_byte_size,
                                   const DataType& data_type, n_t expr_id);
    static void decode_to_memory(void* encoded_p, void* decoded_p, const DataType& data_type);

private:
    std::unique_ptr<Reader> m_reader;
};

}  // namespace fastlanes
