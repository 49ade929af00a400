// reader.cpp -- fastlanes::Connection / Reader / Chunk (include/fls_reader.hpp): a row-group directory read from
// disk, uploaded once and decoded on the GPU by the row-group launch (fls_gpu_rowgroup_*, captured as a HIP graph).
//
// Reference behaviour followed: Reader::Reader reads footer.json, then data.fls into one buffer, then sets up the
// columns (src/reader/reader.cpp:19-50); get_chunk / reset / to_csv (:52-61); Connection::read_fls keeps the
// Reader (src/connection.cpp:56-61). The footer's keys are those of src/json/fls_json.cpp:30-61.
#include "../../include/fls_reader.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <type_traits>

#include "../../include/fastlanes.h"
#include "../../include/fls_gpu.h"

namespace fastlanes {
namespace {

// ---- a small JSON reader (objects, arrays, strings, numbers, true/false/null) -------------------------------------
struct Json {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    bool                        b = false;
    std::string                 text;  // STR: the string; NUM: the number's digits as written
    std::vector<Json>           arr;
    std::map<std::string, Json> obj;

    [[nodiscard]] const Json* find(const std::string& key) const {
        auto it = obj.find(key);
        return kind == OBJ && it != obj.end() ? &it->second : nullptr;
    }
    [[nodiscard]] uint64_t u64(const char* what) const {
        if (kind != NUM || text.empty() || text[0] == '-' || text.find_first_of(".eE") != std::string::npos)
            throw std::runtime_error(std::string("footer: ") + what + " is not a non-negative integer");
        errno = 0;
        const unsigned long long v = std::strtoull(text.c_str(), nullptr, 10);
        if (errno) throw std::runtime_error(std::string("footer: ") + what + " out of range");
        return v;
    }
};

class JsonParser {
public:
    explicit JsonParser(const std::string& s) : s(s) {}
    Json parse() {
        Json v = value(0);
        ws();
        if (i != s.size()) fail("trailing characters");
        return v;
    }
private:
    const std::string& s;
    size_t             i = 0;

    [[noreturn]] void fail(const char* why) const {
        throw std::runtime_error("footer.json: " + std::string(why) + " at byte " + std::to_string(i));
    }
    void ws() {
        while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
    }
    bool eat(char c) {
        ws();
        if (i < s.size() && s[i] == c) return ++i, true;
        return false;
    }
    Json value(int depth) {
        if (depth > 64) fail("nesting too deep");
        ws();
        if (i >= s.size()) fail("unexpected end");
        Json v;
        const char c = s[i];
        if (c == '{') {
            ++i;
            v.kind = Json::OBJ;
            if (eat('}')) return v;
            do {
                ws();
                if (i >= s.size() || s[i] != '"') fail("object key expected");
                std::string k = str();
                if (!eat(':')) fail("':' expected");
                v.obj[k] = value(depth + 1);
            } while (eat(','));
            if (!eat('}')) fail("'}' expected");
        } else if (c == '[') {
            ++i;
            v.kind = Json::ARR;
            if (eat(']')) return v;
            do v.arr.push_back(value(depth + 1));
            while (eat(','));
            if (!eat(']')) fail("']' expected");
        } else if (c == '"') {
            v.kind = Json::STR;
            v.text = str();
        } else if (c == '-' || std::isdigit((unsigned char)c)) {
            v.kind        = Json::NUM;
            const size_t a = i;
            ++i;
            while (i < s.size() && (std::isdigit((unsigned char)s[i]) || s[i] == '.' || s[i] == 'e' || s[i] == 'E' ||
                                    s[i] == '+' || s[i] == '-'))
                ++i;
            v.text = s.substr(a, i - a);
        } else if (s.compare(i, 4, "true") == 0) {
            v.kind = Json::BOOL, v.b = true, i += 4;
        } else if (s.compare(i, 5, "false") == 0) {
            v.kind = Json::BOOL, i += 5;
        } else if (s.compare(i, 4, "null") == 0) {
            i += 4;
        } else {
            fail("unexpected character");
        }
        return v;
    }
    std::string str() {
        std::string out;
        ++i;  // opening quote
        while (i < s.size() && s[i] != '"') {
            char c = s[i++];
            if (c == '\\') {
                if (i >= s.size()) fail("bad escape");
                const char e = s[i++];
                switch (e) {
                case 'n': c = '\n'; break;
                case 't': c = '\t'; break;
                case 'r': c = '\r'; break;
                case 'b': c = '\b'; break;
                case 'f': c = '\f'; break;
                case 'u': {  // keep ASCII code points; others become '?'
                    if (i + 4 > s.size()) fail("bad \\u escape");
                    const unsigned cp = (unsigned)std::strtoul(s.substr(i, 4).c_str(), nullptr, 16);
                    i += 4;
                    c = cp < 128 ? (char)cp : '?';
                    break;
                }
                default: c = e;
                }
            }
            out.push_back(c);
        }
        if (i >= s.size()) fail("unterminated string");
        ++i;
        return out;
    }
};

std::string dump(const Json& v) {
    switch (v.kind) {
    case Json::NUL: return "null";
    case Json::BOOL: return v.b ? "true" : "false";
    case Json::NUM: return v.text;
    case Json::STR: {
        std::string o = "\"";
        for (char c : v.text) o += (c == '"' || c == '\\') ? std::string("\\") + c : std::string(1, c);
        return o + "\"";
    }
    case Json::ARR: {
        std::string o = "[";
        for (size_t k = 0; k < v.arr.size(); ++k) o += (k ? ", " : "") + dump(v.arr[k]);
        return o + "]";
    }
    default: {
        std::string o = "{";
        bool        first = true;
        for (const auto& [k, x] : v.obj) o += (first ? "\"" : ", \"") + k + "\": " + dump(x), first = false;
        return o + "}";
    }
    }
}

std::string read_file(const path& p) {
    std::ifstream f(p, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + p.string());
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
void fls_check(int rc, const char* what) {
    if (rc) throw std::runtime_error(std::string(what) + ": " + fls_gpu_error_string());
}

}  // namespace

uint64_t SizeOf(DataType type) {
    switch (type) {
    case DataType::INT8:
    case DataType::UINT8: return 1;
    case DataType::INT16:
    case DataType::UINT16: return 2;
    case DataType::INT32:
    case DataType::UINT32: return 4;
    case DataType::INT64:
    case DataType::UINT64:
    case DataType::DOUBLE: return 8;
    default: return 0;
    }
}

Footer Footer::read(const path& file) {
    const std::string text = read_file(file);
    const Json        j    = JsonParser(text).parse();
    const Json *      nv = j.find("1: N VEC"), *rs = j.find("2: Rowgroup size"), *cm = j.find("3: Column Metadatas");
    if (!nv || !rs || !cm || cm->kind != Json::ARR)
        throw std::runtime_error("footer: \"1: N VEC\", \"2: Rowgroup size\" and \"3: Column Metadatas\" are required");
    Footer f;
    f.n_vec         = nv->u64("\"1: N VEC\"");
    f.rowgroup_size = rs->u64("\"2: Rowgroup size\"");
    for (const Json& c : cm->arr) {
        ColumnMetadata m;
        const Json*    ofs = c.find("col_ofs");
        const Json*    lt  = c.find("logical_type");
        if (!ofs || !lt) throw std::runtime_error("footer: every column needs \"col_ofs\" and \"logical_type\"");
        m.offset    = ofs->u64("\"col_ofs\"");
        m.data_type = (DataType)lt->u64("\"logical_type\"");
        if (SizeOf(m.data_type) == 0) throw std::runtime_error("footer: column type is not an integer or double type");
        if (const Json* s = c.find("size")) m.size = s->u64("\"size\"");
        if (const Json* p = c.find("pool id")) m.pool_id = (int)p->u64("\"pool id\"");
        if (const Json* r = c.find("Encoding RPN")) m.rpn = dump(*r);
        f.column_metadatas.push_back(m);
    }
    // sizes left out run to the next column's offset (or the end of the row-group)
    for (size_t k = 0; k < f.column_metadatas.size(); ++k) {
        ColumnMetadata& m = f.column_metadatas[k];
        if (m.offset > f.rowgroup_size) throw std::runtime_error("footer: col_ofs past the row-group");
        if (!m.size) {
            uint64_t end = f.rowgroup_size;
            for (const ColumnMetadata& o : f.column_metadatas)
                if (o.offset > m.offset) end = std::min(end, o.offset);
            m.size = end - m.offset;
        }
        if (m.size > f.rowgroup_size - m.offset) throw std::runtime_error("footer: column extends past the row-group");
    }
    return f;
}

void Chunk::copy_to_host(size_t c, void* dst) const {
    hip_check(hipMemcpyAsync(dst, vecs.at(c), 1024 * SizeOf(types.at(c)), hipMemcpyDeviceToHost, (hipStream_t)stream),
              "copy_to_host");
    hip_check(hipStreamSynchronize((hipStream_t)stream), "copy_to_host");
}

struct ReaderState {
    Footer                footer;
    std::vector<uint8_t>  host;
    uint8_t*              d_enc = nullptr;
    uint8_t*              d_out = nullptr;
    fls_gpu_rowgroup*     rg    = nullptr;
    std::vector<uint64_t> out_off;
    hipStream_t           stream  = nullptr;
    bool                  decoded = false;
    Chunk                 chunk;
    ~ReaderState() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (rg) fls_gpu_rowgroup_destroy(rg);
        if (d_enc) (void)hipFree(d_enc);
        if (d_out) (void)hipFree(d_out);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

Reader::Reader(const path& dir_path, Connection& /*fls*/) : m(std::make_unique<ReaderState>()) {
    m->footer = Footer::read(dir_path / FOOTER_FILE_NAME);
    // data.fls into one host buffer, then once to the device
    {
        const std::string data = read_file(dir_path / FASTLANES_FILE_NAME);
        if (data.size() != m->footer.rowgroup_size)
            throw std::runtime_error("data.fls holds " + std::to_string(data.size()) + " bytes, the footer says " +
                                     std::to_string(m->footer.rowgroup_size));
        m->host.assign(data.begin(), data.end());
    }
    hip_check(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipMalloc(&m->d_enc, m->host.size() + 256), "hipMalloc");
    hip_check(hipMemcpy(m->d_enc, m->host.data(), m->host.size(), hipMemcpyHostToDevice), "hipMemcpy");
    // the columns
    const size_t                 n = m->footer.column_metadatas.size();
    std::vector<const uint8_t*>  cols(n);
    std::vector<uint64_t>        sizes(n), offs(n);
    std::vector<uint8_t>         types(n);
    for (size_t c = 0; c < n; ++c) {
        const ColumnMetadata& cm = m->footer.column_metadatas[c];
        cols[c]  = m->host.data() + cm.offset;
        sizes[c] = cm.size;
        offs[c]  = cm.offset;
        types[c] = (uint8_t)cm.data_type;
    }
    fls_check(fls_gpu_rowgroup_create((uint32_t)n, cols.data(), sizes.data(), offs.data(), types.data(), &m->rg),
              "row-group");
    uint64_t total = 0;
    m->out_off.resize(n);
    fls_check(fls_gpu_rowgroup_layout(m->rg, m->out_off.data(), &total), "row-group layout");
    hip_check(hipMalloc(&m->d_out, total + 256), "hipMalloc");
    for (size_t c = 0; c < n; ++c) m->chunk.types.push_back(m->footer.column_metadatas[c].data_type);
    m->chunk.vecs.assign(n, nullptr);
    m->chunk.stream = m->stream;
}

Reader::~Reader() = default;

const Footer& Reader::footer() const { return m->footer; }

void Reader::decode() {
    fls_check(fls_gpu_rowgroup_decode_graph(m->rg, m->d_enc, m->d_out, m->stream), "row-group decode");
    fls_check(fls_gpu_rowgroup_check(m->rg, m->stream), "row-group decode");
    m->decoded = true;
}

Chunk& Reader::get_chunk(n_t vec_idx) {
    if (vec_idx >= m->footer.n_vec)
        throw std::out_of_range("vector " + std::to_string(vec_idx) + " of " + std::to_string(m->footer.n_vec));
    if (!m->decoded) decode();
    m->chunk.m_vec_idx = vec_idx;
    for (size_t c = 0; c < m->chunk.vecs.size(); ++c)
        m->chunk.vecs[c] = m->d_out + m->out_off[c] + vec_idx * 1024 * SizeOf(m->chunk.types[c]);
    return m->chunk;
}

void Reader::reset() { m->decoded = false; }

namespace {
template <typename T>
void put_value(std::string& out, const uint8_t* p) {
    T x;
    std::memcpy(&x, p, sizeof(T));
    if constexpr (std::is_floating_point_v<T>) {
        std::ostringstream ss;
        ss.precision(17);
        ss << x;
        out += ss.str();
    } else {
        out += std::to_string(x);  // 8-bit values as numbers
    }
}
}  // namespace

void Reader::to_csv(const path& dir_path) {
    std::ofstream f(dir_path / CSV_FILE_NAME, std::ios::binary | std::ios::trunc);
    if (!f) throw std::runtime_error("cannot write " + (dir_path / CSV_FILE_NAME).string());
    const size_t                      n = m->chunk.vecs.size();
    std::vector<std::vector<uint8_t>> host(n);
    std::string                       out;
    for (n_t v = 0; v < m->footer.n_vec; ++v) {
        Chunk& ch = get_chunk(v);
        for (size_t c = 0; c < n; ++c) {
            host[c].resize(1024 * SizeOf(ch.type(c)));
            ch.copy_to_host(c, host[c].data());
        }
        out.clear();
        for (unsigned r = 0; r < 1024; ++r) {
            for (size_t c = 0; c < n; ++c) {
                const uint8_t* p = host[c].data() + r * SizeOf(ch.type(c));
                switch (ch.type(c)) {
                case DataType::INT8: put_value<int8_t>(out, p); break;
                case DataType::INT16: put_value<int16_t>(out, p); break;
                case DataType::INT32: put_value<int32_t>(out, p); break;
                case DataType::INT64: put_value<int64_t>(out, p); break;
                case DataType::UINT8: put_value<uint8_t>(out, p); break;
                case DataType::UINT16: put_value<uint16_t>(out, p); break;
                case DataType::UINT32: put_value<uint32_t>(out, p); break;
                case DataType::UINT64: put_value<uint64_t>(out, p); break;
                default: put_value<double>(out, p); break;
                }
                out += c + 1 < n ? '|' : '\n';
            }
        }
        f << out;
    }
    if (!f) throw std::runtime_error("write failed: " + (dir_path / CSV_FILE_NAME).string());
}

Connection::Connection()  = default;
Connection::~Connection() = default;

Reader& Connection::read_fls(const path& dir_path) {
    m_reader.reset();
    m_reader = std::make_unique<Reader>(dir_path, *this);
    return *m_reader;
}

Connection& Connection::reset() {
    m_reader.reset();
    return *this;
}

void Connection::encode_from_memory(void* input_p, n_t n_tup, n_t capacity, void* encoded_p,
                                    uint64_t* encoded_byte_size, const DataType& data_type, n_t expr_id) {
    uint64_t meta = 0;
    if (!::encode_from_memory(input_p, n_tup, capacity, (uint8_t*)encoded_p, encoded_byte_size, &meta,
                              (enum data_t)data_type, (uint8_t)expr_id))
        throw std::runtime_error(std::string("encode_from_memory: ") + fls_gpu_error_string());
}

void Connection::decode_to_memory(void* encoded_p, void* decoded_p, const DataType& data_type) {
    ::decode_to_memory((uint8_t*)encoded_p, decoded_p, (enum data_t)data_type);
}

}  // namespace fastlanes
