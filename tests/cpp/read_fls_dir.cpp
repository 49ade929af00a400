// read_fls_dir -- C++ test program of the row-group reader (include/fls_reader.hpp), run by tests/test_gpu_reader.py.
//
//   read_fls_dir DIR EXPECTED
//
// Reads the row-group directory DIR through fastlanes::Connection::read_fls and compares every vector of every
// column, fetched with Reader::get_chunk, with EXPECTED: the columns' values back to back in footer order (n_vec *
// 1024 values each). Then checks reset() (the next get_chunk decodes again), that get_chunk(n_vec) throws, and
// writes DIR/extracted_from_fls.csv with Reader::to_csv. Prints "ok <n_vec> <columns>" and exits 0 when everything
// matches; prints the first difference and exits 1 otherwise.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <stdexcept>
#include <vector>

#include "../../include/fls_reader.hpp"

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s DIR EXPECTED\n", argv[0]);
        return 2;
    }
    try {
        std::ifstream        f(argv[2], std::ios::binary);
        std::vector<uint8_t> expected((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
        fastlanes::Connection conn;
        fastlanes::Reader&    reader = conn.read_fls(argv[1]);
        const auto&           footer = reader.footer();
        const size_t          n_col  = footer.column_metadatas.size();
        std::vector<uint64_t> base(n_col);
        uint64_t              total = 0;
        for (size_t c = 0; c < n_col; ++c) {
            base[c] = total;
            total += footer.n_vec * 1024 * fastlanes::SizeOf(footer.column_metadatas[c].data_type);
        }
        if (total != expected.size()) {
            std::printf("expected file holds %zu bytes, the footer's columns %llu\n", expected.size(),
                        (unsigned long long)total);
            return 1;
        }
        std::vector<uint8_t> host(8 * 1024);
        for (int pass = 0; pass < 2; ++pass) {
            for (fastlanes::n_t v = 0; v < footer.n_vec; ++v) {
                fastlanes::Chunk& chunk = reader.get_chunk(v);
                if (chunk.size() != n_col || chunk.vec_idx() != v) {
                    std::printf("chunk %llu: %zu columns\n", (unsigned long long)v, chunk.size());
                    return 1;
                }
                for (size_t c = 0; c < n_col; ++c) {
                    const uint64_t vb = 1024 * fastlanes::SizeOf(chunk.type(c));
                    chunk.copy_to_host(c, host.data());
                    if (std::memcmp(host.data(), expected.data() + base[c] + v * vb, vb) != 0) {
                        std::printf("pass %d: column %zu vector %llu differs\n", pass, c, (unsigned long long)v);
                        return 1;
                    }
                }
            }
            reader.reset();  // the second pass decodes the row-group again
        }
        bool threw = false;
        try {
            reader.get_chunk(footer.n_vec);
        } catch (const std::out_of_range&) {
            threw = true;
        }
        if (!threw) {
            std::printf("get_chunk(n_vec) did not throw\n");
            return 1;
        }
        reader.to_csv(argv[1]);
        std::printf("ok %llu %zu\n", (unsigned long long)footer.n_vec, n_col);
        return 0;
    } catch (const std::exception& e) {
        std::printf("error: %s\n", e.what());
        return 1;
    }
}
