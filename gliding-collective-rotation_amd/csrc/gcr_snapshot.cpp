// gcr_snapshot.cpp -- the text table of a .dat snapshot (output_methods.py:
// 420-485 writes it with numpy.savetxt), formatted on several host threads.
// The bytes are numpy.savetxt's for the same columns, formats and header:
// header lines prefixed "# ", one line per row, columns joined by one space,
// "%d" columns printed from their integral values, "%.8e" columns correctly
// rounded to nine significant digits (std::to_chars in scientific form, the
// digits and exponent layout of printf's %.8e and of Python's % operator),
// NaN without a sign as Python prints it.
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "gcr.h"

namespace {

void format_rows(std::string& out, int64_t r0, int64_t r1, int64_t n_rows, int32_t n_cols,
                 const double* cols, const int32_t* is_int)
{
    char buf[64];
    out.reserve((size_t)(r1 - r0) * (size_t)n_cols * 16);
    for (int64_t r = r0; r < r1; ++r) {
        for (int32_t k = 0; k < n_cols; ++k) {
            const double v = cols[(size_t)k * (size_t)n_rows + (size_t)r];
            char* end;
            if (is_int[k]) {
                end = std::to_chars(buf, buf + sizeof buf, (long long)v).ptr;
            } else if (std::isnan(v)) {
                std::memcpy(buf, "nan", 3);
                end = buf + 3;
            } else {  // correctly rounded, as printf's %.8e
                end = std::to_chars(buf, buf + sizeof buf, v, std::chars_format::scientific, 8).ptr;
            }
            if (k) out.push_back(' ');
            out.append(buf, (size_t)(end - buf));
        }
        out.push_back('\n');
    }
}

}  // namespace

// gcr_capi.hip: sets the message gcr_last_error returns
int gcr_internal_fail(int code, const char* msg);

extern "C" int gcr_write_table(const char* path, const char* header, int64_t n_rows, int32_t n_cols,
                               const double* cols, const int32_t* is_int, int32_t n_threads)
{
    if (!path || n_rows < 0 || n_cols < 1 || (n_rows > 0 && (!cols || !is_int)))
        return gcr_internal_fail(GCR_ERR_ARG, "gcr_write_table: bad arguments");
    int T = n_threads > 0 ? n_threads : 1;
    if ((int64_t)T > n_rows / 4096 + 1) T = (int)(n_rows / 4096 + 1);
    std::vector<std::string> parts((size_t)T);
    std::vector<std::thread> pool;
    for (int t = 0; t < T; ++t) {
        const int64_t r0 = n_rows * t / T, r1 = n_rows * (t + 1) / T;
        pool.emplace_back(format_rows, std::ref(parts[(size_t)t]), r0, r1, n_rows, n_cols, cols,
                          is_int);
    }
    for (auto& th : pool) th.join();
    FILE* f = std::fopen(path, "wb");
    if (!f) return gcr_internal_fail(GCR_ERR_ARG, (std::string("cannot open ") + path).c_str());
    bool ok = true;
    if (header && *header) {  // each header line behind "# "
        std::string h = "# ";
        for (const char* c = header; *c; ++c) {
            h.push_back(*c);
            if (*c == '\n') h += "# ";
        }
        h.push_back('\n');
        ok = std::fwrite(h.data(), 1, h.size(), f) == h.size();
    }
    for (const auto& p : parts)
        if (ok && !p.empty()) ok = std::fwrite(p.data(), 1, p.size(), f) == p.size();
    ok = (std::fclose(f) == 0) && ok;
    return ok ? GCR_OK : gcr_internal_fail(GCR_ERR_ARG, (std::string("cannot write ") + path).c_str());
}
