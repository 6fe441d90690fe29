// The text body of an ESRI ASCII grid (host code): every value in the
// shortest decimal form that reads back to the same double, NaN as -9999,
// rows formatted in parallel on host threads.
#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/landschaft_circuit.h"
#include "common.hpp"

namespace lsc {
namespace {

// Appends row y (nx values, one space between them, a newline after) to out.
void format_row(const double* row, int64_t nx, std::string& out) {
    char buf[64];
    for (int64_t x = 0; x < nx; ++x) {
        const double v = std::isnan(row[x]) ? -9999.0 : row[x];
        const std::to_chars_result r = std::to_chars(buf, buf + sizeof buf, v);
        if (x) out.push_back(' ');
        out.append(buf, static_cast<std::size_t>(r.ptr - buf));
    }
    out.push_back('\n');
}

int host_threads() {
    const unsigned h = std::thread::hardware_concurrency();
    return static_cast<int>(std::clamp<unsigned>(h, 1u, 16u));
}

}  // namespace
}  // namespace lsc

extern "C" int lsc_ascii_grid_rows(const double* values, int64_t nx, int64_t ny, char* out, int64_t capacity,
                                   int64_t* size) {
    return lsc::guarded([&] {
        if (nx < 0 || ny < 0) throw std::invalid_argument("negative grid size");
        if (!size || (nx * ny > 0 && !values)) throw std::invalid_argument("null argument");
        // rows in contiguous blocks, one block of text per thread, joined in order
        const int nt = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(lsc::host_threads(), ny)));
        std::vector<std::string> text(static_cast<std::size_t>(nt));
        auto work = [&](int t) {
            const int64_t y0 = ny * t / nt, y1 = ny * (t + 1) / nt;
            std::string& s = text[static_cast<std::size_t>(t)];
            s.reserve(static_cast<std::size_t>((y1 - y0) * (nx * 12 + 1)));
            for (int64_t y = y0; y < y1; ++y) lsc::format_row(values + y * nx, nx, s);
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < nt; ++t) pool.emplace_back(work, t);
        work(0);
        for (auto& th : pool) th.join();
        int64_t total = 0;
        for (const auto& s : text) total += static_cast<int64_t>(s.size());
        *size = total;
        if (!out) return;
        if (capacity < total) throw std::invalid_argument("the text buffer is too small for the grid");
        char* p = out;
        for (const auto& s : text) {
            std::memcpy(p, s.data(), s.size());
            p += s.size();
        }
    });
}
