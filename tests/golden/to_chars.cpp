// Writes each little-endian double read from stdin as C++17's std::to_chars(double)
// (shortest round-trip text, fixed or scientific), one per line, after its hex bits.
#include <charconv>
#include <cinttypes>
#include <cstdint>
#include <cstdio>
#include <cstring>

int main() {
    double v;
    char buf[64];
    while (std::fread(&v, sizeof v, 1, stdin) == 1) {
        std::uint64_t bits;
        std::memcpy(&bits, &v, sizeof bits);
        const auto r = std::to_chars(buf, buf + sizeof buf, v);
        std::printf("%016" PRIx64 " %.*s\n", bits, static_cast<int>(r.ptr - buf), buf);
    }
    return 0;
}
