// io.hip -- text point files as the reference's sandbox reads them (host only).
//
// GPRT_CSV_POINTS follows load_points_fast (gprt-sandbox/src/io.rs:8-30): the
// file is split on '\n'; an empty line, a line that is not UTF-8, or one whose
// first three ','-separated fields (trimmed) do not all parse as numbers is
// skipped; later fields are ignored.  GPRT_CSV_BODIES follows load_galaxy
// (gprt-sandbox/src/bin/barneshut.rs:8-34): blank lines and lines with fewer
// than four fields are skipped, a field that does not parse reads as 0, and a
// body whose mass is not positive is skipped; a line that is not UTF-8 is an
// error there (the reference's reader panics).  Either way the rows keep the
// file's order, so row i is id i.  Numbers are decimal (or inf / infinity /
// nan, any case, with an optional sign), rounded once to binary32 by strtof.
// The file is mapped and parsed in slices by all host cores.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "gprt_internal.hpp"

namespace gprt {
namespace {

bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }

// Bytes of the White_Space character (Rust's char::is_whitespace) that starts
// at b, or that ends at e; 0 if there is none.  The text is valid UTF-8.
// White_Space: U+0009-000D, U+0020, U+0085, U+00A0, U+1680, U+2000-200A,
// U+2028, U+2029, U+202F, U+205F, U+3000.
int ws3(const unsigned char* p) {
    if (p[0] == 0xE1) return p[1] == 0x9A && p[2] == 0x80;
    if (p[0] == 0xE2 && p[1] == 0x80) return p[2] <= 0x8A || p[2] == 0xA8 || p[2] == 0xA9 || p[2] == 0xAF;
    if (p[0] == 0xE2) return p[1] == 0x81 && p[2] == 0x9F;
    return p[0] == 0xE3 && p[1] == 0x80 && p[2] == 0x80;
}
int ws_front(const char* b, const char* e) {
    const unsigned char* p = (const unsigned char*)b;
    if (b >= e) return 0;
    if (p[0] < 0x80) return is_space(b[0]) ? 1 : 0;
    if (e - b >= 2 && p[0] == 0xC2 && (p[1] == 0x85 || p[1] == 0xA0)) return 2;
    return (e - b >= 3 && ws3(p)) ? 3 : 0;
}
int ws_back(const char* b, const char* e) {
    if (b >= e) return 0;
    const unsigned char* q = (const unsigned char*)e;
    if (q[-1] < 0x80) return is_space(e[-1]) ? 1 : 0;
    if (e - b >= 2 && q[-2] == 0xC2 && (q[-1] == 0x85 || q[-1] == 0xA0)) return 2;
    return (e - b >= 3 && ws3(q - 3)) ? 3 : 0;
}
// str::trim: [b, e) without leading and trailing White_Space
void trim(const char*& b, const char*& e) {
    for (int n; (n = ws_front(b, e)) > 0;) b += n;
    for (int n; (n = ws_back(b, e)) > 0;) e -= n;
}

// true if [b, e) is well-formed UTF-8
bool utf8_ok(const unsigned char* b, const unsigned char* e) {
    while (b < e) {
        const unsigned c = *b;
        if (c < 0x80) { ++b; continue; }
        int extra;
        unsigned min;
        if ((c & 0xE0) == 0xC0) { extra = 1; min = 0x80; }
        else if ((c & 0xF0) == 0xE0) { extra = 2; min = 0x800; }
        else if ((c & 0xF8) == 0xF0) { extra = 3; min = 0x10000; }
        else return false;
        if (e - b <= extra) return false;
        unsigned v = c & (0x3F >> extra);
        for (int j = 1; j <= extra; ++j) {
            if ((b[j] & 0xC0) != 0x80) return false;
            v = (v << 6) | (b[j] & 0x3F);
        }
        if (v < min || v > 0x10FFFF || (v >= 0xD800 && v <= 0xDFFF)) return false;
        b += extra + 1;
    }
    return true;
}

// One field, trimmed as str::trim does, as a binary32; false if it is not a
// number in the forms Rust's f32 parser takes (no hex floats, no "nan(...)"
// payloads, any length).
bool parse_f32(const char* b, const char* e, float& out) {
    trim(b, e);
    const size_t n = (size_t)(e - b);
    if (n == 0) return false;
    char small[80];
    std::string big;
    char* buf = small;
    if (n >= sizeof small) {
        big.assign(b, n);
        buf = &big[0];
    } else {
        memcpy(small, b, n);
        small[n] = 0;
    }
    for (size_t j = 0; j < n; ++j) {
        const char c = buf[j];
        if (c == 'x' || c == 'X' || c == '(' || is_space(c)) return false;
    }
    char* end = nullptr;
    const float v = strtof(buf, &end);
    if (end != buf + n) return false;
    out = v;
    return true;
}

struct Slice {
    std::vector<float> rows;
    bool bad_utf8 = false;
};

void parse_slice(const char* b, const char* e, int format, Slice& out) {
    const int width = format == GPRT_CSV_BODIES ? 4 : 3;
    while (b < e) {
        const char* nl = (const char*)memchr(b, '\n', (size_t)(e - b));
        const char* le = nl ? nl : e;
        const char* line = b;
        b = nl ? nl + 1 : e;
        if (le == line) continue;
        if (!utf8_ok((const unsigned char*)line, (const unsigned char*)le)) {
            if (format == GPRT_CSV_BODIES) { out.bad_utf8 = true; return; }
            continue;
        }
        if (format == GPRT_CSV_BODIES) {
            const char* tb = line;
            const char* te = le;
            trim(tb, te);
            if (tb == te) continue;
        }
        // the first `width` fields
        const char* f[5];
        int nf = 0;
        f[nf++] = line;
        for (const char* p = line; p < le && nf < width + 1; ++p)
            if (*p == ',') f[nf++] = p + 1;
        if (nf < width) continue;
        float v[4];
        bool ok = true;
        for (int j = 0; j < width; ++j) {
            const char* fe = (j + 1 < nf) ? f[j + 1] - 1 : le;
            if (!parse_f32(f[j], fe, v[j])) {
                if (format == GPRT_CSV_BODIES) v[j] = 0.0f;
                else { ok = false; break; }
            }
        }
        if (!ok) continue;
        if (format == GPRT_CSV_BODIES && v[3] <= 0.0f) continue;
        out.rows.insert(out.rows.end(), v, v + width);
    }
}

}  // namespace
}  // namespace gprt

using namespace gprt;

extern "C" int gprt_hip_read_points(const char* path, int format, float* out, int64_t capacity, int64_t* rows) {
    if (!path || !rows || (capacity > 0 && !out)) return fail(GPRT_ERR_ARG, "null argument");
    if (format != GPRT_CSV_POINTS && format != GPRT_CSV_BODIES) return fail(GPRT_ERR_ARG, "unknown format");
    *rows = 0;
    const int fd = open(path, O_RDONLY);
    if (fd < 0) return fail(GPRT_ERR_ARG, std::string("cannot open ") + path);
    struct stat st;
    if (fstat(fd, &st) != 0) {
        close(fd);
        return fail(GPRT_ERR_ARG, std::string("cannot stat ") + path);
    }
    const size_t size = (size_t)st.st_size;
    if (size == 0) {
        close(fd);
        return GPRT_OK;
    }
    void* map = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
    close(fd);
    if (map == MAP_FAILED) return fail(GPRT_ERR_ARG, std::string("cannot map ") + path);
    const char* base = (const char*)map;
    const char* end = base + size;
    // slices start after a newline, so no line is split between two of them
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nslices = std::min<size_t>(std::min(hw, 16u), std::max<size_t>(1, size >> 20));
    std::vector<const char*> cut(nslices + 1, end);
    cut[0] = base;
    for (size_t s = 1; s < nslices; ++s) {
        const char* p = base + size * s / nslices;
        if (p < cut[s - 1]) p = cut[s - 1];
        const char* nl = (const char*)memchr(p, '\n', (size_t)(end - p));
        cut[s] = nl ? nl + 1 : end;
    }
    std::vector<Slice> parts(nslices);
    std::vector<std::thread> workers;
    for (size_t s = 1; s < nslices; ++s)
        workers.emplace_back([&, s] { parse_slice(cut[s], cut[s + 1], format, parts[s]); });
    parse_slice(cut[0], cut[1], format, parts[0]);
    for (auto& w : workers) w.join();
    munmap(map, size);
    const int width = format == GPRT_CSV_BODIES ? 4 : 3;
    int64_t total = 0;
    for (auto& p : parts) {
        if (p.bad_utf8) return fail(GPRT_ERR_INPUT, std::string("not UTF-8 text: ") + path);
        total += (int64_t)(p.rows.size() / width);
    }
    *rows = total;
    if (total > capacity)
        return fail(GPRT_ERR_ARG, "capacity " + std::to_string(capacity) + " rows < " + std::to_string(total) +
                                      " rows in " + path);
    float* o = out;
    for (auto& p : parts) {
        if (!p.rows.empty()) memcpy(o, p.rows.data(), p.rows.size() * sizeof(float));
        o += p.rows.size();
    }
    return GPRT_OK;
}
