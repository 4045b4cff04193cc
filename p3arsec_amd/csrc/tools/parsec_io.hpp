// parsec_io.hpp -- the PARSEC blackscholes text formats, read and written on
// several host threads.
//
// Input (blackscholes.c:701-734): "N", then N records
// "s strike r divq v t type divs DGrefval", read by the reference with
// fscanf("%i") and fscanf("%f %f %f %f %f %f %c %f %f").  Output
// (blackscholes.c:923-947): "N\n", then one "%.18f\n" per price.
//
// read_options() maps the file, splits the records after the count at
// line boundaries into one chunk per thread and converts each field as "%f"
// does (parse_float: the exact fast path for short decimals, else strtof;
// leading white space skipped) and each type with the first character after
// white space (" %c").  Records must not
// straddle a chunk boundary; if the chunks do not yield whole records (a
// record spread over several lines), the body is parsed again in one pass.
// Records after the first N are ignored, as the reference ignores them.
// write_prices() formats contiguous ranges on the threads with snprintf and
// writes the buffers in order, so the file is byte-identical to one fprintf
// per price.
#pragma once
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <string>
#include <thread>
#include <vector>

namespace parsec_io {

// The reference's OptionData (blackscholes.c:89-100): 36 bytes.
struct OptionData {
    float s, strike, r, divq, v, t;
    char OptionType;
    float divs, DGrefval;
};
static_assert(sizeof(OptionData) == 36, "OptionData layout");

inline bool is_space(char c) {
    return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f';
}

// One float field, as "%f" reads it (the correctly rounded float, as strtof
// gives).  A plain decimal [sign]digits[.digits] with at most 19 digits, k of
// them after the point, is converted without strtof when that is exact:
//   * m < 2^24 and k <= 10: m / 10^k in float arithmetic, two exact operands
//     and one rounding (Clinger's fast path, "How to Read Floating Point
//     Numbers Accurately", PLDI 1990);
//   * otherwise, k <= 22: d = m / 10^k in double, within 1.5 double ulps of
//     the decimal (10^k is exact; m rounds once past 2^53), then rounded to
//     float; unless the 29 bits of d below float precision lie within 8 ulps
//     of the halfway point, both roundings land on the same float.
// Everything else (longer mantissas, exponents, hex, inf, nan, subnormal or
// overflowing results, near-halfway cases) goes to strtof.
inline float parse_float(const char* q, const char** end) {
    static const float kPow10f[11] = {1e0f, 1e1f, 1e2f, 1e3f, 1e4f, 1e5f,
                                      1e6f, 1e7f, 1e8f, 1e9f, 1e10f};
    static const double kPow10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,
                                      1e8,  1e9,  1e10, 1e11, 1e12, 1e13, 1e14, 1e15,
                                      1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
    const char* start = q;
    while (is_space(*q)) ++q;
    const bool neg = *q == '-';
    if (*q == '-' || *q == '+') ++q;
    uint64_t m = 0;
    int digits = 0, frac = 0;
    bool point = false;
    for (;; ++q) {
        if (*q >= '0' && *q <= '9') {
            if (digits < 19) m = m * 10u + (uint64_t)(*q - '0');
            ++digits;
            if (point) ++frac;
        } else if (*q == '.' && !point) {
            point = true;
        } else {
            break;
        }
    }
    const char c = *q;
    const bool alnum = (c >= '0' && c <= '9') || (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z');
    if (digits > 0 && digits <= 19 && !alnum && c != '.') {
        if (m < (uint64_t(1) << 24) && frac <= 10) {
            const float v = (float)m / kPow10f[frac];
            *end = q;
            return neg ? -v : v;
        }
        if (frac <= 22) {
            const double d = (double)m / kPow10[frac];
            uint64_t bits;
            memcpy(&bits, &d, 8);
            const uint64_t low = bits & ((uint64_t(1) << 29) - 1);
            const uint64_t half = uint64_t(1) << 28;
            const uint64_t gap = low > half ? low - half : half - low;
            if (d >= 0x1p-126 && d < 0x1p127 && gap > 8) {
                const float v = (float)d;
                *end = q;
                return neg ? -v : v;
            }
            if (d == 0.0) {   // m == 0
                *end = q;
                return neg ? -0.0f : 0.0f;
            }
        }
    }
    char* e;
    const float v = strtof(start, &e);
    *end = e;
    return v;
}

// Parses one record at *p (p < end); false if a field does not convert.  The
// buffer is NUL-terminated past `end`, so no conversion reads out of bounds.
inline bool parse_record(const char** p, OptionData* d) {
    const char* q = *p;
    float* f6[6] = {&d->s, &d->strike, &d->r, &d->divq, &d->v, &d->t};
    const char* e;
    for (float* f : f6) {
        *f = parse_float(q, &e);
        if (e == q) return false;
        q = e;
    }
    while (is_space(*q)) ++q;
    if (*q == '\0') return false;
    d->OptionType = *q++;
    d->divs = parse_float(q, &e);
    if (e == q) return false;
    q = e;
    d->DGrefval = parse_float(q, &e);
    if (e == q) return false;
    *p = e;
    return true;
}

// Parses records in [p, end) into out (at most `limit`); returns the pointer
// after the last record, or nullptr on a malformed record.
inline const char* parse_range(const char* p, const char* end, std::vector<OptionData>* out,
                               size_t limit) {
    while (out->size() < limit) {
        while (p < end && is_space(*p)) ++p;
        if (p >= end) break;
        OptionData d;
        memset(&d, 0, sizeof(d));
        if (!parse_record(&p, &d)) return nullptr;
        out->push_back(d);
    }
    return p;
}

// Returns true and fills *out, or false with a message in *err.
inline bool read_options(const char* path, std::vector<OptionData>* out, int threads,
                         std::string* err) {
    // The text is mapped, so the reading threads fault its pages in parallel
    // straight from the page cache; the bytes past the end of a file's last
    // page read as zero, which NUL-terminates the text.  A file whose size is
    // a multiple of the page size (or that cannot be mapped) is read into a
    // buffer instead.
    const int fd = open(path, O_RDONLY);
    struct stat sb;
    if (fd < 0 || fstat(fd, &sb) != 0) {
        if (fd >= 0) close(fd);
        *err = std::string("Unable to open file `") + path + "'.";
        return false;
    }
    const size_t size = (size_t)sb.st_size;
    const size_t page = (size_t)sysconf(_SC_PAGESIZE);
    void* map = MAP_FAILED;
    std::unique_ptr<char[]> copy;
    if (size > 0 && size % page != 0) map = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
    const char* text;
    if (map != MAP_FAILED) {
        text = static_cast<const char*>(map);
    } else {
        copy.reset(new char[size + 1]);
        size_t got = 0;
        while (got < size) {
            const ssize_t k = read(fd, copy.get() + got, size - got);
            if (k <= 0) break;
            got += (size_t)k;
        }
        copy[got] = '\0';
        text = copy.get();
    }
    close(fd);
    struct Unmap {
        void* p;
        size_t n;
        ~Unmap() {
            if (p != MAP_FAILED) munmap(p, n);
        }
    } unmap{map, size};
    const char* p = text;
    const char* end = text + (map != MAP_FAILED ? size : strlen(text));
    char* e;
    const long n = strtol(p, &e, 0);   // "%i": decimal, octal or hex
    if (e == p || n < 0) {
        *err = std::string("Unable to read from file `") + path + "'.";
        return false;
    }
    p = e;
    const size_t want = (size_t)n;
    out->clear();
    out->reserve(want);
    // chunks of the body at line boundaries, one per thread
    const int t = std::max(1, std::min(threads, 64));
    std::vector<const char*> cut{p};
    for (int i = 1; i < t; ++i) {
        const char* c = p + (end - p) * i / t;
        c = std::max(c, cut.back());
        while (c < end && *c != '\n') ++c;
        cut.push_back(c);
    }
    cut.push_back(end);
    std::vector<std::vector<OptionData>> part(t);
    std::vector<int> ok(t, 1);
    std::vector<std::thread> th;
    for (int i = 0; i < t; ++i)
        th.emplace_back([&, i] {
            // each chunk parses to its own end; a record cut in two fails here
            const char* r = parse_range(cut[i], cut[i + 1], &part[i], want);
            ok[i] = r != nullptr && r <= cut[i + 1];
        });
    for (auto& x : th) x.join();
    bool whole = true;
    for (int i = 0; i < t; ++i) whole = whole && ok[i];
    if (whole) {
        for (int i = 0; i < t && out->size() < want; ++i)
            out->insert(out->end(), part[i].begin(),
                        part[i].begin() + (std::ptrdiff_t)std::min(part[i].size(), want - out->size()));
    } else {   // records spread over lines: one sequential pass
        out->clear();
        if (!parse_range(p, end, out, want)) out->clear();
    }
    if (out->size() != want) {
        *err = std::string("Unable to read from file `") + path + "'.";
        return false;
    }
    return true;
}

inline bool write_prices(const char* path, const float* price, int n, int threads,
                         std::string* err) {
    const int t = std::max(1, std::min(threads, 64));
    std::vector<std::string> part(t);
    std::vector<std::thread> th;
    for (int i = 0; i < t; ++i)
        th.emplace_back([&, i] {
            const int64_t a = (int64_t)n * i / t, b = (int64_t)n * (i + 1) / t;
            std::string& s = part[i];
            s.reserve((size_t)(b - a) * 24);
            // a float has at most 39 integer digits: "%.18f\n" fits in 61 bytes
            char line[128];
            for (int64_t j = a; j < b; ++j) {
                const int k = snprintf(line, sizeof(line), "%.18f\n", (double)price[j]);
                s.append(line, (size_t)std::max(0, std::min(k, (int)sizeof(line) - 1)));
            }
        });
    for (auto& x : th) x.join();
    FILE* f = fopen(path, "w");
    if (!f) {
        *err = std::string("Unable to open file `") + path + "'.";
        return false;
    }
    bool good = fprintf(f, "%i\n", n) >= 0;
    for (int i = 0; i < t && good; ++i) good = fwrite(part[i].data(), 1, part[i].size(), f) == part[i].size();
    if (fclose(f) != 0) good = false;
    if (!good) *err = std::string("Unable to write to file `") + path + "'.";
    return good;
}

// The reference's ERR_CHK test (blackscholes.c:333-340 in the FastFlow map,
// :641-647 in bs_thread, the count printed at :949-951).  The reference runs
// it inside the NUM_RUNS loop, so every pass tests every option again: the
// delta DGrefval - price is a float subtraction, fabs of it takes the float
// overload (the file is compiled as C++), and the comparison promotes that
// float to double against the double literal 1e-4, so a delta of exactly
// float(1e-4) = 9.99999974e-5 does not count.  Each failure prints the
// reference's "Error on ..." line (pass by pass, options in order, as the
// sequential build prints them) and counts once per pass.  Returns the count.
inline long long err_chk(const OptionData* data, const float* prices, int n, int passes,
                         FILE* out) {
    std::vector<int> bad;
    for (int i = 0; i < n; ++i) {
        const float delta = data[i].DGrefval - prices[i];
        if ((double)std::fabs(delta) >= 1e-4) bad.push_back(i);
    }
    for (int j = 0; j < passes && out; ++j)
        for (int i : bad) {
            const float delta = data[i].DGrefval - prices[i];
            fprintf(out, "Error on %d. Computed=%.5f, Ref=%.5f, Delta=%.5f\n", i,
                    (double)prices[i], (double)data[i].DGrefval, (double)delta);
        }
    return (long long)bad.size() * passes;
}

}  // namespace parsec_io
