// Test-only: p3arsec_amd/csrc/tools/parsec_io.hpp against the reference's own
// I/O calls (blackscholes.c:701-734 fscanf, :923-947 fprintf) on the CPU.
// Every input file is read by read_options() with 1, 2, 3 and 8 threads and by
// a sequential fscanf loop with the reference's format strings; both must agree
// on success, and on every field's bits.  write_prices() must write the same
// bytes as one fprintf("%.18f\n") per price.  Exit 0 when every case agrees.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <random>
#include <string>
#include <vector>

#include "parsec_io.hpp"

using parsec_io::OptionData;

static int g_bad = 0;

static bool read_reference(const char* path, std::vector<OptionData>* out) {
    FILE* f = fopen(path, "r");
    if (!f) return false;
    int n = 0;
    if (fscanf(f, "%i", &n) != 1 || n < 0) {
        fclose(f);
        return false;
    }
    out->assign((size_t)n, OptionData{});
    for (int i = 0; i < n; ++i) {
        OptionData& d = (*out)[(size_t)i];
        memset(&d, 0, sizeof(d));
        if (fscanf(f, "%f %f %f %f %f %f %c %f %f", &d.s, &d.strike, &d.r, &d.divq, &d.v, &d.t,
                   &d.OptionType, &d.divs, &d.DGrefval) != 9) {
            fclose(f);
            return false;
        }
    }
    fclose(f);
    return true;
}

static void check_file(const char* name, const std::string& text) {
    const std::string path = std::string("/tmp/parsec_io_check_") + name + ".txt";
    FILE* f = fopen(path.c_str(), "wb");
    fwrite(text.data(), 1, text.size(), f);
    fclose(f);
    std::vector<OptionData> want;
    const bool ok_ref = read_reference(path.c_str(), &want);
    for (int t : {1, 2, 3, 8}) {
        std::vector<OptionData> got;
        std::string err;
        const bool ok = parsec_io::read_options(path.c_str(), &got, t, &err);
        if (ok != ok_ref) {
            printf("%s, %d threads: read %s, reference %s\n", name, t, ok ? "ok" : "failed",
                   ok_ref ? "ok" : "failed");
            ++g_bad;
            continue;
        }
        if (ok && (got.size() != want.size() ||
                   memcmp(got.data(), want.data(), got.size() * sizeof(OptionData)) != 0)) {
            printf("%s, %d threads: records differ\n", name, t);
            ++g_bad;
        }
    }
    remove(path.c_str());
    printf("%-22s %s, %zu records\n", name, ok_ref ? "read" : "refused", want.size());
}

static std::string records(int n, const char* fmt, const char* sep, uint64_t seed) {
    std::mt19937_64 r(seed);
    std::uniform_real_distribution<double> u(0, 1);
    std::string s = std::to_string(n) + sep;
    char line[512];
    for (int i = 0; i < n; ++i) {
        const double sp = 10 + 190 * u(r);
        snprintf(line, sizeof(line), fmt, sp, sp * (0.5 + u(r)), 0.01 + 0.09 * u(r), 0.05 + 0.6 * u(r),
                 0.05 + 1.95 * u(r), u(r) < 0.5 ? 'P' : 'C', 10 * u(r));
        s += line;
        s += sep;
    }
    return s;
}

// parsec_io_check --errchk <input> <prices> <passes>: runs err_chk on the
// input's DGrefval column and the prices of a reference output file, printing
// what the reference's ERR_CHK build prints (its "Error on" lines, then
// "Num Errors: N").
static int errchk_mode(const char* input, const char* prices_path, int passes) {
    std::vector<OptionData> d;
    std::string err;
    if (!parsec_io::read_options(input, &d, 1, &err)) {
        printf("read: %s\n", err.c_str());
        return 2;
    }
    FILE* f = fopen(prices_path, "r");
    int n = 0;
    if (!f || fscanf(f, "%i", &n) != 1 || n != (int)d.size()) return 2;
    std::vector<float> p((size_t)n);
    for (int i = 0; i < n; ++i)
        if (fscanf(f, "%f", &p[(size_t)i]) != 1) return 2;
    fclose(f);
    const long long k = parsec_io::err_chk(d.data(), p.data(), n, passes, stdout);
    printf("Num Errors: %lld\n", k);
    return 0;
}

int main(int argc, char** argv) {
    if (argc == 5 && strcmp(argv[1], "--errchk") == 0)
        return errchk_mode(argv[2], argv[3], atoi(argv[4]));
    const char* std9 = "%.9g %.9g %.9g 0 %.9g %.9g %c 0 %.9g";
    check_file("regular", records(10007, std9, "\n", 1));
    check_file("native_fixed", records(5003, "%.2f %.2f %.4f 0.00 %.2f %.2f %c 0.00 %.15f", "\n", 2));
    check_file("scientific_tabs", records(2001, "%e\t%e  %e 0e0 %E\t\t%e %c 0 %e", "\r\n", 3));
    check_file("spread_over_lines", records(999, "%.9g %.9g\n%.9g 0 %.9g\n%.9g %c 0 %.9g", "\n", 4));
    std::string two = records(1000, std9, "\n", 5);
    for (size_t i = two.find('\n') + 1, k = 0; i < two.size(); ++i)   // join pairs of lines
        if (two[i] == '\n' && (k++ % 2 == 0)) two[i] = ' ';
    check_file("two_per_line", two);
    std::string more = records(100, std9, "\n", 6);
    more.replace(0, 3, "90");                                            // N = 90 of 100
    check_file("more_records_than_n", more);
    std::string fewer = records(100, std9, "\n", 7);
    fewer.replace(0, 3, "101");
    check_file("fewer_records_than_n", fewer);
    std::string junk = records(50, std9, "\n", 8);
    junk += "not a number at all\n";
    check_file("trailing_junk", junk);
    std::string bad = records(50, std9, "\n", 9);
    bad.replace(bad.find('\n', 200) + 1, 3, "x.y");
    check_file("malformed_field", bad);
    check_file("empty_book", "0\n");
    std::string hex = records(16, std9, "\n", 10);
    hex.replace(0, 2, "0x10");
    check_file("hex_count", hex);
    check_file("special_values", "3\ninf 100 0.05 0 0.3 1 P 0 0\nnan 100 0.05 0 0.3 1 C 0 0\n"
                                 "1e-40 -0 0.05 0 1e38 1 P 0 -inf\n");

    // field conversion: random decimals of 1 to 19 digits, 0 to 21 of them
    // after the point, with and without signs, against strtof bit for bit
    {
        std::mt19937_64 r(12);
        int64_t diff = 0;
        char num[64];
        for (int i = 0; i < 2000000; ++i) {
            const int nd = 1 + (int)(r() % 19);
            const int fr = (int)(r() % std::min(nd + 3, 22));
            int k = 0;
            if (r() % 3 == 0) num[k++] = r() % 2 ? '-' : '+';
            const int ip = std::max(0, nd - fr);
            for (int j = 0; j < ip; ++j) num[k++] = (char)('0' + r() % 10);
            if (ip == 0 && r() % 2) num[k++] = '0';
            num[k++] = '.';
            for (int j = 0; j < fr; ++j) num[k++] = (char)('0' + r() % 10);
            num[k++] = ' ';
            num[k] = '\0';
            const char* e1;
            char* e2;
            const float a = parsec_io::parse_float(num, &e1);
            const float b = strtof(num, &e2);
            if (memcmp(&a, &b, 4) != 0 || e1 != e2) {
                if (diff < 5) printf("parse_float(\"%s\") = %a, strtof %a\n", num, a, b);
                ++diff;
            }
        }
        // values near float rounding boundaries: the midpoint between two
        // floats printed with 17-19 significant digits, and its neighbours
        for (int i = 0; i < 500000; ++i) {
            const float f = std::uniform_real_distribution<float>(1e-6f, 3e5f)(r);
            const double mid = ((double)f + (double)nextafterf(f, INFINITY)) / 2;
            for (double x : {mid, nextafter(mid, 0.0), nextafter(mid, 1e30)}) {
                snprintf(num, sizeof(num), "%.*g ", 17 + (int)(r() % 3), x);
                const char* e1;
                char* e2;
                const float a = parsec_io::parse_float(num, &e1);
                const float b = strtof(num, &e2);
                if (memcmp(&a, &b, 4) != 0 || e1 != e2) {
                    if (diff < 5) printf("parse_float(\"%s\") = %a, strtof %a\n", num, a, b);
                    ++diff;
                }
            }
        }
        printf("field conversion: 3.5M decimals, %lld differ from strtof\n", (long long)diff);
        if (diff) ++g_bad;
    }

    // output: byte-identical to one fprintf per price
    std::vector<float> p = {0.0f, -0.0f, 1.0f, 3.5e-8f, -42.25f, 3.4028235e38f, 1e-45f,
                            INFINITY, -INFINITY, NAN};
    std::mt19937_64 r(11);
    std::uniform_real_distribution<float> u(-300, 300);
    for (int i = 0; i < 20000; ++i) p.push_back(u(r));
    for (int t : {1, 3, 8}) {
        std::string err;
        const char* a = "/tmp/parsec_io_check_out_a.txt";
        const char* b = "/tmp/parsec_io_check_out_b.txt";
        if (!parsec_io::write_prices(a, p.data(), (int)p.size(), t, &err)) {
            printf("write failed: %s\n", err.c_str());
            ++g_bad;
            continue;
        }
        FILE* f = fopen(b, "w");
        fprintf(f, "%i\n", (int)p.size());
        for (float x : p) fprintf(f, "%.18f\n", x);
        fclose(f);
        std::string ta, tb;
        for (auto [path, s] : {std::pair<const char*, std::string*>{a, &ta}, {b, &tb}}) {
            FILE* g = fopen(path, "rb");
            char c[65536];
            size_t k;
            while ((k = fread(c, 1, sizeof(c), g)) > 0) s->append(c, k);
            fclose(g);
            remove(path);
        }
        if (ta != tb) {
            printf("output with %d threads differs from fprintf\n", t);
            ++g_bad;
        }
    }
    printf("output: %zu prices, byte-identical to fprintf\n", p.size());

    // ERR_CHK (blackscholes.c:333-340): float delta, fabs promoted to double
    // against 1e-4, counted once per pass.  Expected per case from that rule:
    // exactly float(1e-4) either way is not an error (it is below the double
    // 1e-4); the next float up is; 100.0001f - 100 is 9.918e-5 (not an error),
    // 100.00011f - 100 is 1.068e-4 (an error); a NaN price is not counted.
    {
        const float f1e4 = 1e-4f, up = nextafterf(1e-4f, 1.0f);
        struct Case { float price, ref; bool error; };
        const Case cases[] = {{0.0f, f1e4, false},     {0.0f, -f1e4, false},
                              {0.0f, up, true},        {0.0f, -up, true},
                              {100.0f, 100.0001f, false}, {100.0f, 100.00011f, true},
                              {NAN, 1.0f, false},      {5.0f, 5.0f, false}};
        std::vector<OptionData> d;
        std::vector<float> pr;
        int want = 0;
        for (const Case& c : cases) {
            OptionData o;
            memset(&o, 0, sizeof(o));
            o.DGrefval = c.ref;
            d.push_back(o);
            pr.push_back(c.price);
            want += c.error;
        }
        const char* path = "/tmp/parsec_io_check_errchk.txt";
        FILE* f = fopen(path, "w");
        const long long got = parsec_io::err_chk(d.data(), pr.data(), (int)d.size(), 3, f);
        fclose(f);
        std::vector<std::string> lines;
        char buf[256];
        f = fopen(path, "r");
        while (fgets(buf, sizeof(buf), f)) lines.push_back(buf);
        fclose(f);
        remove(path);
        const bool ok = got == 3LL * want && want == 3 && lines.size() == 9 &&
                        lines[0] == "Error on 2. Computed=0.00000, Ref=0.00010, Delta=0.00010\n" &&
                        lines[2] == "Error on 5. Computed=100.00000, Ref=100.00011, Delta=0.00011\n" &&
                        lines[3] == lines[0];
        printf("err_chk: %lld errors over 3 passes (want %d), %zu lines\n", got, 3 * want,
               lines.size());
        if (!ok) ++g_bad;
    }
    return g_bad ? 1 : 0;
}
