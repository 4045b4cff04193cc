// bs_serve_bench -- load generator for the serving loop (libbs_mi355.so).
//
// Registers a catalog of resident books whose sizes are log-uniform on
// [1, 2^20] options (seeded, the same option distributions as the batch
// benchmark), then either
//   * open loop: Poisson arrivals at --rate requests/s for --duration-s,
//     each request a random book, BS_REQ_GREEKS with probability
//     --greeks-frac, else BS_REQ_PRICE; arrivals already due when the client
//     gets to them go in one bs_submit_many call (at most 1024); or
//   * --saturate: closed loop keeping --inflight requests outstanding, to
//     measure the loop's throughput ceiling.
// Latency = scheduled arrival -> completion observed by the dispatcher (the
// request's outputs are resident in HBM and readable), both on
// CLOCK_MONOTONIC, kept in a log-binned histogram (percentiles within
// 0.12%).  With --verify-every K, every K-th completed request's outputs are
// copied back and compared bit for bit with the same book priced alone by
// the batch entry points (bs_price_f32, bs_price_greeks_f32).  Prints one JSON
// object.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "bs_mi355.h"

namespace {

int64_t now_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// Ends the process at once: a failed call can come from either client thread
// while the other and the context's dispatcher are still running, and exit()
// would run static destructors (the GPU runtime's among them) under them.
void die(const char* what, int rc) {
    fprintf(stderr, "%s failed: %d (%s) %s\n", what, rc, bs_status_string(rc), bs_last_error());
    fflush(stdout);
    fflush(stderr);
    _Exit(1);
}
#define CHECK(call)                      \
    do {                                 \
        int rc_ = (call);                \
        if (rc_ != BS_OK) die(#call, rc_); \
    } while (0)

struct Opts {
    double rate = 20000.0, duration_s = 5.0, greeks_frac = 0.3, budget_gb = 32.0;
    int books = 256, inflight = 2048, use_graph = 0, lane_options = 0, verify_every = 0;
    int block_threads = 0;
    int64_t max_book = 1 << 20, small = 65536;
    uint64_t seed = 20261207;
    bool saturate = false;
};

struct Book {
    int64_t id, n;
};

// A request's tag: its arrival (ns after the run's start) in the high bits,
// its book's catalog index and its kind below, so a completion names all three.
constexpr int kTagBookBits = 19;
int64_t make_tag(int64_t arrival_rel, int book, int kind) {
    return (arrival_rel << (kTagBookBits + 1)) | ((int64_t)book << 1) | (kind == BS_REQ_GREEKS);
}
int64_t tag_arrival(int64_t tag) { return tag >> (kTagBookBits + 1); }
int tag_book(int64_t tag) { return (int)((tag >> 1) & ((1 << kTagBookBits) - 1)); }
bool tag_greeks(int64_t tag) { return tag & 1; }

// One book priced alone by the batch entry points: the outputs every serving
// result of that book must equal bit for bit.
struct Alone {
    std::vector<float> price;
    std::vector<float> greeks;   // 5 x n: price, delta, vega, rho, theta
};

// Latency histogram: 1000 log-spaced bins per decade from 100 ns up; a
// percentile is the geometric centre of its bin (within 0.12%).  Constant
// memory and constant time per sample, so a long run never pauses the
// thread that releases results.
struct LatHist {
    static constexpr int kPerDecade = 1000;
    static constexpr double kLoNs = 100.0;
    static constexpr int kBins = 11 * kPerDecade;   // up to 1e13 ns
    std::vector<int64_t> count = std::vector<int64_t>(kBins + 2, 0);
    int64_t n = 0;
    double max_ns = 0.0;
    void add(double ns) {
        int b = ns <= kLoNs ? 0 : (int)std::floor(std::log10(ns / kLoNs) * kPerDecade) + 1;
        b = std::min(b, kBins + 1);
        ++count[b];
        ++n;
        max_ns = std::max(max_ns, ns);
    }
    // the p-quantile in ms: the sample of rank ceil(p * n)
    double pct_ms(double p) const {
        if (n == 0) return 0.0;
        const int64_t rank = std::max<int64_t>(1, (int64_t)std::ceil(p * (double)n));
        int64_t cum = 0;
        for (int b = 0; b < kBins + 2; ++b) {
            cum += count[b];
            if (cum >= rank)
                return std::min(max_ns, b == 0 ? kLoNs : kLoNs * std::pow(10.0, (b - 0.5) / kPerDecade)) * 1e-6;
        }
        return max_ns * 1e-6;
    }
};

}  // namespace

int main(int argc, char** argv) {
    Opts o;
    for (int i = 1; i < argc; ++i) {
        auto val = [&](const char* name) -> const char* {
            if (strcmp(argv[i], name) == 0 && i + 1 < argc) return argv[++i];
            return nullptr;
        };
        const char* v;
        if ((v = val("--rate"))) o.rate = atof(v);
        else if ((v = val("--duration-s"))) o.duration_s = atof(v);
        else if ((v = val("--greeks-frac"))) o.greeks_frac = atof(v);
        else if ((v = val("--budget-gb"))) o.budget_gb = atof(v);
        else if ((v = val("--books"))) o.books = atoi(v);
        else if ((v = val("--inflight"))) o.inflight = atoi(v);
        else if ((v = val("--max-book"))) o.max_book = atoll(v);
        else if ((v = val("--seed"))) o.seed = strtoull(v, nullptr, 10);
        else if ((v = val("--use-graph"))) o.use_graph = atoi(v);
        else if ((v = val("--small-options"))) o.lane_options = atoi(v);
        else if ((v = val("--verify-every"))) o.verify_every = atoi(v);
        else if ((v = val("--block-threads"))) o.block_threads = atoi(v);
        else if (strcmp(argv[i], "--saturate") == 0) o.saturate = true;
        else {
            fprintf(stderr, "unknown argument %s\n", argv[i]);
            return 2;
        }
    }

    // the serving kernels' block size is read when the context is created
    if (o.block_threads) CHECK(bs_config_set(BS_CFG_BLOCK_THREADS, o.block_threads));
    bs_serve_opts so{};
    so.use_graph = o.use_graph;
    so.completion_queue = 1;
    so.small_request_options = o.lane_options;
    bs_ctx_t ctx = nullptr;
    CHECK(bs_ctx_create(0, (size_t)(o.budget_gb * (double)(1ull << 30)), &so, &ctx));

    // catalog: log-uniform sizes, options drawn as in the batch benchmark
    std::mt19937_64 rng(o.seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    if (o.books > (1 << kTagBookBits)) {
        fprintf(stderr, "--books is at most %d\n", 1 << kTagBookBits);
        return 2;
    }
    std::vector<Book> books;
    std::vector<Alone> alone;
    int64_t catalog_options = 0;
    float* dev = nullptr;   // scratch for the alone pricing: 6 inputs + 5 outputs of the largest book
    if (o.verify_every > 0) CHECK(bs_malloc((void**)&dev, (size_t)o.max_book * 11 * sizeof(float)));
    for (int b = 0; b < o.books; ++b) {
        const int64_t n = std::max<int64_t>(
            1, (int64_t)std::floor(std::exp(U(rng) * std::log((double)o.max_book + 1.0))));
        std::vector<float> s(n), k(n), r(n), v(n), t(n);
        std::vector<uint8_t> ty(n);
        for (int64_t i = 0; i < n; ++i) {
            s[i] = (float)(10.0 + 190.0 * U(rng));
            k[i] = (float)(s[i] * (0.5 + U(rng)));
            r[i] = (float)(0.01 + 0.09 * U(rng));
            v[i] = (float)(0.05 + 0.60 * U(rng));
            t[i] = (float)(0.05 + 1.95 * U(rng));
            ty[i] = U(rng) < 0.5 ? 1 : 0;
        }
        bs_book_t id;
        CHECK(bs_book_register(ctx, s.data(), k.data(), r.data(), v.data(), t.data(), ty.data(), n,
                               &id));
        books.push_back({id, n});
        catalog_options += n;
        if (o.verify_every > 0) {
            const size_t nb = (size_t)n * sizeof(float);
            float* in[5] = {dev, dev + o.max_book, dev + 2 * o.max_book, dev + 3 * o.max_book,
                            dev + 4 * o.max_book};
            uint8_t* dty = (uint8_t*)(dev + 5 * o.max_book);
            float* out = dev + 6 * o.max_book;
            const float* src[5] = {s.data(), k.data(), r.data(), v.data(), t.data()};
            for (int j = 0; j < 5; ++j) CHECK(bs_memcpy_h2d(in[j], src[j], nb, nullptr));
            CHECK(bs_memcpy_h2d(dty, ty.data(), (size_t)n, nullptr));
            Alone a;
            a.price.resize(n);
            a.greeks.resize(5 * (size_t)n);
            CHECK(bs_price_f32(in[0], in[1], in[2], in[3], in[4], dty, out, n, nullptr));
            CHECK(bs_memcpy_d2h(a.price.data(), out, nb, nullptr));
            CHECK(bs_price_greeks_f32(in[0], in[1], in[2], in[3], in[4], dty, out,
                                      out + o.max_book, out + 2 * o.max_book,
                                      out + 3 * o.max_book, out + 4 * o.max_book, n, nullptr));
            for (int j = 0; j < 5; ++j)
                CHECK(bs_memcpy_d2h(a.greeks.data() + (size_t)j * n, out + (size_t)j * o.max_book,
                                    nb, nullptr));
            CHECK(bs_device_sync());
            alone.push_back(std::move(a));
        }
    }
    if (dev) CHECK(bs_free(dev));

    // Completions come from the context's completion queue (bs_reap) in the
    // order requests finish; each is released right away.  Each request's tag
    // is its arrival time and its completion carries its size, so the two
    // client threads share nothing but the outstanding count.
    std::atomic<bool> done_submitting{false};
    LatHist lat_all, lat_small;
    int64_t options_done = 0, requests_done = 0;
    std::atomic<int64_t> outstanding{0};
    int64_t c_wait = 0, c_work = 0, s_wait = 0, s_work = 0;   // client thread time, ns
    int64_t verified[2][2] = {{0, 0}, {0, 0}};   // [greeks][n > 64K]
    int64_t mismatched = 0;
    const int64_t t0 = now_ns();

    std::thread completer([&] {
        std::vector<bs_completion> cq(8192);
        std::vector<bs_ticket_t> rel;
        std::vector<float> got(5 * (size_t)o.max_book);
        for (;;) {
            int64_t k = 0;
            const int64_t w0 = now_ns();
            CHECK(bs_reap(ctx, cq.data(), (int64_t)cq.size(), 1000, &k));
            const int64_t w1 = now_ns();
            c_wait += w1 - w0;
            if (k == 0) {
                if (done_submitting.load() && outstanding.load() == 0) break;
                continue;
            }
            rel.clear();
            for (int64_t i = 0; i < k; ++i) {
                const bs_completion& c = cq[i];
                if (c.state != BS_TICKET_DONE) die("request", c.status);
                const double ns = (double)(c.done_ns - (t0 + tag_arrival(c.user)));
                if (o.verify_every > 0 && requests_done % o.verify_every == 0) {
                    const Alone& a = alone[tag_book(c.user)];
                    const bool g = tag_greeks(c.user);
                    const size_t n = (size_t)c.n;
                    float* p[5] = {got.data(), got.data() + n, got.data() + 2 * n,
                                   got.data() + 3 * n, got.data() + 4 * n};
                    if (g) CHECK(bs_result_copy(ctx, c.ticket, p[0], p[1], p[2], p[3], p[4]));
                    else CHECK(bs_result_copy(ctx, c.ticket, p[0], nullptr, nullptr, nullptr, nullptr));
                    const float* want = g ? a.greeks.data() : a.price.data();
                    if (memcmp(got.data(), want, (g ? 5 : 1) * n * sizeof(float)) != 0) ++mismatched;
                    ++verified[g][c.n > o.small];
                }
                lat_all.add(ns);
                if (c.n <= o.small) lat_small.add(ns);
                options_done += c.n;
                ++requests_done;
                rel.push_back(c.ticket);
            }
            CHECK(bs_result_release_many(ctx, rel.data(), (int64_t)rel.size()));
            outstanding.fetch_sub(k);
            c_work += now_ns() - w1;
        }
    });

    std::exponential_distribution<double> gap(o.rate);
    std::uniform_int_distribution<int> pick(0, (int)books.size() - 1);
    const int64_t t_end = t0 + (int64_t)(o.duration_s * 1e9);
    double next = (double)t0 + gap(rng) * 1e9;   // the next scheduled arrival (open loop)
    int64_t submitted = 0;
    std::vector<bs_book_t> ids;
    std::vector<int> kinds;
    std::vector<int64_t> tags;
    std::vector<bs_ticket_t> tks;
    auto add = [&](int64_t arrival) {
        const int b = pick(rng);
        const int kind = U(rng) < o.greeks_frac ? BS_REQ_GREEKS : BS_REQ_PRICE;
        ids.push_back(books[b].id);
        kinds.push_back(kind);
        tags.push_back(make_tag(std::max<int64_t>(0, arrival - t0), b, kind));
    };
    for (;;) {
        ids.clear();
        kinds.clear();
        tags.clear();
        if (o.saturate) {
            // top the window up to --inflight in one call
            int64_t room;
            const int64_t w0 = now_ns();
            while ((room = o.inflight - outstanding.load()) <= 0) std::this_thread::yield();
            const int64_t arrival = now_ns();
            s_wait += arrival - w0;
            if (arrival >= t_end) break;
            for (int64_t i = 0; i < std::min<int64_t>(room, 1024); ++i) add(arrival);
        } else {
            if (next >= (double)t_end) break;
            const int64_t w0 = now_ns();
            while ((double)now_ns() < next) {
            }
            s_wait += now_ns() - w0;
            // every arrival already due goes in this call, each tagged with its
            // own scheduled time (latency counts from there)
            const double now = (double)now_ns();
            while (next <= now && next < (double)t_end && ids.size() < 1024) {
                add((int64_t)next);
                next += gap(rng) * 1e9;
            }
        }
        const int64_t want = (int64_t)ids.size();
        tks.resize(ids.size());
        outstanding.fetch_add(want);
        const int64_t c0 = now_ns();
        CHECK(bs_submit_many(ctx, ids.data(), kinds.data(), tags.data(), want, tks.data()));
        s_work += now_ns() - c0;
        submitted += want;
    }
    done_submitting.store(true);
    completer.join();
    const double wall = (double)(now_ns() - t0) * 1e-9;
    bs_serve_stats st;
    CHECK(bs_ctx_stats(ctx, &st));
    CHECK(bs_ctx_destroy(ctx));

    printf("{\"mode\": \"%s\", \"rate_req_s\": %.1f, \"duration_s\": %.3f, \"requests\": %lld, "
           "\"options\": %lld, \"options_per_s\": %.6e, \"p50_ms\": %.4f, \"p99_ms\": %.4f, "
           "\"p50_ms_le64k\": %.4f, \"p99_ms_le64k\": %.4f, \"max_ms\": %.4f, "
           "\"launches\": %lld, \"mean_requests_per_launch\": %.2f, \"books\": %d, "
           "\"catalog_options\": %lld, \"greeks_frac\": %.2f, \"arena_peak_bytes\": %lld, "
           "\"admission_waits\": %lld, \"use_graph\": %d, \"block_threads\": %d, "
           "\"small_request_options\": %d, "
           "\"mean_queue_us\": %.2f, "
           "\"mean_batch_us\": %.2f, \"mean_options_per_launch\": %.0f, "
           "\"dispatch_ms\": {\"admit\": %.1f, \"launch\": %.1f, \"retire\": %.1f, "
           "\"gpu_idle\": %.1f, \"idle_launches\": %lld}, "
           "\"client_ms\": {\"reap_wait\": %.1f, \"reap_work\": %.1f, \"submit_wait\": %.1f, "
           "\"submit_work\": %.1f}, \"verify_every\": %d, \"verified\": {\"price_le64k\": %lld, "
           "\"price_gt64k\": %lld, \"greeks_le64k\": %lld, \"greeks_gt64k\": %lld}, "
           "\"mismatched\": %lld}\n",
           o.saturate ? "closed-loop saturation" : "open-loop poisson", o.saturate ? 0.0 : o.rate,
           wall, (long long)requests_done, (long long)options_done, (double)options_done / wall,
           lat_all.pct_ms(0.50), lat_all.pct_ms(0.99), lat_small.pct_ms(0.50), lat_small.pct_ms(0.99),
           lat_all.max_ns * 1e-6,
           (long long)st.launches,
           st.launches ? (double)requests_done / (double)st.launches : 0.0, o.books,
           (long long)catalog_options, o.greeks_frac, (long long)st.arena_peak,
           (long long)st.admission_waits, o.use_graph, o.block_threads, o.lane_options,
           st.requests_done ? (double)st.queue_ns_total * 1e-3 / (double)st.requests_done : 0.0,
           st.launches ? (double)st.batch_ns_total * 1e-3 / (double)st.launches : 0.0,
           st.launches ? (double)st.options_priced / (double)st.launches : 0.0,
           st.dispatch_admit_ns * 1e-6, st.dispatch_launch_ns * 1e-6, st.dispatch_retire_ns * 1e-6,
           st.gpu_idle_ns * 1e-6, (long long)st.idle_launches, c_wait * 1e-6, c_work * 1e-6,
           s_wait * 1e-6, s_work * 1e-6, o.verify_every, (long long)verified[0][0],
           (long long)verified[0][1], (long long)verified[1][0], (long long)verified[1][1],
           (long long)mismatched);
    return 0;
}
