// Test-only driver of the serving loop (p3arsec_amd/csrc/serve_loop.hpp) on
// the CPU: the loop's own code with a host executor in place of the GPU, so
// the dispatcher, the clients' calls and shutdown run under ThreadSanitizer
// (tests/native/Makefile builds it plain and with -fsanitize=thread).
//
//   serve_loop_stress <scenario>     exit 0 when every check holds
//
// The host executor completes a batch after a delay proportional to its
// options, can fail every k-th launch or batch, and checks on every launch
// that the batch's outputs lie in the arena and overlap neither each other,
// nor any other batch in flight, nor a resident book.
#include <stdint.h>
#include <stdio.h>

#include <atomic>
#include <map>
#include <mutex>
#include <random>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "../../p3arsec_amd/csrc/serve_loop.hpp"

using namespace bs::serve;

namespace {

std::atomic<int> g_fail{0};
#define CHECK(cond)                                                              \
    do {                                                                         \
        if (!(cond)) {                                                           \
            fprintf(stderr, "check failed: %s (%s:%d)\n", #cond, __FILE__, __LINE__); \
            g_fail.store(1);                                                     \
        }                                                                        \
    } while (0)

struct HostExec {
    static constexpr int kSlots = 6;
    static constexpr int kSmallSlots = 2;
    size_t arena = 0;
    int fail_launch_every = 0, fail_batch_every = 0;
    int64_t fixed_ns = 20000;       // per batch
    double ns_per_option = 0.002;   // ~ 5e11 options/s
    int64_t launches = 0, batches = 0, partial_launches = 0;
    int64_t small_options = 0;   // the loop's lane threshold, checked per launch
    struct Slot {
        bool used = false, fail = false;
        int64_t done_at = 0;
        std::vector<std::pair<size_t, size_t>> outs;
    } slot[kSlots];
    std::map<size_t, size_t> inflight;   // output offset -> bytes, batches in flight
    std::mutex books_mu;                  // books are added/removed by client threads
    std::map<size_t, size_t> books;       // book offset -> bytes

    void thread_start() {}
    int launch(int s, const std::vector<Admitted>& reqs, std::string* err) {
        CHECK(s >= 0 && s < kSlots && !slot[s].used);
        CHECK(!reqs.empty());
        ++launches;
        // an injected launch failure happens after the batch's work was queued
        // (as when a later kernel launch or the event record fails): its
        // outputs stay in flight until quiesce()
        const bool fail_launch = fail_launch_every && launches % fail_launch_every == 0;
        Slot& sl = slot[s];
        sl.outs.clear();
        int64_t options = 0;
        {
            std::lock_guard<std::mutex> lk(books_mu);
            for (const Admitted& r : reqs) {
                // a batch holds one lane's requests, in that lane's slots
                const bool small = small_options > 0 && r.n <= small_options;
                CHECK(small == (s < kSmallSlots));
                const size_t ob = output_bytes(r.n, r.kind);
                CHECK(r.out_off % kAlign == 0 && r.out_off + ob <= arena);
                CHECK(r.book_off + book_bytes(r.book_n) <= arena && r.n == r.book_n);
                auto bk = books.find(r.book_off);
                CHECK(bk != books.end());   // the book is resident while referenced
                // no overlap with a resident book
                auto it = books.upper_bound(r.out_off);
                if (it != books.begin()) {
                    auto pv = std::prev(it);
                    CHECK(pv->first + pv->second <= r.out_off);
                }
                if (it != books.end()) CHECK(r.out_off + ob <= it->first);
                // no overlap with any output in flight
                auto jt = inflight.upper_bound(r.out_off);
                if (jt != inflight.begin()) {
                    auto pv = std::prev(jt);
                    CHECK(pv->first + pv->second <= r.out_off);
                }
                if (jt != inflight.end()) CHECK(r.out_off + ob <= jt->first);
                inflight[r.out_off] = ob;
                sl.outs.push_back({r.out_off, ob});
                options += r.n;
            }
        }
        ++batches;
        sl.used = true;
        sl.fail = fail_batch_every && batches % fail_batch_every == 0;
        sl.done_at = steady_ns() + fixed_ns + (int64_t)(ns_per_option * (double)options);
        if (fail_launch) {
            ++partial_launches;
            *err = "injected launch failure";
            return -1;
        }
        return 0;
    }
    void quiesce(int s) {
        Slot& sl = slot[s];
        if (!sl.used) return;
        while (steady_ns() < sl.done_at) std::this_thread::yield();
        for (const auto& o : sl.outs) inflight.erase(o.first);
        sl.used = false;
    }
    int query(int s, std::string* err) {
        Slot& sl = slot[s];
        CHECK(sl.used);
        if (steady_ns() < sl.done_at) return 0;
        for (const auto& o : sl.outs) inflight.erase(o.first);
        sl.used = false;
        if (sl.fail) {
            *err = "injected batch failure";
            return -1;
        }
        return 1;
    }
    void drain() {}
};

struct World {
    HostExec ex;
    Loop<HostExec>* loop = nullptr;
    std::vector<int64_t> book_ids, book_n;
    size_t book_total = 0;

    World(size_t arena, Limits lim, bool cq, int nbooks, uint64_t seed, int64_t max_n = 1 << 16) {
        ex.arena = arena & ~(kAlign - 1);
        ex.small_options = lim.small_options;
        loop = new Loop<HostExec>(&ex, arena, lim, cq);
        std::mt19937_64 rng(seed);
        for (int b = 0; b < nbooks; ++b) {
            const int64_t n = 1 + (int64_t)(rng() % (uint64_t)max_n);
            add_book(n);
        }
        loop->start();
    }
    ~World() { delete loop; }
    bool add_book(int64_t n) {
        int64_t id;
        size_t off;
        if (loop->register_book(n, &id, &off) != kOk) return false;
        {
            std::lock_guard<std::mutex> lk(ex.books_mu);
            ex.books[off] = book_bytes(n);
        }
        book_ids.push_back(id);
        book_n.push_back(n);
        book_total += book_bytes(n);
        return true;
    }
};

// Clients submit in bulk, reap, occasionally cancel, release; every ticket
// must finish exactly once.  Returns the number of tickets finished.
int64_t run_clients(World& w, int nthreads, int rounds, uint64_t seed, int64_t* failed,
                    int64_t* cancelled) {
    std::atomic<int64_t> fin{0}, fails{0}, cancels{0};
    std::vector<std::thread> th;
    for (int c = 0; c < nthreads; ++c) {
        th.emplace_back([&, c] {
            std::mt19937_64 rng(seed + 977 * c);
            std::vector<int64_t> books, tickets;
            std::vector<int> kinds;
            for (int r = 0; r < rounds; ++r) {
                const int k = 1 + (int)(rng() % 48);
                books.resize(k);
                kinds.resize(k);
                tickets.resize(k);
                for (int i = 0; i < k; ++i) {
                    books[i] = w.book_ids[rng() % w.book_ids.size()];
                    kinds[i] = rng() % 10 < 3 ? kGreeks : kPrice;
                }
                const int rc = w.loop->submit_many(books.data(), kinds.data(), nullptr, k, tickets.data());
                CHECK(rc == kOk);
                if (rc != kOk) return;
                if (rng() % 4 == 0) {   // try to cancel one while it may still be queued
                    const int rc2 = w.loop->cancel(tickets[rng() % k]);
                    CHECK(rc2 == kOk || rc2 == kErrBusy);
                }
                for (int i = 0; i < k; ++i) {
                    std::string msg;
                    const int st = w.loop->wait(tickets[i], 10 * 1000 * 1000, &msg);
                    CHECK(st == kOk || st == kErrCancelled || st == kErrFailed);
                    if (st == kErrFailed) {
                        fails.fetch_add(1);
                        CHECK(msg.find("injected") != std::string::npos);
                    }
                    if (st == kErrCancelled) cancels.fetch_add(1);
                    CHECK(w.loop->release(tickets[i]) == kOk);
                    CHECK(w.loop->release(tickets[i]) == kErrNotFound);   // exactly once
                    fin.fetch_add(1);
                }
            }
        });
    }
    for (auto& t : th) t.join();
    *failed = fails.load();
    *cancelled = cancels.load();
    return fin.load();
}

void check_drained(World& w) {
    Stats st;
    size_t cap, used, peak;
    LoopTimes lt;
    w.loop->stats(&st, &cap, &used, &peak, &lt);
    CHECK(st.requests_queued == 0 && st.requests_running == 0);
    CHECK(used == w.book_total);   // every output returned
    CHECK(peak <= cap);
}

int scenario_stress() {
    Limits lim;
    lim.small_options = 16384;   // about a quarter of the books take the small lane
    World w((size_t)1 << 32, lim, false, 64, 11);
    int64_t failed = 0, cancelled = 0;
    const int64_t fin = run_clients(w, 6, 150, 5, &failed, &cancelled);
    Stats st;
    size_t cap, used, peak;
    LoopTimes lt;
    w.loop->stats(&st, &cap, &used, &peak, &lt);
    CHECK(st.requests_done + st.requests_failed + st.requests_cancelled == fin);
    CHECK(failed == 0 && st.launches > 0);
    check_drained(w);
    printf("stress: %lld requests, %lld launches, %lld cancelled\n", (long long)fin,
           (long long)st.launches, (long long)cancelled);
    return 0;
}

int scenario_faults() {
    Limits lim;
    lim.small_options = 16384;
    World w((size_t)1 << 32, lim, false, 32, 12);
    w.ex.fail_launch_every = 7;
    w.ex.fail_batch_every = 5;
    int64_t failed = 0, cancelled = 0;
    const int64_t fin = run_clients(w, 4, 120, 6, &failed, &cancelled);
    Stats st;
    size_t cap, used, peak;
    LoopTimes lt;
    w.loop->stats(&st, &cap, &used, &peak, &lt);
    CHECK(failed > 0 && st.requests_failed == failed);
    CHECK(st.requests_done + st.requests_failed + st.requests_cancelled == fin);
    CHECK(w.ex.partial_launches > 0);   // launches that failed with work already queued
    check_drained(w);   // failed batches returned their outputs
    printf("faults: %lld requests, %lld failed, %lld launches (%lld failed part-way)\n",
           (long long)fin, (long long)failed, (long long)st.launches,
           (long long)w.ex.partial_launches);
    return 0;
}

int scenario_shutdown() {
    Limits lim;
    lim.max_batch_requests = 8;
    auto* w = new World((size_t)1 << 32, lim, true, 16, 13);
    w->ex.fixed_ns = 200000;   // slow batches: most requests are still queued at stop()
    std::vector<int64_t> books(2000), tickets(2000);
    std::vector<int> kinds(2000, kPrice);
    for (size_t i = 0; i < books.size(); ++i) books[i] = w->book_ids[i % w->book_ids.size()];
    CHECK(w->loop->submit_many(books.data(), kinds.data(), nullptr, 2000, tickets.data()) == kOk);
    std::thread stopper([&] { w->loop->stop(); });
    stopper.join();
    int64_t done = 0, shut = 0;
    for (int64_t t : tickets) {
        std::string msg;
        const int st = w->loop->wait(t, 0, &msg);
        CHECK(st == kOk || st == kErrShutdown);
        (st == kOk ? done : shut) += 1;
    }
    CHECK(shut > 0 && done + shut == 2000);
    int64_t tk;
    CHECK(w->loop->submit(w->book_ids[0], kPrice, &tk) == kErrShutdown);
    std::vector<Completion> cq(4096);
    int64_t k = 0;
    CHECK(w->loop->reap(cq.data(), 4096, 0, &k) == kOk && k == 2000);
    printf("shutdown: %lld done, %lld failed with shutdown\n", (long long)done, (long long)shut);
    delete w;
    return 0;
}

// Books registered and released while requests run against the others.
int scenario_books() {
    Limits lim;
    World w((size_t)1 << 30, lim, false, 24, 14);
    std::atomic<bool> stop{false};
    std::thread churn([&] {
        std::mt19937_64 rng(3);
        HostExec& ex = w.ex;
        while (!stop.load()) {
            const int64_t n = 1 + (int64_t)(rng() % 50000);
            int64_t id;
            size_t off;
            if (w.loop->register_book(n, &id, &off) != kOk) continue;
            {
                std::lock_guard<std::mutex> lk(ex.books_mu);
                ex.books[off] = book_bytes(n);
            }
            {
                std::lock_guard<std::mutex> lk(ex.books_mu);
                ex.books.erase(off);   // unreferenced: nothing can be in flight on it
            }
            CHECK(w.loop->release_book(id) == kOk);
        }
    });
    int64_t failed = 0, cancelled = 0;
    const int64_t fin = run_clients(w, 4, 100, 7, &failed, &cancelled);
    stop.store(true);
    churn.join();
    CHECK(fin > 0 && failed == 0);
    check_drained(w);
    // a referenced book is busy until its requests are released
    int64_t tk;
    CHECK(w.loop->submit(w.book_ids[0], kPrice, &tk) == kOk);
    std::string msg;
    CHECK(w.loop->wait(tk, 10 * 1000 * 1000, &msg) == kOk);
    CHECK(w.loop->release_book(w.book_ids[0]) == kOk);   // done: no longer referenced
    CHECK(w.loop->release(tk) == kOk);
    printf("books: %lld requests with books churning\n", (long long)fin);
    return 0;
}

// An arena that holds the books and a few outputs: admission waits for
// released results; results are held for a while before release.
int scenario_memory() {
    Limits lim;
    const int64_t n = 40000;
    const size_t arena = 4 * book_bytes(n) + 6 * output_bytes(n, kGreeks) + 4096;
    HostExec ex;
    ex.arena = arena & ~(kAlign - 1);
    Loop<HostExec> loop(&ex, arena, lim, true);
    std::vector<int64_t> ids;
    size_t books_total = 0;
    for (int b = 0; b < 4; ++b) {
        int64_t id;
        size_t off;
        CHECK(loop.register_book(n, &id, &off) == kOk);
        ex.books[off] = book_bytes(n);
        ids.push_back(id);
        books_total += book_bytes(n);
    }
    loop.start();
    std::vector<int64_t> books(400), tickets(400);
    std::vector<int> kinds(400);
    for (int i = 0; i < 400; ++i) {
        books[i] = ids[i % 4];
        kinds[i] = i % 3 ? kPrice : kGreeks;
    }
    CHECK(loop.submit_many(books.data(), kinds.data(), nullptr, 400, tickets.data()) == kOk);
    std::vector<Completion> cq(64);
    int64_t got = 0;
    std::mt19937_64 rng(9);
    while (got < 400) {
        int64_t k = 0;
        CHECK(loop.reap(cq.data(), 64, 5 * 1000 * 1000, &k) == kOk);
        CHECK(k > 0);
        if (k == 0) break;
        std::this_thread::sleep_for(std::chrono::microseconds(50 + rng() % 200));   // hold results
        std::vector<int64_t> rel;
        for (int64_t i = 0; i < k; ++i) {
            CHECK(cq[i].state == kDone);
            rel.push_back(cq[i].ticket);
        }
        CHECK(loop.release_many(rel.data(), (int64_t)rel.size()) == kOk);
        got += k;
    }
    Stats st;
    size_t cap, used, peak;
    LoopTimes lt;
    loop.stats(&st, &cap, &used, &peak, &lt);
    CHECK(st.requests_done == 400 && st.admission_waits > 0);
    CHECK(used == books_total && peak <= cap);
    printf("memory: 400 requests, %lld admission waits, peak %zu of %zu bytes\n",
           (long long)st.admission_waits, peak, cap);
    return 0;
}

// Small requests complete beside a slow bulk batch, not behind it: one huge
// request is launched (its batch takes ~200 ms), then small requests are
// submitted; each must finish while the huge one is still running.
int scenario_lanes() {
    Limits lim;
    lim.small_options = 65536;
    HostExec ex;
    ex.arena = ((size_t)1 << 34);
    ex.small_options = lim.small_options;
    ex.ns_per_option = 0.0;
    ex.fixed_ns = 0;
    Loop<HostExec> loop(&ex, (size_t)1 << 34, lim, false);
    int64_t big, small;
    size_t off;
    CHECK(loop.register_book(int64_t(1) << 20, &big, &off) == kOk);
    ex.books[off] = book_bytes(int64_t(1) << 20);
    CHECK(loop.register_book(1000, &small, &off) == kOk);
    ex.books[off] = book_bytes(1000);
    ex.ns_per_option = 200.0;   // the 2^20-option batch: ~210 ms; a 1000-option one: 0.2 ms
    loop.start();
    int64_t tb;
    CHECK(loop.submit(big, kGreeks, &tb) == kOk);
    std::this_thread::sleep_for(std::chrono::milliseconds(5));   // the big batch is in flight
    std::string msg;
    for (int i = 0; i < 20; ++i) {
        int64_t ts;
        CHECK(loop.submit(small, i % 2 ? kGreeks : kPrice, &ts) == kOk);
        CHECK(loop.wait(ts, 10 * 1000 * 1000, &msg) == kOk);
        int st = -1;
        CHECK(loop.poll(tb, &st) == kOk && st == kRunning);   // the big one is still running
        CHECK(loop.release(ts) == kOk);
    }
    CHECK(loop.wait(tb, 10 * 1000 * 1000, &msg) == kOk);
    CHECK(loop.release(tb) == kOk);
    printf("lanes: 20 small requests finished while a bulk batch ran\n");
    return 0;
}

// The loop's ticket lock on its own: eight threads take it in turn; the
// counter it guards is plain memory, so a lost exclusion shows up as a wrong
// total (and as a race under ThreadSanitizer); with one thread sleeping while
// it holds the lock, the others pass from spinning to yielding and still get
// it in arrival order, so every thread's turns are counted.
int scenario_lock() {
    TicketLock mu;
    int64_t counter = 0;
    int64_t turns[8] = {};
    constexpr int kThreadsN = 8, kIters = 20000;
    std::vector<std::thread> th;
    for (int t = 0; t < kThreadsN; ++t) {
        th.emplace_back([&, t] {
            for (int i = 0; i < kIters; ++i) {
                std::lock_guard<TicketLock> lk(mu);
                ++counter;
                ++turns[t];
                if (t == 0 && i % 5000 == 0)
                    std::this_thread::sleep_for(std::chrono::milliseconds(2));
            }
        });
    }
    for (auto& x : th) x.join();
    CHECK(counter == (int64_t)kThreadsN * kIters);
    for (int t = 0; t < kThreadsN; ++t) CHECK(turns[t] == kIters);
    printf("lock: %lld increments under the ticket lock\n", (long long)counter);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) {
        fprintf(stderr, "usage: %s stress|faults|shutdown|books|memory|lanes|lock\n", argv[0]);
        return 2;
    }
    const std::string s = argv[1];
    if (s == "stress") scenario_stress();
    else if (s == "faults") scenario_faults();
    else if (s == "shutdown") scenario_shutdown();
    else if (s == "books") scenario_books();
    else if (s == "memory") scenario_memory();
    else if (s == "lanes") scenario_lanes();
    else if (s == "lock") scenario_lock();
    else return 2;
    return g_fail.load() ? 1 : 0;
}
