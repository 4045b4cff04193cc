// Test-only C front of the serving loop's host bookkeeping (serve_core), so
// the CPU tests can drive admission, queueing, cancellation and failure paths
// without a GPU.  Built by tests/native/Makefile; never part of the product.
#include <stdint.h>

#include <unordered_map>
#include <vector>

#include "../../p3arsec_amd/csrc/serve_core.hpp"

using bs::serve::Core;

extern "C" {

void* sc_create(int64_t arena_bytes, int64_t max_req, int64_t max_opts) {
    bs::serve::Limits lim;
    if (max_req > 0) lim.max_batch_requests = max_req;
    if (max_opts > 0) lim.max_batch_options = max_opts;
    return new Core((size_t)arena_bytes, lim);
}
void sc_destroy(void* h) { delete static_cast<Core*>(h); }

int sc_register(void* h, int64_t n, int64_t* id, int64_t* off) {
    size_t o = 0;
    int rc = static_cast<Core*>(h)->register_book(n, id, &o);
    *off = (int64_t)o;
    return rc;
}
int sc_release_book(void* h, int64_t id) { return static_cast<Core*>(h)->release_book(id); }
int sc_submit(void* h, int64_t book, int kind, int64_t now, int64_t* ticket) {
    return static_cast<Core*>(h)->submit(book, kind, now, ticket);
}
int sc_cancel(void* h, int64_t t, int64_t now) { return static_cast<Core*>(h)->cancel(t, now); }
int sc_submit_many(void* h, const int64_t* books, const int* kinds, const int64_t* users,
                   int64_t count, int64_t now, int64_t* tickets) {
    return static_cast<Core*>(h)->submit_many(books, kinds, users, count, now, tickets);
}
int sc_admit_cap(void* h, int64_t now, int64_t cap, int64_t* tickets, int cn) {
    std::vector<bs::serve::Admitted> b;
    static_cast<Core*>(h)->admit(now, &b, cap);
    int i = 0;
    for (const auto& r : b) {
        if (i < cn) tickets[i] = r.ticket;
        ++i;
    }
    return i;
}
int64_t sc_queued_options(void* h) { return static_cast<Core*>(h)->queued_options(); }
// Two lanes: requests of at most small_options options queue in lane 0.
void* sc_create_lanes(int64_t arena_bytes, int64_t small_options) {
    bs::serve::Limits lim;
    lim.small_options = small_options;
    return new Core((size_t)arena_bytes, lim);
}
int sc_admit_lane(void* h, int64_t now, int lane, int64_t* tickets, int cn) {
    std::vector<bs::serve::Admitted> b;
    static_cast<Core*>(h)->admit(now, &b, 0, lane);
    int i = 0;
    for (const auto& r : b) {
        if (i < cn) tickets[i] = r.ticket;
        ++i;
    }
    return i;
}
int64_t sc_queued_lane(void* h, int lane) { return (int64_t)static_cast<Core*>(h)->queued(lane); }
void sc_enable_completions(void* h, int on) { static_cast<Core*>(h)->enable_completions(on != 0); }
// Pops up to max completions into parallel arrays; returns the count.
int64_t sc_reap(void* h, int64_t* tickets, int* states, int* statuses, int64_t* users,
                int64_t max) {
    std::vector<bs::serve::Completion> v((size_t)(max > 0 ? max : 0));
    const int64_t k = static_cast<Core*>(h)->reap(v.data(), max);
    for (int64_t i = 0; i < k; ++i) {
        tickets[i] = v[i].ticket;
        states[i] = v[i].state;
        statuses[i] = v[i].status;
        users[i] = v[i].user;
    }
    return k;
}
// Admits a batch; writes up to cap tickets and output offsets, returns the count.
int sc_admit(void* h, int64_t now, int64_t* tickets, int64_t* offs, int cap) {
    std::vector<bs::serve::Admitted> b;
    static_cast<Core*>(h)->admit(now, &b);
    int i = 0;
    for (const auto& r : b) {
        if (i < cap) {
            tickets[i] = r.ticket;
            offs[i] = (int64_t)r.out_off;
        }
        ++i;
    }
    return i;
}
void sc_complete(void* h, const int64_t* tickets, int n, int ok, int64_t now) {
    std::vector<int64_t> v(tickets, tickets + n);
    static_cast<Core*>(h)->complete(v, ok != 0, now);
}
void sc_fail_queued(void* h, int status, int64_t now) {
    static_cast<Core*>(h)->fail_queued(status, now);
}
int sc_release_result(void* h, int64_t t) { return static_cast<Core*>(h)->release_result(t); }
int sc_state(void* h, int64_t t, int* state, int* status) {
    auto* r = static_cast<Core*>(h)->request(t);
    if (!r) return bs::serve::kErrNotFound;
    *state = r->state;
    *status = r->status;
    return 0;
}
int64_t sc_arena_used(void* h) { return (int64_t)static_cast<Core*>(h)->arena().used(); }
int64_t sc_arena_peak(void* h) { return (int64_t)static_cast<Core*>(h)->arena().peak(); }
int64_t sc_arena_largest_free(void* h) {
    return (int64_t)static_cast<Core*>(h)->arena().largest_free();
}
int64_t sc_queued(void* h) { return (int64_t)static_cast<Core*>(h)->queued(); }
void sc_stats(void* h, int64_t* out9) {
    auto s = static_cast<Core*>(h)->stats();
    int64_t v[9] = {s.books_resident, s.requests_queued, s.requests_running, s.requests_done,
                    s.requests_failed, s.requests_cancelled, s.launches, s.options_priced,
                    s.admission_waits};
    for (int i = 0; i < 9; ++i) out9[i] = v[i];
}
int64_t sc_book_bytes(int64_t n) { return (int64_t)bs::serve::book_bytes(n); }
int64_t sc_output_bytes(int64_t n, int kind) { return (int64_t)bs::serve::output_bytes(n, kind); }

// Arena alone, for allocator property tests.  The arena names blocks by
// handle; the shim maps offsets to handles so tests can free by offset.
struct ArenaBox {
    bs::serve::Arena a;
    std::unordered_map<int64_t, int64_t> live;  // offset -> handle
    explicit ArenaBox(size_t cap) : a(cap) {}
};
void* sa_create(int64_t cap) { return new ArenaBox((size_t)cap); }
void sa_destroy(void* a) { delete static_cast<ArenaBox*>(a); }
int sa_alloc(void* a, int64_t bytes, int64_t* off) {
    auto* x = static_cast<ArenaBox*>(a);
    size_t o = 0;
    const int64_t h = x->a.alloc((size_t)bytes, &o);
    *off = (int64_t)o;
    if (h < 0) return 0;
    x->live[(int64_t)o] = h;
    return 1;
}
int sa_free(void* a, int64_t off) {
    auto* x = static_cast<ArenaBox*>(a);
    auto it = x->live.find(off);
    if (it == x->live.end()) return 0;
    const bool ok = x->a.free(it->second);
    x->live.erase(it);
    return ok ? 1 : 0;
}
int64_t sa_used(void* a) { return (int64_t)static_cast<ArenaBox*>(a)->a.used(); }
int64_t sa_largest_free(void* a) { return (int64_t)static_cast<ArenaBox*>(a)->a.largest_free(); }
}
