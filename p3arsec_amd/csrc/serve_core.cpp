// serve_core.cpp -- arena, book/request tables and admission of the serving
// loop (see serve_core.hpp).  Pure C++: no GPU calls.
#include "serve_core.hpp"

#include <algorithm>

namespace bs {
namespace serve {

Arena::Arena(size_t capacity) : capacity_(capacity & ~(kAlign - 1)) {
    if (capacity_ > 0) free_[0] = capacity_;
}

bool Arena::alloc(size_t bytes, size_t* off) {
    const size_t need = align_up(bytes == 0 ? 1 : bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < need) continue;
        const size_t o = it->first, sz = it->second;
        free_.erase(it);
        if (sz > need) free_[o + need] = sz - need;
        live_[o] = need;
        used_ += need;
        peak_ = std::max(peak_, used_);
        *off = o;
        return true;
    }
    return false;
}

bool Arena::free(size_t off) {
    auto lv = live_.find(off);
    if (lv == live_.end()) return false;
    size_t o = off, sz = lv->second;
    live_.erase(lv);
    used_ -= sz;
    auto next = free_.lower_bound(o);
    if (next != free_.end() && o + sz == next->first) {
        sz += next->second;
        next = free_.erase(next);
    }
    if (next != free_.begin()) {
        auto prev = std::prev(next);
        if (prev->first + prev->second == o) {
            prev->second += sz;
            return true;
        }
    }
    free_[o] = sz;
    return true;
}

size_t Arena::largest_free() const {
    size_t m = 0;
    for (const auto& f : free_) m = std::max(m, f.second);
    return m;
}

void book_layout(int64_t n, size_t offs[6]) {
    const size_t fb = align_up((size_t)n * 4);
    for (int i = 0; i < 5; ++i) offs[i] = fb * i;
    offs[5] = fb * 5;
}

size_t book_bytes(int64_t n) { return align_up((size_t)n * 4) * 5 + align_up((size_t)n); }

void output_layout(int64_t n, int kind, size_t offs[5]) {
    const size_t fb = align_up((size_t)n * 4);
    for (int i = 0; i < 5; ++i) offs[i] = kind == kGreeks ? fb * i : 0;
}

size_t output_bytes(int64_t n, int kind) {
    return align_up((size_t)n * 4) * (kind == kGreeks ? 5 : 1);
}

int Core::register_book(int64_t n, int64_t* id, size_t* off) {
    if (n <= 0) return kErrInvalid;
    size_t o;
    if (!arena_.alloc(book_bytes(n), &o)) return kErrBudget;
    BookRec b;
    b.id = next_book_++;
    b.n = n;
    b.off = o;
    books_[b.id] = b;
    *id = b.id;
    *off = o;
    ++st_.books_resident;
    return kOk;
}

int Core::release_book(int64_t id) {
    auto it = books_.find(id);
    if (it == books_.end()) return kErrNotFound;
    if (it->second.refs > 0) return kErrBusy;
    arena_.free(it->second.off);
    books_.erase(it);
    --st_.books_resident;
    return kOk;
}

const BookRec* Core::book(int64_t id) const {
    auto it = books_.find(id);
    return it == books_.end() ? nullptr : &it->second;
}

int Core::check_submit(int64_t book, int kind) const {
    if (kind != kPrice && kind != kGreeks) return kErrInvalid;
    auto it = books_.find(book);
    if (it == books_.end()) return kErrNotFound;
    // A request whose outputs cannot fit next to its own (pinned) book is
    // refused now instead of waiting for memory that will never come.
    if (output_bytes(it->second.n, kind) + book_bytes(it->second.n) > arena_.capacity())
        return kErrBudget;
    return kOk;
}

int Core::submit_many(const int64_t* books, const int* kinds, int64_t count, int64_t now,
                      int64_t* tickets) {
    if (count < 0 || (count > 0 && (!books || !kinds || !tickets))) return kErrInvalid;
    for (int64_t i = 0; i < count; ++i) {
        const int rc = check_submit(books[i], kinds[i]);
        if (rc != kOk) return rc;
    }
    for (int64_t i = 0; i < count; ++i) submit(books[i], kinds[i], now, &tickets[i]);
    return kOk;
}

int64_t Core::reap(Completion* out, int64_t max) {
    int64_t k = 0;
    while (k < max && !completed_.empty()) {
        const int64_t t = completed_.front();
        completed_.pop_front();
        auto it = reqs_.find(t);
        if (it == reqs_.end()) continue;  // released before it was reaped
        const Request& r = it->second;
        Completion& c = out[k++];
        c.ticket = r.ticket;
        c.state = r.state;
        c.status = r.status;
        c.t_submit = r.t_submit;
        c.t_done = r.t_done;
    }
    return k;
}

int Core::submit(int64_t book, int kind, int64_t now, int64_t* ticket) {
    const int rc = check_submit(book, kind);
    if (rc != kOk) return rc;
    auto it = books_.find(book);
    Request r;
    r.ticket = next_ticket_++;
    r.book = book;
    r.kind = kind;
    r.n = it->second.n;
    r.t_submit = now;
    reqs_[r.ticket] = r;
    pending_.push_back(r.ticket);
    pending_options_ += r.n;
    ++it->second.refs;
    ++st_.requests_queued;
    *ticket = r.ticket;
    return kOk;
}

int Core::cancel(int64_t ticket, int64_t now) {
    auto it = reqs_.find(ticket);
    if (it == reqs_.end()) return kErrNotFound;
    Request& r = it->second;
    if (r.state != kQueued) return kErrBusy;
    pending_.erase(std::find(pending_.begin(), pending_.end(), ticket));
    pending_options_ -= r.n;
    r.state = kCancelled;
    r.status = kErrCancelled;
    r.t_done = now;
    books_[r.book].refs--;
    --st_.requests_queued;
    ++st_.requests_cancelled;
    finished(r);
    return kOk;
}

std::vector<Request*> Core::admit(int64_t now, int64_t option_cap) {
    std::vector<Request*> batch;
    int64_t options = 0;
    const int64_t cap =
        option_cap > 0 ? std::min(option_cap, lim_.max_batch_options) : lim_.max_batch_options;
    while (!pending_.empty() && (int64_t)batch.size() < lim_.max_batch_requests) {
        Request& r = reqs_[pending_.front()];
        if (!batch.empty() && options + r.n > cap) break;
        size_t o;
        if (!arena_.alloc(output_bytes(r.n, r.kind), &o)) {
            ++st_.admission_waits;  // strict FIFO: wait for memory
            break;
        }
        pending_.pop_front();
        pending_options_ -= r.n;
        r.out_off = o;
        r.has_out = true;
        r.state = kRunning;
        r.t_launch = now;
        st_.queue_ns += now - r.t_submit;
        options += r.n;
        batch.push_back(&r);
        --st_.requests_queued;
        ++st_.requests_running;
    }
    if (!batch.empty()) {
        ++st_.launches;
        st_.options_priced += options;
    }
    return batch;
}

void Core::complete(const std::vector<int64_t>& tickets, bool ok, int64_t now) {
    for (int64_t t : tickets) {
        auto it = reqs_.find(t);
        if (it == reqs_.end()) continue;
        Request& r = it->second;
        if (r.state != kRunning) continue;
        r.t_done = now;
        --st_.requests_running;
        books_[r.book].refs--;
        if (ok) {
            r.state = kDone;
            ++st_.requests_done;
        } else {
            r.state = kFailed;
            r.status = kErrFailed;
            if (r.has_out) arena_.free(r.out_off);
            r.has_out = false;
            ++st_.requests_failed;
        }
        finished(r);
    }
}

void Core::fail_queued(int status, int64_t now) {
    while (!pending_.empty()) {
        Request& r = reqs_[pending_.front()];
        pending_.pop_front();
        pending_options_ -= r.n;
        r.state = kFailed;
        r.status = status;
        r.t_done = now;
        books_[r.book].refs--;
        --st_.requests_queued;
        ++st_.requests_failed;
        finished(r);
    }
}

int Core::release_result(int64_t ticket) {
    auto it = reqs_.find(ticket);
    if (it == reqs_.end()) return kErrNotFound;
    Request& r = it->second;
    if (r.state == kQueued || r.state == kRunning) return kErrBusy;
    if (r.has_out) arena_.free(r.out_off);
    reqs_.erase(it);
    return kOk;
}

Request* Core::request(int64_t ticket) {
    auto it = reqs_.find(ticket);
    return it == reqs_.end() ? nullptr : &it->second;
}

Stats Core::stats() const { return st_; }

}  // namespace serve
}  // namespace bs
