// serve_core.cpp -- arena, book/request tables and admission of the serving
// loop (see serve_core.hpp).  Pure C++: no GPU calls.
#include "serve_core.hpp"

#include <algorithm>
#include <thread>

namespace bs {
namespace serve {

// ---- Arena (TLSF) ----------------------------------------------------------

Arena::Arena(size_t capacity) : capacity_(capacity & ~(kAlign - 1)) {
    for (auto& row : head_)
        for (auto& h : row) h = -1;
    if (capacity_ > 0) {
        const int32_t b = new_record();
        blk_[b].off = 0;
        blk_[b].units = capacity_ / kAlign;
        insert_free(b);
    }
}

// Size class of a block of `units`: classes 0..15 of the first row hold one
// size each; above that, row fl splits [2^(fl+3), 2^(fl+4)) into 16 classes.
void Arena::mapping(uint64_t units, int* fl, int* sl) {
    if (units < (uint64_t)kSL) {
        *fl = 0;
        *sl = (int)units;
        return;
    }
    const int l = 63 - __builtin_clzll(units);
    *fl = l - 3;
    *sl = (int)(units >> (l - 4)) - kSL;
}

int32_t Arena::new_record() {
    if (!spare_.empty()) {
        const int32_t b = spare_.back();
        spare_.pop_back();
        blk_[b] = Block();
        return b;
    }
    blk_.emplace_back();
    return (int32_t)blk_.size() - 1;
}

void Arena::insert_free(int32_t b) {
    int fl, sl;
    mapping(blk_[b].units, &fl, &sl);
    blk_[b].free = true;
    blk_[b].prev_free = -1;
    blk_[b].next_free = head_[fl][sl];
    if (head_[fl][sl] >= 0) blk_[head_[fl][sl]].prev_free = b;
    head_[fl][sl] = b;
    fl_map_ |= uint64_t(1) << fl;
    sl_map_[fl] |= 1u << sl;
}

void Arena::remove_free(int32_t b) {
    int fl, sl;
    mapping(blk_[b].units, &fl, &sl);
    const int32_t p = blk_[b].prev_free, n = blk_[b].next_free;
    if (p >= 0) blk_[p].next_free = n;
    else head_[fl][sl] = n;
    if (n >= 0) blk_[n].prev_free = p;
    if (head_[fl][sl] < 0) {
        sl_map_[fl] &= ~(1u << sl);
        if (sl_map_[fl] == 0) fl_map_ &= ~(uint64_t(1) << fl);
    }
    blk_[b].free = false;
    blk_[b].prev_free = blk_[b].next_free = -1;
}

int64_t Arena::alloc(size_t bytes, size_t* off) {
    const uint64_t units = align_up(bytes == 0 ? 1 : bytes) / kAlign;
    int32_t b = -1;
    // O(1) good fit: the first non-empty class at or above units rounded up
    // to its class's upper edge, whose every block is large enough
    uint64_t r = units;
    if (r >= (uint64_t)kSL) r += (uint64_t(1) << (63 - __builtin_clzll(r) - 4)) - 1;
    int fl, sl;
    mapping(r, &fl, &sl);
    if (fl < kFL) {
        uint32_t sm = sl_map_[fl] & (~0u << sl);
        if (sm == 0) {
            const uint64_t fm = fl + 1 < 64 ? fl_map_ & (~uint64_t(0) << (fl + 1)) : 0;
            if (fm != 0) {
                fl = __builtin_ctzll(fm);
                sm = sl_map_[fl];
            }
        }
        if (sm != 0) b = head_[fl][__builtin_ctz(sm)];
    }
    if (b < 0) {
        // the request's own class may still hold a block that fits
        mapping(units, &fl, &sl);
        for (int32_t x = head_[fl][sl]; x >= 0; x = blk_[x].next_free)
            if (blk_[x].units >= units) {
                b = x;
                break;
            }
        if (b < 0) return -1;
    }
    remove_free(b);
    if (blk_[b].units > units) {
        const int32_t rest = new_record();  // may move blk_: index, not reference
        blk_[rest].off = blk_[b].off + units;
        blk_[rest].units = blk_[b].units - units;
        blk_[rest].prev_phys = b;
        blk_[rest].next_phys = blk_[b].next_phys;
        if (blk_[b].next_phys >= 0) blk_[blk_[b].next_phys].prev_phys = rest;
        blk_[b].next_phys = rest;
        blk_[b].units = units;
        insert_free(rest);
    }
    blk_[b].live = true;
    used_ += units * kAlign;
    peak_ = std::max(peak_, used_);
    *off = blk_[b].off * kAlign;
    return b;
}

bool Arena::free(int64_t handle) {
    if (handle < 0 || handle >= (int64_t)blk_.size() || !blk_[handle].live) return false;
    int32_t b = (int32_t)handle;
    blk_[b].live = false;
    used_ -= blk_[b].units * kAlign;
    const int32_t n = blk_[b].next_phys;
    if (n >= 0 && blk_[n].free) {
        remove_free(n);
        blk_[b].units += blk_[n].units;
        blk_[b].next_phys = blk_[n].next_phys;
        if (blk_[b].next_phys >= 0) blk_[blk_[b].next_phys].prev_phys = b;
        spare_.push_back(n);
    }
    const int32_t p = blk_[b].prev_phys;
    if (p >= 0 && blk_[p].free) {
        remove_free(p);
        blk_[p].units += blk_[b].units;
        blk_[p].next_phys = blk_[b].next_phys;
        if (blk_[p].next_phys >= 0) blk_[blk_[p].next_phys].prev_phys = p;
        spare_.push_back(b);
        b = p;
    }
    insert_free(b);
    return true;
}

size_t Arena::largest_free() const {
    if (fl_map_ == 0) return 0;
    const int fl = 63 - __builtin_clzll(fl_map_);
    const int sl = 31 - __builtin_clz(sl_map_[fl]);
    uint64_t m = 0;
    for (int32_t x = head_[fl][sl]; x >= 0; x = blk_[x].next_free) m = std::max(m, blk_[x].units);
    return m * kAlign;
}

// ---- layouts ----------------------------------------------------------------

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

// ---- request table ----------------------------------------------------------

Request* Core::slot_of(int64_t ticket) {
    if (ticket <= 0) return nullptr;
    const int64_t slot = ticket & ((int64_t(1) << kSlotBits) - 1);
    if (slot >= nslots_) return nullptr;
    Request* r = &chunks_[slot / kChunk][slot % kChunk];
    return r->ticket == ticket ? r : nullptr;
}

int64_t Core::new_ticket(Request** r) {
    int64_t slot;
    if (!spare_slots_.empty()) {
        slot = spare_slots_.back();
        spare_slots_.pop_back();
    } else {
        if (nslots_ >= (int64_t(1) << kSlotBits)) return 0;
        slot = nslots_++;
        if (slot / kChunk >= (int64_t)chunks_.size())
            chunks_.emplace_back(new Request[kChunk]);
    }
    Request* q = &chunks_[slot / kChunk][slot % kChunk];
    *q = Request();
    q->ticket = (next_serial_++ << kSlotBits) | slot;
    *r = q;
    return q->ticket;
}

void Core::free_slot(Request* r) {
    const int64_t slot = r->ticket & ((int64_t(1) << kSlotBits) - 1);
    r->ticket = 0;
    spare_slots_.push_back((int32_t)slot);
}

// ---- books and requests -----------------------------------------------------

int Core::register_book(int64_t n, int64_t* id, size_t* off) {
    if (n <= 0) return kErrInvalid;
    size_t o;
    const int64_t h = arena_.alloc(book_bytes(n), &o);
    if (h < 0) return kErrBudget;
    BookRec b;
    b.id = next_book_++;
    b.n = n;
    b.off = o;
    b.blk = h;
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
    arena_.free(it->second.blk);
    books_.erase(it);
    --st_.books_resident;
    return kOk;
}

const BookRec* Core::book(int64_t id) const {
    auto it = books_.find(id);
    return it == books_.end() ? nullptr : &it->second;
}

int Core::check_submit(int64_t book, int kind, BookRec** rec) {
    if (kind != kPrice && kind != kGreeks) return kErrInvalid;
    auto it = books_.find(book);
    if (it == books_.end()) return kErrNotFound;
    // A request whose outputs cannot fit next to its own (pinned) book is
    // refused now instead of waiting for memory that will never come.
    if (output_bytes(it->second.n, kind) + book_bytes(it->second.n) > arena_.capacity())
        return kErrBudget;
    *rec = &it->second;
    return kOk;
}

// Queues a request on a checked book (b from check_submit) with a ticket from
// new_ticket.
void Core::enqueue(BookRec* b, int kind, int64_t now, int64_t user, Request* r) {
    r->bookp = b;
    r->user = user;
    r->kind = (int8_t)kind;
    r->n = b->n;
    r->lane = (int8_t)lane_of(r->n);
    r->t_submit = now;
    pending_[r->lane].push_back(r->ticket);
    pending_options_[r->lane] += r->n;
    ++b->refs;
    ++st_.requests_queued;
}

int Core::submit(int64_t book, int kind, int64_t now, int64_t* ticket, int64_t user) {
    BookRec* b;
    const int rc = check_submit(book, kind, &b);
    if (rc != kOk) return rc;
    Request* r;
    if (new_ticket(&r) == 0) return kErrNoMem;
    enqueue(b, kind, now, user, r);
    *ticket = r->ticket;
    return kOk;
}

int Core::submit_many(const int64_t* books, const int* kinds, const int64_t* users,
                      int64_t count, int64_t now, int64_t* tickets) {
    if (count < 0 || (count > 0 && (!books || !kinds || !tickets))) return kErrInvalid;
    // Book records live in a node-based map, so the pointers stay valid while
    // this call (which holds the owner's lock) runs.
    scratch_books_.resize((size_t)count);
    for (int64_t i = 0; i < count; ++i) {
        const int rc = check_submit(books[i], kinds[i], &scratch_books_[(size_t)i]);
        if (rc != kOk) return rc;
    }
    if (nslots_ - (int64_t)spare_slots_.size() + count > (int64_t(1) << kSlotBits))
        return kErrNoMem;
    for (int64_t i = 0; i < count; ++i) {
        Request* r;
        new_ticket(&r);   // cannot fail: the slots were counted above
        enqueue(scratch_books_[(size_t)i], kinds[i], now, users ? users[i] : 0, r);
        tickets[i] = r->ticket;
    }
    return kOk;
}

int Core::cancel(int64_t ticket, int64_t now) {
    Request* r = slot_of(ticket);
    if (!r) return kErrNotFound;
    if (r->state != kQueued) return kErrBusy;
    std::deque<int64_t>& q = pending_[r->lane];
    q.erase(std::find(q.begin(), q.end(), ticket));
    pending_options_[r->lane] -= r->n;
    r->state = kCancelled;
    r->status = kErrCancelled;
    r->t_done = now;
    r->bookp->refs--;
    --st_.requests_queued;
    ++st_.requests_cancelled;
    finished(*r);
    return kOk;
}

size_t Core::admit(int64_t now, std::vector<Admitted>* out, int64_t option_cap, int lane) {
    out->clear();
    if (lane != kLaneSmall && lane != kLaneBulk) return 0;
    std::deque<int64_t>& pending = pending_[lane];
    int64_t options = 0;
    const int64_t cap =
        option_cap > 0 ? std::min(option_cap, lim_.max_batch_options) : lim_.max_batch_options;
    size_t ahead = 0;   // pending[0, ahead) are prefetched
    while (!pending.empty() && (int64_t)out->size() < lim_.max_batch_requests) {
        for (; ahead < pending.size() && ahead < 8; ++ahead) prefetch(pending[ahead]);
        Request* r = slot_of(pending.front());
        if (!out->empty() && options + r->n > cap) break;
        size_t o;
        const int64_t h = arena_.alloc(output_bytes(r->n, r->kind), &o);
        if (h < 0) {
            ++st_.admission_waits;  // strict FIFO: wait for memory
            break;
        }
        pending.pop_front();
        --ahead;
        pending_options_[lane] -= r->n;
        r->out_off = o;
        r->out_blk = (int32_t)h;
        r->state = kRunning;
        st_.queue_ns += now - r->t_submit;
        options += r->n;
        out->push_back({r->ticket, r->n, r->bookp->off, o, r->bookp->n, r->kind});
        --st_.requests_queued;
        ++st_.requests_running;
    }
    if (!out->empty()) {
        ++st_.launches;
        st_.options_priced += options;
    }
    return out->size();
}

void Core::complete(const std::vector<int64_t>& tickets, bool ok, int64_t now) {
    for (size_t i = 0; i < tickets.size(); ++i) {
        if (i + 8 < tickets.size()) prefetch(tickets[i + 8]);
        const int64_t t = tickets[i];
        Request* r = slot_of(t);
        if (!r || r->state != kRunning) continue;
        r->t_done = now;
        --st_.requests_running;
        r->bookp->refs--;
        if (ok) {
            r->state = kDone;
            ++st_.requests_done;
        } else {
            r->state = kFailed;
            r->status = kErrFailed;
            if (r->out_blk >= 0) arena_.free(r->out_blk);
            r->out_blk = -1;
            ++st_.requests_failed;
        }
        finished(*r);
    }
}

void Core::fail_queued(int status, int64_t now) {
    for (int lane = 0; lane < kLanes; ++lane) {
        while (!pending_[lane].empty()) {
            Request* r = slot_of(pending_[lane].front());
            pending_[lane].pop_front();
            pending_options_[lane] -= r->n;
            r->state = kFailed;
            r->status = (int8_t)status;
            r->t_done = now;
            r->bookp->refs--;
            --st_.requests_queued;
            ++st_.requests_failed;
            finished(*r);
        }
    }
}

int Core::release_result(int64_t ticket) {
    Request* r = slot_of(ticket);
    if (!r) return kErrNotFound;
    if (r->state == kQueued || r->state == kRunning) return kErrBusy;
    if (r->out_blk >= 0) arena_.free(r->out_blk);
    free_slot(r);
    return kOk;
}

Request* Core::request(int64_t ticket) { return slot_of(ticket); }

// About 0.5 ms of pause-hinted spinning (a hand-over under load comes within
// microseconds), then yields so a waiter whose turn is far off gives up its CPU.
void TicketLock::wait_turn(uint32_t me) {
    constexpr int kSpins = 20000;
    for (int i = 0; serving_.load(std::memory_order_acquire) != me; ++i) {
        if (i < kSpins) __builtin_ia32_pause();
        else std::this_thread::yield();
    }
}

int64_t Core::reap(Completion* out, int64_t max) {
    int64_t k = 0;
    size_t ahead = 0;   // completed_[0, ahead) are prefetched
    while (k < max && !completed_.empty()) {
        for (; ahead < completed_.size() && ahead < 8; ++ahead) prefetch(completed_[ahead]);
        const int64_t t = completed_.front();
        completed_.pop_front();
        --ahead;
        const Request* r = slot_of(t);
        if (!r) continue;  // released before it was reaped
        Completion& c = out[k++];
        c.ticket = r->ticket;
        c.user = r->user;
        c.n = r->n;
        c.state = r->state;
        c.status = r->status;
        c.t_submit = r->t_submit;
        c.t_done = r->t_done;
    }
    return k;
}

Stats Core::stats() const { return st_; }

}  // namespace serve
}  // namespace bs
