// serve_core.hpp -- host-side state of the serving loop, free of HIP.
//
// The serving loop keeps books resident in one HBM arena and prices requests
// against them.  This header holds everything that is bookkeeping: the arena
// allocator, the book and request tables, FIFO admission by memory and batch
// formation.  It has no GPU dependency so it is unit-tested on the CPU; the
// HIP executor (bs_serve.hip) drives it under the context's mutex, so every
// operation here is O(1) per request (no scans, no per-request heap
// allocation in steady state) to keep that mutex's hold times short.
//
// Arena layout (offsets from the arena base, every block 256-byte aligned):
//   book:    s | k | r | v | t (float, n each) | type (uint8, n)
//   request: price (float, n)                      for BS_REQ_PRICE
//            price | delta | vega | rho | theta     for BS_REQ_GREEKS
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <atomic>
#include <deque>
#include <memory>
#include <unordered_map>
#include <vector>

namespace bs {
namespace serve {

// A FIFO ticket lock for the serving loop's state.  The loop's threads
// (clients submitting and releasing, the dispatcher admitting and retiring)
// each hold it for microseconds at a time, thousands of times a second; with
// a sleeping mutex every contended hand-over is a futex wake-up of 5-30 us,
// and at 1.9M requests/s those wake-ups, not the work under the lock, set the
// small requests' latency.  A waiter spins with the CPU's pause hint for a
// while, then yields; hand-overs go in arrival order, so no thread starves.
// Waits that may be long (a client waiting for a ticket, an idle dispatcher)
// sleep on a condition variable instead, which releases the lock.
class TicketLock {
   public:
    void lock() {
        const uint32_t me = next_.fetch_add(1, std::memory_order_relaxed);
        if (serving_.load(std::memory_order_acquire) != me) wait_turn(me);
    }
    void unlock() {   // only the holder writes serving_
        serving_.store(serving_.load(std::memory_order_relaxed) + 1, std::memory_order_release);
    }

   private:
    void wait_turn(uint32_t me);
    alignas(64) std::atomic<uint32_t> next_{0};
    alignas(64) std::atomic<uint32_t> serving_{0};
};

constexpr size_t kAlign = 256;
inline size_t align_up(size_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

// Two-level segregated-fit allocator (TLSF, Masmano et al., ECRTS 2004) over
// [0, capacity) in kAlign units.  Free blocks sit in per-size-class lists
// indexed by two bitmaps: 16 classes per power of two.  alloc() takes the
// first non-empty class whose every block fits (O(1)), else searches the
// request's own class (so it fails only when no free block is large enough);
// free() coalesces with both address neighbours in O(1).  Block records live
// in host memory (the arena itself is device memory) and are named by handle.
class Arena {
   public:
    explicit Arena(size_t capacity);
    // Returns a handle >= 0 and the block's offset, or -1 when nothing fits.
    int64_t alloc(size_t bytes, size_t* off);
    bool free(int64_t handle);  // false if handle is not a live block
    size_t capacity() const { return capacity_; }
    size_t used() const { return used_; }
    size_t peak() const { return peak_; }
    size_t largest_free() const;

   private:
    static constexpr int kSL = 16, kFL = 48;
    struct Block {
        uint64_t off = 0, units = 0;
        int32_t prev_phys = -1, next_phys = -1, prev_free = -1, next_free = -1;
        bool free = false, live = false;
    };
    static void mapping(uint64_t units, int* fl, int* sl);
    int32_t new_record();
    void insert_free(int32_t b);
    void remove_free(int32_t b);
    std::vector<Block> blk_;
    std::vector<int32_t> spare_;  // unused record indices
    int32_t head_[kFL][kSL];
    uint64_t fl_map_ = 0;
    uint32_t sl_map_[kFL] = {};
    size_t capacity_, used_ = 0, peak_ = 0;
};

// Status and state codes match include/bs_mi355.h.
enum : int {
    kOk = 0,
    kErrInvalid = -1,
    kErrNoMem = -3,
    kErrBudget = -4,
    kErrNotFound = -5,
    kErrCancelled = -6,
    kErrTimeout = -7,
    kErrBusy = -8,
    kErrFailed = -10,
    kErrShutdown = -11
};
enum : int { kPrice = 1, kGreeks = 2 };
enum : int { kQueued = 0, kRunning = 1, kDone = 2, kFailed = 3, kCancelled = 4 };
// Admission lanes: requests of at most Limits::small_options options queue in
// the small lane and form batches of their own; the rest in the bulk lane.
enum : int { kLaneSmall = 0, kLaneBulk = 1, kLanes = 2 };

struct BookRec {
    int64_t id = 0;
    int64_t n = 0;
    size_t off = 0;
    int64_t blk = -1;
    int refs = 0;  // queued + running requests
};

// One cache line: a request is written by the submitting thread, the
// dispatcher and the releasing thread in turn, so each hand-over moves one
// line between cores.
struct alignas(64) Request {
    int64_t ticket = 0;        // 0: slot unused
    BookRec* bookp = nullptr;  // stable: a referenced book cannot be released
    int64_t user = 0;
    int64_t n = 0;
    int64_t t_submit = 0, t_done = 0;
    size_t out_off = 0;
    int32_t out_blk = -1;      // arena handle of the outputs
    int8_t kind = kPrice;
    int8_t lane = kLaneBulk;
    int8_t state = kQueued;
    int8_t status = kOk;       // kOk or a (small, negative) error code
};
static_assert(sizeof(Request) == 64, "Request is one cache line");

// What the executor needs of an admitted request to build its descriptor;
// a copy, so the table can be written after the lock is released.
struct Admitted {
    int64_t ticket;
    int64_t n;
    size_t book_off, out_off;
    int64_t book_n;
    int kind;
};

// One finished request, as the completion queue reports it (the layout of
// bs_completion).
struct Completion {
    int64_t ticket = 0;
    int64_t user = 0;
    int64_t n = 0;
    int32_t state = kDone;
    int32_t status = kOk;
    int64_t t_submit = 0, t_done = 0;
};

struct Limits {
    int64_t max_batch_requests = 4096;
    int64_t max_batch_options = int64_t(1) << 30;
    int64_t small_options = 0;  // 0: one lane (every request in the bulk lane)
};

struct Stats {
    int64_t books_resident = 0, requests_queued = 0, requests_running = 0;
    int64_t requests_done = 0, requests_failed = 0, requests_cancelled = 0;
    int64_t launches = 0, options_priced = 0, admission_waits = 0;
    int64_t queue_ns = 0;  // sum over admitted requests of (admission - submit)
};

size_t book_bytes(int64_t n);
size_t output_bytes(int64_t n, int kind);
// Byte offsets of a book's arrays relative to its block: s, k, r, v, t, type.
void book_layout(int64_t n, size_t offs[6]);
// Byte offsets of a request's outputs relative to its block.
void output_layout(int64_t n, int kind, size_t offs[5]);

// Not thread-safe: the owner serialises access.
class Core {
   public:
    Core(size_t arena_bytes, Limits lim) : arena_(arena_bytes), lim_(lim) {}

    int register_book(int64_t n, int64_t* id, size_t* off);
    int release_book(int64_t id);
    const BookRec* book(int64_t id) const;

    int submit(int64_t book, int kind, int64_t now, int64_t* ticket, int64_t user = 0);
    // Submits count requests, all or none: the first invalid one (unknown
    // book, bad kind, outputs that can never fit) fails the call before any
    // is queued.  users may be NULL (tags 0).
    int submit_many(const int64_t* books, const int* kinds, const int64_t* users, int64_t count,
                    int64_t now, int64_t* tickets);
    int cancel(int64_t ticket, int64_t now);
    // Admits the lane's queued requests in FIFO order while their outputs fit
    // in the arena and the batch limits allow; admitted requests become
    // Running and are appended to *out (cleared first).  option_cap > 0
    // lowers the batch's option limit for this call (the first request is
    // always admitted if it fits).  Returns the number admitted.
    size_t admit(int64_t now, std::vector<Admitted>* out, int64_t option_cap = 0,
                 int lane = kLaneBulk);
    // Marks a launched batch finished.  Failed requests release their outputs.
    void complete(const std::vector<int64_t>& tickets, bool ok, int64_t now);
    // Fails every queued request (shutdown).
    void fail_queued(int status, int64_t now);
    int release_result(int64_t ticket);
    Request* request(int64_t ticket);
    // Starts loading a ticket's record into this core's cache (records are
    // written by the other threads of the loop in turn).
    void prefetch(int64_t ticket) const {
        const int64_t slot = ticket & ((int64_t(1) << kSlotBits) - 1);
        if (ticket > 0 && slot < nslots_) __builtin_prefetch(&chunks_[slot / kChunk][slot % kChunk], 1);
    }

    // Completion queue: when enabled, every request that finishes (done,
    // failed or cancelled) is appended in the order it finished; reap() pops
    // up to max of them, skipping tickets already released.
    void enable_completions(bool on) { cq_ = on; }
    bool completions_enabled() const { return cq_; }
    int64_t reap(Completion* out, int64_t max);
    size_t completions_pending() const { return completed_.size(); }

    size_t queued() const { return pending_[0].size() + pending_[1].size(); }
    size_t queued(int lane) const { return pending_[lane].size(); }
    int64_t queued_options(int lane = kLaneBulk) const { return pending_options_[lane]; }
    int lane_of(int64_t n) const {
        return lim_.small_options > 0 && n <= lim_.small_options ? kLaneSmall : kLaneBulk;
    }
    const Arena& arena() const { return arena_; }
    Stats stats() const;

   private:
    // Tickets are (serial << kSlotBits) | slot: O(1) lookup, and a ticket
    // whose slot was reused no longer matches.
    static constexpr int kSlotBits = 24;
    static constexpr int kChunk = 4096;
    Request* slot_of(int64_t ticket);
    int64_t new_ticket(Request** r);
    void free_slot(Request* r);
    int check_submit(int64_t book, int kind, BookRec** rec);
    void enqueue(BookRec* b, int kind, int64_t now, int64_t user, Request* r);
    void finished(const Request& r) {
        if (cq_) completed_.push_back(r.ticket);
    }

    Arena arena_;
    Limits lim_;
    int64_t next_book_ = 1, next_serial_ = 1;
    std::unordered_map<int64_t, BookRec> books_;
    std::vector<std::unique_ptr<Request[]>> chunks_;  // stable request storage
    std::vector<int32_t> spare_slots_;
    int64_t nslots_ = 0;
    std::deque<int64_t> pending_[kLanes];
    std::deque<int64_t> completed_;
    int64_t pending_options_[kLanes] = {0, 0};
    bool cq_ = false;
    Stats st_;
    std::vector<BookRec*> scratch_books_;   // submit_many's checked books
};

}  // namespace serve
}  // namespace bs
