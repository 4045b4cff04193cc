// serve_loop.hpp -- the serving loop's concurrency, free of HIP.
//
// Loop<Exec> owns the bookkeeping (serve::Core), the client-facing calls
// (submit, wait, reap, release, cancel) and the dispatcher thread that admits
// queued requests into batches and hands them to an executor.  The executor
// is the only part that touches the device: bs_serve.hip supplies the HIP one
// (descriptor tables, streams, fused kernels, events); the CPU tests supply a
// host one that completes batches after a delay and can fail launches or
// batches on demand, so the loop runs under ThreadSanitizer without a GPU.
//
// Executor interface:
//   static constexpr int kSlots;                       batches in flight at most
//   static constexpr int kSmallSlots;                  slots [0, kSmallSlots) carry the
//                                                      small lane's batches, the rest
//                                                      the bulk lane's
//   void thread_start();                               on the dispatcher thread
//   int launch(int slot, const std::vector<Admitted>&, std::string* err);   0 = ok
//   int query(int slot, std::string* err);             0 running, 1 done, -1 failed
//   void quiesce(int slot);                            after a failed launch: waits
//                                                      until nothing the launch
//                                                      enqueued before failing can
//                                                      still run
//   void drain();                                      waits for every launched batch
//
// Lanes: with Limits::small_options > 0, requests of at most that many
// options queue in the small lane and are batched only with each other, in
// the small lane's slots, so their completion does not wait for a large
// request's batch; the executor runs the two lanes on different streams.
// Each lane admits in FIFO order.
//
// Locking: `mu_` (a FIFO ticket lock, serve_core.hpp) guards the Core,
// `fatal_` and `batch_ns_`.  The dispatcher
// takes it only to admit a batch or retire one; while batches run it queries
// them with the lock released, so submitting and reaping threads never wait
// behind the poll loop.  `gen_` counts the changes that can unblock admission
// (a submit, a freed result or book, a cancel, shutdown); `queued_` mirrors
// each lane's queue length; `finished_` counts finished requests (reapers
// spin on it).  All three are written under `mu_` and read without it.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "serve_core.hpp"

namespace bs {
namespace serve {

inline int64_t steady_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// Dispatcher time accounting (ns) and the launches that ended an idle stretch.
struct LoopTimes {
    int64_t admit = 0, launch = 0, retire = 0, idle = 0, idle_launches = 0, batch = 0;
};

template <class Exec>
class Loop {
   public:
    // With fewer than two bulk batches in flight and more than this many
    // options queued in the bulk lane, the dispatcher admits half of that
    // queue so that the other half's launch (and the clients' turnaround of
    // the first half) overlaps the first.
    static constexpr int64_t kSplitOptions = int64_t(4) << 20;
    // A reaper with nothing to reap spins this long on `finished_` before it
    // sleeps: under load a batch finishes every few hundred microseconds, and a
    // sleeping reaper costs a scheduler wake-up per batch.
    static constexpr int64_t kReapSpinUs = 200;

    Loop(Exec* ex, size_t arena_bytes, Limits lim, bool completions)
        : ex_(ex), core_(arena_bytes, lim) {
        core_.enable_completions(completions);
    }
    ~Loop() { stop(); }
    Loop(const Loop&) = delete;
    Loop& operator=(const Loop&) = delete;

    void start() { worker_ = std::thread([this] { dispatcher(); }); }

    // Finishes the batches in flight, fails the queued requests with
    // kErrShutdown and joins the dispatcher.  Idempotent.
    void stop() {
        if (!worker_.joinable()) return;
        {
            std::lock_guard<TicketLock> lk(mu_);
            shutdown_.store(true);
            gen_.fetch_add(1);
        }
        cv_work_.notify_all();
        worker_.join();
    }
    bool shutting_down() const { return shutdown_.load(); }

    // ---- books ---------------------------------------------------------------
    int register_book(int64_t n, int64_t* id, size_t* off) {
        std::lock_guard<TicketLock> lk(mu_);
        if (shutdown_.load()) return kErrShutdown;
        return core_.register_book(n, id, off);
    }
    int release_book(int64_t id) {
        int rc;
        {
            std::lock_guard<TicketLock> lk(mu_);
            rc = core_.release_book(id);
            if (rc == kOk) changed();
        }
        if (rc == kOk) cv_work_.notify_all();
        return rc;
    }

    // ---- requests ------------------------------------------------------------
    int submit(int64_t book, int kind, int64_t* ticket) {
        int rc;
        {
            std::lock_guard<TicketLock> lk(mu_);
            if (shutdown_.load()) return kErrShutdown;
            rc = core_.submit(book, kind, steady_ns(), ticket);
            if (rc == kOk) changed();
        }
        if (rc == kOk) cv_work_.notify_one();
        return rc;
    }
    int submit_many(const int64_t* books, const int* kinds, const int64_t* users, int64_t count,
                    int64_t* tickets) {
        if (count == 0) return kOk;
        int rc;
        {
            std::lock_guard<TicketLock> lk(mu_);
            if (shutdown_.load()) return kErrShutdown;
            rc = core_.submit_many(books, kinds, users, count, steady_ns(), tickets);
            if (rc == kOk) changed();
        }
        if (rc == kOk) cv_work_.notify_one();
        return rc;
    }
    int poll(int64_t ticket, int* state) {
        std::lock_guard<TicketLock> lk(mu_);
        const Request* r = core_.request(ticket);
        if (!r) return kErrNotFound;
        *state = r->state;
        return kOk;
    }
    // Waits until the ticket is no longer queued or running (timeout_us < 0:
    // forever).  Returns kOk for a done request, else its failure; *msg gets
    // the executor's error text for a failed batch.
    int wait(int64_t ticket, int64_t timeout_us, std::string* msg) {
        std::unique_lock<TicketLock> lk(mu_);
        auto finished = [&]() {
            const Request* r = core_.request(ticket);
            return !r || (r->state != kQueued && r->state != kRunning);
        };
        if (timeout_us < 0) {
            cv_done_.wait(lk, finished);
        } else if (!cv_done_.wait_for(lk, std::chrono::microseconds(timeout_us), finished)) {
            return kErrTimeout;
        }
        const Request* r = core_.request(ticket);
        if (!r) return kErrNotFound;
        if (r->state == kDone) return kOk;
        if (r->state == kCancelled) return kErrCancelled;
        if (msg) *msg = fatal_.empty() ? "request failed" : fatal_;
        return r->status == kOk ? kErrFailed : r->status;
    }
    // Pops up to max finished requests (completion queue), waiting up to
    // timeout_us for the first one (0 polls, < 0 waits forever).
    int reap(Completion* out, int64_t max, int64_t timeout_us, int64_t* count) {
        *count = 0;
        const auto deadline =
            std::chrono::steady_clock::now() + std::chrono::microseconds(timeout_us);
        std::unique_lock<TicketLock> lk(mu_);
        if (!core_.completions_enabled()) return kErrInvalid;
        if (max == 0) return kOk;
        for (;;) {
            const int64_t seen = finished_.load(std::memory_order_acquire);
            const int64_t k = core_.reap(out, max);
            if (k > 0 || timeout_us == 0) {
                *count = k;
                return kOk;
            }
            lk.unlock();
            const auto now = std::chrono::steady_clock::now();
            const auto spin_end = std::min(now + std::chrono::microseconds(kReapSpinUs),
                                           timeout_us < 0 ? now + std::chrono::hours(1) : deadline);
            while (finished_.load(std::memory_order_acquire) == seen &&
                   std::chrono::steady_clock::now() < spin_end)
                std::this_thread::yield();
            lk.lock();
            if (finished_.load(std::memory_order_acquire) != seen || core_.completions_pending() > 0)
                continue;
            auto ready = [&] { return core_.completions_pending() > 0; };
            if (timeout_us < 0) {
                cv_done_.wait(lk, ready);
            } else if (!cv_done_.wait_until(lk, deadline, ready)) {
                return kOk;
            }
        }
    }
    int release(int64_t ticket) {
        int rc;
        {
            std::lock_guard<TicketLock> lk(mu_);
            rc = core_.release_result(ticket);
            if (rc == kOk) changed();
        }
        if (rc == kOk) cv_work_.notify_all();
        return rc;
    }
    // Every ticket is tried; returns the first failure or kOk.
    int release_many(const int64_t* tickets, int64_t count) {
        int first = kOk;
        {
            std::lock_guard<TicketLock> lk(mu_);
            for (int64_t i = 0; i < count; ++i) {
                if (i + 8 < count) core_.prefetch(tickets[i + 8]);
                const int rc = core_.release_result(tickets[i]);
                if (rc != kOk && first == kOk) first = rc;
            }
            changed();
        }
        cv_work_.notify_all();
        return first;
    }
    int cancel(int64_t ticket) {
        int rc;
        {
            std::lock_guard<TicketLock> lk(mu_);
            rc = core_.cancel(ticket, steady_ns());
            if (rc == kOk) {
                changed();
                finished_.fetch_add(1, std::memory_order_release);
            }
        }
        if (rc == kOk) {
            cv_done_.notify_all();
            cv_work_.notify_all();   // the cancelled request may have blocked the queue head
        }
        return rc;
    }
    int times(int64_t ticket, int64_t* submit_ns, int64_t* done_ns) {
        std::lock_guard<TicketLock> lk(mu_);
        const Request* r = core_.request(ticket);
        if (!r) return kErrNotFound;
        if (submit_ns) *submit_ns = r->t_submit;
        if (done_ns) *done_ns = r->t_done;
        return kOk;
    }
    // A done request's size, kind and output offset (for result access).
    int result_info(int64_t ticket, int64_t* n, int* kind, size_t* out_off) {
        std::lock_guard<TicketLock> lk(mu_);
        const Request* r = core_.request(ticket);
        if (!r) return kErrNotFound;
        if (r->state != kDone) return kErrInvalid;
        *n = r->n;
        *kind = r->kind;
        *out_off = r->out_off;
        return kOk;
    }
    void stats(Stats* st, size_t* cap, size_t* used, size_t* peak, LoopTimes* t) {
        std::lock_guard<TicketLock> lk(mu_);
        *st = core_.stats();
        *cap = core_.arena().capacity();
        *used = core_.arena().used();
        *peak = core_.arena().peak();
        t->admit = t_admit_.load();
        t->launch = t_launch_.load();
        t->retire = t_retire_.load();
        t->idle = t_idle_.load();
        t->idle_launches = idle_launches_.load();
        t->batch = batch_ns_;
    }

   private:
    struct Batch {
        int slot;
        int64_t t_launch;
        std::vector<int64_t> tickets;
    };

    // Called under `mu_` after any change to the queue or the arena.
    void changed() {
        for (int lane = 0; lane < kLanes; ++lane)
            queued_[lane].store((int64_t)core_.queued(lane), std::memory_order_release);
        gen_.fetch_add(1, std::memory_order_acq_rel);
    }

    // Marks a batch finished (ok or failed) and wakes the clients.
    void finish_batch(const std::vector<int64_t>& tickets, bool ok, const std::string& err,
                      int64_t t_launch) {
        const int64_t t = steady_ns();
        {
            std::lock_guard<TicketLock> lk(mu_);
            if (!ok) fatal_ = err;
            core_.complete(tickets, ok, t);
            if (t_launch >= 0) batch_ns_ += t - t_launch;
            changed();
            finished_.fetch_add((int64_t)tickets.size(), std::memory_order_release);
        }
        cv_done_.notify_all();
    }

    void dispatcher() {
        static_assert(Exec::kSmallSlots >= 0 && Exec::kSmallSlots < Exec::kSlots,
                      "the bulk lane needs a slot");
        ex_->thread_start();
        std::vector<Batch> inflight;        // batches on the device
        bool busy[Exec::kSlots] = {};
        int64_t blocked_gen[kLanes] = {-1, -1};   // gen at the lane's last admission that found no room
        int64_t t_empty = -1;               // when the last batch in flight retired
        std::vector<Admitted> reqs;         // reused: no per-batch allocation
        std::string err;
        for (;;) {
            bool progressed = false;
            // retire every finished batch (batches may finish out of launch
            // order; each is independent of the others)
            for (size_t bi = 0; bi < inflight.size();) {
                const int q = ex_->query(inflight[bi].slot, &err);
                if (q == 0) {
                    ++bi;
                    continue;
                }
                const int64_t t = steady_ns();
                finish_batch(inflight[bi].tickets, q > 0, "batch failed: " + err, inflight[bi].t_launch);
                busy[inflight[bi].slot] = false;
                inflight.erase(inflight.begin() + (std::ptrdiff_t)bi);
                const int64_t t2 = steady_ns();
                t_retire_.fetch_add(t2 - t, std::memory_order_relaxed);
                if (inflight.empty()) t_empty = t2;
                blocked_gen[0] = blocked_gen[1] = -1;   // a failed batch frees its outputs
                progressed = true;
            }
            // launch the next batch of each lane into a free slot of that lane
            for (int lane = 0; lane < kLanes; ++lane) {
                const int lo = lane == kLaneSmall ? 0 : Exec::kSmallSlots;
                const int hi = lane == kLaneSmall ? Exec::kSmallSlots : Exec::kSlots;
                int s = -1;
                for (int j = lo; j < hi && s < 0; ++j)
                    if (!busy[j]) s = j;
                if (s < 0 || shutdown_.load(std::memory_order_acquire) ||
                    queued_[lane].load(std::memory_order_acquire) == 0 ||
                    gen_.load(std::memory_order_acquire) == blocked_gen[lane])
                    continue;
                int bulk_inflight = 0;
                for (const Batch& b : inflight) bulk_inflight += b.slot >= Exec::kSmallSlots;
                const int64_t ta = steady_ns();
                {
                    std::lock_guard<TicketLock> lk(mu_);
                    const int64_t g = gen_.load(std::memory_order_relaxed);
                    const int64_t qo = core_.queued_options(lane);
                    const bool split = lane == kLaneBulk && bulk_inflight < 2 && qo > kSplitOptions;
                    core_.admit(steady_ns(), &reqs, split ? qo / 2 : 0, lane);
                    if (reqs.empty()) blocked_gen[lane] = g;   // the lane's head does not fit
                    else changed();
                }
                if (reqs.empty()) continue;
                Batch b{s, steady_ns(), {}};
                b.tickets.reserve(reqs.size());
                for (const auto& r : reqs) b.tickets.push_back(r.ticket);
                t_admit_.fetch_add(b.t_launch - ta, std::memory_order_relaxed);
                if (inflight.empty() && t_empty >= 0) {
                    t_idle_.fetch_add(b.t_launch - t_empty, std::memory_order_relaxed);
                    idle_launches_.fetch_add(1, std::memory_order_relaxed);
                }
                const int e = ex_->launch(s, reqs, &err);
                t_launch_.fetch_add(steady_ns() - b.t_launch, std::memory_order_relaxed);
                if (e != 0) {
                    // part of the batch may be queued on the device already: it
                    // must be done before its outputs and slot are handed out again
                    ex_->quiesce(s);
                    finish_batch(b.tickets, false, "launch failed: " + err, -1);
                } else {
                    busy[s] = true;
                    inflight.push_back(std::move(b));
                }
                progressed = true;
            }
            if (progressed) continue;
            if (!inflight.empty()) {
                std::this_thread::yield();   // batches on the device: keep querying them
                continue;
            }
            if (shutdown_.load(std::memory_order_acquire)) break;
            std::unique_lock<TicketLock> lk(mu_);
            cv_work_.wait(lk, [&] {
                if (shutdown_.load(std::memory_order_relaxed)) return true;
                const int64_t g = gen_.load(std::memory_order_relaxed);
                for (int lane = 0; lane < kLanes; ++lane)
                    if (core_.queued(lane) > 0 && g != blocked_gen[lane]) return true;
                return false;
            });
        }
        ex_->drain();
        {
            std::lock_guard<TicketLock> lk(mu_);
            core_.fail_queued(kErrShutdown, steady_ns());
            changed();
            finished_.fetch_add(1, std::memory_order_release);
        }
        cv_done_.notify_all();
    }

    Exec* ex_;
    Core core_;
    TicketLock mu_;
    std::condition_variable_any cv_work_;   // dispatcher: new requests / memory / shutdown
    std::condition_variable_any cv_done_;   // clients: ticket state changes
    std::string fatal_;
    int64_t batch_ns_ = 0;              // sum of launch -> completion seen, over batches
    std::atomic<int64_t> gen_{0};
    std::atomic<int64_t> queued_[kLanes] = {};
    std::atomic<int64_t> finished_{0};
    std::atomic<int64_t> t_admit_{0}, t_launch_{0}, t_retire_{0}, t_idle_{0}, idle_launches_{0};
    std::atomic<bool> shutdown_{false};
    std::thread worker_;
};

}  // namespace serve
}  // namespace bs
