// bs_serve.hip -- the serving loop: resident books, continuous batching,
// fused launches of every queued request.
//
// A context owns
//   * one device arena (hipMalloc of the HBM budget) holding resident books
//     and request outputs, carved by serve::Core's first-fit allocator;
//   * four descriptor slots, each a pinned host table plus its device copy;
//   * per batch, a stage kernel that copies the slot's table from pinned host
//     memory to the device, then one fused kernel per section (the batch's
//     price requests, its Greeks requests): a block prices one tile of 512
//     options of one request, found by a 64-ary wave search of the table.
//     Launched directly (the default) each section's grid is one block per
//     tile; with use_graph the three kernels are captured once per slot over
//     a fixed persistent grid and each batch is one hipGraphLaunch (on ROCm
//     7.2 the direct form measured 15 us launch-to-completion against 31 us);
//   * a dispatcher thread: it admits queued requests as arena memory allows,
//     writes a table into a free slot and launches it behind the batch still
//     running (up to four batches in flight), and retires a batch when its event
//     completes.  Requests that arrive while a batch runs go into the next one.
// Results are bit-identical to bs_price_f32 / bs_price_greeks_f32: the fused
// kernels run the batch map's own per-group code (bs_ops.hpp).
#include <stddef.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "bs_internal.hpp"
#include "bs_math.hpp"
#include "bs_ops.hpp"
#include "serve_core.hpp"

namespace bs {
namespace {

constexpr int kVW = 2;                 // options per work-item (the batch map's width)
constexpr int kTile = kThreads * kVW;  // options per tile: one block step
constexpr int kMaxDesc = 4096;         // requests per batch: 64 x 64 for find_desc
constexpr int64_t kMaxGrid = int64_t(1) << 20;  // direct-launch grid cap (blocks loop)
constexpr int kStageBlocks = 32;
// Batches in flight at once.  While one runs, the next ones are queued behind
// it on the stream, so the GPU stays busy while clients turn finished
// requests around into new ones.
constexpr int kSlots = 4;
constexpr int kReapSpinUs = 200;
// With fewer than two batches on the GPU and more than this many options
// queued, the dispatcher admits half of the queue so that the other half's
// launch (and the clients' turnaround of the first half) overlaps the first.
constexpr int64_t kSplitOptions = int64_t(4) << 20;

// One request of a batch: 112 bytes, a multiple of the stage kernel's 16-byte words.
struct Desc {
    const float* in[5];   // s, k, r, v, t
    const uint8_t* type;
    float* out[5];        // price, delta, vega, rho, theta (Greeks only use 1..4)
    int64_t n;
    int64_t tile_begin;   // first tile of this request within its section
    int32_t kind;
    int32_t pad;
};
static_assert(sizeof(Desc) == 112, "Desc layout");

// A batch's table: the header, then the price requests' descriptors, then the
// Greeks requests' descriptors.  Section 0 is price, section 1 Greeks.
struct SlotHeader {
    int32_t ndesc[2];
    int64_t tiles[2];
    int64_t seq;
    int64_t pad[4];
};
static_assert(sizeof(SlotHeader) == 64, "SlotHeader layout");

constexpr size_t kSlotBytes = sizeof(SlotHeader) + sizeof(Desc) * kMaxDesc;

// Copies header + descriptors from the pinned table to device memory.
__global__ __launch_bounds__(kThreads) void k_stage(const uint4* __restrict__ host,
                                                    uint4* __restrict__ dev) {
    const SlotHeader* h = reinterpret_cast<const SlotHeader*>(host);
    const int np = h->ndesc[0], ng = h->ndesc[1];
    const int ndesc = (np < 0 || ng < 0 || np + ng > kMaxDesc) ? 0 : np + ng;
    const int64_t words = (int64_t)(sizeof(SlotHeader) + sizeof(Desc) * ndesc) / 16;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < words;
         i += (int64_t)gridDim.x * kThreads)
        dev[i] = host[i];
}

// The request of tile t among d[0, nd): the last descriptor whose tile_begin
// <= t.  A 64-ary search over tile_begin (increasing, d[0].tile_begin == 0):
// each wave resolves 64 x 64 = 4096 descriptors with two loads per lane.
__device__ __forceinline__ int find_desc(const Desc* d, int nd, int64_t t) {
    const int lane = threadIdx.x & 63;
    int j = lane * 64;
    const int base = (__popcll(__ballot(j < nd && d[j].tile_begin <= t)) - 1) * 64;
    j = base + lane;
    return base + __popcll(__ballot(j < nd && d[j].tile_begin <= t)) - 1;
}

template <class Op>
__device__ __forceinline__ Op op_of(const Desc& x);
template <>
__device__ __forceinline__ PriceF32<Ref32, kVW> op_of(const Desc& x) {
    return {SoaIn<float>{x.in[0], x.in[1], x.in[2], x.in[3], x.in[4]}, x.type, x.out[0]};
}
template <>
__device__ __forceinline__ GreeksF32<kVW> op_of(const Desc& x) {
    return {SoaIn<float>{x.in[0], x.in[1], x.in[2], x.in[3], x.in[4]}, x.type,
            GreeksOut{x.out[0], x.out[1], x.out[2], x.out[3], x.out[4]}};
}

// Fused kernel of one section: a block prices tile t (kTile options of one
// request, one kVW-group per work-item) for t = blockIdx.x, + gridDim.x, ...
// Launched directly, the grid has one block per tile; replayed from a graph it
// is a fixed persistent grid.  Every access is bounded by the request's n.
// The per-group code is the batch map's own (bs_ops.hpp).
template <class Op, int kSection>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(waves_for<Op, false>())))
void k_serve(const uint8_t* __restrict__ slot) {
    const SlotHeader* h = reinterpret_cast<const SlotHeader*>(slot);
    const int np = h->ndesc[0], ng = h->ndesc[1];
    if (np < 0 || ng < 0 || np + ng > kMaxDesc) return;
    const int nd = kSection == 0 ? np : ng;
    if (nd <= 0) return;
    const Desc* d = reinterpret_cast<const Desc*>(slot + sizeof(SlotHeader)) + (kSection ? np : 0);
    const int64_t tiles = h->tiles[kSection];
    for (int64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const Desc& x = d[find_desc(d, nd, t)];
        const Op op = op_of<Op>(x);
        const int64_t n = x.n;
        const int64_t g = (t - x.tile_begin) * kThreads + threadIdx.x;
        if ((g + 1) * kVW <= n) {
            op.template finish<true>(g, op.template load<true>(g));
        } else {
            for (int64_t i = g * kVW; i < n; ++i) op.scalar(i);
        }
    }
}

using ServePrice = PriceF32<Ref32, kVW>;
using ServeGreeks = GreeksF32<kVW>;

int64_t now_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct Batch {
    int slot;
    hipEvent_t done;
    int64_t t_launch;
    std::vector<int64_t> tickets;
};

}  // namespace
}  // namespace bs

// Locking: `mu` guards `core`, `fatal` and `batch_ns`.  The dispatcher takes it
// only to admit a batch or retire one; while batches run it polls their
// events with the lock released, so submitting and reaping threads never wait
// behind the poll loop.  `gen` counts the changes that can unblock admission
// (a submit, a freed result or book, a cancel, shutdown); `queued` mirrors
// the queue length.  Both are written under `mu` and read without it.
struct bs_ctx_s {
    int device = 0;
    bs_serve_opts opts{};
    std::mutex mu;
    std::condition_variable cv_work;   // dispatcher: new requests / memory / shutdown
    std::condition_variable cv_done;   // clients: ticket state changes
    bs::serve::Core* core = nullptr;
    char* arena = nullptr;
    hipStream_t stream = nullptr;      // capture stream; batches alternate over streams[]
    hipStream_t streams[2] = {nullptr, nullptr};
    uint8_t* host_slot[bs::kSlots] = {};      // pinned, written by the dispatcher
    uint8_t* host_slot_dev[bs::kSlots] = {};  // the same memory, device address
    uint8_t* dev_slot[bs::kSlots] = {};
    hipGraphExec_t graph[bs::kSlots] = {};
    hipEvent_t ev[bs::kSlots] = {};
    int grid = 0;
    int64_t seq = 0;
    int64_t batch_ns = 0;              // sum of launch -> completion seen, over batches
    std::atomic<int64_t> gen{0};
    std::atomic<int64_t> queued{0};
    std::atomic<int64_t> finished{0};  // requests finished so far (done, failed, cancelled)
    // Dispatcher time accounting (written by the dispatcher only).
    std::atomic<int64_t> t_admit{0}, t_launch{0}, t_retire{0}, t_idle{0}, idle_launches{0};
    std::atomic<bool> shutdown{false};
    std::thread worker;
    std::string fatal;

    // Called under `mu` after any change to the queue or the arena.
    void changed() {
        queued.store((int64_t)core->queued(), std::memory_order_release);
        gen.fetch_add(1, std::memory_order_acq_rel);
    }
};

using namespace bs;

namespace {

static_assert(serve::kErrBudget == BS_ERR_BUDGET && serve::kErrBusy == BS_ERR_BUSY &&
                  serve::kErrNotFound == BS_ERR_NOT_FOUND &&
                  serve::kErrShutdown == BS_ERR_SHUTDOWN && serve::kDone == BS_TICKET_DONE &&
                  serve::kCancelled == BS_TICKET_CANCELLED && serve::kGreeks == BS_REQ_GREEKS,
              "serve_core codes mirror bs_mi355.h");

// Launches one slot on stream st: graph replay, or the stage kernel and the
// fused kernel of each non-empty section directly, each section's grid one
// block per tile.
hipError_t launch_slot(bs_ctx_s* c, int s, hipStream_t st) {
    if (c->graph[s]) return hipGraphLaunch(c->graph[s], st);
    const SlotHeader* h = reinterpret_cast<const SlotHeader*>(c->host_slot[s]);
    hipLaunchKernelGGL(k_stage, dim3(kStageBlocks), dim3(kThreads), 0, st,
                       reinterpret_cast<const uint4*>(c->host_slot_dev[s]),
                       reinterpret_cast<uint4*>(c->dev_slot[s]));
    if (h->tiles[0] > 0)
        hipLaunchKernelGGL((k_serve<ServePrice, 0>), dim3((unsigned)std::min(h->tiles[0], kMaxGrid)),
                           dim3(kThreads), 0, st, c->dev_slot[s]);
    if (h->tiles[1] > 0)
        hipLaunchKernelGGL((k_serve<ServeGreeks, 1>), dim3((unsigned)std::min(h->tiles[1], kMaxGrid)),
                           dim3(kThreads), 0, st, c->dev_slot[s]);
    return hipGetLastError();
}

// Captures slot s as one graph: stage, price section, Greeks section, each
// fused kernel over the fixed persistent grid c->grid.
hipError_t capture_slot(bs_ctx_s* c, int s) {
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_stage, dim3(kStageBlocks), dim3(kThreads), 0, c->stream,
                       reinterpret_cast<const uint4*>(c->host_slot_dev[s]),
                       reinterpret_cast<uint4*>(c->dev_slot[s]));
    hipLaunchKernelGGL((k_serve<ServePrice, 0>), dim3(c->grid), dim3(kThreads), 0, c->stream,
                       c->dev_slot[s]);
    hipLaunchKernelGGL((k_serve<ServeGreeks, 1>), dim3(c->grid), dim3(kThreads), 0, c->stream,
                       c->dev_slot[s]);
    e = hipStreamEndCapture(c->stream, &g);
    if (e != hipSuccess) return e;
    e = hipGraphInstantiate(&c->graph[s], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    return e;
}

// Fills slot s with the batch's descriptors, price requests first, each
// section in admission order.  Runs without the lock: it reads only the
// admitted requests' copies, and the slot is the dispatcher's own.
void write_slot(bs_ctx_s* c, int s, const std::vector<serve::Admitted>& reqs) {
    SlotHeader* h = reinterpret_cast<SlotHeader*>(c->host_slot[s]);
    Desc* d = reinterpret_cast<Desc*>(c->host_slot[s] + sizeof(SlotHeader));
    int nd[2] = {0, 0};
    for (const serve::Admitted& r : reqs) ++nd[r.kind == serve::kGreeks ? 1 : 0];
    int next[2] = {0, nd[0]};
    int64_t tiles[2] = {0, 0};
    for (const serve::Admitted& r : reqs) {
        const int sec = r.kind == serve::kGreeks ? 1 : 0;
        size_t bo[6], oo[5];
        serve::book_layout(r.book_n, bo);
        serve::output_layout(r.n, r.kind, oo);
        Desc& x = d[next[sec]++];
        for (int j = 0; j < 5; ++j) x.in[j] = reinterpret_cast<const float*>(c->arena + r.book_off + bo[j]);
        x.type = reinterpret_cast<const uint8_t*>(c->arena + r.book_off + bo[5]);
        for (int j = 0; j < 5; ++j) x.out[j] = reinterpret_cast<float*>(c->arena + r.out_off + oo[j]);
        x.n = r.n;
        x.tile_begin = tiles[sec];
        x.kind = r.kind;
        x.pad = 0;
        tiles[sec] += (r.n + kTile - 1) / kTile;
    }
    h->ndesc[0] = nd[0];
    h->ndesc[1] = nd[1];
    h->tiles[0] = tiles[0];
    h->tiles[1] = tiles[1];
    h->seq = ++c->seq;
}

void dispatcher(bs_ctx_s* c) {
    (void)hipSetDevice(c->device);
    std::vector<Batch> inflight;        // batches on the GPU; owned by this thread
    bool busy[kSlots] = {};
    int64_t blocked_gen = -1;           // gen at the last admission that found no room
    int64_t t_empty = -1;               // when the last batch in flight retired
    int64_t nlaunch = 0;                // batches alternate over the two streams
    std::vector<serve::Admitted> reqs;  // reused: no per-batch allocation
    reqs.reserve(kMaxDesc);
    for (;;) {
        bool progressed = false;
        // retire every finished batch (batches on the two streams may finish
        // out of launch order; each is independent of the others)
        for (size_t bi = 0; bi < inflight.size();) {
            Batch& b = inflight[bi];
            const hipError_t q = hipEventQuery(b.done);
            if (q == hipErrorNotReady) {
                ++bi;
                continue;
            }
            const int64_t t = now_ns();
            {
                std::lock_guard<std::mutex> lk(c->mu);
                if (q != hipSuccess) c->fatal = std::string("batch failed: ") + hipGetErrorString(q);
                c->core->complete(b.tickets, q == hipSuccess, t);
                c->batch_ns += t - b.t_launch;
                c->changed();
                c->finished.fetch_add((int64_t)b.tickets.size(), std::memory_order_release);
            }
            c->cv_done.notify_all();
            busy[b.slot] = false;
            inflight.erase(inflight.begin() + (std::ptrdiff_t)bi);
            const int64_t t2 = now_ns();
            c->t_retire.fetch_add(t2 - t, std::memory_order_relaxed);
            if (inflight.empty()) t_empty = t2;
            blocked_gen = -1;   // a failed batch frees its outputs
            progressed = true;
        }
        // launch the next batch into a free slot
        int s = -1;
        for (int j = 0; j < kSlots && s < 0; ++j)
            if (!busy[j]) s = j;
        if (s >= 0 && !c->shutdown.load(std::memory_order_acquire) &&
            c->queued.load(std::memory_order_acquire) > 0 &&
            c->gen.load(std::memory_order_acquire) != blocked_gen) {
            Batch b{s, c->ev[s], 0, {}};
            const int64_t ta = now_ns();
            {
                std::lock_guard<std::mutex> lk(c->mu);
                const int64_t g = c->gen.load(std::memory_order_relaxed);
                const int64_t qo = c->core->queued_options();
                c->core->admit(now_ns(), &reqs, inflight.size() < 2 && qo > kSplitOptions ? qo / 2 : 0);
                if (reqs.empty()) blocked_gen = g;   // the queue head does not fit: wait for a change
                else c->changed();
            }
            if (!reqs.empty()) {
                write_slot(c, s, reqs);
                b.tickets.reserve(reqs.size());
                for (const auto& r : reqs) b.tickets.push_back(r.ticket);
                b.t_launch = now_ns();
                c->t_admit.fetch_add(b.t_launch - ta, std::memory_order_relaxed);
                if (inflight.empty() && t_empty >= 0) {
                    c->t_idle.fetch_add(b.t_launch - t_empty, std::memory_order_relaxed);
                    c->idle_launches.fetch_add(1, std::memory_order_relaxed);
                }
                hipStream_t st = c->streams[nlaunch++ & 1];
                hipError_t e = launch_slot(c, s, st);
                if (e == hipSuccess) e = hipEventRecord(c->ev[s], st);
                c->t_launch.fetch_add(now_ns() - b.t_launch, std::memory_order_relaxed);
                if (e != hipSuccess) {
                    {
                        std::lock_guard<std::mutex> lk(c->mu);
                        c->fatal = std::string("launch failed: ") + hipGetErrorString(e);
                        c->core->complete(b.tickets, false, now_ns());
                        c->changed();
                        c->finished.fetch_add((int64_t)b.tickets.size(), std::memory_order_release);
                    }
                    c->cv_done.notify_all();
                } else {
                    busy[s] = true;
                    inflight.push_back(std::move(b));
                }
                progressed = true;
            }
        }
        if (progressed) continue;
        if (!inflight.empty()) {
            std::this_thread::yield();   // a batch is on the GPU: keep polling its event
            continue;
        }
        if (c->shutdown.load(std::memory_order_acquire)) break;
        std::unique_lock<std::mutex> lk(c->mu);
        c->cv_work.wait(lk, [&] {
            return c->shutdown.load(std::memory_order_relaxed) ||
                   (c->core->queued() > 0 && c->gen.load(std::memory_order_relaxed) != blocked_gen);
        });
    }
    {
        std::lock_guard<std::mutex> lk(c->mu);
        c->core->fail_queued(serve::kErrShutdown, now_ns());
        c->changed();
        c->finished.fetch_add(1, std::memory_order_release);
    }
    c->cv_done.notify_all();
}

}  // namespace

extern "C" {

int bs_ctx_create(int device, size_t hbm_budget_bytes, const bs_serve_opts* opts, bs_ctx_t* out) {
    if (!out) return set_error(BS_ERR_INVALID, "ctx is NULL");
    *out = nullptr;
    if (hbm_budget_bytes < serve::kAlign * 16)
        return set_error(BS_ERR_INVALID, "hbm_budget_bytes too small");
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || device < 0 || device >= ndev)
        return set_error(BS_ERR_NODEVICE, "device index out of range");
    auto* c = new bs_ctx_s();
    c->device = device;
    c->opts.max_batch_requests = kMaxDesc;
    c->opts.max_batch_options = int64_t(1) << 31;
    c->opts.use_graph = 0;
    c->opts.grid_blocks = 0;
    if (opts) {
        if (opts->max_batch_requests > 0)
            c->opts.max_batch_requests = std::min<int64_t>(opts->max_batch_requests, kMaxDesc);
        if (opts->max_batch_options > 0) c->opts.max_batch_options = opts->max_batch_options;
        c->opts.use_graph = opts->use_graph;
        c->opts.grid_blocks = opts->grid_blocks;
    }
    int rc = BS_OK;
    e = hipSetDevice(device);
    if (e == hipSuccess) e = hipMalloc((void**)&c->arena, hbm_budget_bytes);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    for (int j = 0; j < 2 && e == hipSuccess; ++j)
        e = hipStreamCreateWithFlags(&c->streams[j], hipStreamNonBlocking);
    for (int s = 0; s < kSlots && e == hipSuccess; ++s) {
        e = hipHostMalloc((void**)&c->host_slot[s], kSlotBytes, hipHostMallocCoherent);
        if (e == hipSuccess)
            e = hipHostGetDevicePointer((void**)&c->host_slot_dev[s], c->host_slot[s], 0);
        if (e == hipSuccess) e = hipMalloc((void**)&c->dev_slot[s], kSlotBytes);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev[s], hipEventDisableTiming);
        if (e == hipSuccess) ::memset(c->host_slot[s], 0, sizeof(SlotHeader));
    }
    if (e == hipSuccess) {
        int cus = 0;
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
        c->grid = c->opts.grid_blocks > 0 ? c->opts.grid_blocks : std::max(1, cus) * 8;
    }
    for (int s = 0; s < kSlots && e == hipSuccess && c->opts.use_graph; ++s) e = capture_slot(c, s);
    if (e != hipSuccess) {
        rc = hip_error(e, "bs_ctx_create");
        bs_ctx_destroy(c);
        return rc;
    }
    serve::Limits lim;
    lim.max_batch_requests = c->opts.max_batch_requests;
    lim.max_batch_options = c->opts.max_batch_options;
    c->core = new serve::Core(hbm_budget_bytes, lim);
    c->core->enable_completions(opts && opts->completion_queue != 0);
    c->worker = std::thread(dispatcher, c);
    *out = c;
    return BS_OK;
}

int bs_ctx_destroy(bs_ctx_t ctx) {
    bs_ctx_s* c = ctx;
    if (!c) return set_error(BS_ERR_INVALID, "ctx is NULL");
    if (c->worker.joinable()) {
        {
            std::lock_guard<std::mutex> lk(c->mu);
            c->shutdown.store(true);
            c->gen.fetch_add(1);
        }
        c->cv_work.notify_all();
        c->worker.join();
    }
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (hipStream_t st : c->streams)
        if (st) (void)hipStreamSynchronize(st);
    for (int s = 0; s < kSlots; ++s) {
        if (c->graph[s]) (void)hipGraphExecDestroy(c->graph[s]);
        if (c->ev[s]) (void)hipEventDestroy(c->ev[s]);
        if (c->host_slot[s]) (void)hipHostFree(c->host_slot[s]);
        if (c->dev_slot[s]) (void)hipFree(c->dev_slot[s]);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    for (hipStream_t st : c->streams)
        if (st) (void)hipStreamDestroy(st);
    if (c->arena) (void)hipFree(c->arena);
    delete c->core;
    delete c;
    return BS_OK;
}

int bs_book_register(bs_ctx_t ctx, const float* s, const float* k, const float* r, const float* v,
                     const float* t, const uint8_t* otype, int64_t n, bs_book_t* book) {
    bs_ctx_s* c = ctx;
    if (!c || !book) return set_error(BS_ERR_INVALID, "ctx or book is NULL");
    if (n <= 0) return set_error(BS_ERR_INVALID, "n <= 0");
    if (!s || !k || !r || !v || !t || !otype) return set_error(BS_ERR_INVALID, "NULL array");
    int64_t id;
    size_t off;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        if (c->shutdown.load()) return set_error(BS_ERR_SHUTDOWN, "context shutting down");
        const int rc = c->core->register_book(n, &id, &off);
        if (rc != BS_OK) return set_error(rc, "book does not fit in the arena");
    }
    size_t bo[6];
    serve::book_layout(n, bo);
    const void* src[6] = {s, k, r, v, t, otype};
    hipError_t e = hipSetDevice(c->device);
    for (int j = 0; j < 6 && e == hipSuccess; ++j)
        e = hipMemcpy(c->arena + off + bo[j], src[j], (size_t)n * (j < 5 ? 4 : 1),
                      hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        std::lock_guard<std::mutex> lk(c->mu);
        c->core->release_book(id);
        c->changed();
        return hip_error(e, "bs_book_register");
    }
    *book = id;
    return BS_OK;
}

int bs_book_release(bs_ctx_t ctx, bs_book_t book) {
    bs_ctx_s* c = ctx;
    if (!c) return set_error(BS_ERR_INVALID, "ctx is NULL");
    int rc;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        rc = c->core->release_book(book);
        if (rc == BS_OK) c->changed();
    }
    if (rc == BS_OK) c->cv_work.notify_all();
    return rc == BS_OK ? BS_OK : set_error(rc, "bs_book_release");
}

int bs_submit(bs_ctx_t ctx, bs_book_t book, int kind, bs_ticket_t* ticket) {
    bs_ctx_s* c = ctx;
    if (!c || !ticket) return set_error(BS_ERR_INVALID, "ctx or ticket is NULL");
    int rc;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        if (c->shutdown.load()) return set_error(BS_ERR_SHUTDOWN, "context shutting down");
        rc = c->core->submit(book, kind, now_ns(), ticket);
        if (rc == BS_OK) c->changed();
    }
    if (rc != BS_OK) return set_error(rc, "bs_submit");
    c->cv_work.notify_one();
    return BS_OK;
}

int bs_submit_many(bs_ctx_t ctx, const bs_book_t* books, const int* kinds, const int64_t* users,
                   int64_t count, bs_ticket_t* tickets) {
    bs_ctx_s* c = ctx;
    if (!c) return set_error(BS_ERR_INVALID, "ctx is NULL");
    if (count < 0 || (count > 0 && (!books || !kinds || !tickets)))
        return set_error(BS_ERR_INVALID, "bad request arrays");
    if (count == 0) return BS_OK;
    int rc;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        if (c->shutdown.load()) return set_error(BS_ERR_SHUTDOWN, "context shutting down");
        rc = c->core->submit_many(books, kinds, users, count, now_ns(), tickets);
        if (rc == BS_OK) c->changed();
    }
    if (rc != BS_OK) return set_error(rc, "bs_submit_many");
    c->cv_work.notify_one();
    return BS_OK;
}

int bs_reap(bs_ctx_t ctx, bs_completion* out, int64_t max, int64_t timeout_us, int64_t* count) {
    static_assert(sizeof(bs_completion) == sizeof(serve::Completion) &&
                      offsetof(bs_completion, user) == offsetof(serve::Completion, user) &&
                      offsetof(bs_completion, state) == offsetof(serve::Completion, state) &&
                      offsetof(bs_completion, done_ns) == offsetof(serve::Completion, t_done),
                  "bs_completion layout");
    bs_ctx_s* c = ctx;
    if (!c || !count) return set_error(BS_ERR_INVALID, "ctx or count is NULL");
    *count = 0;
    if (max < 0 || (max > 0 && !out)) return set_error(BS_ERR_INVALID, "bad output array");
    auto* dst = reinterpret_cast<serve::Completion*>(out);
    const auto t0 = std::chrono::steady_clock::now();
    const auto deadline = t0 + std::chrono::microseconds(timeout_us);
    std::unique_lock<std::mutex> lk(c->mu);
    if (!c->core->completions_enabled())
        return set_error(BS_ERR_INVALID, "context has no completion queue");
    if (max == 0) return BS_OK;
    for (;;) {
        const int64_t seen = c->finished.load(std::memory_order_acquire);
        const int64_t k = c->core->reap(dst, max);
        if (k > 0 || timeout_us == 0) {
            *count = k;
            return BS_OK;
        }
        // Spin for up to kReapSpinUs on the finished counter without the lock
        // (a batch completes every few hundred microseconds under load, and a
        // sleeping reaper costs a scheduler wake-up per batch), then sleep.
        lk.unlock();
        const auto now = std::chrono::steady_clock::now();
        const auto spin_end = std::min(now + std::chrono::microseconds(kReapSpinUs),
                                       timeout_us < 0 ? now + std::chrono::hours(1) : deadline);
        while (c->finished.load(std::memory_order_acquire) == seen &&
               std::chrono::steady_clock::now() < spin_end)
            std::this_thread::yield();
        lk.lock();
        if (c->finished.load(std::memory_order_acquire) != seen ||
            c->core->completions_pending() > 0)
            continue;
        auto ready = [&] { return c->core->completions_pending() > 0; };
        if (timeout_us < 0) {
            c->cv_done.wait(lk, ready);
        } else if (!c->cv_done.wait_until(lk, deadline, ready)) {
            return BS_OK;
        }
    }
}

int bs_result_release_many(bs_ctx_t ctx, const bs_ticket_t* tickets, int64_t count) {
    bs_ctx_s* c = ctx;
    if (!c) return set_error(BS_ERR_INVALID, "ctx is NULL");
    if (count < 0 || (count > 0 && !tickets)) return set_error(BS_ERR_INVALID, "bad ticket array");
    int first = BS_OK;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        for (int64_t i = 0; i < count; ++i) {
            const int rc = c->core->release_result(tickets[i]);
            if (rc != BS_OK && first == BS_OK) first = rc;
        }
        c->changed();
    }
    c->cv_work.notify_all();
    return first == BS_OK ? BS_OK : set_error(first, "bs_result_release_many");
}

int bs_poll(bs_ctx_t ctx, bs_ticket_t ticket, int* state) {
    bs_ctx_s* c = ctx;
    if (!c || !state) return set_error(BS_ERR_INVALID, "ctx or state is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    const serve::Request* r = c->core->request(ticket);
    if (!r) return set_error(BS_ERR_NOT_FOUND, "unknown ticket");
    *state = r->state;
    return BS_OK;
}

int bs_wait(bs_ctx_t ctx, bs_ticket_t ticket, int64_t timeout_us) {
    bs_ctx_s* c = ctx;
    if (!c) return set_error(BS_ERR_INVALID, "ctx is NULL");
    std::unique_lock<std::mutex> lk(c->mu);
    auto finished = [&]() {
        const serve::Request* r = c->core->request(ticket);
        return !r || (r->state != serve::kQueued && r->state != serve::kRunning);
    };
    if (timeout_us < 0) {
        c->cv_done.wait(lk, finished);
    } else if (!c->cv_done.wait_for(lk, std::chrono::microseconds(timeout_us), finished)) {
        return set_error(BS_ERR_TIMEOUT, "bs_wait timed out");
    }
    const serve::Request* r = c->core->request(ticket);
    if (!r) return set_error(BS_ERR_NOT_FOUND, "unknown ticket");
    if (r->state == serve::kDone) return BS_OK;
    if (r->state == serve::kCancelled) return set_error(BS_ERR_CANCELLED, "ticket cancelled");
    return set_error(r->status == BS_OK ? BS_ERR_FAILED : r->status,
                     c->fatal.empty() ? "request failed" : c->fatal);
}

int bs_result_device(bs_ctx_t ctx, bs_ticket_t ticket, const float** price, const float** delta,
                     const float** vega, const float** rho, const float** theta) {
    bs_ctx_s* c = ctx;
    if (!c) return set_error(BS_ERR_INVALID, "ctx is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    const serve::Request* r = c->core->request(ticket);
    if (!r) return set_error(BS_ERR_NOT_FOUND, "unknown ticket");
    if (r->state != serve::kDone) return set_error(BS_ERR_INVALID, "ticket not done");
    size_t oo[5];
    serve::output_layout(r->n, r->kind, oo);
    const float** dst[5] = {price, delta, vega, rho, theta};
    for (int j = 0; j < 5; ++j) {
        if (!dst[j]) continue;
        *dst[j] = (j == 0 || r->kind == serve::kGreeks)
                      ? reinterpret_cast<const float*>(c->arena + r->out_off + oo[j])
                      : nullptr;
    }
    return BS_OK;
}

int bs_result_copy(bs_ctx_t ctx, bs_ticket_t ticket, float* price, float* delta, float* vega,
                   float* rho, float* theta) {
    const float* src[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    int64_t n = 0;
    int kind = 0;
    bs_ctx_s* c = ctx;
    if (!c) return set_error(BS_ERR_INVALID, "ctx is NULL");
    {
        std::lock_guard<std::mutex> lk(c->mu);
        const serve::Request* r = c->core->request(ticket);
        if (!r) return set_error(BS_ERR_NOT_FOUND, "unknown ticket");
        if (r->state != serve::kDone) return set_error(BS_ERR_INVALID, "ticket not done");
        n = r->n;
        kind = r->kind;
    }
    int rc = bs_result_device(ctx, ticket, &src[0], &src[1], &src[2], &src[3], &src[4]);
    if (rc != BS_OK) return rc;
    float* dst[5] = {price, delta, vega, rho, theta};
    hipError_t e = hipSetDevice(c->device);
    for (int j = 0; j < 5 && e == hipSuccess; ++j) {
        if (!dst[j]) continue;
        if (j > 0 && kind != serve::kGreeks) return set_error(BS_ERR_INVALID, "not a Greeks ticket");
        e = hipMemcpy(dst[j], src[j], (size_t)n * 4, hipMemcpyDeviceToHost);
    }
    return e == hipSuccess ? BS_OK : hip_error(e, "bs_result_copy");
}

int bs_result_release(bs_ctx_t ctx, bs_ticket_t ticket) {
    bs_ctx_s* c = ctx;
    if (!c) return set_error(BS_ERR_INVALID, "ctx is NULL");
    int rc;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        rc = c->core->release_result(ticket);
        if (rc == BS_OK) c->changed();
    }
    if (rc == BS_OK) c->cv_work.notify_all();
    return rc == BS_OK ? BS_OK : set_error(rc, "bs_result_release");
}

int bs_cancel(bs_ctx_t ctx, bs_ticket_t ticket) {
    bs_ctx_s* c = ctx;
    if (!c) return set_error(BS_ERR_INVALID, "ctx is NULL");
    int rc;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        rc = c->core->cancel(ticket, now_ns());
        if (rc == BS_OK) {
            c->changed();
            c->finished.fetch_add(1, std::memory_order_release);
        }
    }
    if (rc == BS_OK) {
        c->cv_done.notify_all();
        c->cv_work.notify_all();   // the cancelled request may have blocked the queue head
    }
    return rc == BS_OK ? BS_OK : set_error(rc, "bs_cancel: ticket not queued");
}

int bs_ticket_times(bs_ctx_t ctx, bs_ticket_t ticket, int64_t* submit_ns, int64_t* done_ns) {
    bs_ctx_s* c = ctx;
    if (!c) return set_error(BS_ERR_INVALID, "ctx is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    const serve::Request* r = c->core->request(ticket);
    if (!r) return set_error(BS_ERR_NOT_FOUND, "unknown ticket");
    if (submit_ns) *submit_ns = r->t_submit;
    if (done_ns) *done_ns = r->t_done;
    return BS_OK;
}

int bs_ctx_stats(bs_ctx_t ctx, bs_serve_stats* out) {
    bs_ctx_s* c = ctx;
    if (!c || !out) return set_error(BS_ERR_INVALID, "ctx or stats is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    const serve::Stats st = c->core->stats();
    out->arena_bytes = (int64_t)c->core->arena().capacity();
    out->arena_used = (int64_t)c->core->arena().used();
    out->arena_peak = (int64_t)c->core->arena().peak();
    out->books_resident = st.books_resident;
    out->requests_queued = st.requests_queued;
    out->requests_running = st.requests_running;
    out->requests_done = st.requests_done;
    out->requests_failed = st.requests_failed;
    out->requests_cancelled = st.requests_cancelled;
    out->launches = st.launches;
    out->options_priced = st.options_priced;
    out->admission_waits = st.admission_waits;
    out->queue_ns_total = st.queue_ns;
    out->batch_ns_total = c->batch_ns;
    out->dispatch_admit_ns = c->t_admit.load();
    out->dispatch_launch_ns = c->t_launch.load();
    out->dispatch_retire_ns = c->t_retire.load();
    out->gpu_idle_ns = c->t_idle.load();
    out->idle_launches = c->idle_launches.load();
    return BS_OK;
}

}  // extern "C"
