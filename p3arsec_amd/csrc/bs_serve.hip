// bs_serve.hip -- the serving loop: resident books, continuous batching,
// persistent fused launches replayed from captured hipGraphs.
//
// A context owns
//   * one device arena (hipMalloc of the HBM budget) holding resident books
//     and request outputs, carved by serve::Core's first-fit allocator;
//   * two descriptor slots, each a pinned host table plus its device copy;
//   * per slot, one hipGraph: a stage kernel that copies the slot's table from
//     pinned host memory to the device, then the fused persistent kernel that
//     prices every queued request of the batch (price or price + Greeks) over
//     a fixed grid.  The graphs never change shape; a batch's size lives in the
//     table header, so each launch is one hipGraphLaunch;
//   * a dispatcher thread: it admits queued requests as arena memory allows,
//     writes a table into a free slot and launches that slot's graph behind the
//     one still running (two batches in flight), and retires a batch when its
//     event completes.  Requests that arrive while a batch runs go into the
//     next one.
// Results are bit-identical to bs_price_f32 / bs_price_greeks_f32: the fused
// kernel calls the same bs_price<Ref32> and bs_greeks<Ref32>.
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
#include "serve_core.hpp"

namespace bs {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kTile = kThreads * 4;   // options per block step: 256 threads x 4
constexpr int kMaxDesc = 4096;
constexpr int kStageBlocks = 32;

// One request of a batch: 112 bytes, a multiple of the stage kernel's 16-byte words.
struct Desc {
    const float* in[5];   // s, k, r, v, t
    const uint8_t* type;
    float* out[5];        // price, delta, vega, rho, theta (Greeks only use 1..4)
    int64_t n;
    int64_t tile_begin;   // first global tile of this request
    int32_t kind;
    int32_t pad;
};
static_assert(sizeof(Desc) == 112, "Desc layout");

struct SlotHeader {
    int32_t ndesc;
    int32_t pad;
    int64_t total_tiles;
    int64_t seq;
    int64_t pad2[5];
};
static_assert(sizeof(SlotHeader) == 64, "SlotHeader layout");

constexpr size_t kSlotBytes = sizeof(SlotHeader) + sizeof(Desc) * kMaxDesc;

// Copies header + ndesc descriptors from the pinned table to device memory.
__global__ __launch_bounds__(kThreads) void k_stage(const uint4* __restrict__ host,
                                                    uint4* __restrict__ dev) {
    const SlotHeader* h = reinterpret_cast<const SlotHeader*>(host);
    const int ndesc = h->ndesc < 0 ? 0 : (h->ndesc > kMaxDesc ? kMaxDesc : h->ndesc);
    const int64_t words = (int64_t)(sizeof(SlotHeader) + sizeof(Desc) * ndesc) / 16;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < words;
         i += (int64_t)gridDim.x * kThreads)
        dev[i] = host[i];
}

__device__ __forceinline__ bool put_of(uint32_t w, int j) { return ((w >> (8 * j)) & 0xffu) != 0u; }

// Prices options [o0, o0 + kTile) of one request (clipped to n).
__device__ __forceinline__ void run_tile(const Desc& d, int64_t o0) {
    const int64_t g = o0 / 4 + threadIdx.x;   // 4-option group of this thread
    const int64_t i0 = g * 4;
    if (i0 >= d.n) return;
    const bool greeks = d.kind == serve::kGreeks;
    if (i0 + 4 <= d.n) {
        // arena blocks are 256-byte aligned: the group loads are 16-byte aligned
        const f4 s = __builtin_nontemporal_load(reinterpret_cast<const f4*>(d.in[0]) + g);
        const f4 k = __builtin_nontemporal_load(reinterpret_cast<const f4*>(d.in[1]) + g);
        const f4 r = __builtin_nontemporal_load(reinterpret_cast<const f4*>(d.in[2]) + g);
        const f4 v = __builtin_nontemporal_load(reinterpret_cast<const f4*>(d.in[3]) + g);
        const f4 t = __builtin_nontemporal_load(reinterpret_cast<const f4*>(d.in[4]) + g);
        const uint32_t ty = __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(d.type) + g);
        if (!greeks) {
            f4 p;
#pragma unroll
            for (int j = 0; j < 4; ++j) p[j] = bs_price<Ref32>(s[j], k[j], r[j], v[j], t[j], put_of(ty, j));
            __builtin_nontemporal_store(p, reinterpret_cast<f4*>(d.out[0]) + g);
        } else {
            f4 o[5];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const Greeks32 q = bs_greeks<Ref32>(s[j], k[j], r[j], v[j], t[j], put_of(ty, j));
                o[0][j] = q.price;
                o[1][j] = q.delta;
                o[2][j] = q.vega;
                o[3][j] = q.rho;
                o[4][j] = q.theta;
            }
#pragma unroll
            for (int q = 0; q < 5; ++q)
                __builtin_nontemporal_store(o[q], reinterpret_cast<f4*>(d.out[q]) + g);
        }
        return;
    }
    for (int64_t i = i0; i < d.n; ++i) {
        const float s = d.in[0][i], k = d.in[1][i], r = d.in[2][i], v = d.in[3][i], t = d.in[4][i];
        const bool put = d.type[i] != 0;
        if (!greeks) {
            d.out[0][i] = bs_price<Ref32>(s, k, r, v, t, put);
        } else {
            const Greeks32 q = bs_greeks<Ref32>(s, k, r, v, t, put);
            d.out[0][i] = q.price;
            d.out[1][i] = q.delta;
            d.out[2][i] = q.vega;
            d.out[3][i] = q.rho;
            d.out[4][i] = q.theta;
        }
    }
}

// Fused persistent kernel: block b takes the contiguous tile range
// [b*T/G, (b+1)*T/G) of the batch's T tiles, finds the first tile's request
// by binary search over tile_begin and walks forward from there.
__global__ __launch_bounds__(kThreads) void k_fused(const uint8_t* __restrict__ slot) {
    const SlotHeader* h = reinterpret_cast<const SlotHeader*>(slot);
    const Desc* desc = reinterpret_cast<const Desc*>(slot + sizeof(SlotHeader));
    const int ndesc = h->ndesc;
    const int64_t total = h->total_tiles;
    if (ndesc <= 0 || ndesc > kMaxDesc || total <= 0) return;
    const int64_t G = gridDim.x;
    const int64_t t0 = total * blockIdx.x / G, t1 = total * (blockIdx.x + 1) / G;
    if (t0 >= t1) return;
    int lo = 0, hi = ndesc - 1;   // last desc with tile_begin <= t0
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (desc[mid].tile_begin <= t0) lo = mid;
        else hi = mid - 1;
    }
    int di = lo;
    for (int64_t t = t0; t < t1; ++t) {
        while (di + 1 < ndesc && desc[di + 1].tile_begin <= t) ++di;
        const Desc& d = desc[di];
        run_tile(d, (t - d.tile_begin) * kTile);
    }
}

int64_t now_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct Batch {
    int slot;
    hipEvent_t done;
    std::vector<int64_t> tickets;
};

}  // namespace
}  // namespace bs

struct bs_ctx_s {
    int device = 0;
    bs_serve_opts opts{};
    std::mutex mu;
    std::condition_variable cv_work;   // dispatcher: new requests / memory / shutdown
    std::condition_variable cv_done;   // clients: ticket state changes
    bs::serve::Core* core = nullptr;
    char* arena = nullptr;
    hipStream_t stream = nullptr;
    uint8_t* host_slot[2] = {nullptr, nullptr};    // pinned, written by the dispatcher
    uint8_t* host_slot_dev[2] = {nullptr, nullptr};  // the same memory, device address
    uint8_t* dev_slot[2] = {nullptr, nullptr};
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool slot_busy[2] = {false, false};
    int grid = 0;
    int64_t seq = 0;
    std::vector<bs::Batch> inflight;
    std::thread worker;
    bool shutdown = false;
    std::string fatal;
};

using namespace bs;

namespace {

static_assert(serve::kErrBudget == BS_ERR_BUDGET && serve::kErrBusy == BS_ERR_BUSY &&
                  serve::kErrNotFound == BS_ERR_NOT_FOUND &&
                  serve::kErrShutdown == BS_ERR_SHUTDOWN && serve::kDone == BS_TICKET_DONE &&
                  serve::kCancelled == BS_TICKET_CANCELLED && serve::kGreeks == BS_REQ_GREEKS,
              "serve_core codes mirror bs_mi355.h");

// Launches one slot: graph replay or the two kernels directly.
hipError_t launch_slot(bs_ctx_s* c, int s) {
    if (c->graph[s]) return hipGraphLaunch(c->graph[s], c->stream);
    hipLaunchKernelGGL(k_stage, dim3(kStageBlocks), dim3(kThreads), 0, c->stream,
                       reinterpret_cast<const uint4*>(c->host_slot_dev[s]),
                       reinterpret_cast<uint4*>(c->dev_slot[s]));
    hipLaunchKernelGGL(k_fused, dim3(c->grid), dim3(kThreads), 0, c->stream, c->dev_slot[s]);
    return hipGetLastError();
}

hipError_t capture_slot(bs_ctx_s* c, int s) {
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_stage, dim3(kStageBlocks), dim3(kThreads), 0, c->stream,
                       reinterpret_cast<const uint4*>(c->host_slot_dev[s]),
                       reinterpret_cast<uint4*>(c->dev_slot[s]));
    hipLaunchKernelGGL(k_fused, dim3(c->grid), dim3(kThreads), 0, c->stream, c->dev_slot[s]);
    e = hipStreamEndCapture(c->stream, &g);
    if (e != hipSuccess) return e;
    e = hipGraphInstantiate(&c->graph[s], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    return e;
}

// Fills slot s with the batch's descriptors (caller holds the lock).
void write_slot(bs_ctx_s* c, int s, const std::vector<serve::Request*>& reqs) {
    SlotHeader* h = reinterpret_cast<SlotHeader*>(c->host_slot[s]);
    Desc* d = reinterpret_cast<Desc*>(c->host_slot[s] + sizeof(SlotHeader));
    int64_t tiles = 0;
    for (size_t i = 0; i < reqs.size(); ++i) {
        const serve::Request* r = reqs[i];
        const serve::BookRec* b = c->core->book(r->book);
        size_t bo[6], oo[5];
        serve::book_layout(b->n, bo);
        serve::output_layout(r->n, r->kind, oo);
        Desc& x = d[i];
        for (int j = 0; j < 5; ++j) x.in[j] = reinterpret_cast<const float*>(c->arena + b->off + bo[j]);
        x.type = reinterpret_cast<const uint8_t*>(c->arena + b->off + bo[5]);
        for (int j = 0; j < 5; ++j) x.out[j] = reinterpret_cast<float*>(c->arena + r->out_off + oo[j]);
        x.n = r->n;
        x.tile_begin = tiles;
        x.kind = r->kind;
        x.pad = 0;
        tiles += (r->n + kTile - 1) / kTile;
    }
    h->ndesc = (int32_t)reqs.size();
    h->total_tiles = tiles;
    h->seq = ++c->seq;
}

void dispatcher(bs_ctx_s* c) {
    (void)hipSetDevice(c->device);
    std::unique_lock<std::mutex> lk(c->mu);
    for (;;) {
        bool progressed = false;
        // retire finished batches in launch order
        while (!c->inflight.empty()) {
            Batch& b = c->inflight.front();
            lk.unlock();
            const hipError_t q = hipEventQuery(b.done);
            lk.lock();
            if (q == hipErrorNotReady) break;
            const bool ok = q == hipSuccess;
            if (!ok) c->fatal = std::string("batch failed: ") + hipGetErrorString(q);
            c->core->complete(b.tickets, ok, now_ns());
            c->slot_busy[b.slot] = false;
            c->inflight.erase(c->inflight.begin());
            progressed = true;
            c->cv_done.notify_all();
        }
        // launch the next batch into a free slot
        int s = !c->slot_busy[0] ? 0 : (!c->slot_busy[1] ? 1 : -1);
        if (s >= 0 && !c->shutdown && c->core->queued() > 0) {
            std::vector<serve::Request*> reqs = c->core->admit(now_ns());
            if (!reqs.empty()) {
                write_slot(c, s, reqs);
                Batch b{s, c->ev[s], {}};
                for (auto* r : reqs) b.tickets.push_back(r->ticket);
                hipError_t e = launch_slot(c, s);
                if (e == hipSuccess) e = hipEventRecord(c->ev[s], c->stream);
                if (e != hipSuccess) {
                    c->fatal = std::string("launch failed: ") + hipGetErrorString(e);
                    c->core->complete(b.tickets, false, now_ns());
                    c->cv_done.notify_all();
                } else {
                    c->slot_busy[s] = true;
                    c->inflight.push_back(std::move(b));
                }
                progressed = true;
            }
        }
        if (c->shutdown && c->inflight.empty()) break;
        if (progressed) continue;
        if (!c->inflight.empty()) {
            // a batch is on the GPU: poll its event without holding the lock
            lk.unlock();
            std::this_thread::yield();
            lk.lock();
        } else {
            c->cv_work.wait(lk);
        }
    }
    c->core->fail_queued(serve::kErrShutdown, now_ns());
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
    c->opts.use_graph = 1;
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
    for (int s = 0; s < 2 && e == hipSuccess; ++s) {
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
    for (int s = 0; s < 2 && e == hipSuccess && c->opts.use_graph; ++s) e = capture_slot(c, s);
    if (e != hipSuccess) {
        rc = hip_error(e, "bs_ctx_create");
        bs_ctx_destroy(c);
        return rc;
    }
    serve::Limits lim;
    lim.max_batch_requests = c->opts.max_batch_requests;
    lim.max_batch_options = c->opts.max_batch_options;
    c->core = new serve::Core(hbm_budget_bytes, lim);
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
            c->shutdown = true;
        }
        c->cv_work.notify_all();
        c->worker.join();
    }
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (int s = 0; s < 2; ++s) {
        if (c->graph[s]) (void)hipGraphExecDestroy(c->graph[s]);
        if (c->ev[s]) (void)hipEventDestroy(c->ev[s]);
        if (c->host_slot[s]) (void)hipHostFree(c->host_slot[s]);
        if (c->dev_slot[s]) (void)hipFree(c->dev_slot[s]);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
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
        if (c->shutdown) return set_error(BS_ERR_SHUTDOWN, "context shutting down");
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
        if (c->shutdown) return set_error(BS_ERR_SHUTDOWN, "context shutting down");
        rc = c->core->submit(book, kind, now_ns(), ticket);
    }
    if (rc != BS_OK) return set_error(rc, "bs_submit");
    c->cv_work.notify_one();
    return BS_OK;
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
    }
    if (rc == BS_OK) c->cv_done.notify_all();
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
    return BS_OK;
}

}  // extern "C"
