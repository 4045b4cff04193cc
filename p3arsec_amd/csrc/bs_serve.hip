// bs_serve.hip -- the serving loop: resident books, continuous batching,
// fused launches of every queued request.
//
// A context owns
//   * one device arena (hipMalloc of the HBM budget) holding resident books
//     and request outputs, carved by serve::Core's segregated-fit allocator;
//   * six descriptor slots, each a pinned host table plus its device copy;
//   * per batch, a stage kernel that copies the slot's table from pinned host
//     memory to the device, then one fused kernel per section (the batch's
//     price requests, its Greeks requests): a block prices one tile of 512
//     options of one request, found by a 64-ary wave search of the table.
//     Launched directly (the default) each section's grid is one block per
//     tile.  With use_graph the three kernels are captured once per slot and
//     each batch is one hipGraphLaunch; before it, each section's node gets
//     the batch's grid (hipGraphExecKernelNodeSetParams).  On ROCm 7.2 the
//     direct form measured 15 us launch-to-completion against 31 us;
//   * a dispatcher thread (serve_loop.hpp): it admits queued requests as arena
//     memory allows, writes a table into a free slot and launches it, and
//     retires each batch when its event completes.  Requests of at most 64K
//     options (bs_serve_opts.small_request_options) form batches of their own
//     on a stream of their own (two in flight), so they complete beside large
//     requests rather than behind them; the rest alternate over two streams
//     (four in flight).  Requests that arrive while batches run go into the
//     next batch of their lane.
// Results are bit-identical to bs_price_f32 / bs_price_greeks_f32: the fused
// kernels run the batch map's own per-group code (bs_ops.hpp).
#include <stddef.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <string>
#include <vector>

#include "bs_internal.hpp"
#include "bs_math.hpp"
#include "bs_ops.hpp"
#include "serve_core.hpp"
#include "serve_loop.hpp"

namespace bs {
namespace {

constexpr int kVW = 2;                 // options per work-item (the batch map's width)
// Threads per block of the fused kernels, chosen per context from
// BS_CFG_BLOCK_THREADS (64, 128 or 256; 0 = kServeThreads); a tile is one
// block step, threads x kVW options.
constexpr int kServeThreads = 256;
constexpr int kMaxDesc = 4096;         // requests per batch: 64 x 64 for find_desc
constexpr int64_t kMaxGrid = int64_t(1) << 20;  // default section grid cap (blocks loop)
constexpr int kStageBlocks = 32;

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

// Fused kernel of one section: a block prices tile t (kThr x kVW options of
// one request, one kVW-group per work-item) for t = blockIdx.x, + gridDim.x, ...
// The grid has one block per tile up to the context's cap (kMaxGrid unless
// bs_serve_opts.grid_blocks sets it), in both launch forms.
// Every access is bounded by the request's n.
// The per-group code is the batch map's own (bs_ops.hpp).
template <class Op, int kSection, int kThr>
__global__ __launch_bounds__(kThr) __attribute__((amdgpu_waves_per_eu(waves_for<Op, false>())))
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
        const int64_t g = (t - x.tile_begin) * kThr + threadIdx.x;
        if ((g + 1) * kVW <= n) {
            op.template finish<true>(g, op.template load<true>(g));
        } else {
            for (int64_t i = g * kVW; i < n; ++i) op.scalar(i);
        }
    }
}

using ServePrice = PriceF32<Ref32, kVW>;
using ServeGreeks = GreeksF32<kVW>;

// The fused kernel of section sec with kThr threads per block.
template <int kThr>
const void* serve_fn(int sec) {
    return sec == 0 ? reinterpret_cast<const void*>(&k_serve<ServePrice, 0, kThr>)
                    : reinterpret_cast<const void*>(&k_serve<ServeGreeks, 1, kThr>);
}
const void* serve_fn(int threads, int sec) {
    return threads == 64 ? serve_fn<64>(sec) : threads == 128 ? serve_fn<128>(sec)
                                                                : serve_fn<256>(sec);
}

}  // namespace

// The HIP executor of the serving loop (serve_loop.hpp): descriptor tables,
// the stage and fused kernels, three streams and one event per slot.
struct HipExec {
    // Batches in flight at once: two of the small lane on their own stream,
    // four of the bulk lane alternating over two streams.  While one runs, the
    // next ones are queued behind it, so the GPU stays busy while clients turn
    // finished requests around into new ones; the small lane's batches run
    // beside the bulk lane's instead of behind them.
    static constexpr int kSlots = 6;
    static constexpr int kSmallSlots = 2;

    int device = 0;
    int use_graph = 0;
    int threads = kServeThreads;                // fused kernels' block size
    int64_t tile = kServeThreads * kVW;         // options per tile
    int64_t grid_cap = kMaxGrid;                // blocks per section at most
    char* arena = nullptr;
    hipStream_t stream = nullptr;               // capture stream
    hipStream_t streams[3] = {};                // bulk batches alternate over 0 and 1;
                                                // the small lane's run on 2
    uint8_t* host_slot[kSlots] = {};            // pinned, written by the dispatcher
    uint8_t* host_slot_dev[kSlots] = {};        // the same memory, device address
    uint8_t* dev_slot[kSlots] = {};
    hipGraphExec_t graph[kSlots] = {};
    hipGraph_t graph_src[kSlots] = {};          // kept: its nodes name the exec's nodes
    hipGraphNode_t section_node[kSlots][2] = {};
    unsigned section_grid[kSlots][2] = {};      // grid the exec's section node has now
    void* section_args[kSlots][1] = {};         // kernel arguments: &dev_slot[s]
    hipEvent_t ev[kSlots] = {};
    hipStream_t slot_stream[kSlots] = {};        // the stream slot s was last launched on
    int64_t seq = 0, nlaunch = 0;

    hipError_t init(int dev, size_t arena_bytes, int graph_on, int grid_blocks, int block_threads) {
        device = dev;
        use_graph = graph_on;
        if (block_threads == 64 || block_threads == 128 || block_threads == 256)
            threads = block_threads;
        tile = (int64_t)threads * kVW;
        hipError_t e = hipSetDevice(device);
        if (e == hipSuccess) e = hipMalloc((void**)&arena, arena_bytes);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
        for (int j = 0; j < 3 && e == hipSuccess; ++j)
            e = hipStreamCreateWithFlags(&streams[j], hipStreamNonBlocking);
        for (int s = 0; s < kSlots && e == hipSuccess; ++s) {
            e = hipHostMalloc((void**)&host_slot[s], kSlotBytes, hipHostMallocCoherent);
            if (e == hipSuccess)
                e = hipHostGetDevicePointer((void**)&host_slot_dev[s], host_slot[s], 0);
            if (e == hipSuccess) e = hipMalloc((void**)&dev_slot[s], kSlotBytes);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&ev[s], hipEventDisableTiming);
            if (e == hipSuccess) ::memset(host_slot[s], 0, sizeof(SlotHeader));
        }
        if (grid_blocks > 0) grid_cap = std::min<int64_t>(grid_blocks, kMaxGrid);
        for (int s = 0; s < kSlots && e == hipSuccess && use_graph; ++s) e = capture(s);
        return e;
    }
    void release() {
        (void)hipSetDevice(device);
        drain();
        for (int s = 0; s < kSlots; ++s) {
            if (graph[s]) (void)hipGraphExecDestroy(graph[s]);
            if (graph_src[s]) (void)hipGraphDestroy(graph_src[s]);
            if (ev[s]) (void)hipEventDestroy(ev[s]);
            if (host_slot[s]) (void)hipHostFree(host_slot[s]);
            if (dev_slot[s]) (void)hipFree(dev_slot[s]);
        }
        if (stream) (void)hipStreamDestroy(stream);
        for (hipStream_t st : streams)
            if (st) (void)hipStreamDestroy(st);
        if (arena) (void)hipFree(arena);
    }

    // Captures slot s as one graph: stage, price section, Greeks section.  The
    // section nodes are found by kernel so each launch can set their grids.
    hipError_t capture(int s) {
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_stage, dim3(kStageBlocks), dim3(kThreads), 0, stream,
                           reinterpret_cast<const uint4*>(host_slot_dev[s]),
                           reinterpret_cast<uint4*>(dev_slot[s]));
        section_args[s][0] = &dev_slot[s];
        for (int sec = 0; sec < 2; ++sec)
            (void)hipLaunchKernel(serve_fn(threads, sec), dim3(1), dim3(threads), section_args[s],
                                  0, stream);
        e = hipStreamEndCapture(stream, &g);
        if (e != hipSuccess) return e;
        graph_src[s] = g;
        const void* fn[2] = {serve_fn(threads, 0), serve_fn(threads, 1)};
        hipGraphNode_t nodes[3];
        size_t count = 3;
        e = hipGraphGetNodes(g, nodes, &count);
        for (size_t i = 0; i < count && e == hipSuccess; ++i) {
            hipKernelNodeParams p{};
            e = hipGraphKernelNodeGetParams(nodes[i], &p);
            for (int sec = 0; sec < 2 && e == hipSuccess; ++sec)
                if (p.func == fn[sec]) section_node[s][sec] = nodes[i];
        }
        if (e == hipSuccess && (count != 3 || !section_node[s][0] || !section_node[s][1]))
            e = hipErrorInvalidValue;
        if (e == hipSuccess) e = hipGraphInstantiate(&graph[s], g, nullptr, nullptr, 0);
        section_grid[s][0] = section_grid[s][1] = 1;
        return e;
    }

    // Sets section sec of slot s's graph to a grid of `blocks`.
    hipError_t set_section_grid(int s, int sec, unsigned blocks) {
        if (section_grid[s][sec] == blocks) return hipSuccess;
        hipKernelNodeParams p{};
        p.func = const_cast<void*>(serve_fn(threads, sec));
        p.gridDim = dim3(blocks);
        p.blockDim = dim3(threads);
        p.sharedMemBytes = 0;
        p.kernelParams = section_args[s];
        p.extra = nullptr;
        const hipError_t e = hipGraphExecKernelNodeSetParams(graph[s], section_node[s][sec], &p);
        if (e == hipSuccess) section_grid[s][sec] = blocks;
        return e;
    }

    // Fills slot s with the batch's descriptors, price requests first, each
    // section in admission order.  The slot is the dispatcher's own while no
    // batch uses it, so no lock is held.
    void write_slot(int s, const std::vector<serve::Admitted>& reqs) {
        SlotHeader* h = reinterpret_cast<SlotHeader*>(host_slot[s]);
        Desc* d = reinterpret_cast<Desc*>(host_slot[s] + sizeof(SlotHeader));
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
            for (int j = 0; j < 5; ++j) x.in[j] = reinterpret_cast<const float*>(arena + r.book_off + bo[j]);
            x.type = reinterpret_cast<const uint8_t*>(arena + r.book_off + bo[5]);
            for (int j = 0; j < 5; ++j) x.out[j] = reinterpret_cast<float*>(arena + r.out_off + oo[j]);
            x.n = r.n;
            x.tile_begin = tiles[sec];
            x.kind = r.kind;
            x.pad = 0;
            tiles[sec] += (r.n + tile - 1) / tile;
        }
        h->ndesc[0] = nd[0];
        h->ndesc[1] = nd[1];
        h->tiles[0] = tiles[0];
        h->tiles[1] = tiles[1];
        h->seq = ++seq;
    }

    // ---- the executor interface ----
    void thread_start() { (void)hipSetDevice(device); }

    // Launches slot s on its lane's stream (the bulk lane alternates between
    // two): graph replay, or the stage kernel
    // and the fused kernel of each non-empty section directly; each section's
    // grid is one block per tile.  Then records the slot's event.
    int launch(int s, const std::vector<serve::Admitted>& reqs, std::string* err) {
        write_slot(s, reqs);
        hipStream_t st = s < kSmallSlots ? streams[2] : streams[nlaunch++ & 1];
        slot_stream[s] = st;
        const SlotHeader* h = reinterpret_cast<const SlotHeader*>(host_slot[s]);
        hipError_t e = hipSuccess;
        if (graph[s]) {
            // an empty section keeps one block, which returns at once
            for (int sec = 0; sec < 2 && e == hipSuccess; ++sec) {
                const int64_t blocks = std::max<int64_t>(1, std::min(h->tiles[sec], grid_cap));
                e = set_section_grid(s, sec, (unsigned)blocks);
            }
            if (e == hipSuccess) e = hipGraphLaunch(graph[s], st);
        } else {
            hipLaunchKernelGGL(k_stage, dim3(kStageBlocks), dim3(kThreads), 0, st,
                               reinterpret_cast<const uint4*>(host_slot_dev[s]),
                               reinterpret_cast<uint4*>(dev_slot[s]));
            e = hipGetLastError();
            void* args[1] = {&dev_slot[s]};
            for (int sec = 0; sec < 2 && e == hipSuccess; ++sec)
                if (h->tiles[sec] > 0)
                    e = hipLaunchKernel(serve_fn(threads, sec),
                                        dim3((unsigned)std::min(h->tiles[sec], grid_cap)),
                                        dim3(threads), args, 0, st);
        }
        if (e == hipSuccess) e = hipEventRecord(ev[s], st);
        if (e != hipSuccess) *err = hipGetErrorString(e);
        return e == hipSuccess ? 0 : -1;
    }
    int query(int s, std::string* err) {
        const hipError_t q = hipEventQuery(ev[s]);
        if (q == hipErrorNotReady) return 0;
        if (q == hipSuccess) return 1;
        *err = hipGetErrorString(q);
        return -1;
    }
    // A launch that failed part-way may have queued the stage kernel or a
    // section on the slot's stream: wait for that stream before the loop frees
    // the batch's outputs and reuses the slot's tables.
    void quiesce(int s) {
        if (slot_stream[s]) (void)hipStreamSynchronize(slot_stream[s]);
    }
    void drain() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (hipStream_t st : streams)
            if (st) (void)hipStreamSynchronize(st);
    }
};

}  // namespace bs

struct bs_ctx_s {
    bs::HipExec exec;
    bs::serve::Loop<bs::HipExec>* loop = nullptr;
};

using namespace bs;

namespace {

static_assert(serve::kErrBudget == BS_ERR_BUDGET && serve::kErrBusy == BS_ERR_BUSY &&
                  serve::kErrNotFound == BS_ERR_NOT_FOUND && serve::kErrTimeout == BS_ERR_TIMEOUT &&
                  serve::kErrCancelled == BS_ERR_CANCELLED && serve::kErrFailed == BS_ERR_FAILED &&
                  serve::kErrShutdown == BS_ERR_SHUTDOWN && serve::kDone == BS_TICKET_DONE &&
                  serve::kCancelled == BS_TICKET_CANCELLED && serve::kGreeks == BS_REQ_GREEKS,
              "serve_core codes mirror bs_mi355.h");

int status(int rc, const char* what) { return rc == BS_OK ? BS_OK : set_error(rc, what); }

}  // namespace

extern "C" {

int bs_ctx_create(int device, size_t hbm_budget_bytes, const bs_serve_opts* opts, bs_ctx_t* out) {
    return guarded("bs_ctx_create", [&]() -> int {
        if (!out) return set_error(BS_ERR_INVALID, "ctx is NULL");
        *out = nullptr;
        if (hbm_budget_bytes < serve::kAlign * 16)
            return set_error(BS_ERR_INVALID, "hbm_budget_bytes too small");
        int ndev = 0;
        hipError_t e = hipGetDeviceCount(&ndev);
        if (e != hipSuccess || device < 0 || device >= ndev)
            return set_error(BS_ERR_NODEVICE, "device index out of range");
        serve::Limits lim;
        lim.max_batch_requests = kMaxDesc;
        lim.max_batch_options = int64_t(1) << 31;
        lim.small_options = 65536;
        int use_graph = 0, grid_blocks = 0;
        const int block_threads = config().block_threads;
        if (opts) {
            if (opts->small_request_options != 0)
                lim.small_options = opts->small_request_options < 0 ? 0 : opts->small_request_options;
            if (opts->max_batch_requests > 0)
                lim.max_batch_requests = std::min<int64_t>(opts->max_batch_requests, kMaxDesc);
            if (opts->max_batch_options > 0) lim.max_batch_options = opts->max_batch_options;
            use_graph = opts->use_graph;
            grid_blocks = opts->grid_blocks;
        }
        auto* c = new bs_ctx_s();
        e = c->exec.init(device, hbm_budget_bytes, use_graph, grid_blocks, block_threads);
        if (e != hipSuccess) {
            const int rc = hip_error(e, "bs_ctx_create");
            c->exec.release();
            delete c;
            return rc;
        }
        try {
            c->loop = new serve::Loop<HipExec>(&c->exec, hbm_budget_bytes, lim,
                                               opts && opts->completion_queue != 0);
            c->loop->start();
        } catch (...) {   // free the device resources, then report the failure
            delete c->loop;
            c->exec.release();
            delete c;
            throw;
        }
        *out = c;
        return BS_OK;
    });
}

int bs_ctx_destroy(bs_ctx_t ctx) {
    return guarded("bs_ctx_destroy", [&]() -> int {
        bs_ctx_s* c = ctx;
        if (!c) return set_error(BS_ERR_INVALID, "ctx is NULL");
        if (c->loop) c->loop->stop();
        c->exec.release();
        delete c->loop;
        delete c;
        return BS_OK;
    });
}

int bs_book_register(bs_ctx_t ctx, const float* s, const float* k, const float* r, const float* v,
                     const float* t, const uint8_t* otype, int64_t n, bs_book_t* book) {
    return guarded("bs_book_register", [&]() -> int {
        bs_ctx_s* c = ctx;
        if (!c || !book) return set_error(BS_ERR_INVALID, "ctx or book is NULL");
        if (n <= 0) return set_error(BS_ERR_INVALID, "n <= 0");
        if (!s || !k || !r || !v || !t || !otype) return set_error(BS_ERR_INVALID, "NULL array");
        int64_t id;
        size_t off;
        const int rc = c->loop->register_book(n, &id, &off);
        if (rc == BS_ERR_SHUTDOWN) return set_error(rc, "context shutting down");
        if (rc != BS_OK) return set_error(rc, "book does not fit in the arena");
        size_t bo[6];
        serve::book_layout(n, bo);
        const void* src[6] = {s, k, r, v, t, otype};
        hipError_t e = hipSetDevice(c->exec.device);
        for (int j = 0; j < 6 && e == hipSuccess; ++j)
            e = hipMemcpy(c->exec.arena + off + bo[j], src[j], (size_t)n * (j < 5 ? 4 : 1),
                          hipMemcpyHostToDevice);
        // The batches run on non-blocking streams: the book must be in device
        // memory before its id is handed out.
        if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
        if (e != hipSuccess) {
            c->loop->release_book(id);
            return hip_error(e, "bs_book_register");
        }
        *book = id;
        return BS_OK;
    });
}

int bs_book_release(bs_ctx_t ctx, bs_book_t book) {
    return guarded("bs_book_release", [&]() -> int {
        bs_ctx_s* c = ctx;
        if (!c) return set_error(BS_ERR_INVALID, "ctx is NULL");
        return status(c->loop->release_book(book), "bs_book_release");
    });
}

int bs_submit(bs_ctx_t ctx, bs_book_t book, int kind, bs_ticket_t* ticket) {
    return guarded("bs_submit", [&]() -> int {
        bs_ctx_s* c = ctx;
        if (!c || !ticket) return set_error(BS_ERR_INVALID, "ctx or ticket is NULL");
        return status(c->loop->submit(book, kind, ticket), "bs_submit");
    });
}

int bs_submit_many(bs_ctx_t ctx, const bs_book_t* books, const int* kinds, const int64_t* users,
                   int64_t count, bs_ticket_t* tickets) {
    return guarded("bs_submit_many", [&]() -> int {
        bs_ctx_s* c = ctx;
        if (!c) return set_error(BS_ERR_INVALID, "ctx is NULL");
        if (count < 0 || (count > 0 && (!books || !kinds || !tickets)))
            return set_error(BS_ERR_INVALID, "bad request arrays");
        return status(c->loop->submit_many(books, kinds, users, count, tickets), "bs_submit_many");
    });
}

int bs_reap(bs_ctx_t ctx, bs_completion* out, int64_t max, int64_t timeout_us, int64_t* count) {
    return guarded("bs_reap", [&]() -> int {
        static_assert(sizeof(bs_completion) == sizeof(serve::Completion) &&
                          offsetof(bs_completion, user) == offsetof(serve::Completion, user) &&
                          offsetof(bs_completion, state) == offsetof(serve::Completion, state) &&
                          offsetof(bs_completion, done_ns) == offsetof(serve::Completion, t_done),
                      "bs_completion layout");
        bs_ctx_s* c = ctx;
        if (!c || !count) return set_error(BS_ERR_INVALID, "ctx or count is NULL");
        *count = 0;
        if (max < 0 || (max > 0 && !out)) return set_error(BS_ERR_INVALID, "bad output array");
        const int rc = c->loop->reap(reinterpret_cast<serve::Completion*>(out), max, timeout_us, count);
        return rc == BS_OK ? BS_OK : set_error(rc, "context has no completion queue");
    });
}

int bs_result_release_many(bs_ctx_t ctx, const bs_ticket_t* tickets, int64_t count) {
    return guarded("bs_result_release_many", [&]() -> int {
        bs_ctx_s* c = ctx;
        if (!c) return set_error(BS_ERR_INVALID, "ctx is NULL");
        if (count < 0 || (count > 0 && !tickets)) return set_error(BS_ERR_INVALID, "bad ticket array");
        return status(c->loop->release_many(tickets, count), "bs_result_release_many");
    });
}

int bs_poll(bs_ctx_t ctx, bs_ticket_t ticket, int* state) {
    return guarded("bs_poll", [&]() -> int {
        bs_ctx_s* c = ctx;
        if (!c || !state) return set_error(BS_ERR_INVALID, "ctx or state is NULL");
        return status(c->loop->poll(ticket, state), "unknown ticket");
    });
}

int bs_wait(bs_ctx_t ctx, bs_ticket_t ticket, int64_t timeout_us) {
    return guarded("bs_wait", [&]() -> int {
        bs_ctx_s* c = ctx;
        if (!c) return set_error(BS_ERR_INVALID, "ctx is NULL");
        std::string msg;
        const int rc = c->loop->wait(ticket, timeout_us, &msg);
        switch (rc) {
            case BS_OK: return BS_OK;
            case BS_ERR_TIMEOUT: return set_error(rc, "bs_wait timed out");
            case BS_ERR_NOT_FOUND: return set_error(rc, "unknown ticket");
            case BS_ERR_CANCELLED: return set_error(rc, "ticket cancelled");
            default: return set_error(rc, msg);
        }
    });
}

int bs_result_device(bs_ctx_t ctx, bs_ticket_t ticket, const float** price, const float** delta,
                     const float** vega, const float** rho, const float** theta) {
    return guarded("bs_result_device", [&]() -> int {
        bs_ctx_s* c = ctx;
        if (!c) return set_error(BS_ERR_INVALID, "ctx is NULL");
        int64_t n;
        int kind;
        size_t off;
        const int rc = c->loop->result_info(ticket, &n, &kind, &off);
        if (rc == BS_ERR_NOT_FOUND) return set_error(rc, "unknown ticket");
        if (rc != BS_OK) return set_error(rc, "ticket not done");
        size_t oo[5];
        serve::output_layout(n, kind, oo);
        const float** dst[5] = {price, delta, vega, rho, theta};
        for (int j = 0; j < 5; ++j) {
            if (!dst[j]) continue;
            *dst[j] = (j == 0 || kind == serve::kGreeks)
                          ? reinterpret_cast<const float*>(c->exec.arena + off + oo[j])
                          : nullptr;
        }
        return BS_OK;
    });
}

int bs_result_copy(bs_ctx_t ctx, bs_ticket_t ticket, float* price, float* delta, float* vega,
                   float* rho, float* theta) {
    return guarded("bs_result_copy", [&]() -> int {
        bs_ctx_s* c = ctx;
        if (!c) return set_error(BS_ERR_INVALID, "ctx is NULL");
        int64_t n;
        int kind;
        size_t off;
        const int rc = c->loop->result_info(ticket, &n, &kind, &off);
        if (rc == BS_ERR_NOT_FOUND) return set_error(rc, "unknown ticket");
        if (rc != BS_OK) return set_error(rc, "ticket not done");
        size_t oo[5];
        serve::output_layout(n, kind, oo);
        float* dst[5] = {price, delta, vega, rho, theta};
        hipError_t e = hipSetDevice(c->exec.device);
        for (int j = 0; j < 5 && e == hipSuccess; ++j) {
            if (!dst[j]) continue;
            if (j > 0 && kind != serve::kGreeks) return set_error(BS_ERR_INVALID, "not a Greeks ticket");
            e = hipMemcpy(dst[j], c->exec.arena + off + oo[j], (size_t)n * 4, hipMemcpyDeviceToHost);
        }
        return e == hipSuccess ? BS_OK : hip_error(e, "bs_result_copy");
    });
}

int bs_result_release(bs_ctx_t ctx, bs_ticket_t ticket) {
    return guarded("bs_result_release", [&]() -> int {
        bs_ctx_s* c = ctx;
        if (!c) return set_error(BS_ERR_INVALID, "ctx is NULL");
        return status(c->loop->release(ticket), "bs_result_release");
    });
}

int bs_cancel(bs_ctx_t ctx, bs_ticket_t ticket) {
    return guarded("bs_cancel", [&]() -> int {
        bs_ctx_s* c = ctx;
        if (!c) return set_error(BS_ERR_INVALID, "ctx is NULL");
        return status(c->loop->cancel(ticket), "bs_cancel: ticket not queued");
    });
}

int bs_ticket_times(bs_ctx_t ctx, bs_ticket_t ticket, int64_t* submit_ns, int64_t* done_ns) {
    return guarded("bs_ticket_times", [&]() -> int {
        bs_ctx_s* c = ctx;
        if (!c) return set_error(BS_ERR_INVALID, "ctx is NULL");
        return status(c->loop->times(ticket, submit_ns, done_ns), "unknown ticket");
    });
}

int bs_ctx_stats(bs_ctx_t ctx, bs_serve_stats* out) {
    return guarded("bs_ctx_stats", [&]() -> int {
        bs_ctx_s* c = ctx;
        if (!c || !out) return set_error(BS_ERR_INVALID, "ctx or stats is NULL");
        serve::Stats st;
        serve::LoopTimes lt;
        size_t cap, used, peak;
        c->loop->stats(&st, &cap, &used, &peak, &lt);
        out->arena_bytes = (int64_t)cap;
        out->arena_used = (int64_t)used;
        out->arena_peak = (int64_t)peak;
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
        out->batch_ns_total = lt.batch;
        out->dispatch_admit_ns = lt.admit;
        out->dispatch_launch_ns = lt.launch;
        out->dispatch_retire_ns = lt.retire;
        out->gpu_idle_ns = lt.idle;
        out->idle_launches = lt.idle_launches;
        return BS_OK;
    });
}

}  // extern "C"
