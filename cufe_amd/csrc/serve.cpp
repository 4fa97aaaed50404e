// serve.cpp -- continuous-batching prediction server over resident keys.
//
// The reference serves one row at a time: svm_predict_gpu (main.c:270-390)
// ships the full master secret key and the full ciphertext to the device for
// every row and regenerates every functional key.  Here the keys of every
// loaded model stay in HBM; requests carry rows in compact ciphertext form
// (c_L and coefficient 0 of c_0..c_{l-1}), and a worker thread packs whatever
// requests are queued into the next batch, as many as a batch slot holds.
//
// Device memory is fixed at start: the models' keys plus 2 to 4 batch slots,
// each of a row capacity the HBM budget affords (every row of a batch holds
// its row store, staging, contraction scratch and key values, so a request of
// n rows occupies n times that many bytes while it is in flight).  A request
// larger than a slot is split into slot-sized parts that run as separate
// batches (possibly side by side) and complete the request together.  While
// one slot's batch runs on its stream, the next is staged and launched on
// another slot's stream.  A row's result depends only on the row and its
// model (the contraction of a row never mixes other rows), so a request
// returns the same bits whether it runs alone, among others or in parts.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "resident.h"

using namespace ipfe;

namespace {

using Clock = std::chrono::steady_clock;
constexpr size_t kPoly = (size_t)3 * IPFE_N;

double now_s() { return std::chrono::duration<double>(Clock::now().time_since_epoch()).count(); }

// A submitted request: done when all of its parts are.
struct Ticket {
    uint64_t id = 0;
    double t_arrival = 0, t_done = 0;
    int parts_left = 0;
    int status = 0;  // 0 queued or running, 1 done, -1 failed
    std::string err;
};
using TicketP = std::shared_ptr<Ticket>;

// One part of a request (the whole request unless it exceeds a slot): n rows
// of the caller's buffers, offset to the part.
struct Req {
    TicketP tk;
    int model = 0;
    const Model *m = nullptr;
    int n = 0, l = 0;
    const uint32_t *cl = nullptr, *c0 = nullptr;
    double *labels = nullptr;
    uint32_t *kvalues = nullptr;
};
using ReqP = std::shared_ptr<Req>;

// Rows of one model inside a batch: [r0, r0+R) of the slot's row store.
struct Group {
    const Model *m;
    int r0, R;
    size_t kv_off;  // offset of its [R][S] key values in the pinned buffer
    bool want_kv;
};

struct Slot {
    hipStream_t st = nullptr;
    hipEvent_t done = nullptr;
    RowStore rows;
    Scratch sc;
    DevBuf<uint32_t> d_stage;  // cl [cap][3][N], then packed c0 of the batch
    uint32_t *h_stage = nullptr;
    double *h_labels = nullptr;
    uint32_t *h_kval = nullptr;
    int *h_flags = nullptr;  // one per group
    std::vector<ReqP> reqs;
    std::vector<int> req_row, req_group;
    std::vector<Group> groups;
    bool busy = false;
    int batch_rows = 0;  // rows of the batch in flight (its aligned layout)
    bool big = false;    // a group of the batch in flight is a large range (full-GPU launches)
    size_t batch_bytes = 0;  // HBM its rows occupy (rows x bytes per row)
    bool failed = false;
    std::string err;
};

}  // namespace

struct cufe_server {
    cufe_ctx *ctx = nullptr;
    size_t budget = 0;
    int max_batch = 0;
    int batching = 1;
    // worker's poll interval while a batch is in flight (us; 0 = query the
    // events continuously, yielding the core between queries: a batch is
    // retired as soon as it ends, at the cost of a core while the GPU works);
    // CUFE_SERVE_POLL_US overrides
    int poll_us = 0;
    std::vector<std::unique_ptr<Model>> models;
    size_t model_bytes = 0;
    int S_max = 0, l_max = 0, cap = 0;
    size_t slot_bytes = 0, row_bytes = 0;
    std::vector<std::unique_ptr<Slot>> slots;  // 2 to 4, sized at start
    std::mutex mu;
    std::condition_variable cv_submit, cv_done;
    std::deque<ReqP> pending;
    std::unordered_map<uint64_t, TicketP> live;
    uint64_t next_ticket = 1;
    std::thread worker;
    bool started = false, stop = false;
    long long st_requests = 0, st_rows = 0, st_batches = 0, st_failed = 0, st_batch_rows_max = 0;
    long long st_split = 0;           // requests split into parts
    size_t inflight_bytes = 0, st_inflight_peak = 0;  // HBM held by batches in flight
};

namespace ipfe {
const uint32_t *ctx_msk(const cufe_ctx *c);
int ctx_key_len(const cufe_ctx *c);
}  // namespace ipfe

namespace {

inline int round32(int x) { return (x + 31) & ~31; }

// Rows a batch gathers from the queue: the largest batch whose contraction
// still runs as the single split-K launch (predict.cpp, kSplitMaxRows).
constexpr long long kBatchRows = 256;

// Stages the batch in sl (host copies, H2D, row-store puts) and launches its
// prediction; returns -1 with the error set if a call fails.
int launch_batch(cufe_server *s, Slot &sl)
{
    const int mode = contract_mode();
    // layout: requests grouped by model in order of first appearance, each
    // group starting on a 32-row boundary (the digit store's row blocks)
    std::vector<const Model *> gm;
    for (size_t i = 0; i < sl.reqs.size(); ++i) {
        const Model *m = sl.reqs[i]->m;
        if (std::find(gm.begin(), gm.end(), m) == gm.end()) gm.push_back(m);
    }
    sl.req_row.assign(sl.reqs.size(), 0);
    sl.req_group.assign(sl.reqs.size(), 0);
    sl.groups.clear();
    int row = 0;
    size_t kv = 0;
    for (size_t g = 0; g < gm.size(); ++g) {
        Group gr{gm[g], row, 0, kv, false};
        for (size_t i = 0; i < sl.reqs.size(); ++i) {
            if (sl.reqs[i]->m != gm[g]) continue;
            sl.req_row[i] = row + gr.R;
            sl.req_group[i] = (int)g;
            gr.R += sl.reqs[i]->n;
            gr.want_kv |= sl.reqs[i]->kvalues != nullptr;
        }
        kv += (size_t)gr.R * gm[g]->S;
        row = round32(row + gr.R);
        sl.groups.push_back(gr);
    }
    if (row > s->cap) {
        ipfe_set_error("server: batch layout of %d rows exceeds the slot capacity %d", row, s->cap);
        return -1;
    }
    // upload: c_L rows to their slot rows, each request's c0 packed after
    // them.  Buffers in pinned memory from cufe_host_alloc are copied by DMA
    // directly; others are first copied into the slot's pinned staging.
    uint32_t *h_c0 = sl.h_stage + (size_t)s->cap * kPoly;
    uint32_t *d_c0 = sl.d_stage.get() + (size_t)s->cap * kPoly;
    std::vector<size_t> c0_off(sl.reqs.size());
    size_t c0_words = 0;
    for (size_t i = 0; i < sl.reqs.size(); ++i) {
        const Req &r = *sl.reqs[i];
        const size_t cl_bytes = (size_t)r.n * kPoly * 4, c0_bytes = (size_t)r.n * r.l * 12;
        const size_t cl_at = (size_t)sl.req_row[i] * kPoly;
        const uint32_t *src = r.cl;
        if (!host_pinned(r.cl, cl_bytes)) {
            memcpy(sl.h_stage + cl_at, r.cl, cl_bytes);
            src = sl.h_stage + cl_at;
        }
        HIP_OK(hipMemcpyAsync(sl.d_stage.get() + cl_at, src, cl_bytes, hipMemcpyHostToDevice, sl.st));
        c0_off[i] = c0_words;
        src = r.c0;
        if (!host_pinned(r.c0, c0_bytes)) {
            memcpy(h_c0 + c0_words, r.c0, c0_bytes);
            src = h_c0 + c0_words;
        }
        HIP_OK(hipMemcpyAsync(d_c0 + c0_words, src, c0_bytes, hipMemcpyHostToDevice, sl.st));
        c0_words += (size_t)r.n * r.l * 3;
    }
    // one conversion launch per run of requests whose rows and c0 words follow
    // each other and whose rows have the same length (every request of a
    // one-model batch)
    for (size_t i = 0; i < sl.reqs.size();) {
        const int l = sl.reqs[i]->l;
        int n = sl.reqs[i]->n;
        size_t j = i + 1;
        while (j < sl.reqs.size() && sl.reqs[j]->l == l && sl.req_row[j] == sl.req_row[i] + n &&
               c0_off[j] == c0_off[i] + (size_t)n * l * 3) {
            n += sl.reqs[j]->n;
            ++j;
        }
        if (sl.rows.put(sl.req_row[i], n, sl.d_stage.get() + (size_t)sl.req_row[i] * kPoly, d_c0 + c0_off[i], sl.st,
                        l))
            return -1;
        i = j;
    }
    // small groups take the low-latency split-K launch unless another slot
    // runs a large range (full-GPU launches, which under load the split launch
    // would slow more than it gains).  A batch of at most kBatchRows rows has
    // none, whatever its aligned layout (up to 31 padding rows per model) and
    // whether or not its models can use the split-K launch: only a request
    // part larger than that does.
    sl.batch_rows = row;
    sl.batch_bytes = (size_t)row * s->row_bytes;
    sl.big = false;
    for (const Group &gr : sl.groups) sl.big = sl.big || large_range(*gr.m, gr.R);
    bool large_other = false;
    for (const auto &o : s->slots)
        if (o.get() != &sl && o->busy && o->big) large_other = true;
    sl.sc.split_ok = !large_other;
    for (size_t g = 0; g < sl.groups.size(); ++g) {
        const Group &gr = sl.groups[g];
        if (predict_rows(*gr.m, sl.rows, gr.r0, gr.R, sl.sc, s->cap, mode, sl.h_labels + gr.r0,
                         gr.want_kv ? sl.h_kval + gr.kv_off : nullptr, nullptr, sl.st))
            return -1;
        // the flag is raised only by the separate key-value fold (unfused keys)
        if (gr.m->fused)
            sl.h_flags[g] = 0;
        else
            HIP_OK(hipMemcpyAsync(sl.h_flags + g, sl.sc.flag.get(), 4, hipMemcpyDeviceToHost, sl.st));
    }
    return 0;
}

// Copies results of a finished slot to the callers' buffers and completes the
// parts it ran (a request completes with its last part).
void retire(cufe_server *s, Slot &sl, hipError_t q)
{
    if (q != hipSuccess && !sl.failed) {
        sl.failed = true;
        sl.err = std::string("server batch failed: ") + hipGetErrorString(q);
    }
    std::vector<std::string> errs(sl.reqs.size());
    for (size_t i = 0; i < sl.reqs.size() && !sl.failed; ++i) {
        Req &r = *sl.reqs[i];
        const Group &gr = sl.groups[sl.req_group[i]];
        if (sl.h_flags[sl.req_group[i]]) {
            errs[i] = "server: a decryption sum ended on the representative q with a zero product; this model "
                      "needs the full-polynomial path";
            continue;
        }
        memcpy(r.labels, sl.h_labels + sl.req_row[i], sizeof(double) * r.n);
        if (r.kvalues)
            memcpy(r.kvalues, sl.h_kval + gr.kv_off + (size_t)(sl.req_row[i] - gr.r0) * gr.m->S,
                   sizeof(uint32_t) * r.n * gr.m->S);
    }
    const double t = now_s();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        int rows = 0;
        for (size_t i = 0; i < sl.reqs.size(); ++i) {
            Req &r = *sl.reqs[i];
            Ticket &tk = *r.tk;
            if (sl.failed || !errs[i].empty()) {
                if (tk.err.empty()) tk.err = sl.failed ? sl.err : errs[i];
            }
            rows += r.n;
            if (--tk.parts_left == 0) {
                tk.t_done = t;
                tk.status = tk.err.empty() ? 1 : -1;
                s->st_requests++;
                if (tk.status < 0) s->st_failed++;
            }
        }
        s->st_rows += rows;
        s->st_batches++;
        s->st_batch_rows_max = std::max<long long>(s->st_batch_rows_max, rows);
        s->inflight_bytes -= sl.batch_bytes;
        sl.batch_bytes = 0;
        sl.reqs.clear();
        sl.busy = false;
        sl.failed = false;
        sl.err.clear();
    }
    s->cv_done.notify_all();
}

void worker_main(cufe_server *s)
{
    (void)hipSetDevice(0);
    for (;;) {
        bool progressed = false;
        for (auto &slp : s->slots) {
            Slot &sl = *slp;
            if (!sl.busy) continue;
            const hipError_t q = hipEventQuery(sl.done);
            if (q == hipErrorNotReady) continue;
            retire(s, sl, q);
            progressed = true;
        }
        for (auto &slp : s->slots) {
            Slot &sl = *slp;
            if (sl.busy) continue;
            {
                std::lock_guard<std::mutex> lk(s->mu);
                // admit queued parts in arrival order while the aligned layout
                // fits the slot and the batch stays within kBatchRows (a part
                // that is larger runs alone): such a batch takes the one-launch
                // split-K contraction, so a slot that frees under load takes a
                // short batch and the backlog drains on every slot at once,
                // instead of one slot taking the whole backlog into a long
                // launch that every request in it waits for
                long long taken_rows = 0;
                std::vector<std::pair<const Model *, int>> per_model;
                while (!s->pending.empty() &&
                       (taken_rows == 0 || taken_rows + s->pending.front()->n <= kBatchRows)) {
                    const ReqP &r = s->pending.front();
                    auto it = std::find_if(per_model.begin(), per_model.end(),
                                           [&](const std::pair<const Model *, int> &p) { return p.first == r->m; });
                    int need = 0;
                    for (auto &p : per_model) need += round32(p.second + (p.first == r->m ? r->n : 0));
                    if (it == per_model.end()) need += round32(r->n);
                    if (need > s->cap) break;
                    if (it == per_model.end())
                        per_model.push_back({r->m, r->n});
                    else
                        it->second += r->n;
                    taken_rows += r->n;
                    sl.reqs.push_back(r);
                    s->pending.pop_front();
                    if (!s->batching) break;
                }
                if (!sl.reqs.empty()) sl.busy = true;
            }
            if (!sl.busy) break;  // queue empty
            if (launch_batch(s, sl)) {
                sl.failed = true;
                sl.err = cufe_last_error();
            }
            {
                std::lock_guard<std::mutex> lk(s->mu);
                s->inflight_bytes += sl.batch_bytes;
                s->st_inflight_peak = std::max(s->st_inflight_peak, s->inflight_bytes);
            }
            const hipError_t e = hipEventRecord(sl.done, sl.st);
            if (e != hipSuccess) {
                (void)hipStreamSynchronize(sl.st);
                retire(s, sl, e);
            }
            progressed = true;
        }
        if (progressed) continue;
        bool any_busy = false, any_free = false;
        for (auto &slp : s->slots) {
            any_busy |= slp->busy;
            any_free |= !slp->busy;
        }
        if (!any_free) {
            // every slot runs: query them all again (yielding the core), so
            // that whichever batch ends first is retired and its slot refilled
            // at once -- a blocking wait on one of them can sit behind a
            // larger batch while another slot is already free
            std::this_thread::yield();
            continue;
        }
        std::unique_lock<std::mutex> lk(s->mu);
        if (s->pending.empty()) {
            if (s->stop && !any_busy) return;
            if (any_busy && s->poll_us == 0) {
                // a batch in flight: poll its event without sleeping
                lk.unlock();
                std::this_thread::yield();
            } else {
                s->cv_submit.wait_for(lk, std::chrono::microseconds(any_busy ? s->poll_us : 2000));
            }
        }
    }
}

int submit_req(cufe_server *s, int model, int n, int l, const uint32_t *cl, const uint32_t *c0, double *labels,
               uint32_t *kvalues, double t_arrival, uint64_t *ticket)
{
    if (!s || !s->started) {
        ipfe_set_error("cufe_server_submit: server not started");
        return -1;
    }
    if (model < 0 || model >= (int)s->models.size()) {
        ipfe_set_error("cufe_server_submit: unknown model %d", model);
        return -1;
    }
    const Model *m = s->models[model].get();
    if (n < 1) {
        ipfe_set_error("cufe_server_submit: %d rows", n);
        return -1;
    }
    if (l <= m->maxidx || l > s->l_max) {
        ipfe_set_error("cufe_server_submit: rows of %d slots; model %d needs %d..%d", l, model, m->maxidx + 1, s->l_max);
        return -1;
    }
    if (!cl || !c0 || !labels) {
        ipfe_set_error("cufe_server_submit: null buffer");
        return -1;
    }
    // parts of at most one slot's capacity (a multiple of 64 rows)
    const int part = s->cap;
    const int nparts = (n + part - 1) / part;
    auto tk = std::make_shared<Ticket>();
    tk->t_arrival = t_arrival;
    tk->parts_left = nparts;
    std::vector<ReqP> parts;
    for (int off = 0; off < n; off += part) {
        auto r = std::make_shared<Req>();
        r->tk = tk;
        r->model = model;
        r->m = m;
        r->n = std::min(part, n - off);
        r->l = l;
        r->cl = cl + (size_t)off * kPoly;
        r->c0 = c0 + (size_t)off * l * 3;
        r->labels = labels + off;
        r->kvalues = kvalues ? kvalues + (size_t)off * m->S : nullptr;
        parts.push_back(std::move(r));
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        tk->id = s->next_ticket++;
        for (auto &r : parts) s->pending.push_back(r);
        s->live[tk->id] = tk;
        if (nparts > 1) s->st_split++;
    }
    s->cv_submit.notify_one();
    *ticket = tk->id;
    return 0;
}

void free_slot_host(Slot &sl)
{
    if (sl.h_stage) (void)hipHostFree(sl.h_stage);
    if (sl.h_labels) (void)hipHostFree(sl.h_labels);
    if (sl.h_kval) (void)hipHostFree(sl.h_kval);
    if (sl.h_flags) (void)hipHostFree(sl.h_flags);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.st) (void)hipStreamDestroy(sl.st);
    sl.h_stage = nullptr;
    sl.h_labels = nullptr;
    sl.h_kval = nullptr;
    sl.h_flags = nullptr;
    sl.done = nullptr;
    sl.st = nullptr;
}

}  // namespace

// hbm_budget: device bytes for the models' keys plus both batch slots (0 =
// 16 GiB); max_batch_rows: upper bound of a batch (0 = 4096); batching 0
// runs one request per batch (the per-request baseline).
extern "C" int cufe_server_create(cufe_ctx *ctx, size_t hbm_budget, int max_batch_rows, int batching, cufe_server **out)
{
    if (!ctx || !out || ctx_key_len(ctx) < 1) {
        ipfe_set_error("cufe_server_create: context without keys");
        return -1;
    }
    if (!device()) return -1;
    cufe_server *s = new cufe_server();
    s->ctx = ctx;
    s->budget = hbm_budget ? hbm_budget : ((size_t)16 << 30);
    s->max_batch = max_batch_rows > 0 ? max_batch_rows : 4096;
    s->batching = batching ? 1 : 0;
    if (const char *e = getenv("CUFE_SERVE_POLL_US")) s->poll_us = std::max(0, atoi(e));
    s->l_max = ctx_key_len(ctx);
    *out = s;
    return 0;
}

extern "C" int cufe_server_add_model(cufe_server *s, const cufe_svm_model *m, int *model_id)
{
    if (!s || s->started) {
        ipfe_set_error("cufe_server_add_model: models are added before start");
        return -1;
    }
    Device *dev = device();
    if (!dev) return -1;
    auto md = std::make_unique<Model>();
    if (md->build(ctx_msk(s->ctx), s->l_max, m, dev->stream)) return -1;
    s->model_bytes += md->bytes();
    s->S_max = std::max(s->S_max, md->S);
    if (model_id) *model_id = (int)s->models.size();
    s->models.push_back(std::move(md));
    return 0;
}

// Sizes the slots from the budget (a row capacity of up to max_batch_rows,
// and 2 to 4 slots of it), allocates them and starts the worker.
extern "C" int cufe_server_start(cufe_server *s)
{
    if (!s || s->started || s->models.empty()) {
        ipfe_set_error("cufe_server_start: needs a created server with at least one model");
        return -1;
    }
    // device bytes one batch row occupies: row store, staging (c_L and c0),
    // key values of the largest model and the vote
    const size_t per_row = RowStore::bytes_per_row(s->l_max) + (kPoly + (size_t)s->l_max * 3) * 4 +
                           (size_t)s->S_max * 16 + 8;
    const size_t split = Scratch::split_bytes(s->max_batch, s->S_max);  // per slot, fixed size
    const size_t headroom = (size_t)64 << 20;
    if (s->budget < s->model_bytes + headroom + 2 * (split + 64 * per_row)) {
        ipfe_set_error("cufe_server_start: HBM budget %zu below models (%zu) + two 64-row slots (%zu)", s->budget,
                       s->model_bytes, 2 * (split + 64 * per_row));
        return -1;
    }
    const size_t avail = s->budget - s->model_bytes - headroom;
    size_t cap = (avail / 2 - split) / per_row;  // the capacity two slots afford
    cap = std::min(cap, (size_t)((s->max_batch + 63) / 64 * 64));
    cap = std::min(cap, (size_t)1 << 20);
    s->cap = (int)(cap / 64 * 64);
    s->row_bytes = per_row;
    s->slot_bytes = (size_t)s->cap * per_row + split;
    const int nslots = (int)std::max<size_t>(2, std::min<size_t>(4, avail / s->slot_bytes));
    const size_t stage_words = (size_t)s->cap * (kPoly + (size_t)s->l_max * 3);
    for (int i = 0; i < nslots; ++i) {
        s->slots.push_back(std::make_unique<Slot>());
        Slot &sl = *s->slots.back();
        if (sl.rows.reserve(s->cap, s->l_max) || sl.sc.reserve(s->cap, s->S_max) || sl.d_stage.alloc(stage_words))
            return -1;
        HIP_OK(hipHostMalloc((void **)&sl.h_stage, stage_words * 4, hipHostMallocDefault));
        HIP_OK(hipHostMalloc((void **)&sl.h_labels, sizeof(double) * s->cap, hipHostMallocDefault));
        HIP_OK(hipHostMalloc((void **)&sl.h_kval, sizeof(uint32_t) * (size_t)s->cap * s->S_max, hipHostMallocDefault));
        HIP_OK(hipHostMalloc((void **)&sl.h_flags, sizeof(int) * (s->models.size() + 1), hipHostMallocDefault));
        HIP_OK(hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking));
        HIP_OK(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
    }
    HIP_OK(hipDeviceSynchronize());
    // Run every model once per power-of-two batch size on the (zeroed) rows of
    // slot 0, with and without the split-K launch for small batches (a batch
    // may take either), so the contraction's block shapes are tuned here and
    // not on the first requests.
    Slot &w = *s->slots[0];
    const int mode = contract_mode();
    for (int split_ok = 1; split_ok >= 0; --split_ok) {
        w.sc.split_ok = split_ok != 0;
        for (auto &m : s->models)
            for (int R = 1; R <= s->cap; R *= 2)
                if (predict_rows(*m, w.rows, 0, R, w.sc, s->cap, mode, w.h_labels, nullptr, nullptr, w.st))
                    return -1;
    }
    w.sc.split_ok = true;
    HIP_OK(hipStreamSynchronize(w.st));
    s->started = true;
    s->worker = std::thread(worker_main, s);
    return 0;
}

// Queues n rows of l slots for model `model`: cl [n][3][4096] canonical and
// c0 [n][l][3] (coefficient 0 of c_0..c_{l-1}).  The buffers, and labels [n]
// and kvalues [n][S] (optional) that receive the results, stay owned by the
// caller and must live until the request completes.
extern "C" int cufe_server_submit(cufe_server *s, int model, int n, int l, const uint32_t *cl, const uint32_t *c0,
                                  double *labels, uint32_t *kvalues, uint64_t *ticket)
{
    if (!ticket) {
        ipfe_set_error("cufe_server_submit: null ticket");
        return -1;
    }
    return submit_req(s, model, n, l, cl, c0, labels, kvalues, now_s(), ticket);
}

// Waits up to timeout_ms (< 0: without limit) for a request: 1 done (the
// ticket is released), 0 still pending, -1 failed (released; error set).
// latency_ms (optional): submit-to-completion time.
extern "C" int cufe_server_wait(cufe_server *s, uint64_t ticket, double timeout_ms, double *latency_ms)
{
    if (!s) {
        ipfe_set_error("cufe_server_wait: null server");
        return -1;
    }
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->live.find(ticket);
    if (it == s->live.end()) {
        ipfe_set_error("cufe_server_wait: unknown ticket %llu", (unsigned long long)ticket);
        return -1;
    }
    TicketP r = it->second;
    auto ready = [&] { return r->status != 0; };
    if (timeout_ms < 0)
        s->cv_done.wait(lk, ready);
    else if (!s->cv_done.wait_for(lk, std::chrono::microseconds((long long)(timeout_ms * 1000)), ready))
        return 0;
    s->live.erase(ticket);
    if (latency_ms) *latency_ms = (r->t_done - r->t_arrival) * 1e3;
    if (r->status < 0) {
        ipfe_set_error("%s", r->err.c_str());
        return -1;
    }
    return 1;
}

// Open-loop replay: request i is submitted at arrival_s[i] seconds after the
// call starts, for model[i] with nrows[i] rows of l[i] slots (cl[i], c0[i]);
// labels go to labels[i].  Returns after every request completes, with
// latency_ms[i] = completion - scheduled arrival and makespan_s = last
// completion - start.  The number of failed requests is returned (0 = all ok).
extern "C" int cufe_server_replay(cufe_server *s, int nreq, const double *arrival_s, const int *model,
                                  const int *nrows, const int *l, const uint32_t *const *cl,
                                  const uint32_t *const *c0, double *const *labels, double *latency_ms,
                                  double *makespan_s)
{
    if (!s || nreq < 0) {
        ipfe_set_error("cufe_server_replay: bad arguments");
        return -1;
    }
    std::vector<uint64_t> tickets(nreq);
    const auto t0c = Clock::now();
    const double t0 = now_s();
    for (int i = 0; i < nreq; ++i) {
        const auto when = t0c + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(arrival_s[i]));
        std::this_thread::sleep_until(when);
        if (submit_req(s, model[i], nrows[i], l[i], cl[i], c0[i], labels[i], nullptr, t0 + arrival_s[i], &tickets[i])) {
            // drain what was submitted before reporting
            for (int j = 0; j < i; ++j) (void)cufe_server_wait(s, tickets[j], -1, nullptr);
            return -1;
        }
    }
    int failed = 0;
    double last = t0;
    for (int i = 0; i < nreq; ++i) {
        double lat = 0;
        const int rc = cufe_server_wait(s, tickets[i], -1, &lat);
        if (rc < 0) failed++;
        latency_ms[i] = lat;
        last = std::max(last, t0 + arrival_s[i] + lat * 1e-3);
    }
    if (makespan_s) *makespan_s = last - t0;
    return failed;
}

// requests, rows, batches, failed requests, largest batch (rows), slot
// capacity (rows), model bytes, slot bytes.
extern "C" int cufe_server_stats(cufe_server *s, long long *out8)
{
    if (!s || !out8) {
        ipfe_set_error("cufe_server_stats: null argument");
        return -1;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    out8[0] = s->st_requests;
    out8[1] = s->st_rows;
    out8[2] = s->st_batches;
    out8[3] = s->st_failed;
    out8[4] = s->st_batch_rows_max;
    out8[5] = s->cap;
    out8[6] = (long long)s->model_bytes;
    out8[7] = (long long)s->slot_bytes;
    return 0;
}

// Zeroes the counters cufe_server_stats reports (requests, rows, batches,
// failed requests, largest batch), e.g. after an untimed warm-up.
extern "C" int cufe_server_reset_stats(cufe_server *s)
{
    if (!s) {
        ipfe_set_error("cufe_server_reset_stats: null server");
        return -1;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->st_requests = s->st_rows = s->st_batches = s->st_failed = s->st_batch_rows_max = 0;
    s->st_split = 0;
    s->st_inflight_peak = s->inflight_bytes;
    return 0;
}

// Slots, requests split into parts, peak HBM bytes held by batches in flight
// and device bytes one batch row occupies.
extern "C" int cufe_server_info(cufe_server *s, int *nslots, long long *split_requests, long long *peak_inflight_bytes,
                                long long *bytes_per_row)
{
    if (!s) {
        ipfe_set_error("cufe_server_info: null server");
        return -1;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (nslots) *nslots = (int)s->slots.size();
    if (split_requests) *split_requests = s->st_split;
    if (peak_inflight_bytes) *peak_inflight_bytes = (long long)s->st_inflight_peak;
    if (bytes_per_row) *bytes_per_row = (long long)s->row_bytes;
    return 0;
}

// Completes every queued request, stops the worker and frees the server.
extern "C" void cufe_server_destroy(cufe_server *s)
{
    if (!s) return;
    if (s->started) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
        }
        s->cv_submit.notify_all();
        s->worker.join();
    }
    for (auto &sl : s->slots) free_slot_host(*sl);
    delete s;
}
