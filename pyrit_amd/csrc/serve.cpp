// serve.cpp - continuous-batching build server with resident chunk DAGs
// (the request / job pipeline of the reference's src/world.zig:755-1124,
// reshaped for many concurrent requests on one MI355X).
//
// Worker threads (one HIP stream each) loop: take every pending request that
// fits the budget and the batch limits, skipping chunks that are busy in
// another batch so that requests on one chunk run in submission order; stage
// their voxel records as packed segments (builds: device copy of the request's
// records; edits: the chunk's resident sorted voxels decoded back to records,
// followed by the edit records, so later records win as in
// src/gpu_build.zig:90-96); build the whole batch with pyrhip_build_chunks
// (mixed chunk edges); copy the outputs into exact-size buffers (a
// "generation" shared by the batch's chunks, freed when its last chunk is
// replaced); complete the requests. A small batch (at most kGraphChunks
// chunks and kGraphVoxels records, edges up to 2^kGraphEdge) replays a HIP
// graph of the batch build for its largest edge instead (one per worker and
// edge, captured when the server is created), padded with empty chunks.
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "internal.h"

namespace pyr {
namespace {

constexpr uint32_t kGraphChunks = 64;
constexpr uint32_t kGraphVoxels = 1u << 20;

// Graphs are captured by pyrhip_server_create, before the workers start, and
// never while serving: during a capture HIP fails the legacy-stream calls of
// every other thread of the process ("operation would make the legacy stream
// depend on a capturing blocking stream"), the caller's uploads included.
// A replay costs more the larger the graph's edge (one graph of edge 2^8 for
// every small batch: p50 0.67 ms against 0.36 ms with one graph per largest
// edge on the 2000 requests/s workload), so each edge up to 2^kGraphEdge has
// its own.
constexpr uint32_t kGraphEdge = 8;

double steady_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Generation {
    void* nodes = nullptr;
    void* leaves = nullptr;
    void* keys = nullptr;
    void* attrs = nullptr;
    uint32_t log2_size = 0;  // edge the batch was built with (decoding keys)
    uint32_t node_words = 0, leaf_count = 0;
    uint64_t bytes = 0;
    uint32_t live = 0;       // chunks referring to it
};

struct ChunkState {
    std::shared_ptr<Generation> gen;  // null: empty chunk
    uint32_t root = 0, first = 0, count = 0, log2 = 0;
};

struct Req {
    uint64_t ticket = 0, chunk = 0, user = 0;
    uint32_t kind = 0, log2 = 0, count = 0;
    uint64_t dev_voxels = 0;
    std::vector<PyrVoxel> edits;
    double t_submit = 0;
    uint64_t records = 0, est = 0;  // set when the request joins a batch
};

// device bytes a batch record may need: what a build requests per record
// (PyrhipBuildTimes.device_bytes, all of it held at once; measured 23 B on
// the 1024^3 grid, 21-42 B on the 1024-chunk batch, 57 B on the sparse
// 256^3 sphere surface, the worst case measured) plus the exact-size copy of
// the outputs the server keeps (12 B per unique voxel, nodes and leaves)
constexpr uint64_t kBytesPerRecord = 72;
constexpr uint64_t kBytesPerRequest = 1ull << 20;

}  // namespace
}  // namespace pyr

using namespace pyr;

struct PyrhipServer {
    PyrhipServerInfo info{};
    int device = 0;
    std::mutex m;
    std::vector<PyrhipGraph*> graphs;  // [worker * (kGraphEdge + 1) + edge] (null: none)
    std::condition_variable cv_work, cv_done;
    std::deque<Req> pending;
    std::unordered_map<uint64_t, ChunkState> chunks;
    std::unordered_set<uint64_t> busy;
    std::vector<PyrhipCompletion> done;
    uint64_t next_ticket = 1, submitted = 0, completed = 0, rejected = 0, batches = 0, batched_voxels = 0;
    uint64_t resident_bytes = 0, reserved_bytes = 0;
    uint32_t in_flight = 0;
    double busy_s = 0;
    char error[512] = {};  // message of the last failed batch or rejected request
    bool stop = false;
    std::vector<std::thread> workers;

    uint64_t records_of(const Req& r) const {
        if (r.kind == PYRHIP_REQUEST_BUILD) return r.count;
        auto it = chunks.find(r.chunk);
        uint64_t old = it != chunks.end() && it->second.gen ? it->second.count : 0;
        return old + r.edits.size();
    }

    void complete_locked(const Req& r, PyrResult res, uint32_t voxels, double t_start, uint32_t batch_chunks) {
        PyrhipCompletion c{};
        c.ticket = r.ticket;
        c.chunk = r.chunk;
        c.user = r.user;
        c.result = res;
        c.voxel_count = voxels;
        c.submit_s = r.t_submit;
        c.start_s = t_start;
        c.done_s = steady_s();
        c.batch_chunks = batch_chunks;
        done.push_back(c);
        completed++;
        if (res != PYR_OK) rejected++;
    }

    void release_generation_locked(std::shared_ptr<Generation>& g, hipStream_t s) {
        if (!g) return;
        if (--g->live == 0) {
            pool_free(g->nodes, s);
            pool_free(g->leaves, s);
            pool_free(g->keys, s);
            pool_free(g->attrs, s);
            resident_bytes -= g->bytes;
        }
        g.reset();
    }

    // Takes the requests of the next batch out of `pending` (lock held).
    std::vector<Req> form_batch_locked() {
        std::vector<Req> batch;
        std::unordered_set<uint64_t> taken, skipped;
        uint64_t est = 0, voxels = 0;
        const uint64_t budget = info.memory_budget;
        for (auto it = pending.begin(); it != pending.end();) {
            Req& r = *it;
            if (busy.count(r.chunk) || taken.count(r.chunk) || skipped.count(r.chunk)) {
                skipped.insert(r.chunk);  // keeps later requests of the chunk behind this one
                ++it;
                continue;
            }
            if (r.kind == PYRHIP_REQUEST_EDIT && (r.log2 < 3 || r.log2 > 20)) {
                // an edit without an edge relies on the chunk's earlier build;
                // when that build was rejected the edit has no edge either
                auto c = chunks.find(r.chunk);
                if (c == chunks.end() || !c->second.log2) {
                    snprintf(error, sizeof(error), "edit of chunk %llu, whose build failed, has no log2_size",
                             (unsigned long long)r.chunk);
                    complete_locked(r, PYR_ERROR_INVALID_ARGUMENT, 0, steady_s(), 0);
                    it = pending.erase(it);
                    cv_done.notify_all();
                    continue;
                }
            }
            const uint64_t rec = records_of(r);
            const uint64_t e = rec * kBytesPerRecord + kBytesPerRequest;
            const bool alone_fits = rec <= info.max_batch_voxels && resident_bytes + e <= budget;
            if (!alone_fits && batch.empty() && in_flight == 0) {
                // can never run: reject (the reference reports pool exhaustion as OUT_OF_MEMORY)
                snprintf(error, sizeof(error), "request of %llu records on chunk %llu exceeds the batch limit or budget",
                         (unsigned long long)rec, (unsigned long long)r.chunk);
                complete_locked(r, PYR_ERROR_OUT_OF_MEMORY, 0, steady_s(), 0);
                it = pending.erase(it);
                cv_done.notify_all();
                continue;
            }
            if (batch.size() >= info.max_batch_chunks || voxels + rec > info.max_batch_voxels ||
                resident_bytes + reserved_bytes + est + e > budget) {
                skipped.insert(r.chunk);
                ++it;
                continue;
            }
            r.records = rec;
            r.est = e;
            est += e;
            voxels += rec;
            taken.insert(r.chunk);
            batch.push_back(std::move(r));
            it = pending.erase(it);
        }
        for (const Req& r : batch) busy.insert(r.chunk);
        reserved_bytes += est;
        if (!batch.empty()) in_flight++;
        return batch;
    }

    void worker_main(uint32_t w);
    PyrResult run_batch(std::vector<Req>& batch, hipStream_t s, void** input, uint64_t* input_cap, double* dev_ms,
                        std::shared_ptr<Generation>* gen, std::vector<uint32_t>& roots, std::vector<uint32_t>& first,
                        std::vector<uint32_t>& counts, PyrhipGraph* const* graphs);
};

PyrResult PyrhipServer::run_batch(std::vector<Req>& batch, hipStream_t s, void** input, uint64_t* input_cap,
                                  double* dev_ms, std::shared_ptr<Generation>* gen_out, std::vector<uint32_t>& roots,
                                  std::vector<uint32_t>& first, std::vector<uint32_t>& counts, PyrhipGraph* const* graphs) {
    const uint32_t k = (uint32_t)batch.size();
    std::vector<uint32_t> offs(k + 1, 0);
    std::vector<uint8_t> l2(k, 0);
    std::vector<ChunkState> prior(k);
    {
        std::lock_guard<std::mutex> lk(m);  // chunk states of busy chunks only change in this worker
        for (uint32_t i = 0; i < k; i++) {
            auto it = chunks.find(batch[i].chunk);
            if (it != chunks.end()) prior[i] = it->second;
        }
    }
    uint32_t top = 3;
    for (uint32_t i = 0; i < k; i++) {
        const Req& r = batch[i];
        uint32_t lg = r.kind == PYRHIP_REQUEST_EDIT && prior[i].log2 ? prior[i].log2 : r.log2;
        l2[i] = (uint8_t)lg;
        top = std::max(top, lg);
        offs[i + 1] = offs[i] + (uint32_t)r.records;
    }
    const uint64_t total = offs[k];
    // a small batch replays the worker's graph of its largest edge (outputs
    // stay in the graph's buffers: copied below like the eager build's, never freed)
    PyrhipGraph* graph = k <= kGraphChunks && total <= kGraphVoxels && top <= kGraphEdge ? graphs[top] : nullptr;
    const uint64_t need = std::max<uint64_t>(total, 1) * sizeof(PyrVoxel) + (k + 1) * 4 + k + 256;
    if (*input_cap < need) {
        pool_free(*input, s);
        *input = nullptr;
        *input_cap = 0;
        if (pool_alloc(input, need, s) != hipSuccess) {
            (void)hipGetLastError();
            return fail(PYR_ERROR_OUT_OF_MEMORY, "server batch input of %llu records", (unsigned long long)total);
        }
        *input_cap = need;
    }
    char* base = (char*)*input;
    PyrVoxel* vox = (PyrVoxel*)base;
    uint32_t* d_offs = (uint32_t*)(base + std::max<uint64_t>(total, 1) * sizeof(PyrVoxel));
    uint8_t* d_l2 = (uint8_t*)(d_offs + k + 1);
    PyrResult r = PYR_OK;
    for (uint32_t i = 0; i < k && r == PYR_OK; i++) {
        const Req& q = batch[i];
        PyrVoxel* dst = vox + offs[i];
        if (q.kind == PYRHIP_REQUEST_BUILD) {
            if (q.count && hipMemcpyAsync(dst, (void*)(uintptr_t)q.dev_voxels, (uint64_t)q.count * sizeof(PyrVoxel),
                                          hipMemcpyDeviceToDevice, s) != hipSuccess)
                r = fail(PYR_ERROR_CUDA, "server: staging build input");
        } else {
            const ChunkState& c = prior[i];
            uint32_t old = c.gen ? c.count : 0;
            if (old)
                r = pyrhip_decode_voxels((uint64_t)(uintptr_t)c.gen->keys + (uint64_t)c.first * 8,
                                         (uint64_t)(uintptr_t)c.gen->attrs + (uint64_t)c.first * 4, old,
                                         c.gen->log2_size, (uint64_t)(uintptr_t)dst, s);
            if (r == PYR_OK && !q.edits.empty() &&
                hipMemcpyAsync(dst + old, q.edits.data(), q.edits.size() * sizeof(PyrVoxel), hipMemcpyHostToDevice,
                               s) != hipSuccess)
                r = fail(PYR_ERROR_CUDA, "server: staging edits");
        }
    }
    if (r == PYR_OK && hipMemcpyAsync(d_offs, offs.data(), (k + 1) * 4, hipMemcpyHostToDevice, s) != hipSuccess)
        r = fail(PYR_ERROR_CUDA, "server: offsets");
    if (r == PYR_OK && hipMemcpyAsync(d_l2, l2.data(), k, hipMemcpyHostToDevice, s) != hipSuccess)
        r = fail(PYR_ERROR_CUDA, "server: chunk edges");
    if (r != PYR_OK) return r;

    PyrhipChunkBatch cb{};
    cb.count = k;
    cb.capacity = 0;
    cb.voxels = (uint64_t)(uintptr_t)vox;
    cb.offsets = (uint64_t)(uintptr_t)d_offs;
    cb.total = (uint32_t)total;
    cb.chunk_log2 = (uint64_t)(uintptr_t)d_l2;
    roots.assign(k, 0);
    first.assign(k, 0);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, s);
    PyrhipBuilt b{};
    if (graph) {
        std::vector<uint32_t> po(kGraphChunks + 1, (uint32_t)total), gr(kGraphChunks), gf(kGraphChunks);
        std::vector<uint8_t> pl(kGraphChunks, 3);
        std::copy(offs.begin(), offs.end(), po.begin());
        std::copy(l2.begin(), l2.end(), pl.begin());
        r = pyrhip_graph_launch(graph, (uint64_t)(uintptr_t)vox, po.data(), pl.data(), s, &b, gr.data(), gf.data());
        std::copy(gr.begin(), gr.begin() + k, roots.begin());
        std::copy(gf.begin(), gf.begin() + k, first.begin());
    } else {
        r = pyrhip_build_chunks(top, &cb, &info.tuning, s, &b, roots.data(), first.data());
    }
    (void)hipEventRecord(e1, s);
    // copy the outputs into exact-size buffers: the build allocated them by
    // upper bounds, and the generation stays resident
    auto g = std::make_shared<Generation>();
    if (r == PYR_OK) {
        g->log2_size = b.log2_size;  // the graph's edge when replayed: its keys carry that layout
        g->node_words = b.node_words;
        g->leaf_count = b.leaf_count;
        const uint64_t nb = (uint64_t)b.node_words * 4, lb = (uint64_t)b.leaf_count * 8;
        const uint64_t kb = (uint64_t)b.voxel_count * 8, ab = (uint64_t)b.voxel_count * 4;
        bool ok = pool_alloc(&g->nodes, std::max<uint64_t>(nb, 16), s) == hipSuccess;
        ok &= pool_alloc(&g->leaves, std::max<uint64_t>(lb, 16), s) == hipSuccess;
        ok &= pool_alloc(&g->keys, std::max<uint64_t>(kb, 16), s) == hipSuccess;
        ok &= pool_alloc(&g->attrs, std::max<uint64_t>(ab, 16), s) == hipSuccess;
        if (ok) {
            ok &= hipMemcpyAsync(g->nodes, (void*)(uintptr_t)b.nodes, nb, hipMemcpyDeviceToDevice, s) == hipSuccess;
            ok &= hipMemcpyAsync(g->leaves, (void*)(uintptr_t)b.leaves, lb, hipMemcpyDeviceToDevice, s) == hipSuccess;
            ok &= hipMemcpyAsync(g->keys, (void*)(uintptr_t)b.keys, kb, hipMemcpyDeviceToDevice, s) == hipSuccess;
            ok &= hipMemcpyAsync(g->attrs, (void*)(uintptr_t)b.attrs, ab, hipMemcpyDeviceToDevice, s) == hipSuccess;
        }
        g->bytes = nb + lb + kb + ab;
        const uint32_t batch_voxels = b.voxel_count;
        if (!graph) pyrhip_built_free(&b, s);  // zeroes b
        if (ok && hipStreamSynchronize(s) != hipSuccess) ok = false;
        if (!ok) {
            (void)hipGetLastError();
            pool_free(g->nodes, s);
            pool_free(g->leaves, s);
            pool_free(g->keys, s);
            pool_free(g->attrs, s);
            g.reset();
            r = fail(PYR_ERROR_OUT_OF_MEMORY, "server: resident copy of a batch");
        }
        // per-chunk voxel counts: ranks are contiguous per chunk in batch order
        counts.assign(k, 0);
        if (r == PYR_OK) {
            uint32_t next = batch_voxels;
            for (int64_t i = (int64_t)k - 1; i >= 0; i--) {
                if (first[i] == 0xFFFFFFFFu) continue;
                counts[i] = next - first[i];
                next = first[i];
            }
        }
    }
    float ms = 0;
    if (r == PYR_OK) {
        if (hipEventElapsedTime(&ms, e0, e1) == hipSuccess)
            *dev_ms = ms;
        else
            (void)hipGetLastError();  // timing is best effort
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *gen_out = g;
    return r;
}

void PyrhipServer::worker_main(uint32_t w) {
    (void)hipSetDevice(device);
    hipStream_t s = nullptr;
    (void)hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    void* input = nullptr;
    uint64_t input_cap = 0;
    std::unique_lock<std::mutex> lk(m);
    for (;;) {
        std::vector<Req> batch;
        for (;;) {
            if (stop) break;
            batch = form_batch_locked();
            if (!batch.empty()) break;
            cv_work.wait(lk);
        }
        if (batch.empty()) break;
        const double t_start = steady_s();
        lk.unlock();
        std::shared_ptr<Generation> g;
        std::vector<uint32_t> roots, first, counts;
        double dev_ms = 0;
        PyrResult r = run_batch(batch, s, &input, &input_cap, &dev_ms, &g, roots, first, counts,
                                graphs.data() + (size_t)w * (kGraphEdge + 1));
        lk.lock();
        uint64_t est = 0;
        for (const Req& q : batch) est += q.est;
        reserved_bytes -= est;
        in_flight--;
        batches++;
        busy_s += dev_ms * 1e-3;
        if (r == PYR_OK) {
            uint32_t live = 0;
            for (uint32_t i = 0; i < batch.size(); i++) live += roots[i] != 0xFFFFFFFFu;
            g->live = live;
            if (live) resident_bytes += g->bytes;
            for (uint32_t i = 0; i < batch.size(); i++) {
                const Req& q = batch[i];
                batched_voxels += q.records;
                ChunkState& c = chunks[q.chunk];
                const uint32_t lg = q.kind == PYRHIP_REQUEST_EDIT && c.log2 ? c.log2 : q.log2;
                release_generation_locked(c.gen, s);
                c.log2 = lg;
                c.count = 0;
                if (roots[i] != 0xFFFFFFFFu) {
                    c.gen = g;
                    c.root = roots[i];
                    c.first = first[i];
                    c.count = counts[i];
                }
                complete_locked(q, PYR_OK, c.count, t_start, (uint32_t)batch.size());
            }
            if (!live && g) {
                g->live = 1;
                std::shared_ptr<Generation> tmp = g;
                g->bytes = 0;
                release_generation_locked(tmp, s);
            }
        } else {
            snprintf(error, sizeof(error), "batch of %u requests: %s", (uint32_t)batch.size(), g_diag);
            for (const Req& q : batch) complete_locked(q, r, 0, t_start, (uint32_t)batch.size());
        }
        for (const Req& q : batch) busy.erase(q.chunk);
        cv_done.notify_all();
        cv_work.notify_all();
    }
    lk.unlock();
    pool_free(input, s);
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
}

extern "C" {

PYR_API double pyrhip_server_clock(void) { return steady_s(); }

PYR_API PyrResult pyrhip_server_create(const PyrhipServerInfo* info, PyrhipServer** out) {
    diag_clear();
    if (!out) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return fail(PYR_ERROR_NOT_FOUND, "no HIP device");
    }
    PyrhipServer* sv = new PyrhipServer();
    if (info) sv->info = *info;
    if (!sv->info.memory_budget) sv->info.memory_budget = 64ull << 30;
    if (!sv->info.max_batch_chunks) sv->info.max_batch_chunks = 1024;
    if (!sv->info.max_batch_voxels) sv->info.max_batch_voxels = 1u << 28;
    if (!sv->info.workers) sv->info.workers = 2;
    (void)hipGetDevice(&sv->device);
    sv->graphs.assign((size_t)sv->info.workers * (kGraphEdge + 1), nullptr);
    if (!(sv->info.flags & PYRHIP_SERVER_NO_GRAPHS))
        for (uint32_t w = 0; w < sv->info.workers; w++)
            for (uint32_t e = 3; e <= kGraphEdge; e++) {
                PyrhipGraph*& g = sv->graphs[(size_t)w * (kGraphEdge + 1) + e];
                if (pyrhip_graph_create(e, kGraphChunks, 0, kGraphVoxels, 1, &sv->info.tuning, &g) != PYR_OK) {
                    g = nullptr;  // not enough memory for it: batches of that edge are built eagerly
                    diag_clear();
                }
            }
    for (uint32_t w = 0; w < sv->info.workers; w++) sv->workers.emplace_back([sv, w] { sv->worker_main(w); });
    *out = sv;
    return PYR_OK;
}

PYR_API PyrResult pyrhip_server_submit(PyrhipServer* sv, const PyrhipRequest* q, uint64_t* ticket) {
    diag_clear();
    if (!sv || !q) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    if (q->kind > PYRHIP_REQUEST_EDIT) return fail(PYR_ERROR_INVALID_ARGUMENT, "request kind");
    if (q->count && !q->voxels) return fail(PYR_ERROR_INVALID_ARGUMENT, "null voxels");
    Req r;
    r.chunk = q->chunk;
    r.user = q->user;
    r.kind = q->kind;
    r.log2 = q->log2_size;
    r.count = q->count;
    r.t_submit = steady_s();
    if (q->kind == PYRHIP_REQUEST_BUILD) {
        if (q->log2_size < 3 || q->log2_size > 20)
            return fail(PYR_ERROR_INVALID_ARGUMENT, "log2_size must be in [3, 20]");
        r.dev_voxels = q->voxels;
    } else {
        const PyrVoxel* e = (const PyrVoxel*)(uintptr_t)q->voxels;
        r.edits.assign(e, e + q->count);
    }
    std::lock_guard<std::mutex> lk(sv->m);
    if (q->kind == PYRHIP_REQUEST_EDIT && (q->log2_size < 3 || q->log2_size > 20)) {
        auto it = sv->chunks.find(q->chunk);
        bool known = it != sv->chunks.end() && it->second.log2;
        for (const Req& p : sv->pending) known |= p.chunk == q->chunk;
        if (!known) return fail(PYR_ERROR_INVALID_ARGUMENT, "edit of an unknown chunk needs log2_size");
    }
    r.ticket = sv->next_ticket++;
    if (ticket) *ticket = r.ticket;
    sv->submitted++;
    sv->pending.push_back(std::move(r));
    sv->cv_work.notify_one();
    return PYR_OK;
}

PYR_API uint32_t pyrhip_server_poll(PyrhipServer* sv, PyrhipCompletion* out, uint32_t max) {
    if (!sv || !out) return 0;
    std::lock_guard<std::mutex> lk(sv->m);
    uint32_t n = (uint32_t)std::min<size_t>(max, sv->done.size());
    std::copy(sv->done.begin(), sv->done.begin() + n, out);
    sv->done.erase(sv->done.begin(), sv->done.begin() + n);
    return n;
}

PYR_API PyrResult pyrhip_server_drain(PyrhipServer* sv) {
    diag_clear();
    if (!sv) return fail(PYR_ERROR_INVALID_ARGUMENT, "null server");
    std::unique_lock<std::mutex> lk(sv->m);
    sv->cv_done.wait(lk, [sv] { return sv->completed == sv->submitted; });
    return PYR_OK;
}

PYR_API PyrResult pyrhip_server_chunk(PyrhipServer* sv, uint64_t chunk, PyrhipDag* out, uint32_t* voxel_count,
                                      uint32_t* node_words, uint32_t* leaf_count) {
    diag_clear();
    if (!sv || !out) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(sv->m);
    auto it = sv->chunks.find(chunk);
    if (it == sv->chunks.end()) return fail(PYR_ERROR_INVALID_HANDLE, "chunk %llu is not resident",
                                            (unsigned long long)chunk);
    const ChunkState& c = it->second;
    memset(out, 0, sizeof(*out));
    out->log2_size = c.log2;
    out->default_attribute = 1;
    if (voxel_count) *voxel_count = c.count;
    if (node_words) *node_words = c.gen ? c.gen->node_words : 0;
    if (leaf_count) *leaf_count = c.gen ? c.gen->leaf_count : 0;
    if (!c.gen) return PYR_OK;  // empty chunk: nodes == 0
    out->nodes = (uint64_t)(uintptr_t)c.gen->nodes;
    out->leaves = (uint64_t)(uintptr_t)c.gen->leaves;
    out->attrs = (uint64_t)(uintptr_t)c.gen->attrs + (uint64_t)c.first * 4;
    out->root = c.root;
    return PYR_OK;
}

PYR_API void pyrhip_server_stats(PyrhipServer* sv, PyrhipServerStats* out) {
    if (!sv || !out) return;
    std::lock_guard<std::mutex> lk(sv->m);
    memset(out, 0, sizeof(*out));
    out->submitted = sv->submitted;
    out->completed = sv->completed;
    out->rejected = sv->rejected;
    out->batches = sv->batches;
    uint64_t live = 0;
    for (auto& kv : sv->chunks) live += kv.second.gen != nullptr;
    out->resident_chunks = live;
    out->resident_bytes = sv->resident_bytes;
    out->reserved_bytes = sv->reserved_bytes;
    out->pending = sv->pending.size();
    out->batched_voxels = sv->batched_voxels;
    out->busy_s = sv->busy_s;
}

PYR_API uint32_t pyrhip_server_error(PyrhipServer* sv, char* out, uint32_t size) {
    if (!sv) return 0;
    std::lock_guard<std::mutex> lk(sv->m);
    const uint32_t n = (uint32_t)strlen(sv->error);
    if (out && size) snprintf(out, size, "%s", sv->error);
    return n;
}

PYR_API void pyrhip_server_destroy(PyrhipServer* sv) {
    if (!sv) return;
    {
        std::lock_guard<std::mutex> lk(sv->m);
        sv->stop = true;
        sv->cv_work.notify_all();
    }
    for (std::thread& t : sv->workers) t.join();
    for (PyrhipGraph* g : sv->graphs) pyrhip_graph_destroy(g);
    for (auto& kv : sv->chunks) sv->release_generation_locked(kv.second.gen, nullptr);
    (void)hipDeviceSynchronize();
    delete sv;
}

}  // extern "C"
