// tune.cpp - build tuning per workload class and the autotuner that picks it.
//
// A workload class is (log2_size, bit length of the record count, bit length
// of the chunk count): builds of one class have the same number of radix
// passes and tiles within a factor of two, so one setting serves the class.
// pyrhip_autotune measures settings on a real workload by coordinate descent
// (tile sizes of the sort and of the scans, digit width / record layout, hash
// load) and installs the fastest one whose output is identical to the
// defaults'. Every setting must give the same output; the tuner checks the
// counts, the roots and a checksum of the node and leaf arrays, and never
// picks a setting that differs.
#include <hip/hip_runtime.h>

#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "internal.h"

namespace pyr {
namespace {

constexpr PyrhipBuildTuning kDefaults = {8, 20, 2, 12, 80, 2, 0, 1, 0};

struct Table {
    std::mutex m;
    std::map<uint64_t, PyrhipBuildTuning> classes;
};

Table& table() {
    static Table* t = new Table();
    return *t;
}

uint32_t bit_length(uint64_t v) {
    uint32_t b = 0;
    while (v) {
        b++;
        v >>= 1;
    }
    return b;
}

uint64_t class_key(uint32_t log2_size, uint32_t n, uint32_t chunk_count) {
    return ((uint64_t)log2_size << 16) | ((uint64_t)bit_length(n) << 8) | bit_length(chunk_count);
}

void overlay(PyrhipBuildTuning& t, const PyrhipBuildTuning& o) {
    if (o.radix_bits) t.radix_bits = o.radix_bits;
    if (o.sort_items) t.sort_items = o.sort_items;
    if (o.pair_layout) t.pair_layout = o.pair_layout;
    if (o.scan_items) t.scan_items = o.scan_items;
    if (o.hash_load_pct) t.hash_load_pct = o.hash_load_pct;
    if (o.status_words) t.status_words = o.status_words;
    if (o.finish) t.finish = o.finish;
    if (o.counted_pass) t.counted_pass = o.counted_pass;
    if (o.levels) t.levels = o.levels;
}

bool same_tuning(const PyrhipBuildTuning& a, const PyrhipBuildTuning& b) {
    return a.radix_bits == b.radix_bits && a.sort_items == b.sort_items && a.pair_layout == b.pair_layout &&
           a.scan_items == b.scan_items && a.hash_load_pct == b.hash_load_pct && a.status_words == b.status_words &&
           a.finish == b.finish && a.counted_pass == b.counted_pass && a.levels == b.levels;
}

// what a build produced, enough to tell two settings' outputs apart
struct Signature {
    uint32_t root = 0, node_words = 0, leaf_count = 0, voxel_count = 0;
    uint64_t sum = 0;
    std::vector<uint32_t> roots, first;
    bool operator==(const Signature& o) const {
        return root == o.root && node_words == o.node_words && leaf_count == o.leaf_count &&
               voxel_count == o.voxel_count && sum == o.sum && roots == o.roots && first == o.first;
    }
};

uint64_t mix(uint64_t h, uint64_t v) {
    h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    return h * 0xFF51AFD7ED558CCDull;
}

struct Workload {
    uint32_t log2_size;
    uint64_t voxels;
    uint32_t n;
    const PyrhipChunkBatch* batch;
    hipStream_t s;
};

uint64_t count_of(const Workload& w) { return w.batch ? w.batch->total : w.n; }
uint32_t chunks_of(const Workload& w) { return w.batch ? w.batch->count : 0u; }

PyrResult build_once(const Workload& w, const PyrhipBuildTuning& t, PyrhipBuilt* b, Signature* sig) {
    if (!w.batch) return pyrhip_build(w.log2_size, 0, 0, 0, w.voxels, w.n, &t, w.s, b);
    std::vector<uint32_t> roots(w.batch->count), first(w.batch->count);
    PyrResult r = pyrhip_build_chunks(w.log2_size, w.batch, &t, w.s, b, roots.data(), first.data());
    if (r == PYR_OK && sig) {
        sig->roots.swap(roots);
        sig->first.swap(first);
    }
    return r;
}

PyrResult signature_of(const PyrhipBuilt& b, hipStream_t s, Signature* sig) {
    sig->root = b.root;
    sig->node_words = b.node_words;
    sig->leaf_count = b.leaf_count;
    sig->voxel_count = b.voxel_count;
    std::vector<uint32_t> nodes(b.node_words);
    std::vector<uint64_t> leaves(b.leaf_count);
    if (!nodes.empty() &&
        hipMemcpyAsync(nodes.data(), (void*)(uintptr_t)b.nodes, nodes.size() * 4, hipMemcpyDeviceToHost, s) != hipSuccess)
        return fail(PYR_ERROR_CUDA, "autotune: node download");
    if (!leaves.empty() && hipMemcpyAsync(leaves.data(), (void*)(uintptr_t)b.leaves, leaves.size() * 8,
                                          hipMemcpyDeviceToHost, s) != hipSuccess)
        return fail(PYR_ERROR_CUDA, "autotune: leaf download");
    if (hipStreamSynchronize(s) != hipSuccess) return fail(PYR_ERROR_CUDA, "autotune: download");
    uint64_t h = 0;
    for (uint32_t v : nodes) h = mix(h, v);
    for (uint64_t v : leaves) h = mix(h, v);
    sig->sum = h;
    return PYR_OK;
}

// median of `repeats` timed builds after one untimed one; *sig from the first
PyrResult measure(const Workload& w, const PyrhipBuildTuning& t, uint32_t repeats, hipEvent_t e0, hipEvent_t e1,
                  float* ms, Signature* sig) {
    std::vector<float> times;
    for (uint32_t i = 0; i <= repeats; i++) {
        PyrhipBuilt b{};
        Signature tmp;
        if (hipEventRecord(e0, w.s) != hipSuccess) return fail(PYR_ERROR_CUDA, "autotune: event");
        PyrResult r = build_once(w, t, &b, i == 0 ? &tmp : nullptr);
        if (r != PYR_OK) return r;
        if (hipEventRecord(e1, w.s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
            pyrhip_built_free(&b, w.s);
            return fail(PYR_ERROR_CUDA, "autotune: event");
        }
        float t_ms = 0;
        (void)hipEventElapsedTime(&t_ms, e0, e1);
        if (i == 0) {
            r = signature_of(b, w.s, &tmp);
            if (r == PYR_OK) *sig = std::move(tmp);
        } else {
            times.push_back(t_ms);
        }
        pyrhip_built_free(&b, w.s);
        if (r != PYR_OK) return r;
    }
    std::sort(times.begin(), times.end());
    *ms = times[times.size() / 2];
    return PYR_OK;
}

}  // namespace

PyrhipBuildTuning resolve_tuning(const PyrhipBuildTuning* caller, uint32_t log2_size, uint32_t n,
                                 uint32_t chunk_count) {
    PyrhipBuildTuning t = kDefaults;
    {
        Table& tb = table();
        std::lock_guard<std::mutex> lk(tb.m);
        auto it = tb.classes.find(class_key(log2_size, n, chunk_count));
        if (it != tb.classes.end()) overlay(t, it->second);
    }
    if (caller) overlay(t, *caller);
    return t;
}

}  // namespace pyr

using namespace pyr;

extern "C" {

PYR_API PyrResult pyrhip_tuning_set(uint32_t log2_size, uint32_t n, uint32_t chunk_count,
                                    const PyrhipBuildTuning* tuning) {
    diag_clear();
    Table& tb = table();
    std::lock_guard<std::mutex> lk(tb.m);
    const uint64_t key = class_key(log2_size, n, chunk_count);
    if (!tuning) tb.classes.erase(key);
    else tb.classes[key] = *tuning;
    return PYR_OK;
}

PYR_API PyrResult pyrhip_tuning_get(uint32_t log2_size, uint32_t n, uint32_t chunk_count, PyrhipBuildTuning* out) {
    diag_clear();
    if (!out) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    *out = resolve_tuning(nullptr, log2_size, n, chunk_count);
    return PYR_OK;
}

PYR_API void pyrhip_tuning_clear(void) {
    Table& tb = table();
    std::lock_guard<std::mutex> lk(tb.m);
    tb.classes.clear();
}

PYR_API PyrResult pyrhip_autotune(uint32_t log2_size, uint64_t voxels, uint32_t n, const PyrhipChunkBatch* batch,
                                  uint32_t repeats, void* stream, PyrhipTuneReport* out) {
    diag_clear();
    if (!out) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    memset(out, 0, sizeof(*out));
    if (!batch && n && !voxels) return fail(PYR_ERROR_INVALID_ARGUMENT, "null voxels");
    const Workload w{log2_size, voxels, n, batch, (hipStream_t)stream};
    const uint32_t rep = std::max<uint32_t>(1, std::min<uint32_t>(repeats ? repeats : 3, 15));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        (void)hipGetLastError();
        return fail(PYR_ERROR_CUDA, "autotune: events");
    }
    PyrhipBuildTuning best = kDefaults;
    float best_ms = 0;
    Signature ref;
    PyrResult r = measure(w, best, rep, e0, e1, &best_ms, &ref);
    if (r == PYR_OK) {
        out->default_ms = best_ms;
        out->candidates = 1;
        // coordinate descent: one knob at a time, the others at the best so far;
        // a change must win by 2 % to count as more than noise
        const uint32_t sorts[] = {12, 16, 20, 24}, scans[] = {8, 12, 16, 20}, loads[] = {50, 65, 80, 90};
        // (11-bit digits are left out: 2048 look-back words per tile made a
        // pass ten times slower than 8-bit digits on every workload measured)
        const uint32_t layouts[][2] = {{8, 2}, {8, 1}, {10, 2}, {9, 2}};
        for (int dim = 0; dim < 8; dim++) {
            for (int v = 0; v < (dim >= 4 ? 2 : 4); v++) {
                PyrhipBuildTuning c = best;
                if (dim == 0) c.sort_items = sorts[v];
                if (dim == 1) c.scan_items = scans[v];
                if (dim == 2) c.radix_bits = layouts[v][0], c.pair_layout = layouts[v][1];
                if (dim == 3) c.hash_load_pct = loads[v];
                if (dim == 4) c.status_words = 2 - v;
                if (dim == 5) c.finish = 2 - v;
                if (dim == 6) c.counted_pass = 1 + v;
                if (dim == 7) c.levels = 1 + v;
                if (same_tuning(c, best)) continue;
                // the level path only matters to small builds: one workgroup
                // for every node level pays below a few tens of thousands of
                // bricks (at most the records), far above it a build would
                // take seconds, and there the default is the launch path anyway
                if (dim == 7 && count_of(w) > (1u << 17)) continue;
                // the scan tile only matters to the unique scan, not to the bucket finish
                if (dim == 1 && bucket_finish(best, log2_size, count_of(w), chunks_of(w))) continue;
                float ms = 0;
                Signature sig;
                out->candidates++;
                if (measure(w, c, rep, e0, e1, &ms, &sig) != PYR_OK || !(sig == ref)) {
                    out->rejected++;
                    (void)hipGetLastError();
                    diag_clear();
                    continue;
                }
                if (ms < best_ms * 0.98f) {
                    best = c;
                    best_ms = ms;
                }
            }
        }
        out->best = best;
        out->best_ms = best_ms;
        const uint32_t count = batch ? batch->count : 0, total = batch ? batch->total : n;
        pyrhip_tuning_set(log2_size, total, count, &best);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return r;
}

}  // extern "C"
