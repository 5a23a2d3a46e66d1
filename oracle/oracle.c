/*
 * oracle.c - CPU restatement of Pyrit's SVDAG build and software traversal.
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker, never the product.
 *
 * Built with -ffp-contract=off and no fast-math: the reference compiles with
 * strict IEEE float semantics (no @setFloatMode anywhere), so every float
 * expression below is evaluated as written, one rounding per operation.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define FLT_MAX_F 3.402823466e+38f

/* ------------------------------------------------------------------------ */
/* Morton codes (dag_builder.zig:73-100, gbuild.zig:98-121)                  */
/* ------------------------------------------------------------------------ */

static uint64_t spread3(uint32_t v) {
    uint64_t x = v & 0x1FFFFFu;
    x = (x | (x << 32)) & 0x1F00000000FFFFull;
    x = (x | (x << 16)) & 0x1F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

static uint32_t compact3(uint64_t v) {
    uint64_t x = v & 0x1249249249249249ull;
    x = (x | (x >> 2)) & 0x10C30C30C30C30C3ull;
    x = (x | (x >> 4)) & 0x100F00F00F00F00Full;
    x = (x | (x >> 8)) & 0x1F0000FF0000FFull;
    x = (x | (x >> 16)) & 0x1F00000000FFFFull;
    x = (x | (x >> 32)) & 0x1FFFFFull;
    return (uint32_t)x;
}

uint64_t or_morton(uint32_t x, uint32_t y, uint32_t z) { return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2); }

void or_morton_decode(uint64_t m, uint32_t out[3]) {
    out[0] = compact3(m);
    out[1] = compact3(m >> 1);
    out[2] = compact3(m >> 2);
}

static inline uint32_t brick_bit(uint32_t lx, uint32_t ly, uint32_t lz) { return lx + 4 * ly + 16 * lz; }

/* key of a voxel inside one chunk: brick Morton << 6 | bit (gbuild.zig:98-101) */
static inline uint64_t local_key(uint32_t x, uint32_t y, uint32_t z) {
    return (or_morton(x >> 2, y >> 2, z >> 2) << 6) | brick_bit(x & 3, y & 3, z & 3);
}

/* ------------------------------------------------------------------------ */
/* small growable arrays and a stable LSD radix sort                          */
/* ------------------------------------------------------------------------ */

typedef struct { uint32_t* p; uint64_t n, cap; } VecU32;
typedef struct { uint64_t* p; uint64_t n, cap; } VecU64;

static int vu32_push(VecU32* v, uint32_t x) {
    if (v->n == v->cap) {
        uint64_t nc = v->cap ? v->cap * 2 : 1024;
        uint32_t* np = (uint32_t*)realloc(v->p, nc * sizeof(uint32_t));
        if (!np) return -1;
        v->p = np;
        v->cap = nc;
    }
    v->p[v->n++] = x;
    return 0;
}

static int vu64_push(VecU64* v, uint64_t x) {
    if (v->n == v->cap) {
        uint64_t nc = v->cap ? v->cap * 2 : 1024;
        uint64_t* np = (uint64_t*)realloc(v->p, nc * sizeof(uint64_t));
        if (!np) return -1;
        v->p = np;
        v->cap = nc;
    }
    v->p[v->n++] = x;
    return 0;
}

/* Stable sort of (key, val) by key: equivalent to sorting by (key, input
 * order), the comparator of dag_builder.zig:342-346 and the stable LSD passes of
 * gbuild.zig:277-315. */
static int sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t n, uint32_t key_bits) {
    if (n < 2) return 0;
    uint64_t* k2 = (uint64_t*)malloc(n * 8);
    uint32_t* v2 = (uint32_t*)malloc(n * 4);
    if (!k2 || !v2) { free(k2); free(v2); return -1; }
    uint64_t* ka = keys; uint32_t* va = vals; uint64_t* kb = k2; uint32_t* vb = v2;
    uint32_t passes = (key_bits + 10) / 11;
    for (uint32_t p = 0; p < passes; p++) {
        uint32_t sh = 11 * p;
        uint64_t cnt[2048];
        memset(cnt, 0, sizeof(cnt));
        for (uint64_t i = 0; i < n; i++) cnt[(ka[i] >> sh) & 2047]++;
        uint64_t s = 0;
        for (int d = 0; d < 2048; d++) { uint64_t c = cnt[d]; cnt[d] = s; s += c; }
        for (uint64_t i = 0; i < n; i++) {
            uint64_t dst = cnt[(ka[i] >> sh) & 2047]++;
            kb[dst] = ka[i];
            vb[dst] = va[i];
        }
        uint64_t* tk = ka; ka = kb; kb = tk;
        uint32_t* tv = va; va = vb; vb = tv;
    }
    if (ka != keys) { memcpy(keys, ka, n * 8); memcpy(vals, va, n * 4); }
    free(k2);
    free(v2);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* CPU builder: dag_builder.zig:107-285                                      */
/* ------------------------------------------------------------------------ */

static uint64_t hash_words(const uint32_t* w, uint32_t len) {
    /* any hash works: dag_builder's Wyhash only locates duplicates */
    uint64_t h = 0x9E3779B97F4A7C15ull ^ len;
    for (uint32_t i = 0; i < len; i++) {
        h ^= w[i];
        h *= 0xff51afd7ed558ccdull;
        h ^= h >> 32;
    }
    return h;
}

/* NodeTable (dag_builder.zig:107-155): open addressing over offsets into nodes */
typedef struct { uint32_t* slots; uint64_t cap, used; } NodeTable;

static uint32_t node_len_at(const uint32_t* nodes, uint32_t off) {
    return 2 + (uint32_t)__builtin_popcount(nodes[off] & 0xFF);
}

static int nt_grow(NodeTable* t, const uint32_t* nodes) {
    uint64_t nc = t->cap ? t->cap * 2 : 1024;
    uint32_t* ns = (uint32_t*)malloc(nc * 4);
    if (!ns) return -1;
    memset(ns, 0xFF, nc * 4);
    for (uint64_t i = 0; i < t->cap; i++) {
        uint32_t off = t->slots[i];
        if (off == 0xFFFFFFFFu) continue;
        uint64_t h = hash_words(nodes + off, node_len_at(nodes, off)) & (nc - 1);
        while (ns[h] != 0xFFFFFFFFu) h = (h + 1) & (nc - 1);
        ns[h] = off;
    }
    free(t->slots);
    t->slots = ns;
    t->cap = nc;
    return 0;
}

/* offset of an equal node, or append words (dag_builder.zig:139-154) */
static int64_t nt_intern(NodeTable* t, VecU32* nodes, const uint32_t* words, uint32_t len) {
    if ((t->used + 1) * 4 > t->cap * 3 && nt_grow(t, nodes->p) != 0) return -1;
    uint64_t h = hash_words(words, len) & (t->cap - 1);
    while (t->slots[h] != 0xFFFFFFFFu) {
        uint32_t off = t->slots[h];
        if (node_len_at(nodes->p, off) == len && memcmp(nodes->p + off, words, len * 4) == 0) return off;
        h = (h + 1) & (t->cap - 1);
    }
    if (nodes->n + len > 0xFFFFFFFFull) return -1;
    uint32_t off = (uint32_t)nodes->n;
    for (uint32_t i = 0; i < len; i++)
        if (vu32_push(nodes, words[i])) return -1;
    t->slots[h] = off;
    t->used++;
    return off;
}

/* leaf table: mask -> leaf index */
typedef struct { uint64_t* keys; uint32_t* vals; uint64_t cap, used; } LeafTable;

static uint64_t mix64(uint64_t x) {
    x ^= x >> 33; x *= 0xff51afd7ed558ccdull; x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ull; x ^= x >> 33;
    return x;
}

static int lt_grow(LeafTable* t) {
    uint64_t nc = t->cap ? t->cap * 2 : 1024;
    uint64_t* nk = (uint64_t*)calloc(nc, 8);
    uint32_t* nv = (uint32_t*)malloc(nc * 4);
    if (!nk || !nv) { free(nk); free(nv); return -1; }
    for (uint64_t i = 0; i < t->cap; i++) {
        if (!t->keys[i]) continue;
        uint64_t h = mix64(t->keys[i]) & (nc - 1);
        while (nk[h]) h = (h + 1) & (nc - 1);
        nk[h] = t->keys[i];
        nv[h] = t->vals[i];
    }
    free(t->keys); free(t->vals);
    t->keys = nk; t->vals = nv; t->cap = nc;
    return 0;
}

/* index of mask, inserting `next` if absent; *found tells which (mask != 0) */
static int64_t lt_get_or_put(LeafTable* t, uint64_t mask, uint32_t next, int* found) {
    if ((t->used + 1) * 4 > t->cap * 3 && lt_grow(t)) return -1;
    uint64_t h = mix64(mask) & (t->cap - 1);
    while (t->keys[h]) {
        if (t->keys[h] == mask) { *found = 1; return t->vals[h]; }
        h = (h + 1) & (t->cap - 1);
    }
    t->keys[h] = mask;
    t->vals[h] = next;
    t->used++;
    *found = 0;
    return next;
}

typedef struct { int valid; uint64_t key; uint32_t mask; uint64_t count; uint32_t refs[8]; } Pending;

typedef struct {
    uint32_t log2_size; int store_attrs;
    VecU32 nodes; VecU64 leaves; VecU32 attrs;
    NodeTable nt; LeafTable lt;
    Pending pending[21];
    int have_last; uint64_t last_brick;
    uint64_t voxel_count;
} Builder;

static void add_child(Pending* p, uint32_t idx, uint32_t ref, uint64_t count) {
    /* children arrive in increasing index order (dag_builder.zig:214-220) */
    uint32_t pos = (uint32_t)__builtin_popcount(p->mask);
    p->refs[pos] = ref;
    p->mask |= 1u << idx;
    p->count += count;
}

/* dag_builder.zig:196-212 */
static int finalize_level(Builder* b, uint32_t s) {
    Pending* p = &b->pending[s];
    uint32_t words[10];
    uint32_t n = (uint32_t)__builtin_popcount(p->mask);
    words[0] = p->mask;
    words[1] = p->count > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)p->count;
    memcpy(words + 2, p->refs, n * 4);
    int64_t off = nt_intern(&b->nt, &b->nodes, words, 2 + n);
    if (off < 0) return -1;
    if (s + 1 < b->log2_size) {
        Pending* parent = &b->pending[s + 1];
        if (!parent->valid || parent->key != (p->key >> 3)) return -2;
        add_child(parent, (uint32_t)(p->key & 7), (uint32_t)off, p->count);
    } else {
        p->refs[0] = (uint32_t)off; /* root */
    }
    p->valid = 0;
    return 0;
}

/* dag_builder.zig:224-257 */
static int add_brick(Builder* b, uint64_t m, uint64_t mask, const uint32_t* attrs, uint32_t cnt) {
    if (b->have_last && m <= b->last_brick) return -1;
    b->have_last = 1;
    b->last_brick = m;
    uint32_t top = b->log2_size - 1;
    for (uint32_t s = 2; s <= top; s++) {
        uint64_t key = m >> (3 * (s - 1));
        if (b->pending[s].valid && b->pending[s].key != key) {
            if (finalize_level(b, s)) return -1;
        } else break;
    }
    for (uint32_t s = 2; s <= top; s++) {
        if (!b->pending[s].valid) {
            memset(&b->pending[s], 0, sizeof(Pending));
            b->pending[s].valid = 1;
            b->pending[s].key = m >> (3 * (s - 1));
        }
    }
    int found = 0;
    int64_t li = lt_get_or_put(&b->lt, mask, (uint32_t)b->leaves.n, &found);
    if (li < 0) return -1;
    if (!found && vu64_push(&b->leaves, mask)) return -1;
    add_child(&b->pending[2], (uint32_t)(m & 7), (uint32_t)li, cnt);
    b->voxel_count += cnt;
    if (b->store_attrs)
        for (uint32_t i = 0; i < cnt; i++)
            if (vu32_push(&b->attrs, attrs[i])) return -1;
    return 0;
}

static void builder_release(Builder* b) {
    free(b->nodes.p); free(b->leaves.p); free(b->attrs.p);
    free(b->nt.slots); free(b->lt.keys); free(b->lt.vals);
}

/* dag_builder.zig:260-284 */
static int builder_finish(Builder* b, OrDag* out) {
    uint32_t top = b->log2_size - 1;
    if (!b->pending[top].valid) {
        memset(&b->pending[top], 0, sizeof(Pending));
        b->pending[top].valid = 1;
    }
    for (uint32_t s = 2; s <= top; s++)
        if (b->pending[s].valid && finalize_level(b, s)) return -1;
    out->log2_size = b->log2_size;
    out->root = b->pending[top].refs[0];
    out->voxel_count = b->voxel_count;
    out->nodes = b->nodes.p; out->node_words = b->nodes.n;
    out->leaves = b->leaves.p; out->leaf_count = b->leaves.n;
    if (b->store_attrs) {
        out->attrs = b->attrs.p; out->attr_count = b->attrs.n;
        if (!out->attrs) out->attrs = (uint32_t*)malloc(4);
    } else {
        free(b->attrs.p);
        out->attrs = NULL; out->attr_count = 0;
    }
    if (!out->leaves) out->leaves = (uint64_t*)malloc(8);
    b->nodes.p = NULL; b->leaves.p = NULL; b->attrs.p = NULL;
    free(b->nt.slots); free(b->lt.keys); free(b->lt.vals);
    return 0;
}

static int builder_init(Builder* b, uint32_t log2_size, int store_attrs) {
    memset(b, 0, sizeof(*b));
    if (log2_size < 3 || log2_size > 20) return -1;
    b->log2_size = log2_size;
    b->store_attrs = store_attrs;
    return 0;
}

/* buildPoints (dag_builder.zig:325-368): last entry wins on duplicates */
int or_build_points(uint32_t log2_size, const OrVoxel* pts, size_t count, int store_attrs, OrDag* out) {
    Builder b;
    memset(out, 0, sizeof(*out));
    if (builder_init(&b, log2_size, store_attrs)) return -1;
    int64_t n = (int64_t)1 << log2_size;
    uint64_t* keys = (uint64_t*)malloc((count ? count : 1) * 8);
    uint32_t* vals = (uint32_t*)malloc((count ? count : 1) * 4);
    if (!keys || !vals) { free(keys); free(vals); return -3; }
    for (size_t i = 0; i < count; i++) {
        const OrVoxel* p = &pts[i];
        if (p->x < 0 || p->y < 0 || p->z < 0 || p->x >= n || p->y >= n || p->z >= n || p->attribute == 0) {
            free(keys); free(vals);
            return -1;
        }
        keys[i] = local_key((uint32_t)p->x, (uint32_t)p->y, (uint32_t)p->z);
        vals[i] = p->attribute;
    }
    if (sort_pairs(keys, vals, count, 3 * (log2_size - 2) + 6)) { free(keys); free(vals); return -3; }
    uint32_t attrs[64];
    size_t i = 0;
    int rc = 0;
    while (i < count && rc == 0) {
        uint64_t brick = keys[i] >> 6;
        uint64_t mask = 0;
        uint32_t k = 0;
        for (; i < count && (keys[i] >> 6) == brick; i++) {
            uint64_t bm = 1ull << (keys[i] & 63);
            if (mask & bm) attrs[k - 1] = vals[i];
            else { mask |= bm; attrs[k++] = vals[i]; }
        }
        rc = add_brick(&b, brick, mask, attrs, k);
    }
    free(keys); free(vals);
    if (rc || builder_finish(&b, out)) { builder_release(&b); return -1; }
    return 0;
}

/* buildDense (dag_builder.zig:292-320): voxels[x + n*(y + n*z)] */
int or_build_dense(uint32_t log2_size, const uint32_t* voxels, int store_attrs, OrDag* out) {
    Builder b;
    memset(out, 0, sizeof(*out));
    if (builder_init(&b, log2_size, store_attrs)) return -1;
    uint64_t n = 1ull << log2_size;
    uint64_t bricks = (n / 4) * (n / 4) * (n / 4);
    uint32_t attrs[64];
    for (uint64_t m = 0; m < bricks; m++) {
        uint32_t c[3];
        or_morton_decode(m, c);
        uint64_t mask = 0;
        uint32_t k = 0;
        for (uint32_t bit = 0; bit < 64; bit++) {
            uint64_t x = c[0] * 4 + (bit & 3), y = c[1] * 4 + ((bit >> 2) & 3), z = c[2] * 4 + (bit >> 4);
            uint32_t v = voxels[x + n * (y + n * z)];
            if (v) { mask |= 1ull << bit; attrs[k++] = v; }
        }
        if (mask && add_brick(&b, m, mask, attrs, k)) { builder_release(&b); return -1; }
    }
    if (builder_finish(&b, out)) { builder_release(&b); return -1; }
    return 0;
}

void or_dag_free(OrDag* d) {
    free(d->nodes); free(d->leaves); free(d->attrs);
    memset(d, 0, sizeof(*d));
}

/* Dag.lookup (dag_builder.zig:35-61); 0 = empty (attributes are never 0) */
uint32_t or_lookup(const uint32_t* nodes, const uint64_t* leaves, const uint32_t* attrs, uint32_t root,
                   uint32_t log2_size, uint32_t x, uint32_t y, uint32_t z) {
    uint32_t n = 1u << log2_size;
    if (x >= n || y >= n || z >= n) return 0;
    uint32_t node = root;
    uint64_t rank = 0;
    for (uint32_t s = log2_size - 1;; s--) {
        uint32_t mask = nodes[node] & 0xFF;
        uint32_t idx = ((x >> s) & 1) | (((y >> s) & 1) << 1) | (((z >> s) & 1) << 2);
        if (!((mask >> idx) & 1)) return 0;
        uint32_t before = (uint32_t)__builtin_popcount(mask & ((1u << idx) - 1));
        for (uint32_t k = 0; k < before; k++) {
            uint32_t ref = nodes[node + 2 + k];
            rank += s == 2 ? (uint64_t)__builtin_popcountll(leaves[ref]) : nodes[ref + 1];
        }
        uint32_t child = nodes[node + 2 + before];
        if (s == 2) {
            uint64_t brick = leaves[child];
            uint32_t bit = brick_bit(x & 3, y & 3, z & 3);
            if (!((brick >> bit) & 1)) return 0;
            rank += (uint64_t)__builtin_popcountll(brick & ((1ull << bit) - 1));
            return attrs ? attrs[rank] : 1;
        }
        node = child;
    }
}

/* ------------------------------------------------------------------------ */
/* Serial GPU build: the gbuild.zig ops in the order of gpu_build.zig:135-360 */
/* ------------------------------------------------------------------------ */

static inline uint32_t local_bits(uint32_t log2_size) { return 3 * (log2_size - 2) + 6; } /* gbuild.zig:81-83 */

static inline uint32_t sat_add(uint32_t a, uint32_t b) { /* gbuild.zig:139-142 */
    uint32_t r = a + b;
    return r < a ? 0xFFFFFFFFu : r;
}

/* tuple table for per-level dedup: words -> first index (owner = min index,
 * gbuild.zig:450-454, 527-571) */
typedef struct { int64_t* slots; uint64_t cap; } TupleTable;

/* children of one level in key order */
typedef struct { uint64_t* key; uint32_t* ref; uint32_t* count; uint64_t n; } Level;

static int gb_core(uint32_t log2_size, uint32_t chunk_bits, const uint8_t* chunk_log2, uint32_t chunk_count,
                   uint64_t* keys, uint32_t* vals, uint64_t n_total, OrBuilt* out, uint32_t* roots,
                   uint32_t* first_voxel) {
    memset(out, 0, sizeof(*out));
    out->log2_size = log2_size;
    uint32_t lbits = local_bits(log2_size);
    uint64_t invalid = 1ull << (lbits + chunk_bits); /* gbuild.zig:94-96 */
    if (sort_pairs(keys, vals, n_total, lbits + chunk_bits + 1)) return -3;

    /* unique_flags / unique_scatter (gbuild.zig:366-388) */
    uint64_t* ukey = (uint64_t*)malloc((n_total ? n_total : 1) * 8);
    uint32_t* uval = (uint32_t*)malloc((n_total ? n_total : 1) * 4);
    if (!ukey || !uval) { free(ukey); free(uval); return -3; }
    uint64_t u = 0;
    for (uint64_t i = 0; i < n_total; i++) {
        int last = i + 1 == n_total || keys[i + 1] != keys[i];
        int valid = (keys[i] & invalid) == 0 && vals[i] != 0;
        if (last && valid) { ukey[u] = keys[i]; uval[u] = vals[i]; u++; }
    }
    out->keys = ukey; out->attrs = uval; out->voxel_count = (uint32_t)u;

    /* chunk_first (gbuild.zig:237-242) */
    if (first_voxel) {
        for (uint32_t c = 0; c < chunk_count; c++) first_voxel[c] = 0xFFFFFFFFu;
        for (uint64_t i = 0; i < u; i++) {
            uint32_t c = (uint32_t)(ukey[i] >> lbits);
            if (i == 0 || (uint32_t)(ukey[i - 1] >> lbits) != c) first_voxel[c] = (uint32_t)i;
        }
    }
    if (roots) for (uint32_t c = 0; c < chunk_count; c++) roots[c] = 0xFFFFFFFFu;

    /* bricks (gbuild.zig:394-428) */
    Level ch = {0};
    ch.key = (uint64_t*)malloc((u ? u : 1) * 8);
    ch.ref = (uint32_t*)malloc((u ? u : 1) * 4);
    ch.count = (uint32_t*)malloc((u ? u : 1) * 4);
    uint64_t* masks = (uint64_t*)malloc((u ? u : 1) * 8);
    for (uint64_t i = 0; i < u;) {
        uint64_t brick = ukey[i] >> 6, mask = 0;
        uint64_t j = i;
        for (; j < u && (ukey[j] >> 6) == brick; j++) mask |= 1ull << (ukey[j] & 63);
        ch.key[ch.n] = brick;
        ch.count[ch.n] = (uint32_t)(j - i);
        masks[ch.n] = mask;
        ch.n++;
        i = j;
    }

    /* leaves: first appearance order (gbuild.zig:436-469) */
    LeafTable lt = {0};
    VecU64 leaves = {0};
    for (uint64_t b = 0; b < ch.n; b++) {
        int found = 0;
        int64_t li = lt_get_or_put(&lt, masks[b], (uint32_t)leaves.n, &found);
        if (!found) vu64_push(&leaves, masks[b]);
        ch.ref[b] = (uint32_t)li;
    }
    free(lt.keys); free(lt.vals); free(masks);

    /* node levels (gbuild.zig:475-571, gpu_build.zig:313-335) */
    VecU32 nodes = {0};
    uint32_t cursor = 0, root = 0;
    uint32_t* words = (uint32_t*)malloc((ch.n ? ch.n : 1) * 10 * 4);
    uint32_t* wlen = (uint32_t*)malloc((ch.n ? ch.n : 1) * 4);
    for (uint32_t lv = 2; lv < log2_size; lv++) {
        Level par = {0};
        par.key = (uint64_t*)malloc((ch.n ? ch.n : 1) * 8);
        par.ref = (uint32_t*)malloc((ch.n ? ch.n : 1) * 4);
        par.count = (uint32_t*)malloc((ch.n ? ch.n : 1) * 4);
        /* node_heads / node_build: one parent per run of equal key >> 3 */
        for (uint64_t i = 0; i < ch.n;) {
            uint64_t pk = ch.key[i] >> 3;
            uint32_t* w = words + par.n * 10;
            uint32_t mask = 0, count = 0, k = 0;
            uint64_t j = i;
            for (; j < ch.n && (ch.key[j] >> 3) == pk; j++) {
                mask |= 1u << (ch.key[j] & 7);
                w[2 + k++] = ch.ref[j];
                count = sat_add(count, ch.count[j]);
            }
            w[0] = mask;
            w[1] = count;
            wlen[par.n] = 2 + k;
            par.key[par.n] = pk;
            par.count[par.n] = count;
            par.n++;
            i = j;
        }
        /* node_insert/owner/first/write: owner = first equal tuple; word
         * offsets = cursor + exclusive scan of owner lengths */
        TupleTable tt;
        tt.cap = 1024;
        while (tt.cap < par.n * 2) tt.cap *= 2;
        tt.slots = (int64_t*)malloc(tt.cap * 8);
        memset(tt.slots, 0xFF, tt.cap * 8);
        uint32_t level_words = 0;
        for (uint64_t p = 0; p < par.n; p++) {
            const uint32_t* w = words + p * 10;
            uint64_t h = hash_words(w, wlen[p]) & (tt.cap - 1);
            int64_t owner = -1;
            while (tt.slots[h] >= 0) {
                int64_t o = tt.slots[h];
                if (wlen[o] == wlen[p] && memcmp(words + o * 10, w, wlen[p] * 4) == 0) { owner = o; break; }
                h = (h + 1) & (tt.cap - 1);
            }
            if (owner < 0) {
                tt.slots[h] = (int64_t)p;
                par.ref[p] = cursor + level_words;
                for (uint32_t k = 0; k < wlen[p]; k++) vu32_push(&nodes, w[k]);
                level_words += wlen[p];
            } else {
                par.ref[p] = par.ref[owner];
            }
        }
        free(tt.slots);
        /* roots: uniform batches at the top level (gbuild.zig:245-248); mixed
         * sizes finish chunk c at level chunk_log2[c] - 1 and drop it */
        uint64_t keep = 0;
        for (uint64_t p = 0; p < par.n; p++) {
            int finished = 0;
            if (roots) {
                uint32_t c = (uint32_t)(par.key[p] >> (3 * (log2_size - 1 - lv)));
                uint32_t cl = chunk_log2 ? chunk_log2[c] : log2_size;
                if (cl == lv + 1) { roots[c] = par.ref[p]; finished = 1; }
            }
            if (finished && chunk_log2) continue;
            par.key[keep] = par.key[p];
            par.ref[keep] = par.ref[p];
            par.count[keep] = par.count[p];
            keep++;
        }
        if (lv + 1 == log2_size) root = par.n > 0 ? par.ref[0] : 0; /* level_advance, gbuild.zig:607-612 */
        par.n = keep;
        cursor += level_words;
        free(ch.key); free(ch.ref); free(ch.count);
        ch = par;
    }
    free(ch.key); free(ch.ref); free(ch.count);
    free(words); free(wlen);
    /* finalize: empty geometry is a root without children (gbuild.zig:615-624) */
    if (u == 0) {
        vu32_push(&nodes, 0);
        vu32_push(&nodes, 0);
        root = cursor;
        cursor += 2;
    }
    out->nodes = nodes.p ? nodes.p : (uint32_t*)malloc(4);
    out->node_words = cursor;
    out->leaves = leaves.p ? leaves.p : (uint64_t*)malloc(8);
    out->leaf_count = (uint32_t)leaves.n;
    out->root = root;
    return 0;
}

/* build (gpu_build.zig:94-96) with encode (gbuild.zig:184-199) */
int or_gb_build(uint32_t log2_size, const uint64_t* old_keys, const uint32_t* old_vals, uint32_t n_old,
                const OrVoxel* edits, uint32_t n_new, OrBuilt* out) {
    if (log2_size < 3 || log2_size > 20) return -1;
    uint64_t n = (uint64_t)n_old + n_new;
    uint64_t* keys = (uint64_t*)malloc((n ? n : 1) * 8);
    uint32_t* vals = (uint32_t*)malloc((n ? n : 1) * 4);
    if (!keys || !vals) { free(keys); free(vals); return -3; }
    if (n_old) { memcpy(keys, old_keys, (uint64_t)n_old * 8); memcpy(vals, old_vals, (uint64_t)n_old * 4); }
    int32_t nn = 1 << log2_size;
    uint64_t invalid = 1ull << local_bits(log2_size);
    for (uint32_t i = 0; i < n_new; i++) {
        const OrVoxel* v = &edits[i];
        uint64_t dst = (uint64_t)n_old + i;
        if (v->x < 0 || v->y < 0 || v->z < 0 || v->x >= nn || v->y >= nn || v->z >= nn) {
            keys[dst] = invalid;
            vals[dst] = 0;
        } else {
            keys[dst] = local_key((uint32_t)v->x, (uint32_t)v->y, (uint32_t)v->z);
            vals[dst] = v->attribute;
        }
    }
    int rc = gb_core(log2_size, 0, NULL, 1, keys, vals, n, out, NULL, NULL);
    free(keys); free(vals);
    return rc;
}

/* buildChunks (gpu_build.zig:119-122) with pack_chunks (gbuild.zig:203-230);
 * capacity 0 reads chunk c's voxels at voxels[offsets[c] ..] (packed layout,
 * this repository's extension) instead of voxels[c * capacity ..] */
int or_gb_build_chunks(uint32_t log2_size, const uint8_t* chunk_log2, uint32_t count, uint32_t capacity,
                       const OrVoxel* voxels, const uint32_t* offsets, OrBuilt* out, uint32_t* roots,
                       uint32_t* first_voxel) {
    if (log2_size < 3 || log2_size > 20 || count == 0) return -1;
    uint32_t bits = 0;
    while ((1u << bits) < count) bits++;
    if (local_bits(log2_size) + bits + 1 > 64) return -1;
    uint64_t total = offsets[count];
    uint64_t* keys = (uint64_t*)malloc((total ? total : 1) * 8);
    uint32_t* vals = (uint32_t*)malloc((total ? total : 1) * 4);
    if (!keys || !vals) { free(keys); free(vals); return -3; }
    uint32_t lbits = local_bits(log2_size);
    uint64_t invalid = 1ull << (lbits + bits);
    for (uint64_t i = 0; i < total; i++) {
        uint32_t lo = 0, hi = count;
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) / 2;
            if (offsets[mid] <= i) lo = mid; else hi = mid;
        }
        uint32_t c = lo;
        uint64_t j = i - offsets[c];
        uint64_t key = invalid;
        uint32_t val = 0;
        int32_t n = 1 << (chunk_log2 ? chunk_log2[c] : log2_size);
        if (capacity == 0 || j < capacity) {
            const OrVoxel* v = capacity ? &voxels[(uint64_t)c * capacity + j] : &voxels[i];
            if (v->x >= 0 && v->y >= 0 && v->z >= 0 && v->x < n && v->y < n && v->z < n) {
                key = ((uint64_t)c << lbits) | local_key((uint32_t)v->x, (uint32_t)v->y, (uint32_t)v->z);
                val = v->attribute;
            }
        }
        keys[i] = key;
        vals[i] = val;
    }
    int rc = gb_core(log2_size, bits, chunk_log2, count, keys, vals, total, out, roots, first_voxel);
    free(keys); free(vals);
    return rc;
}

void or_built_free(OrBuilt* b) {
    free(b->nodes); free(b->leaves); free(b->keys); free(b->attrs);
    memset(b, 0, sizeof(*b));
}

/* ------------------------------------------------------------------------ */
/* Canonical form                                                            */
/* ------------------------------------------------------------------------ */

typedef struct {
    const uint32_t* nodes; const uint64_t* leaves; uint32_t log2_size;
    VecU32 out_nodes; VecU64 out_leaves; NodeTable nt;
    /* memo: (level, input offset) -> output offset */
    uint64_t* mkeys; uint32_t* mvals; uint64_t mcap, mused;
    /* leaf mask -> output leaf (the full mask ~0 is the map's empty key: kept apart) */
    uint64_t* lkeys; uint32_t* lvals; uint64_t lcap, lused;
    int64_t full_leaf;
} Canon;

static int map_get(uint64_t* keys, uint32_t* vals, uint64_t cap, uint64_t key, uint32_t* val) {
    uint64_t h = mix64(key) & (cap - 1);
    while (keys[h] != ~0ull) {
        if (keys[h] == key) { *val = vals[h]; return 1; }
        h = (h + 1) & (cap - 1);
    }
    return 0;
}

static int map_put(uint64_t** keys, uint32_t** vals, uint64_t* cap, uint64_t* used, uint64_t key, uint32_t val) {
    if ((*used + 1) * 2 > *cap) {
        uint64_t nc = *cap ? *cap * 2 : 1024;
        uint64_t* nk = (uint64_t*)malloc(nc * 8);
        uint32_t* nv = (uint32_t*)malloc(nc * 4);
        if (!nk || !nv) return -1;
        memset(nk, 0xFF, nc * 8);
        for (uint64_t i = 0; i < *cap; i++) {
            if ((*keys)[i] == ~0ull) continue;
            uint64_t h = mix64((*keys)[i]) & (nc - 1);
            while (nk[h] != ~0ull) h = (h + 1) & (nc - 1);
            nk[h] = (*keys)[i];
            nv[h] = (*vals)[i];
        }
        free(*keys); free(*vals);
        *keys = nk; *vals = nv; *cap = nc;
    }
    uint64_t h = mix64(key) & (*cap - 1);
    while ((*keys)[h] != ~0ull) h = (h + 1) & (*cap - 1);
    (*keys)[h] = key;
    (*vals)[h] = val;
    (*used)++;
    return 0;
}

/* node at input offset `off` whose children have edge 2^s */
static int64_t canon_node(Canon* c, uint32_t off, uint32_t s) {
    uint64_t mk = ((uint64_t)s << 40) | off;
    uint32_t got;
    if (c->mcap && map_get(c->mkeys, c->mvals, c->mcap, mk, &got)) return got;
    uint32_t mask = c->nodes[off] & 0xFF;
    uint32_t n = (uint32_t)__builtin_popcount(mask);
    uint32_t words[10];
    words[0] = c->nodes[off];
    words[1] = c->nodes[off + 1];
    for (uint32_t k = 0; k < n; k++) {
        uint32_t ref = c->nodes[off + 2 + k];
        if (s == 2) {
            /* equal masks are one leaf, whatever input slots hold them */
            const uint64_t m = c->leaves[ref];
            uint32_t li = 0;
            int found;
            if (m == ~0ull) {
                found = c->full_leaf >= 0;
                if (found) li = (uint32_t)c->full_leaf;
            } else {
                found = c->lcap && map_get(c->lkeys, c->lvals, c->lcap, m, &li);
            }
            if (!found) {
                li = (uint32_t)c->out_leaves.n;
                if (vu64_push(&c->out_leaves, m)) return -1;
                if (m == ~0ull) c->full_leaf = li;
                else if (map_put(&c->lkeys, &c->lvals, &c->lcap, &c->lused, m, li)) return -1;
            }
            words[2 + k] = li;
        } else {
            int64_t r = canon_node(c, ref, s - 1);
            if (r < 0) return -1;
            words[2 + k] = (uint32_t)r;
        }
    }
    int64_t o = nt_intern(&c->nt, &c->out_nodes, words, 2 + n);
    if (o < 0) return -1;
    if (map_put(&c->mkeys, &c->mvals, &c->mcap, &c->mused, mk, (uint32_t)o)) return -1;
    return o;
}

int or_canon(const uint32_t* nodes, const uint64_t* leaves, uint32_t root, uint32_t log2_size, OrDag* out) {
    Canon c;
    memset(&c, 0, sizeof(c));
    memset(out, 0, sizeof(*out));
    c.nodes = nodes; c.leaves = leaves; c.log2_size = log2_size;
    c.full_leaf = -1;
    int64_t r = canon_node(&c, root, log2_size - 1);
    free(c.nt.slots); free(c.mkeys); free(c.mvals); free(c.lkeys); free(c.lvals);
    if (r < 0) { free(c.out_nodes.p); free(c.out_leaves.p); return -1; }
    out->log2_size = log2_size;
    out->root = (uint32_t)r;
    out->voxel_count = nodes[root + 1];
    out->nodes = c.out_nodes.p; out->node_words = c.out_nodes.n;
    out->leaves = c.out_leaves.p ? c.out_leaves.p : (uint64_t*)malloc(8);
    out->leaf_count = c.out_leaves.n;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Traversal: dag.zig:48-351                                                 */
/* ------------------------------------------------------------------------ */

static inline uint32_t bit_of(uint32_t v, uint32_t s) { return (v >> s) & 1; }

/* Optional per-thread record of the distinct node words, leaves and
 * attributes one ray's traversal reads (instrumentation for the traversal's
 * algorithmic bytes, or_trace_touch; off otherwise). Entries are addresses,
 * so several geometries of a scene never alias. */
#define TOUCH_SLOTS 16384u
typedef struct Touch {
    uintptr_t key[TOUCH_SLOTS];
    uint32_t gen[TOUCH_SLOTS];
    uint32_t cur;
    uint64_t count[3]; /* node words, leaves, attributes */
} Touch;
static __thread Touch* g_touch;

static inline void touch(const void* addr, int kind) {
    Touch* t = g_touch;
    if (!t) return;
    uintptr_t k = (uintptr_t)addr;
    uint32_t h = (uint32_t)((k >> 2) * 0x9E3779B1u) & (TOUCH_SLOTS - 1);
    for (uint32_t probe = 0; probe < TOUCH_SLOTS; probe++, h = (h + 1) & (TOUCH_SLOTS - 1)) {
        if (t->gen[h] != t->cur) {
            t->gen[h] = t->cur;
            t->key[h] = k;
            t->count[kind]++;
            return;
        }
        if (t->key[h] == k) return;
    }
}

/* attributeRank (dag.zig:70-87); stk[s] = node whose children have edge 2^s */
static uint32_t attribute_rank(const OrDagView* g, const uint32_t* stk, const uint32_t v[3], uint64_t brick) {
    uint32_t rank = 0;
    for (uint32_t s = g->log2_size - 1; s >= 2; s--) {
        uint32_t node = stk[s];
        uint32_t mask = g->nodes[node] & 0xFF;
        touch(&g->nodes[node], 0);
        uint32_t idx = bit_of(v[0], s) | (bit_of(v[1], s) << 1) | (bit_of(v[2], s) << 2);
        uint32_t before = (uint32_t)__builtin_popcount(mask & ((1u << idx) - 1));
        for (uint32_t k = 0; k < before; k++) {
            uint32_t ref = g->nodes[node + 2 + k];
            touch(&g->nodes[node + 2 + k], 0);
            if (s == 2) touch(&g->leaves[ref], 1);
            else touch(&g->nodes[ref + 1], 0);
            rank += s == 2 ? (uint32_t)__builtin_popcountll(g->leaves[ref]) : g->nodes[ref + 1];
        }
    }
    uint32_t bit = brick_bit(v[0] & 3, v[1] & 3, v[2] & 3);
    rank += (uint32_t)__builtin_popcountll(brick & ((1ull << bit) - 1));
    return rank;
}

/* skipped (dag.zig:153-157): bit (attribute & 0xFF) of a 256-bit material set */
static int skipped(const uint64_t* set, uint32_t attribute) {
    if (!set) return 0;
    const uint32_t idx = attribute & 0xFF;
    return (int)((set[idx >> 6] >> (idx & 63)) & 1);
}

/* cutHole (dag.zig:133-151): is the ray's entry point p on face `axis` of the
 * voxel in a hole of the material's fixed 3x3 pattern? */
static int cut_hole(const uint64_t* cut, const int32_t base[3], uint32_t attribute, const int32_t voxel[3],
                    const float p[3], uint32_t axis) {
    if (!skipped(cut, attribute)) return 0;
    uint32_t sub[2] = {0, 0}, k = 0;
    for (uint32_t a = 0; a < 3; a++) {
        if (a == axis) continue;
        const float f = fminf(fmaxf(p[a] - (float)voxel[a], 0.0f), 0.999f);
        sub[k++] = (uint32_t)(f * 3.0f);
    }
    uint32_t h = (uint32_t)(voxel[0] + base[0]) * 0x8da6b343u + (uint32_t)(voxel[1] + base[1]) * 0xd8163841u +
                 (uint32_t)(voxel[2] + base[2]) * 0xcb1ab31fu + (attribute >> 8) * 0x2545f491u + axis * 0x9e3779b9u +
                 (sub[0] * 3 + sub[1]) * 0x632be5abu;
    h ^= h >> 15;
    h *= 0x2c1b3c6du;
    h ^= h >> 12;
    return (h & 0xFF) < 77;
}

/* walk with exit_mode = false (dag.zig:160-351): skip / cut may be NULL */
static int walk_impl(const OrDagView* g, const float o[3], const float d[3], float tmin, float tmax, int want_attr,
                     const uint64_t* skip, const uint64_t* cut, const int32_t base[3], OrDagHit* hit,
                     uint32_t* steps_out) {
    const uint32_t n = 1u << g->log2_size;
    const float nf = (float)n;
    uint32_t oct = 0;
    float od[3], dd[3], inv[3];
    for (int a = 0; a < 3; a++) {
        float oa = o[a], da = d[a];
        if (da < 0) { oa = nf - oa; da = -da; oct |= 1u << a; }
        od[a] = oa;
        dd[a] = da;
        inv[a] = 1.0f / fmaxf(da, 1e-30f);
    }
    uint32_t axis = 0;
    float tn = -od[0] * inv[0];
    for (int a = 1; a < 3; a++) {
        float ta = -od[a] * inv[a];
        if (ta > tn) { tn = ta; axis = (uint32_t)a; }
    }
    float tf = fminf(fminf((nf - od[0]) * inv[0], (nf - od[1]) * inv[1]), (nf - od[2]) * inv[2]);
    float t = fmaxf(tn, tmin);
    const float tend = fminf(tf, tmax);
    if (steps_out) *steps_out = 0;
    if (!(t < tend)) return 0;
    const int starts_inside = tn < tmin;
    int moved = 0;

    uint32_t scale = g->log2_size - 1;
    uint32_t cell[3];
    for (int a = 0; a < 3; a++) {
        float p = fminf(fmaxf(floorf(od[a] + dd[a] * t), 0.0f), nf - 1.0f);
        cell[a] = (uint32_t)p & ~((1u << scale) - 1);
    }
    uint32_t stk[21];
    stk[scale] = g->root;
    uint64_t brick = 0;

    for (uint32_t step = 0; step < 65536; step++) {
        if (steps_out) *steps_out = step + 1;
        int exists;
        uint32_t child = 0;
        if (scale >= 2) {
            uint32_t node = stk[scale];
            uint32_t mask = g->nodes[node] & 0xFF;
            touch(&g->nodes[node], 0);
            uint32_t idx = (bit_of(cell[0], scale) | (bit_of(cell[1], scale) << 1) | (bit_of(cell[2], scale) << 2)) ^ oct;
            uint32_t bit = 1u << idx;
            exists = (mask & bit) != 0;
            if (exists) {
                child = g->nodes[node + 2 + __builtin_popcount(mask & (bit - 1))];
                touch(&g->nodes[node + 2 + __builtin_popcount(mask & (bit - 1))], 0);
            }
        } else {
            uint32_t size = 1u << scale, l[3];
            for (int a = 0; a < 3; a++) l[a] = (((oct >> a) & 1) ? n - size - cell[a] : cell[a]) & 3;
            if (scale == 1) exists = (brick & (0x00330033ull << brick_bit(l[0], l[1], l[2]))) != 0;
            else exists = ((brick >> brick_bit(l[0], l[1], l[2])) & 1) != 0;
        }
        if (exists) {
            if (scale == 0) {
                uint32_t v[3];
                for (int a = 0; a < 3; a++) {
                    v[a] = ((oct >> a) & 1) ? n - 1 - cell[a] : cell[a];
                    hit->voxel[a] = (int32_t)v[a];
                }
                uint32_t attribute = g->default_attribute;
                if ((want_attr || skip || cut) && g->attrs) {
                    const uint32_t r = attribute_rank(g, stk, v, brick);
                    touch(&g->attrs[g->attr_base + r], 2);
                    attribute = g->attrs[g->attr_base + r];
                }
                /* transparent material, or a hole of a cutout material: go on
                 * as if the voxel were empty (a ray that starts inside a
                 * cutout voxel passes it) */
                int through = 0;
                if (starts_inside && !moved) {
                    through = skipped(cut, attribute);
                } else if (cut) {
                    float p[3];
                    for (int a = 0; a < 3; a++) p[a] = o[a] + d[a] * t;
                    through = cut_hole(cut, base, attribute, hit->voxel, p, axis);
                }
                if (skipped(skip, attribute) || through) {
                    scale = 0;
                    float tnext0 = ((float)(cell[0] + 1) - od[0]) * inv[0];
                    axis = 0;
                    for (int a = 1; a < 3; a++) {
                        float ta = ((float)(cell[a] + 1) - od[a]) * inv[a];
                        if (ta < tnext0) { tnext0 = ta; axis = (uint32_t)a; }
                    }
                    if (!(tnext0 < tend)) return 0;
                    t = fmaxf(t, tnext0);
                    moved = 1;
                    uint32_t old0 = cell[axis], next0 = old0 + 1;
                    if (next0 >= n) return 0;
                    cell[axis] = next0;
                    uint32_t h0 = 31 - (uint32_t)__builtin_clz(old0 ^ next0);
                    if (h0 > scale) {
                        scale = h0;
                        uint32_t m0 = ~((1u << scale) - 1);
                        for (int a = 0; a < 3; a++) cell[a] &= m0;
                    }
                    continue;
                }
                int mirrored = ((oct >> axis) & 1) != 0;
                hit->t = t;
                hit->face = 2 * axis + (mirrored ? 0u : 1u);
                hit->attribute = attribute;
                hit->inside = (uint32_t)(starts_inside && !moved);
                return 1;
            }
            if (scale >= 3) stk[scale - 1] = child;
            else if (scale == 2) {
                brick = g->leaves[child];
                touch(&g->leaves[child], 1);
            }
            scale -= 1;
            uint32_t half = 1u << scale;
            float halff = (float)half;
            for (int a = 0; a < 3; a++) {
                float p = od[a] + dd[a] * t;
                if (p - (float)cell[a] >= halff) cell[a] += half;
            }
            continue;
        }
        uint32_t size = 1u << scale;
        float tnext = ((float)(cell[0] + size) - od[0]) * inv[0];
        axis = 0;
        for (int a = 1; a < 3; a++) {
            float ta = ((float)(cell[a] + size) - od[a]) * inv[a];
            if (ta < tnext) { tnext = ta; axis = (uint32_t)a; }
        }
        if (!(tnext < tend)) return 0;
        t = fmaxf(t, tnext);
        moved = 1;
        uint32_t old = cell[axis];
        uint32_t next = old + size;
        if (next >= n) return 0;
        cell[axis] = next;
        uint32_t h = 31 - (uint32_t)__builtin_clz(old ^ next);
        if (h > scale) {
            scale = h;
            uint32_t m = ~((1u << scale) - 1);
            for (int a = 0; a < 3; a++) cell[a] &= m;
        }
    }
    return 0;
}

int or_walk(const OrDagView* g, const float o[3], const float d[3], float tmin, float tmax, int want_attr,
            OrDagHit* hit, uint32_t* steps_out) {
    const int32_t base[3] = {0, 0, 0};
    return walk_impl(g, o, d, tmin, tmax, want_attr, NULL, NULL, base, hit, steps_out);
}

/* traceSkipping (dag.zig:113-115) */
int or_walk_skipping(const OrDagView* g, const float o[3], const float d[3], float tmin, float tmax, int want_attr,
                     const uint64_t skip[4], const uint64_t cut[4], const int32_t base[3], OrDagHit* hit) {
    return walk_impl(g, o, d, tmin, tmax, want_attr, skip, cut, base, hit, NULL);
}

/* anyCutout / cutoutHere (trace.zig:161-173): cutouts only where a voxel
 * edge is at most about 1.4 world units (row 0 of world_to_object) */
static int any_cutout(const OrScene* s) {
    return (s->cutout[0] | s->cutout[1] | s->cutout[2] | s->cutout[3]) != 0;
}
static int cutout_here(const OrInstance* in) {
    float w2o = 0.0f;
    for (int a = 0; a < 3; a++) w2o += in->world_to_object[a] * in->world_to_object[a];
    return w2o > 0.5f;
}

/* safeRcp (trace.zig:56-58) */
static inline float safe_rcp(float x) {
    return 1.0f / (fabsf(x) < 1e-30f ? (x < 0 ? -1e-30f : 1e-30f) : x);
}

/* xformPoint / xformVector (vec.zig:26-40), evaluated left to right */
static void xform_point(const float* m, const float p[3], float r[3]) {
    r[0] = m[0] * p[0] + m[1] * p[1] + m[2] * p[2] + m[3];
    r[1] = m[4] * p[0] + m[5] * p[1] + m[6] * p[2] + m[7];
    r[2] = m[8] * p[0] + m[9] * p[1] + m[10] * p[2] + m[11];
}

static void xform_vector(const float* m, const float v[3], float r[3]) {
    r[0] = m[0] * v[0] + m[1] * v[1] + m[2] * v[2];
    r[1] = m[4] * v[0] + m[5] * v[1] + m[6] * v[2];
    r[2] = m[8] * v[0] + m[9] * v[1] + m[10] * v[2];
}

/* traceScene + toHit (trace.zig:72-118); p_object (may be NULL) gets the
 * hit point in the hit instance's object space (trace.zig:107) */
static void trace_one(const OrScene* s, const float o[3], const float d[3], float tmin, float tmax,
                      uint32_t ray_mask, uint32_t flags, OrHit* out, float* p_object) {
    float inv[3] = {safe_rcp(d[0]), safe_rcp(d[1]), safe_rcp(d[2])};
    int want_attr = (flags & 0x2u) == 0;
    float best = tmax;
    out->t = FLT_MAX_F; out->instance = 0xFFFFFFFFu; out->attribute = 0; out->meta = 0;
    for (uint32_t i = 0; i < s->instance_count; i++) {
        const OrInstance* in = &s->instances[i];
        if (!(in->flags & 1u) || !(in->mask & ray_mask)) continue;
        float t0 = -FLT_MAX_F, t1 = FLT_MAX_F;
        float lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            lo[a] = (in->bounds_min[a] - o[a]) * inv[a];
            hi[a] = (in->bounds_max[a] - o[a]) * inv[a];
        }
        t0 = fmaxf(fmaxf(fminf(lo[0], hi[0]), fminf(lo[1], hi[1])), fminf(lo[2], hi[2]));
        t1 = fminf(fminf(fmaxf(lo[0], hi[0]), fmaxf(lo[1], hi[1])), fmaxf(lo[2], hi[2]));
        if (!(t0 <= t1)) continue;
        if (!(fmaxf(t0, tmin) < fminf(t1, best))) continue;
        const OrGeometry* g = &s->geometries[in->geometry];
        OrDagView v;
        v.nodes = s->nodes + g->node_offset;
        v.leaves = s->leaves + g->leaf_offset;
        v.attrs = (g->flags & 1u) ? s->attrs + g->attribute_offset : NULL;
        v.root = g->root; v.log2_size = g->log2_size; v.default_attribute = g->default_attribute; v.attr_base = 0;
        float oo[3], dv[3];
        xform_point(in->world_to_object, o, oo);
        xform_vector(in->world_to_object, d, dv);
        OrDagHit h;
        /* trace.zig:80, 93-98 */
        const uint64_t* skip = (flags & 0x8u) ? s->transparent : NULL;
        const uint64_t* cut = any_cutout(s) && cutout_here(in) ? s->cutout : NULL;
        const int32_t base[3] = {0, 0, 0};
        const int found = skip || cut ? walk_impl(&v, oo, dv, tmin, best, want_attr, skip, cut, base, &h, NULL)
                                      : or_walk(&v, oo, dv, tmin, best, want_attr, &h, NULL);
        if (found) {
            best = h.t;
            out->t = h.t; out->instance = i; out->attribute = h.attribute;
            out->meta = h.face | (h.inside ? 0x20u : 0u);
            if (p_object)
                for (int a = 0; a < 3; a++) p_object[a] = oo[a] + dv[a] * h.t;
            if (flags & 0x1u) return;
        }
    }
}

typedef struct {
    const OrScene* s; const OrRay* rays; OrHit* hits; uint64_t begin, end; uint32_t mask, flags;
} TraceJob;

static void* trace_worker(void* arg) {
    TraceJob* j = (TraceJob*)arg;
    for (uint64_t i = j->begin; i < j->end; i++) {
        const OrRay* r = &j->rays[i];
        trace_one(j->s, r->origin, r->direction, r->tmin, r->tmax, j->mask, j->flags, &j->hits[i], NULL);
    }
    return NULL;
}

void or_trace_rays(const OrScene* s, const OrRay* rays, OrHit* hits, uint64_t count, uint32_t ray_mask,
                   uint32_t flags, uint32_t threads) {
    if (ray_mask == 0) ray_mask = 0xFFFFFFFFu; /* context.zig:1877 */
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    TraceJob jobs[256];
    for (uint32_t k = 0; k < threads; k++) {
        jobs[k].s = s; jobs[k].rays = rays; jobs[k].hits = hits;
        jobs[k].begin = count * k / threads; jobs[k].end = count * (k + 1) / threads;
        jobs[k].mask = ray_mask; jobs[k].flags = flags;
        if (threads == 1) trace_worker(&jobs[k]);
        else pthread_create(&th[k], NULL, trace_worker, &jobs[k]);
    }
    if (threads > 1)
        for (uint32_t k = 0; k < threads; k++) pthread_join(th[k], NULL);
}

/* extendedHit and hitNormal (trace.zig:125-157): the voxel from the hit
 * point moved half a voxel inward on the hit face's axis (unless the ray
 * started inside), the world hit point o + d t, and the world normal: the
 * face normal through the transpose of world_to_object (vec.zig:43-49),
 * normalised as v * (1 / sqrt((x x + y y) + z z)) (vec.zig:13-23). A miss
 * is all zero but for the miss hit. */
static void extended_hit(const OrScene* s, const float o[3], const float d[3], const OrHit* h, const float po[3],
                         OrHitEx* r) {
    memset(r, 0, sizeof(*r));
    r->hit = *h;
    if (h->instance == 0xFFFFFFFFu) return;
    const uint32_t face = h->meta & 7u, axis = face >> 1;
    const int inside = (h->meta & 0x20u) != 0;
    for (uint32_t a = 0; a < 3; a++) {
        float p = po[a];
        if (!inside && a == axis) p += (face & 1u) ? 0.5f : -0.5f;
        r->voxel[a] = (int32_t)floorf(p);
    }
    for (int a = 0; a < 3; a++) r->position[a] = o[a] + d[a] * h->t;
    const float sign = (face & 1u) ? -1.0f : 1.0f;
    const float n[3] = {axis == 0 ? sign : 0.0f, axis == 1 ? sign : 0.0f, axis == 2 ? sign : 0.0f};
    const float* m = s->instances[h->instance].world_to_object;
    const float v[3] = {m[0] * n[0] + m[4] * n[1] + m[8] * n[2], m[1] * n[0] + m[5] * n[1] + m[9] * n[2],
                        m[2] * n[0] + m[6] * n[1] + m[10] * n[2]};
    const float il = 1.0f / sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int a = 0; a < 3; a++) r->normal[a] = v[a] * il;
}

typedef struct {
    const OrScene* s; const OrRay* rays; OrHitEx* hits; uint64_t begin, end; uint32_t mask, flags;
} TraceExJob;

static void* trace_ex_worker(void* arg) {
    TraceExJob* j = (TraceExJob*)arg;
    for (uint64_t i = j->begin; i < j->end; i++) {
        const OrRay* r = &j->rays[i];
        OrHit h;
        float po[3] = {0, 0, 0};
        trace_one(j->s, r->origin, r->direction, r->tmin, r->tmax, j->mask, j->flags, &h, po);
        extended_hit(j->s, r->origin, r->direction, &h, po, &j->hits[i]);
    }
    return NULL;
}

void or_trace_rays_ex(const OrScene* s, const OrRay* rays, OrHitEx* hits, uint64_t count, uint32_t ray_mask,
                      uint32_t flags, uint32_t threads) {
    if (ray_mask == 0) ray_mask = 0xFFFFFFFFu;
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    TraceExJob jobs[256];
    for (uint32_t k = 0; k < threads; k++) {
        jobs[k].s = s; jobs[k].rays = rays; jobs[k].hits = hits;
        jobs[k].begin = count * k / threads; jobs[k].end = count * (k + 1) / threads;
        jobs[k].mask = ray_mask; jobs[k].flags = flags;
        if (threads == 1) trace_ex_worker(&jobs[k]);
        else pthread_create(&th[k], NULL, trace_ex_worker, &jobs[k]);
    }
    if (threads > 1)
        for (uint32_t k = 0; k < threads; k++) pthread_join(th[k], NULL);
}

void or_trace_touch(const OrScene* s, const OrRay* rays, uint64_t count, uint32_t ray_mask, uint32_t flags,
                    uint64_t out[4]) {
    if (ray_mask == 0) ray_mask = 0xFFFFFFFFu;
    Touch* t = (Touch*)calloc(1, sizeof(Touch));
    memset(out, 0, 4 * sizeof(uint64_t));
    if (!t) return;
    g_touch = t;
    for (uint64_t i = 0; i < count; i++) {
        t->cur++; /* a new ray: every slot reads as empty */
        OrHit h;
        trace_one(s, rays[i].origin, rays[i].direction, rays[i].tmin, rays[i].tmax, ray_mask, flags, &h, NULL);
        out[3] += h.instance != 0xFFFFFFFFu;
    }
    g_touch = NULL;
    out[0] = t->count[0];
    out[1] = t->count[1];
    out[2] = t->count[2];
    free(t);
}

/* ------------------------------------------------------------------------ */
/* Camera (root.zig:449-481, render.zig:16-32, 133-141)                      */
/* ------------------------------------------------------------------------ */

static void normalize3(const float v[3], float r[3]) {
    float l = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    r[0] = v[0] / l; r[1] = v[1] / l; r[2] = v[2] / l;
}

static void cross3(const float a[3], const float b[3], float r[3]) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

void or_camera_look_at(OrCamera* c, const float eye[3], const float target[3], const float up[3]) {
    float b0[3] = {eye[0] - target[0], eye[1] - target[1], eye[2] - target[2]};
    float back[3], r0[3], right[3], tu[3];
    normalize3(b0, back);
    cross3(up, back, r0);
    normalize3(r0, right);
    cross3(back, right, tu);
    float m[12] = {right[0], tu[0], back[0], eye[0], right[1], tu[1], back[1], eye[1], right[2], tu[2], back[2], eye[2]};
    memcpy(c->view_to_world, m, sizeof(m));
}

void or_camera_perspective(OrCamera* c, float fov_y, uint32_t width, uint32_t height, float near_plane) {
    float ty = tanf(fov_y * 0.5f);
    c->projection = 0;
    c->width = width;
    c->height = height;
    c->scale[0] = ty * (float)width / (float)height;
    c->scale[1] = ty;
    c->near_plane = near_plane;
}

static void camera_ray(const OrCamera* c, float px, float py, OrRay* r) {
    float nx = 2.0f * px / (float)c->width - 1.0f;
    float ny = 1.0f - 2.0f * py / (float)c->height;
    float vx = (nx + c->shift[0]) * c->scale[0];
    float vy = (ny + c->shift[1]) * c->scale[1];
    int ortho = c->projection == 1;
    float ov[3] = {ortho ? vx : 0.0f, ortho ? vy : 0.0f, 0.0f};
    float dv[3] = {ortho ? 0.0f : vx, ortho ? 0.0f : vy, -1.0f};
    float dw[3];
    xform_vector(c->view_to_world, dv, dw);
    float len = sqrtf(dw[0] * dw[0] + dw[1] * dw[1] + dw[2] * dw[2]);
    float il = 1.0f / len;
    xform_point(c->view_to_world, ov, r->origin);
    r->direction[0] = dw[0] * il;
    r->direction[1] = dw[1] * il;
    r->direction[2] = dw[2] * il;
    r->tmin = c->near_plane * len;
    r->tmax = c->far_plane > 0 ? c->far_plane * len : FLT_MAX_F;
}

void or_camera_rays(const OrCamera* c, OrRay* rays) {
    for (uint32_t y = 0; y < c->height; y++)
        for (uint32_t x = 0; x < c->width; x++) {
            float px = (float)x + 0.5f + c->jitter[0];
            float py = (float)y + 0.5f + c->jitter[1];
            camera_ray(c, px, py, &rays[(uint64_t)y * c->width + x]);
        }
}

/* inverse of an affine 3x4 matrix in f64 (xform.zig:8-34); 0 when singular */
int or_inverse(const float m[12], float out[12]) {
    const double a[9] = {m[0], m[1], m[2], m[4], m[5], m[6], m[8], m[9], m[10]};
    const double t[3] = {m[3], m[7], m[11]};
    double c00 = a[4] * a[8] - a[5] * a[7];
    double c01 = a[5] * a[6] - a[3] * a[8];
    double c02 = a[3] * a[7] - a[4] * a[6];
    double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    if (!(fabs(det) > 1e-30) || !isfinite(det)) return 0;
    double id = 1.0 / det;
    double r[9] = {
        c00 * id, (a[2] * a[7] - a[1] * a[8]) * id, (a[1] * a[5] - a[2] * a[4]) * id,
        c01 * id, (a[0] * a[8] - a[2] * a[6]) * id, (a[2] * a[3] - a[0] * a[5]) * id,
        c02 * id, (a[1] * a[6] - a[0] * a[7]) * id, (a[0] * a[4] - a[1] * a[3]) * id,
    };
    for (int row = 0; row < 3; row++) {
        double r0 = r[row * 3 + 0], r1 = r[row * 3 + 1], r2 = r[row * 3 + 2];
        out[row * 4 + 0] = (float)r0;
        out[row * 4 + 1] = (float)r1;
        out[row * 4 + 2] = (float)r2;
        out[row * 4 + 3] = (float)(-(r0 * t[0] + r1 * t[1] + r2 * t[2]));
    }
    return 1;
}

/* world AABB of the cube [0, size]^3 under m, with the 1e-5 margin
 * (xform.zig:37-58) */
void or_transformed_box(const float m[12], float size, float lo[3], float hi[3]) {
    for (int row = 0; row < 3; row++) {
        lo[row] = m[row * 4 + 3];
        hi[row] = m[row * 4 + 3];
        for (int col = 0; col < 3; col++) {
            float e = m[row * 4 + col] * size;
            lo[row] += fminf(e, 0.0f);
            hi[row] += fmaxf(e, 0.0f);
        }
    }
    for (int i = 0; i < 3; i++) {
        float eps = 1e-5f * fmaxf(1.0f, fmaxf(fabsf(lo[i]), fabsf(hi[i])));
        lo[i] -= eps;
        hi[i] += eps;
    }
}

/* primary hits and linear depth: depth = -(world_to_view * (o + d t)).z
 * (render.zig:143-166; world_to_view = inverse(view_to_world), context.zig:1771) */
void or_render_primary(const OrScene* s, const OrCamera* c, OrHit* hits, float* depth, uint32_t threads) {
    uint64_t n = (uint64_t)c->width * c->height;
    OrRay* rays = (OrRay*)malloc(n * sizeof(OrRay));
    or_camera_rays(c, rays);
    or_trace_rays(s, rays, hits, n, 0xFFFFFFFFu, 0, threads);
    if (depth) {
        float w2v[12];
        if (!or_inverse(c->view_to_world, w2v)) memset(w2v, 0, sizeof(w2v));
        for (uint64_t i = 0; i < n; i++) {
            if (hits[i].instance == 0xFFFFFFFFu) { depth[i] = FLT_MAX_F; continue; }
            float p[3], q[3];
            for (int a = 0; a < 3; a++) p[a] = rays[i].origin[a] + rays[i].direction[a] * hits[i].t;
            xform_point(w2v, p, q);
            depth[i] = -q[2];
        }
    }
    free(rays);
}

/* ------------------------------------------------------------------------ */
/* f64 DDA reference (tests/common.zig:104-153)                              */
/* ------------------------------------------------------------------------ */

int or_ref_dda(const uint32_t* grid, uint32_t n, const double o[3], const double d[3], double tmin, double tmax,
               double* t_out, int32_t v[3], int32_t* face) {
    double nf = (double)n;
    double t0 = tmin, t1 = tmax;
    int entry = -1;
    for (int a = 0; a < 3; a++) {
        if (d[a] == 0) {
            if (o[a] < 0 || o[a] >= nf) return 0;
            continue;
        }
        double ta = (0 - o[a]) / d[a], tb = (nf - o[a]) / d[a];
        if (ta > tb) { double tmp = ta; ta = tb; tb = tmp; }
        if (ta > t0) { t0 = ta; entry = a; }
        if (tb < t1) t1 = tb;
    }
    if (!(t0 < t1)) return 0;
    int32_t ni = (int32_t)n, cell[3], step[3];
    double tnext[3], tdelta[3];
    for (int a = 0; a < 3; a++) {
        double p = o[a] + d[a] * t0;
        double f = floor(p);
        if (f > 1e9) f = 1e9;
        if (f < -1e9) f = -1e9;
        int32_t c = (int32_t)f;
        if (c < 0) c = 0;
        if (c > ni - 1) c = ni - 1;
        cell[a] = c;
        step[a] = d[a] > 0 ? 1 : -1;
        tdelta[a] = d[a] != 0 ? fabs(1.0 / d[a]) : INFINITY;
        double bound = (double)(cell[a] + (d[a] > 0 ? 1 : 0));
        tnext[a] = d[a] != 0 ? (bound - o[a]) / d[a] : INFINITY;
    }
    int axis = entry < 0 ? 0 : entry;
    double t = t0;
    for (;;) {
        uint64_t idx = (uint64_t)cell[0] + (uint64_t)n * ((uint64_t)cell[1] + (uint64_t)n * (uint64_t)cell[2]);
        if (grid[idx] != 0) {
            *face = (entry < 0 && t == t0) ? -1 : 2 * axis + (d[axis] > 0 ? 1 : 0);
            *t_out = t;
            v[0] = cell[0]; v[1] = cell[1]; v[2] = cell[2];
            return 1;
        }
        axis = tnext[0] < tnext[1] ? (tnext[0] < tnext[2] ? 0 : 2) : (tnext[1] < tnext[2] ? 1 : 2);
        t = tnext[axis];
        if (t >= t1) return 0;
        cell[axis] += step[axis];
        if (cell[axis] < 0 || cell[axis] >= ni) return 0;
        tnext[axis] += tdelta[axis];
    }
}

/* ------------------------------------------------------------------------ */
/* RangeAlloc (range_alloc.zig:10-59)                                        */
/* ------------------------------------------------------------------------ */

void or_range_init(OrRangeAlloc* r, uint64_t capacity) {
    memset(r, 0, sizeof(*r));
    r->capacity = capacity;
    if (capacity > 0) {
        r->cap = 16;
        r->free = (OrRange*)malloc(r->cap * sizeof(OrRange));
        r->free[0].offset = 0;
        r->free[0].size = capacity;
        r->n = 1;
    }
}

int64_t or_range_alloc(OrRangeAlloc* r, uint64_t size) {
    if (size == 0) return 0;
    for (uint32_t i = 0; i < r->n; i++) {
        OrRange* f = &r->free[i];
        if (f->size < size) continue;
        uint64_t off = f->offset;
        f->offset += size;
        f->size -= size;
        if (f->size == 0) {
            memmove(&r->free[i], &r->free[i + 1], (r->n - i - 1) * sizeof(OrRange));
            r->n--;
        }
        r->used += size;
        return (int64_t)off;
    }
    return -1;
}

void or_range_release(OrRangeAlloc* r, uint64_t offset, uint64_t size) {
    if (size == 0) return;
    r->used -= size;
    uint32_t i = 0;
    while (i < r->n && r->free[i].offset < offset) i++;
    int mp = i > 0 && r->free[i - 1].offset + r->free[i - 1].size == offset;
    int mn = i < r->n && offset + size == r->free[i].offset;
    if (mp && mn) {
        r->free[i - 1].size += size + r->free[i].size;
        memmove(&r->free[i], &r->free[i + 1], (r->n - i - 1) * sizeof(OrRange));
        r->n--;
    } else if (mp) {
        r->free[i - 1].size += size;
    } else if (mn) {
        r->free[i].offset = offset;
        r->free[i].size += size;
    } else {
        if (r->n == r->cap) {
            r->cap = r->cap ? r->cap * 2 : 16;
            r->free = (OrRange*)realloc(r->free, r->cap * sizeof(OrRange));
        }
        memmove(&r->free[i + 1], &r->free[i], (r->n - i) * sizeof(OrRange));
        r->free[i].offset = offset;
        r->free[i].size = size;
        r->n++;
    }
}

void or_range_free(OrRangeAlloc* r) { free(r->free); memset(r, 0, sizeof(*r)); }

/* palette.zig:65-152 scan + pack, context.zig:1352-1357 word counts */
void or_palette_pack(const uint32_t* attrs, const uint32_t* first, uint32_t k, uint32_t* count,
                     uint64_t* word_offset, uint32_t* out) {
    uint64_t words = 0;
    for (uint32_t j = 0; j < k; j++) {
        const uint64_t lo = first[j], n = (uint64_t)first[j + 1] - lo;
        uint32_t pal[17];
        uint32_t m = 0;
        int full = 0;
        for (uint64_t i = 0; i < n && !full; i++) {
            const uint32_t v = attrs[lo + i];
            if (v == 0xFFFFFFFFu) {
                full = 1;
                break;
            }
            uint32_t q = 0;
            while (q < m && pal[q] != v) q++;
            if (q < m) continue;
            if (m == 16) {
                full = 1;
                break;
            }
            /* insertion into ascending order */
            uint32_t p = m++;
            while (p > 0 && pal[p - 1] > v) {
                pal[p] = pal[p - 1];
                p--;
            }
            pal[p] = v;
        }
        word_offset[j] = words;
        count[j] = full ? 0xFFFFFFFFu : m;
        if (full) {
            if (out) memcpy(out + words, attrs + lo, n * 4);
            words += n;
            continue;
        }
        for (uint32_t q = m; q < 16; q++) pal[q] = m ? pal[0] : 0u;
        if (out) {
            memcpy(out + words, pal, 64);
            for (uint64_t w = 0; w < (n + 7) / 8; w++) {
                uint32_t word = 0;
                for (uint32_t b = 0; b < 8 && w * 8 + b < n; b++) {
                    const uint32_t v = attrs[lo + w * 8 + b];
                    uint32_t idx = 0;
                    while (pal[idx] != v) idx++;
                    word |= idx << (4 * b);
                }
                out[words + 16 + w] = word;
            }
        }
        words += 16 + (n + 7) / 8;
    }
    word_offset[k] = words;
}

/* ------------------------------------------------------------------------ */
/* RT primitives: rt_prims.zig:47-137                                        */
/* ------------------------------------------------------------------------ */

/* voxel box, max exclusive (rt_prims.zig:31-44) */
typedef struct OrBox { uint32_t lo[3], hi[3]; } OrBox;

static OrBox box_empty(void) {
    OrBox b = {{0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, {0, 0, 0}};
    return b;
}

static void box_merge(OrBox* a, const OrBox* b, const uint32_t off[3]) {
    for (int i = 0; i < 3; i++) {
        if (b->lo[i] + off[i] < a->lo[i]) a->lo[i] = b->lo[i] + off[i];
        if (b->hi[i] + off[i] > a->hi[i]) a->hi[i] = b->hi[i] + off[i];
    }
}

/* rt_prims.zig:46-55: bit = x + 4y + 16z */
static OrBox leaf_box(uint64_t mask) {
    OrBox b = box_empty();
    const uint32_t zero[3] = {0, 0, 0};
    for (uint64_t m = mask; m; m &= m - 1) {
        const uint32_t bit = (uint32_t)__builtin_ctzll(m);
        OrBox v = {{bit & 3, (bit >> 2) & 3, bit >> 4}, {(bit & 3) + 1, ((bit >> 2) & 3) + 1, (bit >> 4) + 1}};
        box_merge(&b, &v, zero);
    }
    return b;
}

/* memo of boxes per (level, node), rt_prims.zig:61 */
typedef struct BoxMemo { uint64_t* key; OrBox* box; uint64_t cap, n; } BoxMemo;

static uint64_t memo_slot(const BoxMemo* m, uint64_t key) {
    uint64_t h = (key * 0x9E3779B97F4A7C15ull) >> 20;
    for (;; h++) {
        const uint64_t s = h & (m->cap - 1);
        if (m->key[s] == 0 || m->key[s] == key) return s;
    }
}

static int memo_grow(BoxMemo* m) {
    BoxMemo g = {calloc(m->cap * 2, 8), malloc(m->cap * 2 * sizeof(OrBox)), m->cap * 2, m->n};
    if (!g.key || !g.box) {
        free(g.key);
        free(g.box);
        return -1;
    }
    for (uint64_t s = 0; s < m->cap; s++)
        if (m->key[s]) {
            const uint64_t t = memo_slot(&g, m->key[s]);
            g.key[t] = m->key[s];
            g.box[t] = m->box[s];
        }
    free(m->key);
    free(m->box);
    *m = g;
    return 0;
}

typedef struct Extract {
    const uint32_t* nodes;
    const uint64_t* leaves;
    uint32_t k;
    BoxMemo memo;
    OrRtPrim* prims;
    float* aabbs;
    uint64_t capacity, count;
    int oom;
} Extract;

/* rt_prims.zig:73-90: box of a node covering a cube of edge 2^level */
static OrBox ex_box(Extract* e, uint32_t node, uint32_t level) {
    const uint64_t key = ((uint64_t)level << 32 | node) + 1; /* +1: 0 marks an empty memo slot */
    uint64_t s = memo_slot(&e->memo, key);
    if (e->memo.key[s]) return e->memo.box[s];
    OrBox b = box_empty();
    const uint32_t mask = e->nodes[node] & 0xFF, child = 1u << (level - 1);
    uint32_t j = 0;
    for (uint32_t idx = 0; idx < 8; idx++) {
        if (!(mask >> idx & 1)) continue;
        const uint32_t ref = e->nodes[node + 2 + j++];
        const OrBox cb = level == 3 ? leaf_box(e->leaves[ref]) : ex_box(e, ref, level - 1);
        const uint32_t off[3] = {(idx & 1) * child, (idx >> 1 & 1) * child, (idx >> 2) * child};
        box_merge(&b, &cb, off);
    }
    if (2 * (e->memo.n + 1) > e->memo.cap && memo_grow(&e->memo) != 0) {
        e->oom = 1;
        return b;
    }
    s = memo_slot(&e->memo, key);
    e->memo.key[s] = key;
    e->memo.box[s] = b;
    e->memo.n++;
    return b;
}

/* rt_prims.zig:92-124: DFS to level k, one primitive per non-empty subtree */
static void ex_visit(Extract* e, uint32_t node, uint32_t level, const uint32_t origin[3], uint32_t rank) {
    if (level == e->k) {
        const OrBox b = ex_box(e, node, level);
        if (b.lo[0] >= b.hi[0]) return; /* empty */
        if (e->count < e->capacity && e->prims) {
            OrRtPrim* p = &e->prims[e->count];
            p->node = node;
            p->attr_base = rank;
            p->cell = (uint64_t)(origin[0] >> e->k) | (uint64_t)(origin[1] >> e->k) << 21 |
                      (uint64_t)(origin[2] >> e->k) << 42; /* rt.zig:16-18 packCell */
            const float eps = 1.0f / 256.0f;
            for (int i = 0; i < 3; i++) {
                e->aabbs[e->count * 6 + i] = (float)(origin[i] + b.lo[i]) - eps;
                e->aabbs[e->count * 6 + 3 + i] = (float)(origin[i] + b.hi[i]) + eps;
            }
        }
        e->count++;
        return;
    }
    const uint32_t mask = e->nodes[node] & 0xFF, child = 1u << (level - 1);
    uint32_t r = rank, j = 0;
    for (uint32_t idx = 0; idx < 8; idx++) {
        if (!(mask >> idx & 1)) continue;
        const uint32_t ref = e->nodes[node + 2 + j++];
        const uint32_t o[3] = {origin[0] + (idx & 1) * child, origin[1] + (idx >> 1 & 1) * child,
                               origin[2] + (idx >> 2) * child};
        ex_visit(e, ref, level - 1, o, r);
        r += e->nodes[ref + 1];
    }
}

/* rt_prims.zig:127-137 extract */
int64_t or_rt_prims(const uint32_t* nodes, const uint64_t* leaves, uint32_t root, uint32_t log2_size,
                    uint32_t requested_log2, OrRtPrim* prims, float* aabbs, uint64_t capacity, uint32_t* out_k) {
    const uint32_t k = requested_log2 < 3 ? 3 : requested_log2 > log2_size ? log2_size : requested_log2;
    Extract e = {nodes, leaves, k, {calloc(1024, 8), malloc(1024 * sizeof(OrBox)), 1024, 0}, prims, aabbs, capacity,
                 0, 0};
    if (!e.memo.key || !e.memo.box) e.oom = 1;
    const uint32_t origin[3] = {0, 0, 0};
    if (!e.oom) ex_visit(&e, root, log2_size, origin, 0);
    free(e.memo.key);
    free(e.memo.box);
    if (out_k) *out_k = k;
    return e.oom ? -1 : (int64_t)e.count;
}
