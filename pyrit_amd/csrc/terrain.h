// terrain.h - seeded voxel-terrain workload generator, identical on host and
// device (integer arithmetic only, so CPU and GPU emit the same voxels bit for
// bit). It plays the role of the reference's example generator behind
// PyrWorldInfo.generate (demo/terrain.zig:115-170, 436-503: hashed value noise,
// fBm, skin-only columns) but is this repository's own design: fixed-point
// quintic value noise, octave decorrelation by hashed offsets, integer
// spherical caves and hashed or depth-banded materials.
//
// A chunk's voxels are emitted column by column (x fastest, then z) and
// bottom-up inside a column, in chunk-local coordinates.
#pragma once
#include <stdint.h>

#ifdef __HIPCC__
#define PYRGEN_HD __host__ __device__ __forceinline__
#else
#define PYRGEN_HD static inline
#endif

// Material ids (attribute values; 0 means empty): grass 1, water 2, stone 3,
// dirt 8 (the block ids of demo/terrain.zig:27-35).
#define PYRGEN_GRASS 1u
#define PYRGEN_WATER 2u
#define PYRGEN_STONE 3u
#define PYRGEN_DIRT 8u

typedef struct PyrTerrainCave { int32_t x, y, z, r; } PyrTerrainCave;

// One chunk to generate. Noise and caves use world coordinates
// (origin + local), so neighbouring chunks of one world line up.
typedef struct PyrTerrainChunk {
    int32_t origin[3];
    uint32_t log2_size;       // edge 2^log2_size voxels
    uint32_t seed;
    int32_t mean;             // mean surface height (world y)
    int32_t amplitude;        // surface amplitude
    uint32_t octaves;         // fBm octaves, >= 1
    uint32_t wavelength_log2; // base wavelength 2^k voxels, k in [1, 16]
    uint32_t material_mode;   // 0: hashed categorical, 1: depth bands
    uint32_t skin;            // 0: solid below the surface, 1: visible skin only
    int32_t water_level;      // < 0: none; else water voxel at y = water_level over lower ground
    uint32_t cave_first;      // caves[cave_first .. cave_first + cave_count)
    uint32_t cave_count;
    uint32_t grid_log2;       // cave bin edge 2^grid_log2 (local xz)
    uint32_t grid_dim;        // bins per axis; 0 = test every cave of the chunk
    uint32_t grid_first;      // bins at bin_start[grid_first .. + grid_dim^2 + 1)
} PyrTerrainChunk;

PYRGEN_HD uint32_t pyrgen_mix(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

PYRGEN_HD uint32_t pyrgen_hash3(int32_t a, int32_t b, int32_t c, uint32_t seed) {
    uint32_t h = pyrgen_mix((uint32_t)a * 0x9E3779B1u + seed);
    h = pyrgen_mix(h ^ ((uint32_t)b * 0x85EBCA77u));
    return pyrgen_mix(h ^ ((uint32_t)c * 0xC2B2AE3Du));
}

// Quintic fade 6t^5 - 15t^4 + 10t^3 in Q16 (t in [0, 65536)).
PYRGEN_HD int64_t pyrgen_fade(int64_t t) {
    int64_t t2 = (t * t) >> 16;
    int64_t t3 = (t2 * t) >> 16;
    int64_t a = 6 * t - 15 * 65536;
    int64_t b = ((t * a) >> 16) + 10 * 65536;
    return (t3 * b) >> 16;
}

// Value noise in Q16 [-65536, 65536) with lattice spacing 2^k.
PYRGEN_HD int64_t pyrgen_noise(int32_t x, int32_t z, uint32_t k, uint32_t seed) {
    int32_t ix = x >> k, iz = z >> k;
    int64_t tx = (int64_t)(x & ((1 << k) - 1)) << (16 - k);
    int64_t tz = (int64_t)(z & ((1 << k) - 1)) << (16 - k);
    int64_t sx = pyrgen_fade(tx), sz = pyrgen_fade(tz);
    int64_t a = pyrgen_hash3(ix, iz, 0, seed) >> 16;
    int64_t b = pyrgen_hash3(ix + 1, iz, 0, seed) >> 16;
    int64_t c = pyrgen_hash3(ix, iz + 1, 0, seed) >> 16;
    int64_t d = pyrgen_hash3(ix + 1, iz + 1, 0, seed) >> 16;
    int64_t top = a + (((b - a) * sx) >> 16);
    int64_t bot = c + (((d - c) * sx) >> 16);
    int64_t r = top + (((bot - top) * sz) >> 16);
    return 2 * r - 65536;
}

// fBm in Q16, roughly [-65536, 65536]: octave o has wavelength 2^(k0-o) (at
// least 2) and weight 2^-o; each octave is shifted by a hashed offset so the
// lattices of different octaves do not line up.
PYRGEN_HD int64_t pyrgen_fbm(int32_t x, int32_t z, uint32_t octaves, uint32_t k0, uint32_t seed) {
    int64_t sum = 0, norm = 0;
    for (uint32_t o = 0; o < octaves; o++) {
        int64_t amp = (int64_t)65536 >> (o < 16 ? o : 16);
        uint32_t k = k0 > o + 1 ? k0 - o : 1;
        uint32_t os = seed + o * 0x9E3779B9u;
        int32_t ox = (int32_t)(pyrgen_hash3((int32_t)o, 1, 0, os) & 0xFFFFu);
        int32_t oz = (int32_t)(pyrgen_hash3((int32_t)o, 2, 0, os) & 0xFFFFu);
        sum += (pyrgen_noise(x + ox, z + oz, k, os) * amp) >> 16;
        norm += amp;
    }
    return norm ? sum * 65536 / norm : 0;
}

// Surface height at world column (x, z): solid cells have y < height.
PYRGEN_HD int32_t pyrgen_height(const PyrTerrainChunk* c, int32_t x, int32_t z) {
    int64_t f = pyrgen_fbm(x, z, c->octaves, c->wavelength_log2, c->seed);
    return c->mean + (int32_t)(((int64_t)c->amplitude * f) >> 16);
}

PYRGEN_HD uint32_t pyrgen_isqrt(uint32_t v) {
    uint32_t r = 0;
    for (uint32_t bit = 1u << 30; bit; bit >>= 2) {
        if (v >= r + bit) {
            v -= r + bit;
            r = (r >> 1) + bit;
        } else {
            r >>= 1;
        }
    }
    return r;
}

// Carved vertical interval [lo, hi] of one cave in world column (x, z); false
// if the column misses it. A cell is carved when dx^2+dy^2+dz^2 < r^2.
PYRGEN_HD bool pyrgen_cave_span(const PyrTerrainCave& cv, int32_t x, int32_t z, int32_t* lo, int32_t* hi) {
    int64_t dx = x - cv.x, dz = z - cv.z;
    int64_t rem = (int64_t)cv.r * cv.r - dx * dx - dz * dz;
    if (rem < 1) return false;
    int32_t m = (int32_t)pyrgen_isqrt((uint32_t)(rem - 1));
    *lo = cv.y - m;
    *hi = cv.y + m;
    return true;
}

PYRGEN_HD uint32_t pyrgen_material(const PyrTerrainChunk* c, int32_t x, int32_t y, int32_t z, int32_t h) {
    if (c->material_mode == 0) {
        uint32_t r = pyrgen_hash3(x, y, z, c->seed ^ 0xA5A5A5A5u) % 100u;
        return r < 60 ? PYRGEN_STONE : r < 85 ? PYRGEN_DIRT : r < 95 ? PYRGEN_GRASS : PYRGEN_WATER;
    }
    int32_t depth = h - 1 - y;
    return depth < 1 ? PYRGEN_GRASS : depth < 4 ? PYRGEN_DIRT : PYRGEN_STONE;
}

// Walks one local column (lx, lz) of chunk c and calls emit(lx, ly, lz, attr)
// bottom-up. Returns the number of voxels emitted.
template <class Emit>
PYRGEN_HD uint32_t pyrgen_column(const PyrTerrainChunk* c, const PyrTerrainCave* caves, const uint32_t* bin_start,
                                 const uint32_t* bin_caves, uint32_t lx, uint32_t lz, Emit emit) {
    const int32_t n = 1 << c->log2_size;
    const int32_t wx = c->origin[0] + (int32_t)lx, wz = c->origin[2] + (int32_t)lz;
    const int32_t y0 = c->origin[1];
    const int32_t h = pyrgen_height(c, wx, wz);
    int32_t lo = 0;
    if (c->skin) {
        int32_t m = pyrgen_height(c, wx - 1, wz);
        int32_t t = pyrgen_height(c, wx + 1, wz);
        m = t < m ? t : m;
        t = pyrgen_height(c, wx, wz - 1);
        m = t < m ? t : m;
        t = pyrgen_height(c, wx, wz + 1);
        m = t < m ? t : m;
        lo = m - 2;
        if (lo > h - 1) lo = h - 1;
    }
    // world y range of solid cells inside this chunk
    int32_t ya = lo - y0, yb = h - y0; // local [ya, yb)
    if (ya < 0) ya = 0;
    if (yb > n) yb = n;
    uint32_t first = 0, count = c->cave_count;
    const uint32_t* list = nullptr;
    if (c->grid_dim) {
        uint32_t b = c->grid_first + (lz >> c->grid_log2) * c->grid_dim + (lx >> c->grid_log2);
        first = bin_start[b];
        count = bin_start[b + 1] - first;
        list = bin_caves;
    }
    // carved spans of the caves that reach this column (the first 8 are kept;
    // with more, every cell re-tests all caves)
    int32_t sa[8], sb[8];
    uint32_t ns = 0;
    bool overflow = false;
    for (uint32_t k = 0; k < count; k++) {
        const PyrTerrainCave& cv = caves[list ? list[first + k] : c->cave_first + k];
        int32_t a, b;
        if (pyrgen_cave_span(cv, wx, wz, &a, &b)) {
            if (ns < 8) {
                sa[ns] = a;
                sb[ns] = b;
                ns++;
            } else {
                overflow = true;
            }
        }
    }
    uint32_t emitted = 0;
    for (int32_t ly = ya; ly < yb; ly++) {
        const int32_t wy = y0 + ly;
        bool carved = false;
        for (uint32_t k = 0; k < ns; k++) carved |= wy >= sa[k] && wy <= sb[k];
        for (uint32_t k = 0; overflow && k < count && !carved; k++) {
            const PyrTerrainCave& cv = caves[list ? list[first + k] : c->cave_first + k];
            int32_t a, b;
            if (pyrgen_cave_span(cv, wx, wz, &a, &b) && wy >= a && wy <= b) carved = true;
        }
        if (carved) continue;
        emit(lx, (uint32_t)ly, lz, pyrgen_material(c, wx, wy, wz, h));
        emitted++;
    }
    if (c->water_level >= 0 && h <= c->water_level) {
        int32_t ly = c->water_level - y0;
        if (ly >= 0 && ly < n) {
            emit(lx, (uint32_t)ly, lz, PYRGEN_WATER);
            emitted++;
        }
    }
    return emitted;
}

// ---------------------------------------------------------------------------
// World chunks: the same heightfield sampled per level of detail. A chunk of
// level `lod` has cells of 2^lod base voxels; a cell is solid when its lowest
// base voxel is below the surface, and only the skin is emitted: in every
// column the cells from one below the lowest neighbouring surface up to the
// column's own surface. Materials by depth band in cells.
// ---------------------------------------------------------------------------
typedef struct PyrTerrainWorld {
    uint32_t seed;
    int32_t mean;             // mean surface height (base voxels)
    int32_t amplitude;
    uint32_t octaves;
    uint32_t wavelength_log2;
    uint32_t reserved;
} PyrTerrainWorld;

// surface of cell column (X, Z) of level lod, in cells: cells y < result are solid
PYRGEN_HD int64_t pyrgen_world_cell_height(const PyrTerrainWorld* w, int64_t X, int64_t Z, uint32_t lod) {
    const int32_t bx = (int32_t)(X << lod), bz = (int32_t)(Z << lod);
    const int64_t f = pyrgen_fbm(bx, bz, w->octaves, w->wavelength_log2, w->seed);
    const int64_t h = (int64_t)w->mean + (((int64_t)w->amplitude * f) >> 16);
    return (h + ((int64_t)1 << lod) - 1) >> lod;  // ceil(h / 2^lod)
}

// Emits the skin of local column (lx, lz) of chunk (kx, ky, kz, lod) with
// edge 2^chunk_log2 through emit(x, y, z, attribute); returns the count.
template <class Emit>
PYRGEN_HD uint32_t pyrgen_world_column(const PyrTerrainWorld* w, int32_t kx, int32_t ky, int32_t kz, uint32_t lod,
                                       uint32_t chunk_log2, uint32_t lx, uint32_t lz, Emit emit) {
    const int64_t n = (int64_t)1 << chunk_log2;
    const int64_t X = (int64_t)kx * n + lx, Z = (int64_t)kz * n + lz, Y0 = (int64_t)ky * n;
    const int64_t h = pyrgen_world_cell_height(w, X, Z, lod);
    int64_t m = pyrgen_world_cell_height(w, X - 1, Z, lod);
    int64_t t = pyrgen_world_cell_height(w, X + 1, Z, lod);
    m = t < m ? t : m;
    t = pyrgen_world_cell_height(w, X, Z - 1, lod);
    m = t < m ? t : m;
    t = pyrgen_world_cell_height(w, X, Z + 1, lod);
    m = t < m ? t : m;
    int64_t lo = m - 1;
    if (lo > h - 1) lo = h - 1;
    int64_t ya = lo - Y0, yb = h - Y0;
    if (ya < 0) ya = 0;
    if (yb > n) yb = n;
    uint32_t emitted = 0;
    for (int64_t ly = ya; ly < yb; ly++) {
        const int64_t depth = h - 1 - (Y0 + ly);
        emit(lx, (uint32_t)ly, lz, depth < 1 ? PYRGEN_GRASS : depth < 4 ? PYRGEN_DIRT : PYRGEN_STONE);
        emitted++;
    }
    return emitted;
}
