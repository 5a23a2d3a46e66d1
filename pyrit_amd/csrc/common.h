// common.h - shared device helpers for the gfx950 build and trace kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pyr {

constexpr int kWave = 64;    // CDNA wavefront width
constexpr int kBlock = 256;  // 4 waves per workgroup for the tile kernels

__device__ __forceinline__ uint64_t spread3(uint32_t v) {
    uint64_t x = v & 0x1FFFFFu;
    x = (x | (x << 32)) & 0x1F00000000FFFFull;
    x = (x | (x << 16)) & 0x1F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

// Key of a voxel inside a chunk: brick Morton << 6 | bit in the 4^3 brick
// (bit = x + 4y + 16z), the key of gbuild.zig:98-101.
__device__ __forceinline__ uint64_t local_key(uint32_t x, uint32_t y, uint32_t z) {
    uint64_t brick = spread3(x >> 2) | (spread3(y >> 2) << 1) | (spread3(z >> 2) << 2);
    return (brick << 6) | ((x & 3) + 4 * (y & 3) + 16 * (z & 3));
}

__host__ __device__ __forceinline__ uint32_t local_bits(uint32_t log2_size) { return 3 * (log2_size - 2) + 6; }

// 64-bit finalizer used to place masks and node tuples in the hash tables; the
// output never depends on it (owner = smallest index, positions from scans).
__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return x;
}

__device__ __forceinline__ uint32_t sat_add(uint32_t a, uint32_t b) {
    uint32_t r = a + b;
    return r < a ? 0xFFFFFFFFu : r;
}

// Loads that bypass the non-coherent per-CU vector L1, for data another
// workgroup of the same launch published.
template <class T>
__device__ __forceinline__ T load_agent(const T* p) {
    return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void store_agent(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// value of the lowest active lane (readfirstlane returns int: convert through
// uint32_t so nothing is sign-extended)
__device__ __forceinline__ uint32_t first_lane32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}
__device__ __forceinline__ uint64_t first_lane64(uint64_t v) {
    return ((uint64_t)first_lane32((uint32_t)(v >> 32)) << 32) | (uint64_t)first_lane32((uint32_t)v);
}

// ---------------------------------------------------------------------------
// Lane moves with DPP (the data-parallel-primitive operand modifier of gfx9:
// the VALU reads another lane's register, no LDS instruction and no LDS
// latency; shuffles compile to ds_bpermute, which goes through the LDS).
// Every lane of the wave must be active. A lane whose source lane does not
// exist, or whose row the row mask leaves out, gets `old`.
// ---------------------------------------------------------------------------
constexpr int kDppRowShr1 = 0x111;     // lane i <- lane i - 1 inside its row of 16
constexpr int kDppRowShr2 = 0x112;
constexpr int kDppRowShr4 = 0x114;
constexpr int kDppRowShr8 = 0x118;
constexpr int kDppWaveShl1 = 0x130;    // lane i <- lane i + 1
constexpr int kDppWaveShr1 = 0x138;    // lane i <- lane i - 1
constexpr int kDppRowBcast15 = 0x142;  // with row mask 0xa: rows 1 and 3 <- lanes 15 and 47
constexpr int kDppRowBcast31 = 0x143;  // with row mask 0xc: rows 2 and 3 <- lane 31

template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ uint32_t dpp32(uint32_t v, uint32_t old = 0) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ uint64_t dpp64(uint64_t v, uint64_t old = 0) {
    const uint32_t lo = dpp32<CTRL, ROWS>((uint32_t)v, (uint32_t)old);
    const uint32_t hi = dpp32<CTRL, ROWS>((uint32_t)(v >> 32), (uint32_t)(old >> 32));
    return ((uint64_t)hi << 32) | lo;
}
// value of lane `lane` (a constant or wave-uniform index) in every lane
__device__ __forceinline__ uint32_t lane_read32(uint32_t v, int lane) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}
__device__ __forceinline__ uint64_t lane_read64(uint64_t v, int lane) {
    return ((uint64_t)lane_read32((uint32_t)(v >> 32), lane) << 32) | lane_read32((uint32_t)v, lane);
}

// the same for keys of either width
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ uint32_t dppx(uint32_t v, uint32_t old = 0) { return dpp32<CTRL, ROWS>(v, old); }
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ uint64_t dppx(uint64_t v, uint64_t old = 0) { return dpp64<CTRL, ROWS>(v, old); }
__device__ __forceinline__ uint32_t lane_read(uint32_t v, int lane) { return lane_read32(v, lane); }
__device__ __forceinline__ uint64_t lane_read(uint64_t v, int lane) { return lane_read64(v, lane); }

// inclusive prefix sum over the 64 lanes: four steps inside rows of 16,
// then the row totals
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t x) {
    x += dpp32<kDppRowShr1>(x);
    x += dpp32<kDppRowShr2>(x);
    x += dpp32<kDppRowShr4>(x);
    x += dpp32<kDppRowShr8>(x);
    x += dpp32<kDppRowBcast15, 0xa>(x);
    x += dpp32<kDppRowBcast31, 0xc>(x);
    return x;
}

__device__ __forceinline__ uint64_t wave_sum64(uint64_t x) {
    x += dpp64<kDppRowShr1>(x);
    x += dpp64<kDppRowShr2>(x);
    x += dpp64<kDppRowShr4>(x);
    x += dpp64<kDppRowShr8>(x);
    x += dpp64<kDppRowBcast15, 0xa>(x);
    x += dpp64<kDppRowBcast31, 0xc>(x);
    return lane_read64(x, kWave - 1);
}

// ---------------------------------------------------------------------------
// Decoupled look-back scan across the tiles of one launch.
//
// Status word: [63:62] flag (0 = not ready, 1 = aggregate, 2 = inclusive
// prefix), [61:40] epoch of the launch, [39:0] value. A word from another
// launch (other epoch) reads as "not ready", so one zero-initialised status
// array serves every launch of a build without clearing. Waits are bounded:
// a wait gives up after kWaitTicks of the constant 100 MHz clock, or as soon
// as any other wait of the build has given up (its error word is non-zero),
// records an error and goes on, so every wave terminates quickly even if a
// predecessor never publishes.
// ---------------------------------------------------------------------------
constexpr uint64_t kFlagAgg = 1ull << 62;
constexpr uint64_t kFlagPre = 2ull << 62;
constexpr uint64_t kValueMask = (1ull << 40) - 1;
constexpr uint64_t kWaitTicks = 100000000ull;  // 1 s at 100 MHz

struct Wait {
    uint64_t t0 = 0;
    uint32_t polls = 0;
};

// One more poll of a bounded wait: false = give up (error bit `bit` is set
// in *err unless another wait set an error first).
__device__ __forceinline__ bool wait_more(Wait& w, uint32_t* err, uint32_t bit) {
    if (w.polls == 0) w.t0 = wall_clock64();
    if ((++w.polls & 31) == 0) {
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return false;
        if (wall_clock64() - w.t0 > kWaitTicks) {
            atomicOr(err, bit);
            return false;
        }
    }
    __builtin_amdgcn_s_sleep(1);
    return true;
}

__device__ __forceinline__ uint64_t status_word(uint64_t flag, uint32_t epoch, uint64_t v) {
    return flag | ((uint64_t)(epoch & 0x3FFFFFu) << 40) | (v & kValueMask);
}

__device__ __forceinline__ uint32_t status_flag(uint64_t w, uint32_t epoch) {
    if (((w >> 40) & 0x3FFFFFu) != (epoch & 0x3FFFFFu)) return 0;
    return (uint32_t)(w >> 62);
}

// Exclusive prefix of this tile's `aggregate`. Called by all 64 lanes of one
// wave; returns the same value in every lane. Predecessors are polled 64 at a
// time.
__device__ __forceinline__ uint64_t lookback_wave(uint64_t* status, uint32_t tile, uint64_t aggregate,
                                                  uint32_t epoch, uint32_t* err) {
    const uint32_t lane = lane_id();
    if (tile == 0) {
        if (lane == 0)
            __hip_atomic_store(&status[0], status_word(kFlagPre, epoch, aggregate), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        return 0;
    }
    if (lane == 0)
        __hip_atomic_store(&status[tile], status_word(kFlagAgg, epoch, aggregate), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    uint64_t excl = 0;
    int64_t pos = (int64_t)tile - 1;
    Wait wt;
    while (true) {
        int64_t idx = pos - (int64_t)lane;
        uint64_t w = idx >= 0 ? __hip_atomic_load(&status[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                              : status_word(kFlagPre, epoch, 0);
        uint32_t f = status_flag(w, epoch);
        const uint64_t pre = __ballot(f == 2);
        // only the lanes up to the newest inclusive prefix are needed
        const uint64_t need = pre ? (pre & (~pre + 1)) * 2 - 1 : ~0ull;
        if (__ballot(f == 0) & need) {
            if (__ballot(!wait_more(wt, err, 1u)) != 0) break;
            continue;
        }
        uint64_t v = w & kValueMask;
        if (pre) {
            uint32_t first = (uint32_t)__ffsll((unsigned long long)pre) - 1;
            excl += wave_sum64(lane <= first ? v : 0);
            break;
        }
        excl += wave_sum64(v);
        pos -= kWave;
    }
    if (lane == 0)
        __hip_atomic_store(&status[tile], status_word(kFlagPre, epoch, excl + aggregate), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    return excl;
}

// Tile-local exclusive scan of per-item values laid out wave-striped: item i
// of lane l in wave w is element w*64*ITEMS + i*64 + l of the tile. Returns
// the tile total; s_wave must hold 4 words of LDS.
template <int ITEMS>
__device__ __forceinline__ uint32_t tile_scan(const uint32_t (&v)[ITEMS], uint32_t (&excl)[ITEMS],
                                              uint32_t* s_wave) {
    const uint32_t lane = lane_id(), wave = threadIdx.x / kWave;
    uint32_t carry = 0;
#pragma unroll
    for (int i = 0; i < ITEMS; i++) {
        uint32_t inc = wave_inclusive_scan(v[i]);
        excl[i] = carry + inc - v[i];
        carry += lane_read32(inc, kWave - 1);
    }
    if (lane == 0) s_wave[wave] = carry;
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kBlock / kWave; w++) {
        uint32_t t = s_wave[w];
        before += (uint32_t)w < wave ? t : 0;
        total += t;
    }
#pragma unroll
    for (int i = 0; i < ITEMS; i++) excl[i] += before;
    return total;
}

} // namespace pyr
