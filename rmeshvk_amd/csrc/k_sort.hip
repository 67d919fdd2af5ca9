// Stable LSD radix sort of (u64 or u32 key, u32 value) pairs, 8-bit digits
// over the low `sorting_bits` bits of the key — the semantics of rmesh-sort's
// DeviceRadixSort (crates/rmesh-sort/src/drs.rs:23-26,180-279): bits above
// sorting_bits are ignored and equal masked keys keep their input order.  The
// passes may also start above bit 0 (input already ordered by the bits below:
// LSD sorting over the remaining digits only, still stable).
//
// MI355X design (not the reference's upsweep/scan/downsweep port):
//  * fixed grid of G blocks, each owning one contiguous chunk of the input
//    whose length is derived on the device from the pair count — the launch
//    shape never depends on a host readback, so a whole step can be captured
//    in a hipGraph, and no block ever waits on another (no look-back spins);
//  * stable in-tile ranking with wave64 ballots: each wave matches the digit
//    bits of its 64 keys (one ballot per bit), ranks by popcount, and keeps a
//    private per-digit running count in LDS — no barriers inside the ranking
//    rounds;
//  * 8-bit digits (256 buckets) or 4-bit digits (16 buckets, twice the passes,
//    half the ballots per key and far smaller histograms), picked per renderer
//    by the autotuner;
//  * the tile (2048 keys) is re-ordered by digit in LDS before the scatter,
//    so the global writes are runs per digit instead of 8-byte scatters.
#include <mutex>

#include "rmesh_common.hpp"
#include "rmesh_kernels.hpp"

namespace rmesh {

constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / kWave;
constexpr int kRounds = 8;                          // 64-key rounds per wave per tile
constexpr int kWaveKeys = kRounds * kWave;          // 512
constexpr int kTileKeys = kWaveKeys * kSortWaves;   // 2048 (LDS 31 KB: 5 waves/SIMD)

__device__ __forceinline__ void chunk_of(uint32_t n, uint32_t G, uint32_t b, uint32_t& lo, uint32_t& hi) {
    uint32_t per = (n + G - 1) / G;
    per = (per + kTileKeys - 1) / kTileKeys * kTileKeys;
    uint64_t l = (uint64_t)per * b;
    lo = (uint32_t)min<uint64_t>(l, n);
    hi = (uint32_t)min<uint64_t>(l + per, n);
}

// DB: digit width (8: 256 buckets, 6 passes over 48 key bits; 4: 16 buckets,
// twice the passes with a quarter of the ballots per key each); chosen by the
// autotuner.
template <int DB, typename K>
__global__ __launch_bounds__(kSortThreads) void k_rs_upsweep(const K* __restrict__ keys, const uint32_t* __restrict__ n_ptr,
                                                             uint32_t cap, uint32_t shift, uint32_t* __restrict__ hist) {
    constexpr uint32_t RADIX = 1u << DB, DMASK = RADIX - 1u;
    __shared__ uint32_t h[kSortWaves][RADIX];
    uint32_t G = gridDim.x, b = blockIdx.x;
    uint32_t n = min(*n_ptr, cap);
    uint32_t lo, hi;
    chunk_of(n, G, b, lo, hi);
    for (int i = threadIdx.x; i < kSortWaves * (int)RADIX; i += kSortThreads) (&h[0][0])[i] = 0;
    __syncthreads();
    uint32_t w = threadIdx.x / kWave;
    // eight keys in flight per thread before their histogram adds
    constexpr int U = 8;
    uint32_t i0 = lo + threadIdx.x;
    for (; i0 + (U - 1) * kSortThreads < hi; i0 += U * kSortThreads) {
        K kk[U];
#pragma unroll
        for (int u = 0; u < U; ++u) kk[u] = keys[i0 + u * kSortThreads];
#pragma unroll
        for (int u = 0; u < U; ++u) atomicAdd(&h[w][(uint32_t)(kk[u] >> shift) & DMASK], 1u);
    }
    for (; i0 < hi; i0 += kSortThreads) atomicAdd(&h[w][(uint32_t)(keys[i0] >> shift) & DMASK], 1u);
    __syncthreads();
    uint32_t d = threadIdx.x;
    if (d < RADIX) {
        uint32_t s = 0;
#pragma unroll
        for (int k = 0; k < kSortWaves; ++k) s += h[k][d];
        hist[(uint64_t)d * G + b] = s;
    }
}

// Exclusive scan of hist[d][0..G) for one digit d per workgroup (256
// workgroups in parallel); the digit total goes to totals[d].
__global__ __launch_bounds__(256) void k_rs_scan(uint32_t* __restrict__ hist, uint32_t G, uint32_t* __restrict__ totals) {
    __shared__ uint32_t red[256 / kWave];
    uint32_t* h = hist + (uint64_t)blockIdx.x * G;
    uint32_t per = (G + blockDim.x - 1) / blockDim.x;
    uint32_t lo = threadIdx.x * per;
    uint32_t hi = min(lo + per, G);
    uint32_t s = 0;
    for (uint32_t i = lo; i < hi; ++i) s += h[i];
    uint32_t tot;
    uint32_t ex = block_excl_scan(s, red, tot);
    for (uint32_t i = lo; i < hi; ++i) {
        uint32_t x = h[i];
        h[i] = ex;
        ex += x;
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = tot;
}

template <int DB, typename K>
__global__ __launch_bounds__(kSortThreads) void k_rs_downsweep(const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                               K* __restrict__ kout, uint32_t* __restrict__ vout,
                                                               const uint32_t* __restrict__ n_ptr, uint32_t cap, uint32_t shift,
                                                               const uint32_t* __restrict__ hist, const uint32_t* __restrict__ totals) {
    constexpr uint32_t RADIX = 1u << DB, DMASK = RADIX - 1u;
    static_assert(RADIX <= (uint32_t)kSortThreads, "one thread per digit");
    __shared__ K s_keys[kTileKeys];
    __shared__ uint32_t s_vals[kTileKeys];
    __shared__ uint32_t s_wcnt[kSortWaves][RADIX];
    __shared__ uint32_t s_loc[RADIX];
    __shared__ uint32_t s_glob[RADIX];
    __shared__ uint32_t s_run[RADIX];
    __shared__ uint32_t s_red[kSortWaves];

    uint32_t G = gridDim.x, b = blockIdx.x;
    uint32_t n = min(*n_ptr, cap);
    uint32_t lo, hi;
    chunk_of(n, G, b, lo, hi);
    const uint32_t tid = threadIdx.x, w = tid / kWave, lane = __lane_id();
    {
        // global start of digit `tid` = exclusive scan of the digit totals
        uint32_t all;
        uint32_t dstart = block_excl_scan(tid < RADIX ? totals[tid] : 0u, s_red, all);
        if (tid < RADIX) s_run[tid] = dstart + hist[(uint64_t)tid * G + b];
    }
    const uint64_t lt = lanemask_lt();

    for (uint32_t base = lo; base < hi; base += kTileKeys) {
        for (int i = tid; i < kSortWaves * (int)RADIX; i += kSortThreads) (&s_wcnt[0][0])[i] = 0;
        __syncthreads();
        K key[kRounds];
        uint32_t val[kRounds];
        uint32_t dr[kRounds];  // digit | rank << 9 ; digit 256 marks an empty slot
        // all of the tile's loads first (one memory latency per tile, not one
        // per round: the rounds' wave barriers would otherwise serialise them)
#pragma unroll
        for (int j = 0; j < kRounds; ++j) {
            uint32_t i = base + w * kWaveKeys + j * kWave + lane;
            bool valid = i < hi;
            key[j] = valid ? kin[i] : K(0);
            val[j] = valid ? vin[i] : 0u;
        }
        // Phase 1: per-wave stable ranking, 64 keys per round.
#pragma unroll
        for (int j = 0; j < kRounds; ++j) {
            uint32_t i = base + w * kWaveKeys + j * kWave + lane;
            bool valid = i < hi;
            uint32_t d = (uint32_t)(key[j] >> shift) & DMASK;
            uint64_t peers = __ballot(valid);
#pragma unroll
            for (int bit = 0; bit < DB; ++bit) {
                uint64_t bb = __ballot((d >> bit) & 1u);
                peers &= ((d >> bit) & 1u) ? bb : ~bb;
            }
            uint32_t r = 0;
            if (valid) r = s_wcnt[w][d] + (uint32_t)__popcll(peers & lt);
            __builtin_amdgcn_wave_barrier();
            bool leader = valid && ((peers & lt) == 0ull);
            if (leader) s_wcnt[w][d] += (uint32_t)__popcll(peers);
            __builtin_amdgcn_wave_barrier();
            dr[j] = valid ? (d | (r << 9)) : 256u;
        }
        __syncthreads();
        // Phase 2: per digit — wave prefixes, tile counts, local starts, global bases.
        {
            const uint32_t d = tid;
            uint32_t tot = 0;
            if (d < RADIX) {
#pragma unroll
                for (int k = 0; k < kSortWaves; ++k) {
                    const uint32_t c = s_wcnt[k][d];
                    s_wcnt[k][d] = tot;
                    tot += c;
                }
            }
            uint32_t all;
            uint32_t loc = block_excl_scan(tot, s_red, all);
            if (d < RADIX) {
                s_loc[d] = loc;
                s_glob[d] = s_run[d];
                s_run[d] += tot;
            }
        }
        __syncthreads();
        // Phase 3: stage in digit order.
#pragma unroll
        for (int j = 0; j < kRounds; ++j) {
            uint32_t d = dr[j] & 511u;
            if (d < 256u) {
                uint32_t lp = s_loc[d] + s_wcnt[w][d] + (dr[j] >> 9);
                s_keys[lp] = key[j];
                s_vals[lp] = val[j];
            }
        }
        __syncthreads();
        // Phase 4: coalesced scatter, runs per digit.
        uint32_t tile_n = min((uint32_t)kTileKeys, hi - base);
        for (uint32_t i = tid; i < tile_n; i += kSortThreads) {
            const K k = s_keys[i];
            uint32_t d = (uint32_t)(k >> shift) & DMASK;
            uint32_t gp = s_glob[d] + (i - s_loc[d]);
            kout[gp] = k;
            vout[gp] = s_vals[i];
        }
        __syncthreads();
    }
}

// Blocks of the downsweep that are resident at once on the current device
// (CUs x blocks per CU at its LDS and register use, for the key width K),
// cached per device.  A grid of at most that many blocks runs every block of
// a pass in one round: no second, partly filled round of blocks at the end of
// each pass.
template <typename K>
static uint32_t resident_sort_blocks() {
    static std::mutex mu;
    static uint32_t cached[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 2048u;
    std::lock_guard<std::mutex> lk(mu);
    if (!cached[dev]) {
        int cus = 0, b8 = 0, b4 = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&b8, reinterpret_cast<const void*>(k_rs_downsweep<8, K>),
                                                         kSortThreads, 0) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&b4, reinterpret_cast<const void*>(k_rs_downsweep<4, K>),
                                                         kSortThreads, 0) != hipSuccess ||
            cus <= 0 || b8 <= 0 || b4 <= 0) {
            cached[dev] = 2048u;
        } else {
            cached[dev] = (uint32_t)cus * (uint32_t)(b8 < b4 ? b8 : b4);
        }
    }
    return cached[dev];
}

template <typename K>
static uint32_t sort_grid(uint32_t cap) {
    uint32_t g = (cap + kTileKeys - 1) / kTileKeys;
    const uint32_t r = resident_sort_blocks<K>();
    if (g > r) g = r;
    if (g > 2048) g = 2048;
    if (g < 1) g = 1;
    return g;
}

// Sorts the pairs in (ka, va) on key bits [first_bit, end_bit); returns true
// when the result ended in (kb, vb).
template <int DB, typename K>
static bool radix_passes(K* ka, uint32_t* va, K* kb, uint32_t* vb, const uint32_t* n_ptr, uint32_t cap,
                         uint32_t first_bit, uint32_t end_bit, uint32_t* hist, hipStream_t s) {
    constexpr uint32_t RADIX = 1u << DB;
    uint32_t G = sort_grid<K>(cap);
    bool in_b = false;
    for (uint32_t shift = first_bit; shift < end_bit; shift += DB) {
        const K* kin = in_b ? kb : ka;
        const uint32_t* vin = in_b ? vb : va;
        K* kout = in_b ? ka : kb;
        uint32_t* vout = in_b ? va : vb;
        hipLaunchKernelGGL((k_rs_upsweep<DB, K>), dim3(G), dim3(kSortThreads), 0, s, kin, n_ptr, cap, shift, hist);
        hipLaunchKernelGGL(k_rs_scan, dim3(RADIX), dim3(256), 0, s, hist, G, hist + (uint64_t)RADIX * G);
        hipLaunchKernelGGL((k_rs_downsweep<DB, K>), dim3(G), dim3(kSortThreads), 0, s, kin, vin, kout, vout, n_ptr, cap,
                           shift, hist, hist + (uint64_t)RADIX * G);
        in_b = !in_b;
    }
    return in_b;
}

bool launch_radix_sort(uint64_t* ka, uint32_t* va, uint64_t* kb, uint32_t* vb, const uint32_t* n_ptr, uint32_t cap,
                       uint32_t sorting_bits, uint32_t* hist, hipStream_t s, uint32_t digit_bits, uint32_t first_bit) {
    if (digit_bits == 4) return radix_passes<4>(ka, va, kb, vb, n_ptr, cap, first_bit, sorting_bits, hist, s);
    return radix_passes<8>(ka, va, kb, vb, n_ptr, cap, first_bit, sorting_bits, hist, s);
}

bool launch_radix_sort32(uint32_t* ka, uint32_t* va, uint32_t* kb, uint32_t* vb, const uint32_t* n_ptr, uint32_t cap,
                         uint32_t* hist, hipStream_t s, uint32_t digit_bits, uint32_t end_bit) {
    if (digit_bits == 4) return radix_passes<4>(ka, va, kb, vb, n_ptr, cap, 0u, end_bit, hist, s);
    return radix_passes<8>(ka, va, kb, vb, n_ptr, cap, 0u, end_bit, hist, s);
}

size_t radix_hist_words(uint32_t cap) { return 256ull * sort_grid<uint64_t>(cap) + 256; }
size_t radix_hist_words32(uint32_t cap) { return 256ull * sort_grid<uint32_t>(cap) + 256; }

}  // namespace rmesh
