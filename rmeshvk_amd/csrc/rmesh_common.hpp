// Shared device helpers for the rmesh HIP kernels (gfx950, wave64).
//
// Float conventions are fixed so the binning kernels reproduce the FP32
// arithmetic of the reference kernels bit for bit (the library is compiled
// with -ffp-contract=off; division and sqrt are IEEE-correct):
//   dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z
//   cross(a,b) = (a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x)
//   normalize(v) = v / sqrt(dot(v,v))
//   min/max = IEEE minNum/maxNum (fminf/fmaxf)
//   float -> u32/i32 conversions truncate and saturate, NaN -> 0
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rmesh {

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Per-view camera/tile parameters, read by every stage of a batched step.
// vp is column-major (crates/rmesh-python/src/lib.rs:117-125); c2w holds
// the three rotation columns and intr = (fx, fy, cx, cy) (lib.rs:127-136).
// ---------------------------------------------------------------------------
struct alignas(16) ViewParams {
    float vp[16];
    float c2w[9];
    float intr[4];
    float cam[3];
    float min_t;
    float Wf, Hf;
    uint32_t width, height;
    uint32_t tiles_x, tiles_y;
    uint32_t tile_base;   // first global tile id of this view (ranges, compositing grids)
    uint32_t tile_size;
    // Sort-key tile ids.  One view: the tile itself and the reference's
    // sentinel key ~0.  Several views in one sort: each view owns tiles_x *
    // tiles_y + 1 key tiles, the last one its sentinel tile, so a view's
    // sentinel pairs sort between its own pairs and the next view's, as the
    // reference's per-view sort leaves them after its pairs.
    uint32_t key_base;       // first key tile of this view
    uint32_t sentinel_tile;  // key tile of this view's sentinels (0xFFFFFFFF: the reference's ~0 key)
    uint32_t scene;          // the caller's index of this view's scene
    // Per-view buffer offsets (several views, several scenes of uneven size and
    // resolution in one step): this view's colours (and colour gradients) are
    // colors[color_off + t] for the global tet ids t of its scene, and its
    // pixels start at pix_base in the image buffers (rgba, aux, targets,
    // image gradients, transmittance finals).
    int64_t color_off;
    uint64_t pix_base;
};

// One scene of a step: tets [tet_base, tet_base + M), views [view_base,
// view_base + V); its binning items are item_base + local tet * V + local
// view (tet-major: the reference's pair order within every view).  Scenes
// are listed in tet order; item_base and tet_base never decrease.
struct SceneRange {
    uint64_t item_base;
    uint32_t tet_base, M;
    uint32_t view_base, V;
};

// Last scene s with key(s) <= x (binary search over the n scenes).
template <typename F>
__device__ __forceinline__ uint32_t scene_search(uint32_t n, F key, uint64_t x) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (key(mid) <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}
__device__ __forceinline__ uint32_t scene_of_item(const SceneRange* __restrict__ sc, uint32_t n, uint64_t item) {
    return scene_search(n, [&](uint32_t s) { return sc[s].item_base; }, item);
}
__device__ __forceinline__ uint32_t scene_of_tet(const SceneRange* __restrict__ sc, uint32_t n, uint32_t t) {
    return scene_search(n, [&](uint32_t s) { return (uint64_t)sc[s].tet_base; }, t);
}

struct V3 {
    float x, y, z;
};

__device__ __forceinline__ V3 mk3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 sub3(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 add3(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 scl3(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 neg3(V3 a) { return V3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ float dot3(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ V3 cross3(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ V3 normalize3(V3 v) {
    float l = sqrtf(dot3(v, v));
    return V3{v.x / l, v.y / l, v.z / l};
}

// safe math — crates/rmesh-util/src/wgsl/math.wgsl:3-65
constexpr float kTiny = 1.0754944e-20f;
constexpr float kMinVal = -1e+20f;
constexpr float kMaxVal = 1e+20f;
constexpr float kLogMaxVal = 46.0517f;

__device__ __forceinline__ float safe_clip(float v, float lo, float hi) { return fmaxf(fminf(v, hi), lo); }
__device__ __forceinline__ V3 safe_clip3(V3 v, float lo, float hi) {
    return V3{safe_clip(v.x, lo, hi), safe_clip(v.y, lo, hi), safe_clip(v.z, lo, hi)};
}
__device__ __forceinline__ float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }
__device__ __forceinline__ float safe_div(float a, float b) {
    if (fabsf(b) < kTiny) return safe_clip(a / kTiny, kMinVal, kMaxVal);
    return safe_clip(a / b, kMinVal, kMaxVal);
}
__device__ __forceinline__ float safe_exp(float v) { return expf(clampf(v, -88.0f, kLogMaxVal)); }
__device__ __forceinline__ float phi_f(float x) {
    if (fabsf(x) < 0.02f) return 1.0f + x * (-0.5f + x * ((float)(1.0 / 6.0) + x * ((float)(-1.0 / 24.0))));
    return safe_div(1.0f - expf(-x), x);
}
__device__ __forceinline__ float dphi_dx_f(float x) {
    if (fabsf(x) < 0.02f) return -0.5f + x * ((float)(1.0 / 3.0) + x * ((float)(-1.0 / 8.0) + x * ((float)(1.0 / 30.0))));
    float ex = expf(-x);
    return safe_div(ex * (x + 1.0f) - 1.0f, x * x);
}
__device__ __forceinline__ float softplus_f(float x) {
    if (x * 10.0f > 20.0f) return x;
    return 0.1f * logf(1.0f + expf(10.0f * x));
}
__device__ __forceinline__ float dsoftplus_f(float x) {
    if (x * 10.0f > 20.0f) return 1.0f;
    float e = expf(10.0f * x);
    return safe_div(e, 1.0f + e);
}

// Fast variants for the per-pair compositing math, where the result is
// compared within tolerance (images 1e-4, gradients 1e-3): hardware
// reciprocal (v_rcp_f32, 1 ulp) and base-2 exponential (v_exp_f32) instead of
// the IEEE division / accurate expf sequences.  Coverage decisions (scanline
// spans) never use these.
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fast_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float fast_exp(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }
__device__ __forceinline__ float fast_safe_div(float a, float b) {
    if (fabsf(b) < kTiny) return safe_clip(a * (1.0f / kTiny), kMinVal, kMaxVal);
    return safe_clip(fast_div(a, b), kMinVal, kMaxVal);
}
// Clamps for the per-pair math as one v_med3_f32 (no NaN canonicalisation).
__device__ __forceinline__ float med_clip(float v, float lo, float hi) { return __builtin_amdgcn_fmed3f(v, lo, hi); }
__device__ __forceinline__ V3 med_clip3(V3 v, float lo, float hi) {
    return V3{med_clip(v.x, lo, hi), med_clip(v.y, lo, hi), med_clip(v.z, lo, hi)};
}
__device__ __forceinline__ float relu(float v) { return __builtin_amdgcn_fmed3f(v, 0.0f, __builtin_huge_valf()); }
__device__ __forceinline__ float fast_safe_exp(float v) { return fast_exp(clampf(v, -88.0f, kLogMaxVal)); }
__device__ __forceinline__ float fast_phi(float x, float ex) {  // ex = exp(-x)
    if (fabsf(x) < 0.02f) return 1.0f + x * (-0.5f + x * ((float)(1.0 / 6.0) + x * ((float)(-1.0 / 24.0))));
    return fast_safe_div(1.0f - ex, x);
}
__device__ __forceinline__ float fast_dphi(float x, float ex) {
    if (fabsf(x) < 0.02f) return -0.5f + x * ((float)(1.0 / 3.0) + x * ((float)(-1.0 / 8.0) + x * ((float)(1.0 / 30.0))));
    return fast_safe_div(ex * (x + 1.0f) - 1.0f, x * x);
}

// WGSL-style saturating conversions (explicit: no reliance on UB rules).
__device__ __forceinline__ uint32_t sat_u32(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}
__device__ __forceinline__ int32_t sat_i32(float x) {
    if (x != x) return 0;
    if (x <= -2147483648.0f) return INT32_MIN;
    if (x >= 2147483648.0f) return INT32_MAX;
    return (int32_t)x;
}
__device__ __forceinline__ uint32_t uclamp(uint32_t v, uint32_t lo, uint32_t hi) {
    uint32_t m = v > lo ? v : lo;
    return m < hi ? m : hi;
}

// clip = vp * vec4(p, 1), columns summed in order (project_compute.wgsl:77-81)
__device__ __forceinline__ void mul_vp(const float* vp, V3 p, float out[4]) {
#pragma unroll
    for (int k = 0; k < 4; ++k) out[k] = ((vp[k] * p.x + vp[4 + k] * p.y) + vp[8 + k] * p.z) + vp[12 + k];
}

// Tet face table (a, b, c, opposite) — crates/rmesh-util/src/wgsl/intersect.wgsl:4-9
__device__ __forceinline__ void face_idx(int fi, int& a, int& b, int& c, int& o) {
    // FACES = {0,2,1,3}, {1,2,3,0}, {0,3,2,1}, {3,0,1,2}
    a = fi == 0 ? 0 : (fi == 1 ? 1 : (fi == 2 ? 0 : 3));
    b = fi == 0 ? 2 : (fi == 1 ? 2 : (fi == 2 ? 3 : 0));
    c = fi == 0 ? 1 : (fi == 1 ? 3 : (fi == 2 ? 2 : 1));
    o = fi == 0 ? 3 : (fi == 1 ? 0 : (fi == 2 ? 1 : 2));
}

// Pixel ray direction (rasterize_compute.wgsl:344-349)
__device__ __forceinline__ V3 pixel_ray(const ViewParams& v, uint32_t px, uint32_t py) {
    V3 dc = mk3((((float)px + 0.5f) - v.intr[2]) / v.intr[0], (((float)py + 0.5f) - v.intr[3]) / v.intr[1], 1.0f);
    dc = normalize3(dc);
    V3 d = mk3((v.c2w[0] * dc.x + v.c2w[3] * dc.y) + v.c2w[6] * dc.z,
               (v.c2w[1] * dc.x + v.c2w[4] * dc.y) + v.c2w[7] * dc.z,
               (v.c2w[2] * dc.x + v.c2w[5] * dc.y) + v.c2w[8] * dc.z);
    return normalize3(d);
}

// ---------------------------------------------------------------------------
// wave64 / block primitives
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() {
    uint32_t l = __lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Inclusive scan of a u32 across the wave.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    uint32_t l = __lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o, 64);
        if (l >= (uint32_t)o) v += t;
    }
    return v;
}

// Exclusive block scan of per-thread u32 values; returns the block total.
// `smem` needs blockDim/64 entries.  Contains two barriers.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* smem, uint32_t& total) {
    uint32_t inc = wave_incl_scan(v);
    uint32_t wid = threadIdx.x / kWave;
    uint32_t nw = blockDim.x / kWave;
    if (__lane_id() == kWave - 1) smem[wid] = inc;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
    for (uint32_t w = 0; w < nw; ++w) {
        uint32_t s = smem[w];
        if (w < wid) pre += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return pre + inc - v;
}

}  // namespace rmesh
