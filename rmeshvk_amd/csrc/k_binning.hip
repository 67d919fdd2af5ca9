// Projection, tile counting, pair generation and per-tile ranges.
//
// Reference kernels restated here (crates/rmesh-render, crates/rmesh-tile):
//   project_compute.wgsl:88-262        -> k_project  (cull + conservative tile count)
//   prepare_dispatch.wgsl + rts.wgsl   -> block sums in k_project + k_spine_scan
//   tile_fill_compute.wgsl             -> sentinel padding inside k_pairs
//   tile_gen_scan_compute.wgsl:61-203  -> k_pairs_sorted (k_pairs)
//   tile_ranges_compute.wgsl:10-45     -> k_ranges
//
// MI355X design: one thread per (tet, view) item of every view of the step,
// tet-major, so the views of a tet share its record (64 B, 4 x float4) in one
// wave.  Pair offsets come from per-block sums, one spine scan and a wave64
// block scan redone in the writing kernel — no offsets array and no atomics
// on the ordering path, so the pair order is deterministic.  The items are
// sorted by depth key before pair generation (k_project writes the keys,
// k_pairs_sorted emits in that order), so the pair sort needs only the tile
// bits; k_pairs keeps the reference's item-order pairs with 64-bit keys.
#include "rmesh_common.hpp"
#include "rmesh_kernels.hpp"

namespace rmesh {

// Tet record: r0 = (v0, density), r1 = (v1, g.x), r2 = (v2, g.y), r3 = (v3, g.z)
__global__ __launch_bounds__(256) void k_prep_records(const float* __restrict__ verts, const uint32_t* __restrict__ idx,
                                                      const float* __restrict__ dens, const float* __restrict__ grads,
                                                      uint32_t M, float4* __restrict__ rec) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M) return;
    uint4 q = reinterpret_cast<const uint4*>(idx)[t];
    uint32_t vi[4] = {q.x, q.y, q.z, q.w};
    float extra[4] = {dens[t], grads[3ull * t], grads[3ull * t + 1], grads[3ull * t + 2]};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float* p = verts + 3ull * vi[i];
        rec[4ull * t + i] = make_float4(p[0], p[1], p[2], extra[i]);
    }
}

__global__ __launch_bounds__(256) void k_copy_colors(const float* __restrict__ base, uint32_t M, float4* __restrict__ col) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M) return;
    col[t] = make_float4(base[3ull * t], base[3ull * t + 1], base[3ull * t + 2], 0.0f);
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t value, uint64_t n) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = value;
}

// Float atomic adds of known count for the atomic-rate ceiling: every lane
// adds 1 to its own 64-byte line, the lines of a wave's lanes scattered over
// `lines` lines (the adjoint flush's pattern: consecutive lanes on terms of
// different tets), no return value.
__global__ __launch_bounds__(256) void k_atomic_probe(float* __restrict__ p, uint32_t lines, uint32_t per_thread) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t line = (i * 7919u) % lines;
    for (uint32_t k = 0; k < per_thread; ++k) {
        unsafeAtomicAdd(p + 16ull * line, 1.0f);
        line += 977u;
        if (line >= lines) line -= lines;
    }
}
void launch_atomic_probe(float* p, uint32_t lines, uint32_t blocks, uint32_t per_thread, hipStream_t s) {
    hipLaunchKernelGGL(k_atomic_probe, dim3(blocks), dim3(256), 0, s, p, lines, per_thread);
}

// Plain streaming copy used to calibrate HBM byte counters (FETCH_SIZE /
// WRITE_SIZE) against a kernel of known traffic.
__global__ __launch_bounds__(256) void k_copy_probe(const float4* __restrict__ src, float4* __restrict__ dst, uint64_t n) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

__device__ __forceinline__ void load_rec(const float4* __restrict__ rec, uint32_t t, V3 v[4], float& dens, V3& g) {
    float4 r0 = rec[4ull * t], r1 = rec[4ull * t + 1], r2 = rec[4ull * t + 2], r3 = rec[4ull * t + 3];
    v[0] = mk3(r0.x, r0.y, r0.z);
    v[1] = mk3(r1.x, r1.y, r1.z);
    v[2] = mk3(r2.x, r2.y, r2.z);
    v[3] = mk3(r3.x, r3.y, r3.z);
    dens = r0.w;
    g = mk3(r1.w, r2.w, r3.w);
}

// Scanline tile-row extent at tile granularity (project_compute.wgsl:219-258,
// tile_gen_scan_compute.wgsl:151-182): a row's x range is the min / max over
// the edge crossings of its top line y = ty T and its bottom line
// y = (ty + 1) T, and over the vertices inside the band.  RowScan walks the
// rows top to bottom and carries each bottom line's crossings over as the
// next row's top line (the same expression, so the same values), which halves
// the crossing divisions of every row after the first.  fminf / fmaxf give
// the same extent in any order.
struct RowScan {
    float top_lo, top_hi;

    // x range of the edge crossings of line y (1e10 / -1e10: none).  Branch
    // free: every edge's crossing is computed and kept only where the edge
    // spans the line (across a wave of 64 (tet, view) items nearly every
    // edge spans it somewhere, so a branch would rarely skip the division).
    __device__ __forceinline__ static void line(const float proj[4][2], float y, float& lo, float& hi) {
        constexpr int EI[6] = {0, 0, 0, 1, 1, 2};
        constexpr int EJ[6] = {1, 2, 3, 2, 3, 3};
        lo = 1e10f;
        hi = -1e10f;
#pragma unroll
        for (int e = 0; e < 6; ++e) {
            const float yi = proj[EI[e]][1], yj = proj[EJ[e]][1];
            const float xi = proj[EI[e]][0], xj = proj[EJ[e]][0];
            const bool cross = (yi <= y && yj > y) || (yj <= y && yi > y);
            const float tt = (y - yi) / (yj - yi);
            const float x = xi + tt * (xj - xi);
            lo = cross ? fminf(lo, x) : lo;
            hi = cross ? fmaxf(hi, x) : hi;
        }
    }
    __device__ __forceinline__ void start(const float proj[4][2], uint32_t ty0, float T) {
        line(proj, (float)ty0 * T, top_lo, top_hi);
    }
    // extent of row ty (rows visited in increasing order from start's ty0)
    __device__ __forceinline__ bool row(const float proj[4][2], uint32_t ty, float T, float& xl, float& xr) {
        float bot_lo, bot_hi;
        line(proj, (float)(ty + 1u) * T, bot_lo, bot_hi);
        xl = fminf(top_lo, bot_lo);
        xr = fmaxf(top_hi, bot_hi);
        top_lo = bot_lo;
        top_hi = bot_hi;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float vy = proj[k][1];
            if (vy >= (float)ty * T && vy <= (float)(ty + 1u) * T) {
                xl = fminf(xl, proj[k][0]);
                xr = fmaxf(xr, proj[k][0]);
            }
        }
        return xl <= xr;
    }
};

// x / T.  For a power-of-two T the product with the exact reciprocal is the
// same correctly rounded value, without the division sequence.
template <bool POW2>
__device__ __forceinline__ float div_tile(float x, float T, float invT) {
    if constexpr (POW2) return x * invT;
    else return x / T;
}

struct TileBox {
    uint32_t x0, x1, y0, y1;
};

// AABB in tiles (project_compute.wgsl:202-216).  The tile counts per axis,
// ceil(W / T) and ceil(H / T) there, are the view's tiles_x / tiles_y:
// (W + T - 1) / T in integers, the same numbers for any image below 2^24
// pixels a side (a non-integer W / T is at least 1 / T from an integer, far
// more than its rounding error).
template <bool POW2>
__device__ __forceinline__ bool tile_box(const float proj[4][2], const ViewParams& vw, float T, float invT, TileBox& b) {
    const float W = vw.Wf, H = vw.Hf;
    float pix_min_x = fmaxf(fminf(fminf(proj[0][0], proj[1][0]), fminf(proj[2][0], proj[3][0])), 0.0f);
    float pix_max_x = fminf(fmaxf(fmaxf(proj[0][0], proj[1][0]), fmaxf(proj[2][0], proj[3][0])), W - 1.0f);
    float pix_min_y = fmaxf(fminf(fminf(proj[0][1], proj[1][1]), fminf(proj[2][1], proj[3][1])), 0.0f);
    float pix_max_y = fminf(fmaxf(fmaxf(proj[0][1], proj[1][1]), fmaxf(proj[2][1], proj[3][1])), H - 1.0f);
    if (pix_min_x > pix_max_x || pix_min_y > pix_max_y) return false;
    b.x0 = sat_u32(div_tile<POW2>(pix_min_x, T, invT));
    b.x1 = min(sat_u32(div_tile<POW2>(pix_max_x, T, invT)), vw.tiles_x - 1u);
    b.y0 = sat_u32(div_tile<POW2>(pix_min_y, T, invT));
    b.y1 = min(sat_u32(div_tile<POW2>(pix_max_y, T, invT)), vw.tiles_y - 1u);
    return true;
}

template <bool POW2>
__device__ __forceinline__ void row_tiles(float xl, float xr, float T, float invT, const TileBox& b, uint32_t& tx0,
                                          uint32_t& tx1) {
    tx0 = uclamp(sat_u32(fmaxf(floorf(div_tile<POW2>(xl, T, invT)), 0.0f)), b.x0, b.x1);
    tx1 = uclamp(sat_u32(fmaxf(floorf(div_tile<POW2>(xr, T, invT)), 0.0f)), b.x0, b.x1);
}

// K1 for one (tet, view): tile count; `visible` is the reference's vis_idx
// flag (project_compute.wgsl:111-262).
template <bool POW2>
__device__ __forceinline__ uint32_t project_count(const V3 v[4], float dens, const ViewParams& vw, bool& visible) {
    float p[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float clip[4];
        mul_vp(vw.vp, v[i], clip);
        float inv_w = 1.0f / (clip[3] + 1e-6f);
        p[i][0] = clip[0] * inv_w;
        p[i][1] = clip[1] * inv_w;
        p[i][2] = clip[2] * inv_w;
        p[i][3] = clip[3];
    }
    bool any_behind = (p[0][3] <= 0.0f) || (p[1][3] <= 0.0f) || (p[2][3] <= 0.0f) || (p[3][3] <= 0.0f);
    float W = vw.Wf, H = vw.Hf;
    bool vis = true;
    if (!any_behind) {
        float min_x = fminf(fminf(p[0][0], p[1][0]), fminf(p[2][0], p[3][0]));
        float max_x = fmaxf(fmaxf(p[0][0], p[1][0]), fmaxf(p[2][0], p[3][0]));
        float min_y = fminf(fminf(p[0][1], p[1][1]), fminf(p[2][1], p[3][1]));
        float max_y = fmaxf(fmaxf(p[0][1], p[1][1]), fmaxf(p[2][1], p[3][1]));
        float min_z = fminf(fminf(p[0][2], p[1][2]), fminf(p[2][2], p[3][2]));
        if (max_x < -1.0f || min_x > 1.0f || max_y < -1.0f || min_y > 1.0f || min_z > 1.0f) vis = false;
        float ext_x = (max_x - min_x) * W;
        float ext_y = (max_y - min_y) * H;
        if (ext_x * ext_y < 1.0f) vis = false;
    }
    if (!vis) {
        visible = false;
        return 0;
    }
    {
        V3 e01 = sub3(v[1], v[0]), e02 = sub3(v[2], v[0]), e03 = sub3(v[3], v[0]);
        V3 e12 = sub3(v[2], v[1]), e13 = sub3(v[3], v[1]), e23 = sub3(v[3], v[2]);
        float m = fmaxf(fmaxf(dot3(e01, e01), fmaxf(dot3(e02, e02), dot3(e03, e03))),
                        fmaxf(dot3(e12, e12), fmaxf(dot3(e13, e13), dot3(e23, e23))));
        float max_od = dens * sqrtf(m);
        if (max_od < 0.01f) {
            visible = false;
            return 0;
        }
    }
    visible = true;
    if (any_behind) return 0;
    float proj[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        proj[i][0] = ((p[i][0] + 1.0f) * 0.5f) * W;
        proj[i][1] = ((1.0f - p[i][1]) * 0.5f) * H;
    }
    const float T = (float)vw.tile_size, invT = 1.0f / T;
    TileBox b;
    if (!tile_box<POW2>(proj, vw, T, invT, b)) return 0;
    uint32_t count = 0;
    RowScan rs;
    rs.start(proj, b.y0, T);
    for (uint32_t ty = b.y0; ty <= b.y1; ++ty) {
        float xl, xr;
        if (rs.row(proj, ty, T, xl, xr)) {
            uint32_t tx0, tx1;
            row_tiles<POW2>(xl, xr, T, invT, b, tx0, tx1);
            count += tx1 - tx0 + 1u;
        }
    }
    return count;
}

// (tet, view) of a binning item: scene by binary search, then tet-major
// inside the scene.  One scene: its range is read at a uniform address
// (scalar loads), not through the lane's search result.
__device__ __forceinline__ void item_tet_view(const SceneRange* __restrict__ sc, uint32_t n_scenes, uint64_t item,
                                              uint32_t& t, uint32_t& vv) {
    if (n_scenes == 1) {
        const uint32_t V = sc[0].V;
        const uint64_t local = item - sc[0].item_base;
        t = sc[0].tet_base + (uint32_t)(local / V);
        vv = sc[0].view_base + (uint32_t)(local % V);
        return;
    }
    const SceneRange& r = sc[scene_of_item(sc, n_scenes, item)];
    const uint64_t local = item - r.item_base;
    t = r.tet_base + (uint32_t)(local / r.V);
    vv = r.view_base + (uint32_t)(local % r.V);
}

// Depth sort values: item << 5 | count, the count saturating at 31 (then read
// from counts[item]).
constexpr uint32_t kCountBits = 5, kCountEsc = (1u << kCountBits) - 1u;
__device__ __forceinline__ uint32_t sorted_count(uint32_t sv, const uint32_t* __restrict__ counts) {
    const uint32_t c = sv & kCountEsc;
    return c == kCountEsc ? counts[sv >> kCountBits] : c;
}

// Sort key of a tet's depth in a view: ~bits(|c - cam|^2 - r^2) over its
// circumsphere (c, r^2) (tile_gen_scan_compute.wgsl:135-143): ascending key =
// descending power distance, the reference's back-to-front order.
__device__ __forceinline__ uint32_t depth_key(const float4& cd, const ViewParams& vw) {
    const V3 diff = sub3(mk3(cd.x, cd.y, cd.z), mk3(vw.cam[0], vw.cam[1], vw.cam[2]));
    const float depth = dot3(diff, diff) - cd.w;
    return ~__float_as_uint(depth);
}

// K1, batched: one thread per (tet, view) item of every scene (item =
// scene's item base + local tet * V + local view: the pair order of the
// reference within every view is tet-major).  counts[item]; per-block pair
// totals for the scan; visible tets of view 0 (the reference's
// indirect_args.instance_count, project_compute.wgsl:181).  With dkey set,
// also the item's depth sort key (0xFFFFFFFF for items without pairs) and
// its sort value item << 5 | min(count, 31) (31: look the count up; items
// < 2^27), the input of the depth sort of items, and the item count for it.
template <bool POW2>
__global__ __launch_bounds__(256) void k_project(const float4* __restrict__ rec, const float4* __restrict__ circ,
                                                 const SceneRange* __restrict__ scenes, uint32_t n_scenes,
                                                 uint64_t n_items, const ViewParams* __restrict__ views,
                                                 uint32_t* __restrict__ counts, uint32_t* __restrict__ block_sums,
                                                 uint32_t* __restrict__ block_vis, uint32_t* __restrict__ dkey,
                                                 uint32_t* __restrict__ ditem, uint32_t* __restrict__ n_items_out) {
    __shared__ uint32_t red[256 / kWave];
    __shared__ uint32_t s_vis;
    if (threadIdx.x == 0) s_vis = 0;
    const uint64_t item = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = item < n_items;
    uint32_t c = 0;
    bool vis0 = false;
    if (live) {
        uint32_t t, vv;
        item_tet_view(scenes, n_scenes, item, t, vv);
        V3 v[4];
        float dens;
        V3 g;
        load_rec(rec, t, v, dens, g);
        bool vis = false;
        const ViewParams& vw = views[vv];
        c = project_count<POW2>(v, dens, vw, vis);
        counts[item] = c;
        vis0 = vis && vv == 0;
        if (dkey) {
            dkey[item] = c ? depth_key(circ[t], vw) : 0xFFFFFFFFu;
            ditem[item] = ((uint32_t)item << kCountBits) | min(c, kCountEsc);
        }
    }
    if (dkey && blockIdx.x == 0 && threadIdx.x == 0) *n_items_out = (uint32_t)n_items;
    const uint64_t bal = __ballot(vis0);
    __syncthreads();
    if (__lane_id() == 0 && bal) atomicAdd(&s_vis, (uint32_t)__popcll(bal));
    uint32_t total;
    (void)block_excl_scan(c, red, total);  // ends with a barrier: s_vis is complete
    if (threadIdx.x == 0) {
        // per-block counts, summed by the spine scan: one global atomic per
        // block on a single address serialised in L2 and held the kernel up
        block_sums[blockIdx.x] = total;
        block_vis[blockIdx.x] = s_vis;
    }
}

// Exclusive scan of block sums in place (single workgroup of 1024 threads,
// 16384 sums per round, each thread 16 consecutive ones: four 16-byte loads);
// writes the grand total (number of pairs P) to *total.  The running total is
// kept in 64 bits: past 2^32 - 16 pairs the offsets would wrap, so *total
// saturates at 0xFFFFFFFF (the host refuses that count) and the overflow flag
// is raised (a captured step then leaves the parameters untouched).  With
// aux set, also the plain sum of aux[0, n) into *aux_total.
constexpr uint32_t kSpinePer = 16;
__global__ __launch_bounds__(1024) void k_spine_scan(uint32_t* __restrict__ sums, uint32_t n, uint32_t* __restrict__ total,
                                                     uint32_t* __restrict__ overflow, const uint32_t* __restrict__ aux,
                                                     uint32_t* __restrict__ aux_total) {
    __shared__ uint32_t red[1024 / kWave];
    __shared__ unsigned long long s_round;  // the round's total in 64 bits
    __shared__ uint32_t s_aux;
    if (threadIdx.x == 0) s_aux = 0u;
    uint64_t carry = 0;
    for (uint32_t base = 0; base < n; base += kSpinePer * 1024u) {
        const uint32_t i0 = base + kSpinePer * threadIdx.x;
        uint32_t x[kSpinePer];
        if (i0 + kSpinePer <= n) {
            const uint4* p = reinterpret_cast<const uint4*>(sums + i0);  // i0 is a multiple of 16: aligned
#pragma unroll
            for (uint32_t q = 0; q < kSpinePer / 4; ++q) {
                const uint4 v = p[q];
                x[4 * q] = v.x;
                x[4 * q + 1] = v.y;
                x[4 * q + 2] = v.z;
                x[4 * q + 3] = v.w;
            }
        } else {
#pragma unroll
            for (uint32_t k = 0; k < kSpinePer; ++k) x[k] = i0 + k < n ? sums[i0 + k] : 0u;
        }
        uint64_t loc = 0;
#pragma unroll
        for (uint32_t k = 0; k < kSpinePer; ++k) loc += x[k];
        if (threadIdx.x == 0) s_round = 0ull;
        __syncthreads();  // also publishes s_aux = 0 before the first adds
        if (aux) {
            uint32_t a = 0;
            if (i0 + kSpinePer <= n) {
                const uint4* p = reinterpret_cast<const uint4*>(aux + i0);
#pragma unroll
                for (uint32_t q = 0; q < kSpinePer / 4; ++q) {
                    const uint4 v = p[q];
                    a += (v.x + v.y) + (v.z + v.w);
                }
            } else {
                for (uint32_t k = 0; k < kSpinePer; ++k) a += i0 + k < n ? aux[i0 + k] : 0u;
            }
            a = wave_sum_u32(a);
            if (__lane_id() == 0 && a) atomicAdd(&s_aux, a);
        }
        atomicAdd(&s_round, (unsigned long long)loc);
        uint32_t tot;
        uint32_t ex = block_excl_scan((uint32_t)loc, red, tot) + (uint32_t)carry;  // ends with a barrier
#pragma unroll
        for (uint32_t k = 0; k < kSpinePer; ++k) {
            if (i0 + k < n) sums[i0 + k] = ex;
            ex += x[k];
        }
        carry += s_round;
        __syncthreads();  // every thread has read s_round before the next round clears it
    }
    if (threadIdx.x == 0) {
        const bool wrap = carry > 0xFFFFFFF0ull;
        *total = wrap ? 0xFFFFFFFFu : (uint32_t)carry;
        if (wrap) *overflow = 1u;
        if (aux) *aux_total = s_aux;  // the rounds' barriers ordered every wave's add before this read
    }
}

// Two-level form of the spine scan for many block sums: each workgroup scans
// kSpineTile sums in place (local offsets) and leaves its exact total (64
// bits) and its aux sum in part / aux_part; one workgroup turns the totals
// into offsets (and the grand totals, saturated as above); a third pass adds
// each workgroup's offset.  Three short launches instead of one workgroup
// walking every sum in rounds.
constexpr uint32_t kSpineTile = 4096;
__global__ __launch_bounds__(1024) void k_spine_local(uint32_t* __restrict__ sums, uint32_t n,
                                                      unsigned long long* __restrict__ part,
                                                      const uint32_t* __restrict__ aux,
                                                      uint32_t* __restrict__ aux_part) {
    __shared__ uint32_t red[1024 / kWave];
    __shared__ unsigned long long s_tot;
    __shared__ uint32_t s_aux;
    if (threadIdx.x == 0) {
        s_tot = 0ull;
        s_aux = 0u;
    }
    const uint32_t i0 = blockIdx.x * kSpineTile + 4u * threadIdx.x;
    uint32_t x[4], a = 0u;
#pragma unroll
    for (int k = 0; k < 4; ++k) x[k] = i0 + k < n ? sums[i0 + k] : 0u;
    if (aux) {
#pragma unroll
        for (int k = 0; k < 4; ++k) a += i0 + k < n ? aux[i0 + k] : 0u;
    }
    __syncthreads();  // s_tot, s_aux cleared
    atomicAdd(&s_tot, ((unsigned long long)x[0] + x[1]) + ((unsigned long long)x[2] + x[3]));
    if (aux) {
        a = wave_sum_u32(a);
        if (__lane_id() == 0 && a) atomicAdd(&s_aux, a);
    }
    uint32_t tot;
    uint32_t ex = block_excl_scan((x[0] + x[1]) + (x[2] + x[3]), red, tot);  // ends with a barrier
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (i0 + k < n) sums[i0 + k] = ex;
        ex += x[k];
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = s_tot;
        if (aux) aux_part[blockIdx.x] = s_aux;
    }
}

__global__ __launch_bounds__(64) void k_spine_top(unsigned long long* __restrict__ part, uint32_t nb,
                                                  uint32_t* __restrict__ total, uint32_t* __restrict__ overflow,
                                                  const uint32_t* __restrict__ aux_part,
                                                  uint32_t* __restrict__ aux_total) {
    if (threadIdx.x != 0) return;
    unsigned long long carry = 0ull;
    uint32_t a = 0u;
    for (uint32_t b = 0; b < nb; ++b) {
        const unsigned long long p = part[b];
        part[b] = carry;
        carry += p;
        if (aux_part) a += aux_part[b];
    }
    const bool wrap = carry > 0xFFFFFFF0ull;
    *total = wrap ? 0xFFFFFFFFu : (uint32_t)carry;
    if (wrap) *overflow = 1u;
    if (aux_part) *aux_total = a;
}

__global__ __launch_bounds__(1024) void k_spine_add(uint32_t* __restrict__ sums, uint32_t n,
                                                    const unsigned long long* __restrict__ part) {
    const uint32_t add = (uint32_t)part[blockIdx.x];
    const uint32_t i0 = blockIdx.x * kSpineTile + 4u * threadIdx.x;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (i0 + k < n) sums[i0 + k] += add;
}

// K7 (+ K6 sentinel padding) for one item: its max_write pairs at
// [start, start + max_write).  64-bit keys: (key tile << 32) | ~bits(depth);
// 32-bit keys: the key tile alone (pairs already in depth order).
template <bool POW2, typename KT>
__device__ __forceinline__ void emit_item_pairs(uint64_t item, uint32_t start, uint32_t max_write,
                                                const float4* __restrict__ rec, const float4* __restrict__ circ,
                                                const SceneRange* __restrict__ scenes, uint32_t n_scenes,
                                                const ViewParams* __restrict__ views, KT* __restrict__ keys,
                                                uint32_t* __restrict__ vals, uint32_t cap,
                                                uint32_t* __restrict__ overflow) {
    constexpr bool FULL = sizeof(KT) == 8;
    uint32_t t, vv;
    item_tet_view(scenes, n_scenes, item, t, vv);
    V3 v[4];
    float dens;
    V3 g;
    load_rec(rec, t, v, dens, g);
    const ViewParams& vw = views[vv];
    const float W = vw.Wf, H = vw.Hf, T = (float)vw.tile_size, invT = 1.0f / T;
    float proj[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float c[4];
        mul_vp(vw.vp, v[i], c);
        float nx = c[0] / (c[3] + 1e-6f);
        float ny = c[1] / (c[3] + 1e-6f);
        proj[i][0] = ((nx + 1.0f) * 0.5f) * W;
        proj[i][1] = ((1.0f - ny) * 0.5f) * H;
    }
    uint32_t local = 0;
    TileBox b;
    if (tile_box<POW2>(proj, vw, T, invT, b)) {
        uint32_t inv_depth = 0u;
        if constexpr (FULL) inv_depth = depth_key(circ[t], vw);
        RowScan rs;
        rs.start(proj, b.y0, T);
        for (uint32_t ty = b.y0; ty <= b.y1 && local < max_write; ++ty) {
            float xl, xr;
            if (!rs.row(proj, ty, T, xl, xr)) continue;
            uint32_t tx0, tx1;
            row_tiles<POW2>(xl, xr, T, invT, b, tx0, tx1);
            for (uint32_t tx = tx0; tx <= tx1 && local < max_write; ++tx) {
                uint32_t gt = vw.key_base + ty * vw.tiles_x + tx;
                uint32_t wi = start + local;
                if (wi < cap) {
                    if constexpr (FULL) keys[wi] = ((uint64_t)gt << 32) | inv_depth;
                    else keys[wi] = gt;
                    vals[wi] = t;
                } else {
                    *overflow = 1u;
                }
                local++;
            }
        }
    }
    // Slots counted by K1 but not produced by the division-based
    // re-projection keep the tile_fill sentinel (tile_gen_scan_compute.wgsl:202).
    KT sentinel;
    if constexpr (FULL) sentinel = ((uint64_t)vw.sentinel_tile << 32) | 0xFFFFFFFFull;
    else sentinel = vw.sentinel_tile;
    for (; local < max_write; ++local) {
        uint32_t wi = start + local;
        if (wi < cap) {
            keys[wi] = sentinel;
            vals[wi] = 0u;
        } else {
            *overflow = 1u;
        }
    }
}

// Pairs in item order (the reference's pre-sort order): thread per item,
// offsets from the item-order block sums and a block scan redone here.
template <bool POW2>
__global__ __launch_bounds__(256) void k_pairs(const float4* __restrict__ rec, const float4* __restrict__ circ,
                                               const SceneRange* __restrict__ scenes, uint32_t n_scenes, uint64_t n_items,
                                               const ViewParams* __restrict__ views,
                                               const uint32_t* __restrict__ counts, const uint32_t* __restrict__ block_off,
                                               uint64_t* __restrict__ keys, uint32_t* __restrict__ vals, uint32_t cap,
                                               uint32_t* __restrict__ overflow) {
    __shared__ uint32_t red[256 / kWave];
    const uint64_t item = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = item < n_items;
    const uint32_t max_write = live ? counts[item] : 0u;
    uint32_t tot;
    const uint32_t start = block_excl_scan(max_write, red, tot) + block_off[blockIdx.x];
    if (max_write == 0) return;
    emit_item_pairs<POW2, uint64_t>(item, start, max_write, rec, circ, scenes, n_scenes, views, keys, vals, cap,
                                    overflow);
}

// Depth-ordered pair generation.  The items, sorted stably by depth key
// (sorted values sval), emit their pairs in that order, so a stable sort of
// the pairs on the tile bits alone gives, inside every tile, the order by
// (depth key, item) — exactly the order of the stable (tile, depth) sort of
// the pairs in item order.  Two passes over the sorted items: per-block pair
// totals, then (after the spine scan) the pairs at their offsets.  The pair
// writes are sequential; the tet records are read in depth order (gathers
// from a table that stays in the last-level cache), which measured cheaper
// than writing the pairs by item to scattered offsets.
__global__ __launch_bounds__(256) void k_sorted_sums(const uint32_t* __restrict__ sval, uint32_t n_items,
                                                     const uint32_t* __restrict__ counts,
                                                     uint32_t* __restrict__ block_sums) {
    __shared__ uint32_t red[256 / kWave];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t c = i < n_items ? sorted_count(sval[i], counts) : 0u;
    uint32_t tot;
    (void)block_excl_scan(c, red, tot);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

template <bool POW2>
__global__ __launch_bounds__(256) void k_pairs_sorted(const float4* __restrict__ rec,
                                                      const SceneRange* __restrict__ scenes, uint32_t n_scenes,
                                                      const uint32_t* __restrict__ sval, uint32_t n_items,
                                                      const ViewParams* __restrict__ views,
                                                      const uint32_t* __restrict__ counts,
                                                      const uint32_t* __restrict__ block_off,
                                                      uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                      uint32_t cap, uint32_t* __restrict__ overflow) {
    __shared__ uint32_t red[256 / kWave];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t sv = i < n_items ? sval[i] : 0u;
    const uint32_t max_write = i < n_items ? sorted_count(sv, counts) : 0u;
    uint32_t tot;
    const uint32_t start = block_excl_scan(max_write, red, tot) + block_off[blockIdx.x];
    if (max_write == 0) return;
    emit_item_pairs<POW2, uint32_t>(sv >> kCountBits, start, max_write, rec, nullptr, scenes, n_scenes, views, keys,
                                    vals, cap, overflow);
}

// The 64-bit keys of depth-ordered pairs (key tile, sorted values): the depth
// key recomputed from the tet's circumsphere and its view's camera; the
// sentinels' full key.  For inspection (the 64-bit key lists of the
// reference's sort), not on the step path.
__global__ __launch_bounds__(256) void k_full_keys(const uint32_t* __restrict__ ktile, const uint32_t* __restrict__ vals,
                                                   const uint32_t* __restrict__ n_ptr, uint32_t cap,
                                                   const float4* __restrict__ circ, const ViewParams* __restrict__ views,
                                                   const uint32_t* __restrict__ ktile_gtile, uint32_t n_ktiles,
                                                   const uint32_t* __restrict__ tile_view, uint64_t* __restrict__ out) {
    const uint32_t n = min(*n_ptr, cap);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t kt = ktile[i];
    const uint32_t gt = kt < n_ktiles ? ktile_gtile[kt] : 0xFFFFFFFFu;
    const uint32_t d = gt == 0xFFFFFFFFu ? 0xFFFFFFFFu : depth_key(circ[vals[i]], views[tile_view[gt]]);
    out[i] = ((uint64_t)kt << 32) | d;
}

// K11: per-tile [start, end) in the sorted pair list, including the
// reference's behaviour for trailing sentinels (tile_ranges_compute.wgsl:10-45):
// a tile whose pairs are followed by sentinels keeps end = 0.  Key tiles map
// to global tiles through ktile_gtile (0xFFFFFFFF: a view's sentinel tile;
// the key tile 0xFFFFFFFF is the reference's own sentinel), so each view's
// ranges are those of the reference's per-view sort.
// The key tile: the high word of a 64-bit key, a 32-bit key itself.
template <typename KT>
__device__ __forceinline__ uint32_t range_tile(KT key, const uint32_t* __restrict__ ktile_gtile, uint32_t n_ktiles) {
    const uint32_t kt = sizeof(KT) == 8 ? (uint32_t)((uint64_t)key >> 32) : (uint32_t)key;
    return kt < n_ktiles ? ktile_gtile[kt] : 0xFFFFFFFFu;
}
template <typename KT>
__global__ void k_ranges(const KT* __restrict__ keys, const uint32_t* __restrict__ n_ptr, uint32_t cap,
                         uint32_t* __restrict__ ranges, const uint32_t* __restrict__ ktile_gtile, uint32_t n_ktiles) {
    uint32_t n = min(*n_ptr, cap);
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint32_t tile = range_tile(keys[i], ktile_gtile, n_ktiles);
        if (tile == 0xFFFFFFFFu) continue;
        if (i == 0) {
            ranges[2ull * tile] = i;
        } else {
            uint32_t prev = range_tile(keys[i - 1], ktile_gtile, n_ktiles);
            if (tile != prev) {
                ranges[2ull * tile] = i;
                if (prev != 0xFFFFFFFFu) ranges[2ull * prev + 1] = i;
            }
        }
        if (i == n - 1) ranges[2ull * tile + 1] = i + 1;
    }
}

// k_ranges over 32-bit keys, four consecutive keys per thread (one 16-byte
// load; the key before them comes from one more load), same per-key rule.
__global__ void k_ranges32x4(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ n_ptr, uint32_t cap,
                             uint32_t* __restrict__ ranges, const uint32_t* __restrict__ ktile_gtile,
                             uint32_t n_ktiles) {
    const uint32_t n = min(*n_ptr, cap);
    const uint32_t stride = 4u * gridDim.x * blockDim.x;
    for (uint32_t i0 = 4u * (blockIdx.x * blockDim.x + threadIdx.x); i0 < n; i0 += stride) {
        // the key buffer holds 2 cap words (it is the 64-bit keys' buffer): the 16-byte load stays inside it
        const uint4 q = *reinterpret_cast<const uint4*>(keys + i0);
        const uint32_t kk[4] = {q.x, q.y, q.z, q.w};
        uint32_t prev = i0 > 0 ? range_tile(keys[i0 - 1], ktile_gtile, n_ktiles) : 0xFFFFFFFFu;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t i = i0 + (uint32_t)u;
            if (i >= n) break;
            const uint32_t tile = range_tile(kk[u], ktile_gtile, n_ktiles);
            if (tile != 0xFFFFFFFFu) {
                if (i == 0) {
                    ranges[2ull * tile] = i;
                } else if (tile != prev) {
                    ranges[2ull * tile] = i;
                    if (prev != 0xFFFFFFFFu) ranges[2ull * prev + 1] = i;
                }
                if (i == n - 1) ranges[2ull * tile + 1] = i + 1;
            }
            prev = tile;
        }
    }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
void launch_prep_records(const float* verts, const uint32_t* idx, const float* dens, const float* grads, uint32_t M,
                         float4* rec, hipStream_t s) {
    if (!M) return;
    hipLaunchKernelGGL(k_prep_records, dim3((M + 255) / 256), dim3(256), 0, s, verts, idx, dens, grads, M, rec);
}
void launch_copy_colors(const float* base, uint32_t M, float4* col, hipStream_t s) {
    if (!M) return;
    hipLaunchKernelGGL(k_copy_colors, dim3((M + 255) / 256), dim3(256), 0, s, base, M, col);
}
void launch_fill_u32(uint32_t* p, uint32_t value, uint64_t n, hipStream_t s) {
    if (!n) return;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)blocks), dim3(256), 0, s, p, value, n);
}
void launch_copy_probe(const float4* src, float4* dst, uint64_t n16, hipStream_t s) {
    hipLaunchKernelGGL(k_copy_probe, dim3(4096), dim3(256), 0, s, src, dst, n16);
}
uint32_t item_blocks(uint64_t n_items) { return (uint32_t)((n_items + 255) / 256); }
static bool pow2(uint32_t ts) { return ts && !(ts & (ts - 1u)); }
void launch_project(const float4* rec, const float4* circ, const SceneRange* scenes, uint32_t n_scenes,
                    uint64_t n_items, const ViewParams* views, uint32_t tile_size, uint32_t* counts,
                    uint32_t* block_sums, uint32_t* block_vis, uint32_t* dkey, uint32_t* ditem,
                    uint32_t* n_items_out, hipStream_t s) {
    if (!n_items) return;
    if (pow2(tile_size))
        hipLaunchKernelGGL(k_project<true>, dim3(item_blocks(n_items)), dim3(256), 0, s, rec, circ, scenes, n_scenes,
                           n_items, views, counts, block_sums, block_vis, dkey, ditem, n_items_out);
    else
        hipLaunchKernelGGL(k_project<false>, dim3(item_blocks(n_items)), dim3(256), 0, s, rec, circ, scenes, n_scenes,
                           n_items, views, counts, block_sums, block_vis, dkey, ditem, n_items_out);
}
uint32_t spine_parts(uint32_t n) { return (n + kSpineTile - 1) / kSpineTile + 1; }
void launch_spine_scan(uint32_t* sums, uint32_t n, uint32_t* total, uint32_t* overflow, hipStream_t s,
                       const uint32_t* aux, uint32_t* aux_total, unsigned long long* part, uint32_t* aux_part) {
    if (part && n > kSpineTile) {
        const uint32_t nb = (n + kSpineTile - 1) / kSpineTile;
        hipLaunchKernelGGL(k_spine_local, dim3(nb), dim3(1024), 0, s, sums, n, part, aux, aux ? aux_part : nullptr);
        hipLaunchKernelGGL(k_spine_top, dim3(1), dim3(64), 0, s, part, nb, total, overflow, aux ? aux_part : nullptr,
                           aux_total);
        hipLaunchKernelGGL(k_spine_add, dim3(nb), dim3(1024), 0, s, sums, n, part);
        return;
    }
    hipLaunchKernelGGL(k_spine_scan, dim3(1), dim3(1024), 0, s, sums, n, total, overflow, aux, aux_total);
}
void launch_pairs(const float4* rec, const float4* circ, const SceneRange* scenes, uint32_t n_scenes, uint64_t n_items,
                  const ViewParams* views, uint32_t tile_size, const uint32_t* counts, const uint32_t* block_off,
                  uint64_t* keys, uint32_t* vals, uint32_t cap, uint32_t* overflow, hipStream_t s) {
    if (!n_items) return;
    if (pow2(tile_size))
        hipLaunchKernelGGL(k_pairs<true>, dim3(item_blocks(n_items)), dim3(256), 0, s, rec, circ, scenes, n_scenes,
                           n_items, views, counts, block_off, keys, vals, cap, overflow);
    else
        hipLaunchKernelGGL(k_pairs<false>, dim3(item_blocks(n_items)), dim3(256), 0, s, rec, circ, scenes, n_scenes,
                           n_items, views, counts, block_off, keys, vals, cap, overflow);
}
void launch_pairs_depth_ordered(const float4* rec, const SceneRange* scenes, uint32_t n_scenes, const uint32_t* sval,
                                uint32_t n_items, const ViewParams* views, uint32_t tile_size, const uint32_t* counts,
                                uint32_t* block_sums, uint32_t* n_pairs, uint32_t* keys, uint32_t* vals, uint32_t cap,
                                uint32_t* overflow, unsigned long long* spine_part, hipStream_t s) {
    if (!n_items) return;
    const dim3 gr(item_blocks(n_items)), bl(256);
    hipLaunchKernelGGL(k_sorted_sums, gr, bl, 0, s, sval, n_items, counts, block_sums);
    launch_spine_scan(block_sums, item_blocks(n_items), n_pairs, overflow, s, nullptr, nullptr, spine_part,
                      nullptr);  // the same total P
    if (pow2(tile_size))
        hipLaunchKernelGGL(k_pairs_sorted<true>, gr, bl, 0, s, rec, scenes, n_scenes, sval, n_items, views, counts,
                           block_sums, keys, vals, cap, overflow);
    else
        hipLaunchKernelGGL(k_pairs_sorted<false>, gr, bl, 0, s, rec, scenes, n_scenes, sval, n_items, views, counts,
                           block_sums, keys, vals, cap, overflow);
}
// Tiles heaviest first: the compositing grids take their tiles in this order,
// so the few very dense tiles (tens of times the median pair count) start
// early instead of running alone at the end of the grid (longest-processing-
// time-first).  A bucket sort by the pair count's bit length and its next two
// bits (four buckets per octave), descending; the order inside a bucket is
// whatever the LDS atomics give (every tile's result is independent of when
// it runs).
constexpr int kOrderBuckets = 33 * 4;
__device__ __forceinline__ uint32_t tile_bucket(const uint32_t* __restrict__ ranges, uint32_t t) {
    const uint32_t rs = ranges[2ull * t], re = ranges[2ull * t + 1];
    const uint32_t c = re > rs ? re - rs : 0u;
    if (c < 4u) return c;
    const uint32_t len = 32u - (uint32_t)__clz((int)c);  // >= 3
    return len * 4u + ((c >> (len - 3u)) & 3u);           // 3 leading bits: 1xx
}


__global__ __launch_bounds__(1024) void k_tile_order(const uint32_t* __restrict__ ranges, uint32_t n,
                                                      uint32_t* __restrict__ order) {
    __shared__ uint32_t cnt[kOrderBuckets];
    __shared__ uint32_t off[kOrderBuckets];
    const uint32_t tid = threadIdx.x;
    if (tid < kOrderBuckets) cnt[tid] = 0u;
    __syncthreads();
    for (uint32_t t = tid; t < n; t += blockDim.x) atomicAdd(&cnt[tile_bucket(ranges, t)], 1u);
    __syncthreads();
    if (tid == 0) {
        uint32_t run = 0;
        for (int b = kOrderBuckets - 1; b >= 0; --b) {
            off[b] = run;
            run += cnt[b];
        }
    }
    __syncthreads();
    for (uint32_t t = tid; t < n; t += blockDim.x) order[atomicAdd(&off[tile_bucket(ranges, t)], 1u)] = t;
}

// Scene-major order for batches of several scenes: the scenes one after
// another, the one with the fewest pairs first (ties: the caller's order), and
// the tiles of each scene heaviest first as above.  A scene's tiles then finish
// together early in the step instead of all scenes finishing at its end; the
// tile grid is the same, so the total work is too.  Counts per (scene, bucket)
// in LDS, then offsets by one thread.
constexpr uint32_t kOrderMaxScenes = 64;
__global__ __launch_bounds__(1024) void k_tile_order_scenes(const uint32_t* __restrict__ ranges, uint32_t n,
                                                            const uint32_t* __restrict__ tile_view,
                                                            const ViewParams* __restrict__ views, uint32_t n_scenes,
                                                            uint32_t* __restrict__ order) {
    __shared__ uint32_t cnt[kOrderMaxScenes * kOrderBuckets];  // counts, then running offsets
    __shared__ unsigned long long s_pairs[kOrderMaxScenes];
    __shared__ uint32_t s_by_rank[kOrderMaxScenes];
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < n_scenes * kOrderBuckets; i += blockDim.x) cnt[i] = 0u;
    if (tid < n_scenes) s_pairs[tid] = 0ull;
    __syncthreads();
    for (uint32_t t = tid; t < n; t += blockDim.x) {
        const uint32_t sc = views[tile_view[t]].scene;
        atomicAdd(&cnt[sc * kOrderBuckets + tile_bucket(ranges, t)], 1u);
        const uint32_t rs = ranges[2ull * t], re = ranges[2ull * t + 1];
        atomicAdd(&s_pairs[sc], (unsigned long long)(re > rs ? re - rs : 0u));
    }
    __syncthreads();
    if (tid < n_scenes) {
        uint32_t r = 0;
        for (uint32_t o = 0; o < n_scenes; ++o) r += (s_pairs[o] < s_pairs[tid] || (s_pairs[o] == s_pairs[tid] && o < tid)) ? 1u : 0u;
        s_by_rank[r] = tid;
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t run = 0;
        for (uint32_t r = 0; r < n_scenes; ++r) {
            uint32_t* c = cnt + s_by_rank[r] * kOrderBuckets;
            for (int b = kOrderBuckets - 1; b >= 0; --b) {
                const uint32_t k = c[b];
                c[b] = run;
                run += k;
            }
        }
    }
    __syncthreads();
    for (uint32_t t = tid; t < n; t += blockDim.x) {
        const uint32_t sc = views[tile_view[t]].scene;
        order[atomicAdd(&cnt[sc * kOrderBuckets + tile_bucket(ranges, t)], 1u)] = t;
    }
}

__global__ void k_stamp_start(uint64_t* __restrict__ p, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0ull;
    if (threadIdx.x == 0) p[n] = wall_clock64();
}
void launch_stamp_start(uint64_t* p, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_stamp_start, dim3(1), dim3(64), 0, s, p, n);
}

void launch_tile_order(const uint32_t* ranges, uint32_t n_tiles, uint32_t* order, hipStream_t s) {
    if (n_tiles) hipLaunchKernelGGL(k_tile_order, dim3(1), dim3(1024), 0, s, ranges, n_tiles, order);
}

bool launch_tile_order_scenes(const uint32_t* ranges, uint32_t n_tiles, const uint32_t* tile_view,
                              const ViewParams* views, uint32_t n_scenes, uint32_t* order, hipStream_t s) {
    if (n_scenes == 0 || n_scenes > kOrderMaxScenes) return false;
    if (n_tiles)
        hipLaunchKernelGGL(k_tile_order_scenes, dim3(1), dim3(1024), 0, s, ranges, n_tiles, tile_view, views, n_scenes,
                           order);
    return true;
}

static uint32_t range_blocks(uint32_t cap) {
    uint32_t blocks = (cap + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks == 0) blocks = 1;
    return blocks;
}
void launch_ranges(const uint64_t* keys, const uint32_t* n_ptr, uint32_t cap, uint32_t* ranges,
                   const uint32_t* ktile_gtile, uint32_t n_ktiles, hipStream_t s) {
    hipLaunchKernelGGL(k_ranges<uint64_t>, dim3(range_blocks(cap)), dim3(256), 0, s, keys, n_ptr, cap, ranges,
                       ktile_gtile, n_ktiles);
}
void launch_ranges32(const uint32_t* keys, const uint32_t* n_ptr, uint32_t cap, uint32_t* ranges,
                     const uint32_t* ktile_gtile, uint32_t n_ktiles, hipStream_t s) {
    hipLaunchKernelGGL(k_ranges32x4, dim3(range_blocks((cap + 3) / 4)), dim3(256), 0, s, keys, n_ptr, cap, ranges,
                       ktile_gtile, n_ktiles);
}
void launch_full_keys(const uint32_t* ktile, const uint32_t* vals, const uint32_t* n_ptr, uint32_t cap,
                      const float4* circ, const ViewParams* views, const uint32_t* ktile_gtile, uint32_t n_ktiles,
                      const uint32_t* tile_view, uint64_t* out, uint32_t n_host, hipStream_t s) {
    if (!n_host) return;
    hipLaunchKernelGGL(k_full_keys, dim3((n_host + 255) / 256), dim3(256), 0, s, ktile, vals, n_ptr, cap, circ, views,
                       ktile_gtile, n_ktiles, tile_view, out);
}

}  // namespace rmesh
