// Tiled ray–tetrahedron compositing: forward (front to back) and adjoint
// (back to front with state undo).
//
// Reference kernels restated:
//   crates/rmesh-render/src/wgsl/rasterize_compute.wgsl:92-511        -> k_forward
//   crates/rmesh-train/src/wgsl/loss_compute.wgsl:24-72               -> loss terms in k_backward init
//   crates/rmesh-backward/src/wgsl/backward_tiled_compute.wgsl:154-587 -> k_backward
//
// MI355X design.  One workgroup per (view, tile).  The tile's depth-sorted tet
// list is consumed in batches; per batch:
//   1. one thread per tet gathers the 64-byte tet record and derives the inward
//      face planes, the base colour at the camera and the tile-local projected
//      vertices into LDS;
//   2. all lanes compute the per-row scanline spans of the batch; one wave
//      turns them into a prefix over (tet, row), which enumerates every covered
//      (pixel, tet) pair of the batch in a compact range [0, C);
//   3. pair-parallel phase: lanes take consecutive pairs, intersect the pixel
//      ray with the tet and evaluate the volume integral (all lanes busy, no
//      per-tet divergence) and write a small record per pair to LDS;
//   4. pixel-sequential phase: each lane owns one pixel and applies the
//      order-dependent recurrence of its pairs in depth order (compositing in
//      the forward; transmittance undo and the d(log T) chain in the adjoint);
//   5. adjoint only, pair-parallel again: gradient terms per pair, a segmented
//      wave64 scan keyed by tet sums each tet's terms, segment tails add them
//      into an LDS accumulator, and the batch is flushed with one hardware
//      float atomic per term per (tile, tet) — the reference's atomic count.
// Per-pixel state lives in VGPRs; ray directions and d(image) live in LDS.
//
// Per-pixel semantics (coverage rules, skip thresholds, clamps, undo order)
// are the reference's; the spans use the reference's tile-local arithmetic so
// coverage decisions are identical.  The adjoint's colour undo
// (backward_tiled_compute.wgsl:426-429) feeds no output and is not computed.
#include "rmesh_common.hpp"
#include "rmesh_kernels.hpp"

namespace rmesh {

struct TileCtx {
    uint32_t tile, tile_x, tile_y;
    float tile_ox, tile_oy;
};

__device__ __forceinline__ TileCtx tile_ctx(const ViewParams& vw, uint32_t gtile) {
    TileCtx c;
    c.tile = gtile - vw.tile_base;
    c.tile_x = c.tile % vw.tiles_x;
    c.tile_y = c.tile / vw.tiles_x;
    c.tile_ox = (float)(c.tile_x * vw.tile_size);
    c.tile_oy = (float)(c.tile_y * vw.tile_size);
    return c;
}

// Row span of the tet in tile-local pixels (rasterize_compute.wgsl:170-275;
// FORWARD adds the vertex-in-row rule, the adjoint omits it:
// backward_tiled_compute.wgsl:242-299).  Packed as (xl & 0xffff) | (xr << 16).
template <bool FORWARD>
__device__ __forceinline__ uint32_t row_span(const float* px_, const float* py_, uint32_t row, int32_t TS) {
    float xl_f = 1e10f, xr_f = -1e10f;
    float yc = (float)row + 0.5f;
    const int EI[6] = {0, 0, 0, 1, 1, 2};
    const int EJ[6] = {1, 2, 3, 2, 3, 3};
#pragma unroll
    for (int e = 0; e < 6; ++e) {
        float yi = py_[EI[e]], yj = py_[EJ[e]];
        if ((yi <= yc && yj > yc) || (yj <= yc && yi > yc)) {
            float t = (yc - yi) / (yj - yi);
            float x = px_[EI[e]] + t * (px_[EJ[e]] - px_[EI[e]]);
            xl_f = fminf(xl_f, x);
            xr_f = fmaxf(xr_f, x);
        }
    }
    if (FORWARD) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            float vy = py_[k];
            if (vy >= yc - 0.5f && vy < yc + 0.5f) {
                xl_f = fminf(xl_f, px_[k]);
                xr_f = fmaxf(xr_f, px_[k]);
            }
        }
    }
    int32_t xl = 0, xr = -1;
    if (xl_f <= xr_f) {
        const float eps = 0.001f;
        int32_t l = max(sat_i32(ceilf(xl_f - 0.5f - eps)), 0);
        int32_t r = min(sat_i32(floorf(xr_f - 0.5f + eps)), TS - 1);
        if (l <= r) {
            xl = l;
            xr = r;
        }
    }
    return ((uint32_t)xl & 0xffffu) | ((uint32_t)xr << 16);
}

__device__ __forceinline__ int32_t span_xl(uint32_t sp) { return (int32_t)(int16_t)(sp & 0xffffu); }
__device__ __forceinline__ int32_t span_xr(uint32_t sp) { return (int32_t)(int16_t)(sp >> 16); }

// LDS image of one batch of tets.  Row arrays use an odd stride (TS + 1) so
// the pixel-sequential phase (fixed tet, lanes on different rows) is free of
// bank conflicts.
template <int TS, int B, bool BWD>
struct BatchLDS {
    static constexpr int SP = TS + 1;
    float n[4][3][B];       // inward face normals
    float num[4][B];        // n . (va - cam)
    float bc[3][B];         // colour + g . (cam - v0)
    float g[3][B];
    float dens[B];
    float pxy[8][B];        // tile-local projected vertices (x0..x3, y0..y3)
    uint32_t behind[B];
    uint32_t span[B * SP];  // [tet][row] packed span
    uint32_t rpref[B * SP]; // [tet][row] exclusive prefix of row widths; [tet][TS] = tet total
    uint32_t tpref[B + 1];  // exclusive prefix of tet totals
    uint32_t nbp;           // tets processed this batch (capacity cut)
    // adjoint only
    float v[BWD ? 4 : 1][3][B];
    uint32_t vi[BWD ? 4 : 1][B];
    uint32_t tid[BWD ? B : 1];
    uint32_t flip[BWD ? B : 1];
};

template <int TS, int B, bool BWD>
__device__ __forceinline__ void load_tet(BatchLDS<TS, B, BWD>& L, int k, uint32_t t, const float4* __restrict__ rec,
                                         const uint32_t* __restrict__ idx, const float4* __restrict__ colors,
                                         const ViewParams& vw, const TileCtx& tc) {
    float4 r0 = rec[4ull * t], r1 = rec[4ull * t + 1], r2 = rec[4ull * t + 2], r3 = rec[4ull * t + 3];
    V3 v[4] = {mk3(r0.x, r0.y, r0.z), mk3(r1.x, r1.y, r1.z), mk3(r2.x, r2.y, r2.z), mk3(r3.x, r3.y, r3.z)};
    V3 g = mk3(r1.w, r2.w, r3.w);
    V3 cam = mk3(vw.cam[0], vw.cam[1], vw.cam[2]);
    float4 col = colors[t];
    uint32_t flips = 0;
#pragma unroll
    for (int fi = 0; fi < 4; ++fi) {
        int a, b, c, o;
        face_idx(fi, a, b, c, o);
        V3 nn = cross3(sub3(v[c], v[a]), sub3(v[b], v[a]));
        if (dot3(nn, sub3(v[o], v[a])) < 0.0f) {
            nn = neg3(nn);
            flips |= 1u << fi;
        }
        L.n[fi][0][k] = nn.x;
        L.n[fi][1][k] = nn.y;
        L.n[fi][2][k] = nn.z;
        L.num[fi][k] = dot3(nn, sub3(v[a], cam));
    }
    float bo = dot3(g, sub3(cam, v[0]));
    L.bc[0][k] = col.x + bo;
    L.bc[1][k] = col.y + bo;
    L.bc[2][k] = col.z + bo;
    L.g[0][k] = g.x;
    L.g[1][k] = g.y;
    L.g[2][k] = g.z;
    L.dens[k] = r0.w;
    float clip[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i) mul_vp(vw.vp, v[i], clip[i]);
    bool behind = (clip[0][3] <= 0.0f) || (clip[1][3] <= 0.0f) || (clip[2][3] <= 0.0f) || (clip[3][3] <= 0.0f);
    L.behind[k] = behind ? 1u : 0u;
    if (!behind) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            L.pxy[i][k] = ((clip[i][0] / clip[i][3] + 1.0f) * 0.5f) * vw.Wf - tc.tile_ox;
            L.pxy[4 + i][k] = ((1.0f - clip[i][1] / clip[i][3]) * 0.5f) * vw.Hf - tc.tile_oy;
        }
    }
    if (BWD) {
        uint4 q = reinterpret_cast<const uint4*>(idx)[t];
        uint32_t vv[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            L.v[i][0][k] = v[i].x;
            L.v[i][1][k] = v[i].y;
            L.v[i][2][k] = v[i].z;
            L.vi[i][k] = vv[i];
        }
        L.tid[k] = t;
        L.flip[k] = flips;
    }
}

// Spans of the batch, then (wave 0) per-tet row prefixes, the tet prefix and
// the capacity cut.  Ends with the batch's pair count in L.tpref[L.nbp].
template <int TS, int B, bool BWD, bool FORWARD_RULE, int CAP, int NT>
__device__ __forceinline__ void batch_layout(BatchLDS<TS, B, BWD>& L, int nb) {
    constexpr int SP = TS + 1;
    for (int it = threadIdx.x; it < nb * TS; it += NT) {
        int k = it / TS, row = it % TS;
        uint32_t sp = 0xffff0000u;  // xl = 0, xr = -1 (empty)
        if (!L.behind[k]) {
            float px_[4] = {L.pxy[0][k], L.pxy[1][k], L.pxy[2][k], L.pxy[3][k]};
            float py_[4] = {L.pxy[4][k], L.pxy[5][k], L.pxy[6][k], L.pxy[7][k]};
            sp = row_span<FORWARD_RULE>(px_, py_, (uint32_t)row, TS);
        }
        L.span[k * SP + row] = sp;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const int k = threadIdx.x;
        uint32_t tot = 0;
        if (k < nb) {
#pragma unroll
            for (int r = 0; r < TS; ++r) {
                uint32_t sp = L.span[k * SP + r];
                L.rpref[k * SP + r] = tot;
                tot += (uint32_t)max(span_xr(sp) - span_xl(sp) + 1, 0);
            }
            L.rpref[k * SP + TS] = tot;
        }
        uint32_t inc = wave_incl_scan(tot);
        if (k < nb) L.tpref[k + 1] = inc;
        if (k == 0) L.tpref[0] = 0;
        uint64_t fits = __ballot(k < nb && inc <= (uint32_t)CAP);
        if (k == 0) L.nbp = max((uint32_t)__popcll(fits), 1u);
    }
    __syncthreads();
}

// (tet, row, col) of pair `idx` of the batch.
template <int TS, int B, bool BWD>
__device__ __forceinline__ void pair_locate(const BatchLDS<TS, B, BWD>& L, uint32_t idx, uint32_t nbp, int& k, int& row,
                                            int& col) {
    constexpr int SP = TS + 1;
    int lo = 0, hi = (int)nbp;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (L.tpref[mid] <= idx) lo = mid;
        else hi = mid;
    }
    k = lo;
    uint32_t local = idx - L.tpref[k];
    int rl = 0, rh = TS;
    while (rh - rl > 1) {
        int mid = (rl + rh) >> 1;
        if (L.rpref[k * SP + mid] <= local) rl = mid;
        else rh = mid;
    }
    row = rl;
    col = span_xl(L.span[k * SP + row]) + (int)(local - L.rpref[k * SP + row]);
}

// Ray–tet slab intersection against the batch's inward planes
// (rasterize_compute.wgsl:354-390).
template <int TS, int B, bool BWD>
__device__ __forceinline__ bool intersect(const BatchLDS<TS, B, BWD>& L, int k, V3 rd, float min_t, float& t_min,
                                          float& t_max, int& fmin, int& fmax) {
    t_min = -3.402823e38f;
    t_max = 3.402823e38f;
    fmin = 4;
    fmax = 0;
    bool valid = true;
#pragma unroll
    for (int fi = 0; fi < 4; ++fi) {
        V3 nn = mk3(L.n[fi][0][k], L.n[fi][1][k], L.n[fi][2][k]);
        float den = dot3(nn, rd);
        float num = L.num[fi][k];
        if (min_t != 0.0f) num = num - min_t * den;
        if (fabsf(den) < 1e-20f) {
            if (num > 0.0f) valid = false;
            continue;
        }
        float t = num / den;
        if (den > 0.0f) {
            if (t > t_min) {
                t_min = t;
                fmin = fi;
            }
        } else if (t < t_max) {
            t_max = t;
            fmax = fi;
        }
    }
    return valid && t_min < t_max;
}

// ---------------------------------------------------------------------------
// Forward
// ---------------------------------------------------------------------------
template <int TS>
struct FwdCfg {
    static constexpr int NT = (TS * TS + 63) / 64 * 64;
    static constexpr int B = 64;
    static constexpr int CAP = 1024;
};

template <int TS, bool AUX>
__global__ __launch_bounds__(FwdCfg<TS>::NT) void k_forward(
    const float4* __restrict__ rec, const float4* __restrict__ colors, uint64_t color_view_stride,
    const uint32_t* __restrict__ sorted_vals, const uint32_t* __restrict__ ranges, const ViewParams* __restrict__ views,
    uint32_t tiles_per_view, float* __restrict__ rgba, float* __restrict__ aux) {
    using C = FwdCfg<TS>;
    constexpr int NT = C::NT, B = C::B, CAP = C::CAP, SP = TS + 1;
    constexpr int NA = AUX ? CAP : 1;
    __shared__ BatchLDS<TS, B, false> L;
    __shared__ float s_od[CAP];
    __shared__ float s_cp[3][CAP];
    __shared__ float s_aux[3][NA];     // t_min, t_max, w0 (w1 = 1 - phi recomputed from w0 + alpha_t)
    __shared__ uint32_t s_ent[NA];
    __shared__ float s_ray[3][TS * TS];

    const uint32_t gtile = blockIdx.x;
    const uint32_t view = gtile / tiles_per_view;
    const ViewParams& vw = views[view];
    const TileCtx tc = tile_ctx(vw, gtile);
    const float4* col_v = colors + color_view_stride * view;
    const int tid = threadIdx.x;
    const int row = tid / TS, colp = tid % TS;
    const uint32_t px = tc.tile_x * TS + colp, py = tc.tile_y * TS + row;
    const bool active = tid < TS * TS && px < vw.width && py < vw.height;
    const float min_t = vw.min_t;
    const V3 cam = mk3(vw.cam[0], vw.cam[1], vw.cam[2]);

    if (tid < TS * TS) {
        V3 rd = mk3(0.f, 0.f, 1.f);
        if (active) rd = pixel_ray(vw, px, py);
        s_ray[0][tid] = rd.x;
        s_ray[1][tid] = rd.y;
        s_ray[2][tid] = rd.z;
    }
    float cr = 0.f, cg = 0.f, cb = 0.f, log_t = 0.f;
    float nx = 0.f, ny = 0.f, nz = 0.f, dep = 0.f, e0 = 0.f, e1 = 0.f, e2 = 0.f, e3 = 0.f;

    const uint32_t rs = ranges[2ull * gtile], re = ranges[2ull * gtile + 1];
    const uint32_t count = re > rs ? re - rs : 0u;
    uint32_t done = 0;
    while (done < count) {
        if (__syncthreads_and(!active || log_t < -13.8f)) break;
        const int nb = (int)min<uint32_t>(B, count - done);
        if (tid < nb) load_tet<TS, B, false>(L, tid, sorted_vals[re - 1 - done - tid], rec, nullptr, col_v, vw, tc);
        __syncthreads();
        batch_layout<TS, B, false, true, CAP, NT>(L, nb);
        const uint32_t nbp = L.nbp;
        const uint32_t npairs = L.tpref[nbp];
        // pair-parallel: intersection + volume integral
        for (uint32_t i = tid; i < npairs; i += NT) {
            int k, r, c;
            pair_locate<TS, B, false>(L, i, nbp, k, r, c);
            const uint32_t gx = tc.tile_x * TS + c, gy = tc.tile_y * TS + r;
            float od_rec = -1.0f;
            float cp0 = 0.f, cp1 = 0.f, cp2 = 0.f;
            float tmn = 0.f, tmx = 0.f, w0 = 0.f;
            int ent = 4;
            if (gx < vw.width && gy < vw.height) {
                const int p = r * TS + c;
                V3 rd = mk3(s_ray[0][p], s_ray[1][p], s_ray[2][p]);
                float t_min, t_max;
                int fmin, fmax;
                if (intersect<TS, B, false>(L, k, rd, min_t, t_min, t_max, fmin, fmax)) {
                    V3 g = mk3(L.g[0][k], L.g[1][k], L.g[2][k]);
                    float dc_dt = dot3(g, rd);
                    float bo = min_t != 0.0f ? min_t * dc_dt : 0.0f;
                    float b0 = L.bc[0][k] + bo, b1 = L.bc[1][k] + bo, b2 = L.bc[2][k] + bo;
                    float a0 = dc_dt * t_min, a1 = dc_dt * t_max;
                    float cs0 = fmaxf(b0 + a0, 0.f), cs1 = fmaxf(b1 + a0, 0.f), cs2 = fmaxf(b2 + a0, 0.f);
                    float ce0 = fmaxf(b0 + a1, 0.f), ce1 = fmaxf(b1 + a1, 0.f), ce2 = fmaxf(b2 + a1, 0.f);
                    float dist = t_max - t_min;
                    float od = clampf(L.dens[k] * dist, 1e-8f, 88.0f);
                    float alpha_t = safe_exp(-od);
                    float ph = phi_f(od);
                    w0 = ph - alpha_t;
                    float w1 = 1.0f - ph;
                    cp0 = safe_clip(ce0 * w0 + cs0 * w1, 0.0f, kMaxVal);
                    cp1 = safe_clip(ce1 * w0 + cs1 * w1, 0.0f, kMaxVal);
                    cp2 = safe_clip(ce2 * w0 + cs2 * w1, 0.0f, kMaxVal);
                    od_rec = od;
                    tmn = t_min;
                    tmx = t_max;
                    ent = fmin;
                }
            }
            s_od[i] = od_rec;
            s_cp[0][i] = cp0;
            s_cp[1][i] = cp1;
            s_cp[2][i] = cp2;
            if (AUX) {
                s_aux[0][i] = tmn;
                s_aux[1][i] = tmx;
                s_aux[2][i] = w0;
                s_ent[i] = (uint32_t)ent;
            }
        }
        __syncthreads();
        // pixel-sequential compositing in depth order (rasterize_compute.wgsl:336-463)
        if (active) {
            for (uint32_t k = 0; k < nbp; ++k) {
                const uint32_t sp = L.span[k * SP + row];
                const int xl = span_xl(sp);
                if (colp < xl || colp > span_xr(sp)) continue;
                if (log_t < -13.8f) continue;
                const uint32_t slot = L.tpref[k] + L.rpref[k * SP + row] + (uint32_t)(colp - xl);
                const float od = s_od[slot];
                if (od < 0.0f) continue;
                const float Tj = safe_exp(log_t);
                if (Tj >= 1e-6f) {
                    cr = cr + s_cp[0][slot] * Tj;
                    cg = cg + s_cp[1][slot] * Tj;
                    cb = cb + s_cp[2][slot] * Tj;
                    log_t = log_t - od;
                    if (AUX) {
                        // rasterize_compute.wgsl:434-456
                        float alpha_t = safe_exp(-od);
                        float alpha = 1.0f - alpha_t;
                        float w0 = s_aux[2][slot];
                        float w1 = 1.0f - (w0 + alpha_t);
                        float t_min = s_aux[0][slot], t_max = s_aux[1][slot];
                        int entry = (int)s_ent[slot];
                        V3 en = mk3(0.f, 0.f, 0.f);
                        if (entry < 4) en = normalize3(neg3(mk3(L.n[entry][0][k], L.n[entry][1][k], L.n[entry][2][k])));
                        float depth_premul = w0 * t_max + w1 * t_min;
                        float ta = Tj * alpha;
                        nx = nx + ta * en.x;
                        ny = ny + ta * en.y;
                        nz = nz + ta * en.z;
                        dep = dep + Tj * depth_premul;
                        float wt = alpha * Tj;
                        float c_mid = (t_min + t_max) * 0.5f;
                        float wc = wt * c_mid;
                        e0 = e0 + wt;
                        e1 = e1 + (wt > 1e-20f ? wt * logf(wt) : 0.0f);
                        e2 = e2 + wt * c_mid;
                        e3 = e3 + (wc > 1e-20f ? wc * logf(wc) : 0.0f);
                    }
                }
            }
        }
        done += nbp;
    }
    if (active) {
        uint64_t pix = ((uint64_t)view * vw.height + py) * vw.width + px;
        reinterpret_cast<float4*>(rgba)[pix] = make_float4(cr, cg, cb, 1.0f - safe_exp(log_t));
        if (AUX) {
            reinterpret_cast<float4*>(aux)[2 * pix] = make_float4(nx, ny, nz, dep);
            reinterpret_cast<float4*>(aux)[2 * pix + 1] = make_float4(e0, e1, e2, e3);
        }
    }
}

// ---------------------------------------------------------------------------
// Adjoint
// ---------------------------------------------------------------------------
constexpr int kTerms = 19;  // d_density, d_v[4][3], d_grad[3], d_colour[3]

template <int TS>
struct BwdCfg {
    static constexpr int NT = (TS * TS + 63) / 64 * 64;
    static constexpr int B = 32;
    static constexpr int CAP = 768;
};

// Segmented inclusive sum across the wave; lanes with equal `key` form
// contiguous segments (keys non-decreasing with the lane index).
struct SegMasks {
    bool m[6];
};
__device__ __forceinline__ SegMasks seg_masks(int key) {
    SegMasks s;
    const int lane = (int)__lane_id();
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        const int o = 1 << j;
        int up = __shfl_up(key, o, 64);
        s.m[j] = lane >= o && up == key;
    }
    return s;
}
__device__ __forceinline__ float seg_scan(float v, const SegMasks& s) {
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        float up = __shfl_up(v, 1 << j, 64);
        if (s.m[j]) v += up;
    }
    return v;
}

template <int TS>
__global__ __launch_bounds__(BwdCfg<TS>::NT) void k_backward(
    const float4* __restrict__ rec, const uint32_t* __restrict__ idx, const float4* __restrict__ colors,
    uint64_t color_view_stride, const uint32_t* __restrict__ sorted_vals, const uint32_t* __restrict__ ranges,
    const ViewParams* __restrict__ views, uint32_t tiles_per_view, const float* __restrict__ rgba,
    const float* __restrict__ dl_img, const float* __restrict__ gt, uint32_t loss_type,
    float* __restrict__ loss_partial, float* __restrict__ d_verts, float* __restrict__ d_dens,
    float* __restrict__ d_grads, float* __restrict__ d_colors, uint64_t dcolor_view_stride) {
    using C = BwdCfg<TS>;
    constexpr int NT = C::NT, B = C::B, CAP = C::CAP, SP = TS + 1, NW = NT / 64;
    __shared__ BatchLDS<TS, B, true> L;
    // per-pair record: od -> t_prev, cp0 -> d_log_t before, cp1, cp2, t_min, t_max, meta
    __shared__ float s_r[6][CAP];
    __shared__ uint32_t s_meta[CAP];   // fmin | fmax << 2 | valid << 4 | did << 5
    __shared__ float s_ray[3][TS * TS];
    __shared__ float s_dc[3][TS * TS];
    __shared__ float s_acc[B][kTerms];
    __shared__ float s_loss[NW];

    const uint32_t gtile = blockIdx.x;
    const uint32_t view = gtile / tiles_per_view;
    const ViewParams& vw = views[view];
    const TileCtx tc = tile_ctx(vw, gtile);
    const float4* col_v = colors + color_view_stride * view;
    float* dcol_v = d_colors + 3ull * dcolor_view_stride * view;
    const int tid = threadIdx.x, wid = tid / 64, lane = (int)__lane_id();
    const int row = tid / TS, colp = tid % TS;
    const uint32_t px = tc.tile_x * TS + colp, py = tc.tile_y * TS + row;
    const bool active = tid < TS * TS && px < vw.width && py < vw.height;
    const uint64_t pix = ((uint64_t)view * vw.height + py) * vw.width + px;
    const V3 cam = mk3(vw.cam[0], vw.cam[1], vw.cam[2]);

    // Initial state from the rendered image (backward_tiled_compute.wgsl:186-204);
    // with a ground-truth image the loss gradient is formed here
    // (loss_compute.wgsl:24-72): dL/dalpha = 0.
    float log_t = 0.f, d_log_t = 0.f, lsum = 0.f;
    if (tid < TS * TS) {
        float dcv[3] = {0.f, 0.f, 0.f};
        V3 rd = mk3(0.f, 0.f, 1.f);
        if (active) {
            rd = pixel_ray(vw, px, py);
            float4 im = reinterpret_cast<const float4*>(rgba)[pix];
            log_t = logf(fmaxf(1.0f - im.w, 1e-20f));
            if (gt) {
                float npx = (float)(vw.width * vw.height);
                float df[3] = {im.x - gt[3 * pix], im.y - gt[3 * pix + 1], im.z - gt[3 * pix + 2]};
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    if (loss_type == 0) {
                        float sg = df[c] > 0.f ? 1.f : (df[c] < 0.f ? -1.f : 0.f);
                        dcv[c] = sg / npx;
                        lsum += fabsf(df[c]);
                    } else {
                        dcv[c] = (2.0f * df[c]) / npx;
                        lsum += df[c] * df[c];
                    }
                }
            } else {
                float4 d = reinterpret_cast<const float4*>(dl_img)[pix];
                dcv[0] = d.x;
                dcv[1] = d.y;
                dcv[2] = d.z;
                d_log_t = -d.w * (1.0f - im.w);
            }
        }
        s_ray[0][tid] = rd.x;
        s_ray[1][tid] = rd.y;
        s_ray[2][tid] = rd.z;
        s_dc[0][tid] = dcv[0];
        s_dc[1][tid] = dcv[1];
        s_dc[2][tid] = dcv[2];
    }
    if (gt && loss_partial) {
        float ws = wave_sum(lsum);
        if (lane == 0) s_loss[wid] = ws;
        __syncthreads();
        if (tid == 0) {
            float s = 0.f;
            for (int w = 0; w < NW; ++w) s += s_loss[w];
            loss_partial[gtile] = s;
        }
    }

    const uint32_t rs = ranges[2ull * gtile], re = ranges[2ull * gtile + 1];
    const uint32_t count = re > rs ? re - rs : 0u;
    uint32_t done = 0;
    while (done < count) {
        const int nb = (int)min<uint32_t>(B, count - done);
        __syncthreads();
        if (tid < nb) load_tet<TS, B, true>(L, tid, sorted_vals[rs + done + tid], rec, idx, col_v, vw, tc);
        for (int i = tid; i < B * kTerms; i += NT) (&s_acc[0][0])[i] = 0.0f;
        __syncthreads();
        batch_layout<TS, B, true, false, CAP, NT>(L, nb);
        const uint32_t nbp = L.nbp;
        const uint32_t npairs = L.tpref[nbp];
        // phase A (pair-parallel): intersection + integral
        for (uint32_t i = tid; i < npairs; i += NT) {
            int k, r, c;
            pair_locate<TS, B, true>(L, i, nbp, k, r, c);
            const uint32_t gx = tc.tile_x * TS + c, gy = tc.tile_y * TS + r;
            uint32_t meta = 0;
            float od = 0.f, cp0 = 0.f, cp1 = 0.f, cp2 = 0.f, t_min = 0.f, t_max = 0.f;
            if (gx < vw.width && gy < vw.height) {
                const int p = r * TS + c;
                V3 rd = mk3(s_ray[0][p], s_ray[1][p], s_ray[2][p]);
                int fmin, fmax;
                if (intersect<TS, B, true>(L, k, rd, 0.0f, t_min, t_max, fmin, fmax)) {
                    // the reference's adjoint defaults both faces to 0 (backward_tiled_compute.wgsl:365-366)
                    if (fmin == 4) fmin = 0;
                    V3 g = mk3(L.g[0][k], L.g[1][k], L.g[2][k]);
                    float dc_dt = dot3(g, rd);
                    float b0 = L.bc[0][k], b1 = L.bc[1][k], b2 = L.bc[2][k];
                    float a0 = dc_dt * t_min, a1 = dc_dt * t_max;
                    float cs0 = fmaxf(b0 + a0, 0.f), cs1 = fmaxf(b1 + a0, 0.f), cs2 = fmaxf(b2 + a0, 0.f);
                    float ce0 = fmaxf(b0 + a1, 0.f), ce1 = fmaxf(b1 + a1, 0.f), ce2 = fmaxf(b2 + a1, 0.f);
                    float dist = t_max - t_min;
                    od = clampf(L.dens[k] * dist, 1e-8f, 88.0f);
                    float alpha_t = safe_exp(-od);
                    float ph = phi_f(od);
                    float w0 = ph - alpha_t, w1 = 1.0f - ph;
                    cp0 = safe_clip(ce0 * w0 + cs0 * w1, kMinVal, kMaxVal);
                    cp1 = safe_clip(ce1 * w0 + cs1 * w1, kMinVal, kMaxVal);
                    cp2 = safe_clip(ce2 * w0 + cs2 * w1, kMinVal, kMaxVal);
                    meta = (uint32_t)fmin | ((uint32_t)fmax << 2) | (1u << 4);
                }
            }
            s_r[0][i] = od;
            s_r[1][i] = cp0;
            s_r[2][i] = cp1;
            s_r[3][i] = cp2;
            s_r[4][i] = t_min;
            s_r[5][i] = t_max;
            s_meta[i] = meta;
        }
        __syncthreads();
        // phase B (pixel-sequential, back to front): undo and the d(log T) chain
        // (backward_tiled_compute.wgsl:419-451)
        if (active) {
            const float dc0 = s_dc[0][tid], dc1 = s_dc[1][tid], dc2 = s_dc[2][tid];
            for (uint32_t k = 0; k < nbp; ++k) {
                const uint32_t sp = L.span[k * SP + row];
                const int xl = span_xl(sp);
                if (colp < xl || colp > span_xr(sp)) continue;
                const uint32_t slot = L.tpref[k] + L.rpref[k * SP + row] + (uint32_t)(colp - xl);
                const uint32_t meta = s_meta[slot];
                if (!(meta & 16u)) continue;
                const float od = s_r[0][slot];
                const float prev_log_t = fminf(log_t + od, 0.0f);
                const float t_prev = safe_exp(prev_log_t);
                log_t = prev_log_t;
                if (t_prev < 1e-6f) continue;
                const float cp0 = s_r[1][slot], cp1 = s_r[2][slot], cp2 = s_r[3][slot];
                s_r[0][slot] = t_prev;
                s_r[1][slot] = d_log_t;
                s_meta[slot] = meta | 32u;
                d_log_t = safe_clip(d_log_t + ((dc0 * cp0 + dc1 * cp1) + dc2 * cp2) * t_prev, kMinVal, kMaxVal);
            }
        }
        __syncthreads();
        // phase C (pair-parallel): gradient terms, segmented reduction by tet
        // (backward_tiled_compute.wgsl:446-543)
        for (uint32_t base = (uint32_t)wid * 64; base < npairs; base += NT) {
            const uint32_t i = base + lane;
            float acc[kTerms];
#pragma unroll
            for (int q = 0; q < kTerms; ++q) acc[q] = 0.0f;
            int key = 0x7fffffff;
            if (i < npairs) {
                int k, r, c;
                pair_locate<TS, B, true>(L, i, nbp, k, r, c);
                key = k;
                const uint32_t meta = s_meta[i];
                if (meta & 32u) {
                    const int p = r * TS + c;
                    const V3 rd = mk3(s_ray[0][p], s_ray[1][p], s_ray[2][p]);
                    const float dc0 = s_dc[0][p], dc1 = s_dc[1][p], dc2 = s_dc[2][p];
                    const float t_prev = s_r[0][i], d_log_t_b = s_r[1][i];
                    const float t_min = s_r[4][i], t_max = s_r[5][i];
                    const int min_face = (int)(meta & 3u), max_face = (int)((meta >> 2) & 3u);
                    const float dens = L.dens[k];
                    V3 g = mk3(L.g[0][k], L.g[1][k], L.g[2][k]);
                    float dc_dt = dot3(g, rd);
                    float b0 = L.bc[0][k], b1 = L.bc[1][k], b2 = L.bc[2][k];
                    float a0 = dc_dt * t_min, a1 = dc_dt * t_max;
                    float csr0 = b0 + a0, csr1 = b1 + a0, csr2 = b2 + a0;
                    float cer0 = b0 + a1, cer1 = b1 + a1, cer2 = b2 + a1;
                    float cs0 = fmaxf(csr0, 0.f), cs1 = fmaxf(csr1, 0.f), cs2 = fmaxf(csr2, 0.f);
                    float ce0 = fmaxf(cer0, 0.f), ce1 = fmaxf(cer1, 0.f), ce2 = fmaxf(cer2, 0.f);
                    float dist = t_max - t_min;
                    float od = clampf(dens * dist, 1e-8f, 88.0f);
                    float alpha_t = safe_exp(-od);
                    float ph = phi_f(od);
                    float w0 = ph - alpha_t, w1 = 1.0f - ph;
                    float dcp0 = safe_clip(dc0 * t_prev, kMinVal, kMaxVal);
                    float dcp1 = safe_clip(dc1 * t_prev, kMinVal, kMaxVal);
                    float dcp2 = safe_clip(dc2 * t_prev, kMinVal, kMaxVal);
                    float d_od_state = -d_log_t_b;
                    float dphi = dphi_dx_f(od);
                    float dw0 = dphi + alpha_t, dw1 = -dphi;
                    float d_od_int = (dcp0 * (ce0 * dw0 + cs0 * dw1) + dcp1 * (ce1 * dw0 + cs1 * dw1)) +
                                     dcp2 * (ce2 * dw0 + cs2 * dw1);
                    float d_od = safe_clip(d_od_state + d_od_int, kMinVal, kMaxVal);
                    float od_raw = dens * dist;
                    bool act = (od_raw >= 1e-8f) && (od_raw <= 88.0f);
                    float d_od_raw = act ? d_od : 0.0f;
                    float d_dens_l = safe_clip(d_od_raw * dist, kMinVal, kMaxVal);
                    float d_dist = safe_clip(d_od_raw * dens, kMinVal, kMaxVal);
                    float d_tmin = -d_dist, d_tmax = d_dist;
                    float dcs0 = csr0 > 0.f ? dcp0 * w1 : 0.f, dcs1 = csr1 > 0.f ? dcp1 * w1 : 0.f;
                    float dcs2 = csr2 > 0.f ? dcp2 * w1 : 0.f;
                    float dce0 = cer0 > 0.f ? dcp0 * w0 : 0.f, dce1 = cer1 > 0.f ? dcp1 * w0 : 0.f;
                    float dce2 = cer2 > 0.f ? dcp2 * w0 : 0.f;
                    float dbc0 = dcs0 + dce0, dbc1 = dcs1 + dce1, dbc2 = dcs2 + dce2;
                    float sum_s = (dcs0 + dcs1) + dcs2, sum_e = (dce0 + dce1) + dce2;
                    float d_dcdt = sum_s * t_min + sum_e * t_max;
                    d_tmin = d_tmin + sum_s * dc_dt;
                    d_tmax = d_tmax + sum_e * dc_dt;
                    float dbo = (dbc0 + dbc1) + dbc2;
                    V3 v0 = mk3(L.v[0][0][k], L.v[0][1][k], L.v[0][2][k]);
                    V3 dg = add3(scl3(sub3(cam, v0), dbo), scl3(rd, d_dcdt));
                    V3 dv[4] = {mk3(0.f, 0.f, 0.f), mk3(0.f, 0.f, 0.f), mk3(0.f, 0.f, 0.f), mk3(0.f, 0.f, 0.f)};
                    const uint32_t flips = L.flip[k];
#pragma unroll
                    for (int side = 0; side < 2; ++side) {
                        int fi = side == 0 ? min_face : max_face;
                        float tt = side == 0 ? t_min : t_max;
                        float dt = side == 0 ? d_tmin : d_tmax;
                        int a, b, cc, o;
                        face_idx(fi, a, b, cc, o);
                        V3 va = mk3(L.v[a][0][k], L.v[a][1][k], L.v[a][2][k]);
                        V3 vb = mk3(L.v[b][0][k], L.v[b][1][k], L.v[b][2][k]);
                        V3 vc = mk3(L.v[cc][0][k], L.v[cc][1][k], L.v[cc][2][k]);
                        V3 nn = mk3(L.n[fi][0][k], L.n[fi][1][k], L.n[fi][2][k]);
                        if ((flips >> fi) & 1u) nn = neg3(nn);  // raw (unflipped) normal
                        float den = dot3(nn, rd);
                        float inv_den = safe_div(1.0f, den);
                        V3 hit = add3(cam, scl3(rd, tt));
                        V3 ah = sub3(va, hit);
                        V3 dta = scl3(add3(cross3(ah, sub3(vb, vc)), nn), inv_den);
                        V3 dtb = scl3(cross3(ah, sub3(vc, va)), inv_den);
                        V3 dtc = scl3(cross3(ah, sub3(va, vb)), inv_den);
#pragma unroll
                        for (int s = 0; s < 4; ++s) {
                            if (s == a) dv[s] = add3(dv[s], safe_clip3(scl3(dta, dt), kMinVal, kMaxVal));
                            if (s == b) dv[s] = add3(dv[s], safe_clip3(scl3(dtb, dt), kMinVal, kMaxVal));
                            if (s == cc) dv[s] = add3(dv[s], safe_clip3(scl3(dtc, dt), kMinVal, kMaxVal));
                        }
                    }
                    dv[0] = add3(dv[0], scl3(neg3(g), dbo));
                    acc[0] = safe_clip(d_dens_l, kMinVal, kMaxVal);
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        V3 c3 = safe_clip3(dv[s], kMinVal, kMaxVal);
                        acc[1 + 3 * s] = c3.x;
                        acc[2 + 3 * s] = c3.y;
                        acc[3 + 3 * s] = c3.z;
                    }
                    V3 dgc = safe_clip3(dg, kMinVal, kMaxVal);
                    acc[13] = dgc.x;
                    acc[14] = dgc.y;
                    acc[15] = dgc.z;
                    V3 dbcc = safe_clip3(mk3(dbc0, dbc1, dbc2), kMinVal, kMaxVal);
                    acc[16] = dbcc.x;
                    acc[17] = dbcc.y;
                    acc[18] = dbcc.z;
                }
            }
            // segmented sum over consecutive pairs of the same tet
            const SegMasks sm = seg_masks(key);
            const int next_key = __shfl_down(key, 1, 64);
            const bool tail = key != 0x7fffffff && (lane == 63 || next_key != key);
#pragma unroll
            for (int q = 0; q < kTerms; ++q) {
                float s = seg_scan(acc[q], sm);
                if (tail && s != 0.0f) atomicAdd(&s_acc[key][q], s);
            }
        }
        __syncthreads();
        // flush: one atomic per non-zero term per (tile, tet)
        for (int it = tid; it < (int)nbp * kTerms; it += NT) {
            const int k = it / kTerms, q = it % kTerms;
            const float s = s_acc[k][q];
            if (s == 0.0f) continue;
            const uint32_t t = L.tid[k];
            float* dst;
            if (q == 0) dst = d_dens + t;
            else if (q <= 12) dst = d_verts + 3ull * L.vi[(q - 1) / 3][k] + (q - 1) % 3;
            else if (q <= 15) dst = d_grads + 3ull * t + (q - 13);
            else dst = dcol_v + 3ull * t + (q - 16);
            unsafeAtomicAdd(dst, s);
        }
        done += nbp;
    }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
template <int TS>
static void fwd_launch(const FwdArgs& a, hipStream_t s) {
    constexpr int NT = FwdCfg<TS>::NT;
    if (a.aux)
        hipLaunchKernelGGL((k_forward<TS, true>), dim3(a.num_gtiles), dim3(NT), 0, s, a.rec, a.colors, a.color_view_stride,
                           a.sorted_vals, a.ranges, a.views, a.tiles_per_view, a.rgba, a.aux);
    else
        hipLaunchKernelGGL((k_forward<TS, false>), dim3(a.num_gtiles), dim3(NT), 0, s, a.rec, a.colors, a.color_view_stride,
                           a.sorted_vals, a.ranges, a.views, a.tiles_per_view, a.rgba, a.aux);
}

bool launch_forward(const FwdArgs& a, uint32_t tile_size, hipStream_t s) {
    switch (tile_size) {
        case 8: fwd_launch<8>(a, s); return true;
        case 12: fwd_launch<12>(a, s); return true;
        case 16: fwd_launch<16>(a, s); return true;
        default: return false;
    }
}

template <int TS>
static void bwd_launch(const BwdArgs& a, hipStream_t s) {
    constexpr int NT = BwdCfg<TS>::NT;
    hipLaunchKernelGGL((k_backward<TS>), dim3(a.num_gtiles), dim3(NT), 0, s, a.rec, a.idx, a.colors, a.color_view_stride,
                       a.sorted_vals, a.ranges, a.views, a.tiles_per_view, a.rgba, a.dl_img, a.gt, a.loss_type,
                       a.loss_partial, a.d_verts, a.d_dens, a.d_grads, a.d_colors, a.dcolor_view_stride);
}

bool launch_backward(const BwdArgs& a, uint32_t tile_size, hipStream_t s) {
    switch (tile_size) {
        case 8: bwd_launch<8>(a, s); return true;
        case 12: bwd_launch<12>(a, s); return true;
        case 16: bwd_launch<16>(a, s); return true;
        default: return false;
    }
}

}  // namespace rmesh
