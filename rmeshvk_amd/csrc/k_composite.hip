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

#include <type_traits>

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

// Row span packing: tiles up to 16 pixels wide keep a span in a byte
// (xl | xr << 4), wider tiles (32) in 16 bits (xl | xr << 8); the empty span
// has xl = all ones > xr = 0.
template <int TS>
struct SpanFmt {
    using T = typename std::conditional<(TS <= 16), uint8_t, uint16_t>::type;
    static constexpr uint32_t SH = TS <= 16 ? 4u : 8u;
    static constexpr uint32_t MASK = (1u << SH) - 1u;
    static constexpr uint32_t EMPTY = MASK;
};
template <int TS>
__device__ __forceinline__ int32_t span_xl(uint32_t sp) { return (int32_t)(sp & SpanFmt<TS>::MASK); }
template <int TS>
__device__ __forceinline__ int32_t span_xr(uint32_t sp) { return (int32_t)(sp >> SpanFmt<TS>::SH); }

// Scanline spans of one projected tet (rasterize_compute.wgsl:170-275; FORWARD
// adds the vertex-in-row rule, the adjoint omits it:
// backward_tiled_compute.wgsl:242-299), packed per SpanFmt.
//
// The reference tests each of the six edges, min(yi, yj) <= yc < max(yi, yj),
// and takes the min / max of the crossing edges' x.  With the vertices sorted
// by y (y0 <= y1 <= y2 <= y3) that test selects exactly the edges
//   {01, 02, 03} for yc in [y0, y1), {02, 03, 12, 13} in [y1, y2),
//   {03, 13, 23} in [y2, y3), and none elsewhere,
// so a row's extent is a min3 / max3 over a fixed set, chosen by two compares.
// An edge's x is formed from its lower vertex with a 1-ulp reciprocal slope;
// the span is widened by 0.001 px on both sides, so this rounding can only
// disagree with the reference's about a pixel centre within ~1e-6 px of the
// widened boundary, 0.001 px outside the silhouette, whose ray misses the tet
// (such a pixel is intersection-tested or not; it never contributes).
struct TetRows {
    float y[4], x[4];  // projected vertices sorted by y
    float s[6];        // dx/dy of the sorted edges 01 02 03 12 13 23

    __device__ __forceinline__ void cswap(int a, int b) {
        const bool sw = y[b] < y[a];
        const float ya = y[a], xa = x[a];
        y[a] = sw ? y[b] : ya;
        x[a] = sw ? x[b] : xa;
        y[b] = sw ? ya : y[b];
        x[b] = sw ? xa : x[b];
    }
    __device__ __forceinline__ float slope(int a, int b) const {
        return (x[b] - x[a]) * __builtin_amdgcn_rcpf(y[b] - y[a]);
    }
    __device__ __forceinline__ void init(const float* px_, const float* py_) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            y[k] = py_[k];
            x[k] = px_[k];
        }
        cswap(0, 1);
        cswap(2, 3);
        cswap(0, 2);
        cswap(1, 3);
        cswap(1, 2);
        s[0] = slope(0, 1);
        s[1] = slope(0, 2);
        s[2] = slope(0, 3);
        s[3] = slope(1, 2);
        s[4] = slope(1, 3);
        s[5] = slope(2, 3);
    }
    template <bool FORWARD, int TS>
    __device__ __forceinline__ uint32_t span(uint32_t row) const {
        const float yc = (float)row + 0.5f;
        const float d0 = yc - y[0], d1 = yc - y[1], d2 = yc - y[2];
        const float x01 = x[0] + d0 * s[0], x02 = x[0] + d0 * s[1], x03 = x[0] + d0 * s[2];
        const float x12 = x[1] + d1 * s[3], x13 = x[1] + d1 * s[4], x23 = x[2] + d2 * s[5];
        const bool lo = yc < y[1], mid = yc < y[2];
        // edge 03 crosses every band; the other two operands per band:
        // lo {01, 02}, mid {02 with 12, 13}, hi {13, 23}
        const float pl = lo ? x01 : (mid ? fminf(x02, x12) : x13), ql = lo ? x02 : (mid ? x13 : x23);
        const float pr = lo ? x01 : (mid ? fmaxf(x02, x12) : x13);
        float xl_f = fminf(fminf(x03, pl), ql);
        float xr_f = fmaxf(fmaxf(x03, pr), ql);
        if (!(y[0] <= yc && yc < y[3])) {
            xl_f = 1e10f;
            xr_f = -1e10f;
        }
        if (FORWARD) {
            // vertices inside the row's band [row, row + 1)
            const float r0 = (float)row, r1 = r0 + 1.0f;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (y[k] >= r0 && y[k] < r1) {
                    xl_f = fminf(xl_f, x[k]);
                    xr_f = fmaxf(xr_f, x[k]);
                }
            }
        }
        if (xl_f <= xr_f) {
            // clamped in float: in [0, TS - 1] when the span is not empty
            const float eps = 0.001f;
            const float l = fmaxf(ceilf(xl_f - 0.5f - eps), 0.0f);
            const float r = fminf(floorf(xr_f - 0.5f + eps), (float)(TS - 1));
            if (l <= r) return (uint32_t)l | ((uint32_t)r << SpanFmt<TS>::SH);
        }
        return SpanFmt<TS>::EMPTY;
    }
};

// LDS image of one batch of tets.  Row arrays use an odd stride (TS + 1) so
// the pixel-sequential phase (fixed tet, lanes on different rows) is free of
// bank conflicts.
template <int TS, int B, bool BWD>
struct BatchLDS {
    static constexpr int SP = TS + 1;
    // per tet, 7 float4 (112 B: lanes reading different tets spread over the
    // banks): 4 x (inward face normal, n . (va - cam)), (colour + g . (cam - v0),
    // density), (g, 0), (v0, 0)
    static constexpr int TR = 7;
    float4 tr[B * TR];
    float pxy[8][B];        // tile-local projected vertices (x0..x3, y0..y3)
    uint32_t behind[B];
    uint32_t rows[B];       // tile rows the tet's projected y-range touches: r0 | r1 << 8 (r0 > r1: none)
    typename SpanFmt<TS>::T span[B * SP];  // [tet][row] span (SpanFmt)
    uint16_t rpref[B * SP];  // [tet][row] record base of the row - its first column (mod 2^16)
    uint32_t wsum[(TS * TS + 63) / 64];  // layout scan carry per wave (one wave per 64 pixels of the tile)
    uint32_t nbp;           // chunk count << 8 | tets processed this batch (capacity cut)
    // adjoint only: per face (grad lambda_a, lambda_a(v0), grad lambda_b,
    // lambda_b(v0)) in face order, then per face the sign s of the stored
    // inward normal n = s * raw normal; 36 floats per tet so that lanes reading
    // different tets start on different banks; tet id
    static constexpr int FS = 4 * 8 + 4;
    alignas(16) float face[BWD ? B * FS : 1];
    uint32_t tid[BWD ? B : 1];
};

// Global loads of one tet (issued one batch ahead, consumed by store_tet).
struct TetFetch {
    float4 r[4];
    float4 col;
    uint32_t t;
};

template <bool BWD>
__device__ __forceinline__ void fetch_tet(TetFetch& f, uint32_t t, const float4* __restrict__ rec,
                                          const uint32_t* __restrict__ idx, const float4* __restrict__ colors) {
    f.t = t;
#pragma unroll
    for (int i = 0; i < 4; ++i) f.r[i] = rec[4ull * t + i];
    f.col = colors[t];
}

// LDS staging of a batch's tet data, filled by direct global -> LDS loads
// (global_load_lds_dwordx4: lane l's 16 bytes land at base + 16 l), so the
// next batch's gather is in flight during this batch without holding VGPRs.
template <int B>
struct Staging {
    float4 rec[4][B];
    float4 col[B];
    uint32_t t[B];
};

// Wave 0 only (lane == tet slot): start loading tet t for the lanes with `on`.
template <bool BWD, int B>
__device__ __forceinline__ void stage_async(Staging<B>& S, uint32_t t, bool on, const float4* __restrict__ rec,
                                            const uint32_t* __restrict__ idx, const float4* __restrict__ colors) {
    typedef __attribute__((address_space(3))) void* lds_ptr;
    if (on) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
            __builtin_amdgcn_global_load_lds((const void*)(rec + 4ull * t + i), (lds_ptr)&S.rec[i][0], 16, 0, 0);
        __builtin_amdgcn_global_load_lds((const void*)(colors + t), (lds_ptr)&S.col[0], 16, 0, 0);
        S.t[__lane_id()] = t;
    }
}

template <bool BWD, int B>
__device__ __forceinline__ void staged_tet(const Staging<B>& S, int k, TetFetch& f) {
#pragma unroll
    for (int i = 0; i < 4; ++i) f.r[i] = S.rec[i][k];
    f.col = S.col[k];
    f.t = S.t[k];
}

// Face fi of tet k: inward plane (and, for the adjoint, the face table entry).
template <int TS, int B, bool BWD>
__device__ __forceinline__ void store_face(BatchLDS<TS, B, BWD>& L, int k, int fi, V3 va, V3 vb, V3 vc, V3 vo, V3 v0,
                                           V3 cam) {
    V3 raw = cross3(sub3(vc, va), sub3(vb, va));
    V3 nn = raw;
    if (dot3(nn, sub3(vo, va)) < 0.0f) nn = neg3(nn);
    L.tr[k * L.TR + fi] = make_float4(nn.x, nn.y, nn.z, dot3(nn, sub3(va, cam)));
    if (BWD) {
        // barycentric weights of a point h of the face plane, affine in h:
        // lambda_a(h) = lambda_a(v0) + grad_a . (h - v0), grad_a = ((c - b) x raw) / |raw|^2
        // (and cyclic: (a - c) for b); the three gradients sum to zero, so
        // lambda_c = 1 - lambda_a - lambda_b everywhere
        const float inv = 1.0f / dot3(raw, raw);
        const V3 ga = scl3(cross3(sub3(vc, vb), raw), inv);
        const V3 gb = scl3(cross3(sub3(va, vc), raw), inv);
        const V3 pa = sub3(va, v0), pb = sub3(vb, v0), pc = sub3(vc, v0);
        float4* F = reinterpret_cast<float4*>(&L.face[k * L.FS + fi * 8]);
        F[0] = make_float4(ga.x, ga.y, ga.z, dot3(cross3(pc, pb), raw) * inv);
        F[1] = make_float4(gb.x, gb.y, gb.z, dot3(cross3(pa, pc), raw) * inv);
        L.face[k * L.FS + 32 + fi] = nn.x == raw.x && nn.y == raw.y && nn.z == raw.z ? 1.0f : -1.0f;
    }
}

// Per-tet data of tet k: base colour at the camera, density, g, v0, the
// tile-local projected vertices, vertex ids and tet id.
template <int TS, int B, bool BWD>
__device__ __forceinline__ void store_rest(BatchLDS<TS, B, BWD>& L, int k, const TetFetch& f, const V3* v,
                                           const ViewParams& vw, const TileCtx& tc) {
    const V3 g = mk3(f.r[1].w, f.r[2].w, f.r[3].w);
    const V3 cam = mk3(vw.cam[0], vw.cam[1], vw.cam[2]);
    float bo = dot3(g, sub3(cam, v[0]));
    L.tr[k * L.TR + 4] = make_float4(f.col.x + bo, f.col.y + bo, f.col.z + bo, f.r[0].w);
    L.tr[k * L.TR + 5] = make_float4(g.x, g.y, g.z, 0.0f);
    L.tr[k * L.TR + 6] = make_float4(v[0].x, v[0].y, v[0].z, 0.0f);
    float clip[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i) mul_vp(vw.vp, v[i], clip[i]);
    bool behind = (clip[0][3] <= 0.0f) || (clip[1][3] <= 0.0f) || (clip[2][3] <= 0.0f) || (clip[3][3] <= 0.0f);
    L.behind[k] = behind ? 1u : 0u;
    uint32_t rr = 0xFFu;
    if (!behind) {
        float py_[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            L.pxy[i][k] = ((clip[i][0] / clip[i][3] + 1.0f) * 0.5f) * vw.Wf - tc.tile_ox;
            py_[i] = ((1.0f - clip[i][1] / clip[i][3]) * 0.5f) * vw.Hf - tc.tile_oy;
            L.pxy[4 + i][k] = py_[i];
        }
        // rows whose pixel band [row, row + 1] meets [ymin, ymax]: the only rows
        // where an edge can cross or a vertex can lie (row + 1 >= ymin, row <= ymax)
        const float ymin = fminf(fminf(py_[0], py_[1]), fminf(py_[2], py_[3]));
        const float ymax = fmaxf(fmaxf(py_[0], py_[1]), fmaxf(py_[2], py_[3]));
        const float r0 = fmaxf(ceilf(ymin) - 1.0f, 0.0f);
        const float r1 = fminf(floorf(ymax), (float)(TS - 1));
        if (r0 <= r1) rr = (uint32_t)r0 | ((uint32_t)r1 << 8);
    }
    L.rows[k] = rr;
    if (BWD) {
        L.tid[k] = f.t;
    }
}

// One lane stages a whole tet (records in registers).
template <int TS, int B, bool BWD>
__device__ __forceinline__ void store_tet(BatchLDS<TS, B, BWD>& L, int k, const TetFetch& f, const ViewParams& vw,
                                          const TileCtx& tc) {
    V3 v[4] = {mk3(f.r[0].x, f.r[0].y, f.r[0].z), mk3(f.r[1].x, f.r[1].y, f.r[1].z), mk3(f.r[2].x, f.r[2].y, f.r[2].z),
               mk3(f.r[3].x, f.r[3].y, f.r[3].z)};
    V3 cam = mk3(vw.cam[0], vw.cam[1], vw.cam[2]);
#pragma unroll
    for (int fi = 0; fi < 4; ++fi) {
        int a, b, c, o;
        face_idx(fi, a, b, c, o);
        store_face(L, k, fi, v[a], v[b], v[c], v[o], v[0], cam);
    }
    store_rest(L, k, f, v, vw, tc);
}

// Quad prefetch: lane j of the four lanes 4k..4k+3 of tet slot k holds the
// tet's vertex record j (vertex j; w = density for j = 0, else g component
// j - 1), its colour and the tet id.
struct QuadFetch {
    float4 r;
    float4 col;
    uint32_t t;
};

template <bool BWD>
__device__ __forceinline__ void fetch_quad(QuadFetch& f, uint32_t t, int j, const float4* __restrict__ rec,
                                           const uint32_t* __restrict__ idx, const float4* __restrict__ colors) {
    f.t = t;
    f.r = rec[4ull * t + j];
    f.col = colors[t];
}

// quad_perm: lane j of each quad reads lane P_j of the quad (CTRL = P0 | P1 << 2 | P2 << 4 | P3 << 6).
template <int CTRL>
__device__ __forceinline__ float qperm(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ V3 qperm3(float4 v) {
    return mk3(qperm<CTRL>(v.x), qperm<CTRL>(v.y), qperm<CTRL>(v.z));
}

// Four lanes per tet store the batch from quad-prefetched records: lane j
// builds face j (its vertices a, b, c, o come from the quad by one quad_perm
// each: FACES = {0,2,1,3}, {1,2,3,0}, {0,3,2,1}, {3,0,1,2}) and projects
// vertex j; lanes 0-2 write the per-tet rows.  Every lane of the block calls
// it (DPP reads the whole quad); `on` marks the lanes of the batch's tets.
template <int TS, int B, bool BWD>
__device__ __forceinline__ void store_quad(BatchLDS<TS, B, BWD>& L, const QuadFetch& f, int k, int j, bool on,
                                           const ViewParams& vw, const TileCtx& tc) {
    const V3 va = qperm3<0xC4>(f.r);  // [0,1,0,3]
    const V3 vb = qperm3<0x3A>(f.r);  // [2,2,3,0]
    const V3 vc = qperm3<0x6D>(f.r);  // [1,3,2,1]
    const V3 vo = qperm3<0x93>(f.r);  // [3,0,1,2]
    const V3 v0 = qperm3<0x00>(f.r);
    const float g1 = qperm<0x55>(f.r.w), g2 = qperm<0xAA>(f.r.w), g3 = qperm<0xFF>(f.r.w);
    // vertex j's projection; the quad's behind flag and y-range
    const V3 vj = mk3(f.r.x, f.r.y, f.r.z);
    float clip[4];
    mul_vp(vw.vp, vj, clip);
    const float pxj = ((clip[0] / clip[3] + 1.0f) * 0.5f) * vw.Wf - tc.tile_ox;
    const float pyj = ((1.0f - clip[1] / clip[3]) * 0.5f) * vw.Hf - tc.tile_oy;
    float bh = clip[3] <= 0.0f ? 1.0f : 0.0f;
    bh = fmaxf(bh, qperm<0xB1>(bh));  // [1,0,3,2]
    bh = fmaxf(bh, qperm<0x4E>(bh));  // [2,3,0,1]
    float ymin = fminf(pyj, qperm<0xB1>(pyj)), ymax = fmaxf(pyj, qperm<0xB1>(pyj));
    ymin = fminf(ymin, qperm<0x4E>(ymin));
    ymax = fmaxf(ymax, qperm<0x4E>(ymax));
    if (!on) return;
    const V3 cam = mk3(vw.cam[0], vw.cam[1], vw.cam[2]);
    store_face(L, k, j, va, vb, vc, vo, v0, cam);
    const bool behind = bh != 0.0f;
    if (!behind) {
        L.pxy[j][k] = pxj;
        L.pxy[4 + j][k] = pyj;
    }
    if (j == 0) {
        const V3 g = mk3(g1, g2, g3);
        const float bo = dot3(g, sub3(cam, v0));
        L.tr[k * L.TR + 4] = make_float4(f.col.x + bo, f.col.y + bo, f.col.z + bo, f.r.w);
        L.behind[k] = behind ? 1u : 0u;
        uint32_t rr = 0xFFu;
        if (!behind) {
            // rows whose pixel band [row, row + 1] meets [ymin, ymax] (as store_rest)
            const float r0 = fmaxf(ceilf(ymin) - 1.0f, 0.0f);
            const float r1 = fminf(floorf(ymax), (float)(TS - 1));
            if (r0 <= r1) rr = (uint32_t)r0 | ((uint32_t)r1 << 8);
        }
        L.rows[k] = rr;
        if (BWD) L.tid[k] = f.t;
    } else if (j == 1) {
        L.tr[k * L.TR + 5] = make_float4(g1, g2, g3, 0.0f);
    } else if (j == 2) {
        L.tr[k * L.TR + 6] = make_float4(v0.x, v0.y, v0.z, 0.0f);
    }
}

// Five lanes per tet stage the batch from the LDS staging area: lanes 0-3 a
// face each, lane 4 the per-tet data.
template <int TS, int B, bool BWD, int NT>
__device__ __forceinline__ void store_batch_staged(BatchLDS<TS, B, BWD>& L, const Staging<B>& S, int nb,
                                                   const ViewParams& vw, const TileCtx& tc) {
    const V3 cam = mk3(vw.cam[0], vw.cam[1], vw.cam[2]);
    for (int it = threadIdx.x; it < 5 * nb; it += NT) {
        const int k = it / 5, part = it % 5;
        auto vtx = [&](int i) {
            const float4 r = S.rec[i][k];
            return mk3(r.x, r.y, r.z);
        };
        if (part < 4) {
            int a, b, c, o;
            face_idx(part, a, b, c, o);
            store_face(L, k, part, vtx(a), vtx(b), vtx(c), vtx(o), vtx(0), cam);
        } else {
            TetFetch f;
            staged_tet<BWD>(S, k, f);
            const V3 v[4] = {vtx(0), vtx(1), vtx(2), vtx(3)};
            store_rest(L, k, f, v, vw, tc);
        }
    }
}

// Entry i < 128 of a wave's two-register id window (entries 0-63 in w0, 64-127 in w1).
template <int B>
__device__ __forceinline__ uint32_t window_id(uint32_t w0, uint32_t w1, int i) {
    const uint32_t a = (uint32_t)__shfl((int)w0, i & 63, 64);
    if constexpr (B > 32) {
        const uint32_t b = (uint32_t)__shfl((int)w1, i & 63, 64);
        return i < 64 ? a : b;
    }
    return a;
}

template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t dpp_u(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, 0xF, ROWS == 0xF);
}

// Inclusive sum over the wave: row_shr 1/2/4/8 within 16-lane rows, then
// row_bcast:15 (rows 1, 3) and row_bcast:31 (rows 2, 3).
__device__ __forceinline__ uint32_t wave_scan_dpp(uint32_t v) {
    v += dpp_u<0x111, 0xF>(v);
    v += dpp_u<0x112, 0xF>(v);
    v += dpp_u<0x114, 0xF>(v);
    v += dpp_u<0x118, 0xF>(v);
    v += dpp_u<0x142, 0xA>(v);
    v += dpp_u<0x143, 0xC>(v);
    return v;
}

// Value of the last lane of each aligned group of G lanes (quad_perm; G = 8: bpermute).
template <int G>
__device__ __forceinline__ uint32_t group_last(uint32_t v) {
    if constexpr (G == 8) return (uint32_t)__shfl((int)v, (int)(__lane_id() | 7u), 64);
    else if constexpr (G == 4) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xFF, 0xF, 0xF, false);
    else if constexpr (G == 2) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xF5, 0xF, 0xF, false);
    else return v;
}

// Chunk table entry: up to R consecutive covered pixels of one tet (in row,
// col order), packed as
//   tet k (6 bits) | row (5) | first column (5) | first record slot (11) | count (5).
// With R = 1 the table holds one 16-bit entry (tet, row, first column) per
// row segment, at the segment's first record slot, and a bit mask marks those
// slots: record j belongs to the segment starting at the last marked slot <= j.
constexpr int chunk_capacity(int CAP, int R, int B) { return CAP / R + B; }
template <int R>
using chunk_t = typename std::conditional<R == 1, uint16_t, uint32_t>::type;
template <int R>
__device__ __forceinline__ chunk_t<R> chunk_entry(uint32_t k, uint32_t row, uint32_t col, uint32_t slot0,
                                                  uint32_t cnt) {
    const uint32_t lo = k | (row << 6) | (col << 11);
    if constexpr (R == 1) return (uint16_t)lo;
    else return lo | (slot0 << 16) | (cnt << 27);
}

// Layout of a batch in one pass: spans, record slots, chunk table and the
// capacity cut (records <= CAP).  LPT lanes per tet, each over a contiguous
// block of the rows of the tet's projected y-range (rows outside it are empty
// and never read).  One wave scan of the lanes' pixel counts, packed with the
// per-tet chunk counts in the high half-word and carried across waves through
// L.wsum, gives every lane its record base and its tet's chunk base.  Leaves
//   span[tet][row], rpref[tet][row] = record base of the row - its first column,
//   tab[] = chunk table, nbp = chunk count << 8 | tets kept,
// published by the caller's barrier.  Contains one barrier; every thread of
// the block calls it.
template <int TS, int B, bool BWD, bool FORWARD_RULE, int CAP, int R, int NT>
__device__ __forceinline__ void batch_layout(BatchLDS<TS, B, BWD>& L, int nb, chunk_t<R>* __restrict__ tab,
                                             uint32_t* __restrict__ starts) {
    constexpr int SP = TS + 1;
    constexpr int LPT = NT >= 8 * B ? 8 : NT >= 4 * B ? 4 : NT >= 2 * B ? 2 : 1;
    constexpr int QMAX = (TS + LPT - 1) / LPT;
    static_assert(LPT * B <= NT && B <= 64, "one pass over the batch");
    static_assert(NT / 64 <= (int)(sizeof(L.wsum) / sizeof(L.wsum[0])), "a scan carry per wave");
    static_assert(B * TS * TS < 65536, "packed scan fields");
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    const int k = t / LPT, j = t % LPT;
    const bool valid = k < nb;
    uint32_t sp[QMAX];
    int ra = 0, nr = 0;  // this lane's first row and row count
    uint32_t s = 0;      // covered pixels in those rows
    if (valid) {
        const uint32_t rr = L.rows[k], r0 = rr & 0xFFu, r1 = rr >> 8;
        if (r0 <= r1) {
            const int q = ((int)(r1 - r0) + LPT) / LPT;
            ra = (int)r0 + j * q;
            nr = min(q, (int)r1 + 1 - ra);
        }
        if (nr > 0) {
            const float px_[4] = {L.pxy[0][k], L.pxy[1][k], L.pxy[2][k], L.pxy[3][k]};
            const float py_[4] = {L.pxy[4][k], L.pxy[5][k], L.pxy[6][k], L.pxy[7][k]};
            TetRows tr;
            tr.init(px_, py_);
#pragma unroll
            for (int i = 0; i < QMAX; ++i) {
                if (i < nr) {
                    sp[i] = tr.span<FORWARD_RULE, TS>((uint32_t)(ra + i));
                    L.span[k * SP + ra + i] = (typename SpanFmt<TS>::T)sp[i];
                    s += (uint32_t)max(span_xr<TS>(sp[i]) - span_xl<TS>(sp[i]) + 1, 0);
                }
            }
        }
    }
    if (t == 0) L.nbp = 0u;
    // pixels of the tet before this lane, and the tet's total
    uint32_t g = s;
    if constexpr (LPT >= 2) {
        const uint32_t u = dpp_u<0x111, 0xF>(g);  // row_shr:1
        if (j >= 1) g += u;
    }
    if constexpr (LPT >= 4) {
        const uint32_t u = dpp_u<0x112, 0xF>(g);  // row_shr:2
        if (j >= 2) g += u;
    }
    if constexpr (LPT >= 8) {
        const uint32_t u = dpp_u<0x114, 0xF>(g);  // row_shr:4
        if (j >= 4) g += u;
    }
    const uint32_t tot = group_last<LPT>(g);
    const uint32_t pl = g - s;
    const uint32_t cc = (tot + R - 1) / R;
    const bool last = j == LPT - 1;
    uint32_t inc = wave_scan_dpp(s | (valid && last ? cc << 16 : 0u));
    if (lane == 63) L.wsum[wid] = inc;
    __syncthreads();
    for (int w = 0; w < wid; ++w) inc += L.wsum[w];
    const uint32_t texc = (inc & 0xFFFFu) - s - pl;               // records of the tets before k
    const uint32_t cexc = (inc >> 16) - (valid && last ? cc : 0u);  // chunks of the tets before k
    const bool kept = valid && texc + tot <= (uint32_t)CAP;
    if (kept) {
        uint32_t base = pl;  // tet-local record index of the row's first pixel
#pragma unroll
        for (int i = 0; i < QMAX; ++i) {
            if (i < nr) {
                const uint32_t row = (uint32_t)(ra + i);
                const int xl = span_xl<TS>(sp[i]), w = span_xr<TS>(sp[i]) - xl + 1;
                if (w > 0) {
                    if constexpr (R == 1) {
                        const uint32_t b = texc + base;  // the row segment's first record slot
                        tab[b] = chunk_entry<1>((uint32_t)k, row, (uint32_t)xl, 0u, 0u);
                        atomicOr(&starts[b >> 5], 1u << (b & 31u));
                    } else {
                        const uint32_t c0 = (base + R - 1) / R, c1 = (base + (uint32_t)w - 1) / R;
                        for (uint32_t c = c0; c <= c1; ++c) {
                            const uint32_t l = c * R;
                            tab[cexc + c] = chunk_entry<R>((uint32_t)k, row, (uint32_t)(xl + (int)(l - base)),
                                                           texc + l, min((uint32_t)R, tot - l));
                        }
                    }
                    L.rpref[k * SP + row] = (uint16_t)(texc + base - (uint32_t)xl);
                    base += (uint32_t)w;
                }
            }
        }
    }
    // kept tets are a prefix of the batch: the highest one carries the totals
    const uint64_t m = __ballot(kept && last);
    if (m && lane == 63 - __clzll((long long)m)) atomicMax(&L.nbp, ((cexc + cc) << 8) | (uint32_t)(k + 1));
}

// Tets to stage for the next batch.  Tets staged past the capacity cut are
// dropped and staged again with the following batch (their fetch, face
// setup and spans are redone), so after a batch that did not fit whole the
// next one stages the kept count plus half again (at least 4 more), and the
// full B again after one that fitted.  Batches only partition the list: every
// pixel still composites the same tets in the same order.
template <int B>
__device__ __forceinline__ int next_batch(uint32_t nbp, int nb, uint32_t left) {
    const uint32_t want = (int)nbp == nb ? (uint32_t)B : min((uint32_t)B, nbp + max(nbp >> 1, 4u));
    return (int)min(want, left);
}

struct ChunkRef {
    int k, row, col;
    uint32_t slot0, cnt;
};
// R = 1: record slot j.  A row segment spans at most 16 slots, so its start
// lies in j's mask word or the one before.
__device__ __forceinline__ ChunkRef record_at(const uint16_t* seg, const uint32_t* starts, uint32_t j) {
    const uint32_t wd = j >> 5;
    const uint32_t m = starts[wd] & ((2u << (j & 31u)) - 1u);
    const uint32_t s0 = m ? (wd << 5) + 31u - (uint32_t)__clz((int)m)
                          : ((wd - 1u) << 5) + 31u - (uint32_t)__clz((int)starts[wd - 1u]);
    const uint32_t e = seg[s0];
    return ChunkRef{(int)(e & 63u), (int)((e >> 6) & 31u), (int)((e >> 11) & 31u) + (int)(j - s0), j, 1u};
}
template <int R>
__device__ __forceinline__ ChunkRef chunk_at(const chunk_t<R>* tab, uint32_t j) {
    const uint32_t e = tab[j];
    if constexpr (R == 1) return ChunkRef{(int)(e & 63u), (int)((e >> 6) & 31u), (int)((e >> 11) & 31u), j, 1u};
    return ChunkRef{(int)(e & 63u), (int)((e >> 6) & 31u), (int)((e >> 11) & 31u), (e >> 16) & 2047u, e >> 27};
}

// Walks the covered pixels of one tet in (row, col) order starting at local index l.
template <int TS>
struct PixWalk {
    int row, col, xr, r1;
    template <int B, bool BWD>
    __device__ __forceinline__ void init(const BatchLDS<TS, B, BWD>& L, const ChunkRef& c) {
        row = c.row;
        col = c.col;
        xr = span_xr<TS>(L.span[c.k * (TS + 1) + c.row]);
        r1 = (int)(L.rows[c.k] >> 8);
    }
    template <int B, bool BWD>
    __device__ __forceinline__ void next(const BatchLDS<TS, B, BWD>& L, int k) {
        constexpr int SP = TS + 1;
        if (++col > xr) {
            for (int r = row + 1; r <= r1; ++r) {
                uint32_t sp = L.span[k * SP + r];
                if (span_xr<TS>(sp) >= span_xl<TS>(sp)) {
                    row = r;
                    col = span_xl<TS>(sp);
                    xr = span_xr<TS>(sp);
                    return;
                }
            }
            row = TS;
        }
    }
};

// Per-tet data held in registers while a lane walks its chunk.
struct TetRegs {
    V3 n[4];
    float num[4];
    float bc[3];
    V3 g;
    float dens;
};

template <int TS, int B, bool BWD>
__device__ __forceinline__ void tet_regs(const BatchLDS<TS, B, BWD>& L, int k, TetRegs& T) {
#pragma unroll
    for (int fi = 0; fi < 4; ++fi) {
        const float4 q = L.tr[k * L.TR + fi];
        T.n[fi] = mk3(q.x, q.y, q.z);
        T.num[fi] = q.w;
    }
    const float4 c = L.tr[k * L.TR + 4], gg = L.tr[k * L.TR + 5];
    T.bc[0] = c.x;
    T.bc[1] = c.y;
    T.bc[2] = c.z;
    T.dens = c.w;
    T.g = mk3(gg.x, gg.y, gg.z);
}

// Fused dot product for the adjoint's own terms (the slab intersection and
// the volume integral keep the unfused order of the reference's arithmetic:
// vertex gradients magnify a one-ulp change of t through grazing faces).
__device__ __forceinline__ float fdot3(V3 a, V3 b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }

// Ray–tet slab intersection against inward planes (rasterize_compute.wgsl:354-390).
__device__ __forceinline__ bool intersect(const TetRegs& T, V3 rd, float min_t, float& t_min, float& t_max, int& fmin,
                                          int& fmax) {
#pragma clang fp contract(fast)
    t_min = -3.402823e38f;
    t_max = 3.402823e38f;
    fmin = 4;
    fmax = 0;
    bool valid = true;
#pragma unroll
    for (int fi = 0; fi < 4; ++fi) {
        float den = dot3(T.n[fi], rd);
        float num = T.num[fi];
        if (min_t != 0.0f) num = num - min_t * den;
        if (fabsf(den) < 1e-20f) {
            if (num > 0.0f) valid = false;
            continue;
        }
        float t = fast_div(num, den);
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

// Volume integral of the linear colour field over [t_min, t_max]
// (rasterize_compute.wgsl:394-410): optical depth, premultiplied colour.
struct Integral {
    float od, alpha_t, w0, w1;
    float cs[3], ce[3], csr[3], cer[3];
    float cp[3];
};
// EXTRA = false: bo_extra is 0 and not added (x + 0 only differs from x in
// the sign of a zero, which no use of b sees).
template <bool EXTRA = true>
__device__ __forceinline__ void integral(const TetRegs& T, V3 rd, float t_min, float t_max, float bo_extra, float clip_lo,
                                         Integral& I) {
#pragma clang fp contract(fast)
    float dc_dt = dot3(T.g, rd);
    float a0 = dc_dt * t_min, a1 = dc_dt * t_max;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float b = EXTRA ? T.bc[c] + bo_extra : T.bc[c];
        I.csr[c] = b + a0;
        I.cer[c] = b + a1;
        I.cs[c] = relu(I.csr[c]);
        I.ce[c] = relu(I.cer[c]);
    }
    float dist = t_max - t_min;
    I.od = med_clip(T.dens * dist, 1e-8f, 88.0f);
    I.alpha_t = fast_safe_exp(-I.od);
    float ph = fast_phi(I.od, I.alpha_t);
    I.w0 = ph - I.alpha_t;
    I.w1 = 1.0f - ph;
#pragma unroll
    for (int c = 0; c < 3; ++c) I.cp[c] = med_clip(I.ce[c] * I.w0 + I.cs[c] * I.w1, clip_lo, kMaxVal);
}

// ---------------------------------------------------------------------------
// Forward
// ---------------------------------------------------------------------------
// Batch geometry of the forward: B tets per LDS batch (<= 64, one mask bit
// each), CAP (pixel, tet) records per batch, R pixels per lane in the pair
// phase, PF = fetch the next batch's records while this batch runs.
// WPE_: minimum waves per SIMD the register allocation must allow (0: the
// compiler's choice); a cap on VGPRs that trades recomputation for occupancy.
template <int TS_, int B_, int CAP_, int R_, int PF_, int WPE_ = 0>
struct FwdCfg {
    static constexpr int TS = TS_;
    static constexpr int NT = (TS * TS + 63) / 64 * 64;
    static constexpr int B = B_;
    static constexpr int CAP = CAP_;
    static constexpr int R = R_;
    // next batch's records: 0 at batch start, 1 in registers, 2 LDS DMA, 3 in
    // registers by quads (4 lanes per tet: falls back to 1 where 4 B > NT)
    static constexpr int PF = (PF_ == 3 && 4 * B_ > NT) ? 1 : PF_;
    static constexpr int WPE = WPE_ > 0 ? WPE_ : 1;
};

// CK: also store the transmittance checkpoints of the checkpointed adjoint:
// a pixel still compositing writes its log T before every tile-list position
// 32 j it reaches (coalesced: the tile's pixels write one 4 TS^2-byte slot),
// and at the end its final log T and stop position (1 + the last position it
// composited).
template <class C, bool AUX, bool CK>
__global__ __launch_bounds__(C::NT) __attribute__((amdgpu_waves_per_eu(C::WPE))) void k_forward(
    const float4* __restrict__ rec, const float4* __restrict__ colors,
    const uint32_t* __restrict__ sorted_vals, const uint32_t* __restrict__ ranges, const ViewParams* __restrict__ views,
    const uint32_t* __restrict__ tile_view, const uint32_t* __restrict__ tile_order, float* __restrict__ rgba,
    float* __restrict__ aux, float* __restrict__ ck, float2* __restrict__ fin) {
    constexpr int TS = C::TS, NT = C::NT, B = C::B, CAP = C::CAP, R = C::R, SP = TS + 1;
    constexpr int NA = AUX ? CAP : 1;
    static_assert(B <= 64, "pixel masks hold 64 tets; the layout scan is one wave");
    static_assert(CAP >= TS * TS, "one tet's pixels of a tile must fit a batch");
    __shared__ BatchLDS<TS, B, false> L;
    __shared__ float s_od[CAP];
    __shared__ float s_cp[3][CAP];
    __shared__ float s_aux[3][NA];     // t_min, t_max, w0
    __shared__ uint32_t s_ent[NA];
    __shared__ float s_ray[3][TS * TS];
    __shared__ uint32_t s_live[TS * TS];  // pixel inside the image and not yet saturated (log T >= -13.8)
    __shared__ unsigned long long s_mask[TS * TS];  // covered tets of the batch per pixel
    __shared__ chunk_t<R> s_chunk[chunk_capacity(CAP, R, B)];
    __shared__ uint32_t s_start[R == 1 ? CAP / 32 : 1];  // R = 1: record slots that start a row segment
    static_assert(CAP <= 2048 && TS <= 32 && R <= 31, "chunk table packing");
    static_assert(R != 1 || (CAP % 32 == 0 && CAP / 32 <= NT), "segment start mask");

    const uint32_t gtile = tile_order ? tile_order[blockIdx.x] : blockIdx.x;
    const uint32_t view = tile_view[gtile];
    const ViewParams& vw = views[view];
    const TileCtx tc = tile_ctx(vw, gtile);
    const float4* col_v = colors + vw.color_off;
    const int tid = threadIdx.x;
    const int row = tid / TS, colp = tid % TS;
    const uint32_t px = tc.tile_x * TS + colp, py = tc.tile_y * TS + row;
    const bool active = tid < TS * TS && px < vw.width && py < vw.height;
    const float min_t = vw.min_t;

    if (tid < TS * TS) {
        V3 rd = mk3(0.f, 0.f, 1.f);
        if (active) rd = pixel_ray(vw, px, py);
        s_ray[0][tid] = rd.x;
        s_ray[1][tid] = rd.y;
        s_ray[2][tid] = rd.z;
        s_live[tid] = active ? 1u : 0u;
    }
    float cr = 0.f, cg = 0.f, cb = 0.f, log_t = 0.f;
    float nx = 0.f, ny = 0.f, nz = 0.f, dep = 0.f, e0 = 0.f, e1 = 0.f, e2 = 0.f, e3 = 0.f;

    const uint32_t rs = ranges[2ull * gtile], re = ranges[2ull * gtile + 1];
    const uint32_t count = re > rs ? re - rs : 0u;
    uint32_t done = 0;
    int nb = (int)min<uint32_t>(B, count);
    // checkpoints: next position to store, stop position, slot base
    uint32_t next_ck = kCkGroup, stop = 0u;
    const uint64_t ck_base = ((uint64_t)(rs >> 5) + gtile - 1u) * (uint64_t)(TS * TS) + (uint64_t)tid;
    TetFetch f;
    static_assert(C::PF != 2 || B <= 64, "LDS-DMA staging: one wave, next batch inside a 128-entry id window");
    __shared__ Staging<C::PF == 2 ? B : 1> S;
    const int wid = tid / 64, lane = (int)__lane_id();
    QuadFetch qf{};
    if (C::PF == 1 && tid < nb) fetch_tet<false>(f, sorted_vals[re - 1 - tid], rec, nullptr, col_v);
    if (C::PF == 3 && tid < 4 * nb) fetch_quad<false>(qf, sorted_vals[re - 1 - (tid >> 2)], tid & 3, rec, nullptr, col_v);
    if (C::PF == 2 && wid == 0) {
        const uint32_t t = lane < nb ? sorted_vals[re - 1 - lane] : 0u;
        stage_async<false>(S, t, lane < nb, rec, nullptr, col_v);
    }
    // MERGE: a batch's LDS image is built right after the previous batch's
    // pixel-sequential phase, without a barrier between them (that phase
    // reads only the row bases and the pair records, which the image does not
    // touch; each lane clears its own pixel's mask after reading it), so one
    // barrier phase per batch goes.  Not with the aux outputs (their phase
    // reads the face planes) or LDS-DMA staging (its loads need a barrier).
    constexpr bool MERGE = !AUX && C::PF != 2;
    auto stage = [&](int nbs, uint32_t d) {
        if (C::PF == 0 && tid < nbs) fetch_tet<false>(f, sorted_vals[re - 1 - d - tid], rec, nullptr, col_v);
        if constexpr (C::PF == 3) store_quad<TS, B, false>(L, qf, tid >> 2, tid & 3, tid < 4 * nbs, vw, tc);
        else if (tid < nbs) store_tet<TS, B, false>(L, tid, f, vw, tc);
        if (tid < TS * TS) s_mask[tid] = 0ull;
        if (R == 1 && tid < CAP / 32) s_start[tid] = 0u;
    };
    if constexpr (MERGE) {
        if (count) stage(nb, 0u);
    }
    while (done < count) {
        // staged records of this batch have landed (and none is in flight at an early exit)
        if (C::PF == 2) __builtin_amdgcn_s_waitcnt(0);
        if (__syncthreads_and(!active || log_t < -13.8f)) break;
        // wave 0: tet ids of this and the next batch (back to front of the
        // sorted list), 64 per register
        uint32_t win = 0, win1 = 0;
        if (C::PF == 2 && wid == 0) {
            win = done + lane < count ? sorted_vals[re - 1 - done - lane] : 0u;
            if constexpr (B > 32) win1 = done + 64 + lane < count ? sorted_vals[re - 1 - done - 64 - lane] : 0u;
        }
        if constexpr (!MERGE) {
            if constexpr (C::PF == 2) {
                store_batch_staged<TS, B, false, NT>(L, S, nb, vw, tc);
                if (tid < TS * TS) s_mask[tid] = 0ull;
                if (R == 1 && tid < CAP / 32) s_start[tid] = 0u;
            } else {
                stage(nb, done);
            }
            __syncthreads();
        }
        batch_layout<TS, B, false, true, CAP, R, NT>(L, nb, s_chunk, s_start);
        __syncthreads();
        const uint32_t nbp = L.nbp & 0xFFu;
        const uint32_t nchunks = L.nbp >> 8;
        // prefetch the next batch's tet records while this one is processed
        const uint32_t next = done + nbp;
        const int nb_next = next_batch<B>(nbp, nb, count - next);
        if (C::PF == 1 && tid < nb_next) fetch_tet<false>(f, sorted_vals[re - 1 - next - tid], rec, nullptr, col_v);
        if (C::PF == 3 && tid < 4 * nb_next)
            fetch_quad<false>(qf, sorted_vals[re - 1 - next - (tid >> 2)], tid & 3, rec, nullptr, col_v);
        if (C::PF == 2 && wid == 0) {
            const uint32_t tn = window_id<B>(win, win1, (int)nbp + lane);
            stage_async<false>(S, tn, lane < nb_next, rec, nullptr, col_v);
        }
        // record-parallel: intersection + volume integral of one (pixel, tet)
        // record per lane (R = 1), or of a chunk of R pixels of one tet
        for (uint32_t j = tid; j < nchunks; j += NT) {
            ChunkRef c;
            if constexpr (R == 1) c = record_at(s_chunk, s_start, j);
            else c = chunk_at<R>(s_chunk, j);
            const int k = c.k;
            TetRegs T;
            tet_regs(L, k, T);
            auto one = [&](uint32_t slot, int p) {
                // saturated pixels skip the rest of the tile's tets (their compositing loop would)
                if (s_live[p]) {
                    V3 rd = mk3(s_ray[0][p], s_ray[1][p], s_ray[2][p]);
                    float t_min, t_max;
                    int fmin, fmax;
                    if (intersect(T, rd, min_t, t_min, t_max, fmin, fmax)) {
                        Integral I;
                        integral(T, rd, t_min, t_max, min_t != 0.0f ? min_t * dot3(T.g, rd) : 0.0f, 0.0f, I);
                        s_od[slot] = I.od;
                        s_cp[0][slot] = I.cp[0];
                        s_cp[1][slot] = I.cp[1];
                        s_cp[2][slot] = I.cp[2];
                        if (AUX) {
                            s_aux[0][slot] = t_min;
                            s_aux[1][slot] = t_max;
                            s_aux[2][slot] = I.w0;
                            s_ent[slot] = (uint32_t)fmin;
                        }
                        atomicOr(&s_mask[p], 1ull << k);
                    }
                }
            };
            if constexpr (R == 1) {
                one(j, c.row * TS + c.col);
            } else {
                PixWalk<TS> w;
                w.init(L, c);
                for (uint32_t q = 0; q < c.cnt; ++q) {
                    one(c.slot0 + q, w.row * TS + w.col);
                    w.next(L, k);
                }
            }
        }
        __syncthreads();
        // pixel-sequential compositing in depth order over this pixel's covered
        // tets (rasterize_compute.wgsl:336-463)
        if (active) {
            unsigned long long m = s_mask[tid];
            while (m) {
                const uint32_t k = (uint32_t)__ffsll((long long)m) - 1u;
                m &= m - 1ull;
                if (log_t < -13.8f) break;
                if constexpr (CK) {
                    const uint32_t q = done + k;
                    for (; next_ck <= q; next_ck += kCkGroup) ck[ck_base + (uint64_t)(next_ck / kCkGroup) * (TS * TS)] = log_t;
                }
                const uint32_t slot = (L.rpref[k * SP + row] + (uint32_t)colp) & 0xFFFFu;
                const float od = s_od[slot];
                const float Tj = fast_safe_exp(log_t);
                if (Tj >= 1e-6f) {
                    cr = cr + s_cp[0][slot] * Tj;
                    cg = cg + s_cp[1][slot] * Tj;
                    cb = cb + s_cp[2][slot] * Tj;
                    log_t = log_t - od;
                    if constexpr (CK) stop = done + k + 1u;
                    if (AUX) {
                        // rasterize_compute.wgsl:434-456
                        float alpha_t = fast_safe_exp(-od);
                        float alpha = 1.0f - alpha_t;
                        float w0 = s_aux[2][slot];
                        float w1 = 1.0f - (w0 + alpha_t);
                        float t_min = s_aux[0][slot], t_max = s_aux[1][slot];
                        int entry = (int)s_ent[slot];
                        V3 en = mk3(0.f, 0.f, 0.f);
                        if (entry < 4) {
                            const float4 q = L.tr[k * L.TR + entry];
                            en = normalize3(neg3(mk3(q.x, q.y, q.z)));
                        }
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
            if (log_t < -13.8f) s_live[tid] = 0u;
            else if constexpr (CK)  // positions up to the batch's end that this pixel passed uncovered
                for (; next_ck <= next && next_ck < count; next_ck += kCkGroup)
                    ck[ck_base + (uint64_t)(next_ck / kCkGroup) * (TS * TS)] = log_t;
        }
        if constexpr (MERGE) {
            if (nb_next > 0) stage(nb_next, next);
        }
        done = next;
        nb = nb_next;
    }
    if (C::PF == 2) __builtin_amdgcn_s_waitcnt(0);  // no LDS DMA outlives the workgroup
    if (active) {
        const uint64_t pix = vw.pix_base + (uint64_t)py * vw.width + px;
        if constexpr (CK) fin[pix] = make_float2(log_t, __uint_as_float(stop));
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

// Batch geometry of the adjoint (B <= 32, one mask bit each); SCAN = how a
// wave sums each tet's terms before the LDS adds: 2 segmented scan on DPP,
// 1 segmented scan on ds_bpermute, 0 none (every lane adds its own terms).
template <int TS_, int B_, int CAP_, int R_, int PF_, int SCAN_, int WPE_ = 0>
struct BwdCfg {
    static constexpr int TS = TS_;
    static constexpr int NT = (TS * TS + 63) / 64 * 64;
    static constexpr int B = B_;
    static constexpr int CAP = CAP_;
    static constexpr int R = R_;
    static constexpr int PF = (PF_ == 3 && 4 * B_ > NT) ? 1 : PF_;  // as FwdCfg::PF
    static constexpr int SCAN = SCAN_;
    static constexpr int WPE = WPE_ > 0 ? WPE_ : 1;
    static constexpr int PF_ARG = PF_;
    static constexpr int WPE_ARG = WPE_;
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

// The same segmented inclusive sum on DPP: four row_shr steps inside each
// 16-lane row, then row_bcast:15 (rows 1, 3) and row_bcast:31 (rows 2, 3)
// carry the previous rows' partial sums.  Lanes whose DPP source lies outside
// the row or the row mask read 0.  Masks are 0/1 floats (the summands are
// finite: every gradient term is clamped to +-1e20 before it is summed).
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_i(int v, int old) {
    return __builtin_amdgcn_update_dpp(old, v, CTRL, ROWS, 0xF, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_f(float v) {
    // full row mask: bound_ctrl (out-of-row lanes read 0) lets the shift fold into the FMA
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xF, ROWS == 0xF));
}
struct SegMasksDpp {
    float m[6];
};
__device__ __forceinline__ SegMasksDpp seg_masks_dpp(int key) {
    SegMasksDpp s;
    s.m[0] = dpp_i<0x111, 0xF>(key, -1) == key ? 1.0f : 0.0f;  // row_shr:1
    s.m[1] = dpp_i<0x112, 0xF>(key, -1) == key ? 1.0f : 0.0f;  // row_shr:2
    s.m[2] = dpp_i<0x114, 0xF>(key, -1) == key ? 1.0f : 0.0f;  // row_shr:4
    s.m[3] = dpp_i<0x118, 0xF>(key, -1) == key ? 1.0f : 0.0f;  // row_shr:8
    s.m[4] = dpp_i<0x142, 0xA>(key, -1) == key ? 1.0f : 0.0f;  // row_bcast:15
    s.m[5] = dpp_i<0x143, 0xC>(key, -1) == key ? 1.0f : 0.0f;  // row_bcast:31
    return s;
}
// All N sums level by level (independent DPP reads back to back, no hazard stalls).
template <int N>
__device__ __forceinline__ void seg_scan_dpp(float* v, const SegMasksDpp& s) {
#pragma unroll
    for (int q = 0; q < N; ++q) v[q] = fmaf(dpp_f<0x111, 0xF>(v[q]), s.m[0], v[q]);
#pragma unroll
    for (int q = 0; q < N; ++q) v[q] = fmaf(dpp_f<0x112, 0xF>(v[q]), s.m[1], v[q]);
#pragma unroll
    for (int q = 0; q < N; ++q) v[q] = fmaf(dpp_f<0x114, 0xF>(v[q]), s.m[2], v[q]);
#pragma unroll
    for (int q = 0; q < N; ++q) v[q] = fmaf(dpp_f<0x118, 0xF>(v[q]), s.m[3], v[q]);
#pragma unroll
    for (int q = 0; q < N; ++q) v[q] = fmaf(dpp_f<0x142, 0xA>(v[q]), s.m[4], v[q]);
#pragma unroll
    for (int q = 0; q < N; ++q) v[q] = fmaf(dpp_f<0x143, 0xC>(v[q]), s.m[5], v[q]);
}

// The same scan of 19 sums with the shift folded into the accumulate: one
// v_fmac_f32_dpp (v += dpp(v) * m) per sum and level.  Each level is one asm
// block opened by s_nop 1, the two wait states a DPP read needs after a VALU
// write of its source; inside a level every register is written once, and
// its next DPP read is 18 instructions later.
#define RMESH_FMAC_DPP(i, CTRL) "v_fmac_f32_dpp %" #i ", %" #i ", %19 " CTRL "\n"
#define RMESH_SCAN_LEVEL(M, CTRL)                                                                              \
    asm volatile("s_nop 1\n" RMESH_FMAC_DPP(0, CTRL) RMESH_FMAC_DPP(1, CTRL) RMESH_FMAC_DPP(2, CTRL)            \
                     RMESH_FMAC_DPP(3, CTRL) RMESH_FMAC_DPP(4, CTRL) RMESH_FMAC_DPP(5, CTRL) RMESH_FMAC_DPP(6, CTRL) \
                         RMESH_FMAC_DPP(7, CTRL) RMESH_FMAC_DPP(8, CTRL) RMESH_FMAC_DPP(9, CTRL)                    \
                             RMESH_FMAC_DPP(10, CTRL) RMESH_FMAC_DPP(11, CTRL) RMESH_FMAC_DPP(12, CTRL)             \
                                 RMESH_FMAC_DPP(13, CTRL) RMESH_FMAC_DPP(14, CTRL) RMESH_FMAC_DPP(15, CTRL)         \
                                     RMESH_FMAC_DPP(16, CTRL) RMESH_FMAC_DPP(17, CTRL) RMESH_FMAC_DPP(18, CTRL)     \
                 : "+v"(v[0]), "+v"(v[1]), "+v"(v[2]), "+v"(v[3]), "+v"(v[4]), "+v"(v[5]), "+v"(v[6]), "+v"(v[7]),  \
                   "+v"(v[8]), "+v"(v[9]), "+v"(v[10]), "+v"(v[11]), "+v"(v[12]), "+v"(v[13]), "+v"(v[14]),         \
                   "+v"(v[15]), "+v"(v[16]), "+v"(v[17]), "+v"(v[18])                                             \
                 : "v"(M))
__device__ __forceinline__ void seg_scan_dpp19(float* v, const SegMasksDpp& s) {
    RMESH_SCAN_LEVEL(s.m[0], "row_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1");
    RMESH_SCAN_LEVEL(s.m[1], "row_shr:2 row_mask:0xf bank_mask:0xf bound_ctrl:1");
    RMESH_SCAN_LEVEL(s.m[2], "row_shr:4 row_mask:0xf bank_mask:0xf bound_ctrl:1");
    RMESH_SCAN_LEVEL(s.m[3], "row_shr:8 row_mask:0xf bank_mask:0xf bound_ctrl:1");
    RMESH_SCAN_LEVEL(s.m[4], "row_bcast:15 row_mask:0xa bank_mask:0xf");
    RMESH_SCAN_LEVEL(s.m[5], "row_bcast:31 row_mask:0xc bank_mask:0xf");
}
#undef RMESH_SCAN_LEVEL
#undef RMESH_FMAC_DPP

// Hit barycentrics of face fi's roles (a, b, c) placed at the tet's vertex
// slots, 0 at the opposite vertex.  FACES = {0,2,1}, {1,2,3}, {0,3,2},
// {3,0,1}: slot s of face f takes role
//   f0: a c b -   f1: - a b c   f2: a - c b   f3: b c - a
// Three face tests shared by the four slots (three selects each), in place
// of a per-slot search over the face's role slots.
__device__ __forceinline__ void face_slot_weights(int fi, float la, float lb, float lc, float* w) {
    const bool f0 = fi == 0, f1 = fi == 1, f2 = fi == 2;
    w[0] = f0 ? la : (f1 ? 0.0f : (f2 ? la : lb));
    w[1] = f0 ? lc : (f1 ? la : (f2 ? 0.0f : lc));
    w[2] = f0 ? lb : (f1 ? lb : (f2 ? lc : 0.0f));
    w[3] = f0 ? 0.0f : (f1 ? lc : (f2 ? lb : la));
}

// Gradient terms of one covered (pixel, tet) pair, accumulated into acc
// (backward_tiled_compute.wgsl:446-543, per-pixel clipping as there).  The
// intersection derivatives use dt/dva = ((vc - hit) x (vb - hit)) / den (and
// cyclic), algebraically identical to the reference's
// (cross(va - hit, vb - vc) + n) / den; since the three crosses are parallel
// to the face normal n they equal lambda_j * n with the hit's barycentric
// weights lambda_j.  The weights are affine in the hit point, so each face
// carries their gradients and their values at v0 (store_tet) and a pair
// evaluates them with three dot products; they are scattered to the face's
// vertex slots.
//
// The reference clips every vertex term, each side's term and the colour-
// gradient term to +-1e20 per pixel.  A pair whose terms are bounded below
// 1e19 in magnitude (|G| max-norm x the largest weight per side, plus
// |g| |dbo|; |cam - v0| |dbo| + |rd| |d_dcdt| for dg) cannot reach a clip,
// so it accumulates n (k w) with fused multiply-adds and no clips; any other
// pair (and NaN) takes the clipped path, term by term as the reference.
// FIRST: acc holds nothing yet (one pair per lane): terms are assigned, not
// added to zeros.
template <int TS, int B, bool FIRST>
__device__ __forceinline__ void pair_grad(const BatchLDS<TS, B, true>& L, int k, const TetRegs& T, V3 v0, V3 cam,
                                          V3 rd, float dc0, float dc1, float dc2, float t_prev, float d_log_t_b,
                                          float t_min, float t_max, int min_face, int max_face, float* acc) {
#pragma clang fp contract(fast)
    Integral I;
    integral<false>(T, rd, t_min, t_max, 0.0f, kMinVal, I);
    const float dist = t_max - t_min;
    const float dens = T.dens;
    const float dc_dt = fdot3(T.g, rd);
    float dcp0 = med_clip(dc0 * t_prev, kMinVal, kMaxVal);
    float dcp1 = med_clip(dc1 * t_prev, kMinVal, kMaxVal);
    float dcp2 = med_clip(dc2 * t_prev, kMinVal, kMaxVal);
    float d_od_state = -d_log_t_b;
    float dphi = fast_dphi(I.od, I.alpha_t);
    float dw0 = dphi + I.alpha_t, dw1 = -dphi;
    float d_od_int = fmaf(dcp0, fmaf(I.ce[0], dw0, I.cs[0] * dw1),
                          fmaf(dcp1, fmaf(I.ce[1], dw0, I.cs[1] * dw1), dcp2 * fmaf(I.ce[2], dw0, I.cs[2] * dw1)));
    float d_od = med_clip(d_od_state + d_od_int, kMinVal, kMaxVal);
    float od_raw = dens * dist;
    bool act = (od_raw >= 1e-8f) && (od_raw <= 88.0f);
    float d_od_raw = act ? d_od : 0.0f;
    float d_dens_l = med_clip(d_od_raw * dist, kMinVal, kMaxVal);
    float d_dist = med_clip(d_od_raw * dens, kMinVal, kMaxVal);
    float d_tmin = -d_dist, d_tmax = d_dist;
    float dcs0 = I.csr[0] > 0.f ? dcp0 * I.w1 : 0.f, dcs1 = I.csr[1] > 0.f ? dcp1 * I.w1 : 0.f;
    float dcs2 = I.csr[2] > 0.f ? dcp2 * I.w1 : 0.f;
    float dce0 = I.cer[0] > 0.f ? dcp0 * I.w0 : 0.f, dce1 = I.cer[1] > 0.f ? dcp1 * I.w0 : 0.f;
    float dce2 = I.cer[2] > 0.f ? dcp2 * I.w0 : 0.f;
    float dbc0 = dcs0 + dce0, dbc1 = dcs1 + dce1, dbc2 = dcs2 + dce2;
    float sum_s = (dcs0 + dcs1) + dcs2, sum_e = (dce0 + dce1) + dce2;
    float d_dcdt = fmaf(sum_s, t_min, sum_e * t_max);
    d_tmin = fmaf(sum_s, dc_dt, d_tmin);
    d_tmax = fmaf(sum_e, dc_dt, d_tmax);
    float dbo = (dbc0 + dbc1) + dbc2;
    const V3 cam_v0 = sub3(cam, v0);
    float wv[2][4];   // per side: barycentric weight of each vertex slot
    V3 nn[2];         // per side: raw face normal
    float kd[2];      // per side: dt / den
    float bound = 0.0f;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const int fi = side == 0 ? min_face : max_face;
        const float tt = side == 0 ? t_min : t_max;
        const float dt = side == 0 ? d_tmin : d_tmax;
        const float4* F = reinterpret_cast<const float4*>(&L.face[k * L.FS + fi * 8]);
        const float4 f0 = F[0], f1 = F[1];
        const float4 nq = L.tr[k * L.TR + fi];
        const V3 ga = mk3(f0.x, f0.y, f0.z), gb = mk3(f1.x, f1.y, f1.z);
        const float sg = L.face[k * L.FS + 32 + fi];
        nn[side] = mk3(nq.x * sg, nq.y * sg, nq.z * sg);
        const float inv_den = fast_safe_div(1.0f, fdot3(nn[side], rd));
        const V3 h = mk3(fmaf(rd.x, tt, cam_v0.x), fmaf(rd.y, tt, cam_v0.y), fmaf(rd.z, tt, cam_v0.z));  // hit - v0
        const float la = f0.w + fdot3(ga, h);
        const float lb = f1.w + fdot3(gb, h);
        const float lc = (1.0f - la) - lb;
        kd[side] = dt * inv_den;
        face_slot_weights(fi, la, lb, lc, wv[side]);
        const float nmax = fmaxf(fmaxf(fabsf(nq.x), fabsf(nq.y)), fabsf(nq.z));
        const float wmax = fmaxf(fmaxf(fabsf(la), fabsf(lb)), fabsf(lc));
        bound = fmaf(fabsf(kd[side]) * nmax, wmax, bound);
    }
    const float gmax = fmaxf(fmaxf(fabsf(T.g.x), fabsf(T.g.y)), fabsf(T.g.z));
    const float cmax = fmaxf(fmaxf(fabsf(cam_v0.x), fabsf(cam_v0.y)), fabsf(cam_v0.z));
    const float rmax = fmaxf(fmaxf(fabsf(rd.x), fabsf(rd.y)), fabsf(rd.z));
    bound = fmaf(gmax, fabsf(dbo), bound);
    const float bound_g = fmaf(cmax, fabsf(dbo), rmax * fabsf(d_dcdt));
    auto put = [&](int q, float x) { acc[q] = FIRST ? x : acc[q] + x; };
    // x1 k1 + x0 k0 (+ acc)
    auto put2 = [&](int q, float x0, float k0, float x1, float k1) {
        acc[q] = FIRST ? fmaf(x1, k1, x0 * k0) : fmaf(x1, k1, fmaf(x0, k0, acc[q]));
    };
    put(0, med_clip(d_dens_l, kMinVal, kMaxVal));
    if (bound <= 1e19f && bound_g <= 1e19f) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const float k0 = kd[0] * wv[0][s], k1 = kd[1] * wv[1][s];
            put2(1 + 3 * s, nn[0].x, k0, nn[1].x, k1);
            put2(2 + 3 * s, nn[0].y, k0, nn[1].y, k1);
            put2(3 + 3 * s, nn[0].z, k0, nn[1].z, k1);
        }
        acc[1] = fmaf(-T.g.x, dbo, acc[1]);
        acc[2] = fmaf(-T.g.y, dbo, acc[2]);
        acc[3] = fmaf(-T.g.z, dbo, acc[3]);
        put2(13, rd.x, d_dcdt, cam_v0.x, dbo);
        put2(14, rd.y, d_dcdt, cam_v0.y, dbo);
        put2(15, rd.z, d_dcdt, cam_v0.z, dbo);
    } else {
        const V3 G0 = scl3(nn[0], kd[0]), G1 = scl3(nn[1], kd[1]);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            V3 d = add3(med_clip3(scl3(G0, wv[0][s]), kMinVal, kMaxVal), med_clip3(scl3(G1, wv[1][s]), kMinVal, kMaxVal));
            if (s == 0) d = add3(d, scl3(neg3(T.g), dbo));
            d = med_clip3(d, kMinVal, kMaxVal);
            put(1 + 3 * s, d.x);
            put(2 + 3 * s, d.y);
            put(3 + 3 * s, d.z);
        }
        V3 dgc = med_clip3(add3(scl3(cam_v0, dbo), scl3(rd, d_dcdt)), kMinVal, kMaxVal);
        put(13, dgc.x);
        put(14, dgc.y);
        put(15, dgc.z);
    }
    V3 dbcc = med_clip3(mk3(dbc0, dbc1, dbc2), kMinVal, kMaxVal);
    put(16, dbcc.x);
    put(17, dbcc.y);
    put(18, dbcc.z);
}

// CK: the checkpointed adjoint.  The undo restarts at every 32-position group
// of the tile list from the log T the forward stored at the group's end (the
// pixel's final log T for its last group), and pairs at or after the pixel's
// stop position — never composited — are skipped: the exact adjoint of the
// forward.  Without CK: the reference's undo from log(1 - alpha).
template <class C, bool CK>
__global__ __launch_bounds__(C::NT) __attribute__((amdgpu_waves_per_eu(C::WPE))) void k_backward(
    const float4* __restrict__ rec, const uint32_t* __restrict__ idx, const float4* __restrict__ colors,
    const uint32_t* __restrict__ sorted_vals, const uint32_t* __restrict__ ranges,
    const ViewParams* __restrict__ views, const uint32_t* __restrict__ tile_view, const uint32_t* __restrict__ tile_order,
    const float* __restrict__ rgba, const float* __restrict__ dl_img, const float* __restrict__ gt, uint32_t loss_type,
    float* __restrict__ loss_partial, float* __restrict__ tet_g, float* __restrict__ d_colors,
    const float* __restrict__ ck, const float2* __restrict__ fin, uint64_t* __restrict__ scene_done) {
    constexpr int TS = C::TS, NT = C::NT, B = C::B, CAP = C::CAP, R = C::R, SP = TS + 1, NW = NT / 64;
    __shared__ BatchLDS<TS, B, true> L;
    // per-pair record: od -> t_prev, dc . cp -> d_log_t before, t_min, t_max
    // (dc . cp, the pixel's image gradient dotted with the pair's premultiplied
    // colour, is all the d(log T) chain needs of the colour)
    __shared__ float s_r[4][CAP];
    __shared__ uint8_t s_meta[CAP];    // fmin | fmax << 2 | valid << 4 | did << 5
    __shared__ float s_ray[3][TS * TS];
    __shared__ float s_dc[3][TS * TS];
    __shared__ float s_acc[B][kTerms];
    __shared__ uint32_t s_ftid[B];  // tet ids of the batch whose terms s_acc holds
    // covered tets of the batch per pixel, one bit per tet
    using MaskT = typename std::conditional<(B > 32), unsigned long long, unsigned int>::type;
    __shared__ MaskT s_mask[TS * TS];
    __shared__ chunk_t<R> s_chunk[chunk_capacity(CAP, R, B)];
    static_assert(CAP <= 2048 && TS <= 32 && R <= 31, "chunk table packing");
    __shared__ float s_loss[NW];
    static_assert(B <= 64, "pixel masks hold 64 tets");
    __shared__ uint32_t s_start[R == 1 ? CAP / 32 : 1];  // R = 1: record slots that start a row segment
    static_assert(R != 1 || (CAP % 32 == 0 && CAP / 32 <= NT), "segment start mask");
    static_assert(CAP >= TS * TS, "one tet's pixels of a tile must fit a batch");

    const uint32_t gtile = tile_order ? tile_order[blockIdx.x] : blockIdx.x;
    const uint32_t view = tile_view[gtile];
    const ViewParams& vw = views[view];
    const TileCtx tc = tile_ctx(vw, gtile);
    const float4* col_v = colors + vw.color_off;
    float* dcol_v = d_colors + 3 * vw.color_off;
    const int tid = threadIdx.x, wid = tid / 64, lane = (int)__lane_id();
    const int row = tid / TS, colp = tid % TS;
    const uint32_t px = tc.tile_x * TS + colp, py = tc.tile_y * TS + row;
    const bool active = tid < TS * TS && px < vw.width && py < vw.height;
    const uint64_t pix = vw.pix_base + (uint64_t)py * vw.width + px;
    const V3 cam = mk3(vw.cam[0], vw.cam[1], vw.cam[2]);

    // Initial state from the rendered image (backward_tiled_compute.wgsl:186-204);
    // with a ground-truth image the loss gradient is formed here
    // (loss_compute.wgsl:24-72): dL/dalpha = 0.
    float log_t = 0.f, d_log_t = 0.f, lsum = 0.f;
    float fin_lt = 0.f;   // CK: the pixel's final log T
    uint32_t stop = 0u;   // CK: pairs at positions >= stop were never composited
    if (tid < TS * TS) {
        float dcv[3] = {0.f, 0.f, 0.f};
        V3 rd = mk3(0.f, 0.f, 1.f);
        if (active) {
            rd = pixel_ray(vw, px, py);
            float4 im = reinterpret_cast<const float4*>(rgba)[pix];
            if constexpr (CK) {
                const float2 fs = fin[pix];
                fin_lt = fs.x;
                stop = __float_as_uint(fs.y);
                log_t = fin_lt;
            } else {
                log_t = logf(fmaxf(1.0f - im.w, 1e-20f));
            }
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
                d_log_t = -d.w * (CK ? safe_exp(fin_lt) : 1.0f - im.w);
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
    int nb = (int)min<uint32_t>(B, count);
    // CK: slot of checkpoint j + 1 (the end of group j) is ck_base + j
    const uint64_t ck_base = ((uint64_t)(rs >> 5) + gtile) * (uint64_t)(TS * TS) + (uint64_t)tid;
    int32_t cur = -1;  // CK: group the pixel's undo is in
    TetFetch f;
    static_assert(C::PF != 2 || B <= 64, "LDS-DMA staging: one wave, next batch inside a 128-entry id window");
    static_assert(!CK || B <= (int)kCkGroup, "a batch spans at most two checkpoint groups");
    __shared__ Staging<C::PF == 2 ? B : 1> S;
    QuadFetch qf{};
    if (C::PF == 1 && tid < nb) fetch_tet<true>(f, sorted_vals[rs + tid], rec, idx, col_v);
    if (C::PF == 3 && tid < 4 * nb) fetch_quad<true>(qf, sorted_vals[rs + (tid >> 2)], tid & 3, rec, idx, col_v);
    if (C::PF == 2 && wid == 0) {
        const uint32_t t = lane < nb ? sorted_vals[rs + lane] : 0u;
        stage_async<true>(S, t, lane < nb, rec, idx, col_v);
    }
    // The flush of a batch's terms runs at the top of the next iteration, in
    // the same phase as the next batch's LDS image (they touch disjoint LDS):
    // one barrier phase fewer per batch, and the atomics overlap the staging.
    uint32_t flush_n = 0;  // tets of the previous batch to flush
    auto flush = [&]() {
        // one atomic per non-zero term per (tile, tet), the first 16 into the
        // tet's 64-byte staging record (one L2 request for the lanes of a
        // tet), the colour terms into the view's colour gradients
        for (int it = tid; it < (int)flush_n * kTerms; it += NT) {
            const int k = it / kTerms, q = it % kTerms;
            const float s = s_acc[k][q];
            if (s == 0.0f) continue;
            const uint32_t t = s_ftid[k];
            float* dst = q < 16 ? tet_g + 16ull * t + q : dcol_v + 3ull * t + (q - 16);
            unsafeAtomicAdd(dst, s);
        }
    };
    while (done < count) {
        if (C::PF == 2) __builtin_amdgcn_s_waitcnt(0);  // staged records of this batch have landed
        __syncthreads();
        flush();
        uint32_t win = 0, win1 = 0;  // wave 0: tet ids of this and the next batch, 64 per register
        if (C::PF == 2 && wid == 0) {
            win = done + lane < count ? sorted_vals[rs + done + lane] : 0u;
            if constexpr (B > 32) win1 = done + 64 + lane < count ? sorted_vals[rs + done + 64 + lane] : 0u;
        }
        if (C::PF == 0 && tid < nb) fetch_tet<true>(f, sorted_vals[rs + done + tid], rec, idx, col_v);
        if constexpr (C::PF == 2) store_batch_staged<TS, B, true, NT>(L, S, nb, vw, tc);
        else if constexpr (C::PF == 3) store_quad<TS, B, true>(L, qf, tid >> 2, tid & 3, tid < 4 * nb, vw, tc);
        else if (tid < nb) store_tet<TS, B, true>(L, tid, f, vw, tc);
        if (tid < TS * TS) s_mask[tid] = 0u;
        if (R == 1 && tid < CAP / 32) s_start[tid] = 0u;
        __syncthreads();
        batch_layout<TS, B, true, false, CAP, R, NT>(L, nb, s_chunk, s_start);
        __syncthreads();
        const uint32_t nbp = L.nbp & 0xFFu;
        const uint32_t nchunks = L.nbp >> 8;
        const uint32_t next = done + nbp;
        const int nb_next = next_batch<B>(nbp, nb, count - next);
        // CK: log T at the end of the (at most two) groups this batch spans,
        // positions [count - next, count - 1 - done] in compositing order
        float ck_hi = 0.f, ck_lo = 0.f;
        int32_t j_hi = 0;
        if (CK && active) {
            j_hi = (int32_t)((count - 1u - done) / kCkGroup);
            const int32_t j_lo = (int32_t)((count - next) / kCkGroup);
            ck_hi = (uint32_t)(j_hi + 1) * kCkGroup < stop ? ck[ck_base + (uint64_t)j_hi * (TS * TS)] : fin_lt;
            ck_lo = (uint32_t)(j_lo + 1) * kCkGroup < stop ? ck[ck_base + (uint64_t)j_lo * (TS * TS)] : fin_lt;
        }
        if (C::PF == 1 && tid < nb_next) fetch_tet<true>(f, sorted_vals[rs + next + tid], rec, idx, col_v);
        if (C::PF == 3 && tid < 4 * nb_next)
            fetch_quad<true>(qf, sorted_vals[rs + next + (tid >> 2)], tid & 3, rec, idx, col_v);
        if (C::PF == 2 && wid == 0) {
            const uint32_t tn = window_id<B>(win, win1, (int)nbp + lane);
            stage_async<true>(S, tn, lane < nb_next, rec, idx, col_v);
        }
        // the flush above has read s_acc and s_ftid (two barriers ago)
        for (int i = tid; i < B * kTerms; i += NT) (&s_acc[0][0])[i] = 0.0f;
        if (tid < (int)nbp) s_ftid[tid] = L.tid[tid];
        // phase A (record- or chunk-parallel): intersection + integral
        for (uint32_t j = tid; j < nchunks; j += NT) {
            ChunkRef c;
            if constexpr (R == 1) c = record_at(s_chunk, s_start, j);
            else c = chunk_at<R>(s_chunk, j);
            const int k = c.k;
            TetRegs T;
            tet_regs(L, k, T);
            auto one = [&](uint32_t slot, int prow, int pcol) {
                const uint32_t gx = tc.tile_x * TS + pcol, gy = tc.tile_y * TS + prow;
                uint32_t meta = 0;
                if (gx < vw.width && gy < vw.height) {
                    const int p = prow * TS + pcol;
                    V3 rd = mk3(s_ray[0][p], s_ray[1][p], s_ray[2][p]);
                    float t_min, t_max;
                    int fmin, fmax;
                    if (intersect(T, rd, 0.0f, t_min, t_max, fmin, fmax)) {
                        // the reference's adjoint defaults both faces to 0 (backward_tiled_compute.wgsl:365-366)
                        if (fmin == 4) fmin = 0;
                        Integral I;
                        integral<false>(T, rd, t_min, t_max, 0.0f, kMinVal, I);
                        s_r[0][slot] = I.od;
                        s_r[1][slot] = (s_dc[0][p] * I.cp[0] + s_dc[1][p] * I.cp[1]) + s_dc[2][p] * I.cp[2];
                        s_r[2][slot] = t_min;
                        s_r[3][slot] = t_max;
                        meta = (uint32_t)fmin | ((uint32_t)fmax << 2) | (1u << 4);
                        atomicOr(&s_mask[p], (MaskT)1 << k);
                    }
                }
                s_meta[slot] = (uint8_t)meta;
            };
            if constexpr (R == 1) {
                one(j, c.row, c.col);
            } else {
                PixWalk<TS> w;
                w.init(L, c);
                for (uint32_t q = 0; q < c.cnt; ++q) {
                    one(c.slot0 + q, w.row, w.col);
                    w.next(L, k);
                }
            }
        }
        __syncthreads();
        // phase B (pixel-sequential, back to front over this pixel's covered
        // tets): undo and the d(log T) chain (backward_tiled_compute.wgsl:419-451)
        if (active) {
            MaskT m = s_mask[tid];
            while (m) {
                const uint32_t k = B > 32 ? (uint32_t)__ffsll((long long)m) - 1u : (uint32_t)__ffs((int)m) - 1u;
                m &= m - (MaskT)1;
                if constexpr (CK) {
                    const uint32_t q = count - 1u - done - k;  // position in compositing order
                    if (q >= stop) continue;
                    const int32_t j = (int32_t)(q / kCkGroup);
                    if (j != cur) {
                        log_t = j == j_hi ? ck_hi : ck_lo;
                        cur = j;
                    }
                }
                const uint32_t slot = (L.rpref[k * SP + row] + (uint32_t)colp) & 0xFFFFu;
                const float od = s_r[0][slot];
                const float prev_log_t = fminf(log_t + od, 0.0f);
                const float t_prev = fast_safe_exp(prev_log_t);
                log_t = prev_log_t;
                if (t_prev < 1e-6f) continue;
                const float dcp = s_r[1][slot];
                s_r[0][slot] = t_prev;
                s_r[1][slot] = d_log_t;
                s_meta[slot] = (uint8_t)(s_meta[slot] | 32u);
                d_log_t = safe_clip(d_log_t + dcp * t_prev, kMinVal, kMaxVal);
            }
        }
        __syncthreads();
        // phase C (chunk-parallel): gradient terms of R pixels of one tet per
        // lane, then a segmented wave sum keyed by tet
        for (uint32_t base = (uint32_t)wid * 64; base < nchunks; base += NT) {
            const uint32_t j = base + lane;
            float acc[kTerms];
            if constexpr (R != 1) {
#pragma unroll
                for (int q = 0; q < kTerms; ++q) acc[q] = 0.0f;
            }
            bool have = false;  // R = 1: acc assigned by the lane's one pair
            int key = 0x7fffffff;
            if (j < nchunks) {
                ChunkRef c;
                if constexpr (R == 1) c = record_at(s_chunk, s_start, j);
                else c = chunk_at<R>(s_chunk, j);
                const int k = c.k;
                key = k;
                TetRegs T;
                tet_regs(L, k, T);
                const float4 v0q = L.tr[k * L.TR + 6];
                const V3 v0 = mk3(v0q.x, v0q.y, v0q.z);
                auto one = [&](uint32_t slot, int p) {
                    const uint32_t meta = s_meta[slot];
                    if (meta & 32u) {
                        const V3 rd = mk3(s_ray[0][p], s_ray[1][p], s_ray[2][p]);
                        pair_grad<TS, B, R == 1>(L, k, T, v0, cam, rd, s_dc[0][p], s_dc[1][p], s_dc[2][p],
                                                 s_r[0][slot], s_r[1][slot], s_r[2][slot], s_r[3][slot],
                                                 (int)(meta & 3u), (int)((meta >> 2) & 3u), acc);
                        have = true;
                    }
                };
                if constexpr (R == 1) {
                    one(j, c.row * TS + c.col);
                } else {
                    PixWalk<TS> w;
                    w.init(L, c);
                    for (uint32_t q = 0; q < c.cnt; ++q) {
                        one(c.slot0 + q, w.row * TS + w.col);
                        w.next(L, k);
                    }
                }
            }
            if constexpr (R == 1) {
                if (!have) {
#pragma unroll
                    for (int q = 0; q < kTerms; ++q) acc[q] = 0.0f;
                }
            }
            if (C::SCAN >= 2) {
                // lanes of one tet are contiguous: a segmented wave scan leaves
                // each tet's sum in its last lane, which adds it to LDS
                const SegMasksDpp sm = seg_masks_dpp(key);
                const int next_key = dpp_i<0x130, 0xF>(key, -1);  // wave_shl:1 (lane 63 reads -1)
                const bool tail = key != 0x7fffffff && next_key != key;
                if constexpr (C::SCAN == 3) {
                    static_assert(kTerms == 19, "the asm scan holds 19 sums");
                    seg_scan_dpp19(acc, sm);
                } else {
                    seg_scan_dpp<kTerms>(acc, sm);
                }
                // no zero test per term: a tail's adds are one LDS atomic per
                // term for the few tail lanes, cheaper than 19 compares and branches
                if (tail) {
#pragma unroll
                    for (int q = 0; q < kTerms; ++q) atomicAdd(&s_acc[key][q], acc[q]);
                }
            } else if (C::SCAN == 1) {
                const SegMasks sm = seg_masks(key);
                const int next_key = __shfl_down(key, 1, 64);
                const bool tail = key != 0x7fffffff && (lane == 63 || next_key != key);
#pragma unroll
                for (int q = 0; q < kTerms; ++q) {
                    float s = seg_scan(acc[q], sm);
                    if (tail && s != 0.0f) atomicAdd(&s_acc[key][q], s);
                }
            } else if (key != 0x7fffffff) {
#pragma unroll
                for (int q = 0; q < kTerms; ++q)
                    if (acc[q] != 0.0f) atomicAdd(&s_acc[key][q], acc[q]);
            }
        }
        flush_n = nbp;
        done = next;
        nb = nb_next;
    }
    __syncthreads();
    flush();
    if (C::PF == 2) __builtin_amdgcn_s_waitcnt(0);  // no LDS DMA outlives the workgroup
    if (scene_done && tid == 0) atomicMax(reinterpret_cast<unsigned long long*>(scene_done) + vw.scene,
                                          (unsigned long long)wall_clock64());
}

// ---------------------------------------------------------------------------
// host launchers: kFwdVariants x kBwdVariants batch geometries per tile size,
// picked per renderer by rmesh_autotune (variant 0 is the measured default)
// ---------------------------------------------------------------------------
template <class C>
static void fwd_launch(const FwdArgs& a, hipStream_t s) {
    if (a.aux)  // the aux outputs serve the forward-only API: no checkpoints with them
        hipLaunchKernelGGL((k_forward<C, true, false>), dim3(a.num_gtiles), dim3(C::NT), 0, s, a.rec, a.colors,
                           a.sorted_vals, a.ranges, a.views, a.tile_view, a.tile_order,
                           a.rgba, a.aux, nullptr, nullptr);
    else if (a.ck)
        hipLaunchKernelGGL((k_forward<C, false, true>), dim3(a.num_gtiles), dim3(C::NT), 0, s, a.rec, a.colors,
                           a.sorted_vals, a.ranges, a.views, a.tile_view, a.tile_order,
                           a.rgba, nullptr, a.ck, a.fin);
    else
        hipLaunchKernelGGL((k_forward<C, false, false>), dim3(a.num_gtiles), dim3(C::NT), 0, s, a.rec, a.colors,
                           a.sorted_vals, a.ranges, a.views, a.tile_view, a.tile_order,
                           a.rgba, nullptr, nullptr, nullptr);
}

template <int TS>
static bool fwd_variant(const FwdArgs& a, int v, hipStream_t s) {
    switch (v) {
        case 0: fwd_launch<FwdCfg<TS, 48, 768, 1, 3>>(a, s); return true;
        case 1: fwd_launch<FwdCfg<TS, 48, 704, 1, 3, 5>>(a, s); return true;
        case 2: fwd_launch<FwdCfg<TS, 32, 704, 1, 3, 5>>(a, s); return true;
        case 3: fwd_launch<FwdCfg<TS, 64, 1024, 1, 3>>(a, s); return true;
        case 4: fwd_launch<FwdCfg<TS, 64, 768, 1, 3, 5>>(a, s); return true;
        case 5: fwd_launch<FwdCfg<TS, 56, 896, 1, 3>>(a, s); return true;
        case 6: fwd_launch<FwdCfg<TS, 56, 768, 1, 3, 5>>(a, s); return true;
        case 7: fwd_launch<FwdCfg<TS, 64, 896, 1, 3>>(a, s); return true;
        default: return false;
    }
}

// Tile 32: 1024-thread workgroups; a batch must hold one tet's 1024 pixels
// (CAP >= TS^2), so two geometries only.
static bool fwd_variant32(const FwdArgs& a, int v, hipStream_t s) {
    switch (v) {
        case 0: fwd_launch<FwdCfg<32, 32, 1024, 1, 3>>(a, s); return true;
        case 1: fwd_launch<FwdCfg<32, 48, 1536, 1, 3>>(a, s); return true;
        default: return false;
    }
}

bool launch_forward(const FwdArgs& a, uint32_t tile_size, int variant, hipStream_t s) {
    switch (tile_size) {
        case 8: return fwd_variant<8>(a, variant, s);
        case 12: return fwd_variant<12>(a, variant, s);
        case 16: return fwd_variant<16>(a, variant, s);
        case 32: return fwd_variant32(a, variant, s);
        default: return false;
    }
}

template <class C>
static void bwd_launch(const BwdArgs& a, hipStream_t s) {
    if (a.ck) {
        if constexpr (C::B > 32) {
            // the checkpointed adjoint's batches span at most two checkpoint groups (B <= 32)
            bwd_launch<BwdCfg<C::TS, 32, C::CAP, C::R, C::PF_ARG, C::SCAN, C::WPE_ARG>>(a, s);
        } else {
            hipLaunchKernelGGL((k_backward<C, true>), dim3(a.num_gtiles), dim3(C::NT), 0, s, a.rec, a.idx, a.colors,
                               a.sorted_vals, a.ranges, a.views, a.tile_view, a.tile_order,
                               a.rgba, a.dl_img, a.gt, a.loss_type, a.loss_partial, a.tet_g, a.d_colors, a.ck, a.fin,
                               a.scene_done);
        }
    } else
        hipLaunchKernelGGL((k_backward<C, false>), dim3(a.num_gtiles), dim3(C::NT), 0, s, a.rec, a.idx, a.colors,
                           a.sorted_vals, a.ranges, a.views, a.tile_view, a.tile_order,
                           a.rgba, a.dl_img, a.gt, a.loss_type, a.loss_partial, a.tet_g, a.d_colors, nullptr,
                           nullptr, a.scene_done);
}

template <int TS>
static bool bwd_variant(const BwdArgs& a, int v, hipStream_t s) {
    switch (v) {
        case 0: bwd_launch<BwdCfg<TS, 32, 480, 1, 3, 3, 5>>(a, s); return true;
        case 1: bwd_launch<BwdCfg<TS, 32, 512, 1, 3, 3, 4>>(a, s); return true;
        case 2: bwd_launch<BwdCfg<TS, 32, 512, 1, 2, 3, 4>>(a, s); return true;
        case 3: bwd_launch<BwdCfg<TS, 32, 576, 1, 3, 3, 4>>(a, s); return true;
        case 4: bwd_launch<BwdCfg<TS, 32, 640, 1, 2, 3, 4>>(a, s); return true;
        case 5: bwd_launch<BwdCfg<TS, 32, 640, 1, 3, 3, 4>>(a, s); return true;
        // 31.5 KB of LDS and 96 VGPRs at tile 16: 5 workgroups (20 waves) per CU
        case 6: bwd_launch<BwdCfg<TS, 32, 544, 1, 3, 3, 5>>(a, s); return true;
        // more tets per batch where they are small (the checkpointed mode runs these with B 32)
        case 7: bwd_launch<BwdCfg<TS, 48, 576, 1, 3, 3, 4>>(a, s); return true;
        case 8: bwd_launch<BwdCfg<TS, 64, 512, 1, 3, 3, 4>>(a, s); return true;
        case 9: bwd_launch<BwdCfg<TS, 40, 576, 1, 3, 3, 4>>(a, s); return true;
        case 10: bwd_launch<BwdCfg<TS, 48, 512, 1, 3, 3, 4>>(a, s); return true;
        default: return false;
    }
}

static bool bwd_variant32(const BwdArgs& a, int v, hipStream_t s) {
    switch (v) {
        case 0: bwd_launch<BwdCfg<32, 32, 1024, 1, 3, 3, 4>>(a, s); return true;
        case 1: bwd_launch<BwdCfg<32, 16, 1024, 1, 3, 3, 4>>(a, s); return true;
        default: return false;
    }
}

bool launch_backward(const BwdArgs& a, uint32_t tile_size, int variant, hipStream_t s) {
    switch (tile_size) {
        case 8: return bwd_variant<8>(a, variant, s);
        case 12: return bwd_variant<12>(a, variant, s);
        case 16: return bwd_variant<16>(a, variant, s);
        case 32: return bwd_variant32(a, variant, s);
        default: return false;
    }
}


// ---------------------------------------------------------------------------
// Per-tet error statistics for densification (forward_error):
//   crates/rmesh-error/src/wgsl/error_tiled_compute.wgsl:88-373
// The forward's tiled traversal on the last forward's tile lists (one view):
// batches of B tets front to back, the forward's spans and slab test.  Phase A
// (record-parallel) keeps each covered pair's slab and optical depth; phase B
// (pixel-sequential, depth order) applies the -5.5 log-transmittance cut-off
// and turns a record into its weight T_val = exp(log T) od (0: no
// contribution); phase C (record-parallel) forms the 14 weighted moments and
// the pixel count, sums them per tet with the segmented DPP scan, and the batch
// adds one atomic per statistic per (tile, tet); the max weight goes through an
// LDS atomicMax.
// ---------------------------------------------------------------------------
constexpr int kErrStats = 15;  // tet_err columns 0 1 2 5 6..15 (14), then the pixel count
__device__ __forceinline__ int err_column(int q) { return q < 3 ? q : q + 2; }

template <class C>
__global__ __launch_bounds__(C::NT) void k_error(const float4* __restrict__ rec, const uint32_t* __restrict__ sorted_vals,
                                                 const uint32_t* __restrict__ ranges,
                                                 const ViewParams* __restrict__ views, const float* __restrict__ pixel_err,
                                                 const float* __restrict__ ssim_err, float* __restrict__ tet_err,
                                                 int32_t* __restrict__ tet_count) {
    constexpr int TS = C::TS, NT = C::NT, B = C::B, CAP = C::CAP, SP = TS + 1;
    static_assert(C::R == 1 && B <= 64 && CAP % 32 == 0 && CAP / 32 <= NT && CAP >= TS * TS, "error pass geometry");
    __shared__ BatchLDS<TS, B, false> L;
    __shared__ float s_r[3][CAP];  // t_min, t_max, optical depth -> weight (0: none)
    __shared__ float s_ray[3][TS * TS];
    __shared__ float s_pe[TS * TS], s_se[TS * TS];
    __shared__ uint32_t s_live[TS * TS];            // inside the image, log T >= -5.5
    __shared__ unsigned long long s_mask[TS * TS];  // covered tets of the batch per pixel
    __shared__ chunk_t<1> s_chunk[chunk_capacity(CAP, 1, B)];
    __shared__ uint32_t s_start[CAP / 32];
    __shared__ float4 s_cen[B];
    __shared__ uint32_t s_tid[B];  // tet ids of the batch (the forward's LDS image has none)
    __shared__ float s_acc[B][kErrStats];
    __shared__ int32_t s_max[B];

    const uint32_t gtile = blockIdx.x;
    const ViewParams& vw = views[0];
    const TileCtx tc = tile_ctx(vw, gtile);
    const int tid = threadIdx.x, wid = tid / 64, lane = (int)__lane_id();
    const int row = tid / TS, colp = tid % TS;
    const uint32_t px = tc.tile_x * TS + colp, py = tc.tile_y * TS + row;
    const bool active = tid < TS * TS && px < vw.width && py < vw.height;
    const float min_t = vw.min_t;
    const V3 cam = mk3(vw.cam[0], vw.cam[1], vw.cam[2]);

    if (tid < TS * TS) {
        V3 rd = mk3(0.f, 0.f, 1.f);
        float pe = 0.f, se = 0.f;
        if (active) {
            rd = pixel_ray(vw, px, py);
            const uint64_t pix = (uint64_t)py * vw.width + px;
            pe = pixel_err[pix];
            se = ssim_err[pix];
        }
        s_ray[0][tid] = rd.x;
        s_ray[1][tid] = rd.y;
        s_ray[2][tid] = rd.z;
        s_pe[tid] = pe;
        s_se[tid] = se;
        s_live[tid] = active ? 1u : 0u;
    }
    float log_t = 0.f;

    const uint32_t rs = ranges[2ull * gtile], re = ranges[2ull * gtile + 1];
    const uint32_t count = re > rs ? re - rs : 0u;
    uint32_t done = 0;
    int nb = (int)min<uint32_t>(B, count);
    TetFetch f;
    QuadFetch qf{};
    if (C::PF == 1 && tid < nb) fetch_tet<false>(f, sorted_vals[re - 1 - tid], rec, nullptr, rec);
    if (C::PF == 3 && tid < 4 * nb) fetch_quad<false>(qf, sorted_vals[re - 1 - (tid >> 2)], tid & 3, rec, nullptr, rec);
    while (done < count) {
        if (__syncthreads_and(!active || log_t < -5.5f)) break;
        if constexpr (C::PF == 3) {
            store_quad<TS, B, false>(L, qf, tid >> 2, tid & 3, tid < 4 * nb, vw, tc);
            const V3 a = qperm3<0x00>(qf.r), b = qperm3<0x55>(qf.r), c = qperm3<0xAA>(qf.r), d = qperm3<0xFF>(qf.r);
            const V3 cen = scl3(add3(add3(add3(a, b), c), d), 0.25f);
            if (tid < 4 * nb && (tid & 3) == 0) {
                s_cen[tid >> 2] = make_float4(cen.x, cen.y, cen.z, 0.0f);
                s_tid[tid >> 2] = qf.t;
            }
        } else if (tid < nb) {
            store_tet<TS, B, false>(L, tid, f, vw, tc);
            const V3 a = mk3(f.r[0].x, f.r[0].y, f.r[0].z), b = mk3(f.r[1].x, f.r[1].y, f.r[1].z),
                     c = mk3(f.r[2].x, f.r[2].y, f.r[2].z), d = mk3(f.r[3].x, f.r[3].y, f.r[3].z);
            const V3 cen = scl3(add3(add3(add3(a, b), c), d), 0.25f);
            s_cen[tid] = make_float4(cen.x, cen.y, cen.z, 0.0f);
            s_tid[tid] = f.t;
        }
        if (tid < TS * TS) s_mask[tid] = 0ull;
        if (tid < CAP / 32) s_start[tid] = 0u;
        for (int i = tid; i < B * kErrStats; i += NT) (&s_acc[0][0])[i] = 0.0f;
        if (tid < B) s_max[tid] = 0;
        __syncthreads();
        batch_layout<TS, B, false, true, CAP, 1, NT>(L, nb, s_chunk, s_start);
        __syncthreads();
        const uint32_t nbp = L.nbp & 0xFFu;
        const uint32_t nrec = L.nbp >> 8;
        const uint32_t next = done + nbp;
        const int nb_next = (int)min<uint32_t>(B, count - next);
        if (C::PF == 1 && tid < nb_next) fetch_tet<false>(f, sorted_vals[re - 1 - next - tid], rec, nullptr, rec);
        if (C::PF == 3 && tid < 4 * nb_next)
            fetch_quad<false>(qf, sorted_vals[re - 1 - next - (tid >> 2)], tid & 3, rec, nullptr, rec);
        // phase A: slab and optical depth of every covered pair
        for (uint32_t j = tid; j < nrec; j += NT) {
            const ChunkRef c = record_at(s_chunk, s_start, j);
            const int p = c.row * TS + c.col;
            if (s_live[p]) {
                TetRegs T;
                tet_regs(L, c.k, T);
                const V3 rd = mk3(s_ray[0][p], s_ray[1][p], s_ray[2][p]);
                float t_min, t_max;
                int fmin, fmax;
                if (intersect(T, rd, min_t, t_min, t_max, fmin, fmax)) {
                    s_r[0][j] = t_min;
                    s_r[1][j] = t_max;
                    s_r[2][j] = med_clip(T.dens * (t_max - t_min), 1e-8f, 88.0f);
                    atomicOr(&s_mask[p], 1ull << c.k);
                }
            }
        }
        __syncthreads();
        // phase B: depth order per pixel; weight or 0 for every covered pair
        if (active) {
            unsigned long long m = s_mask[tid];
            while (m) {
                const uint32_t k = (uint32_t)__ffsll((long long)m) - 1u;
                m &= m - 1ull;
                const uint32_t slot = (L.rpref[k * SP + row] + (uint32_t)colp) & 0xFFFFu;
                float tv = 0.0f;
                if (log_t >= -5.5f) {
                    const float trans = safe_exp(log_t);
                    if (trans >= 1e-6f) {
                        const float od = s_r[2][slot];
                        tv = trans * od;
                        log_t = log_t - od;
                    }
                }
                s_r[2][slot] = tv;
            }
            if (log_t < -5.5f) s_live[tid] = 0u;
        }
        __syncthreads();
        // phase C: moments of the weighted pairs, summed per tet
        for (uint32_t base = (uint32_t)wid * 64; base < nrec; base += NT) {
            const uint32_t j = base + lane;
            float acc[kErrStats];
#pragma unroll
            for (int q = 0; q < kErrStats; ++q) acc[q] = 0.0f;
            int key = 0x7fffffff;
            if (j < nrec) {
                const ChunkRef c = record_at(s_chunk, s_start, j);
                key = c.k;
                const int p = c.row * TS + c.col;
                const float tv = (s_mask[p] >> c.k) & 1ull ? s_r[2][j] : 0.0f;
                if (tv > 0.0f) {
                    const V3 rd = mk3(s_ray[0][p], s_ray[1][p], s_ray[2][p]);
                    const V3 ro = add3(cam, scl3(rd, min_t));
                    const V3 en = add3(ro, scl3(rd, s_r[0][j])), ex = add3(ro, scl3(rd, s_r[1][j]));
                    const float pe = s_pe[p], se = s_se[p];
                    const float4 cq = s_cen[c.k];
                    acc[0] = tv;
                    acc[1] = tv * pe;
                    acc[2] = (tv * pe) * pe;
                    acc[3] = tv * se;
                    acc[4] = tv * en.x;
                    acc[5] = tv * en.y;
                    acc[6] = tv * en.z;
                    acc[7] = tv * ex.x;
                    acc[8] = tv * ex.y;
                    acc[9] = tv * ex.z;
                    acc[10] = tv;
                    acc[11] = tv * cq.x;
                    acc[12] = tv * cq.y;
                    acc[13] = tv * cq.z;
                    acc[14] = 1.0f;
                    atomicMax(&s_max[c.k], sat_i32(65535.0f * tv));
                }
            }
            const SegMasksDpp sm = seg_masks_dpp(key);
            const int next_key = dpp_i<0x130, 0xF>(key, -1);  // wave_shl:1 (lane 63 reads -1)
            const bool tail = key != 0x7fffffff && next_key != key;
            seg_scan_dpp<kErrStats>(acc, sm);
#pragma unroll
            for (int q = 0; q < kErrStats; ++q)
                if (tail && acc[q] != 0.0f) atomicAdd(&s_acc[key][q], acc[q]);
        }
        __syncthreads();
        // flush: one atomic per non-zero statistic per (tile, tet)
        for (int it = tid; it < (int)nbp * (kErrStats + 1); it += NT) {
            const int k = it / (kErrStats + 1), q = it % (kErrStats + 1);
            const uint32_t t = s_tid[k];
            if (q < kErrStats - 1) {
                const float v = s_acc[k][q];
                if (v != 0.0f) unsafeAtomicAdd(tet_err + 16ull * t + err_column(q), v);
            } else if (q == kErrStats - 1) {
                const int32_t n = (int32_t)s_acc[k][q];
                if (n) atomicAdd(tet_count + 2ull * t, n);
            } else if (s_max[k] > 0) {
                atomicMax(tet_count + 2ull * t + 1, s_max[k]);
            }
        }
        done = next;
        nb = nb_next;
    }
}

template <int TS>
static void err_launch(const ErrArgs& a, hipStream_t s) {
    using C = FwdCfg<TS, 32, (TS * TS > 512 ? TS * TS : 512), 1, 3>;
    hipLaunchKernelGGL((k_error<C>), dim3(a.num_tiles), dim3(C::NT), 0, s, a.rec, a.sorted_vals, a.ranges, a.views,
                       a.pixel_err, a.ssim_err, a.tet_err, a.tet_count);
}

bool launch_error(const ErrArgs& a, uint32_t tile_size, hipStream_t s) {
    switch (tile_size) {
        case 8: err_launch<8>(a, s); return true;
        case 12: err_launch<12>(a, s); return true;
        case 16: err_launch<16>(a, s); return true;
        case 32: err_launch<32>(a, s); return true;
        default: return false;
    }
}

// ---------------------------------------------------------------------------
// Interval-tiled forward (forward_interval_tiled, crates/rmesh-python/src/
// lib.rs:1218-1390).  Each tet's projection is split into screen-space fan
// triangles around one interior point (interval_generate.wgsl:174-408): with
// one vertex inside the triangle of the other three (case 1) the silhouette
// is that triangle and the interior point is the inner vertex, paired with
// the opposite face's depth at it; otherwise (case 2) the silhouette is the
// quad of the four vertices and the interior point is the crossing of its
// diagonals, carrying both diagonals' depths.  Per pixel the covering fan
// triangle's screen barycentrics interpolate front/back NDC z and colour
// offsets; the chord is the linearized z difference times the ray scale
// (interval_tiled_rasterize.wgsl:96-441).  Outputs: image, xyzd (x, y, z
// stay 0: no vertex normals on this path; w = depth) and the five-float
// distortion state.
// ---------------------------------------------------------------------------
constexpr float kIvlNear = 0.01f, kIvlFar = 1000.0f;  // lib.rs:1241-1242
constexpr int kIvlRec = 7;                            // float4 per tet

__device__ __forceinline__ float ivl_mix(float a, float b, float t) { return a * (1.0f - t) + b * t; }

// interval_generate.wgsl:112-126
__device__ __forceinline__ void ivl_bary(float ptx, float pty, float ax, float ay, float bx, float by, float cx,
                                         float cy, float& b0, float& b1, float& b2) {
    const float v0x = bx - ax, v0y = by - ay, v1x = cx - ax, v1y = cy - ay, v2x = ptx - ax, v2y = pty - ay;
    const float d00 = v0x * v0x + v0y * v0y, d01 = v0x * v1x + v0y * v1y, d11 = v1x * v1x + v1y * v1y;
    const float d20 = v2x * v0x + v2y * v0y, d21 = v2x * v1x + v2y * v1y;
    const float denom = d00 * d11 - d01 * d01;
    const float inv = fabsf(denom) > 1e-10f ? 1.0f / denom : 0.0f;
    b1 = (d11 * d20 - d01 * d21) * inv;
    b2 = (d00 * d21 - d01 * d20) * inv;
    b0 = 1.0f - b1 - b2;
}

// interval_generate.wgsl:129-139: parameter along a of the crossing with line b (-1: parallel)
__device__ __forceinline__ float ivl_line_t(float a1x, float a1y, float a2x, float a2y, float b1x, float b1y,
                                            float b2x, float b2y) {
    const float d1x = a2x - a1x, d1y = a2y - a1y, d2x = b2x - b1x, d2y = b2y - b1y;
    const float denom = d1x * d2y - d1y * d2x;
    if (fabsf(denom) < 1e-10f) return -1.0f;
    const float dx = b1x - a1x, dy = b1y - a1y;
    return (dx * d2y - dy * d2x) / denom;
}

// interval_tiled_rasterize.wgsl:90-97 (edges inclusive)
__device__ __forceinline__ bool ivl_in_tri(float ptx, float pty, float ax, float ay, float bx, float by, float cx,
                                           float cy) {
    const float d0 = (bx - ax) * (pty - ay) - (by - ay) * (ptx - ax);
    const float d1 = (cx - bx) * (pty - by) - (cy - by) * (ptx - bx);
    const float d2 = (ax - cx) * (pty - cy) - (ay - cy) * (ptx - cx);
    const bool has_neg = d0 < 0.0f || d1 < 0.0f || d2 < 0.0f;
    const bool has_pos = d0 > 0.0f || d1 > 0.0f || d2 > 0.0f;
    return !(has_neg && has_pos);
}

// a[i] for i in 0..3 (and c ? a : b) as bit masks: a select chain over one
// array's elements would be folded into an indexed load, which keeps the
// array in scratch memory.
__device__ __forceinline__ float sel4(const float* a, int i) {
    const uint32_t m0 = 0u - (uint32_t)(i == 0), m1 = 0u - (uint32_t)(i == 1), m2 = 0u - (uint32_t)(i == 2),
                   m3 = 0u - (uint32_t)(i == 3);
    return __uint_as_float((__float_as_uint(a[0]) & m0) | (__float_as_uint(a[1]) & m1) |
                           (__float_as_uint(a[2]) & m2) | (__float_as_uint(a[3]) & m3));
}
__device__ __forceinline__ float bsel(bool c, float a, float b) {
    const uint32_t m = 0u - (uint32_t)c;
    return __uint_as_float((__float_as_uint(a) & m) | (__float_as_uint(b) & ~m));
}

// Screen-space geometry of one tet's decomposition: pixel positions, NDC z,
// 1 / w and colour offsets of the vertices, and the discrete choices with
// their weights (interval_generate.wgsl:174-300).  False: the tet is skipped
// (a vertex behind the camera, a flat depth range or density < 1e-6).
struct IvlGeo {
    float px[4], py[4], nz[4], iw[4], off[4];
    int inner;               // vertex inside the triangle of the other three, 4: none (quad)
    int f0, f1, f2, ce;      // 3 points: the face opposite the inner vertex, and that vertex
    float w0, w1, w2;        // the inner vertex's barycentrics in that face
    int a0, a1, b0, b1;      // quad: crossing edges a0-a1 and b0-b1
    float ta, tb;            // the crossing's parameter along each edge
};

__device__ __forceinline__ bool ivl_geometry(const V3 v[4], V3 g, float dens, const ViewParams& vw, IvlGeo& G) {
    const V3 cam = mk3(vw.cam[0], vw.cam[1], vw.cam[2]);
    float clip[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i) mul_vp(vw.vp, v[i], clip[i]);
    bool skip = clip[0][3] <= 0.0f || clip[1][3] <= 0.0f || clip[2][3] <= 0.0f || clip[3][3] <= 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float w = clip[i][3];
        G.iw[i] = fabsf(w) > 1e-10f ? 1.0f / w : 0.0f;
        const float nx = clip[i][0] * G.iw[i], ny = clip[i][1] * G.iw[i];
        G.nz[i] = clip[i][2] * G.iw[i];
        G.px[i] = ((nx + 1.0f) * 0.5f) * vw.Wf;
        G.py[i] = ((1.0f - ny) * 0.5f) * vw.Hf;
        G.off[i] = dot3(g, sub3(v[i], cam));
    }
    const float zmax = fmaxf(fmaxf(G.nz[0], G.nz[1]), fmaxf(G.nz[2], G.nz[3]));
    const float zmin = fminf(fminf(G.nz[0], G.nz[1]), fminf(G.nz[2], G.nz[3]));
    if (skip || zmax - zmin < 1e-7f || dens < 1e-6f) return false;
    const float* px = G.px;
    const float* py = G.py;
    // interval_generate.wgsl:89-109: a vertex inside the triangle of the other three
    G.inner = 4;
#pragma unroll
    for (int fi = 3; fi >= 0; --fi) {
        int a, b, c, q;
        face_idx(fi, a, b, c, q);
        const float d0 = (px[b] - px[a]) * (py[q] - py[a]) - (py[b] - py[a]) * (px[q] - px[a]);
        const float d1 = (px[c] - px[b]) * (py[q] - py[b]) - (py[c] - py[b]) * (px[q] - px[b]);
        const float d2 = (px[a] - px[c]) * (py[q] - py[c]) - (py[a] - py[c]) * (px[q] - px[c]);
        const bool all_pos = d0 >= 0.0f && d1 >= 0.0f && d2 >= 0.0f;
        const bool all_neg = d0 <= 0.0f && d1 <= 0.0f && d2 <= 0.0f;
        if (all_pos || all_neg) G.inner = q;  // descending order: the first face in table order wins
    }
    if (G.inner < 4) {
        face_idx((G.inner + 1) & 3, G.f0, G.f1, G.f2, G.ce);  // the face opposite the inner vertex
        ivl_bary(sel4(px, G.ce), sel4(py, G.ce), sel4(px, G.f0), sel4(py, G.f0), sel4(px, G.f1), sel4(py, G.f1),
                 sel4(px, G.f2), sel4(py, G.f2), G.w0, G.w1, G.w2);
    } else {
        // the first pair of opposite edges whose segments cross (01/23, 02/13,
        // 03/12), else 01/23
        G.a0 = 0; G.a1 = 1; G.b0 = 2; G.b1 = 3;
        G.ta = ivl_line_t(px[0], py[0], px[1], py[1], px[2], py[2], px[3], py[3]);
        G.tb = ivl_line_t(px[2], py[2], px[3], py[3], px[0], py[0], px[1], py[1]);
        if (!(G.ta > 0.0f && G.ta < 1.0f && G.tb > 0.0f && G.tb < 1.0f)) {
            const float ua = ivl_line_t(px[0], py[0], px[2], py[2], px[1], py[1], px[3], py[3]);
            const float ub = ivl_line_t(px[1], py[1], px[3], py[3], px[0], py[0], px[2], py[2]);
            if (ua > 0.0f && ua < 1.0f && ub > 0.0f && ub < 1.0f) {
                G.a1 = 2; G.b0 = 1; G.ta = ua; G.tb = ub;
            } else {
                const float wa = ivl_line_t(px[0], py[0], px[3], py[3], px[1], py[1], px[2], py[2]);
                const float wb = ivl_line_t(px[1], py[1], px[2], py[2], px[0], py[0], px[3], py[3]);
                if (wa > 0.0f && wa < 1.0f && wb > 0.0f && wb < 1.0f) {
                    G.a1 = 3; G.b0 = 1; G.b1 = 2; G.ta = wa; G.tb = wb;
                }
            }
        }
    }
    return true;
}

// One thread per tet: the silhouette decomposition, as 7 float4
//   [0..3] (x, y, ndc z, colour offset) of silhouette points 0-3 (3 points:
//          point 3 repeats point 0), pixel coordinates of the full image,
//   [4]    (x, y, front z, back z) of the interior point,
//   [5]    (front offset, back offset, density, point count: 0 = skipped),
//   [6]    (colour + g . (cam - v0), 0).
__global__ __launch_bounds__(256) void k_interval_generate(const float4* __restrict__ rec,
                                                           const float4* __restrict__ colors,
                                                           const ViewParams* __restrict__ views, uint32_t M,
                                                           float4* __restrict__ out) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M) return;
    const ViewParams& vw = views[0];
    float4 r[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = rec[4ull * t + i];
    const float4 col = colors[t];
    const V3 v[4] = {mk3(r[0].x, r[0].y, r[0].z), mk3(r[1].x, r[1].y, r[1].z), mk3(r[2].x, r[2].y, r[2].z),
                     mk3(r[3].x, r[3].y, r[3].z)};
    const float dens = r[0].w;
    const V3 g = mk3(r[1].w, r[2].w, r[3].w);
    const V3 cam = mk3(vw.cam[0], vw.cam[1], vw.cam[2]);
    float4* o = out + (uint64_t)kIvlRec * t;
    const float bo = dot3(g, sub3(cam, v[0]));
    const float4 base = make_float4(col.x + bo, col.y + bo, col.z + bo, 0.0f);
    IvlGeo G;
    if (!ivl_geometry(v, g, dens, vw, G)) {
        const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
        for (int i = 0; i < 6; ++i) o[i] = zero;
        o[6] = base;
        return;
    }
    const float* px = G.px;
    const float* py = G.py;
    const float* nz = G.nz;
    const float* off = G.off;
    int s0, s1, s2, s3;  // silhouette points
    float cx, cy, zf4, zb4, of4, ob4;
    if (G.inner < 4) {
        const float z_face = G.w0 * sel4(nz, G.f0) + G.w1 * sel4(nz, G.f1) + G.w2 * sel4(nz, G.f2);
        const float z_c = sel4(nz, G.ce);
        const float face_off = G.w0 * sel4(off, G.f0) + G.w1 * sel4(off, G.f1) + G.w2 * sel4(off, G.f2);
        const bool face_first = z_face <= z_c;
        s0 = G.f0; s1 = G.f1; s2 = G.f2; s3 = G.f0;
        cx = sel4(px, G.ce);
        cy = sel4(py, G.ce);
        zf4 = fminf(z_face, z_c);
        zb4 = fmaxf(z_face, z_c);
        of4 = face_first ? face_off : sel4(off, G.ce);
        ob4 = face_first ? sel4(off, G.ce) : face_off;
    } else {
        cx = ivl_mix(sel4(px, G.a0), sel4(px, G.a1), G.ta);
        cy = ivl_mix(sel4(py, G.a0), sel4(py, G.a1), G.ta);
        const float z_a = ivl_mix(sel4(nz, G.a0), sel4(nz, G.a1), G.ta);
        const float z_b = ivl_mix(sel4(nz, G.b0), sel4(nz, G.b1), G.tb);
        const float off_a = ivl_mix(sel4(off, G.a0), sel4(off, G.a1), G.ta);
        const float off_b = ivl_mix(sel4(off, G.b0), sel4(off, G.b1), G.tb);
        s0 = G.a0; s1 = G.b0; s2 = G.a1; s3 = G.b1;
        zf4 = fminf(z_a, z_b);
        zb4 = fmaxf(z_a, z_b);
        of4 = z_a <= z_b ? off_a : off_b;
        ob4 = z_a <= z_b ? off_b : off_a;
    }
    const int sv[4] = {s0, s1, s2, s3};
#pragma unroll
    for (int k = 0; k < 4; ++k)
        o[k] = make_float4(sel4(px, sv[k]), sel4(py, sv[k]), sel4(nz, sv[k]), sel4(off, sv[k]));
    o[4] = make_float4(cx, cy, zf4, zb4);
    o[5] = make_float4(of4, ob4, dens, G.inner < 4 ? 3.0f : 4.0f);
    o[6] = base;
}

// Tiled compositing over the fan triangles: the batch machinery of the
// tiled forward (spans of the tet's four projected points, which bound the
// silhouette; pixels of a span outside it find no fan triangle), with one
// record per covered pixel holding the colour, optical depth, depth and the
// distortion factors, composited per pixel in depth order.
template <class C>
__global__ __launch_bounds__(C::NT) __attribute__((amdgpu_waves_per_eu(C::WPE))) void k_interval_forward(const float4* __restrict__ ivr,
                                                           const uint32_t* __restrict__ sorted_vals,
                                                           const uint32_t* __restrict__ ranges,
                                                           const ViewParams* __restrict__ views,
                                                           const uint32_t* __restrict__ tile_order,
                                                           float* __restrict__ rgba, float* __restrict__ xyzd,
                                                           float* __restrict__ dist) {
    constexpr int TS = C::TS, NT = C::NT, B = C::B, CAP = C::CAP, SP = TS + 1;
    static_assert(B <= 64 && CAP >= TS * TS && CAP % 32 == 0 && CAP / 32 <= NT && CAP <= 2048, "batch geometry");
    static_assert(BatchLDS<TS, B, false>::TR == kIvlRec, "records live in the face rows");
    __shared__ BatchLDS<TS, B, false> L;
    __shared__ float s_r[8][CAP];  // colour rgb, optical depth, depth, 1 - e^-u, m1 / T, self term / T^2
    __shared__ float s_rsc[TS * TS];
    __shared__ uint32_t s_live[TS * TS];
    __shared__ unsigned long long s_mask[TS * TS];
    __shared__ uint16_t s_chunk[chunk_capacity(CAP, 1, B)];
    __shared__ uint32_t s_start[CAP / 32];

    const uint32_t gtile = tile_order ? tile_order[blockIdx.x] : blockIdx.x;
    const ViewParams& vw = views[0];
    const TileCtx tc = tile_ctx(vw, gtile);
    const int tid = threadIdx.x;
    const int row = tid / TS, colp = tid % TS;
    const uint32_t px = tc.tile_x * TS + colp, py = tc.tile_y * TS + row;
    const bool active = tid < TS * TS && px < vw.width && py < vw.height;
    if (tid < TS * TS) {
        float rs = 1.0f;
        if (active) {
            // interval_tiled_rasterize.wgsl:300-306: |(x_cam, y_cam, 1)|
            const float xc = ((float)px + 0.5f - vw.intr[2]) / vw.intr[0];
            const float yc = ((float)py + 0.5f - vw.intr[3]) / vw.intr[1];
            rs = sqrtf((xc * xc + yc * yc) + 1.0f);
        }
        s_rsc[tid] = rs;
        s_live[tid] = active ? 1u : 0u;
    }
    float cr = 0.f, cg = 0.f, cb = 0.f, log_t = 0.f, dep = 0.f;
    float ds0 = 0.f, ds1 = 0.f, ds2 = 0.f, ds3 = 0.f, ds4 = 0.f;
    const float nf = kIvlNear * kIvlFar, range_nf = kIvlFar - kIvlNear;

    const uint32_t rs_ = ranges[2ull * gtile], re = ranges[2ull * gtile + 1];
    const uint32_t count = re > rs_ ? re - rs_ : 0u;
    uint32_t done = 0;
    int nb = (int)min<uint32_t>(B, count);
    float4 f[kIvlRec];
    auto fetch = [&](uint32_t t) {
#pragma unroll
        for (int i = 0; i < kIvlRec; ++i) f[i] = ivr[(uint64_t)kIvlRec * t + i];
    };
    if (tid < nb) fetch(sorted_vals[re - 1 - tid]);
    while (done < count) {
        if (__syncthreads_and(!active || log_t < -13.8f)) break;
        if (tid < nb) {
            const int k = tid;
#pragma unroll
            for (int i = 0; i < kIvlRec; ++i) L.tr[k * kIvlRec + i] = f[i];
            // the tet's four projected points: silhouette 0-2 and the inner
            // vertex (3 points) or silhouette 0-3
            const bool tri = f[5].w == 3.0f;
            const float qx[4] = {f[0].x, f[1].x, f[2].x, tri ? f[4].x : f[3].x};
            const float qy[4] = {f[0].y, f[1].y, f[2].y, tri ? f[4].y : f[3].y};
            uint32_t rr = 0xFFu;
            if (f[5].w != 0.0f) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    L.pxy[i][k] = qx[i] - tc.tile_ox;
                    L.pxy[4 + i][k] = qy[i] - tc.tile_oy;
                }
                const float ymin = fminf(fminf(qy[0], qy[1]), fminf(qy[2], qy[3])) - tc.tile_oy;
                const float ymax = fmaxf(fmaxf(qy[0], qy[1]), fmaxf(qy[2], qy[3])) - tc.tile_oy;
                const float r0 = fmaxf(ceilf(ymin) - 1.0f, 0.0f);
                const float r1 = fminf(floorf(ymax), (float)(TS - 1));
                if (r0 <= r1) rr = (uint32_t)r0 | ((uint32_t)r1 << 8);
            }
            L.rows[k] = rr;
        }
        if (tid < TS * TS) s_mask[tid] = 0ull;
        if (tid < CAP / 32) s_start[tid] = 0u;
        __syncthreads();
        batch_layout<TS, B, false, true, CAP, 1, NT>(L, nb, s_chunk, s_start);
        __syncthreads();
        const uint32_t nbp = L.nbp & 0xFFu;
        const uint32_t nrec = L.nbp >> 8;
        const uint32_t next = done + nbp;
        const int nb_next = (int)min<uint32_t>(B, count - next);
        if (tid < nb_next) fetch(sorted_vals[re - 1 - next - tid]);
        // record-parallel: fan triangle, interpolation and the per-pair terms
        // (interval_tiled_rasterize.wgsl:248-400)
        for (uint32_t j = tid; j < nrec; j += NT) {
            const ChunkRef c = record_at(s_chunk, s_start, j);
            const int p = c.row * TS + c.col;
            if (!s_live[p]) continue;
            const float4* T = &L.tr[c.k * kIvlRec];
            const float4 q0 = T[0], q1 = T[1], q2 = T[2], q3 = T[3], q4 = T[4], q5 = T[5];
            const bool tri3 = q5.w == 3.0f;
            const float pcx = (float)(tc.tile_x * TS + (uint32_t)c.col) + 0.5f;
            const float pcy = (float)(tc.tile_y * TS + (uint32_t)c.row) + 0.5f;
            // first fan triangle (point i, point i + 1 mod n, interior point) holding the pixel centre
            float4 ea = q0, eb = q1;
            bool found = ivl_in_tri(pcx, pcy, q0.x, q0.y, q1.x, q1.y, q4.x, q4.y);
            if (!found && ivl_in_tri(pcx, pcy, q1.x, q1.y, q2.x, q2.y, q4.x, q4.y)) {
                found = true;
                ea = q1;
                eb = q2;
            }
            const float4 e3 = tri3 ? q0 : q3;
            if (!found && ivl_in_tri(pcx, pcy, q2.x, q2.y, e3.x, e3.y, q4.x, q4.y)) {
                found = true;
                ea = q2;
                eb = e3;
            }
            if (!found && !tri3 && ivl_in_tri(pcx, pcy, q3.x, q3.y, q0.x, q0.y, q4.x, q4.y)) {
                found = true;
                ea = q3;
                eb = q0;
            }
            if (!found) continue;
            float b0, b1, b2;
            ivl_bary(pcx, pcy, ea.x, ea.y, eb.x, eb.y, q4.x, q4.y, b0, b1, b2);
            const float zfn = b0 * ea.z + b1 * eb.z + b2 * q4.z;
            const float zbn = b0 * ea.z + b1 * eb.z + b2 * q4.w;
            const float ofn = b0 * ea.w + b1 * eb.w + b2 * q5.x;
            const float obn = b0 * ea.w + b1 * eb.w + b2 * q5.y;
            const float zf = nf / (kIvlFar - clampf(zfn, 0.0f, 1.0f) * range_nf);
            const float zb = nf / (kIvlFar - clampf(zbn, 0.0f, 1.0f) * range_nf);
            const float rsc = s_rsc[p];
            const float dd = fabsf(zb - zf) * rsc;
            const float4 base = T[6];
            const float dens = q5.z;
            const V3 cf = mk3(fmaxf(base.x + ofn, 0.0f), fmaxf(base.y + ofn, 0.0f), fmaxf(base.z + ofn, 0.0f));
            const V3 cbk = mk3(fmaxf(base.x + obn, 0.0f), fmaxf(base.y + obn, 0.0f), fmaxf(base.z + obn, 0.0f));
            const float od = clampf(dens * dd, 1e-8f, 88.0f);
            const float alpha_t = safe_exp(-od);
            const float phv = phi_f(od);
            const float w0 = phv - alpha_t, w1 = 1.0f - phv;
            const V3 cp = safe_clip3(mk3(cbk.x * w0 + cf.x * w1, cbk.y * w0 + cf.y * w1, cbk.z * w0 + cf.z * w1),
                                     0.0f, kMaxVal);
            // distortion factors (interval_tiled_rasterize.wgsl:383-400)
            const float enter = zf * rsc, exit_d = zb * rsc;
            const float u = dens * (exit_d - enter);
            const float en = safe_exp(-u);
            const float ad = fmaxf(1.0f - en, 0.0f);
            s_r[0][j] = cp.x;
            s_r[1][j] = cp.y;
            s_r[2][j] = cp.z;
            s_r[3][j] = od;
            s_r[4][j] = w0 * zb + w1 * zf;
            s_r[5][j] = ad;
            s_r[6][j] = enter * ad + safe_div(1.0f - (1.0f + u) * en, dens);
            s_r[7][j] = safe_div(1.0f - 2.0f * u * en - en * en, dens);
            atomicOr(&s_mask[p], 1ull << c.k);
        }
        __syncthreads();
        // pixel-sequential compositing in depth order (interval_tiled_rasterize.wgsl:350-401)
        if (active) {
            unsigned long long m = s_mask[tid];
            while (m) {
                const uint32_t k = (uint32_t)__ffsll((long long)m) - 1u;
                m &= m - 1ull;
                if (log_t < -13.8f) break;
                const uint32_t slot = (L.rpref[k * SP + row] + (uint32_t)colp) & 0xFFFFu;
                const float Tj = safe_exp(log_t);
                if (Tj >= 1e-6f) {
                    cr = cr + s_r[0][slot] * Tj;
                    cg = cg + s_r[1][slot] * Tj;
                    cb = cb + s_r[2][slot] * Tj;
                    log_t = log_t - s_r[3][slot];
                    dep = dep + Tj * s_r[4][slot];
                    const float m0 = Tj * s_r[5][slot];
                    const float m1 = Tj * s_r[6][slot];
                    const float sd = (Tj * Tj) * s_r[7][slot];
                    ds0 = ds0 + 2.0f * m1 * ds2;
                    ds1 = ds1 + 2.0f * m0 * ds3;
                    ds2 = ds2 + m0;
                    ds3 = ds3 + m1;
                    ds4 = ds4 + sd;
                }
            }
            if (log_t < -13.8f) s_live[tid] = 0u;
        }
        done = next;
        nb = nb_next;
    }
    if (active) {
        const uint64_t pix = (uint64_t)py * vw.width + px;
        reinterpret_cast<float4*>(rgba)[pix] = make_float4(cr, cg, cb, 1.0f - safe_exp(log_t));
        reinterpret_cast<float4*>(xyzd)[pix] = make_float4(0.0f, 0.0f, 0.0f, dep);
        float* d = dist + 5ull * pix;
        d[0] = ds0;
        d[1] = ds1;
        d[2] = ds2;
        d[3] = ds3;
        d[4] = ds4;
    }
}

// Per-pair quantities of one covered pixel (interval_tiled_rasterize.wgsl:
// 248-340, backward_interval_tiled.wgsl:360-430): the fan triangle (slots ai,
// bi of the silhouette and the centre) and barycentrics, the interpolated and
// linearized depths, colours, optical depth and the integral weights.
struct IvlPair {
    float b0, b1, b2;
    int ai, bi;
    float zfn, zbn, zfc, zbc, z_f, z_b, dist, od, alpha_t, w0, w1;
    V3 cf_raw, cb_raw, cf, cb, cp;
};

__device__ __forceinline__ bool ivl_pair(const float4* __restrict__ T, float pcx, float pcy, float rsc, IvlPair& q) {
    const float4 q0 = T[0], q1 = T[1], q2 = T[2], q3 = T[3], q4 = T[4], q5 = T[5];
    const bool tri3 = q5.w == 3.0f;
    float4 ea = q0, eb = q1;
    q.ai = 0;
    q.bi = 1;
    bool found = ivl_in_tri(pcx, pcy, q0.x, q0.y, q1.x, q1.y, q4.x, q4.y);
    if (!found && ivl_in_tri(pcx, pcy, q1.x, q1.y, q2.x, q2.y, q4.x, q4.y)) {
        found = true;
        ea = q1; eb = q2; q.ai = 1; q.bi = 2;
    }
    const float4 e3 = tri3 ? q0 : q3;
    if (!found && ivl_in_tri(pcx, pcy, q2.x, q2.y, e3.x, e3.y, q4.x, q4.y)) {
        found = true;
        ea = q2; eb = e3; q.ai = 2; q.bi = tri3 ? 0 : 3;
    }
    if (!found && !tri3 && ivl_in_tri(pcx, pcy, q3.x, q3.y, q0.x, q0.y, q4.x, q4.y)) {
        found = true;
        ea = q3; eb = q0; q.ai = 3; q.bi = 0;
    }
    if (!found) return false;
    ivl_bary(pcx, pcy, ea.x, ea.y, eb.x, eb.y, q4.x, q4.y, q.b0, q.b1, q.b2);
    q.zfn = q.b0 * ea.z + q.b1 * eb.z + q.b2 * q4.z;
    q.zbn = q.b0 * ea.z + q.b1 * eb.z + q.b2 * q4.w;
    const float ofn = q.b0 * ea.w + q.b1 * eb.w + q.b2 * q5.x;
    const float obn = q.b0 * ea.w + q.b1 * eb.w + q.b2 * q5.y;
    const float nf = kIvlNear * kIvlFar, range_nf = kIvlFar - kIvlNear;
    q.zfc = clampf(q.zfn, 0.0f, 1.0f);
    q.zbc = clampf(q.zbn, 0.0f, 1.0f);
    q.z_f = nf / (kIvlFar - q.zfc * range_nf);
    q.z_b = nf / (kIvlFar - q.zbc * range_nf);
    q.dist = fabsf(q.z_b - q.z_f) * rsc;
    const float4 base = T[6];
    const float dens = q5.z;
    q.cf_raw = mk3(base.x + ofn, base.y + ofn, base.z + ofn);
    q.cb_raw = mk3(base.x + obn, base.y + obn, base.z + obn);
    q.cf = mk3(fmaxf(q.cf_raw.x, 0.0f), fmaxf(q.cf_raw.y, 0.0f), fmaxf(q.cf_raw.z, 0.0f));
    q.cb = mk3(fmaxf(q.cb_raw.x, 0.0f), fmaxf(q.cb_raw.y, 0.0f), fmaxf(q.cb_raw.z, 0.0f));
    q.od = clampf(dens * q.dist, 1e-8f, 88.0f);
    q.alpha_t = safe_exp(-q.od);
    const float phv = phi_f(q.od);
    q.w0 = phv - q.alpha_t;
    q.w1 = 1.0f - phv;
    q.cp = safe_clip3(mk3(q.cb.x * q.w0 + q.cf.x * q.w1, q.cb.y * q.w0 + q.cf.y * q.w1, q.cb.z * q.w0 + q.cf.z * q.w1),
                      0.0f, kMaxVal);
    return true;
}

// Per-tet adjoint accumulators of the interval decomposition, in floats:
//   [0..19]  slot s = 0..4: d front z, d back z, d front offset, d back offset
//   [20..22] d base colour, [23] d density,
//   [24..35] silhouette slots 0-3: d normal (front + back),
//   [36..41] centre: d front normal, d back normal.
constexpr int kIvlGradTerms = 24, kIvlNormalTerms = 18, kIvlGradStride = 48;

// Per-tile adjoint (backward_interval_tiled.wgsl:165-720), back to front in
// batches: phase A evaluates each covered pair, phase B undoes the pixel's
// colour / transmittance / distortion state in order and keeps the pair's
// T_prev, the transmittance gradient before it and the m0 gradient, phase C
// forms the pair's terms and reduces them per tet (segmented wave scan, LDS
// adds, one global add per term per (tile, tet)).
template <class C>
__global__ __launch_bounds__(C::NT) __attribute__((amdgpu_waves_per_eu(C::WPE))) void k_interval_backward(
    const float4* __restrict__ ivr, const uint32_t* __restrict__ sorted_vals, const uint32_t* __restrict__ ranges,
    const ViewParams* __restrict__ views, const uint32_t* __restrict__ tile_order, const float* __restrict__ rgba,
    const float* __restrict__ dist_img, const float* __restrict__ dl_img, const float* __restrict__ dl_xyzd,
    const float* __restrict__ dl_dist, float* __restrict__ d_ivr) {
    constexpr int TS = C::TS, NT = C::NT, B = C::B, CAP = C::CAP, SP = TS + 1;
    constexpr int KT = kIvlGradTerms + kIvlNormalTerms;
    static_assert(B <= 64 && CAP >= TS * TS && CAP % 32 == 0 && CAP / 32 <= NT && CAP <= 2048, "batch geometry");
    __shared__ BatchLDS<TS, B, false> L;
    __shared__ float s_r[6][CAP];  // A: colour rgb, od, 1 - e^-u, m1 / T; B: T_prev, -d log T, d m0
    __shared__ float s_rsc[TS * TS];
    __shared__ float s_dl[7][TS * TS];  // d colour rgb, d xyzd
    __shared__ uint32_t s_live[TS * TS];
    __shared__ unsigned long long s_mask[TS * TS];
    __shared__ uint16_t s_chunk[chunk_capacity(CAP, 1, B)];
    __shared__ uint32_t s_start[CAP / 32];
    __shared__ float s_acc[B][KT];
    __shared__ uint32_t s_tid[B];

    const uint32_t gtile = tile_order ? tile_order[blockIdx.x] : blockIdx.x;
    const ViewParams& vw = views[0];
    const TileCtx tc = tile_ctx(vw, gtile);
    const int tid = threadIdx.x, lane = (int)__lane_id(), wid = tid / 64;
    const int row = tid / TS, colp = tid % TS;
    const uint32_t px = tc.tile_x * TS + colp, py = tc.tile_y * TS + row;
    const bool active = tid < TS * TS && px < vw.width && py < vw.height;
    // pixel state of the forward's end (backward_interval_tiled.wgsl:196-225)
    float log_t = 0.f, d_log_t = 0.f, st_z = 0.f, st_w = 0.f, d_z = 0.f, d_w = 0.f, d_x = 0.f, d_y = 0.f;
    float dl_r = 0.f, dl_g = 0.f, dl_b = 0.f;
    bool has_dn = false;
    if (tid < TS * TS) {
        float rs = 1.0f;
        float dx[4] = {0.f, 0.f, 0.f, 0.f};
        if (active) {
            const uint64_t pix = (uint64_t)py * vw.width + px;
            const float xc = ((float)px + 0.5f - vw.intr[2]) / vw.intr[0];
            const float yc = ((float)py + 0.5f - vw.intr[3]) / vw.intr[1];
            rs = sqrtf((xc * xc + yc * yc) + 1.0f);
            const float4 img = reinterpret_cast<const float4*>(rgba)[pix];
            const float4 dl = reinterpret_cast<const float4*>(dl_img)[pix];
            log_t = logf(fmaxf(1.0f - img.w, 1e-20f));
            d_log_t = -dl.w * (1.0f - img.w);
            dl_r = dl.x;
            dl_g = dl.y;
            dl_b = dl.z;
            st_z = dist_img[5 * pix + 2];
            st_w = dist_img[5 * pix + 3];
            if (dl_dist) {
                d_x = dl_dist[5 * pix];
                d_y = dl_dist[5 * pix + 1];
                d_z = dl_dist[5 * pix + 2];
                d_w = dl_dist[5 * pix + 3];
            }
            if (dl_xyzd) {
                const float4 q = reinterpret_cast<const float4*>(dl_xyzd)[pix];
                dx[0] = q.x; dx[1] = q.y; dx[2] = q.z; dx[3] = q.w;
            }
        }
        s_rsc[tid] = rs;
        s_live[tid] = active ? 1u : 0u;
        s_dl[0][tid] = dl_r;
        s_dl[1][tid] = dl_g;
        s_dl[2][tid] = dl_b;
#pragma unroll
        for (int i = 0; i < 4; ++i) s_dl[3 + i][tid] = dx[i];
        has_dn = dx[0] != 0.0f || dx[1] != 0.0f || dx[2] != 0.0f;
    }
    // the normal terms are formed only where some pixel of the tile has a normal gradient
    has_dn = __syncthreads_or(has_dn);

    const uint32_t rs_ = ranges[2ull * gtile], re = ranges[2ull * gtile + 1];
    const uint32_t count = re > rs_ ? re - rs_ : 0u;
    uint32_t done = 0;
    int nb = (int)min<uint32_t>(B, count);
    float4 f[kIvlRec];
    uint32_t ft = 0;
    auto fetch = [&](uint32_t t) {
        ft = t;
#pragma unroll
        for (int i = 0; i < kIvlRec; ++i) f[i] = ivr[(uint64_t)kIvlRec * t + i];
    };
    if (tid < nb) fetch(sorted_vals[rs_ + tid]);
    while (done < count) {
        __syncthreads();  // the previous batch's flush has read s_acc / s_tid
        if (tid < nb) {
            const int k = tid;
#pragma unroll
            for (int i = 0; i < kIvlRec; ++i) L.tr[k * kIvlRec + i] = f[i];
            s_tid[k] = ft;
            const bool tri = f[5].w == 3.0f;
            const float qx[4] = {f[0].x, f[1].x, f[2].x, tri ? f[4].x : f[3].x};
            const float qy[4] = {f[0].y, f[1].y, f[2].y, tri ? f[4].y : f[3].y};
            uint32_t rr = 0xFFu;
            if (f[5].w != 0.0f) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    L.pxy[i][k] = qx[i] - tc.tile_ox;
                    L.pxy[4 + i][k] = qy[i] - tc.tile_oy;
                }
                const float ymin = fminf(fminf(qy[0], qy[1]), fminf(qy[2], qy[3])) - tc.tile_oy;
                const float ymax = fmaxf(fmaxf(qy[0], qy[1]), fmaxf(qy[2], qy[3])) - tc.tile_oy;
                const float r0 = fmaxf(ceilf(ymin) - 1.0f, 0.0f);
                const float r1 = fminf(floorf(ymax), (float)(TS - 1));
                if (r0 <= r1) rr = (uint32_t)r0 | ((uint32_t)r1 << 8);
            }
            L.rows[k] = rr;
        }
        if (tid < TS * TS) s_mask[tid] = 0ull;
        if (tid < CAP / 32) s_start[tid] = 0u;
        for (int i = tid; i < B * KT; i += NT) (&s_acc[0][0])[i] = 0.0f;
        __syncthreads();
        batch_layout<TS, B, false, true, CAP, 1, NT>(L, nb, s_chunk, s_start);
        __syncthreads();
        const uint32_t nbp = L.nbp & 0xFFu;
        const uint32_t nrec = L.nbp >> 8;
        const uint32_t next = done + nbp;
        const int nb_next = (int)min<uint32_t>(B, count - next);
        if (tid < nb_next) fetch(sorted_vals[rs_ + next + tid]);
        // phase A: the pair's colour, optical depth and distortion factors
        for (uint32_t j = tid; j < nrec; j += NT) {
            const ChunkRef c = record_at(s_chunk, s_start, j);
            const int p = c.row * TS + c.col;
            if (!s_live[p]) continue;
            const float4* T = &L.tr[c.k * kIvlRec];
            const float pcx = (float)(tc.tile_x * TS + (uint32_t)c.col) + 0.5f;
            const float pcy = (float)(tc.tile_y * TS + (uint32_t)c.row) + 0.5f;
            const float rsc = s_rsc[p];
            IvlPair q;
            if (!ivl_pair(T, pcx, pcy, rsc, q)) continue;
            const float dens = T[5].z;
            const float enter = q.z_f * rsc, exit_d = q.z_b * rsc;
            const float u = dens * (exit_d - enter);
            const float en = safe_exp(-u);
            const float ad = fmaxf(1.0f - en, 0.0f);
            s_r[0][j] = q.cp.x;
            s_r[1][j] = q.cp.y;
            s_r[2][j] = q.cp.z;
            s_r[3][j] = q.od;
            s_r[4][j] = ad;
            s_r[5][j] = enter * ad + safe_div(1.0f - (1.0f + u) * en, dens);
            atomicOr(&s_mask[p], 1ull << c.k);
        }
        __syncthreads();
        // phase B: undo back to front (backward_interval_tiled.wgsl:432-460, 602-630)
        if (active) {
            unsigned long long m = s_mask[tid];
            while (m) {
                const uint32_t k = (uint32_t)__ffsll((long long)m) - 1u;
                m &= m - 1ull;
                const uint32_t slot = (L.rpref[k * SP + row] + (uint32_t)colp) & 0xFFFFu;
                const float cr = s_r[0][slot], cg = s_r[1][slot], cb = s_r[2][slot], od = s_r[3][slot];
                const float prev = fminf(log_t + od, 0.0f);
                const float tp = safe_exp(prev);
                log_t = prev;
                if (tp < 1e-6f) {
                    s_r[0][slot] = 0.0f;
                    continue;
                }
                const float dos = -d_log_t;
                d_log_t = safe_clip(d_log_t + ((dl_r * cr + dl_g * cg) + dl_b * cb) * tp, kMinVal, kMaxVal);
                const float m0 = tp * s_r[4][slot];
                const float m1 = tp * s_r[5][slot];
                st_z = st_z - m0;
                st_w = st_w - m1;
                const float dm0 = d_z + 2.0f * st_w * d_y;
                d_z = d_z + 2.0f * m1 * d_x;
                d_w = d_w + 2.0f * m0 * d_y;
                s_r[0][slot] = tp;
                s_r[1][slot] = dos;
                s_r[2][slot] = dm0;
            }
        }
        __syncthreads();
        // phase C: the pair's terms, summed per tet (backward_interval_tiled.wgsl:460-660)
        for (uint32_t base = (uint32_t)wid * 64; base < nrec; base += NT) {
            const uint32_t j = base + lane;
            float acc[kIvlGradTerms], accn[kIvlNormalTerms];
#pragma unroll
            for (int i = 0; i < kIvlGradTerms; ++i) acc[i] = 0.0f;
#pragma unroll
            for (int i = 0; i < kIvlNormalTerms; ++i) accn[i] = 0.0f;
            int key = 0x7fffffff;
            if (j < nrec) {
                const ChunkRef c = record_at(s_chunk, s_start, j);
                key = c.k;
                const int p = c.row * TS + c.col;
                const float tp = (s_mask[p] >> c.k) & 1ull ? s_r[0][j] : 0.0f;
                if (tp > 0.0f) {
                    const float4* T = &L.tr[c.k * kIvlRec];
                    const float pcx = (float)(tc.tile_x * TS + (uint32_t)c.col) + 0.5f;
                    const float pcy = (float)(tc.tile_y * TS + (uint32_t)c.row) + 0.5f;
                    const float rsc = s_rsc[p];
                    IvlPair q;
                    ivl_pair(T, pcx, pcy, rsc, q);
                    const float dens = T[5].z;
                    const float dos = s_r[1][j], dm0 = s_r[2][j];
                    const V3 dcol = mk3(s_dl[0][p], s_dl[1][p], s_dl[2][p]);
                    const V3 d_cp = safe_clip3(scl3(dcol, tp), kMinVal, kMaxVal);
                    const float dphi = dphi_dx_f(q.od);
                    const float dw0 = dphi + q.alpha_t, dw1 = -dphi;
                    const V3 d_cb_int = scl3(d_cp, q.w0), d_cf_int = scl3(d_cp, q.w1);
                    const float d_od_int = dot3(d_cp, add3(scl3(q.cb, dw0), scl3(q.cf, dw1)));
                    const float d_od = safe_clip(dos + d_od_int, kMinVal, kMaxVal);
                    const float od_raw = dens * q.dist;
                    const float d_od_raw = (od_raw >= 1e-8f && od_raw <= 88.0f) ? d_od : 0.0f;
                    const float d_dens_local = safe_clip(d_od_raw * q.dist, kMinVal, kMaxVal);
                    const float d_dist = safe_clip(d_od_raw * dens, kMinVal, kMaxVal);
                    const float sgn = q.z_b >= q.z_f ? 1.0f : -1.0f;
                    const float d_zf_view = -d_dist * rsc * sgn, d_zb_view = d_dist * rsc * sgn;
                    const float range_nf = kIvlFar - kIvlNear;
                    const float den_f = kIvlFar - q.zfc * range_nf, den_b = kIvlFar - q.zbc * range_nf;
                    const float dz_f = kIvlNear * kIvlFar * range_nf / (den_f * den_f);
                    const float dz_b = kIvlNear * kIvlFar * range_nf / (den_b * den_b);
                    const bool zf_act = q.zfn >= 0.0f && q.zfn <= 1.0f, zb_act = q.zbn >= 0.0f && q.zbn <= 1.0f;
                    const float d_zfn = zf_act ? d_zf_view * dz_f : 0.0f;
                    const float d_zbn = zb_act ? d_zb_view * dz_b : 0.0f;
                    const V3 d_cb_raw = mk3(q.cb_raw.x > 0.0f ? d_cb_int.x : 0.0f, q.cb_raw.y > 0.0f ? d_cb_int.y : 0.0f,
                                            q.cb_raw.z > 0.0f ? d_cb_int.z : 0.0f);
                    const V3 d_cf_raw = mk3(q.cf_raw.x > 0.0f ? d_cf_int.x : 0.0f, q.cf_raw.y > 0.0f ? d_cf_int.y : 0.0f,
                                            q.cf_raw.z > 0.0f ? d_cf_int.z : 0.0f);
                    const V3 d_bc = safe_clip3(add3(d_cf_raw, d_cb_raw), kMinVal, kMaxVal);
                    const float d_of = (d_cf_raw.x + d_cf_raw.y) + d_cf_raw.z;
                    const float d_ob = (d_cb_raw.x + d_cb_raw.y) + d_cb_raw.z;
                    // xyzd (backward_interval_tiled.wgsl:544-600); vertex normals are zero
                    const float dxd = s_dl[6][p];
                    const float depth_c = q.w0 * q.z_b + q.w1 * q.z_f;
                    const float d_od_x = -(dxd * depth_c) * tp;
                    const float d_depth = dxd * tp;
                    const float d_od_w = d_depth * (q.z_b * dw0 + q.z_f * dw1);
                    const float d_zf_x = zf_act ? (d_depth * q.w1) * dz_f : 0.0f;
                    const float d_zb_x = zb_act ? (d_depth * q.w0) * dz_b : 0.0f;
                    // distortion (backward_interval_tiled.wgsl:602-650): through 1 - e^-u only
                    const float enter = q.z_f * rsc, dt_d = q.z_b * rsc - enter;
                    const float en = safe_exp(-(dens * dt_d));
                    const float d_u = (dm0 * tp) * en;
                    const float tzf = d_zfn + d_zf_x, tzb = d_zbn + d_zb_x;
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        const float ws = (s == q.ai ? q.b0 : 0.0f) + (s == q.bi ? q.b1 : 0.0f);
                        acc[4 * s] = ws * tzf;
                        acc[4 * s + 1] = ws * tzb;
                        acc[4 * s + 2] = ws * d_of;
                        acc[4 * s + 3] = ws * d_ob;
                    }
                    acc[16] = q.b2 * tzf;
                    acc[17] = q.b2 * tzb;
                    acc[18] = q.b2 * d_of;
                    acc[19] = q.b2 * d_ob;
                    acc[20] = d_bc.x;
                    acc[21] = d_bc.y;
                    acc[22] = d_bc.z;
                    acc[23] = safe_clip(d_dens_local, kMinVal, kMaxVal) +
                              safe_clip((d_od_x + d_od_w) * q.dist, kMinVal, kMaxVal) +
                              safe_clip(d_u * dt_d, kMinVal, kMaxVal);
                    if (has_dn) {
                        const V3 dn = mk3(s_dl[3][p] * tp, s_dl[4][p] * tp, s_dl[5][p] * tp);
                        const V3 dnf = scl3(dn, q.w1), dnb = scl3(dn, q.w0);
                        const V3 dns = add3(dnf, dnb);
#pragma unroll
                        for (int s = 0; s < 4; ++s) {
                            const float ws = (s == q.ai ? q.b0 : 0.0f) + (s == q.bi ? q.b1 : 0.0f);
                            accn[3 * s] = ws * dns.x;
                            accn[3 * s + 1] = ws * dns.y;
                            accn[3 * s + 2] = ws * dns.z;
                        }
                        accn[12] = q.b2 * dnf.x;
                        accn[13] = q.b2 * dnf.y;
                        accn[14] = q.b2 * dnf.z;
                        accn[15] = q.b2 * dnb.x;
                        accn[16] = q.b2 * dnb.y;
                        accn[17] = q.b2 * dnb.z;
                    }
                }
            }
            const SegMasksDpp sm = seg_masks_dpp(key);
            const int next_key = dpp_i<0x130, 0xF>(key, -1);  // wave_shl:1 (lane 63 reads -1)
            const bool tail = key != 0x7fffffff && next_key != key;
            seg_scan_dpp<kIvlGradTerms>(acc, sm);
#pragma unroll
            for (int i = 0; i < kIvlGradTerms; ++i)
                if (tail && acc[i] != 0.0f) atomicAdd(&s_acc[key][i], acc[i]);
            if (has_dn) {
                seg_scan_dpp<kIvlNormalTerms>(accn, sm);
#pragma unroll
                for (int i = 0; i < kIvlNormalTerms; ++i)
                    if (tail && accn[i] != 0.0f) atomicAdd(&s_acc[key][kIvlGradTerms + i], accn[i]);
            }
        }
        __syncthreads();
        // flush: one global add per non-zero term per (tile, tet)
        for (int it = tid; it < (int)nbp * KT; it += NT) {
            const int k = it / KT, i = it % KT;
            const float v = s_acc[k][i];
            if (v != 0.0f) unsafeAtomicAdd(d_ivr + (uint64_t)kIvlGradStride * s_tid[k] + i, v);
        }
        done = next;
        nb = nb_next;
    }
}

__device__ __forceinline__ void add_at(V3* a, int i, V3 v) {
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (k == i) a[k] = add3(a[k], v);
}

// Per-tet chain back (interval_chain_back.wgsl:160-545): the decomposition's
// gradients through each vertex's NDC z and colour offset, with the screen
// geometry (silhouette positions, barycentrics, crossing parameters) fixed.
// Per tet and vertex slot, one 32-byte sector: (d vertex xyz, 0, d vertex
// normal xyz, 0), so a vertex's gather reads one sector per incident tet.
constexpr int kTetVertTerms = 32;

__global__ __launch_bounds__(256) void k_interval_chain(const float4* __restrict__ rec,
                                                        const ViewParams* __restrict__ views, uint32_t M,
                                                        const float* __restrict__ d_ivr, float* __restrict__ tv,
                                                        float* __restrict__ d_dens, float* __restrict__ d_base,
                                                        float* __restrict__ d_grads) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M) return;
    const ViewParams& vw = views[0];
    float4 r[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = rec[4ull * t + i];
    const V3 v[4] = {mk3(r[0].x, r[0].y, r[0].z), mk3(r[1].x, r[1].y, r[1].z), mk3(r[2].x, r[2].y, r[2].z),
                     mk3(r[3].x, r[3].y, r[3].z)};
    const V3 g = mk3(r[1].w, r[2].w, r[3].w);
    IvlGeo G;
    if (!ivl_geometry(v, g, r[0].w, vw, G)) {
        float4* tv4 = reinterpret_cast<float4*>(tv + (uint64_t)kTetVertTerms * t);
#pragma unroll
        for (int i = 0; i < kTetVertTerms / 4; ++i) tv4[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const V3 cam = mk3(vw.cam[0], vw.cam[1], vw.cam[2]);
    float D[kIvlGradStride];
    const float4* D4 = reinterpret_cast<const float4*>(d_ivr + (uint64_t)kIvlGradStride * t);
#pragma unroll
    for (int i = 0; i < kIvlGradStride / 4; ++i) {
        const float4 q = D4[i];
        D[4 * i] = q.x; D[4 * i + 1] = q.y; D[4 * i + 2] = q.z; D[4 * i + 3] = q.w;
    }
    const V3 rz = mk3(vw.vp[2], vw.vp[6], vw.vp[10]), rw = mk3(vw.vp[3], vw.vp[7], vw.vp[11]);
    auto dz = [&](int vi, float d) {
        return safe_clip3(scl3(scl3(sub3(rz, scl3(rw, sel4(G.nz, vi))), sel4(G.iw, vi)), d), kMinVal, kMaxVal);
    };
    const float vx[4] = {v[0].x, v[1].x, v[2].x, v[3].x}, vy[4] = {v[0].y, v[1].y, v[2].y, v[3].y},
                vz[4] = {v[0].z, v[1].z, v[2].z, v[3].z};
    auto vtx = [&](int vi) { return mk3(sel4(vx, vi), sel4(vy, vi), sel4(vz, vi)); };
    V3 dv[4], dn[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) dv[i] = dn[i] = mk3(0.f, 0.f, 0.f);
    const float dbs = (D[20] + D[21]) + D[22];
    V3 dg = scl3(sub3(cam, v[0]), dbs);
    dv[0] = sub3(dv[0], scl3(g, dbs));
    const bool tri = G.inner < 4;
    const int sv[4] = {tri ? G.f0 : G.a0, tri ? G.f1 : G.b0, tri ? G.f2 : G.a1, tri ? G.f0 : G.b1};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int vi = sv[i];
        add_at(dv, vi, dz(vi, D[4 * i] + D[4 * i + 1]));
        const float doff = D[4 * i + 2] + D[4 * i + 3];
        add_at(dv, vi, safe_clip3(scl3(g, doff), kMinVal, kMaxVal));
        dg = add3(dg, safe_clip3(scl3(sub3(vtx(vi), cam), doff), kMinVal, kMaxVal));
        add_at(dn, vi, mk3(D[24 + 3 * i], D[25 + 3 * i], D[26 + 3 * i]));
    }
    const V3 cnf = mk3(D[36], D[37], D[38]), cnb = mk3(D[39], D[40], D[41]);
    if (tri) {
        const float z_face = G.w0 * sel4(G.nz, G.f0) + G.w1 * sel4(G.nz, G.f1) + G.w2 * sel4(G.nz, G.f2);
        const bool ff = z_face <= sel4(G.nz, G.ce);
        const float dzf = bsel(ff, D[16], D[17]), dzc = bsel(ff, D[17], D[16]);
        const float dof = bsel(ff, D[18], D[19]), doc = bsel(ff, D[19], D[18]);
        const int fv[3] = {G.f0, G.f1, G.f2};
        const float fw[3] = {G.w0, G.w1, G.w2};
#pragma unroll
        for (int i = 0; i < 3; ++i) add_at(dv, fv[i], dz(fv[i], fw[i] * dzf));
        add_at(dv, G.ce, dz(G.ce, dzc));
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const float doi = fw[i] * dof;
            add_at(dv, fv[i], safe_clip3(scl3(g, doi), kMinVal, kMaxVal));
            dg = add3(dg, safe_clip3(scl3(sub3(vtx(fv[i]), cam), doi), kMinVal, kMaxVal));
        }
        add_at(dv, G.ce, safe_clip3(scl3(g, doc), kMinVal, kMaxVal));
        dg = add3(dg, safe_clip3(scl3(sub3(vtx(G.ce), cam), doc), kMinVal, kMaxVal));
        const V3 dnf = mk3(bsel(ff, cnf.x, cnb.x), bsel(ff, cnf.y, cnb.y), bsel(ff, cnf.z, cnb.z));
        const V3 dnc = mk3(bsel(ff, cnb.x, cnf.x), bsel(ff, cnb.y, cnf.y), bsel(ff, cnb.z, cnf.z));
#pragma unroll
        for (int i = 0; i < 3; ++i) add_at(dn, fv[i], scl3(dnf, fw[i]));
        add_at(dn, G.ce, dnc);
    } else {
        const float z_a = ivl_mix(sel4(G.nz, G.a0), sel4(G.nz, G.a1), G.ta);
        const float z_b = ivl_mix(sel4(G.nz, G.b0), sel4(G.nz, G.b1), G.tb);
        const bool af = z_a <= z_b;
        const float dza = bsel(af, D[16], D[17]), dzb = bsel(af, D[17], D[16]);
        const float doa = bsel(af, D[18], D[19]), dob = bsel(af, D[19], D[18]);
        add_at(dv, G.a0, dz(G.a0, (1.0f - G.ta) * dza));
        add_at(dv, G.a1, dz(G.a1, G.ta * dza));
        add_at(dv, G.b0, dz(G.b0, (1.0f - G.tb) * dzb));
        add_at(dv, G.b1, dz(G.b1, G.tb * dzb));
        const float oa0 = (1.0f - G.ta) * doa, oa1 = G.ta * doa, ob0 = (1.0f - G.tb) * dob, ob1 = G.tb * dob;
        add_at(dv, G.a0, safe_clip3(scl3(g, oa0), kMinVal, kMaxVal));
        add_at(dv, G.a1, safe_clip3(scl3(g, oa1), kMinVal, kMaxVal));
        dg = add3(dg, safe_clip3(add3(scl3(sub3(vtx(G.a0), cam), oa0), scl3(sub3(vtx(G.a1), cam), oa1)), kMinVal,
                                 kMaxVal));
        add_at(dv, G.b0, safe_clip3(scl3(g, ob0), kMinVal, kMaxVal));
        add_at(dv, G.b1, safe_clip3(scl3(g, ob1), kMinVal, kMaxVal));
        dg = add3(dg, safe_clip3(add3(scl3(sub3(vtx(G.b0), cam), ob0), scl3(sub3(vtx(G.b1), cam), ob1)), kMinVal,
                                 kMaxVal));
        const V3 dna = mk3(bsel(af, cnf.x, cnb.x), bsel(af, cnf.y, cnb.y), bsel(af, cnf.z, cnb.z));
        const V3 dnb = mk3(bsel(af, cnb.x, cnf.x), bsel(af, cnb.y, cnf.y), bsel(af, cnb.z, cnf.z));
        add_at(dn, G.a0, scl3(dna, 1.0f - G.ta));
        add_at(dn, G.a1, scl3(dna, G.ta));
        add_at(dn, G.b0, scl3(dnb, 1.0f - G.tb));
        add_at(dn, G.b1, scl3(dnb, G.tb));
    }
    d_dens[t] = D[23];
    // the tet's per-vertex terms, gathered per vertex by k_vertex_gather
    float4* tv4 = reinterpret_cast<float4*>(tv + (uint64_t)kTetVertTerms * t);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const V3 a = safe_clip3(dv[i], kMinVal, kMaxVal), b = safe_clip3(dn[i], kMinVal, kMaxVal);
        tv4[2 * i] = make_float4(a.x, a.y, a.z, 0.0f);
        tv4[2 * i + 1] = make_float4(b.x, b.y, b.z, 0.0f);
    }
    const V3 a = safe_clip3(dg, kMinVal, kMaxVal), b = safe_clip3(mk3(D[20], D[21], D[22]), kMinVal, kMaxVal);
    d_grads[3ull * t] = a.x;
    d_grads[3ull * t + 1] = a.y;
    d_grads[3ull * t + 2] = a.z;
    d_base[3ull * t] = b.x;
    d_base[3ull * t + 1] = b.y;
    d_base[3ull * t + 2] = b.z;
}

// The tiled adjoint's per-tet staging records (16 floats: density, vertex
// slots 0..3 xyz, colour gradient xyz) out to the parameter gradients: thread
// i < M writes tet i's density and colour-gradient terms, thread i < N sums
// vertex i's slots over its (tet << 2 | slot) incidence list in tet order (no
// atomics).
__global__ __launch_bounds__(256) void k_tet_gather(const uint32_t* __restrict__ off, const uint32_t* __restrict__ ent,
                                                    const float* __restrict__ tet_g, uint32_t N, uint32_t M,
                                                    float* __restrict__ d_verts, float* __restrict__ d_dens,
                                                    float* __restrict__ d_grads) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < M) {
        const float4* g4 = reinterpret_cast<const float4*>(tet_g + 16ull * i);
        const float4 a = g4[0], d = g4[3];
        d_dens[i] = a.x;
        d_grads[3ull * i] = d.y;
        d_grads[3ull * i + 1] = d.z;
        d_grads[3ull * i + 2] = d.w;
    }
    if (i < N) {
        float x = 0.f, y = 0.f, z = 0.f;
        uint32_t e = off[i];
        const uint32_t end = off[i + 1];
        // four incidences in flight at a time (the sum keeps the list order)
        for (; e + 4 <= end; e += 4) {
            uint32_t te[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) te[u] = ent[e + u];
            float q[4][3];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const float* p = tet_g + 16ull * (te[u] >> 2) + 1u + 3u * (te[u] & 3u);
                q[u][0] = p[0];
                q[u][1] = p[1];
                q[u][2] = p[2];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                x += q[u][0];
                y += q[u][1];
                z += q[u][2];
            }
        }
        for (; e < end; ++e) {
            const uint32_t te = ent[e];
            const float* p = tet_g + 16ull * (te >> 2) + 1u + 3u * (te & 3u);
            x += p[0];
            y += p[1];
            z += p[2];
        }
        d_verts[3ull * i] = x;
        d_verts[3ull * i + 1] = y;
        d_verts[3ull * i + 2] = z;
    }
}

void launch_tet_gather(const uint32_t* off, const uint32_t* ent, const float* tet_g, uint32_t N, uint32_t M,
                       float* d_verts, float* d_dens, float* d_grads, hipStream_t s) {
    const uint32_t n = N > M ? N : M;
    if (n) hipLaunchKernelGGL(k_tet_gather, dim3((n + 255) / 256), dim3(256), 0, s, off, ent, tet_g, N, M, d_verts,
                              d_dens, d_grads);
}

// One thread per vertex: the sum of its tets' terms over the vertex -> (tet,
// slot) incidence lists (no atomics; the order of the sum is fixed).
__global__ __launch_bounds__(256) void k_vertex_gather(const uint32_t* __restrict__ off,
                                                       const uint32_t* __restrict__ ent, const float* __restrict__ tv,
                                                       uint32_t N, float* __restrict__ d_verts,
                                                       float* __restrict__ d_vn) {
    const uint32_t vtx = blockIdx.x * blockDim.x + threadIdx.x;
    if (vtx >= N) return;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, n0 = 0.f, n1 = 0.f, n2 = 0.f;
    for (uint32_t e = off[vtx], end = off[vtx + 1]; e < end; ++e) {
        const uint32_t te = ent[e];
        const float4* p = reinterpret_cast<const float4*>(tv + (uint64_t)kTetVertTerms * (te >> 2) + 8u * (te & 3u));
        const float4 a = p[0], n = p[1];
        a0 += a.x;
        a1 += a.y;
        a2 += a.z;
        n0 += n.x;
        n1 += n.y;
        n2 += n.z;
    }
    d_verts[3ull * vtx] = a0;
    d_verts[3ull * vtx + 1] = a1;
    d_verts[3ull * vtx + 2] = a2;
    d_vn[3ull * vtx] = n0;
    d_vn[3ull * vtx + 1] = n1;
    d_vn[3ull * vtx + 2] = n2;
}

template <int TS>
static void ivl_bwd_launch(const IvlBwdArgs& a, hipStream_t s) {
    using C = FwdCfg<TS, (TS > 16 ? 16 : 32), (TS * TS > 512 ? TS * TS : 512), 1, 1>;
    hipLaunchKernelGGL((k_interval_backward<C>), dim3(a.num_tiles), dim3(C::NT), 0, s, a.ivr, a.sorted_vals, a.ranges,
                       a.views, a.tile_order, a.rgba, a.dist, a.dl_img, a.dl_xyzd, a.dl_dist, a.d_ivr);
}

bool launch_interval_backward(const IvlBwdArgs& a, uint32_t tile_size, hipStream_t s) {
    if (tile_size != 8 && tile_size != 12 && tile_size != 16 && tile_size != 32) return false;
    switch (tile_size) {
        case 8: ivl_bwd_launch<8>(a, s); break;
        case 12: ivl_bwd_launch<12>(a, s); break;
        case 16: ivl_bwd_launch<16>(a, s); break;
        default: ivl_bwd_launch<32>(a, s); break;
    }
    if (a.M)
        hipLaunchKernelGGL(k_interval_chain, dim3((a.M + 255) / 256), dim3(256), 0, s, a.rec, a.views, a.M, a.d_ivr,
                           a.tv, a.d_dens, a.d_base, a.d_grads);
    if (a.N)
        hipLaunchKernelGGL(k_vertex_gather, dim3((a.N + 255) / 256), dim3(256), 0, s, a.v2t_off, a.v2t, a.tv, a.N,
                           a.d_verts, a.d_vn);
    return true;
}

template <int TS>
static void ivl_launch(const IvlArgs& a, hipStream_t s) {
    using C = FwdCfg<TS, (TS > 16 ? 16 : 32), (TS * TS > 512 ? TS * TS : 512), 1, 1>;
    hipLaunchKernelGGL((k_interval_forward<C>), dim3(a.num_tiles), dim3(C::NT), 0, s, a.ivr, a.sorted_vals, a.ranges,
                       a.views, a.tile_order, a.rgba, a.xyzd, a.dist);
}

bool launch_interval(const IvlArgs& a, uint32_t tile_size, hipStream_t s) {
    if (tile_size != 8 && tile_size != 12 && tile_size != 16 && tile_size != 32) return false;
    if (a.M) hipLaunchKernelGGL(k_interval_generate, dim3((a.M + 255) / 256), dim3(256), 0, s, a.rec, a.colors, a.views,
                                a.M, a.ivr);
    switch (tile_size) {
        case 8: ivl_launch<8>(a, s); break;
        case 12: ivl_launch<12>(a, s); break;
        case 16: ivl_launch<16>(a, s); break;
        default: ivl_launch<32>(a, s); break;
    }
    return true;
}

}  // namespace rmesh
