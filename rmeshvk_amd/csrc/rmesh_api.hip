// Host orchestration and the C ABI (include/rmesh_hip.h).
//
// Mirrors the call sequence of rmeshvk's RMeshRenderer (crates/rmesh-python/
// src/lib.rs): forward_tiled = project -> scan -> pair generation -> sort ->
// tile ranges -> compositing (lib.rs:1040-1204); backward reuses the sorted
// lists of the last forward (the reference regenerates and re-sorts the same
// pairs, lib.rs:1938-1958); train_step = forward + loss + backward + Adam with
// the step counter incremented once per call (lib.rs:2405-2636).
//
// MI355X specifics: one HIP stream per handle (or the caller's, e.g. torch's
// current stream); every per-step buffer is allocated once at a capacity
// derived from the first run, so a batched step is a fixed launch sequence
// that can be captured into a hipGraph and replayed; all clears inside that
// sequence are fill kernels (no memset nodes); the pair count, the step
// counter and the capacity-overflow flag stay in device memory.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rmesh_hip.h"
#include "rmesh_common.hpp"
#include "rmesh_kernels.hpp"
#include "rmesh_error.hpp"

using namespace rmesh;

namespace {

thread_local std::string g_err;

rmesh_status fail(rmesh_status code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                                     \
    do {                                                                                                  \
        hipError_t _e = (expr);                                                                           \
        if (_e != hipSuccess)                                                                             \
            return fail(RMESH_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e));             \
    } while (0)

template <typename T>
hipError_t dalloc(T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    return hipMalloc((void**)p, n * sizeof(T));
}

}  // namespace

void rmesh::set_last_error(const std::string& msg) { g_err = msg; }

// One scene of a renderer (a multi-scene renderer holds several, each with its
// own tets, resolution, SH degree and view capacity; a plain renderer one).
struct SceneHost {
    uint32_t N = 0, M = 0;           // vertices, tets
    uint32_t vert_base = 0, tet_base = 0;  // offsets in the concatenated arrays
    uint32_t W = 0, H = 0, max_views = 1;
    uint32_t deg = 0, nc = 1;        // SH degree (0 without SH) and coefficients per channel
    uint64_t sh_base = 0;            // first SH float
    uint32_t user = 0;               // the caller's scene index
};
// Tets of one SH degree: a contiguous run of scenes (scenes are kept sorted by degree).
struct ShClass {
    uint32_t deg, t0, n;
    uint64_t sh_off;
};

struct rmesh_renderer {
    int device = 0;
    // scenes in tet order (sorted by SH degree), the caller's index of each, the SH classes
    std::vector<SceneHost> sc;
    std::vector<uint32_t> sc_of_user;
    std::vector<ShClass> sh_classes;
    uint64_t sh_total = 0;        // SH floats of all scenes
    uint64_t pix_cap = 0;         // pixels of all view slots (sum of max_views W H)
    uint64_t color_cap = 0;       // float4 colours of all view slots
    uint64_t item_cap = 0;        // binning items at full view capacity
    uint32_t tile_cap = 0;        // global tiles at full view capacity, tile size 8
    uint32_t view_cap = 0;        // views at full capacity
    // the current view layout: views per scene (caller order) and what follows from it
    std::vector<uint32_t> step_views;
    std::vector<SceneRange> hscenes;
    std::vector<uint32_t> htile_view, hktile;
    uint64_t n_items = 0, color_len = 0, pix_len = 0;
    uint32_t n_gt = 0, n_kt = 0, n_views = 0;
    SceneRange* d_scenes = nullptr;
    uint32_t* d_tile_view = nullptr;   // global tile -> view
    uint32_t* d_ktile = nullptr;       // key tile -> global tile (0xFFFFFFFF: sentinel)
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    uint32_t N = 0, M = 0, W = 0, H = 0, TS = 12, tiles_x = 0, tiles_y = 0, T = 0;
    uint32_t T_alloc = 0;          // per-tile buffers are sized for the smallest tile (8)
    int fwd_var = 0, bwd_var = 0;  // batch-geometry variants of the compositing kernels
    uint32_t sort_db = 8;          // radix-sort digit width (8 or 4)
    uint32_t max_views = 1;
    bool has_sh = false;
    bool sh_half = false;  // forward colours from f16-rounded SH (the viewer path)
    uint32_t sh_deg = 0, nc = 1;
    float min_t = 0.0f;
    uint32_t step = 0;
    float lr_sh = 0.0f, lr_base = 0.0f;

    // parameters
    float *verts = nullptr, *base = nullptr, *dens = nullptr, *grads = nullptr, *sh = nullptr, *circ = nullptr;
    uint32_t* idx = nullptr;
    // optimiser state (allocated on first train step)
    float *m_v = nullptr, *v_v = nullptr, *m_d = nullptr, *v_d = nullptr, *m_g = nullptr, *v_g = nullptr;
    float *m_sh = nullptr, *v_sh = nullptr, *m_b = nullptr, *v_b = nullptr;
    // gradients
    float *d_verts = nullptr, *d_dens = nullptr, *d_grads = nullptr, *d_colors = nullptr, *d_sh = nullptr;
    // the tiled adjoint's per-tet staging: [M][16] = density, 4 vertex slots x xyz, colour gradient xyz
    float* tet_g = nullptr;
    // derived per step
    float4* rec = nullptr;
    float4* colors = nullptr;  // [max_views][M] (SH) or [M]
    // binning
    ViewParams* d_views = nullptr;
    uint32_t *counts = nullptr, *block_sums = nullptr, *vis_count = nullptr, *n_pairs = nullptr, *overflow = nullptr;
    uint32_t* block_vis = nullptr;  // per k_project block: visible items of view 0
    unsigned long long* spine_part = nullptr;  // two-level spine scan: per-workgroup totals
    uint32_t* spine_aux = nullptr;             // and instance-count partials
    uint32_t* d_step = nullptr;
    uint64_t *keys_a = nullptr, *keys_b = nullptr;
    uint32_t *vals_a = nullptr, *vals_b = nullptr;
    uint32_t cap = 0;
    uint32_t* hist = nullptr;
    // depth sort of the (tet, view) items before pair generation: keys, item
    // indices (ping-pong), its histogram, the item count on the device, the
    // block sums of the depth-ordered pair scan
    uint32_t *dkey_a = nullptr, *dkey_b = nullptr, *ditem_a = nullptr, *ditem_b = nullptr;
    uint32_t *ihist = nullptr, *d_nitems = nullptr, *block_sums2 = nullptr;
    bool pairs_depth_ordered = false;  // the last pair generation emitted items in depth order
    uint32_t* ranges = nullptr;
    uint32_t* tile_order = nullptr;  // compositing workgroup -> global tile, heaviest first
    bool sorted_in_b = false;
    // adjoint mode: false = the reference's undo from the image alpha, true =
    // the checkpointed adjoint (transmittance checkpoints stored by the forward)
    bool adj_ck = false;
    bool have_ck = false;        // the last forward stored checkpoints
    float* ck = nullptr;         // [slots][TS*TS]
    uint64_t ck_floats = 0;
    float2* fin = nullptr;       // [max_views][H][W]: final log T, stop position
    // images
    float *rgba = nullptr, *aux = nullptr, *dl = nullptr, *gt = nullptr, *loss_partial = nullptr, *loss_out = nullptr;
    bool aux_alloc = false;
    // forward_error inputs and outputs (allocated on first use)
    float *err_pe = nullptr, *err_se = nullptr, *tet_err = nullptr;
    int32_t* tet_cnt = nullptr;
    // interval forward: per-tet decomposition and the depth / distortion images (allocated on first use)
    float4* ivr = nullptr;
    float *ivl_xyzd = nullptr, *ivl_dist = nullptr;
    // interval adjoint: upstream gradients, per-tet decomposition gradients, vertex-normal gradients
    float *ivl_dxyzd = nullptr, *ivl_ddist = nullptr, *ivl_dgrad = nullptr, *ivl_dvn = nullptr, *ivl_tv = nullptr;
    // vertex -> (tet, slot) incidence lists of the (fixed) tet indices, built on first use
    uint32_t *v2t_off = nullptr, *v2t = nullptr;
    bool have_interval = false;  // the buffers above hold the last interval forward's state
    // last forward
    uint32_t last_views = 0;
    bool have_forward = false;
    // graph
    bool graph_mode = false;
    hipGraph_t graph[8] = {};          // one graph, or one per stage when timing
    hipGraphExec_t graph_exec[8] = {};
    int graph_parts = 0;
    uint32_t graph_views = 0, graph_loss = 0;
    std::vector<uint32_t> graph_layout;  // views per scene the graph was captured for
    // per-scene timing: wall clock at the step start and when each scene's last adjoint tile retired
    bool scene_timing = false;
    uint32_t scene_order = RMESH_SCENE_ORDER_INTERLEAVED;  // compositing grids' tile order across scenes
    uint64_t* scene_clock = nullptr;  // [scenes + 1]
    float graph_lrs[5] = {0, 0, 0, 0, 0};  // lr_v, lr_d, lr_g, lr_sh, lr_base baked into the graph
    std::vector<ViewParams> hviews;
    // optional per-stage timing (HIP events recorded between stages)
    bool timing = false;
    hipEvent_t ev[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};

    // global tiles and sort bits of the current view layout (the V arguments name the layout's view count)
    uint32_t gtiles(uint32_t) const { return n_gt; }
    uint32_t sorting_bits(uint32_t) const { return rmesh_sorting_bits(n_kt); }
    bool multi() const { return sc.size() > 1; }
    // Pairs generated in depth order of their items, then sorted on the tile
    // bits only.  Not when one view's sentinel key (tile 0xFFFFFFFF) masks to
    // the last real tile (tile count 2^(sorting bits - 32)): there the sentinels
    // share a masked tile with real pairs and the full (tile, depth) sort
    // orders them.
    bool depth_first() const {
        if (!dkey_a) return false;
        const uint32_t tb = sorting_bits(0) - 32u;
        return !(n_views == 1 && tb < 32u && n_kt == (1u << tb));
    }
};

extern "C" uint32_t rmesh_sorting_bits(uint32_t num_tiles) {
    uint32_t tile_bits = 1;
    if (num_tiles > 1) {
        uint32_t x = num_tiles - 1;
        tile_bits = 0;
        while (x) {
            tile_bits++;
            x >>= 1;
        }
    }
    return (32u + tile_bits + 7u) & ~7u;
}

extern "C" const char* rmesh_last_error(void) { return g_err.c_str(); }

namespace {

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        (void)hipGetDevice(&prev);
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        int cur = -1;
        (void)hipGetDevice(&cur);
        if (prev >= 0 && cur != prev) (void)hipSetDevice(prev);
    }
};

void make_view(const rmesh_renderer* h, const SceneHost& sh, const float* cam, const float* vp, const float* ci,
               ViewParams& o) {
    std::memset(&o, 0, sizeof(o));
    for (int i = 0; i < 16; ++i) o.vp[i] = vp[i];
    for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) o.c2w[j * 3 + k] = ci[j * 4 + k];
    for (int i = 0; i < 4; ++i) o.intr[i] = ci[12 + i];
    for (int i = 0; i < 3; ++i) o.cam[i] = cam[i];
    o.min_t = h->min_t;
    o.Wf = (float)sh.W;
    o.Hf = (float)sh.H;
    o.width = sh.W;
    o.height = sh.H;
    o.tiles_x = (sh.W + h->TS - 1) / h->TS;
    o.tiles_y = (sh.H + h->TS - 1) / h->TS;
    o.tile_size = h->TS;
}

// The view layout of a step: per_user[u] views of the caller's scene u, in the
// caller's scene order (cams/vps/c2ws in that order).  Views get their global
// tiles, sort-key tiles (one sentinel tile per view when several views share
// the sort, none with one view: the reference's own keys), colour and pixel
// offsets; scenes (tet order) get their binning items.  Uploads the view
// table, the scene table and the tile tables.
rmesh_status upload_layout(rmesh_renderer* h, const uint32_t* per_user, const float* cams, const float* vps,
                           const float* c2ws) {
    const uint32_t ns = (uint32_t)h->sc.size();
    uint32_t V = 0;
    for (uint32_t u = 0; u < ns; ++u) {
        const SceneHost& sh = h->sc[h->sc_of_user[u]];
        if (per_user[u] > sh.max_views)
            return fail(RMESH_ERR_INVALID, "scene " + std::to_string(u) + ": n_views must be at most max_views=" +
                                               std::to_string(sh.max_views));
        V += per_user[u];
    }
    if (V == 0) return fail(RMESH_ERR_INVALID, "no views");
    h->have_interval = false;
    h->step_views.assign(per_user, per_user + ns);
    h->hviews.resize(V);
    h->htile_view.clear();
    h->hktile.clear();
    std::vector<uint32_t> first_view(ns, 0);
    uint64_t color = 0, pix = 0;
    uint32_t gt = 0, v = 0;
    for (uint32_t u = 0; u < ns; ++u) {
        const uint32_t si = h->sc_of_user[u];
        const SceneHost& sh = h->sc[si];
        first_view[si] = v;
        for (uint32_t j = 0; j < per_user[u]; ++j, ++v) {
            ViewParams& o = h->hviews[v];
            make_view(h, sh, cams + 3ull * v, vps + 16ull * v, c2ws + 16ull * v, o);
            const uint32_t T = o.tiles_x * o.tiles_y;
            o.tile_base = gt;
            o.key_base = (uint32_t)h->hktile.size();
            o.sentinel_tile = V == 1 ? 0xFFFFFFFFu : o.key_base + T;
            o.color_off = h->has_sh ? (int64_t)color - (int64_t)sh.tet_base : 0;
            o.pix_base = pix;
            o.scene = u;
            for (uint32_t t = 0; t < T; ++t) {
                h->htile_view.push_back(v);
                h->hktile.push_back(gt + t);
            }
            if (V > 1) h->hktile.push_back(0xFFFFFFFFu);
            gt += T;
            color += sh.M;
            pix += (uint64_t)sh.W * sh.H;
        }
    }
    h->n_views = V;
    h->n_gt = gt;
    h->n_kt = (uint32_t)h->hktile.size();
    h->color_len = h->has_sh ? color : h->M;
    h->pix_len = pix;
    h->hscenes.resize(ns);
    uint64_t item = 0;
    for (uint32_t si = 0; si < ns; ++si) {
        const SceneHost& sh = h->sc[si];
        const uint32_t nv = per_user[sh.user];
        h->hscenes[si] = SceneRange{item, sh.tet_base, sh.M, first_view[si], nv};
        item += (uint64_t)sh.M * nv;
    }
    h->n_items = item;
    HIP_TRY(hipMemcpyAsync(h->d_views, h->hviews.data(), sizeof(ViewParams) * V, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->d_scenes, h->hscenes.data(), sizeof(SceneRange) * ns, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->d_tile_view, h->htile_view.data(), 4ull * h->n_gt, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->d_ktile, h->hktile.data(), 4ull * h->n_kt, hipMemcpyHostToDevice, h->stream));
    // the host vectors stay alive (and unchanged) until the next layout: the
    // asynchronous copies read them in stream order
    return RMESH_OK;
}

// V views of a renderer's single scene.
rmesh_status upload_views(rmesh_renderer* h, uint32_t V, const float* cams, const float* vps, const float* c2ws) {
    if (h->multi()) return fail(RMESH_ERR_UNSUPPORTED, "a multi-scene renderer takes its views per scene");
    return upload_layout(h, &V, cams, vps, c2ws);
}

void drop_graph(rmesh_renderer* h) {
    for (int p = 0; p < 8; ++p) {
        if (h->graph_exec[p]) (void)hipGraphExecDestroy(h->graph_exec[p]);
        if (h->graph[p]) (void)hipGraphDestroy(h->graph[p]);
        h->graph_exec[p] = nullptr;
        h->graph[p] = nullptr;
    }
    h->graph_parts = 0;
}

// Tile grid of tile size ts; invalidates everything derived from the old grid.
void set_tiles(rmesh_renderer* h, uint32_t ts) {
    h->TS = ts;
    h->tiles_x = (h->W + ts - 1) / ts;
    h->tiles_y = (h->H + ts - 1) / ts;
    h->T = h->tiles_x * h->tiles_y;
    h->have_forward = false;
    h->have_interval = false;
    drop_graph(h);
}

rmesh_status ensure_capacity(rmesh_renderer* h, uint64_t need) {
    if (need <= h->cap && h->keys_a && h->hist) return RMESH_OK;
    if (need > 0xFFFFFFF0ull) return fail(RMESH_ERR_CAPACITY, "tile-tet pair count exceeds 2^32");
    uint64_t c = std::max<uint64_t>(need + need / 4 + (1u << 16), 1u << 20);
    if (c > 0xFFFFFFF0ull) c = 0xFFFFFFF0ull;
    HIP_TRY(hipStreamSynchronize(h->stream));
    (void)hipFree(h->keys_a);
    (void)hipFree(h->keys_b);
    (void)hipFree(h->vals_a);
    (void)hipFree(h->vals_b);
    (void)hipFree(h->hist);
    h->keys_a = h->keys_b = nullptr;
    h->vals_a = h->vals_b = nullptr;
    h->hist = nullptr;
    h->cap = 0;
    HIP_TRY(dalloc(&h->keys_a, c));
    HIP_TRY(dalloc(&h->keys_b, c));
    HIP_TRY(dalloc(&h->vals_a, c));
    HIP_TRY(dalloc(&h->vals_b, c));
    HIP_TRY(dalloc(&h->hist, std::max(radix_hist_words((uint32_t)c), radix_hist_words32((uint32_t)c))));
    h->cap = (uint32_t)c;
    drop_graph(h);
    return RMESH_OK;
}

// Checkpoint buffers for V views at the current capacity and tile size
// (allocated on demand in the checkpointed mode; growing them drops the graph).
rmesh_status ensure_checkpoints(rmesh_renderer* h, uint32_t V) {
    if (!h->adj_ck) return RMESH_OK;
    const uint64_t need = checkpoint_slots(h->cap, h->n_gt) * h->TS * h->TS;
    if (need > h->ck_floats || !h->ck) {
        HIP_TRY(hipStreamSynchronize(h->stream));
        (void)hipFree(h->ck);
        h->ck = nullptr;
        h->ck_floats = 0;
        const uint64_t c = need + need / 8;
        HIP_TRY(dalloc(&h->ck, c));
        h->ck_floats = c;
        drop_graph(h);
    }
    if (!h->fin) HIP_TRY(dalloc(&h->fin, (size_t)h->pix_cap));
    return RMESH_OK;
}

void mark(rmesh_renderer* h, int i) {
    if (h->timing) (void)hipEventRecord(h->ev[i], h->stream);
}

// Colours + projection + scan; in eager mode also reads the pair count back
// to size the pair buffers (skipped under graph capture).
rmesh_status stage_binning_front(rmesh_renderer* h, uint32_t V, bool capture) {
    hipStream_t s = h->stream;
    const uint32_t ns = (uint32_t)h->sc.size();
    launch_prep_records(h->verts, h->idx, h->dens, h->grads, h->M, h->rec, s);
    if (h->has_sh) {
        for (const ShClass& c : h->sh_classes)  // one launch per SH degree present
            launch_sh_colors(h->rec, h->sh + c.sh_off, c.deg, c.t0, c.n, h->d_scenes, ns, h->d_views, h->colors,
                             h->sh_half, s);
    } else {
        launch_copy_colors(h->base, h->M, h->colors, s);
    }
    launch_fill_u32(h->overflow, 0u, 1, s);
    const bool df = h->depth_first();
    launch_project(h->rec, reinterpret_cast<const float4*>(h->circ), h->d_scenes, ns, h->n_items, h->d_views, h->TS,
                   h->counts, h->block_sums, h->block_vis, df ? h->dkey_a : nullptr, df ? h->ditem_a : nullptr,
                   h->d_nitems, s);
    launch_spine_scan(h->block_sums, item_blocks(h->n_items), h->n_pairs, h->overflow, s, h->block_vis, h->vis_count,
                      h->spine_part, h->spine_aux);
    if (!capture) {
        uint32_t P = 0;
        HIP_TRY(hipMemcpyAsync(&P, h->n_pairs, 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        rmesh_status st = ensure_capacity(h, P);
        if (st != RMESH_OK) return st;
        st = ensure_checkpoints(h, V);
        if (st != RMESH_OK) return st;
    }
    return RMESH_OK;
}

uint32_t* keys32(uint64_t* k) { return reinterpret_cast<uint32_t*>(k); }

// Pairs in item order (the reference's pre-sort order), offsets from the
// item-order scan of the binning front.
void launch_pair_gen_item_order(rmesh_renderer* h, hipStream_t s) {
    launch_pairs(h->rec, reinterpret_cast<const float4*>(h->circ), h->d_scenes, (uint32_t)h->sc.size(), h->n_items,
                 h->d_views, h->TS, h->counts, h->block_sums, h->keys_a, h->vals_a, h->cap, h->overflow, s);
    h->pairs_depth_ordered = false;
}
// Pair generation.  Depth-first: the items (M V 32-bit keys) are sorted by
// depth key, and their pairs emitted in that order, so the pair sort that
// follows needs only the tile bits (2 passes of 8 bits instead of 6).
void launch_pair_gen(rmesh_renderer* h, hipStream_t s) {
    if (!h->depth_first()) {
        launch_pair_gen_item_order(h, s);
        return;
    }
    const bool in_b = launch_radix_sort32(h->dkey_a, h->ditem_a, h->dkey_b, h->ditem_b, h->d_nitems,
                                          (uint32_t)h->item_cap, h->ihist, s, h->sort_db);
    launch_pairs_depth_ordered(h->rec, h->d_scenes, (uint32_t)h->sc.size(), in_b ? h->ditem_b : h->ditem_a,
                               (uint32_t)h->n_items, h->d_views, h->TS, h->counts, h->block_sums2, h->n_pairs,
                               keys32(h->keys_a), h->vals_a, h->cap, h->overflow, h->spine_part, s);
    h->pairs_depth_ordered = true;
}
// Depth-ordered pairs carry 32-bit keys (the key tile) and are sorted on the
// tile bits; item-ordered pairs carry the 64-bit (tile, depth) keys.
void launch_sort_stage(rmesh_renderer* h, hipStream_t s) {
    if (h->pairs_depth_ordered)
        h->sorted_in_b = launch_radix_sort32(keys32(h->keys_a), h->vals_a, keys32(h->keys_b), h->vals_b, h->n_pairs,
                                             h->cap, h->hist, s, h->sort_db, h->sorting_bits(0) - 32u);
    else
        h->sorted_in_b = launch_radix_sort(h->keys_a, h->vals_a, h->keys_b, h->vals_b, h->n_pairs, h->cap,
                                           h->sorting_bits(0), h->hist, s, h->sort_db);
}
void launch_ranges_stage(rmesh_renderer* h, hipStream_t s) {
    launch_fill_u32(h->ranges, 0u, 2ull * h->n_gt, s);
    uint64_t* sk = h->sorted_in_b ? h->keys_b : h->keys_a;
    if (h->pairs_depth_ordered) launch_ranges32(keys32(sk), h->n_pairs, h->cap, h->ranges, h->d_ktile, h->n_kt, s);
    else launch_ranges(sk, h->n_pairs, h->cap, h->ranges, h->d_ktile, h->n_kt, s);
    if (h->scene_order != RMESH_SCENE_ORDER_BY_SCENE ||
        !launch_tile_order_scenes(h->ranges, h->n_gt, h->d_tile_view, h->d_views, (uint32_t)h->sc.size(),
                                  h->tile_order, s))
        launch_tile_order(h->ranges, h->n_gt, h->tile_order, s);
}

rmesh_status stage_binning_back(rmesh_renderer* h, uint32_t V) {
    hipStream_t s = h->stream;
    launch_pair_gen(h, s);
    launch_sort_stage(h, s);
    launch_ranges_stage(h, s);
    return RMESH_OK;
}

const uint32_t* sorted_vals(const rmesh_renderer* h) { return h->sorted_in_b ? h->vals_b : h->vals_a; }


rmesh_status stage_forward(rmesh_renderer* h, uint32_t V, bool want_aux) {
    FwdArgs a;
    a.rec = h->rec;
    a.colors = h->colors;
    a.sorted_vals = sorted_vals(h);
    a.ranges = h->ranges;
    a.views = h->d_views;
    a.tile_view = h->d_tile_view;
    a.num_gtiles = h->n_gt;
    a.tile_order = h->tile_order;
    a.rgba = h->rgba;
    a.aux = want_aux ? h->aux : nullptr;
    const bool ck = h->adj_ck && !want_aux;  // the aux forward keeps no checkpoints
    a.ck = ck ? h->ck : nullptr;
    a.fin = ck ? h->fin : nullptr;
    if (!launch_forward(a, h->TS, h->fwd_var, h->stream)) return fail(RMESH_ERR_INVALID, "unsupported tile size");
    h->have_ck = ck;
    return RMESH_OK;
}

void zero_grads(rmesh_renderer* h, uint32_t V) {
    hipStream_t s = h->stream;
    launch_fill_u32(reinterpret_cast<uint32_t*>(h->tet_g), 0u, 16ull * h->M, s);
    launch_fill_u32(reinterpret_cast<uint32_t*>(h->d_verts), 0u, 3ull * h->N, s);
    launch_fill_u32(reinterpret_cast<uint32_t*>(h->d_dens), 0u, h->M, s);
    launch_fill_u32(reinterpret_cast<uint32_t*>(h->d_grads), 0u, 3ull * h->M, s);
    launch_fill_u32(reinterpret_cast<uint32_t*>(h->d_colors), 0u, 3ull * h->color_len, s);
}

rmesh_status stage_backward(rmesh_renderer* h, uint32_t V, const float* dl, const float* gt, uint32_t loss_type) {
    BwdArgs a;
    a.rec = h->rec;
    a.idx = h->idx;
    a.colors = h->colors;
    a.sorted_vals = sorted_vals(h);
    a.ranges = h->ranges;
    a.views = h->d_views;
    a.tile_view = h->d_tile_view;
    a.num_gtiles = h->n_gt;
    a.tile_order = h->tile_order;
    a.rgba = h->rgba;
    a.dl_img = dl;
    a.gt = gt;
    a.loss_type = loss_type;
    a.loss_partial = gt ? h->loss_partial : nullptr;
    a.tet_g = h->tet_g;
    a.d_colors = h->d_colors;
    a.ck = nullptr;
    a.fin = nullptr;
    a.scene_done = h->scene_timing ? h->scene_clock : nullptr;
    if (h->adj_ck) {
        if (!h->have_ck)
            return fail(RMESH_ERR_INVALID,
                        "checkpointed adjoint: the last forward kept no checkpoints (a render_aux forward, or the "
                        "mode was switched after it); run forward_tiled again");
        a.ck = h->ck;
        a.fin = h->fin;
    }
    if (!launch_backward(a, h->TS, h->bwd_var, h->stream)) return fail(RMESH_ERR_INVALID, "unsupported tile size");
    return RMESH_OK;
}

// The staged per-tet terms (the compositing adjoint's, then the SH adjoint's)
// out to the gradient arrays: per tet, density and colour gradient; per
// vertex, the sum over its (tet, slot) incidence list.
void launch_gather_grads(rmesh_renderer* h, hipStream_t s) {
    launch_tet_gather(h->v2t_off, h->v2t, h->tet_g, h->N, h->M, h->d_verts, h->d_dens, h->d_grads, s);
}

rmesh_status ensure_optimizer(rmesh_renderer* h) {
    if (h->m_v) return RMESH_OK;
    size_t nv = 3ull * h->N, nd = h->M, ng = 3ull * h->M;
    HIP_TRY(dalloc(&h->m_v, nv));
    HIP_TRY(dalloc(&h->v_v, nv));
    HIP_TRY(dalloc(&h->m_d, nd));
    HIP_TRY(dalloc(&h->v_d, nd));
    HIP_TRY(dalloc(&h->m_g, ng));
    HIP_TRY(dalloc(&h->v_g, ng));
    HIP_TRY(dalloc(&h->m_b, ng));
    HIP_TRY(dalloc(&h->v_b, ng));
    HIP_TRY(hipMemset(h->m_v, 0, nv * 4));
    HIP_TRY(hipMemset(h->v_v, 0, nv * 4));
    HIP_TRY(hipMemset(h->m_d, 0, nd * 4));
    HIP_TRY(hipMemset(h->v_d, 0, nd * 4));
    HIP_TRY(hipMemset(h->m_g, 0, ng * 4));
    HIP_TRY(hipMemset(h->v_g, 0, ng * 4));
    HIP_TRY(hipMemset(h->m_b, 0, ng * 4));
    HIP_TRY(hipMemset(h->v_b, 0, ng * 4));
    if (h->has_sh) {
        size_t ns = h->sh_total;
        HIP_TRY(dalloc(&h->m_sh, ns));
        HIP_TRY(dalloc(&h->v_sh, ns));
        HIP_TRY(hipMemset(h->m_sh, 0, ns * 4));
        HIP_TRY(hipMemset(h->v_sh, 0, ns * 4));
    }
    return RMESH_OK;
}

// SH adjoint (+ fused Adam when adam) of every SH degree class: one launch per
// degree present.
void launch_sh_classes(rmesh_renderer* h, bool adam, float lr, AdamStep st, hipStream_t s) {
    for (const ShClass& c : h->sh_classes) {
        ShBwdArgs a;
        std::memset(&a, 0, sizeof(a));
        a.rec = h->rec;
        a.sh = h->sh + c.sh_off;
        a.deg = c.deg;
        a.t0 = c.t0;
        a.M = c.n;
        a.scenes = h->d_scenes;
        a.n_scenes = (uint32_t)h->sc.size();
        a.views = h->d_views;
        a.d_colors = h->d_colors;
        a.d_sh = h->d_sh + c.sh_off;
        a.tet_g = h->tet_g;
        a.adam = adam;
        a.m_sh = adam ? h->m_sh + c.sh_off : nullptr;
        a.v_sh = adam ? h->v_sh + c.sh_off : nullptr;
        a.lr = lr;
        a.st = st;
        launch_sh_backward(a, s);
    }
}

// Adam stage (adam_compute.wgsl; groups and order of lib.rs:2590-2626):
// vertices, densities, color_grads (+ base_colors / SH when enabled).
rmesh_status stage_adam(rmesh_renderer* h, uint32_t V, float lr_v, float lr_d, float lr_g, float lr_sh) {
    hipStream_t s = h->stream;
    launch_step_inc(h->d_step, h->overflow, s);
    AdamStep st{h->d_step, h->overflow, 0.9f, 0.999f, 1e-8f};
    if (h->has_sh) launch_sh_classes(h, lr_sh != 0.0f, lr_sh, st, s);
    launch_gather_grads(h, s);
    AdamArgs aa;
    std::memset(&aa, 0, sizeof(aa));
    aa.st = st;
    int n = 0;
    aa.grp[n++] = AdamGroup{h->verts, h->d_verts, h->m_v, h->v_v, 3ull * h->N, lr_v};
    aa.grp[n++] = AdamGroup{h->dens, h->d_dens, h->m_d, h->v_d, h->M, lr_d};
    aa.grp[n++] = AdamGroup{h->grads, h->d_grads, h->m_g, h->v_g, 3ull * h->M, lr_g};
    if (!h->has_sh && h->lr_base != 0.0f) aa.grp[n++] = AdamGroup{h->base, h->d_colors, h->m_b, h->v_b, 3ull * h->M, h->lr_base};
    aa.ngroups = n;
    launch_adam(aa, s);
    launch_reduce_loss(h->loss_partial, h->d_views, h->n_views, h->loss_out, s);
    return RMESH_OK;
}

// The train step as 7 stages; timing marks sit between them:
// 0 colours + projection + scan | 1 pair generation | 2 sort | 3 ranges |
// 4 forward | 5 adjoint | 6 SH adjoint + gradient gather + Adam + loss.
constexpr int kStages = 7;

rmesh_status record_stage(rmesh_renderer* h, int stage, uint32_t V, const float* gt, uint32_t loss_type, float lr_v,
                          float lr_d, float lr_g, float lr_sh, bool capture) {
    hipStream_t s = h->stream;
    switch (stage) {
        case 0: {
            if (h->scene_timing) launch_stamp_start(h->scene_clock, (uint32_t)h->sc.size(), s);
            rmesh_status st = stage_binning_front(h, V, capture);
            if (st != RMESH_OK) return st;
            zero_grads(h, V);
            return RMESH_OK;
        }
        case 1:
            launch_pair_gen(h, s);
            return RMESH_OK;
        case 2:
            launch_sort_stage(h, s);
            return RMESH_OK;
        case 3:
            launch_ranges_stage(h, s);
            return RMESH_OK;
        case 4: return stage_forward(h, V, false);
        case 5: return stage_backward(h, V, nullptr, gt, loss_type);
        default: return stage_adam(h, V, lr_v, lr_d, lr_g, lr_sh);
    }
}

// Eager step (or one capture of the whole step), with timing marks.
rmesh_status record_train(rmesh_renderer* h, uint32_t V, const float* gt, uint32_t loss_type, float lr_v, float lr_d,
                          float lr_g, float lr_sh, bool capture) {
    for (int i = 0; i < kStages; ++i) {
        mark(h, i);
        rmesh_status st = record_stage(h, i, V, gt, loss_type, lr_v, lr_d, lr_g, lr_sh, capture);
        if (st != RMESH_OK) return st;
    }
    mark(h, kStages);
    return RMESH_OK;
}

void free_all(rmesh_renderer* h) {
    drop_graph(h);
    void* ptrs[] = {h->verts, h->base, h->dens, h->grads, h->sh, h->circ, h->idx, h->m_v, h->v_v, h->m_d, h->v_d,
                    h->m_g, h->v_g, h->m_sh, h->v_sh, h->m_b, h->v_b, h->d_verts, h->d_dens, h->d_grads, h->d_colors,
                    h->d_sh, h->rec, h->colors, h->d_views, h->counts, h->block_sums, h->vis_count, h->n_pairs,
                    h->overflow, h->d_step, h->keys_a, h->keys_b, h->vals_a, h->vals_b, h->hist, h->ranges, h->rgba,
                    h->aux, h->dl, h->gt, h->loss_partial, h->loss_out, h->err_pe, h->err_se, h->tet_err,
                    h->tet_cnt, h->tile_order, h->ivr, h->ivl_xyzd, h->ivl_dist, h->ivl_dxyzd,
                    h->ivl_ddist, h->ivl_dgrad, h->ivl_dvn, h->ivl_tv, h->v2t_off, h->v2t, h->ck, h->fin,
                    h->d_scenes, h->d_tile_view, h->d_ktile, h->scene_clock, h->tet_g, h->dkey_a, h->dkey_b,
                    h->ditem_a, h->ditem_b, h->ihist, h->d_nitems, h->block_sums2, h->block_vis, h->spine_part,
                    h->spine_aux};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (hipEvent_t e : h->ev)
        if (e) (void)hipEventDestroy(e);
    if (h->own_stream) (void)hipStreamDestroy(h->own_stream);
}

rmesh_status check_views(rmesh_renderer* h, uint32_t V) {
    if (V == 0 || V > h->max_views)
        return fail(RMESH_ERR_INVALID, "n_views must be in [1, max_views=" + std::to_string(h->max_views) + "]");
    return RMESH_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
// Scenes from n descriptors (one: a plain renderer).  Tets, vertices and SH
// are concatenated with the scenes sorted by SH degree (stable), so each
// degree's tets are one contiguous run; the caller's scene order is kept for
// views and images.
static rmesh_status create_impl(const rmesh_scene_desc* descs, uint32_t n_scenes, rmesh_renderer** out) {
    if (!descs || !out || n_scenes == 0) return fail(RMESH_ERR_INVALID, "null argument");
    *out = nullptr;
    uint64_t NT = 0, MT = 0;
    for (uint32_t u = 0; u < n_scenes; ++u) {
        const rmesh_scene_desc* d = descs + u;
        const std::string who = n_scenes > 1 ? "scene " + std::to_string(u) + ": " : "";
        if (!d->vertices || !d->indices || !d->base_colors || !d->densities || !d->color_grads || !d->circumdata)
            return fail(RMESH_ERR_INVALID, who + "scene arrays must be non-null");
        if (d->aux_dim != 0) return fail(RMESH_ERR_UNSUPPORTED, "custom aux channels (aux_dim > 0) are not supported");
        if (d->width == 0 || d->height == 0) return fail(RMESH_ERR_INVALID, who + "width and height must be positive");
        if (d->n_tets > 0xFFFFFFFFull || d->n_vertices > 0x55555555ull) return fail(RMESH_ERR_INVALID, "scene too large");
        if (d->sh_coeffs && d->sh_degree > 3) return fail(RMESH_ERR_INVALID, who + "sh_degree must be 0..3");
        if ((d->sh_coeffs != nullptr) != (descs[0].sh_coeffs != nullptr))
            return fail(RMESH_ERR_INVALID, "either every scene carries SH coefficients or none does");
        if ((d->tile_size ? d->tile_size : 12) != (descs[0].tile_size ? descs[0].tile_size : 12))
            return fail(RMESH_ERR_INVALID, "every scene of a renderer shares one tile size");
        if (d->sh_half != descs[0].sh_half) return fail(RMESH_ERR_INVALID, "every scene shares sh_half");
        for (uint64_t i = 0; i < 4 * d->n_tets; ++i)
            if (d->indices[i] >= d->n_vertices) return fail(RMESH_ERR_INVALID, who + "tet index out of range");
        NT += d->n_vertices;
        MT += d->n_tets;
    }
    if (MT > 0xFFFFFFFFull || NT > 0x55555555ull) return fail(RMESH_ERR_INVALID, "scenes too large together");
    const rmesh_scene_desc* d0 = descs;
    uint32_t ts = d0->tile_size ? d0->tile_size : 12;
    if (ts != 8 && ts != 12 && ts != 16 && ts != 32) return fail(RMESH_ERR_INVALID, "tile_size must be 8, 12, 16 or 32");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(RMESH_ERR_DEVICE, "no HIP device available");
    int dev = d0->device_id;
    if (dev < 0) (void)hipGetDevice(&dev);
    if (dev >= ndev) return fail(RMESH_ERR_INVALID, "device_id out of range");

    rmesh_renderer* h = new rmesh_renderer();
    h->device = dev;
    DeviceGuard g(dev);
    h->has_sh = d0->sh_coeffs != nullptr;
    h->sh_half = h->has_sh && d0->sh_half != 0;
    // scene order: by SH degree, then the caller's order
    std::vector<uint32_t> order(n_scenes);
    for (uint32_t u = 0; u < n_scenes; ++u) order[u] = u;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        return (h->has_sh ? descs[a].sh_degree : 0) < (h->has_sh ? descs[b].sh_degree : 0);
    });
    h->sc.resize(n_scenes);
    h->sc_of_user.resize(n_scenes);
    uint32_t vb = 0, tb = 0;
    uint64_t shb = 0;
    for (uint32_t i = 0; i < n_scenes; ++i) {
        const rmesh_scene_desc* d = descs + order[i];
        SceneHost& c = h->sc[i];
        c.N = (uint32_t)d->n_vertices;
        c.M = (uint32_t)d->n_tets;
        c.vert_base = vb;
        c.tet_base = tb;
        c.W = d->width;
        c.H = d->height;
        c.max_views = d->max_views ? d->max_views : 1;
        c.deg = h->has_sh ? d->sh_degree : 0;
        c.nc = (c.deg + 1) * (c.deg + 1);
        c.sh_base = shb;
        c.user = order[i];
        h->sc_of_user[order[i]] = i;
        vb += c.N;
        tb += c.M;
        if (h->has_sh) shb += 3ull * c.M * c.nc;
        if (h->has_sh && c.M) {
            if (h->sh_classes.empty() || h->sh_classes.back().deg != c.deg)
                h->sh_classes.push_back(ShClass{c.deg, c.tet_base, 0, c.sh_base});
            h->sh_classes.back().n += c.M;
        }
        h->pix_cap += (uint64_t)c.max_views * c.W * c.H;
        h->color_cap += h->has_sh ? (uint64_t)c.max_views * c.M : 0;
        h->item_cap += (uint64_t)c.max_views * c.M;
        h->tile_cap += c.max_views * (((c.W + 7) / 8) * ((c.H + 7) / 8));
        h->view_cap += c.max_views;
    }
    if (!h->has_sh) h->color_cap = tb;
    h->sh_total = shb;
    h->N = vb;
    h->M = tb;
    // the single-scene entry points work on the first scene's image size
    h->W = h->sc[h->sc_of_user[0]].W;
    h->H = h->sc[h->sc_of_user[0]].H;
    set_tiles(h, ts);
    h->T_alloc = ((h->W + 7) / 8) * ((h->H + 7) / 8);
    h->max_views = h->sc[h->sc_of_user[0]].max_views;
    h->sh_deg = h->sc[0].deg;
    h->nc = (h->sh_deg + 1) * (h->sh_deg + 1);
    auto bail = [&](rmesh_status s) {
        free_all(h);
        delete h;
        return s;
    };
#define CREATE_TRY(expr)                                                                                 \
    do {                                                                                                 \
        hipError_t _e = (expr);                                                                          \
        if (_e != hipSuccess) return bail(fail(RMESH_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e))); \
    } while (0)
    CREATE_TRY(hipStreamCreateWithFlags(&h->own_stream, hipStreamNonBlocking));
    h->stream = h->own_stream;
    size_t N3 = 3ull * h->N, M = h->M, M3 = 3ull * h->M;
    CREATE_TRY(dalloc(&h->verts, N3));
    CREATE_TRY(dalloc(&h->base, M3));
    CREATE_TRY(dalloc(&h->dens, M));
    CREATE_TRY(dalloc(&h->grads, M3));
    CREATE_TRY(dalloc(&h->circ, 4 * M));
    CREATE_TRY(dalloc(&h->idx, 4 * M));
    if (h->has_sh) {
        CREATE_TRY(dalloc(&h->sh, h->sh_total));
        CREATE_TRY(dalloc(&h->d_sh, h->sh_total));
    }
    for (const SceneHost& c : h->sc) {
        const rmesh_scene_desc* d = descs + c.user;
        CREATE_TRY(hipMemcpy(h->verts + 3ull * c.vert_base, d->vertices, 12ull * c.N, hipMemcpyHostToDevice));
        CREATE_TRY(hipMemcpy(h->base + 3ull * c.tet_base, d->base_colors, 12ull * c.M, hipMemcpyHostToDevice));
        CREATE_TRY(hipMemcpy(h->dens + c.tet_base, d->densities, 4ull * c.M, hipMemcpyHostToDevice));
        CREATE_TRY(hipMemcpy(h->grads + 3ull * c.tet_base, d->color_grads, 12ull * c.M, hipMemcpyHostToDevice));
        CREATE_TRY(hipMemcpy(h->circ + 4ull * c.tet_base, d->circumdata, 16ull * c.M, hipMemcpyHostToDevice));
        if (h->has_sh)
            CREATE_TRY(hipMemcpy(h->sh + c.sh_base, d->sh_coeffs, 12ull * c.M * c.nc, hipMemcpyHostToDevice));
        // tet indices refer to the concatenated vertex array
        std::vector<uint32_t> idx(d->indices, d->indices + 4ull * c.M);
        if (c.vert_base)
            for (uint32_t& x : idx) x += c.vert_base;
        CREATE_TRY(hipMemcpy(h->idx + 4ull * c.tet_base, idx.data(), 16ull * c.M, hipMemcpyHostToDevice));
    }
    {
        // vertex -> (tet << 2 | slot) incidence lists over all scenes' tets
        // (counting sort of the global indices: each list in tet order)
        std::vector<uint32_t> ix(4ull * h->M), off(h->N + 1ull, 0u), ent(4ull * h->M);
        CREATE_TRY(hipMemcpy(ix.data(), h->idx, 16ull * h->M, hipMemcpyDeviceToHost));
        for (uint32_t v : ix) ++off[v + 1];
        for (uint32_t v = 0; v < h->N; ++v) off[v + 1] += off[v];
        std::vector<uint32_t> cur(off.begin(), off.end() - 1);
        for (uint64_t e = 0; e < ix.size(); ++e) ent[cur[ix[e]]++] = (uint32_t)e;
        CREATE_TRY(dalloc(&h->v2t_off, h->N + 1ull));
        CREATE_TRY(dalloc(&h->v2t, 4ull * h->M));
        CREATE_TRY(hipMemcpy(h->v2t_off, off.data(), 4ull * (h->N + 1ull), hipMemcpyHostToDevice));
        CREATE_TRY(hipMemcpy(h->v2t, ent.data(), 16ull * h->M, hipMemcpyHostToDevice));
    }
    CREATE_TRY(dalloc(&h->tet_g, 16ull * h->M));
    CREATE_TRY(dalloc(&h->d_verts, N3));
    CREATE_TRY(dalloc(&h->d_dens, M));
    CREATE_TRY(dalloc(&h->d_grads, M3));
    CREATE_TRY(dalloc(&h->d_colors, 3 * h->color_cap));
    CREATE_TRY(dalloc(&h->rec, 4 * M));
    CREATE_TRY(dalloc(&h->colors, h->color_cap));
    CREATE_TRY(dalloc(&h->d_views, h->view_cap));
    CREATE_TRY(dalloc(&h->d_scenes, n_scenes));
    CREATE_TRY(dalloc(&h->d_tile_view, h->tile_cap));
    CREATE_TRY(dalloc(&h->d_ktile, h->tile_cap + h->view_cap));
    CREATE_TRY(dalloc(&h->counts, h->item_cap));
    CREATE_TRY(dalloc(&h->block_sums, item_blocks(h->item_cap) + 1));
    CREATE_TRY(dalloc(&h->block_vis, item_blocks(h->item_cap) + 1));
    CREATE_TRY(dalloc(&h->spine_part, spine_parts(item_blocks(h->item_cap) + 1)));
    CREATE_TRY(dalloc(&h->spine_aux, spine_parts(item_blocks(h->item_cap) + 1)));
    if (h->item_cap < (1ull << 27)) {  // item indices fit the depth sort's values (item << 5 | count)
        CREATE_TRY(dalloc(&h->dkey_a, h->item_cap));
        CREATE_TRY(dalloc(&h->dkey_b, h->item_cap));
        CREATE_TRY(dalloc(&h->ditem_a, h->item_cap));
        CREATE_TRY(dalloc(&h->ditem_b, h->item_cap));
        CREATE_TRY(dalloc(&h->ihist, radix_hist_words32((uint32_t)h->item_cap)));
        CREATE_TRY(dalloc(&h->d_nitems, 1));
        CREATE_TRY(dalloc(&h->block_sums2, item_blocks(h->item_cap) + 1));
    }
    CREATE_TRY(dalloc(&h->vis_count, 1));
    CREATE_TRY(dalloc(&h->n_pairs, 1));
    CREATE_TRY(dalloc(&h->overflow, 1));
    CREATE_TRY(dalloc(&h->d_step, 1));
    CREATE_TRY(hipMemset(h->d_step, 0, 4));
    CREATE_TRY(dalloc(&h->ranges, 2ull * h->tile_cap));
    CREATE_TRY(dalloc(&h->tile_order, (size_t)h->tile_cap));
    CREATE_TRY(dalloc(&h->rgba, 4 * h->pix_cap));
    CREATE_TRY(dalloc(&h->dl, 4ull * h->W * h->H));
    CREATE_TRY(dalloc(&h->gt, 3 * h->pix_cap));
    CREATE_TRY(dalloc(&h->loss_partial, (size_t)h->tile_cap));
    CREATE_TRY(dalloc(&h->loss_out, h->view_cap));
#undef CREATE_TRY
    *out = h;
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_create(const rmesh_scene_desc* d, rmesh_renderer** out) {
    return create_impl(d, 1, out);
}

extern "C" rmesh_status rmesh_create_multi(const rmesh_scene_desc* descs, uint32_t n_scenes, rmesh_renderer** out) {
    return create_impl(descs, n_scenes, out);
}

extern "C" void rmesh_destroy(rmesh_renderer* h) {
    if (!h) return;
    DeviceGuard g(h->device);
    (void)hipStreamSynchronize(h->stream);
    free_all(h);
    delete h;
}

extern "C" rmesh_status rmesh_set_min_t(rmesh_renderer* h, float min_t) {
    if (!h) return fail(RMESH_ERR_INVALID, "null handle");
    h->min_t = min_t;
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_set_stream(rmesh_renderer* h, void* stream) {
    if (!h) return fail(RMESH_ERR_INVALID, "null handle");
    hipStream_t s = stream ? (hipStream_t)stream : h->own_stream;
    if (s != h->stream) {
        DeviceGuard g(h->device);
        HIP_TRY(hipStreamSynchronize(h->stream));
        h->stream = s;
    }
    return RMESH_OK;
}

static rmesh_status forward_views(rmesh_renderer* h, uint32_t V, const float* cams, const float* vps, const float* c2ws,
                                  bool want_aux) {
    rmesh_status st = upload_views(h, V, cams, vps, c2ws);
    if (st != RMESH_OK) return st;
    st = stage_binning_front(h, V, false);
    if (st != RMESH_OK) return st;
    st = stage_binning_back(h, V);
    if (st != RMESH_OK) return st;
    if (want_aux && !h->aux_alloc) {
        HIP_TRY(dalloc(&h->aux, 8ull * h->W * h->H * h->max_views));
        h->aux_alloc = true;
    }
    st = stage_forward(h, V, want_aux);
    if (st != RMESH_OK) return st;
    h->last_views = V;
    h->have_forward = true;
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_forward_tiled(rmesh_renderer* h, const float* cam, const float* vp, const float* ci,
                                            float* out_rgba, float* out_aux) {
    if (!h || !cam || !vp || !ci || !out_rgba) return fail(RMESH_ERR_INVALID, "null argument");
    DeviceGuard g(h->device);
    rmesh_status st = forward_views(h, 1, cam, vp, ci, out_aux != nullptr);
    if (st != RMESH_OK) return st;
    size_t npx = (size_t)h->W * h->H;
    HIP_TRY(hipMemcpyAsync(out_rgba, h->rgba, npx * 16, hipMemcpyDeviceToHost, h->stream));
    if (out_aux) HIP_TRY(hipMemcpyAsync(out_aux, h->aux, npx * 32, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    uint32_t ovf = 0;
    HIP_TRY(hipMemcpy(&ovf, h->overflow, 4, hipMemcpyDeviceToHost));
    if (ovf) return fail(RMESH_ERR_CAPACITY, "pair buffer overflow");
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_forward(rmesh_renderer* h, const float* cam, const float* vp, const float* ci,
                                      float* out_rgba) {
    return rmesh_forward_tiled(h, cam, vp, ci, out_rgba, nullptr);
}

extern "C" rmesh_status rmesh_forward_error(rmesh_renderer* h, const float* pixel_err, const float* ssim_err,
                                            float* tet_err, int32_t* tet_count) {
    if (!h || !pixel_err || !ssim_err || !tet_err || !tet_count) return fail(RMESH_ERR_INVALID, "null argument");
    if (!h->have_forward || h->last_views != 1)
        return fail(RMESH_ERR_INVALID, "forward_error requires a preceding forward_tiled");
    DeviceGuard g(h->device);
    hipStream_t s = h->stream;
    const size_t npx = (size_t)h->W * h->H;
    if (!h->tet_err) {
        HIP_TRY(dalloc(&h->err_pe, npx));
        HIP_TRY(dalloc(&h->err_se, npx));
        HIP_TRY(dalloc(&h->tet_err, 16ull * h->M));
        HIP_TRY(dalloc(&h->tet_cnt, 2ull * h->M));
    }
    HIP_TRY(hipMemcpyAsync(h->err_pe, pixel_err, npx * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(h->err_se, ssim_err, npx * 4, hipMemcpyHostToDevice, s));
    launch_fill_u32(reinterpret_cast<uint32_t*>(h->tet_err), 0u, 16ull * h->M, s);
    launch_fill_u32(reinterpret_cast<uint32_t*>(h->tet_cnt), 0u, 2ull * h->M, s);
    ErrArgs a;
    a.rec = h->rec;
    a.sorted_vals = sorted_vals(h);
    a.ranges = h->ranges;
    a.views = h->d_views;
    a.num_tiles = h->T;
    a.pixel_err = h->err_pe;
    a.ssim_err = h->err_se;
    a.tet_err = h->tet_err;
    a.tet_count = h->tet_cnt;
    if (!launch_error(a, h->TS, s)) return fail(RMESH_ERR_UNSUPPORTED, "tile size");
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(tet_err, h->tet_err, 64ull * h->M, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(tet_count, h->tet_cnt, 8ull * h->M, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_forward_interval_tiled(rmesh_renderer* h, const float* cam, const float* vp,
                                                     const float* ci, float* out_rgba, float* out_xyzd,
                                                     float* out_dist) {
    if (!h || !cam || !vp || !ci || !out_rgba || !out_xyzd || !out_dist) return fail(RMESH_ERR_INVALID, "null argument");
    DeviceGuard g(h->device);
    hipStream_t s = h->stream;
    const size_t npx = (size_t)h->W * h->H;
    if (!h->ivr) {
        HIP_TRY(dalloc(&h->ivr, 7ull * h->M));
        HIP_TRY(dalloc(&h->ivl_xyzd, 4 * npx));
        HIP_TRY(dalloc(&h->ivl_dist, 5 * npx));
    }
    rmesh_status st = upload_views(h, 1, cam, vp, ci);
    if (st != RMESH_OK) return st;
    st = stage_binning_front(h, 1, false);
    if (st != RMESH_OK) return st;
    st = stage_binning_back(h, 1);
    if (st != RMESH_OK) return st;
    IvlArgs a;
    a.rec = h->rec;
    a.colors = h->colors;
    a.views = h->d_views;
    a.M = h->M;
    a.ivr = h->ivr;
    a.sorted_vals = sorted_vals(h);
    a.ranges = h->ranges;
    a.tile_order = h->tile_order;
    a.num_tiles = h->T;
    a.rgba = h->rgba;
    a.xyzd = h->ivl_xyzd;
    a.dist = h->ivl_dist;
    if (!launch_interval(a, h->TS, s)) return fail(RMESH_ERR_UNSUPPORTED, "tile size");
    HIP_TRY(hipGetLastError());
    // the tiled forward's state (image, aux) is not this path's
    h->have_forward = false;
    HIP_TRY(hipMemcpyAsync(out_rgba, h->rgba, npx * 16, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(out_xyzd, h->ivl_xyzd, npx * 16, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(out_dist, h->ivl_dist, npx * 20, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    uint32_t ovf = 0;
    HIP_TRY(hipMemcpy(&ovf, h->overflow, 4, hipMemcpyDeviceToHost));
    if (ovf) return fail(RMESH_ERR_CAPACITY, "pair buffer overflow");
    h->have_interval = true;
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_backward_interval_tiled(rmesh_renderer* h, const float* dl_img, const float* dl_xyzd,
                                                      const float* dl_dist, float* d_vertices, float* d_densities,
                                                      float* d_base_colors, float* d_color_grads,
                                                      float* d_vertex_normals) {
    if (!h || !dl_img || !d_vertices || !d_densities || !d_base_colors || !d_color_grads || !d_vertex_normals)
        return fail(RMESH_ERR_INVALID, "null argument");
    if (!h->have_interval)
        return fail(RMESH_ERR_INVALID, "backward_interval_tiled requires a preceding forward_interval_tiled");
    DeviceGuard g(h->device);
    hipStream_t s = h->stream;
    const size_t npx = (size_t)h->W * h->H;
    if (!h->ivl_dgrad) {
        HIP_TRY(dalloc(&h->ivl_dxyzd, 4 * npx));
        HIP_TRY(dalloc(&h->ivl_ddist, 5 * npx));
        HIP_TRY(dalloc(&h->ivl_dgrad, 48ull * h->M));
        HIP_TRY(dalloc(&h->ivl_dvn, 3ull * h->N));
        HIP_TRY(dalloc(&h->ivl_tv, 32ull * h->M));
        // (the vertex incidence lists are built at create)
    }
    HIP_TRY(hipMemcpyAsync(h->dl, dl_img, npx * 16, hipMemcpyHostToDevice, s));
    if (dl_xyzd) HIP_TRY(hipMemcpyAsync(h->ivl_dxyzd, dl_xyzd, npx * 16, hipMemcpyHostToDevice, s));
    if (dl_dist) HIP_TRY(hipMemcpyAsync(h->ivl_ddist, dl_dist, npx * 20, hipMemcpyHostToDevice, s));
    launch_fill_u32(reinterpret_cast<uint32_t*>(h->ivl_dgrad), 0u, 48ull * h->M, s);
    launch_fill_u32(reinterpret_cast<uint32_t*>(h->d_dens), 0u, h->M, s);
    launch_fill_u32(reinterpret_cast<uint32_t*>(h->d_colors), 0u, 3ull * h->M, s);
    launch_fill_u32(reinterpret_cast<uint32_t*>(h->d_grads), 0u, 3ull * h->M, s);
    IvlBwdArgs a;
    a.rec = h->rec;
    a.idx = h->idx;
    a.views = h->d_views;
    a.M = h->M;
    a.ivr = h->ivr;
    a.sorted_vals = sorted_vals(h);
    a.ranges = h->ranges;
    a.tile_order = h->tile_order;
    a.num_tiles = h->T;
    a.rgba = h->rgba;
    a.dist = h->ivl_dist;
    a.dl_img = h->dl;
    a.dl_xyzd = dl_xyzd ? h->ivl_dxyzd : nullptr;
    a.dl_dist = dl_dist ? h->ivl_ddist : nullptr;
    a.d_ivr = h->ivl_dgrad;
    a.tv = h->ivl_tv;
    a.v2t_off = h->v2t_off;
    a.v2t = h->v2t;
    a.N = h->N;
    a.d_verts = h->d_verts;
    a.d_dens = h->d_dens;
    a.d_base = h->d_colors;
    a.d_grads = h->d_grads;
    a.d_vn = h->ivl_dvn;
    if (!launch_interval_backward(a, h->TS, s)) return fail(RMESH_ERR_UNSUPPORTED, "tile size");
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(d_vertices, h->d_verts, 12ull * h->N, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(d_densities, h->d_dens, 4ull * h->M, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(d_base_colors, h->d_colors, 12ull * h->M, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(d_color_grads, h->d_grads, 12ull * h->M, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(d_vertex_normals, h->ivl_dvn, 12ull * h->N, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_read_instance_count(rmesh_renderer* h, uint32_t* out) {
    if (!h || !out) return fail(RMESH_ERR_INVALID, "null argument");
    DeviceGuard g(h->device);
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipMemcpy(out, h->vis_count, 4, hipMemcpyDeviceToHost));
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_backward(rmesh_renderer* h, const float* dl, float* dV, float* dD, float* dB, float* dG) {
    if (!h || !dl || !dV || !dD || !dB || !dG) return fail(RMESH_ERR_INVALID, "null argument");
    if (!h->have_forward || h->last_views != 1) return fail(RMESH_ERR_INVALID, "backward requires a preceding forward_tiled");
    if (h->sh_half) return fail(RMESH_ERR_UNSUPPORTED, "f16-rounded SH (sh_half) serves the forward only");
    DeviceGuard g(h->device);
    hipStream_t s = h->stream;
    size_t npx = (size_t)h->W * h->H;
    HIP_TRY(hipMemcpyAsync(h->dl, dl, npx * 16, hipMemcpyHostToDevice, s));
    zero_grads(h, 1);
    rmesh_status st = stage_backward(h, 1, h->dl, nullptr, 0);
    if (st != RMESH_OK) return st;
    // colour gradient returned as d_base_colors (the per-tet colours of this view)
    HIP_TRY(hipMemcpyAsync(dB, h->d_colors, 3ull * h->M * 4, hipMemcpyDeviceToHost, s));
    if (h->has_sh) launch_sh_classes(h, false, 0.0f, AdamStep{h->d_step, h->overflow, 0.9f, 0.999f, 1e-8f}, s);
    launch_gather_grads(h, s);
    HIP_TRY(hipMemcpyAsync(dV, h->d_verts, 3ull * h->N * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(dD, h->d_dens, 1ull * h->M * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(dG, h->d_grads, 3ull * h->M * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return RMESH_OK;
}

// Device-pointer variants for callers whose tensors already live on the GPU
// (torch): copies are device to device on the renderer's stream (set it to
// the caller's with rmesh_set_stream), no host round trip.
extern "C" rmesh_status rmesh_update_params_device(rmesh_renderer* h, const float* v, const float* b, const float* d,
                                                   const float* gr) {
    if (!h || !v || !b || !d || !gr) return fail(RMESH_ERR_INVALID, "null argument");
    DeviceGuard g(h->device);
    hipStream_t s = h->stream;
    HIP_TRY(hipMemcpyAsync(h->verts, v, 12ull * h->N, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(h->base, b, 12ull * h->M, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(h->dens, d, 4ull * h->M, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(h->grads, gr, 12ull * h->M, hipMemcpyDeviceToDevice, s));
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_forward_tiled_device(rmesh_renderer* h, const float* cam, const float* vp,
                                                   const float* ci, float* d_rgba) {
    if (!h || !cam || !vp || !ci || !d_rgba) return fail(RMESH_ERR_INVALID, "null argument");
    DeviceGuard g(h->device);
    rmesh_status st = forward_views(h, 1, cam, vp, ci, false);
    if (st != RMESH_OK) return st;
    HIP_TRY(hipMemcpyAsync(d_rgba, h->rgba, 16ull * h->W * h->H, hipMemcpyDeviceToDevice, h->stream));
    uint32_t ovf = 0;
    HIP_TRY(hipMemcpyAsync(&ovf, h->overflow, 4, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (ovf) return fail(RMESH_ERR_CAPACITY, "pair buffer overflow");
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_backward_device(rmesh_renderer* h, const float* d_dl, float* dV, float* dD, float* dB,
                                              float* dG) {
    if (!h || !d_dl || !dV || !dD || !dB || !dG) return fail(RMESH_ERR_INVALID, "null argument");
    if (!h->have_forward || h->last_views != 1) return fail(RMESH_ERR_INVALID, "backward requires a preceding forward_tiled");
    if (h->sh_half) return fail(RMESH_ERR_UNSUPPORTED, "f16-rounded SH (sh_half) serves the forward only");
    DeviceGuard g(h->device);
    hipStream_t s = h->stream;
    HIP_TRY(hipMemcpyAsync(h->dl, d_dl, 16ull * h->W * h->H, hipMemcpyDeviceToDevice, s));
    zero_grads(h, 1);
    rmesh_status st = stage_backward(h, 1, h->dl, nullptr, 0);
    if (st != RMESH_OK) return st;
    HIP_TRY(hipMemcpyAsync(dB, h->d_colors, 12ull * h->M, hipMemcpyDeviceToDevice, s));
    if (h->has_sh) launch_sh_classes(h, false, 0.0f, AdamStep{h->d_step, h->overflow, 0.9f, 0.999f, 1e-8f}, s);
    launch_gather_grads(h, s);
    HIP_TRY(hipMemcpyAsync(dV, h->d_verts, 12ull * h->N, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(dD, h->d_dens, 4ull * h->M, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(dG, h->d_grads, 12ull * h->M, hipMemcpyDeviceToDevice, s));
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_backward_hw(rmesh_renderer* h, const float* dl, float* dV, float* dD, float* dB,
                                          float* dG) {
    // the forward() it follows is served by the tiled forward, so its state is the tiled adjoint's
    return rmesh_backward(h, dl, dV, dD, dB, dG);
}

extern "C" rmesh_status rmesh_get_sh_grad(rmesh_renderer* h, float* d_sh) {
    if (!h || !d_sh) return fail(RMESH_ERR_INVALID, "null argument");
    if (!h->has_sh) return fail(RMESH_ERR_INVALID, "renderer has no SH coefficients");
    DeviceGuard g(h->device);
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipMemcpy(d_sh, h->d_sh, 4ull * h->sh_total, hipMemcpyDeviceToHost));
    return RMESH_OK;
}

static rmesh_status finish_train(rmesh_renderer* h, uint32_t V, float* out_loss, float* out_view_losses = nullptr) {
    std::vector<float> lv(V);
    uint32_t ovf = 0;
    HIP_TRY(hipMemcpyAsync(lv.data(), h->loss_out, 4ull * V, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipMemcpyAsync(&ovf, h->overflow, 4, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (ovf) return fail(RMESH_ERR_CAPACITY, "pair buffer overflow");
    h->step += 1;
    h->last_views = V;
    h->have_forward = true;
    if (out_loss) {
        double s = 0.0;
        for (float x : lv) s += x;
        *out_loss = (float)(s / V);
    }
    if (out_view_losses) std::memcpy(out_view_losses, lv.data(), 4ull * V);
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_train_step(rmesh_renderer* h, const float* cam, const float* vp, const float* ci,
                                         const float* gt, float lr_v, float lr_d, float lr_g, uint32_t loss_type,
                                         float* out_loss) {
    if (!h || !cam || !vp || !ci || !gt) return fail(RMESH_ERR_INVALID, "null argument");
    if (h->sh_half) return fail(RMESH_ERR_UNSUPPORTED, "f16-rounded SH (sh_half) serves the forward only");
    DeviceGuard g(h->device);
    rmesh_status st = ensure_optimizer(h);
    if (st != RMESH_OK) return st;
    st = upload_views(h, 1, cam, vp, ci);
    if (st != RMESH_OK) return st;
    HIP_TRY(hipMemcpyAsync(h->gt, gt, 12ull * h->W * h->H, hipMemcpyHostToDevice, h->stream));
    st = record_train(h, 1, h->gt, loss_type, lr_v, lr_d, lr_g, h->lr_sh, false);
    if (st != RMESH_OK) return st;
    return finish_train(h, 1, out_loss);
}

extern "C" rmesh_status rmesh_set_extra_lrs(rmesh_renderer* h, float lr_sh, float lr_base) {
    if (!h) return fail(RMESH_ERR_INVALID, "null handle");
    h->lr_sh = lr_sh;
    h->lr_base = lr_base;
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_set_adjoint_mode(rmesh_renderer* h, int32_t mode) {
    if (!h) return fail(RMESH_ERR_INVALID, "null handle");
    if (mode != RMESH_ADJOINT_UNDO && mode != RMESH_ADJOINT_CHECKPOINT)
        return fail(RMESH_ERR_INVALID, "adjoint mode must be RMESH_ADJOINT_UNDO or RMESH_ADJOINT_CHECKPOINT");
    const bool ck = mode == RMESH_ADJOINT_CHECKPOINT;
    if (ck != h->adj_ck) {
        DeviceGuard g(h->device);
        drop_graph(h);
        h->adj_ck = ck;
        h->have_ck = false;
        h->have_forward = false;  // the next backward needs a forward in this mode
        if (!ck) {
            HIP_TRY(hipStreamSynchronize(h->stream));
            (void)hipFree(h->ck);
            (void)hipFree(h->fin);
            h->ck = nullptr;
            h->fin = nullptr;
            h->ck_floats = 0;
        }
    }
    return RMESH_OK;
}

extern "C" int32_t rmesh_get_adjoint_mode(const rmesh_renderer* h) {
    if (!h) return -1;
    return h->adj_ck ? RMESH_ADJOINT_CHECKPOINT : RMESH_ADJOINT_UNDO;
}

extern "C" rmesh_status rmesh_device_bytes(rmesh_renderer* h, uint64_t* out) {
    if (!h || !out) return fail(RMESH_ERR_INVALID, "null argument");
    DeviceGuard g(h->device);
    const void* ptrs[] = {h->verts, h->base, h->dens, h->grads, h->sh, h->circ, h->idx, h->m_v, h->v_v, h->m_d,
                          h->v_d, h->m_g, h->v_g, h->m_sh, h->v_sh, h->m_b, h->v_b, h->d_verts, h->d_dens,
                          h->d_grads, h->d_colors, h->d_sh, h->rec, h->colors, h->d_views, h->counts, h->block_sums,
                          h->vis_count, h->n_pairs, h->overflow, h->d_step, h->keys_a, h->keys_b, h->vals_a,
                          h->vals_b, h->hist, h->ranges, h->rgba, h->aux, h->dl, h->gt, h->loss_partial,
                          h->loss_out, h->err_pe, h->err_se, h->tet_err, h->tet_cnt, h->tile_order, h->ivr,
                          h->ivl_xyzd, h->ivl_dist, h->ivl_dxyzd, h->ivl_ddist, h->ivl_dgrad, h->ivl_dvn,
                          h->ivl_tv, h->v2t_off, h->v2t, h->ck, h->fin, h->d_scenes, h->d_tile_view, h->d_ktile,
                          h->tet_g, h->dkey_a, h->dkey_b, h->ditem_a, h->ditem_b, h->ihist, h->d_nitems,
                          h->block_sums2, h->block_vis, h->spine_part, h->spine_aux};
    uint64_t total = 0;
    for (const void* p : ptrs) {
        if (!p) continue;
        size_t n = 0;
        HIP_TRY(hipMemPtrGetInfo(const_cast<void*>(p), &n));
        total += n;
    }
    *out = total;
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_update_params(rmesh_renderer* h, const float* v, const float* b, const float* d,
                                            const float* gr) {
    if (!h || !v || !b || !d || !gr) return fail(RMESH_ERR_INVALID, "null argument");
    DeviceGuard g(h->device);
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipMemcpy(h->verts, v, 12ull * h->N, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(h->base, b, 12ull * h->M, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(h->dens, d, 4ull * h->M, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(h->grads, gr, 12ull * h->M, hipMemcpyHostToDevice));
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_get_params(rmesh_renderer* h, float* v, float* b, float* d, float* gr) {
    if (!h || !v || !b || !d || !gr) return fail(RMESH_ERR_INVALID, "null argument");
    DeviceGuard g(h->device);
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipMemcpy(v, h->verts, 12ull * h->N, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(b, h->base, 12ull * h->M, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(d, h->dens, 4ull * h->M, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(gr, h->grads, 12ull * h->M, hipMemcpyDeviceToHost));
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_update_sh(rmesh_renderer* h, const float* sh) {
    if (!h || !sh) return fail(RMESH_ERR_INVALID, "null argument");
    if (!h->has_sh) return fail(RMESH_ERR_INVALID, "renderer has no SH coefficients");
    DeviceGuard g(h->device);
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipMemcpy(h->sh, sh, 4ull * h->sh_total, hipMemcpyHostToDevice));
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_get_sh(rmesh_renderer* h, float* sh) {
    if (!h || !sh) return fail(RMESH_ERR_INVALID, "null argument");
    if (!h->has_sh) return fail(RMESH_ERR_INVALID, "renderer has no SH coefficients");
    DeviceGuard g(h->device);
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipMemcpy(sh, h->sh, 4ull * h->sh_total, hipMemcpyDeviceToHost));
    return RMESH_OK;
}

extern "C" uint32_t rmesh_get_step(const rmesh_renderer* h) { return h ? h->step : 0; }
extern "C" uint32_t rmesh_get_width(const rmesh_renderer* h) { return h ? h->W : 0; }
extern "C" uint32_t rmesh_get_height(const rmesh_renderer* h) { return h ? h->H : 0; }

extern "C" rmesh_status rmesh_forward_batched(rmesh_renderer* h, uint32_t V, const float* cams, const float* vps,
                                              const float* c2ws, float* out_rgba) {
    if (!h || !cams || !vps || !c2ws) return fail(RMESH_ERR_INVALID, "null argument");
    rmesh_status st = check_views(h, V);
    if (st != RMESH_OK) return st;
    DeviceGuard g(h->device);
    st = forward_views(h, V, cams, vps, c2ws, false);
    if (st != RMESH_OK) return st;
    if (out_rgba)
        HIP_TRY(hipMemcpyAsync(out_rgba, h->rgba, 16ull * h->W * h->H * V, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_set_graph_mode(rmesh_renderer* h, int enable) {
    if (!h) return fail(RMESH_ERR_INVALID, "null handle");
    h->graph_mode = enable != 0;
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_set_tuning(rmesh_renderer* h, uint32_t tile_size, int32_t fwd_variant,
                                         int32_t bwd_variant) {
    if (!h) return fail(RMESH_ERR_INVALID, "null handle");
    if (tile_size != 8 && tile_size != 12 && tile_size != 16 && tile_size != 32)
        return fail(RMESH_ERR_INVALID, "tile_size must be 8, 12, 16 or 32");
    if (fwd_variant < 0 || fwd_variant >= fwd_variants(tile_size) || bwd_variant < 0 ||
        bwd_variant >= bwd_variants(tile_size))
        return fail(RMESH_ERR_INVALID, "kernel variant out of range for this tile size");
    if (tile_size != h->TS) set_tiles(h, tile_size);
    if (fwd_variant != h->fwd_var || bwd_variant != h->bwd_var) drop_graph(h);
    h->fwd_var = fwd_variant;
    h->bwd_var = bwd_variant;
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_get_tuning(const rmesh_renderer* h, int32_t* out3) {
    if (!h || !out3) return fail(RMESH_ERR_INVALID, "null argument");
    out3[0] = (int32_t)h->TS;
    out3[1] = h->fwd_var;
    out3[2] = h->bwd_var;
    return RMESH_OK;
}

extern "C" int32_t rmesh_num_variants(int32_t which) { return which == 0 ? kFwdVariants : kBwdVariants; }

extern "C" rmesh_status rmesh_set_sort_digit_bits(rmesh_renderer* h, uint32_t bits) {
    if (!h) return fail(RMESH_ERR_INVALID, "null handle");
    if (bits != 4 && bits != 8) return fail(RMESH_ERR_INVALID, "sort digit width must be 4 or 8 bits");
    if (bits != h->sort_db) drop_graph(h);
    h->sort_db = bits;
    return RMESH_OK;
}

extern "C" uint32_t rmesh_get_sort_digit_bits(const rmesh_renderer* h) { return h ? h->sort_db : 0u; }

// Times the binning + sort at both sort digit widths (8 and 4 bits), every
// forward variant and every adjoint variant for each tile size (8, 12, 16,
// 32) on the given views (eager, HIP events, best of `reps`), then keeps the
// fastest combination.  Parameters and optimiser state are
// untouched; the gradient buffers are cleared afterwards.
static rmesh_status autotune_layout(rmesh_renderer* h, const uint32_t* per_user, const float* cams, const float* vps,
                                    const float* c2ws, const float* gt_dev, uint32_t loss_type, uint32_t reps,
                                    float* out_ms) {
    if (h->sh_half) return fail(RMESH_ERR_UNSUPPORTED, "f16-rounded SH (sh_half) serves the forward only");
    DeviceGuard g(h->device);
    rmesh_status st = upload_layout(h, per_user, cams, vps, c2ws);
    if (st != RMESH_OK) return st;
    const uint32_t V = h->n_views;
    if (gt_dev && gt_dev != h->gt)
        HIP_TRY(hipMemcpyAsync(h->gt, gt_dev, 12ull * h->pix_len, hipMemcpyDeviceToDevice, h->stream));
    if (reps == 0) reps = 3;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    auto timed = [&](auto&& body, float& best) -> rmesh_status {
        best = 3.0e38f;
        for (uint32_t r = 0; r < reps; ++r) {
            HIP_TRY(hipEventRecord(e0, h->stream));
            rmesh_status s2 = body();
            if (s2 != RMESH_OK) return s2;
            HIP_TRY(hipEventRecord(e1, h->stream));
            HIP_TRY(hipEventSynchronize(e1));
            float ms = 0.0f;
            HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
            best = std::min(best, ms);
        }
        return RMESH_OK;
    };
    const uint32_t sizes[4] = {8, 12, 16, 32};
    float best_total = 3.0e38f, best_parts[3] = {0, 0, 0};
    uint32_t best_ts = h->TS, best_db = h->sort_db;
    int best_f = 0, best_b = 0;
    rmesh_status result = RMESH_OK;
    for (uint32_t ts : sizes) {
        set_tiles(h, ts);
        if ((result = upload_layout(h, per_user, cams, vps, c2ws)) != RMESH_OK) break;
        if ((result = stage_binning_front(h, V, false)) != RMESH_OK) break;  // sizes the pair buffers
        float t_bin = 3.0e38f, t_fwd = 3.0e38f, t_bwd = 3.0e38f;
        int bf = 0, bb = 0;
        uint32_t bdb = 8;
        for (uint32_t db : {8u, 4u}) {  // sort digit width
            h->sort_db = db;
            float t = 0.0f;
            result = timed([&]() {
                rmesh_status s2 = stage_binning_front(h, V, true);
                if (s2 != RMESH_OK) return s2;
                return stage_binning_back(h, V);
            }, t);
            if (result != RMESH_OK) break;
            if (t < t_bin) {
                t_bin = t;
                bdb = db;
            }
        }
        if (result != RMESH_OK) break;
        h->sort_db = bdb;
        if ((result = stage_binning_back(h, V)) != RMESH_OK) break;  // lists of the chosen width
        for (int f = 0; f < fwd_variants(ts) && result == RMESH_OK; ++f) {
            h->fwd_var = f;
            float t = 0.0f;
            result = timed([&]() { return stage_forward(h, V, false); }, t);
            if (t < t_fwd) {
                t_fwd = t;
                bf = f;
            }
        }
        if (result != RMESH_OK) break;
        h->fwd_var = bf;
        if ((result = stage_forward(h, V, false)) != RMESH_OK) break;  // image of the chosen variant
        for (int b = 0; b < bwd_variants(ts) && result == RMESH_OK; ++b) {
            h->bwd_var = b;
            float t = 0.0f;
            result = timed([&]() { return stage_backward(h, V, nullptr, h->gt, loss_type); }, t);
            if (t < t_bwd) {
                t_bwd = t;
                bb = b;
            }
        }
        if (result != RMESH_OK) break;
        const float total = t_bin + t_fwd + t_bwd;
        if (total < best_total) {
            best_total = total;
            best_ts = ts;
            best_db = bdb;
            best_f = bf;
            best_b = bb;
            best_parts[0] = t_bin;
            best_parts[1] = t_fwd;
            best_parts[2] = t_bwd;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    set_tiles(h, best_ts);
    h->sort_db = best_db;
    h->fwd_var = best_f;
    h->bwd_var = best_b;
    zero_grads(h, V);
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (result != RMESH_OK) return result;
    if (out_ms) {
        out_ms[0] = best_total;
        out_ms[1] = best_parts[0];
        out_ms[2] = best_parts[1];
        out_ms[3] = best_parts[2];
    }
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_autotune(rmesh_renderer* h, uint32_t V, const float* cams, const float* vps,
                                       const float* c2ws, const float* gt_dev, uint32_t loss_type, uint32_t reps,
                                       float* out_ms) {
    if (!h || !cams || !vps || !c2ws) return fail(RMESH_ERR_INVALID, "null argument");
    if (h->multi()) return fail(RMESH_ERR_UNSUPPORTED, "a multi-scene renderer tunes with rmesh_autotune_multi");
    rmesh_status st = check_views(h, V);
    if (st != RMESH_OK) return st;
    return autotune_layout(h, &V, cams, vps, c2ws, gt_dev, loss_type, reps, out_ms);
}

extern "C" rmesh_status rmesh_autotune_multi(rmesh_renderer* h, const uint32_t* views_per_scene, const float* cams,
                                             const float* vps, const float* c2ws, const float* gt_dev,
                                             uint32_t loss_type, uint32_t reps, float* out_ms) {
    if (!h || !views_per_scene || !cams || !vps || !c2ws) return fail(RMESH_ERR_INVALID, "null argument");
    return autotune_layout(h, views_per_scene, cams, vps, c2ws, gt_dev, loss_type, reps, out_ms);
}

// One train step over a view layout (per_user[u] views of the caller's scene
// u): eager, or a captured hipGraph replayed while the layout, loss type and
// learning rates stay the same.
static rmesh_status train_layout(rmesh_renderer* h, const uint32_t* per_user, const float* cams, const float* vps,
                                 const float* c2ws, const float* gt_dev, uint32_t loss_type, float lr_v, float lr_d,
                                 float lr_g, float lr_sh, float* out_loss, float* out_view_losses) {
    if (h->sh_half) return fail(RMESH_ERR_UNSUPPORTED, "f16-rounded SH (sh_half) serves the forward only");
    DeviceGuard g(h->device);
    rmesh_status st = ensure_optimizer(h);
    if (st != RMESH_OK) return st;
    st = upload_layout(h, per_user, cams, vps, c2ws);
    if (st != RMESH_OK) return st;
    const uint32_t V = h->n_views;
    if (gt_dev && gt_dev != h->gt)
        HIP_TRY(hipMemcpyAsync(h->gt, gt_dev, 12ull * h->pix_len, hipMemcpyDeviceToDevice, h->stream));
    float lrs[5] = {lr_v, lr_d, lr_g, lr_sh, h->lr_base};
    // With timing on, every stage is its own graph and the timing events are
    // recorded between graph launches (never inside a capture).
    const int parts = h->timing ? kStages : 1;
    bool graph_ok = h->graph_parts == parts && h->graph_layout == h->step_views && h->graph_loss == loss_type &&
                    std::memcmp(lrs, h->graph_lrs, sizeof(lrs)) == 0;
    if (h->graph_mode && h->cap > 0) {
        st = ensure_checkpoints(h, V);  // host allocation, never inside a capture (drops a stale graph)
        if (st != RMESH_OK) return st;
        graph_ok = graph_ok && h->graph_parts == parts;
    }
    if (h->graph_mode && h->cap > 0 && !graph_ok) {
        drop_graph(h);
        for (int p = 0; p < parts; ++p) {
            HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
            rmesh_status cst = RMESH_OK;
            for (int i = (parts == 1 ? 0 : p); i < (parts == 1 ? kStages : p + 1) && cst == RMESH_OK; ++i)
                cst = record_stage(h, i, V, h->gt, loss_type, lr_v, lr_d, lr_g, lr_sh, true);
            hipGraph_t graph = nullptr;
            hipError_t e = hipStreamEndCapture(h->stream, &graph);
            if (cst != RMESH_OK || e != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                drop_graph(h);
                if (cst != RMESH_OK) return cst;
                return fail(RMESH_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
            }
            h->graph[p] = graph;
            hipError_t ie = hipGraphInstantiate(&h->graph_exec[p], graph, nullptr, nullptr, 0);
            if (ie != hipSuccess) {
                drop_graph(h);
                return fail(RMESH_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
            }
        }
        h->graph_parts = parts;
        h->graph_views = V;
        h->graph_layout = h->step_views;
        h->graph_loss = loss_type;
        std::memcpy(h->graph_lrs, lrs, sizeof(lrs));
        graph_ok = true;
    }
    if (h->graph_mode && graph_ok) {
        for (int p = 0; p < h->graph_parts; ++p) {
            if (h->graph_parts > 1) mark(h, p);
            HIP_TRY(hipGraphLaunch(h->graph_exec[p], h->stream));
        }
        if (h->graph_parts > 1) mark(h, kStages);
        st = finish_train(h, V, out_loss, out_view_losses);
        if (st == RMESH_ERR_CAPACITY) {
            // Parameters were left untouched on overflow: grow and redo eagerly.
            uint32_t P = 0;
            HIP_TRY(hipMemcpy(&P, h->n_pairs, 4, hipMemcpyDeviceToHost));
            st = ensure_capacity(h, P);
            if (st != RMESH_OK) return st;
            st = ensure_checkpoints(h, V);
            if (st != RMESH_OK) return st;
            st = record_train(h, V, h->gt, loss_type, lr_v, lr_d, lr_g, lr_sh, false);
            if (st != RMESH_OK) return st;
            return finish_train(h, V, out_loss, out_view_losses);
        }
        return st;
    }
    st = record_train(h, V, h->gt, loss_type, lr_v, lr_d, lr_g, lr_sh, false);
    if (st != RMESH_OK) return st;
    return finish_train(h, V, out_loss, out_view_losses);
}

extern "C" rmesh_status rmesh_train_step_batched(rmesh_renderer* h, uint32_t V, const float* cams, const float* vps,
                                                 const float* c2ws, const float* gt_dev, uint32_t loss_type, float lr_v,
                                                 float lr_d, float lr_g, float lr_sh, float* out_loss) {
    if (!h || !cams || !vps || !c2ws) return fail(RMESH_ERR_INVALID, "null argument");
    if (h->multi()) return fail(RMESH_ERR_UNSUPPORTED, "a multi-scene renderer trains with rmesh_train_step_multi");
    rmesh_status st = check_views(h, V);
    if (st != RMESH_OK) return st;
    return train_layout(h, &V, cams, vps, c2ws, gt_dev, loss_type, lr_v, lr_d, lr_g, lr_sh, out_loss, nullptr);
}

extern "C" rmesh_status rmesh_train_step_multi(rmesh_renderer* h, const uint32_t* views_per_scene, const float* cams,
                                               const float* vps, const float* c2ws, const float* gt_dev,
                                               uint32_t loss_type, float lr_v, float lr_d, float lr_g, float lr_sh,
                                               float* out_view_losses, float* out_loss) {
    if (!h || !views_per_scene || !cams || !vps || !c2ws) return fail(RMESH_ERR_INVALID, "null argument");
    return train_layout(h, views_per_scene, cams, vps, c2ws, gt_dev, loss_type, lr_v, lr_d, lr_g, lr_sh, out_loss,
                        out_view_losses);
}

extern "C" rmesh_status rmesh_forward_multi(rmesh_renderer* h, const uint32_t* views_per_scene, const float* cams,
                                            const float* vps, const float* c2ws, float* out_rgba) {
    if (!h || !views_per_scene || !cams || !vps || !c2ws) return fail(RMESH_ERR_INVALID, "null argument");
    DeviceGuard g(h->device);
    rmesh_status st = upload_layout(h, views_per_scene, cams, vps, c2ws);
    if (st != RMESH_OK) return st;
    const uint32_t V = h->n_views;
    if ((st = stage_binning_front(h, V, false)) != RMESH_OK) return st;
    if ((st = stage_binning_back(h, V)) != RMESH_OK) return st;
    if ((st = stage_forward(h, V, false)) != RMESH_OK) return st;
    h->last_views = V;
    h->have_forward = !h->multi();
    if (out_rgba) HIP_TRY(hipMemcpyAsync(out_rgba, h->rgba, 16ull * h->pix_len, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    uint32_t ovf = 0;
    HIP_TRY(hipMemcpy(&ovf, h->overflow, 4, hipMemcpyDeviceToHost));
    if (ovf) return fail(RMESH_ERR_CAPACITY, "pair buffer overflow");
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_set_targets_multi(rmesh_renderer* h, const uint32_t* views_per_scene, const float* gt) {
    if (!h || !views_per_scene || !gt) return fail(RMESH_ERR_INVALID, "null argument");
    uint64_t px = 0;
    for (uint32_t u = 0; u < (uint32_t)h->sc.size(); ++u) {
        const SceneHost& sh = h->sc[h->sc_of_user[u]];
        if (views_per_scene[u] > sh.max_views) return fail(RMESH_ERR_INVALID, "more views than max_views");
        px += (uint64_t)views_per_scene[u] * sh.W * sh.H;
    }
    DeviceGuard g(h->device);
    HIP_TRY(hipMemcpyAsync(h->gt, gt, 12ull * px, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return RMESH_OK;
}

extern "C" uint32_t rmesh_scene_count(const rmesh_renderer* h) { return h ? (uint32_t)h->sc.size() : 0u; }

// Per-scene slice of a parameter or gradient buffer (RMESH_BUF_*), in the
// scene's own indexing.
extern "C" rmesh_status rmesh_read_scene(rmesh_renderer* h, uint32_t scene, int which, void* host, uint64_t nbytes) {
    if (!h || !host) return fail(RMESH_ERR_INVALID, "null argument");
    if (scene >= h->sc.size()) return fail(RMESH_ERR_INVALID, "no such scene");
    const SceneHost& c = h->sc[h->sc_of_user[scene]];
    uint64_t off = 0, cap = 0;  // bytes
    switch (which) {
        case RMESH_BUF_VERTICES: case RMESH_BUF_D_VERTICES: off = 12ull * c.vert_base; cap = 12ull * c.N; break;
        case RMESH_BUF_DENSITIES: case RMESH_BUF_D_DENSITIES: off = 4ull * c.tet_base; cap = 4ull * c.M; break;
        case RMESH_BUF_COLOR_GRADS: case RMESH_BUF_BASE_COLORS: case RMESH_BUF_D_COLOR_GRADS:
            off = 12ull * c.tet_base;
            cap = 12ull * c.M;
            break;
        case RMESH_BUF_SH: case RMESH_BUF_D_SH:
            if (!h->has_sh) return fail(RMESH_ERR_INVALID, "renderer has no SH coefficients");
            off = 4ull * c.sh_base;
            cap = 12ull * c.M * c.nc;
            break;
        default: return fail(RMESH_ERR_INVALID, "buffer has no per-scene slice");
    }
    if (nbytes > cap) return fail(RMESH_ERR_INVALID, "read beyond the scene's slice");
    char* src = static_cast<char*>(rmesh_device_buffer(h, which));
    if (!src) return fail(RMESH_ERR_INVALID, "buffer not allocated");
    DeviceGuard g(h->device);
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (nbytes) HIP_TRY(hipMemcpy(host, src + off, nbytes, hipMemcpyDeviceToHost));
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_set_scene_timing(rmesh_renderer* h, int enable) {
    if (!h) return fail(RMESH_ERR_INVALID, "null handle");
    DeviceGuard g(h->device);
    if (enable && !h->scene_clock) HIP_TRY(dalloc(&h->scene_clock, h->sc.size() + 1));
    if ((enable != 0) != h->scene_timing) {
        drop_graph(h);
        h->scene_timing = enable != 0;
    }
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_set_scene_order(rmesh_renderer* h, uint32_t order) {
    if (!h) return fail(RMESH_ERR_INVALID, "null handle");
    if (order != RMESH_SCENE_ORDER_INTERLEAVED && order != RMESH_SCENE_ORDER_BY_SCENE)
        return fail(RMESH_ERR_INVALID, "scene order must be RMESH_SCENE_ORDER_INTERLEAVED or _BY_SCENE");
    if (order == RMESH_SCENE_ORDER_BY_SCENE && h->sc.size() > 64)
        return fail(RMESH_ERR_UNSUPPORTED, "scene-major order holds at most 64 scenes");
    if (order != h->scene_order) drop_graph(h);
    h->scene_order = order;
    return RMESH_OK;
}

extern "C" uint32_t rmesh_get_scene_order(const rmesh_renderer* h) { return h ? h->scene_order : 0u; }

// Per scene (the caller's order): ms from the last step's start to the end of
// the scene's last adjoint tile.
extern "C" rmesh_status rmesh_scene_latency(rmesh_renderer* h, float* ms) {
    if (!h || !ms) return fail(RMESH_ERR_INVALID, "null argument");
    if (!h->scene_timing) return fail(RMESH_ERR_INVALID, "scene timing is off");
    DeviceGuard g(h->device);
    const size_t n = h->sc.size();
    std::vector<uint64_t> c(n + 1);
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipMemcpy(c.data(), h->scene_clock, 8 * (n + 1), hipMemcpyDeviceToHost));
    int khz = 0;
    HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, h->device));
    if (khz <= 0) return fail(RMESH_ERR_DEVICE, "no wall clock rate");
    for (size_t u = 0; u < n; ++u) ms[u] = c[u] > c[n] ? (float)((double)(c[u] - c[n]) / (double)khz) : 0.0f;
    return RMESH_OK;
}



extern "C" void* rmesh_device_buffer(rmesh_renderer* h, int which) {
    if (!h) return nullptr;
    switch (which) {
        case RMESH_BUF_RGBA: return h->rgba;
        case RMESH_BUF_GT: return h->gt;
        case RMESH_BUF_VERTICES: return h->verts;
        case RMESH_BUF_DENSITIES: return h->dens;
        case RMESH_BUF_COLOR_GRADS: return h->grads;
        case RMESH_BUF_BASE_COLORS: return h->base;
        case RMESH_BUF_SH: return h->sh;
        case RMESH_BUF_D_VERTICES: return h->d_verts;
        case RMESH_BUF_D_DENSITIES: return h->d_dens;
        case RMESH_BUF_D_COLOR_GRADS: return h->d_grads;
        case RMESH_BUF_D_SH: return h->d_sh;
        case RMESH_BUF_D_BASE_COLORS: return h->d_colors;
        default: return nullptr;
    }
}

extern "C" rmesh_status rmesh_synchronize(rmesh_renderer* h) {
    if (!h) return fail(RMESH_ERR_INVALID, "null handle");
    DeviceGuard g(h->device);
    HIP_TRY(hipStreamSynchronize(h->stream));
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_debug_tiles_touched(rmesh_renderer* h, uint32_t view, uint32_t* out) {
    if (!h || !out) return fail(RMESH_ERR_INVALID, "null argument");
    if (!h->have_forward || view >= h->last_views) return fail(RMESH_ERR_INVALID, "no such view in the last forward");
    DeviceGuard g(h->device);
    HIP_TRY(hipStreamSynchronize(h->stream));
    std::vector<uint32_t> all((size_t)h->M * h->last_views);
    HIP_TRY(hipMemcpy(all.data(), h->counts, all.size() * 4, hipMemcpyDeviceToHost));
    for (uint32_t t = 0; t < h->M; ++t) out[t] = all[(size_t)t * h->last_views + view];
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_debug_pairs(rmesh_renderer* h, const float* cam, const float* vp, const float* ci,
                                          uint64_t* keys, uint32_t* vals, uint64_t cap, uint64_t* n_pairs) {
    if (!h || !cam || !vp || !ci || !n_pairs) return fail(RMESH_ERR_INVALID, "null argument");
    DeviceGuard g(h->device);
    rmesh_status st = upload_views(h, 1, cam, vp, ci);
    if (st != RMESH_OK) return st;
    st = stage_binning_front(h, 1, false);
    if (st != RMESH_OK) return st;
    launch_pair_gen_item_order(h, h->stream);
    uint32_t P = 0;
    HIP_TRY(hipMemcpyAsync(&P, h->n_pairs, 4, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    *n_pairs = P;
    uint64_t n = std::min<uint64_t>(P, cap);
    if (keys && n) HIP_TRY(hipMemcpy(keys, h->keys_a, n * 8, hipMemcpyDeviceToHost));
    if (vals && n) HIP_TRY(hipMemcpy(vals, h->vals_a, n * 4, hipMemcpyDeviceToHost));
    h->have_forward = false;
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_debug_sorted(rmesh_renderer* h, uint64_t* keys, uint32_t* vals, uint64_t cap,
                                           uint64_t* n_pairs, uint32_t* ranges, uint64_t ranges_cap) {
    if (!h || !n_pairs) return fail(RMESH_ERR_INVALID, "null argument");
    if (!h->have_forward) return fail(RMESH_ERR_INVALID, "no forward has run");
    DeviceGuard g(h->device);
    HIP_TRY(hipStreamSynchronize(h->stream));
    uint32_t P = 0;
    HIP_TRY(hipMemcpy(&P, h->n_pairs, 4, hipMemcpyDeviceToHost));
    *n_pairs = P;
    uint64_t n = std::min<uint64_t>(P, cap);
    uint64_t* sk = h->sorted_in_b ? h->keys_b : h->keys_a;
    if (keys && n && h->pairs_depth_ordered) {
        // 32-bit key tiles: rebuild the 64-bit (tile, depth) keys in the free key buffer
        uint64_t* full = h->sorted_in_b ? h->keys_a : h->keys_b;
        launch_full_keys(keys32(sk), sorted_vals(h), h->n_pairs, h->cap, reinterpret_cast<const float4*>(h->circ),
                         h->d_views, h->d_ktile, h->n_kt, h->d_tile_view, full, P, h->stream);
        HIP_TRY(hipStreamSynchronize(h->stream));
        sk = full;
    }
    if (keys && n) HIP_TRY(hipMemcpy(keys, sk, n * 8, hipMemcpyDeviceToHost));
    if (vals && n) HIP_TRY(hipMemcpy(vals, sorted_vals(h), n * 4, hipMemcpyDeviceToHost));
    uint64_t nr = std::min<uint64_t>(ranges_cap, 2ull * h->gtiles(h->last_views));
    if (ranges && nr) HIP_TRY(hipMemcpy(ranges, h->ranges, nr * 4, hipMemcpyDeviceToHost));
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t n, uint32_t bits, int32_t device_id) {
    if ((!keys || !vals) && n) return fail(RMESH_ERR_INVALID, "null argument");
    if (bits % 8 != 0 || bits > 64) return fail(RMESH_ERR_INVALID, "sorting_bits must be a multiple of 8 up to 64");
    if (n > 0xFFFFFFF0ull) return fail(RMESH_ERR_INVALID, "too many pairs");
    if (n == 0) return RMESH_OK;
    int dev = device_id;
    if (dev < 0) (void)hipGetDevice(&dev);
    DeviceGuard g(dev);
    uint32_t cap = (uint32_t)n;
    uint64_t *ka = nullptr, *kb = nullptr;
    uint32_t *va = nullptr, *vb = nullptr, *hist = nullptr, *np = nullptr;
    hipStream_t s = nullptr;
    rmesh_status st = RMESH_OK;
    do {
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
            dalloc(&ka, cap) != hipSuccess || dalloc(&kb, cap) != hipSuccess || dalloc(&va, cap) != hipSuccess ||
            dalloc(&vb, cap) != hipSuccess || dalloc(&hist, radix_hist_words(cap)) != hipSuccess ||
            dalloc(&np, 1) != hipSuccess) {
            st = fail(RMESH_ERR_DEVICE, "allocation failed");
            break;
        }
        if (hipMemcpy(ka, keys, n * 8, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(va, vals, n * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(np, &cap, 4, hipMemcpyHostToDevice) != hipSuccess) {
            st = fail(RMESH_ERR_DEVICE, "sort: upload failed");
            break;
        }
        bool in_b = launch_radix_sort(ka, va, kb, vb, np, cap, bits, hist, s);
        hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            st = fail(RMESH_ERR_DEVICE, std::string("sort: ") + hipGetErrorString(e));
            break;
        }
        if (hipMemcpy(keys, in_b ? kb : ka, n * 8, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(vals, in_b ? vb : va, n * 4, hipMemcpyDeviceToHost) != hipSuccess)
            st = fail(RMESH_ERR_DEVICE, "sort: download failed");
    } while (0);
    (void)hipFree(ka);
    (void)hipFree(kb);
    (void)hipFree(va);
    (void)hipFree(vb);
    (void)hipFree(hist);
    (void)hipFree(np);
    if (s) (void)hipStreamDestroy(s);
    return st;
}

extern "C" rmesh_status rmesh_set_targets(rmesh_renderer* h, uint32_t n_views, const float* gt) {
    if (!h || !gt) return fail(RMESH_ERR_INVALID, "null argument");
    rmesh_status st = check_views(h, n_views);
    if (st != RMESH_OK) return st;
    DeviceGuard g(h->device);
    HIP_TRY(hipMemcpyAsync(h->gt, gt, 12ull * h->W * h->H * n_views, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_set_timing(rmesh_renderer* h, int enable) {
    if (!h) return fail(RMESH_ERR_INVALID, "null handle");
    DeviceGuard g(h->device);
    if (enable && !h->ev[0])
        for (hipEvent_t& e : h->ev) HIP_TRY(hipEventCreate(&e));
    if ((enable != 0) != h->timing) {
        h->timing = enable != 0;
        drop_graph(h);
    }
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_stage_times(rmesh_renderer* h, float* ms) {
    if (!h || !ms) return fail(RMESH_ERR_INVALID, "null argument");
    if (!h->timing) return fail(RMESH_ERR_INVALID, "timing is not enabled");
    DeviceGuard g(h->device);
    HIP_TRY(hipStreamSynchronize(h->stream));
    for (int i = 0; i < 7; ++i) HIP_TRY(hipEventElapsedTime(&ms[i], h->ev[i], h->ev[i + 1]));
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_pair_count(rmesh_renderer* h, uint64_t* out) {
    if (!h || !out) return fail(RMESH_ERR_INVALID, "null argument");
    DeviceGuard g(h->device);
    HIP_TRY(hipStreamSynchronize(h->stream));
    uint32_t P = 0;
    HIP_TRY(hipMemcpy(&P, h->n_pairs, 4, hipMemcpyDeviceToHost));
    *out = P;
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_read_buffer(rmesh_renderer* h, int which, void* host, uint64_t nbytes) {
    if (!h || !host) return fail(RMESH_ERR_INVALID, "null argument");
    uint64_t cap = 0;
    switch (which) {
        case RMESH_BUF_RGBA: cap = 16ull * h->pix_cap; break;
        case RMESH_BUF_GT: cap = 12ull * h->pix_cap; break;
        case RMESH_BUF_VERTICES: case RMESH_BUF_D_VERTICES: cap = 12ull * h->N; break;
        case RMESH_BUF_DENSITIES: case RMESH_BUF_D_DENSITIES: cap = 4ull * h->M; break;
        case RMESH_BUF_COLOR_GRADS: case RMESH_BUF_BASE_COLORS: case RMESH_BUF_D_COLOR_GRADS: cap = 12ull * h->M; break;
        case RMESH_BUF_SH: case RMESH_BUF_D_SH: cap = h->has_sh ? 4ull * h->sh_total : 0; break;
        case RMESH_BUF_D_BASE_COLORS: cap = 12ull * h->color_len; break;
        default: return fail(RMESH_ERR_INVALID, "unknown buffer");
    }
    if (nbytes > cap) return fail(RMESH_ERR_INVALID, "read beyond the buffer");
    void* src = rmesh_device_buffer(h, which);
    if (!src) return fail(RMESH_ERR_INVALID, "buffer not allocated");
    DeviceGuard g(h->device);
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipMemcpy(host, src, nbytes, hipMemcpyDeviceToHost));
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_atomic_probe(int32_t device_id, uint32_t lines, uint64_t n_ops, float* ms) {
    if (lines < 977 || lines > (1u << 24)) return fail(RMESH_ERR_INVALID, "lines must be in [977, 2^24]");
    const uint32_t per_thread = 64;
    const uint64_t threads = n_ops / per_thread;
    if (threads < 256 || threads / 256 > 0xFFFFFFFFull) return fail(RMESH_ERR_INVALID, "n_ops out of range");
    const uint32_t blocks = (uint32_t)(threads / 256);
    int dev = device_id;
    if (dev < 0) (void)hipGetDevice(&dev);
    DeviceGuard g(dev);
    float* p = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    rmesh_status st = RMESH_OK;
    if (dalloc(&p, 16ull * lines) != hipSuccess || hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        st = fail(RMESH_ERR_DEVICE, "atomic probe allocation failed");
    } else {
        launch_fill_u32(reinterpret_cast<uint32_t*>(p), 0u, 16ull * lines, s);
        launch_atomic_probe(p, lines, blocks, per_thread, s);  // warm
        (void)hipEventRecord(e0, s);
        launch_atomic_probe(p, lines, blocks, per_thread, s);
        (void)hipEventRecord(e1, s);
        hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = fail(RMESH_ERR_DEVICE, std::string("atomic probe: ") + hipGetErrorString(e));
        else if (ms) (void)hipEventElapsedTime(ms, e0, e1);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    (void)hipFree(p);
    return st;
}

extern "C" rmesh_status rmesh_copy_probe(int32_t device_id, uint64_t nbytes, uint32_t iters, float* ms_per_copy) {
    if (nbytes < 16 || nbytes % 16) return fail(RMESH_ERR_INVALID, "nbytes must be a positive multiple of 16");
    int dev = device_id;
    if (dev < 0) (void)hipGetDevice(&dev);
    DeviceGuard g(dev);
    float4 *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    rmesh_status st = RMESH_OK;
    if (dalloc(&a, nbytes / 16) != hipSuccess || dalloc(&b, nbytes / 16) != hipSuccess ||
        hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        st = fail(RMESH_ERR_DEVICE, "copy probe allocation failed");
    } else {
        launch_fill_u32(reinterpret_cast<uint32_t*>(a), 0x3f800000u, nbytes / 4, s);
        launch_copy_probe(a, b, nbytes / 16, s);
        (void)hipEventRecord(e0, s);
        for (uint32_t i = 0; i < iters; ++i) launch_copy_probe(a, b, nbytes / 16, s);
        (void)hipEventRecord(e1, s);
        hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = fail(RMESH_ERR_DEVICE, std::string("copy probe: ") + hipGetErrorString(e));
        else if (ms_per_copy) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            *ms_per_copy = iters ? ms / iters : 0.f;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    (void)hipFree(a);
    (void)hipFree(b);
    return st;
}
