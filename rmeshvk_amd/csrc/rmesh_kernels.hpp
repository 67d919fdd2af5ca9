// Host-side launchers of the rmesh HIP kernels (internal to the library).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace rmesh {

struct ViewParams;

// binning (k_binning.hip)
void launch_prep_records(const float* verts, const uint32_t* idx, const float* dens, const float* grads, uint32_t M,
                         float4* rec, hipStream_t s);
void launch_copy_colors(const float* base, uint32_t M, float4* col, hipStream_t s);
void launch_fill_u32(uint32_t* p, uint32_t value, uint64_t n, hipStream_t s);
// blocks x 256 lanes x per_thread float atomic adds over `lines` 64-byte lines (lines >= 977)
void launch_atomic_probe(float* p, uint32_t lines, uint32_t blocks, uint32_t per_thread, hipStream_t s);
void launch_copy_probe(const float4* src, float4* dst, uint64_t n16, hipStream_t s);
uint32_t item_blocks(uint64_t n_items);  // workgroups of the (tet, view) item kernels
// tile_size: the views' (one per renderer); a power of two takes the division-free path.
// block_vis: per block, the visible items of view 0 (their sum is the instance count).
// dkey / ditem (optional, with n_items_out): the depth sort's keys and item indices.
void launch_project(const float4* rec, const float4* circ, const SceneRange* scenes, uint32_t n_scenes,
                    uint64_t n_items, const ViewParams* views, uint32_t tile_size, uint32_t* counts,
                    uint32_t* block_sums, uint32_t* block_vis, uint32_t* dkey, uint32_t* ditem,
                    uint32_t* n_items_out, hipStream_t s);
// Exclusive scan of sums[0, n) in place, total to *total; with aux, the sum of aux[0, n) to *aux_total.
// part (spine_parts(n) entries) and aux_part (the same count): scratch of the two-level form.
uint32_t spine_parts(uint32_t n);
void launch_spine_scan(uint32_t* sums, uint32_t n, uint32_t* total, uint32_t* overflow, hipStream_t s,
                       const uint32_t* aux = nullptr, uint32_t* aux_total = nullptr,
                       unsigned long long* part = nullptr, uint32_t* aux_part = nullptr);
void launch_pairs(const float4* rec, const float4* circ, const SceneRange* scenes, uint32_t n_scenes, uint64_t n_items,
                  const ViewParams* views, uint32_t tile_size, const uint32_t* counts, const uint32_t* block_off,
                  uint64_t* keys, uint32_t* vals, uint32_t cap, uint32_t* overflow, hipStream_t s);
// Depth-ordered pair generation from the items' depth-sorted values sval
// (item << 5 | count, n_items < 2^27): block sums into block_sums, the spine
// scan (total to n_pairs), the pairs in that order with 32-bit keys (the key
// tile).
void launch_pairs_depth_ordered(const float4* rec, const SceneRange* scenes, uint32_t n_scenes, const uint32_t* sval,
                                uint32_t n_items, const ViewParams* views, uint32_t tile_size, const uint32_t* counts,
                                uint32_t* block_sums, uint32_t* n_pairs, uint32_t* keys, uint32_t* vals, uint32_t cap,
                                uint32_t* overflow, unsigned long long* spine_part, hipStream_t s);
// 64-bit keys (key tile << 32 | depth key) of sorted 32-bit-key pairs, for inspection
void launch_full_keys(const uint32_t* ktile, const uint32_t* vals, const uint32_t* n_ptr, uint32_t cap,
                      const float4* circ, const ViewParams* views, const uint32_t* ktile_gtile, uint32_t n_ktiles,
                      const uint32_t* tile_view, uint64_t* out, uint32_t n_host, hipStream_t s);
// ranges indexed by global tile; ktile_gtile maps every key tile to its global tile (0xFFFFFFFF: sentinel)
void launch_ranges(const uint64_t* keys, const uint32_t* n_ptr, uint32_t cap, uint32_t* ranges,
                   const uint32_t* ktile_gtile, uint32_t n_ktiles, hipStream_t s);
void launch_ranges32(const uint32_t* keys, const uint32_t* n_ptr, uint32_t cap, uint32_t* ranges,
                     const uint32_t* ktile_gtile, uint32_t n_ktiles, hipStream_t s);
// Permutation of the global tiles, heaviest (most pairs) first.
void launch_tile_order(const uint32_t* ranges, uint32_t n_tiles, uint32_t* order, hipStream_t s);
// Scene-major heaviest-first order (scenes by ascending pair count); false
// (nothing launched) above 64 scenes.
bool launch_tile_order_scenes(const uint32_t* ranges, uint32_t n_tiles, const uint32_t* tile_view,
                              const ViewParams* views, uint32_t n_scenes, uint32_t* order, hipStream_t s);
// p[0..n) = 0, p[n] = the wall clock now (start of a step's per-scene timing)
void launch_stamp_start(uint64_t* p, uint32_t n, hipStream_t s);

// sort (k_sort.hip)
// digit_bits: 8 (default) or 4; sorting_bits and first_bit must be multiples
// of it.  Passes over key bits [first_bit, sorting_bits): input already in
// order of the bits below first_bit stays so among equal upper bits.
bool launch_radix_sort(uint64_t* ka, uint32_t* va, uint64_t* kb, uint32_t* vb, const uint32_t* n_ptr, uint32_t cap,
                       uint32_t sorting_bits, uint32_t* hist, hipStream_t s, uint32_t digit_bits = 8,
                       uint32_t first_bit = 0);
size_t radix_hist_words(uint32_t cap);
// u32 keys, bits [0, end_bit)
bool launch_radix_sort32(uint32_t* ka, uint32_t* va, uint32_t* kb, uint32_t* vb, const uint32_t* n_ptr, uint32_t cap,
                         uint32_t* hist, hipStream_t s, uint32_t digit_bits = 8, uint32_t end_bit = 32);
size_t radix_hist_words32(uint32_t cap);

// compositing (k_composite.hip)
struct FwdArgs {
    const float4* rec;
    const float4* colors;        // view v's colours at colors[views[v].color_off + t]
    const uint32_t* sorted_vals;
    const uint32_t* ranges;
    const ViewParams* views;
    const uint32_t* tile_view;   // global tile -> view
    uint32_t num_gtiles;
    const uint32_t* tile_order;  // nullable: workgroup -> global tile (identity when null)
    float* rgba;
    float* aux;  // nullable: [V][H][W][8]
    // nullable: transmittance checkpoints for the checkpointed adjoint.
    // ck[slot][TS*TS]: a pixel's log T before tile-list position 32 j (j >= 1,
    // 32 j < the tile's pair count) at slot (range start >> 5) + global tile + j - 1;
    // fin[V][H][W]: (final log T, stop position as the float's bits).
    float* ck;
    float2* fin;
};
// Batch-geometry variants of the compositing kernels (rmesh_autotune picks one).
constexpr int kFwdVariants = 8;
constexpr int kBwdVariants = 11;
// tile 32 compiles fewer geometries (a batch must hold one tet's 1024 pixels)
inline int fwd_variants(uint32_t tile_size) { return tile_size == 32 ? 2 : kFwdVariants; }
inline int bwd_variants(uint32_t tile_size) { return tile_size == 32 ? 2 : kBwdVariants; }
bool launch_forward(const FwdArgs& a, uint32_t tile_size, int variant, hipStream_t s);

// Per-tet error statistics (forward_error) over one view's tile lists.
struct ErrArgs {
    const float4* rec;
    const uint32_t* sorted_vals;
    const uint32_t* ranges;
    const ViewParams* views;  // views[0]
    uint32_t num_tiles;
    const float* pixel_err;   // [H][W]
    const float* ssim_err;    // [H][W]
    float* tet_err;           // [M][16], accumulated
    int32_t* tet_count;       // [M][2], accumulated (count, max of i32(65535 T_val))
};
bool launch_error(const ErrArgs& a, uint32_t tile_size, hipStream_t s);

// Interval-tiled forward of view 0: silhouette decomposition of every tet,
// then compositing over the regular binning's tile lists.
struct IvlArgs {
    const float4* rec;
    const float4* colors;        // view 0's colours
    const ViewParams* views;
    uint32_t M;
    float4* ivr;                 // [M][7] decomposition (written, then read)
    const uint32_t* sorted_vals;
    const uint32_t* ranges;
    const uint32_t* tile_order;  // nullable, as FwdArgs::tile_order
    uint32_t num_tiles;
    float* rgba;                 // [H][W][4]
    float* xyzd;                 // [H][W][4]
    float* dist;                 // [H][W][5]
};
bool launch_interval(const IvlArgs& a, uint32_t tile_size, hipStream_t s);

// Interval-tiled adjoint of view 0 over the last interval forward's tile
// lists and decomposition, then the per-tet chain back to the scene.
struct IvlBwdArgs {
    const float4* rec;
    const uint32_t* idx;
    const ViewParams* views;
    uint32_t M;
    const float4* ivr;           // [M][7] decomposition of the forward
    const uint32_t* sorted_vals;
    const uint32_t* ranges;
    const uint32_t* tile_order;  // nullable
    uint32_t num_tiles;
    const float* rgba;           // forward image [H][W][4]
    const float* dist;           // forward distortion [H][W][5]
    const float* dl_img;         // [H][W][4]
    const float* dl_xyzd;        // [H][W][4] or nullptr
    const float* dl_dist;        // [H][W][5] or nullptr
    float* d_ivr;                // [M][48] zeroed scratch
    float* tv;                   // [M][32] per-tet vertex terms (scratch, written by the chain back)
    const uint32_t* v2t_off;     // [N + 1] vertex -> incidence list offsets
    const uint32_t* v2t;         // [4M] incidences: tet << 2 | vertex slot, per vertex
    uint32_t N;
    float* d_verts;              // [N*3] written
    float* d_dens;               // [M] zeroed; written per decomposed tet
    float* d_base;               // [M*3] zeroed; written per decomposed tet
    float* d_grads;              // [M*3] zeroed; written per decomposed tet
    float* d_vn;                 // [N*3] written
};
bool launch_interval_backward(const IvlBwdArgs& a, uint32_t tile_size, hipStream_t s);

struct BwdArgs {
    const float4* rec;
    const uint32_t* idx;
    const float4* colors;
    const uint32_t* sorted_vals;
    const uint32_t* ranges;
    const ViewParams* views;
    const uint32_t* tile_view;
    uint32_t num_gtiles;
    const uint32_t* tile_order;  // nullable, as FwdArgs::tile_order
    const float* rgba;
    const float* dl_img;      // used when gt == nullptr
    const float* gt;          // [V][H][W][3] or nullptr
    uint32_t loss_type;
    float* loss_partial;      // [num_gtiles] or nullptr
    // per tet t, 16 floats at tet_g[16 t]: density, vertex slots 0..3 (xyz),
    // colour gradient (xyz); launch_tet_gather moves them to the parameter
    // gradients
    float* tet_g;
    float* d_colors;          // view v's colour gradients at d_colors[3 (views[v].color_off + t)]
    // nullable: the forward's checkpoints (FwdArgs::ck / fin); set: the
    // checkpointed adjoint, null: the undo from the image alpha
    const float* ck;
    const float2* fin;
    // nullable: per scene, the wall clock (s_memrealtime) when its last tile
    // of this adjoint retired (atomic max over the scene's workgroups)
    uint64_t* scene_done;
};
// Tile-list positions per transmittance checkpoint.
constexpr uint32_t kCkGroup = 32;
// Checkpoint slots for P pairs over G global tiles (each slot TS*TS floats).
inline uint64_t checkpoint_slots(uint64_t P, uint64_t G) { return (P >> 5) + G + 1; }
bool launch_backward(const BwdArgs& a, uint32_t tile_size, int variant, hipStream_t s);
// d_dens / d_grads per tet and d_verts per vertex (sum over the vertex's
// (tet << 2 | slot) incidence list off / ent) out of the adjoint's tet_g.
void launch_tet_gather(const uint32_t* off, const uint32_t* ent, const float* tet_g, uint32_t N, uint32_t M,
                       float* d_verts, float* d_dens, float* d_grads, hipStream_t s);

// colours, SH and optimiser (k_optim.hip)
// half: coefficients rounded to f16 first (the reference viewer's packed SH)
// colours of the tets [t0, t0 + n) of one SH degree (coefficients from sh) in
// every view of each tet's scene
void launch_sh_colors(const float4* rec, const float* sh, uint32_t deg, uint32_t t0, uint32_t n, const SceneRange* scenes,
                      uint32_t n_scenes, const ViewParams* views, float4* colors, bool half, hipStream_t s);

// The step counter and the overflow flag live in device memory so a captured
// graph reads the current values at replay time.
struct AdamStep {
    const uint32_t* step;   // device: Adam step after this update (>= 1)
    const uint32_t* skip;   // device: non-zero -> leave parameters untouched
    float beta1, beta2, eps;
};
// SH adjoint of the tets [t0, t0 + M) of one degree; sh, d_sh, m_sh, v_sh
// point at the first of them.
struct ShBwdArgs {
    const float4* rec;
    float* sh;               // parameters (updated in place when adam)
    uint32_t deg;
    uint32_t t0;
    uint32_t M;
    const SceneRange* scenes;
    uint32_t n_scenes;
    const ViewParams* views;
    const float* d_colors;   // view v's at d_colors[3 (views[v].color_off + t)]
    float* d_sh;             // output gradient (when !adam)
    float* tet_g;            // += per-tet staging record (16 floats, the adjoint's); one quad owns a tet
    bool adam;
    float* m_sh;
    float* v_sh;
    float lr;
    AdamStep st;
};
void launch_sh_backward(const ShBwdArgs& a, hipStream_t s);

struct AdamGroup {
    float* p;
    const float* g;
    float* m;
    float* v;
    uint64_t n;
    float lr;
};
constexpr int kMaxAdamGroups = 6;
struct AdamArgs {
    AdamGroup grp[kMaxAdamGroups];
    int ngroups;
    AdamStep st;
};
void launch_adam(const AdamArgs& a, hipStream_t s);
void launch_step_inc(uint32_t* step, const uint32_t* skip, hipStream_t s);
// per view: the sum of its tiles' partial losses over its pixel count
void launch_reduce_loss(const float* partial, const ViewParams* views, uint32_t V, float* out,
                        hipStream_t s);

}  // namespace rmesh
