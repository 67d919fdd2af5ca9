/*
 * rmesh_hip — C ABI of the MI355X-native differentiable radiance-mesh
 * train step (projection + tile binning, depth radix sort, tiled compositing,
 * reverse-order adjoint, loss and Adam).
 *
 * This is the drop-in boundary for rmeshvk's `RMeshRenderer` pyclass
 * (crates/rmesh-python/src/lib.rs:138-2705): each entry point below names the
 * method it replaces.  Plain pointers and sizes only; "host" pointers are
 * caller-owned CPU memory, "_dev" pointers are device memory on the renderer's
 * GPU.  All calls are synchronous at the boundary unless noted, one renderer
 * handle is not thread-safe (same as the PyO3 `&mut self` borrow).
 *
 * Errors: every call returns rmesh_status; rmesh_last_error() gives the
 * message of the last failure on the calling thread.  The Python shim maps
 * RMESH_ERR_INVALID -> ValueError, RMESH_ERR_DEVICE -> RuntimeError,
 * RMESH_ERR_UNSUPPORTED -> NotImplementedError (lib.rs:311-329,894-904,2381-2389).
 */
#ifndef RMESH_HIP_H
#define RMESH_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rmesh_renderer rmesh_renderer;
typedef int32_t rmesh_status;

#define RMESH_OK 0
#define RMESH_ERR_INVALID 1
#define RMESH_ERR_DEVICE 2
#define RMESH_ERR_UNSUPPORTED 3
#define RMESH_ERR_CAPACITY 4

typedef struct rmesh_scene_desc {
    const float* vertices;     /* [n_vertices*3] */
    uint64_t n_vertices;
    const uint32_t* indices;   /* [n_tets*4] */
    uint64_t n_tets;
    const float* base_colors;  /* [n_tets*3] colours used as given (reference training path) */
    const float* densities;    /* [n_tets] */
    const float* color_grads;  /* [n_tets*3] */
    const float* circumdata;   /* [n_tets*4] (cx, cy, cz, r^2) */
    uint32_t width, height;
    uint32_t aux_dim;          /* custom aux channels: only 0 is supported */
    uint32_t tile_size;        /* 0 -> 12 (lib.rs:442); 8, 12 and 16 are supported */
    int32_t device_id;         /* -1 -> current device */
    const float* sh_coeffs;    /* nullable [n_tets*3*(deg+1)^2], channel-major per tet */
    uint32_t sh_degree;        /* 0..3, used when sh_coeffs != NULL */
    uint32_t max_views;        /* views per batched call (0 -> 1) */
    uint32_t sh_half;          /* non-zero: colours from the SH rounded to f16, as the reference
                                  viewer's packed SH (project_compute.wgsl:267-318); forward only */
} rmesh_scene_desc;

/* RMeshRenderer.__new__ (lib.rs:265-870).  Uploads the scene and allocates
 * per-handle device state. */
rmesh_status rmesh_create(const rmesh_scene_desc* desc, rmesh_renderer** out);
/* Drop of RMeshRenderer. */
void rmesh_destroy(rmesh_renderer* h);
/* Message of the last failing call on this thread ("" if none). */
const char* rmesh_last_error(void);

/* RMeshRenderer.set_min_t (lib.rs:873-875). */
rmesh_status rmesh_set_min_t(rmesh_renderer* h, float min_t);
/* Run on an external HIP stream (e.g. torch's current stream); NULL restores
 * the handle's own stream. */
rmesh_status rmesh_set_stream(rmesh_renderer* h, void* hip_stream);

/* RMeshRenderer.forward_tiled (lib.rs:1040-1204).  cam[3], vp[16]
 * (column-major), c2w_intrinsics[16]; out_rgba [H*W*4] host, out_aux
 * [H*W*8] host or NULL (render_aux=True). */
rmesh_status rmesh_forward_tiled(rmesh_renderer* h, const float* cam, const float* vp, const float* c2w_intrinsics,
                                 float* out_rgba, float* out_aux);
/* RMeshRenderer.forward (lib.rs:932-1025): the hardware-raster path is served
 * by the tiled forward (same outputs, FP32 instead of an Rgba16Float target). */
rmesh_status rmesh_forward(rmesh_renderer* h, const float* cam, const float* vp, const float* c2w_intrinsics,
                           float* out_rgba);
/* RMeshRenderer.read_instance_count (lib.rs:1879-1884): tets that passed
 * culling in the last forward (first view). */
rmesh_status rmesh_read_instance_count(rmesh_renderer* h, uint32_t* out);
/* RMeshRenderer.forward_error (lib.rs:1621-1711, error_tiled_compute.wgsl):
 * per-tet error statistics over the tile lists and camera of the last
 * forward_tiled.  pixel_err, ssim_err [H*W] host; outputs host: tet_err
 * [M*16] (weight, weighted L1, L1^2, -, -, SSIM, entry xyz, exit xyz, weight,
 * centroid xyz; all weighted by T_val = exp(log T) * od) and tet_count [M*2]
 * (covered pixels, max of i32(65535 * T_val)). */
rmesh_status rmesh_forward_error(rmesh_renderer* h, const float* pixel_err, const float* ssim_err, float* tet_err,
                                 int32_t* tet_count);
/* RMeshRenderer.forward_interval_tiled (lib.rs:1218-1390,
 * interval_generate.wgsl, interval_tiled_rasterize.wgsl): silhouette fan
 * triangles per tet, NDC z interpolated in screen space and linearized with
 * near 0.01 / far 1000, composited over the regular tile lists.  Outputs
 * host: rgba [H*W*4], xyzd [H*W*4] (x, y, z = 0: no vertex normals; w =
 * depth), distortion [H*W*5].  Leaves no forward state for backward or
 * forward_error. */
rmesh_status rmesh_forward_interval_tiled(rmesh_renderer* h, const float* cam, const float* vp,
                                          const float* c2w_intrinsics, float* out_rgba, float* out_xyzd,
                                          float* out_distortion);
/* RMeshRenderer.backward_interval_tiled (lib.rs:2049-2240,
 * backward_interval_tiled.wgsl, interval_chain_back.wgsl): adjoint of the
 * last forward_interval_tiled.  dl_d_image [H*W*4] host; dl_d_xyzd [H*W*4]
 * and dl_d_distortion [H*W*5] host or NULL (zero).  Outputs host:
 * d_vertices [N*3], d_densities [M], d_base_colors [M*3], d_color_grads
 * [M*3], d_vertex_normals [N*3].  As the reference, gradients flow through
 * each silhouette slot's NDC depths and colour offsets with the screen
 * geometry fixed, and the depth and distortion terms follow its adjoint. */
rmesh_status rmesh_backward_interval_tiled(rmesh_renderer* h, const float* dl_d_image, const float* dl_d_xyzd,
                                           const float* dl_d_distortion, float* d_vertices, float* d_densities,
                                           float* d_base_colors, float* d_color_grads, float* d_vertex_normals);

/* RMeshRenderer.backward (lib.rs:1894-2031).  Uses the sorted tile lists and
 * image of the last forward_tiled.  dl_d_image [H*W*4] host; outputs host:
 * d_vertices [N*3], d_densities [M], d_base_colors [M*3], d_color_grads [M*3]. */
rmesh_status rmesh_backward(rmesh_renderer* h, const float* dl_d_image, float* d_vertices, float* d_densities,
                            float* d_base_colors, float* d_color_grads);
/* RMeshRenderer.backward_hw (lib.rs:2255-2378): adjoint of the last forward().
 * forward() is served by the tiled forward here, so this is rmesh_backward
 * over that forward's state (the reference reads its hardware-raster image). */
rmesh_status rmesh_backward_hw(rmesh_renderer* h, const float* dl_d_image, float* d_vertices, float* d_densities,
                               float* d_base_colors, float* d_color_grads);
/* SH extension: d_sh [M*3*nc] of the last rmesh_backward (host). */
rmesh_status rmesh_get_sh_grad(rmesh_renderer* h, float* d_sh);
/* Device-pointer variants of update_params / forward_tiled / backward for
 * callers whose arrays live on the GPU (torch tensors: RMeshForward), same
 * layouts: copies are device to device on the renderer's stream (set it to
 * the caller's with rmesh_set_stream).  update_params_device and
 * backward_device only enqueue work; forward_tiled_device waits for its pair
 * count check. */
rmesh_status rmesh_update_params_device(rmesh_renderer* h, const float* vertices, const float* base_colors,
                                        const float* densities, const float* color_grads);
rmesh_status rmesh_forward_tiled_device(rmesh_renderer* h, const float* cam, const float* vp,
                                        const float* c2w_intrinsics, float* d_rgba);
rmesh_status rmesh_backward_device(rmesh_renderer* h, const float* d_dl_d_image, float* d_vertices,
                                   float* d_densities, float* d_base_colors, float* d_color_grads);

/* RMeshRenderer.train_step (lib.rs:2405-2636): forward + loss + backward +
 * Adam (vertices, densities, color_grads; step += 1).  gt [H*W*3] host,
 * loss_type 0 = L1, 1 = L2.  *out_loss receives the loss (mean over pixels of
 * the per-pixel sum over RGB, matching the gradient's 1/(W*H) scaling); the
 * reference never writes its loss buffer and returns 0.0. */
rmesh_status rmesh_train_step(rmesh_renderer* h, const float* cam, const float* vp, const float* c2w_intrinsics,
                              const float* gt, float lr_verts, float lr_dens, float lr_grads, uint32_t loss_type,
                              float* out_loss);
/* Extension learning rates for train steps: SH coefficients and base_colors
 * (0 disables; the reference never optimises base_colors). */
rmesh_status rmesh_set_extra_lrs(rmesh_renderer* h, float lr_sh, float lr_base);

/* Adjoint mode of backward and the train steps (no counterpart in the
 * reference, whose adjoint is the undo).  RMESH_ADJOINT_UNDO (default): the
 * reference's adjoint (backward_tiled_compute.wgsl:186-204,419-437), which
 * rebuilds log T from the image alpha as log(1 - a) and walks the whole tile
 * list back.  RMESH_ADJOINT_CHECKPOINT: the forward also stores each pixel's
 * log T every 32 tets of the tile list (and its final log T and stop
 * position); the adjoint restarts the undo from those and skips what the
 * forward never composited: the exact adjoint of the forward, at
 * 4 (P / 32 + tiles) TS^2 + 8 H W V bytes of device memory.  Switching modes
 * requires a new forward before backward. */
#define RMESH_ADJOINT_UNDO 0
#define RMESH_ADJOINT_CHECKPOINT 1
rmesh_status rmesh_set_adjoint_mode(rmesh_renderer* h, int32_t mode);
int32_t rmesh_get_adjoint_mode(const rmesh_renderer* h);
/* Device memory held by the renderer (bytes, every buffer it allocated). */
rmesh_status rmesh_device_bytes(rmesh_renderer* h, uint64_t* out);

/* RMeshRenderer.update_params / get_params (lib.rs:2641-2686), host arrays. */
rmesh_status rmesh_update_params(rmesh_renderer* h, const float* vertices, const float* base_colors,
                                 const float* densities, const float* color_grads);
rmesh_status rmesh_get_params(rmesh_renderer* h, float* vertices, float* base_colors, float* densities,
                              float* color_grads);
rmesh_status rmesh_update_sh(rmesh_renderer* h, const float* sh);
rmesh_status rmesh_get_sh(rmesh_renderer* h, float* sh);

/* RMeshRenderer.step / width / height getters (lib.rs:2689-2704). */
uint32_t rmesh_get_step(const rmesh_renderer* h);
uint32_t rmesh_get_width(const rmesh_renderer* h);
uint32_t rmesh_get_height(const rmesh_renderer* h);

/* ---- batched multi-view path (no counterpart in the reference: it runs one
 * view per submit) ---- */

/* Forward of n_views cameras (cams [n*3], vps [n*16], c2ws [n*16] host);
 * images land in the renderer's device image buffer [n][H][W][4]
 * (see rmesh_device_buffer); out_rgba (host, nullable) receives a copy. */
rmesh_status rmesh_forward_batched(rmesh_renderer* h, uint32_t n_views, const float* cams, const float* vps,
                                   const float* c2ws, float* out_rgba);
/* One training step over n_views: forward + loss + backward of every view,
 * gradients summed, then one Adam update of every enabled group (step += 1).
 * gt_dev: device [n][H][W][3]; NULL means the renderer's own target buffer
 * (RMESH_BUF_GT).  lr_sh applies when the renderer carries SH.  out_loss
 * (host, nullable) receives the mean per-view loss.  When graph mode is on,
 * the launch sequence is captured once into a hipGraph and replayed. */
rmesh_status rmesh_train_step_batched(rmesh_renderer* h, uint32_t n_views, const float* cams, const float* vps,
                                      const float* c2ws, const float* gt_dev, uint32_t loss_type, float lr_verts,
                                      float lr_dens, float lr_grads, float lr_sh, float* out_loss);
rmesh_status rmesh_set_graph_mode(rmesh_renderer* h, int enable);

/* ---- tuning (no counterpart in the reference, which fixes the tile size at
 * 12 in its Python path, lib.rs:442, and has one kernel shape) ---- */

/* Tile size (8, 12, 16, 32) and batch-geometry variants of the forward and
 * adjoint compositing kernels (0 .. rmesh_num_variants(0|1) - 1; tile 32
 * has variants 0 and 1 only).  Changing them drops a captured graph and the
 * last forward's state. */
rmesh_status rmesh_set_tuning(rmesh_renderer* h, uint32_t tile_size, int32_t fwd_variant, int32_t bwd_variant);
/* out3 = {tile_size, fwd_variant, bwd_variant}. */
rmesh_status rmesh_get_tuning(const rmesh_renderer* h, int32_t* out3);
/* Number of kernel variants: which = 0 forward, 1 adjoint (tiles 8-16). */
int32_t rmesh_num_variants(int32_t which);
/* Radix-sort digit width: 8 (default, 256 buckets) or 4 bits (16 buckets,
 * twice the passes).  The reference sorts 8-bit digits
 * (crates/rmesh-sort/src/lib.rs:85-90); both give the same stable order. */
rmesh_status rmesh_set_sort_digit_bits(rmesh_renderer* h, uint32_t bits);
uint32_t rmesh_get_sort_digit_bits(const rmesh_renderer* h);
/* Autotuner: times binning + sort at both digit widths, every forward
 * variant and every adjoint variant for each tile size (8, 12, 16, 32) on
 * these views (best of `reps` runs, HIP events) and keeps the fastest
 * combination.  Parameters and optimiser state
 * are untouched; gradient buffers are cleared.  gt_dev as in
 * rmesh_train_step_batched.  out_ms (nullable, 4 floats): best total,
 * binning + sort, forward, adjoint. */
rmesh_status rmesh_autotune(rmesh_renderer* h, uint32_t n_views, const float* cams, const float* vps,
                            const float* c2ws, const float* gt_dev, uint32_t loss_type, uint32_t reps, float* out_ms);

#define RMESH_BUF_RGBA 0        /* float [max_views][H][W][4] */
#define RMESH_BUF_GT 1          /* float [max_views][H][W][3] */
#define RMESH_BUF_VERTICES 2    /* float [N*3] */
#define RMESH_BUF_DENSITIES 3   /* float [M] */
#define RMESH_BUF_COLOR_GRADS 4 /* float [M*3] */
#define RMESH_BUF_BASE_COLORS 5 /* float [M*3] */
#define RMESH_BUF_SH 6          /* float [M*3*nc] */
#define RMESH_BUF_D_VERTICES 7
#define RMESH_BUF_D_DENSITIES 8
#define RMESH_BUF_D_COLOR_GRADS 9
#define RMESH_BUF_D_SH 10          /* float [SH floats of all scenes] (written without an SH learning rate) */
#define RMESH_BUF_D_BASE_COLORS 11 /* float [views][M*3] per-view colour gradients (SH), or [M*3] */
/* Device pointer of an internal buffer (zero-copy for torch callers). */
void* rmesh_device_buffer(rmesh_renderer* h, int which);
/* Upload n_views target images (host [n][H][W][3]) into RMESH_BUF_GT. */
rmesh_status rmesh_set_targets(rmesh_renderer* h, uint32_t n_views, const float* gt);
/* Per-stage timing of training steps: when enabled, HIP events are recorded
 * between the stages of every train step (also inside a captured graph);
 * rmesh_stage_times returns the 7 stage durations of the last step in ms:
 * colours+projection+scan, pair generation, sort, ranges, forward, adjoint,
 * SH adjoint + Adam + loss. */
rmesh_status rmesh_set_timing(rmesh_renderer* h, int enable);
rmesh_status rmesh_stage_times(rmesh_renderer* h, float* ms7);
/* Number of tile-tet pairs of the last step (all views). */
rmesh_status rmesh_pair_count(rmesh_renderer* h, uint64_t* out);
/* Copy nbytes of an internal buffer (RMESH_BUF_*) to host memory. */
rmesh_status rmesh_read_buffer(rmesh_renderer* h, int which, void* host, uint64_t nbytes);
/* Streaming copy of nbytes (read nbytes + write nbytes per copy), iters
 * times; reports the average ms per copy.  Used to calibrate HBM byte
 * counters against a kernel of known traffic. */
rmesh_status rmesh_copy_probe(int32_t device_id, uint64_t nbytes, uint32_t iters, float* ms_per_copy);
/* n_ops float atomic adds (no return), every lane on its own 64-byte line,
 * a wave's lanes scattered over `lines` lines (977 .. 2^24); *ms = the
 * launch time.  The atomic-rate ceiling the adjoint's flush is priced
 * against. */
rmesh_status rmesh_atomic_probe(int32_t device_id, uint32_t lines, uint64_t n_ops, float* ms);
/* Synchronise the renderer's stream. */
rmesh_status rmesh_synchronize(rmesh_renderer* h);

/* ---- inspection of the binning state, for parity tests (the reference's
 * kernel tests read the same buffers back: tiles_touched, indirect_args,
 * tile_sort_keys/values, tile_ranges — kernel_tests.rs:96-167,1618-1671) ---- */

/* Per-tet tile counts of view `view` of the last forward [M]. */
rmesh_status rmesh_debug_tiles_touched(rmesh_renderer* h, uint32_t view, uint32_t* out);
/* Pre-sort pairs of one camera: runs projection + pair generation only. */
rmesh_status rmesh_debug_pairs(rmesh_renderer* h, const float* cam, const float* vp, const float* c2w_intrinsics,
                               uint64_t* keys, uint32_t* vals, uint64_t cap, uint64_t* n_pairs);
/* Sorted pairs and per-tile ranges of the last forward (all views). */
rmesh_status rmesh_debug_sorted(rmesh_renderer* h, uint64_t* keys, uint32_t* vals, uint64_t cap, uint64_t* n_pairs,
                                uint32_t* ranges, uint64_t ranges_cap);
/* Standalone sort of host (key, value) pairs on the GPU with the pipeline's
 * radix sort: stable on the low `sorting_bits` bits (multiple of 8). */
rmesh_status rmesh_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t n, uint32_t sorting_bits, int32_t device_id);
/* rmesh_sort_pairs without any allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under hipGraph
 * capture, serving loops, harnesses that forbid hipMalloc).  keys/vals are
 * DEVICE pointers, sorted in place (stable on the low `sorting_bits` bits,
 * multiple of 8) by the same kernels; the call only enqueues work on
 * `hip_stream` (NULL: the null stream) of device `device_id` (-1: current)
 * and returns.  Scratch is a fixed per-device array of the library (14 MB,
 * held until the process ends): n above 2^20 pairs returns
 * RMESH_ERR_INVALID.  Calls that overlap on one GPU share that array and
 * race; serialise them (one stream, or events). */
rmesh_status rmesh_sort_pairs_noalloc(uint64_t* keys, uint32_t* vals, uint64_t n, uint32_t sorting_bits,
                                      int32_t device_id, void* hip_stream);
/* Bits sorted for a given tile count (crates/rmesh-sort/src/drs.rs:23-26). */
uint32_t rmesh_sorting_bits(uint32_t num_tiles);

/* ---- multi-scene batches (no counterpart in the reference, which holds one
 * scene per RMeshRenderer, crates/rmesh-python/src/lib.rs:265-277) ----
 * One renderer over n scenes of uneven tet count, resolution (descs[i].width
 * / height) and SH degree (0..3; every scene with SH coefficients, or none),
 * each with up to descs[i].max_views views per step; every scene shares the
 * tile size, device and sh_half of descs[0].  A step runs one launch per
 * stage over all scenes' views (the SH colour and SH adjoint stages one per
 * SH degree present): binning items per (tet, view) of each scene, one sort,
 * compositing grids over all views' tiles, one Adam update over all
 * parameters.  views_per_scene[i] views of scene i this step (0..max_views),
 * cameras (cams [V][3], vps [V][16], c2ws [V][16]) listed scene by scene in
 * scene order; images and targets are concatenated per view in that order,
 * each view at its scene's resolution.  The single-scene entry points refuse
 * a multi-scene renderer (RMESH_ERR_UNSUPPORTED). */
rmesh_status rmesh_create_multi(const rmesh_scene_desc* descs, uint32_t n_scenes, rmesh_renderer** out);
uint32_t rmesh_scene_count(const rmesh_renderer* h);
/* out_rgba (nullable): the views' images, concatenated. */
rmesh_status rmesh_forward_multi(rmesh_renderer* h, const uint32_t* views_per_scene, const float* cams,
                                 const float* vps, const float* c2ws, float* out_rgba);
/* Targets for the views of a layout (host, [view][H][W][3] concatenated). */
rmesh_status rmesh_set_targets_multi(rmesh_renderer* h, const uint32_t* views_per_scene, const float* gt);
/* One forward + loss + adjoint + SH adjoint + Adam step over the layout,
 * graph-captured in graph mode (recaptured when the layout changes).
 * out_view_losses (nullable): per-view loss; out_loss: their mean. */
rmesh_status rmesh_train_step_multi(rmesh_renderer* h, const uint32_t* views_per_scene, const float* cams,
                                    const float* vps, const float* c2ws, const float* gt_dev, uint32_t loss_type,
                                    float lr_verts, float lr_dens, float lr_grads, float lr_sh,
                                    float* out_view_losses, float* out_loss);
/* rmesh_autotune over a multi-scene layout (one tile size and variant for all). */
rmesh_status rmesh_autotune_multi(rmesh_renderer* h, const uint32_t* views_per_scene, const float* cams,
                                  const float* vps, const float* c2ws, const float* gt_dev, uint32_t loss_type,
                                  uint32_t reps, float* out_ms);
/* Scene `scene`'s slice of a parameter or gradient buffer (RMESH_BUF_VERTICES,
 * _DENSITIES, _COLOR_GRADS, _BASE_COLORS, _SH and their gradients). */
rmesh_status rmesh_read_scene(rmesh_renderer* h, uint32_t scene, int which, void* host, uint64_t nbytes);
/* Per-scene timing of the steps that follow: the wall clock at each step's
 * start and when each scene's last adjoint tile retired. */
rmesh_status rmesh_set_scene_timing(rmesh_renderer* h, int enable);
/* ms[scene]: from the last step's start to the scene's last adjoint tile. */
rmesh_status rmesh_scene_latency(rmesh_renderer* h, float* ms);
/* Order of the compositing grids' tiles across scenes.  INTERLEAVED (the
 * default): all scenes' tiles together, heaviest first (best balance; every
 * scene finishes near the end of the stage).  BY_SCENE: the scenes one after
 * another, fewest pairs first, each heaviest first (early scenes finish
 * early; at most 64 scenes).  Results are the same either way. */
#define RMESH_SCENE_ORDER_INTERLEAVED 0u
#define RMESH_SCENE_ORDER_BY_SCENE 1u
rmesh_status rmesh_set_scene_order(rmesh_renderer* h, uint32_t order);
uint32_t rmesh_get_scene_order(const rmesh_renderer* h);

/* ---- scene files (crates/rmesh-data/src/lib.rs) ----
 * load_rmesh / load_rmesh_raw / load_ply (lib.rs:140-148, 243-394, 711-853):
 * parse a file image held in host memory into the renderer's flat arrays.
 * .rmesh: header (N, M, SH degree, PCA components), start pose, vertices,
 * indices, u8 log densities (exp((v - 100) / 20)), f16 PCA mean / basis /
 * weights of the SH coefficients (decompressed: mean + sum_k w_k basis_k),
 * f16 colour gradients, optional tagged sections (counted, not decoded).
 * PLY: binary little-endian vertices and tetrahedra with interleaved SH.
 * The SH leave channel-major per tet ([M][3][(deg+1)^2], the layout of
 * rmesh_scene_desc.sh_coeffs); circumdata is computed in float64 with the
 * reference's degenerate-tet rules.  Host code, no device needed. */
typedef struct rmesh_scene_file rmesh_scene_file;
#define RMESH_FILE_RMESH 0      /* .rmesh, gzip-compressed or raw (detected) */
#define RMESH_FILE_PLY 1        /* binary little-endian PLY */
#define RMESH_FILE_RMESH_RAW 2  /* .rmesh, uncompressed only (load_rmesh_raw) */
typedef struct rmesh_scene_info {
    uint64_t n_vertices, n_tets;
    uint32_t sh_degree;     /* SH coefficients per tet: 3 * (deg+1)^2 */
    uint32_t k_components;  /* PCA components as stored (.rmesh), 0 for PLY */
    float start_pose[7];    /* x, y, z, qx, qy, qz, qw (.rmesh), zeros for PLY */
    uint32_t n_tagged;      /* tagged extension sections (.rmesh) */
} rmesh_scene_info;
rmesh_status rmesh_scene_parse(const uint8_t* data, uint64_t nbytes, int32_t format, rmesh_scene_file** out,
                               rmesh_scene_info* info);
/* Copies the parsed arrays out (any pointer may be NULL to skip it):
 * vertices [N*3], indices [M*4], densities [M], color_grads [M*3],
 * circumdata [M*4], sh_coeffs [M*3*(deg+1)^2]. */
rmesh_status rmesh_scene_copy(const rmesh_scene_file* f, float* vertices, uint32_t* indices, float* densities,
                              float* color_grads, float* circumdata, float* sh_coeffs);
void rmesh_scene_free(rmesh_scene_file* f);

#ifdef __cplusplus
}
#endif

#endif /* RMESH_HIP_H */
