/*
 * CPU restatement of rmeshvk's tiled train-step kernels — TEST INFRASTRUCTURE ONLY.
 *
 * This header is instantiated twice by rmesh_oracle.c: once with R = float
 * (the reference's own FP32 arithmetic, used as the bit-exact checker for
 * binning/sort/ranges and the tolerance checker for images and gradients) and
 * once with R = double (used for finite-difference gradient checks).
 *
 * What pins each function is listed in oracle.py.  Parity unpinned: the
 * depth and distortion terms of orc_interval_backward.
 *
 * Every function names the reference kernel and lines it follows.  Float
 * semantics chosen where WGSL leaves them to the driver:
 *   dot(a,b)      = (a.x*b.x + a.y*b.y) + a.z*b.z, no FMA contraction
 *   cross(a,b)    = (a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x)
 *   normalize(v)  = v / sqrt(dot(v,v))   (IEEE division and square root)
 *   mat*vec       = column-major, summed column 0 → 3
 *   min/max       = IEEE minNum/maxNum (fmin/fmax)
 *   f32 → u32/i32 = truncation, saturating at the type range, NaN → 0
 */

#ifndef R
#error "define R (float/double) and SFX before including oracle_impl.h"
#endif

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)

/* ------------------------------------------------------------------ */
/* small vector helpers                                                 */
/* ------------------------------------------------------------------ */
typedef struct { R x, y, z; } FN(v3_);
#define V3 FN(v3_)

static inline V3 FN(mk)(R x, R y, R z) { V3 r = {x, y, z}; return r; }
static inline V3 FN(sub)(V3 a, V3 b) { return FN(mk)(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 FN(add)(V3 a, V3 b) { return FN(mk)(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 FN(scl)(V3 a, R s) { return FN(mk)(a.x * s, a.y * s, a.z * s); }
static inline V3 FN(neg)(V3 a) { return FN(mk)(-a.x, -a.y, -a.z); }
static inline R FN(dot)(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline V3 FN(cross)(V3 a, V3 b) {
    return FN(mk)(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline R FN(rmin)(R a, R b) { return RFMIN(a, b); }
static inline R FN(rmax)(R a, R b) { return RFMAX(a, b); }
static inline R FN(clampr)(R v, R lo, R hi) { return FN(rmin)(FN(rmax)(v, lo), hi); }
static inline V3 FN(normalize)(V3 v) {
    R l = RSQRT(FN(dot)(v, v));
    return FN(mk)(v.x / l, v.y / l, v.z / l);
}
static inline V3 FN(load3)(const R* p, uint32_t i) { return FN(mk)(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }

/* safe math — crates/rmesh-util/src/wgsl/math.wgsl:3-65 */
#define TINY_VAL_R ((R)1.0754944e-20)
#define MIN_VAL_R ((R)-1e+20)
#define MAX_VAL_R ((R)1e+20)
#define LOG_MAX_VAL_R ((R)46.0517)

static inline R FN(safe_clip)(R v, R lo, R hi) { return FN(rmax)(FN(rmin)(v, hi), lo); }
static inline V3 FN(safe_clip3)(V3 v, R lo, R hi) {
    return FN(mk)(FN(safe_clip)(v.x, lo, hi), FN(safe_clip)(v.y, lo, hi), FN(safe_clip)(v.z, lo, hi));
}
static inline R FN(safe_div)(R a, R b) {
    if (RABS(b) < TINY_VAL_R) return FN(safe_clip)(a / TINY_VAL_R, MIN_VAL_R, MAX_VAL_R);
    return FN(safe_clip)(a / b, MIN_VAL_R, MAX_VAL_R);
}
static inline R FN(safe_exp)(R v) { return REXP(FN(clampr)(v, (R)-88.0, LOG_MAX_VAL_R)); }
/* math.wgsl:35-42 */
static inline R FN(phi)(R x) {
    if (RABS(x) < (R)0.02)
        return (R)1.0 + x * ((R)-0.5 + x * ((R)(1.0 / 6.0) + x * ((R)(-1.0 / 24.0))));
    return FN(safe_div)((R)1.0 - REXP(-x), x);
}
/* math.wgsl:44-51 */
static inline R FN(dphi_dx)(R x) {
    if (RABS(x) < (R)0.02)
        return (R)-0.5 + x * ((R)(1.0 / 3.0) + x * ((R)(-1.0 / 8.0) + x * ((R)(1.0 / 30.0))));
    R ex = REXP(-x);
    return FN(safe_div)(ex * (x + (R)1.0) - (R)1.0, x * x);
}
/* math.wgsl:20-33 */
static inline R FN(softplus)(R x) {
    if (x * (R)10.0 > (R)20.0) return x;
    return (R)0.1 * RLOG((R)1.0 + REXP((R)10.0 * x));
}
static inline R FN(dsoftplus)(R x) {
    if (x * (R)10.0 > (R)20.0) return (R)1.0;
    R e = REXP((R)10.0 * x);
    return FN(safe_div)(e, (R)1.0 + e);
}

/* clip = vp * vec4(p, 1): column-major, columns summed in order (project_compute.wgsl:77-81) */
static inline void FN(mul_vp)(const R vp[16], V3 p, R out[4]) {
    for (int k = 0; k < 4; ++k)
        out[k] = ((vp[0 * 4 + k] * p.x + vp[1 * 4 + k] * p.y) + vp[2 * 4 + k] * p.z) + vp[3 * 4 + k];
}

/* ------------------------------------------------------------------ */
/* camera / view description                                            */
/* ------------------------------------------------------------------ */
typedef struct {
    R vp[16];    /* column-major view-projection */
    R c2w[9];    /* camera-to-world rotation, columns */
    R intr[4];   /* fx fy cx cy */
    R cam[3];
    R min_t;
    uint32_t width, height, tile_size;
} FN(orc_cam_);
#define CAM FN(orc_cam_)

static void FN(cam_from_flat)(CAM* c, const float* cam_pos, const float* vp, const float* c2w_intr,
                              uint32_t w, uint32_t h, uint32_t ts, float min_t) {
    /* c2w_from_flat / mat4_from_flat — crates/rmesh-python/src/lib.rs:117-136 */
    for (int i = 0; i < 16; ++i) c->vp[i] = (R)vp[i];
    for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) c->c2w[j * 3 + k] = (R)c2w_intr[j * 4 + k];
    for (int i = 0; i < 4; ++i) c->intr[i] = (R)c2w_intr[12 + i];
    for (int i = 0; i < 3; ++i) c->cam[i] = (R)cam_pos[i];
    c->min_t = (R)min_t;
    c->width = w; c->height = h; c->tile_size = ts;
}

/* tet face table (a, b, c, opposite) — crates/rmesh-util/src/wgsl/intersect.wgsl:4-9 */
static const int FN(FACES)[4][4] = {{0, 2, 1, 3}, {1, 2, 3, 0}, {0, 3, 2, 1}, {3, 0, 1, 2}};

/* edges scanned by the tile scanline — project_compute.wgsl:225-226 */
static const int FN(EI)[6] = {0, 0, 0, 1, 1, 2};
static const int FN(EJ)[6] = {1, 2, 3, 2, 3, 3};

/* ------------------------------------------------------------------ */
/* K1: per-tet projection, culling and conservative tile count          */
/* crates/rmesh-render/src/wgsl/project_compute.wgsl:88-262             */
/* ------------------------------------------------------------------ */
/* Returns the tile count for one tet; *visible receives the vis_idx flag
 * (tets with visible=1 get a compact slot even when the count is 0). */
static uint32_t FN(project_one)(const R* verts, const uint32_t* idx, const R* dens, uint32_t t,
                                const CAM* c, int* visible) {
    V3 v[4];
    for (int i = 0; i < 4; ++i) v[i] = FN(load3)(verts, idx[4 * t + i]);
    R p[4][4]; /* ndc.xyz, clip.w */
    for (int i = 0; i < 4; ++i) {
        R clip[4];
        FN(mul_vp)(c->vp, v[i], clip);
        R inv_w = (R)1.0 / (clip[3] + (R)1e-6);
        p[i][0] = clip[0] * inv_w; p[i][1] = clip[1] * inv_w; p[i][2] = clip[2] * inv_w; p[i][3] = clip[3];
    }
    int any_behind = (p[0][3] <= 0) || (p[1][3] <= 0) || (p[2][3] <= 0) || (p[3][3] <= 0);
    R W = (R)c->width, H = (R)c->height;
    int vis = 1;
    if (!any_behind) {
        R min_x = FN(rmin)(FN(rmin)(p[0][0], p[1][0]), FN(rmin)(p[2][0], p[3][0]));
        R max_x = FN(rmax)(FN(rmax)(p[0][0], p[1][0]), FN(rmax)(p[2][0], p[3][0]));
        R min_y = FN(rmin)(FN(rmin)(p[0][1], p[1][1]), FN(rmin)(p[2][1], p[3][1]));
        R max_y = FN(rmax)(FN(rmax)(p[0][1], p[1][1]), FN(rmax)(p[2][1], p[3][1]));
        R min_z = FN(rmin)(FN(rmin)(p[0][2], p[1][2]), FN(rmin)(p[2][2], p[3][2]));
        if (max_x < (R)-1.0 || min_x > (R)1.0 || max_y < (R)-1.0 || min_y > (R)1.0 || min_z > (R)1.0) vis = 0;
        R ext_x = (max_x - min_x) * W;
        R ext_y = (max_y - min_y) * H;
        if (ext_x * ext_y < (R)1.0) vis = 0;
    }
    if (!vis) { *visible = 0; return 0; }
    /* optical-depth cull, :149-162 */
    {
        V3 e01 = FN(sub)(v[1], v[0]), e02 = FN(sub)(v[2], v[0]), e03 = FN(sub)(v[3], v[0]);
        V3 e12 = FN(sub)(v[2], v[1]), e13 = FN(sub)(v[3], v[1]), e23 = FN(sub)(v[3], v[2]);
        R m = FN(rmax)(FN(rmax)(FN(dot)(e01, e01), FN(rmax)(FN(dot)(e02, e02), FN(dot)(e03, e03))),
                       FN(rmax)(FN(dot)(e12, e12), FN(rmax)(FN(dot)(e13, e13), FN(dot)(e23, e23))));
        R max_od = dens[t] * RSQRT(m);
        if (max_od < (R)0.01) { *visible = 0; return 0; }
    }
    *visible = 1;
    if (any_behind) return 0;
    R proj[4][2];
    for (int i = 0; i < 4; ++i) {
        proj[i][0] = ((p[i][0] + (R)1.0) * (R)0.5) * W;
        proj[i][1] = (((R)1.0 - p[i][1]) * (R)0.5) * H;
    }
    R pix_min_x = FN(rmax)(FN(rmin)(FN(rmin)(proj[0][0], proj[1][0]), FN(rmin)(proj[2][0], proj[3][0])), (R)0.0);
    R pix_max_x = FN(rmin)(FN(rmax)(FN(rmax)(proj[0][0], proj[1][0]), FN(rmax)(proj[2][0], proj[3][0])), W - (R)1.0);
    R pix_min_y = FN(rmax)(FN(rmin)(FN(rmin)(proj[0][1], proj[1][1]), FN(rmin)(proj[2][1], proj[3][1])), (R)0.0);
    R pix_max_y = FN(rmin)(FN(rmax)(FN(rmax)(proj[0][1], proj[1][1]), FN(rmax)(proj[2][1], proj[3][1])), H - (R)1.0);
    if (pix_min_x > pix_max_x || pix_min_y > pix_max_y) return 0;
    R T = (R)c->tile_size;
    uint32_t tiles_x_total = sat_u32((double)RCEIL(W / T));
    uint32_t tile_min_x = sat_u32((double)(pix_min_x / T));
    uint32_t tile_max_x = umin32(sat_u32((double)(pix_max_x / T)), tiles_x_total - 1u);
    uint32_t tile_min_y = sat_u32((double)(pix_min_y / T));
    uint32_t tile_max_y = umin32(sat_u32((double)(pix_max_y / T)), sat_u32((double)RCEIL(H / T)) - 1u);
    uint32_t count = 0;
    for (uint32_t ty = tile_min_y; ty <= tile_max_y; ++ty) {
        R xl = (R)1e10f, xr = (R)-1e10f;
        for (int e = 0; e < 6; ++e) {
            R yi = proj[FN(EI)[e]][1], yj = proj[FN(EJ)[e]][1];
            R xi = proj[FN(EI)[e]][0], xj = proj[FN(EJ)[e]][0];
            for (uint32_t b = 0; b < 2; ++b) {
                R ytest = (R)(ty + b) * T;
                if ((yi <= ytest && yj > ytest) || (yj <= ytest && yi > ytest)) {
                    R tt = (ytest - yi) / (yj - yi);
                    R x = xi + tt * (xj - xi);
                    xl = FN(rmin)(xl, x); xr = FN(rmax)(xr, x);
                }
            }
        }
        for (int k = 0; k < 4; ++k) {
            R vy = proj[k][1];
            if (vy >= (R)ty * T && vy <= (R)(ty + 1u) * T) {
                xl = FN(rmin)(xl, proj[k][0]); xr = FN(rmax)(xr, proj[k][0]);
            }
        }
        if (xl <= xr) {
            uint32_t tx0 = uclamp32(sat_u32((double)FN(rmax)(RFLOOR(xl / T), (R)0.0)), tile_min_x, tile_max_x);
            uint32_t tx1 = uclamp32(sat_u32((double)FN(rmax)(RFLOOR(xr / T), (R)0.0)), tile_min_x, tile_max_x);
            count += tx1 - tx0 + 1u;
        }
    }
    return count;
}

void FN(orc_project)(const R* verts, const uint32_t* idx, const R* dens, uint32_t M,
                     const float* cam_pos, const float* vp, const float* c2w_intr,
                     uint32_t w, uint32_t h, uint32_t ts,
                     uint32_t* tiles_touched, uint8_t* visible) {
    CAM c;
    FN(cam_from_flat)(&c, cam_pos, vp, c2w_intr, w, h, ts, 0.0f);
#pragma omp parallel for schedule(static) if (M >= 4096)
    for (int64_t ti = 0; ti < (int64_t)M; ++ti) {
        const uint32_t t = (uint32_t)ti;
        int vis = 0;
        tiles_touched[t] = FN(project_one)(verts, idx, dens, t, &c, &vis);
        visible[t] = (uint8_t)vis;
    }
}

/* ------------------------------------------------------------------ */
/* K7: pair generation with sentinel padding                             */
/* crates/rmesh-tile/src/wgsl/tile_gen_scan_compute.wgsl:61-203         */
/* tile_fill_compute.wgsl:9-19 (sentinel key 0xFFFFFFFF:0xFFFFFFFF,      */
/* value 0) applied to every slot the generator leaves unwritten.       */
/* Keys are u64 = (tile_id << 32) | ~bits(depth).                       */
/* Compact order = tet-id order (the reference's vis_idx order comes    */
/* from an atomic and is not deterministic).                            */
/* ------------------------------------------------------------------ */
uint64_t FN(orc_pairs)(const R* verts, const uint32_t* idx, const R* circum, uint32_t M,
                       const float* cam_pos, const float* vp, const float* c2w_intr,
                       uint32_t w, uint32_t h, uint32_t ts, const uint32_t* tiles_touched,
                       uint64_t* keys, uint32_t* vals, uint64_t cap) {
    CAM c;
    FN(cam_from_flat)(&c, cam_pos, vp, c2w_intr, w, h, ts, 0.0f);
    R W = (R)w, H = (R)h, T = (R)ts;
    uint32_t tiles_x = (w + ts - 1) / ts;
    uint64_t total = 0;
    for (uint32_t t = 0; t < M; ++t) total += tiles_touched[t];
    for (uint64_t i = 0; i < total && i < cap; ++i) { keys[i] = ~(uint64_t)0; vals[i] = 0; }
    uint64_t write_start = 0;
    for (uint32_t t = 0; t < M; ++t) {
        uint32_t max_write = tiles_touched[t];
        uint64_t base = write_start;
        write_start += max_write;
        if (max_write == 0) continue;
        V3 v[4];
        for (int i = 0; i < 4; ++i) v[i] = FN(load3)(verts, idx[4 * t + i]);
        R proj[4][2];
        for (int i = 0; i < 4; ++i) {
            R clip[4];
            FN(mul_vp)(c.vp, v[i], clip);
            R nx = clip[0] / (clip[3] + (R)1e-6);
            R ny = clip[1] / (clip[3] + (R)1e-6);
            proj[i][0] = ((nx + (R)1.0) * (R)0.5) * W;
            proj[i][1] = (((R)1.0 - ny) * (R)0.5) * H;
        }
        R pix_min_x = FN(rmax)(FN(rmin)(FN(rmin)(proj[0][0], proj[1][0]), FN(rmin)(proj[2][0], proj[3][0])), (R)0.0);
        R pix_max_x = FN(rmin)(FN(rmax)(FN(rmax)(proj[0][0], proj[1][0]), FN(rmax)(proj[2][0], proj[3][0])), W - (R)1.0);
        R pix_min_y = FN(rmax)(FN(rmin)(FN(rmin)(proj[0][1], proj[1][1]), FN(rmin)(proj[2][1], proj[3][1])), (R)0.0);
        R pix_max_y = FN(rmin)(FN(rmax)(FN(rmax)(proj[0][1], proj[1][1]), FN(rmax)(proj[2][1], proj[3][1])), H - (R)1.0);
        if (pix_min_x > pix_max_x || pix_min_y > pix_max_y) continue;
        uint32_t tiles_x_total = sat_u32((double)RCEIL(W / T));
        uint32_t tile_min_x = sat_u32((double)(pix_min_x / T));
        uint32_t tile_max_x = umin32(sat_u32((double)(pix_max_x / T)), tiles_x_total - 1u);
        uint32_t tile_min_y = sat_u32((double)(pix_min_y / T));
        uint32_t tile_max_y = umin32(sat_u32((double)(pix_max_y / T)), sat_u32((double)RCEIL(H / T)) - 1u);
        V3 cc = FN(mk)(circum[4 * t], circum[4 * t + 1], circum[4 * t + 2]);
        R r2 = circum[4 * t + 3];
        V3 diff = FN(sub)(cc, FN(mk)(c.cam[0], c.cam[1], c.cam[2]));
        R depth = FN(dot)(diff, diff) - r2;
        uint32_t inv_depth = ~f32_bits((float)depth);
        uint32_t local = 0;
        for (uint32_t ty = tile_min_y; ty <= tile_max_y; ++ty) {
            R xl = (R)1e10f, xr = (R)-1e10f;
            for (int e = 0; e < 6; ++e) {
                R yi = proj[FN(EI)[e]][1], yj = proj[FN(EJ)[e]][1];
                R xi = proj[FN(EI)[e]][0], xj = proj[FN(EJ)[e]][0];
                for (uint32_t b = 0; b < 2; ++b) {
                    R ytest = (R)(ty + b) * T;
                    if ((yi <= ytest && yj > ytest) || (yj <= ytest && yi > ytest)) {
                        R tt = (ytest - yi) / (yj - yi);
                        R x = xi + tt * (xj - xi);
                        xl = FN(rmin)(xl, x); xr = FN(rmax)(xr, x);
                    }
                }
            }
            for (int k = 0; k < 4; ++k) {
                R vy = proj[k][1];
                if (vy >= (R)ty * T && vy <= (R)(ty + 1u) * T) {
                    xl = FN(rmin)(xl, proj[k][0]); xr = FN(rmax)(xr, proj[k][0]);
                }
            }
            if (xl <= xr) {
                uint32_t tx0 = uclamp32(sat_u32((double)FN(rmax)(RFLOOR(xl / T), (R)0.0)), tile_min_x, tile_max_x);
                uint32_t tx1 = uclamp32(sat_u32((double)FN(rmax)(RFLOOR(xr / T), (R)0.0)), tile_min_x, tile_max_x);
                for (uint32_t tx = tx0; tx <= tx1; ++tx) {
                    uint32_t tile_id = ty * tiles_x + tx;
                    if (local < max_write) {
                        uint64_t wi = base + local;
                        if (wi < cap) {
                            keys[wi] = ((uint64_t)tile_id << 32) | inv_depth;
                            vals[wi] = t;
                        }
                        local++;
                    }
                }
            }
        }
    }
    return total;
}

/* ------------------------------------------------------------------ */
/* K12: tiled front-to-back compositing                                  */
/* crates/rmesh-render/src/wgsl/rasterize_compute.wgsl:92-511           */
/* ------------------------------------------------------------------ */
/* Pixel-row span of a projected tet in tile-local pixel coordinates.
 * forward=1 adds the vertex-in-row rule (rasterize_compute.wgsl:236-242);
 * the backward kernel omits it (backward_tiled_compute.wgsl:245-283). */
static void FN(row_span)(const R proj[4][2], uint32_t row, uint32_t TS, int forward, int32_t* xl_i, int32_t* xr_i) {
    R xl_f = (R)1e10f, xr_f = (R)-1e10f;
    R yc = (R)row + (R)0.5;
    for (int e = 0; e < 6; ++e) {
        int a = FN(EI)[e], b = FN(EJ)[e];
        R yi = proj[a][1], yj = proj[b][1];
        if ((yi <= yc && yj > yc) || (yj <= yc && yi > yc)) {
            R tt = (yc - yi) / (yj - yi);
            R x = proj[a][0] + tt * (proj[b][0] - proj[a][0]);
            xl_f = FN(rmin)(xl_f, x); xr_f = FN(rmax)(xr_f, x);
        }
    }
    if (forward) {
        for (int k = 0; k < 4; ++k) {
            R vy = proj[k][1];
            if (vy >= yc - (R)0.5 && vy < yc + (R)0.5) {
                xl_f = FN(rmin)(xl_f, proj[k][0]); xr_f = FN(rmax)(xr_f, proj[k][0]);
            }
        }
    }
    *xl_i = 0; *xr_i = -1;
    if (xl_f <= xr_f) {
        R eps = (R)0.001f;
        int32_t l = imax32(sat_i32((double)RCEIL(xl_f - (R)0.5 - eps)), 0);
        int32_t r = imin32(sat_i32((double)RFLOOR(xr_f - (R)0.5 + eps)), (int32_t)TS - 1);
        if (l <= r) { *xl_i = l; *xr_i = r; }
    }
}

/* Tile-local projection used by both tiled kernels (rasterize_compute.wgsl:153-167). */
static int FN(tile_proj)(const CAM* c, const V3 v[4], R tile_ox, R tile_oy, R proj[4][2]) {
    R W = (R)c->width, H = (R)c->height;
    R clip[4][4];
    for (int i = 0; i < 4; ++i) FN(mul_vp)(c->vp, v[i], clip[i]);
    int any_behind = (clip[0][3] <= 0) || (clip[1][3] <= 0) || (clip[2][3] <= 0) || (clip[3][3] <= 0);
    if (!any_behind) {
        for (int i = 0; i < 4; ++i) {
            proj[i][0] = ((clip[i][0] / clip[i][3] + (R)1.0) * (R)0.5) * W - tile_ox;
            proj[i][1] = (((R)1.0 - clip[i][1] / clip[i][3]) * (R)0.5) * H - tile_oy;
        }
    }
    return any_behind;
}

static V3 FN(pixel_ray)(const CAM* c, uint32_t px, uint32_t py) {
    V3 dc = FN(mk)((((R)px + (R)0.5) - c->intr[2]) / c->intr[0], (((R)py + (R)0.5) - c->intr[3]) / c->intr[1], (R)1.0);
    dc = FN(normalize)(dc);
    V3 col0 = FN(mk)(c->c2w[0], c->c2w[1], c->c2w[2]);
    V3 col1 = FN(mk)(c->c2w[3], c->c2w[4], c->c2w[5]);
    V3 col2 = FN(mk)(c->c2w[6], c->c2w[7], c->c2w[8]);
    V3 d = FN(mk)((col0.x * dc.x + col1.x * dc.y) + col2.x * dc.z,
                  (col0.y * dc.x + col1.y * dc.y) + col2.y * dc.z,
                  (col0.z * dc.x + col1.z * dc.y) + col2.z * dc.z);
    return FN(normalize)(d);
}

/* Per-pixel transmittance checkpoints (the checkpointed adjoint; the
 * reference has none): the forward stores each pixel's log T before every
 * tile-list position 32 j (j >= 1, 32 j < the tile's pair count) at slot
 * (range start >> 5) + tile + j - 1 of ck[slot][TS*TS] (pixel row-major in
 * the tile), and per image pixel the final log T and the stop position
 * (1 + the last tet position it composited, 0 if none).  Slots of distinct
 * tiles never overlap: (P >> 5) + tiles + 1 slots hold every tile's. */
#define CK_GROUP 32u
static inline uint64_t FN(ck_slot)(uint32_t rs, uint32_t tile, uint32_t j) {
    return (uint64_t)(rs >> 5) + tile + j - 1u;
}

/* One tile of the forward (rasterize_compute.wgsl:92-511); st: TS*TS*12,
 * ss: TS*TS*4, stp: TS*TS scratch of the calling thread. */
static void FN(fwd_tile)(const R* verts, const uint32_t* idx, const R* colors, const R* dens, const R* grads,
                         const uint32_t* sorted_vals, const uint32_t* ranges, const CAM* cp, uint32_t w, uint32_t h,
                         uint32_t TS, uint32_t tile, R* rgba, R* aux, uint32_t* stats, R* st, uint32_t* ss,
                         uint32_t* stp, R* ck, R* fin_logt, uint32_t* fin_stop) {
    const CAM c = *cp;
    uint32_t tiles_x = (w + TS - 1) / TS;
    V3 cam = FN(mk)(c.cam[0], c.cam[1], c.cam[2]);
    uint32_t tile_x = tile % tiles_x, tile_y = tile / tiles_x;
    R tile_ox = (R)(tile_x * TS), tile_oy = (R)(tile_y * TS);
    memset(st, 0, sizeof(R) * TS * TS * 12);
    memset(ss, 0, sizeof(uint32_t) * TS * TS * 4);
    memset(stp, 0, sizeof(uint32_t) * TS * TS);
    uint32_t rs = ranges[2 * tile], re = ranges[2 * tile + 1];
    uint32_t cursor = re > rs ? re : rs;
    while (cursor > rs) {
        cursor -= 1;
        const uint32_t q = re - 1 - cursor; /* position in compositing order */
        if (ck && q > 0 && q % CK_GROUP == 0) {
            R* dst = ck + FN(ck_slot)(rs, tile, q / CK_GROUP) * TS * TS;
            for (uint32_t i = 0; i < TS * TS; ++i) dst[i] = st[12 * i + 3];
        }
        uint32_t t = sorted_vals[cursor];
        V3 v[4];
        for (int i = 0; i < 4; ++i) v[i] = FN(load3)(verts, idx[4 * t + i]);
        R proj[4][2];
        int any_behind = FN(tile_proj)(&c, v, tile_ox, tile_oy, proj);
        R density = dens[t];
        V3 col = FN(load3)(colors, t), g = FN(load3)(grads, t);
        V3 fn[4];
        for (int fi = 0; fi < 4; ++fi) {
            const int* f = FN(FACES)[fi];
            V3 n = FN(cross)(FN(sub)(v[f[2]], v[f[0]]), FN(sub)(v[f[1]], v[f[0]]));
            if (FN(dot)(n, FN(sub)(v[f[3]], v[f[0]])) < 0) n = FN(neg)(n);
            fn[fi] = n;
        }
        for (uint32_t row = 0; row < TS; ++row) {
            int32_t xl_i = 0, xr_i = -1;
            if (!any_behind) FN(row_span)(proj, row, TS, 1, &xl_i, &xr_i);
            for (int32_t col_i = xl_i; col_i <= xr_i; ++col_i) {
                uint32_t colu = (uint32_t)col_i;
                uint32_t pl = row * TS + colu;
                uint32_t px = tile_x * TS + colu, py = tile_y * TS + row;
                if (px >= w || py >= h) continue;
                R* s = st + 12 * pl;
                if (s[3] < (R)-13.8) { ss[4 * pl + 2]++; continue; }
                V3 rd = FN(pixel_ray)(&c, px, py);
                V3 ro = FN(add)(cam, FN(scl)(rd, c.min_t));
                R t_min = (R)-3.402823e38f, t_max = (R)3.402823e38f;
                int entry = 4, valid = 1;
                for (int fi = 0; fi < 4; ++fi) {
                    const int* f = FN(FACES)[fi];
                    V3 n = fn[fi];
                    R num = FN(dot)(n, FN(sub)(v[f[0]], ro));
                    R den = FN(dot)(n, rd);
                    if (RABS(den) < (R)1e-20f) { if (num > 0) valid = 0; continue; }
                    R tt = num / den;
                    if (den > 0) { if (tt > t_min) { t_min = tt; entry = fi; } }
                    else { if (tt < t_max) t_max = tt; }
                }
                if (!(valid && t_min < t_max)) { ss[4 * pl + 0]++; continue; }
                R base_offset = FN(dot)(g, FN(sub)(ro, v[0]));
                V3 bc = FN(mk)(col.x + base_offset, col.y + base_offset, col.z + base_offset);
                R dc_dt = FN(dot)(g, rd);
                R a0 = dc_dt * t_min, a1 = dc_dt * t_max;
                V3 cs = FN(mk)(FN(rmax)(bc.x + a0, 0), FN(rmax)(bc.y + a0, 0), FN(rmax)(bc.z + a0, 0));
                V3 ce = FN(mk)(FN(rmax)(bc.x + a1, 0), FN(rmax)(bc.y + a1, 0), FN(rmax)(bc.z + a1, 0));
                R dist = t_max - t_min;
                R od = FN(clampr)(density * dist, (R)1e-8f, (R)88.0);
                R alpha_t = FN(safe_exp)(-od);
                R phi_v = FN(phi)(od);
                R w0 = phi_v - alpha_t, w1 = (R)1.0 - phi_v;
                V3 cpm = FN(safe_clip3)(FN(mk)(ce.x * w0 + cs.x * w1, ce.y * w0 + cs.y * w1, ce.z * w0 + cs.z * w1),
                                        (R)0.0, MAX_VAL_R);
                R alpha = (R)1.0 - alpha_t;
                R Tj = FN(safe_exp)(s[3]);
                if (od < (R)0.01) ss[4 * pl + 1]++;
                else if (Tj < (R)1e-6f) ss[4 * pl + 2]++;
                else ss[4 * pl + 3]++;
                if (Tj >= (R)1e-6f) {
                    s[0] = s[0] + cpm.x * Tj; s[1] = s[1] + cpm.y * Tj; s[2] = s[2] + cpm.z * Tj;
                    s[3] = s[3] - od;
                    stp[pl] = q + 1;
                    /* aux: normals + depth, entropy (rasterize_compute.wgsl:434-456) */
                    V3 en = FN(mk)(0, 0, 0);
                    if (entry < 4) en = FN(normalize)(FN(neg)(fn[entry]));
                    R depth_premul = w0 * t_max + w1 * t_min;
                    R ta = Tj * alpha;
                    s[4] = s[4] + ta * en.x; s[5] = s[5] + ta * en.y; s[6] = s[6] + ta * en.z;
                    s[7] = s[7] + Tj * depth_premul;
                    R wt = alpha * Tj;
                    R c_mid = (t_min + t_max) * (R)0.5;
                    R wc = wt * c_mid;
                    s[8] = s[8] + wt;
                    s[9] = s[9] + (wt > (R)1e-20f ? wt * RLOG(wt) : (R)0.0);
                    s[10] = s[10] + wt * c_mid;
                    s[11] = s[11] + (wc > (R)1e-20f ? wc * RLOG(wc) : (R)0.0);
                }
            }
        }
    }
    for (uint32_t i = 0; i < TS * TS; ++i) {
        uint32_t row = i / TS, colu = i % TS;
        uint32_t px = tile_x * TS + colu, py = tile_y * TS + row;
        if (px < w && py < h) {
            size_t pix = (size_t)py * w + px;
            R* s = st + 12 * i;
            rgba[4 * pix + 0] = s[0]; rgba[4 * pix + 1] = s[1]; rgba[4 * pix + 2] = s[2];
            rgba[4 * pix + 3] = (R)1.0 - FN(safe_exp)(s[3]);
            if (aux) for (int k = 0; k < 8; ++k) aux[8 * pix + k] = s[4 + k];
            if (stats) for (int k = 0; k < 4; ++k) stats[4 * pix + k] = ss[4 * i + k];
            if (fin_logt) fin_logt[pix] = s[3];
            if (fin_stop) fin_stop[pix] = stp[i];
        }
    }
}

/* colors: [M*3] per-tet colour (base_colors or SH-evaluated); aux: [H*W*8] or NULL.
 * Tiles are independent (each writes its own pixels and checkpoint slots), so
 * they run in parallel on the host's cores; results do not depend on the
 * thread count.  ck / fin_logt / fin_stop (NULL: not kept): the checkpoints
 * of the checkpointed adjoint (layout above). */
void FN(orc_forward)(const R* verts, const uint32_t* idx, const R* colors, const R* dens, const R* grads,
                     const uint32_t* sorted_vals, const uint32_t* ranges,
                     const float* cam_pos, const float* vp, const float* c2w_intr,
                     uint32_t w, uint32_t h, uint32_t ts, float min_t,
                     uint32_t tile_begin, uint32_t tile_end, uint32_t tile_step, R* rgba, R* aux, uint32_t* stats,
                     R* ck, R* fin_logt, uint32_t* fin_stop) {
    CAM c;
    FN(cam_from_flat)(&c, cam_pos, vp, c2w_intr, w, h, ts, min_t);
    const uint32_t TS = ts, step = tile_step ? tile_step : 1;
    const int64_t ntiles = tile_end > tile_begin ? ((int64_t)tile_end - tile_begin + step - 1) / step : 0;
#pragma omp parallel if (ntiles >= 32)
    {
        R* st = (R*)malloc(sizeof(R) * TS * TS * 12);
        uint32_t* ss = (uint32_t*)malloc(sizeof(uint32_t) * TS * TS * 4);
        uint32_t* stp = (uint32_t*)malloc(sizeof(uint32_t) * TS * TS);
#pragma omp for schedule(dynamic, 1)
        for (int64_t k = 0; k < ntiles; ++k)
            FN(fwd_tile)(verts, idx, colors, dens, grads, sorted_vals, ranges, &c, w, h, TS,
                         tile_begin + (uint32_t)k * step, rgba, aux, stats, st, ss, stp, ck, fin_logt, fin_stop);
        free(st);
        free(ss);
        free(stp);
    }
}

/* ------------------------------------------------------------------ */
/* Interval-tiled forward and adjoint (forward_interval_tiled,          */
/* lib.rs:1218-1390; backward_interval_tiled, lib.rs:2049-2240):         */
/* per-tet silhouette decomposition (interval_generate.wgsl:174-408),    */
/* tiled compositing over it (interval_tiled_rasterize.wgsl:96-441), the */
/* per-tile adjoint (backward_interval_tiled.wgsl:165-720) and the per-  */
/* tet chain back to the scene (interval_chain_back.wgsl:160-545).       */
/* near/far are the reference's 0.01 / 1000.  The decomposition is kept  */
/* per tet id (the reference stores it per compact visible index but     */
/* reads it back through the tile lists' tet ids: the two agree when the */
/* compaction is the identity).                                          */
/* Parity unpinned: the adjoint's depth and distortion terms, which the  */
/* reference does not form as derivatives of its forward (only the GPU   */
/* is compared with them).                                               */
/* ------------------------------------------------------------------ */
typedef struct {
    R px[5], py[5], zf[5], zb[5], of[5], ob[5]; /* slots 0-3 silhouette, 4 centre */
    R nf[5][3], nb[5][3];                      /* interpolated vertex normals */
    V3 base;                                   /* base colour at the camera */
    R dens;
    uint32_t nsil;                             /* 0: skipped */
    /* the decomposition's discrete choices and weights */
    int inner;                                 /* inner vertex (3 points), 4: quad */
    int ea, eb;                                /* crossing edges (quad) */
    R bw[3];                                   /* inner vertex in its opposite face (3 points) */
    R ta, tb;                                  /* crossing parameters (quad) */
} FN(ivl_);
#define IVL FN(ivl_)
/* flat layout of one decomposition (exported to the tests) */
#define IVL_NF 80

static const int FN(OPPF)[4] = {1, 2, 3, 0};
static const int FN(TE)[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};

static R FN(mixr)(R a, R b, R t) { return a * ((R)1.0 - t) + b * t; }

static void FN(ivl_pack)(const IVL* v, R* o) {
    for (int k = 0; k < 5; ++k) {
        o[k] = v->px[k]; o[5 + k] = v->py[k]; o[10 + k] = v->zf[k]; o[15 + k] = v->zb[k];
        o[20 + k] = v->of[k]; o[25 + k] = v->ob[k];
        for (int c = 0; c < 3; ++c) { o[30 + 3 * k + c] = v->nf[k][c]; o[45 + 3 * k + c] = v->nb[k][c]; }
    }
    o[60] = v->base.x; o[61] = v->base.y; o[62] = v->base.z; o[63] = v->dens;
    o[64] = (R)v->nsil; o[65] = (R)v->inner; o[66] = (R)v->ea; o[67] = (R)v->eb;
    o[68] = v->bw[0]; o[69] = v->bw[1]; o[70] = v->bw[2]; o[71] = v->ta; o[72] = v->tb;
    for (int k = 73; k < IVL_NF; ++k) o[k] = 0;
}
static void FN(ivl_unpack)(const R* o, IVL* v) {
    for (int k = 0; k < 5; ++k) {
        v->px[k] = o[k]; v->py[k] = o[5 + k]; v->zf[k] = o[10 + k]; v->zb[k] = o[15 + k];
        v->of[k] = o[20 + k]; v->ob[k] = o[25 + k];
        for (int c = 0; c < 3; ++c) { v->nf[k][c] = o[30 + 3 * k + c]; v->nb[k][c] = o[45 + 3 * k + c]; }
    }
    v->base = FN(mk)(o[60], o[61], o[62]); v->dens = o[63];
    v->nsil = (uint32_t)o[64]; v->inner = (int)o[65]; v->ea = (int)o[66]; v->eb = (int)o[67];
    v->bw[0] = o[68]; v->bw[1] = o[69]; v->bw[2] = o[70]; v->ta = o[71]; v->tb = o[72];
}

/* interval_generate.wgsl:89-109 */
static uint32_t FN(classify_sil)(const R p[4][2]) {
    for (int v = 0; v < 4; ++v) {
        const int* f = FN(FACES)[v];
        R ax = p[f[0]][0], ay = p[f[0]][1], bx = p[f[1]][0], by = p[f[1]][1], cx = p[f[2]][0], cy = p[f[2]][1];
        R tx = p[f[3]][0], ty = p[f[3]][1];
        R d0 = (bx - ax) * (ty - ay) - (by - ay) * (tx - ax);
        R d1 = (cx - bx) * (ty - by) - (cy - by) * (tx - bx);
        R d2 = (ax - cx) * (ty - cy) - (ay - cy) * (tx - cx);
        int all_pos = d0 >= 0 && d1 >= 0 && d2 >= 0;
        int all_neg = d0 <= 0 && d1 <= 0 && d2 <= 0;
        if (all_pos || all_neg) return (uint32_t)f[3];
    }
    return 4u;
}

/* interval_generate.wgsl:112-126 (also interval_tiled_rasterize.wgsl:74-87) */
static void FN(bary2d)(R ptx, R pty, R ax, R ay, R bx, R by, R cx, R cy, R out[3]) {
    R v0x = bx - ax, v0y = by - ay, v1x = cx - ax, v1y = cy - ay, v2x = ptx - ax, v2y = pty - ay;
    R d00 = v0x * v0x + v0y * v0y, d01 = v0x * v1x + v0y * v1y, d11 = v1x * v1x + v1y * v1y;
    R d20 = v2x * v0x + v2y * v0y, d21 = v2x * v1x + v2y * v1y;
    R denom = d00 * d11 - d01 * d01;
    R inv = RABS(denom) > (R)1e-10f ? (R)1.0 / denom : (R)0.0;
    R v = (d11 * d20 - d01 * d21) * inv;
    R w = (d00 * d21 - d01 * d20) * inv;
    out[0] = (R)1.0 - v - w;
    out[1] = v;
    out[2] = w;
}

/* interval_generate.wgsl:129-139 */
static R FN(line_t)(R a1x, R a1y, R a2x, R a2y, R b1x, R b1y, R b2x, R b2y) {
    R d1x = a2x - a1x, d1y = a2y - a1y, d2x = b2x - b1x, d2y = b2y - b1y;
    R denom = d1x * d2y - d1y * d2x;
    if (RABS(denom) < (R)1e-10f) return (R)-1.0;
    R dx = b1x - a1x, dy = b1y - a1y;
    return (dx * d2y - dy * d2x) / denom;
}

/* Per-vertex NDC z, pixel position and colour offset. */
static int FN(ivl_project)(const CAM* c, const V3 v[4], V3 g, R p[4][2], R nz[4], R off[4], R inv_w[4]) {
    R clip[4][4];
    for (int i = 0; i < 4; ++i) FN(mul_vp)(c->vp, v[i], clip[i]);
    int behind = clip[0][3] <= 0 || clip[1][3] <= 0 || clip[2][3] <= 0 || clip[3][3] <= 0;
    R W = (R)c->width, H = (R)c->height;
    V3 cam = FN(mk)(c->cam[0], c->cam[1], c->cam[2]);
    for (int i = 0; i < 4; ++i) {
        R w = clip[i][3];
        inv_w[i] = RABS(w) > (R)1e-10f ? (R)1.0 / w : (R)0.0;
        R nx = clip[i][0] * inv_w[i], ny = clip[i][1] * inv_w[i];
        nz[i] = clip[i][2] * inv_w[i];
        p[i][0] = ((nx + (R)1.0) * (R)0.5) * W;
        p[i][1] = (((R)1.0 - ny) * (R)0.5) * H;
        off[i] = FN(dot)(g, FN(sub)(v[i], cam));
    }
    return behind;
}

/* The attributes of a decomposition whose discrete choices and weights
 * (inner vertex / crossing edges, bw / ta, tb) are given in o: front/back
 * NDC z, colour offsets, normals of the five slots and the base colour.
 * vn: vertex normals or NULL (zero). */
static void FN(ivl_attrs)(const CAM* c, const V3 v[4], const V3* vn, V3 col, V3 g, IVL* o) {
    R p[4][2], nz[4], off[4], iw[4];
    FN(ivl_project)(c, v, g, p, nz, off, iw);
    V3 cam = FN(mk)(c->cam[0], c->cam[1], c->cam[2]);
    R bo = FN(dot)(g, FN(sub)(cam, v[0]));
    o->base = FN(mk)(col.x + bo, col.y + bo, col.z + bo);
    V3 n[4];
    for (int i = 0; i < 4; ++i) n[i] = vn ? vn[i] : FN(mk)(0, 0, 0);
    int sv[4];
    V3 cnf, cnb;
    if (o->inner < 4) {
        const int* f = FN(FACES)[FN(OPPF)[o->inner]];
        int ce = f[3];
        R z_face = o->bw[0] * nz[f[0]] + o->bw[1] * nz[f[1]] + o->bw[2] * nz[f[2]];
        R z_c = nz[ce];
        R face_off = o->bw[0] * off[f[0]] + o->bw[1] * off[f[1]] + o->bw[2] * off[f[2]];
        V3 n_face = FN(add)(FN(add)(FN(scl)(n[f[0]], o->bw[0]), FN(scl)(n[f[1]], o->bw[1])), FN(scl)(n[f[2]], o->bw[2]));
        int face_first = z_face <= z_c;
        sv[0] = f[0]; sv[1] = f[1]; sv[2] = f[2]; sv[3] = f[0];
        o->px[4] = p[ce][0]; o->py[4] = p[ce][1];
        o->zf[4] = FN(rmin)(z_face, z_c); o->zb[4] = FN(rmax)(z_face, z_c);
        o->of[4] = face_first ? face_off : off[ce];
        o->ob[4] = face_first ? off[ce] : face_off;
        cnf = face_first ? n_face : n[ce];
        cnb = face_first ? n[ce] : n_face;
    } else {
        const int* A = FN(TE)[o->ea];
        const int* Bv = FN(TE)[o->eb];
        R z_a = FN(mixr)(nz[A[0]], nz[A[1]], o->ta), z_b = FN(mixr)(nz[Bv[0]], nz[Bv[1]], o->tb);
        R off_a = FN(mixr)(off[A[0]], off[A[1]], o->ta), off_b = FN(mixr)(off[Bv[0]], off[Bv[1]], o->tb);
        V3 n_a = FN(add)(FN(scl)(n[A[0]], (R)1.0 - o->ta), FN(scl)(n[A[1]], o->ta));
        V3 n_b = FN(add)(FN(scl)(n[Bv[0]], (R)1.0 - o->tb), FN(scl)(n[Bv[1]], o->tb));
        sv[0] = A[0]; sv[1] = Bv[0]; sv[2] = A[1]; sv[3] = Bv[1];
        o->px[4] = FN(mixr)(p[A[0]][0], p[A[1]][0], o->ta);
        o->py[4] = FN(mixr)(p[A[0]][1], p[A[1]][1], o->ta);
        o->zf[4] = FN(rmin)(z_a, z_b); o->zb[4] = FN(rmax)(z_a, z_b);
        o->of[4] = z_a <= z_b ? off_a : off_b;
        o->ob[4] = z_a <= z_b ? off_b : off_a;
        cnf = z_a <= z_b ? n_a : n_b;
        cnb = z_a <= z_b ? n_b : n_a;
    }
    for (int k = 0; k < 4; ++k) {
        o->px[k] = p[sv[k]][0]; o->py[k] = p[sv[k]][1];
        o->zf[k] = o->zb[k] = nz[sv[k]];
        o->of[k] = o->ob[k] = off[sv[k]];
        o->nf[k][0] = o->nb[k][0] = n[sv[k]].x;
        o->nf[k][1] = o->nb[k][1] = n[sv[k]].y;
        o->nf[k][2] = o->nb[k][2] = n[sv[k]].z;
    }
    o->nf[4][0] = cnf.x; o->nf[4][1] = cnf.y; o->nf[4][2] = cnf.z;
    o->nb[4][0] = cnb.x; o->nb[4][1] = cnb.y; o->nb[4][2] = cnb.z;
}

/* interval_generate.wgsl:174-408: the discrete choices, then the attributes. */
static void FN(ivl_generate)(const CAM* c, const V3 v[4], const V3* vn, V3 col, V3 g, R dens, IVL* o) {
    memset(o, 0, sizeof(*o));
    R p[4][2], nz[4], off[4], iw[4];
    if (FN(ivl_project)(c, v, g, p, nz, off, iw)) return;
    o->dens = dens;
    R zmax = FN(rmax)(FN(rmax)(nz[0], nz[1]), FN(rmax)(nz[2], nz[3]));
    R zmin = FN(rmin)(FN(rmin)(nz[0], nz[1]), FN(rmin)(nz[2], nz[3]));
    if (zmax - zmin < (R)1e-7f || dens < (R)1e-6f) return;
    uint32_t iv = FN(classify_sil)((const R(*)[2])p);
    o->inner = (int)iv;
    if (iv < 4u) {
        const int* f = FN(FACES)[FN(OPPF)[iv]];
        FN(bary2d)(p[f[3]][0], p[f[3]][1], p[f[0]][0], p[f[0]][1], p[f[1]][0], p[f[1]][1], p[f[2]][0], p[f[2]][1],
                   o->bw);
        o->nsil = 3u;
    } else {
        static const int PAIRS[3][2] = {{0, 5}, {1, 4}, {2, 3}};
        int ea = 0, eb = 5;
        for (int i = 0; i < 3; ++i) {
            const int* A = FN(TE)[PAIRS[i][0]];
            const int* Bv = FN(TE)[PAIRS[i][1]];
            R ta = FN(line_t)(p[A[0]][0], p[A[0]][1], p[A[1]][0], p[A[1]][1], p[Bv[0]][0], p[Bv[0]][1], p[Bv[1]][0], p[Bv[1]][1]);
            R tb = FN(line_t)(p[Bv[0]][0], p[Bv[0]][1], p[Bv[1]][0], p[Bv[1]][1], p[A[0]][0], p[A[0]][1], p[A[1]][0], p[A[1]][1]);
            if (ta > 0 && ta < 1 && tb > 0 && tb < 1) { ea = PAIRS[i][0]; eb = PAIRS[i][1]; break; }
        }
        const int* A = FN(TE)[ea];
        const int* Bv = FN(TE)[eb];
        o->ea = ea; o->eb = eb;
        o->ta = FN(line_t)(p[A[0]][0], p[A[0]][1], p[A[1]][0], p[A[1]][1], p[Bv[0]][0], p[Bv[0]][1], p[Bv[1]][0], p[Bv[1]][1]);
        o->tb = FN(line_t)(p[Bv[0]][0], p[Bv[0]][1], p[Bv[1]][0], p[Bv[1]][1], p[A[0]][0], p[A[0]][1], p[A[1]][0], p[A[1]][1]);
        o->nsil = 4u;
    }
    FN(ivl_attrs)(c, v, vn, col, g, o);
}

/* interval_tiled_rasterize.wgsl:90-97 */
static int FN(in_tri)(R ptx, R pty, R ax, R ay, R bx, R by, R cx, R cy) {
    R d0 = (bx - ax) * (pty - ay) - (by - ay) * (ptx - ax);
    R d1 = (cx - bx) * (pty - by) - (cy - by) * (ptx - bx);
    R d2 = (ax - cx) * (pty - cy) - (ay - cy) * (ptx - cx);
    int has_neg = d0 < 0 || d1 < 0 || d2 < 0;
    int has_pos = d0 > 0 || d1 > 0 || d2 > 0;
    return !(has_neg && has_pos);
}

static void FN(ivl_all)(const R* verts, const uint32_t* idx, const R* colors, const R* dens, const R* grads,
                        const CAM* c, uint32_t M, IVL* out) {
    for (uint32_t t = 0; t < M; ++t) {
        V3 v[4];
        for (int i = 0; i < 4; ++i) v[i] = FN(load3)(verts, idx[4 * t + i]);
        FN(ivl_generate)(c, v, NULL, FN(load3)(colors, t), FN(load3)(grads, t), dens[t], &out[t]);
    }
}

/* Covered columns of one tile row of the silhouette polygon
 * (interval_tiled_rasterize.wgsl:192-243): 0 when empty. */
static int FN(ivl_row_span)(const IVL* iv, uint32_t row, R tile_ox, R tile_oy, uint32_t TS, int32_t* xl, int32_t* xr) {
    uint32_t ns = iv->nsil;
    R xl_f = (R)1e10f, xr_f = (R)-1e10f, yc = (R)row + (R)0.5;
    for (uint32_t e = 0; e < ns; ++e) {
        uint32_t ej = (e + 1) % ns;
        R yi = iv->py[e] - tile_oy, yj = iv->py[ej] - tile_oy;
        R xi = iv->px[e] - tile_ox, xj = iv->px[ej] - tile_ox;
        if ((yi <= yc && yj > yc) || (yj <= yc && yi > yc)) {
            R tt = (yc - yi) / (yj - yi);
            R x = xi + tt * (xj - xi);
            xl_f = FN(rmin)(xl_f, x); xr_f = FN(rmax)(xr_f, x);
        }
    }
    for (uint32_t k = 0; k < ns; ++k) {
        R vy = iv->py[k] - tile_oy;
        if (vy >= yc - (R)0.5 && vy < yc + (R)0.5) {
            xl_f = FN(rmin)(xl_f, iv->px[k] - tile_ox); xr_f = FN(rmax)(xr_f, iv->px[k] - tile_ox);
        }
    }
    if (!(xl_f <= xr_f)) return 0;
    R eps = (R)0.001f;
    *xl = imax32(sat_i32((double)RCEIL(xl_f - (R)0.5 - eps)), 0);
    *xr = imin32(sat_i32((double)RFLOOR(xr_f - (R)0.5 + eps)), (int32_t)TS - 1);
    return *xl <= *xr;
}

/* The first fan triangle (i, i + 1 mod n, centre) holding the pixel centre. */
static int FN(ivl_find)(const IVL* iv, R pcx, R pcy, uint32_t* ai, uint32_t* bi, R b[3]) {
    uint32_t ns = iv->nsil;
    for (uint32_t tri = 0; tri < ns; ++tri) {
        uint32_t ta = tri, tb = (tri + 1) % ns;
        if (FN(in_tri)(pcx, pcy, iv->px[ta], iv->py[ta], iv->px[tb], iv->py[tb], iv->px[4], iv->py[4])) {
            *ai = ta; *bi = tb;
            FN(bary2d)(pcx, pcy, iv->px[ta], iv->py[ta], iv->px[tb], iv->py[tb], iv->px[4], iv->py[4], b);
            return 1;
        }
    }
    return 0;
}

/* Per-pair quantities shared by the forward and the adjoint
 * (interval_tiled_rasterize.wgsl:300-400, backward_interval_tiled.wgsl:400-440). */
typedef struct {
    R zfn, zbn, zfc, zbc, z_f, z_b, rsc, dist, od, alpha_t, phv, w0, w1;
    V3 cf_raw, cb_raw, cf, cb, cp;
} FN(ivl_pair_);
#define IPAIR FN(ivl_pair_)

static void FN(ivl_pair)(const CAM* c, const IVL* iv, uint32_t ai, uint32_t bi, const R b[3], R pcx, R pcy, IPAIR* q) {
    const R near_p = (R)0.01f, far_p = (R)1000.0f, range_nf = far_p - near_p;
    q->zfn = b[0] * iv->zf[ai] + b[1] * iv->zf[bi] + b[2] * iv->zf[4];
    q->zbn = b[0] * iv->zb[ai] + b[1] * iv->zb[bi] + b[2] * iv->zb[4];
    R ofn = b[0] * iv->of[ai] + b[1] * iv->of[bi] + b[2] * iv->of[4];
    R obn = b[0] * iv->ob[ai] + b[1] * iv->ob[bi] + b[2] * iv->ob[4];
    q->zfc = FN(clampr)(q->zfn, (R)0.0, (R)1.0);
    q->zbc = FN(clampr)(q->zbn, (R)0.0, (R)1.0);
    q->z_f = (near_p * far_p) / (far_p - q->zfc * range_nf);
    q->z_b = (near_p * far_p) / (far_p - q->zbc * range_nf);
    R xc = (pcx - c->intr[2]) / c->intr[0], yc2 = (pcy - c->intr[3]) / c->intr[1];
    q->rsc = RSQRT((xc * xc + yc2 * yc2) + (R)1.0);
    q->dist = RABS(q->z_b - q->z_f) * q->rsc;
    q->cf_raw = FN(mk)(iv->base.x + ofn, iv->base.y + ofn, iv->base.z + ofn);
    q->cb_raw = FN(mk)(iv->base.x + obn, iv->base.y + obn, iv->base.z + obn);
    q->cf = FN(mk)(FN(rmax)(q->cf_raw.x, 0), FN(rmax)(q->cf_raw.y, 0), FN(rmax)(q->cf_raw.z, 0));
    q->cb = FN(mk)(FN(rmax)(q->cb_raw.x, 0), FN(rmax)(q->cb_raw.y, 0), FN(rmax)(q->cb_raw.z, 0));
    q->od = FN(clampr)(iv->dens * q->dist, (R)1e-8f, (R)88.0);
    q->alpha_t = FN(safe_exp)(-q->od);
    q->phv = FN(phi)(q->od);
    q->w0 = q->phv - q->alpha_t;
    q->w1 = (R)1.0 - q->phv;
    q->cp = FN(safe_clip3)(FN(mk)(q->cb.x * q->w0 + q->cf.x * q->w1, q->cb.y * q->w0 + q->cf.y * q->w1,
                                  q->cb.z * q->w0 + q->cf.z * q->w1), (R)0.0, MAX_VAL_R);
}

/* Tiled compositing over given decompositions.  rgba [H*W*4], xyzd [H*W*4],
 * dist [H*W*5] (overwritten for the tiles visited). */
static void FN(ivl_composite)(const IVL* ivs, const uint32_t* sorted_vals, const uint32_t* ranges, const CAM* c,
                              uint32_t w, uint32_t h, uint32_t ts, R* rgba, R* xyzd, R* dist) {
    uint32_t tiles_x = (w + ts - 1) / ts, tiles_y = (h + ts - 1) / ts, TS = ts;
    R* st = (R*)calloc((size_t)TS * TS * 13, sizeof(R)); /* colour 3, log T, xyzd 4, dist 5 */
    for (uint32_t tile = 0; tile < tiles_x * tiles_y; ++tile) {
        uint32_t tile_x = tile % tiles_x, tile_y = tile / tiles_x;
        R tile_ox = (R)(tile_x * TS), tile_oy = (R)(tile_y * TS);
        memset(st, 0, sizeof(R) * TS * TS * 13);
        uint32_t rs = ranges[2 * tile], re = ranges[2 * tile + 1];
        uint32_t cursor = re;
        while (cursor > rs) {
            cursor -= 1;
            const IVL* iv = &ivs[sorted_vals[cursor]];
            if (iv->nsil == 0) continue;
            for (uint32_t row = 0; row < TS; ++row) {
                int32_t xl_i, xr_i;
                if (!FN(ivl_row_span)(iv, row, tile_ox, tile_oy, TS, &xl_i, &xr_i)) continue;
                for (int32_t ci = xl_i; ci <= xr_i; ++ci) {
                    uint32_t colu = (uint32_t)ci, pl = row * TS + colu;
                    uint32_t px = tile_x * TS + colu, py = tile_y * TS + row;
                    if (px >= w || py >= h) continue;
                    R* s = st + 13 * pl;
                    if (s[3] < (R)-13.8) continue;
                    R pcx = (R)px + (R)0.5, pcy = (R)py + (R)0.5;
                    uint32_t ai, bi;
                    R b[3];
                    if (!FN(ivl_find)(iv, pcx, pcy, &ai, &bi, b)) continue;
                    IPAIR q;
                    FN(ivl_pair)(c, iv, ai, bi, b, pcx, pcy, &q);
                    R Tj = FN(safe_exp)(s[3]);
                    if (Tj < (R)1e-6f) continue;
                    s[0] += q.cp.x * Tj; s[1] += q.cp.y * Tj; s[2] += q.cp.z * Tj;
                    s[3] = s[3] - q.od;
                    /* xyzd (interval_tiled_rasterize.wgsl:360-375): normals, depth */
                    R nfv[3], nbv[3];
                    for (int k = 0; k < 3; ++k) {
                        nfv[k] = b[0] * iv->nf[ai][k] + b[1] * iv->nf[bi][k] + b[2] * iv->nf[4][k];
                        nbv[k] = b[0] * iv->nb[ai][k] + b[1] * iv->nb[bi][k] + b[2] * iv->nb[4][k];
                        s[4 + k] += Tj * (nbv[k] * q.w0 + nfv[k] * q.w1);
                    }
                    s[7] += Tj * (q.w0 * q.z_b + q.w1 * q.z_f);
                    /* distortion state (interval_tiled_rasterize.wgsl:383-400) */
                    R enter = q.z_f * q.rsc, exit_d = q.z_b * q.rsc;
                    R u = iv->dens * (exit_d - enter);
                    R en = FN(safe_exp)(-u);
                    R ad = FN(rmax)((R)1.0 - en, 0);
                    R m0 = Tj * ad;
                    R m1 = Tj * (enter * ad + FN(safe_div)((R)1.0 - ((R)1.0 + u) * en, iv->dens));
                    R sd = Tj * Tj * FN(safe_div)((R)1.0 - (R)2.0 * u * en - en * en, iv->dens);
                    R* ds = s + 8;
                    ds[0] += (R)2.0 * m1 * ds[2];
                    ds[1] += (R)2.0 * m0 * ds[3];
                    ds[2] += m0;
                    ds[3] += m1;
                    ds[4] += sd;
                }
            }
        }
        for (uint32_t i = 0; i < TS * TS; ++i) {
            uint32_t row = i / TS, colu = i % TS, px = tile_x * TS + colu, py = tile_y * TS + row;
            if (px >= w || py >= h) continue;
            size_t pix = (size_t)py * w + px;
            R* s = st + 13 * i;
            rgba[4 * pix] = s[0]; rgba[4 * pix + 1] = s[1]; rgba[4 * pix + 2] = s[2];
            rgba[4 * pix + 3] = (R)1.0 - FN(safe_exp)(s[3]);
            for (int k = 0; k < 4; ++k) xyzd[4 * pix + k] = s[4 + k];
            for (int k = 0; k < 5; ++k) dist[5 * pix + k] = s[8 + k];
        }
    }
    free(st);
}

/* rgba [H*W*4], xyzd [H*W*4], dist [H*W*5] (overwritten for the tiles visited). */
void FN(orc_interval_forward)(const R* verts, const uint32_t* idx, const R* colors, const R* dens, const R* grads,
                              const uint32_t* sorted_vals, const uint32_t* ranges, const float* cam_pos,
                              const float* vp, const float* c2w_intr, uint32_t w, uint32_t h, uint32_t ts,
                              uint32_t M, R* rgba, R* xyzd, R* dist) {
    CAM c;
    FN(cam_from_flat)(&c, cam_pos, vp, c2w_intr, w, h, ts, 0.0f);
    IVL* ivs = (IVL*)malloc(sizeof(IVL) * (M ? M : 1));
    FN(ivl_all)(verts, idx, colors, dens, grads, &c, M, ivs);
    FN(ivl_composite)(ivs, sorted_vals, ranges, &c, w, h, ts, rgba, xyzd, dist);
    free(ivs);
}

/* The decompositions [M][IVL_NF] (vn [N*3] vertex normals or NULL). */
void FN(orc_interval_generate)(const R* verts, const uint32_t* idx, const R* colors, const R* dens, const R* grads,
                               const R* vn, const float* cam_pos, const float* vp, const float* c2w_intr, uint32_t w,
                               uint32_t h, uint32_t M, R* out) {
    CAM c;
    FN(cam_from_flat)(&c, cam_pos, vp, c2w_intr, w, h, 12, 0.0f);
    for (uint32_t t = 0; t < M; ++t) {
        V3 v[4], n[4];
        for (int i = 0; i < 4; ++i) {
            v[i] = FN(load3)(verts, idx[4 * t + i]);
            if (vn) n[i] = FN(load3)(vn, idx[4 * t + i]);
        }
        IVL iv;
        FN(ivl_generate)(&c, v, vn ? n : NULL, FN(load3)(colors, t), FN(load3)(grads, t), dens[t], &iv);
        FN(ivl_pack)(&iv, out + (size_t)IVL_NF * t);
    }
}

/* The attributes of (possibly moved) vertices under the discrete choices and
 * weights of the given decompositions (in: [M][IVL_NF]; out likewise). */
void FN(orc_interval_attrs)(const R* verts, const uint32_t* idx, const R* colors, const R* grads, const R* vn,
                            const float* cam_pos, const float* vp, const float* c2w_intr, uint32_t w, uint32_t h,
                            uint32_t M, const R* in, R* out) {
    CAM c;
    FN(cam_from_flat)(&c, cam_pos, vp, c2w_intr, w, h, 12, 0.0f);
    for (uint32_t t = 0; t < M; ++t) {
        IVL iv;
        FN(ivl_unpack)(in + (size_t)IVL_NF * t, &iv);
        if (iv.nsil) {
            V3 v[4], n[4];
            for (int i = 0; i < 4; ++i) {
                v[i] = FN(load3)(verts, idx[4 * t + i]);
                if (vn) n[i] = FN(load3)(vn, idx[4 * t + i]);
            }
            FN(ivl_attrs)(&c, v, vn ? n : NULL, FN(load3)(colors, t), FN(load3)(grads, t), &iv);
        }
        FN(ivl_pack)(&iv, out + (size_t)IVL_NF * t);
    }
}

/* Compositing over given decompositions [M][IVL_NF]. */
void FN(orc_interval_composite)(const R* ivf, const uint32_t* sorted_vals, const uint32_t* ranges,
                                const float* cam_pos, const float* vp, const float* c2w_intr, uint32_t w, uint32_t h,
                                uint32_t ts, uint32_t M, R* rgba, R* xyzd, R* dist) {
    CAM c;
    FN(cam_from_flat)(&c, cam_pos, vp, c2w_intr, w, h, ts, 0.0f);
    IVL* ivs = (IVL*)malloc(sizeof(IVL) * (M ? M : 1));
    for (uint32_t t = 0; t < M; ++t) FN(ivl_unpack)(ivf + (size_t)IVL_NF * t, &ivs[t]);
    FN(ivl_composite)(ivs, sorted_vals, ranges, &c, w, h, ts, rgba, xyzd, dist);
    free(ivs);
}

/* Per-tile adjoint (backward_interval_tiled.wgsl:165-720) over given
 * decompositions: accumulates d_ivl [M][IVL_NF] in the flat layout (slot
 * attributes, normals, base colour, density).  rgba / dist: the forward's
 * outputs; dl_xyzd / dl_dist may be NULL (zero). */
void FN(orc_interval_backward)(const R* ivf, const uint32_t* sorted_vals, const uint32_t* ranges,
                               const float* cam_pos, const float* vp, const float* c2w_intr, uint32_t w, uint32_t h,
                               uint32_t ts, uint32_t M, const R* rgba, const R* dist_img, const R* dl_img,
                               const R* dl_xyzd, const R* dl_dist, R* d_ivl) {
    CAM c;
    FN(cam_from_flat)(&c, cam_pos, vp, c2w_intr, w, h, ts, 0.0f);
    const R near_p = (R)0.01f, far_p = (R)1000.0f, range_nf = far_p - near_p;
    IVL* ivs = (IVL*)malloc(sizeof(IVL) * (M ? M : 1));
    for (uint32_t t = 0; t < M; ++t) FN(ivl_unpack)(ivf + (size_t)IVL_NF * t, &ivs[t]);
    uint32_t tiles_x = (w + ts - 1) / ts, tiles_y = (h + ts - 1) / ts, TS = ts;
    /* per pixel: colour 3, log T, d log T, dist z, dist w, d dist z, d dist w */
    R* st = (R*)calloc((size_t)TS * TS * 9, sizeof(R));
    for (uint32_t tile = 0; tile < tiles_x * tiles_y; ++tile) {
        uint32_t tile_x = tile % tiles_x, tile_y = tile / tiles_x;
        R tile_ox = (R)(tile_x * TS), tile_oy = (R)(tile_y * TS);
        for (uint32_t i = 0; i < TS * TS; ++i) {
            uint32_t row = i / TS, colu = i % TS, px = tile_x * TS + colu, py = tile_y * TS + row;
            R* s = st + 9 * i;
            memset(s, 0, sizeof(R) * 9);
            if (px >= w || py >= h) continue;
            size_t pix = (size_t)py * w + px;
            R a = rgba[4 * pix + 3];
            s[0] = rgba[4 * pix]; s[1] = rgba[4 * pix + 1]; s[2] = rgba[4 * pix + 2];
            s[3] = RLOG(FN(rmax)((R)1.0 - a, (R)1e-20f));
            s[4] = -dl_img[4 * pix + 3] * ((R)1.0 - a);
            s[5] = dist_img[5 * pix + 2];
            s[6] = dist_img[5 * pix + 3];
            s[7] = dl_dist ? dl_dist[5 * pix + 2] : 0;
            s[8] = dl_dist ? dl_dist[5 * pix + 3] : 0;
        }
        uint32_t rs = ranges[2 * tile], re = ranges[2 * tile + 1];
        for (uint32_t cursor = rs; cursor < re; ++cursor) {
            uint32_t t = sorted_vals[cursor];
            const IVL* iv = &ivs[t];
            if (iv->nsil == 0) continue;
            R* D = d_ivl + (size_t)IVL_NF * t;
            for (uint32_t row = 0; row < TS; ++row) {
                int32_t xl_i, xr_i;
                if (!FN(ivl_row_span)(iv, row, tile_ox, tile_oy, TS, &xl_i, &xr_i)) continue;
                for (int32_t ci = xl_i; ci <= xr_i; ++ci) {
                    uint32_t colu = (uint32_t)ci, pl = row * TS + colu;
                    uint32_t px = tile_x * TS + colu, py = tile_y * TS + row;
                    if (px >= w || py >= h) continue;
                    size_t pix = (size_t)py * w + px;
                    R* s = st + 9 * pl;
                    R pcx = (R)px + (R)0.5, pcy = (R)py + (R)0.5;
                    uint32_t ai, bi;
                    R b[3];
                    if (!FN(ivl_find)(iv, pcx, pcy, &ai, &bi, b)) continue;
                    IPAIR q;
                    FN(ivl_pair)(&c, iv, ai, bi, b, pcx, pcy, &q);
                    /* undo (backward_interval_tiled.wgsl:432-447) */
                    R d_log_t = s[4];
                    R prev_log_t = FN(rmin)(s[3] + q.od, (R)0.0);
                    R t_prev = FN(safe_exp)(prev_log_t);
                    s[0] -= q.cp.x * t_prev; s[1] -= q.cp.y * t_prev; s[2] -= q.cp.z * t_prev;
                    s[3] = prev_log_t;
                    if (t_prev < (R)1e-6f) continue;
                    V3 d_color = FN(mk)(dl_img[4 * pix], dl_img[4 * pix + 1], dl_img[4 * pix + 2]);
                    V3 d_cp = FN(safe_clip3)(FN(scl)(d_color, t_prev), MIN_VAL_R, MAX_VAL_R);
                    R d_od_state = -d_log_t;
                    s[4] = FN(safe_clip)(d_log_t + FN(dot)(d_color, q.cp) * t_prev, MIN_VAL_R, MAX_VAL_R);
                    R dphi = FN(dphi_dx)(q.od);
                    R dw0 = dphi + q.alpha_t, dw1 = -dphi;
                    V3 d_cb_int = FN(scl)(d_cp, q.w0), d_cf_int = FN(scl)(d_cp, q.w1);
                    R d_od_int = FN(dot)(d_cp, FN(add)(FN(scl)(q.cb, dw0), FN(scl)(q.cf, dw1)));
                    R d_od = FN(safe_clip)(d_od_state + d_od_int, MIN_VAL_R, MAX_VAL_R);
                    R od_raw = iv->dens * q.dist;
                    R d_od_raw = (od_raw >= (R)1e-8f && od_raw <= (R)88.0) ? d_od : (R)0.0;
                    R d_dens_local = FN(safe_clip)(d_od_raw * q.dist, MIN_VAL_R, MAX_VAL_R);
                    R d_dist = FN(safe_clip)(d_od_raw * iv->dens, MIN_VAL_R, MAX_VAL_R);
                    R sgn = q.z_b >= q.z_f ? (R)1.0 : (R)-1.0;
                    R d_zf_view = -d_dist * q.rsc * sgn, d_zb_view = d_dist * q.rsc * sgn;
                    R den_f = far_p - q.zfc * range_nf, den_b = far_p - q.zbc * range_nf;
                    R dz_f = near_p * far_p * range_nf / (den_f * den_f);
                    R dz_b = near_p * far_p * range_nf / (den_b * den_b);
                    int zf_act = q.zfn >= 0 && q.zfn <= 1, zb_act = q.zbn >= 0 && q.zbn <= 1;
                    R d_zfn = zf_act ? d_zf_view * dz_f : (R)0.0;
                    R d_zbn = zb_act ? d_zb_view * dz_b : (R)0.0;
                    V3 d_cb_raw = FN(mk)(q.cb_raw.x > 0 ? d_cb_int.x : 0, q.cb_raw.y > 0 ? d_cb_int.y : 0,
                                         q.cb_raw.z > 0 ? d_cb_int.z : 0);
                    V3 d_cf_raw = FN(mk)(q.cf_raw.x > 0 ? d_cf_int.x : 0, q.cf_raw.y > 0 ? d_cf_int.y : 0,
                                         q.cf_raw.z > 0 ? d_cf_int.z : 0);
                    V3 d_bc = FN(add)(d_cf_raw, d_cb_raw);
                    R d_of = (d_cf_raw.x + d_cf_raw.y) + d_cf_raw.z;
                    R d_ob = (d_cb_raw.x + d_cb_raw.y) + d_cb_raw.z;
                    const uint32_t sl[3] = {ai, bi, 4u};
                    for (int k = 0; k < 3; ++k) {
                        D[10 + sl[k]] += b[k] * d_zfn;
                        D[15 + sl[k]] += b[k] * d_zbn;
                        D[20 + sl[k]] += b[k] * d_of;
                        D[25 + sl[k]] += b[k] * d_ob;
                    }
                    V3 cbc = FN(safe_clip3)(d_bc, MIN_VAL_R, MAX_VAL_R);
                    D[60] += cbc.x; D[61] += cbc.y; D[62] += cbc.z;
                    D[63] += FN(safe_clip)(d_dens_local, MIN_VAL_R, MAX_VAL_R);
                    /* xyzd (backward_interval_tiled.wgsl:544-600) */
                    R dx4[4] = {0, 0, 0, 0};
                    if (dl_xyzd) for (int k = 0; k < 4; ++k) dx4[k] = dl_xyzd[4 * pix + k];
                    R nfv[3], nbv[3], ncon[3];
                    for (int k = 0; k < 3; ++k) {
                        nfv[k] = b[0] * iv->nf[ai][k] + b[1] * iv->nf[bi][k] + b[2] * iv->nf[4][k];
                        nbv[k] = b[0] * iv->nb[ai][k] + b[1] * iv->nb[bi][k] + b[2] * iv->nb[4][k];
                        ncon[k] = nbv[k] * q.w0 + nfv[k] * q.w1;
                    }
                    R depth_c = q.w0 * q.z_b + q.w1 * q.z_f;
                    R d_od_x = -(((dx4[0] * ncon[0] + dx4[1] * ncon[1]) + dx4[2] * ncon[2]) + dx4[3] * depth_c) * t_prev;
                    R dn[3] = {dx4[0] * t_prev, dx4[1] * t_prev, dx4[2] * t_prev};
                    R d_depth = dx4[3] * t_prev;
                    R d_od_w = (dn[0] * (nbv[0] * dw0 + nfv[0] * dw1) + dn[1] * (nbv[1] * dw0 + nfv[1] * dw1)) +
                               dn[2] * (nbv[2] * dw0 + nfv[2] * dw1) + d_depth * (q.z_b * dw0 + q.z_f * dw1);
                    D[63] += FN(safe_clip)((d_od_x + d_od_w) * q.dist, MIN_VAL_R, MAX_VAL_R);
                    R d_zf_x = zf_act ? (d_depth * q.w1) * dz_f : (R)0.0;
                    R d_zb_x = zb_act ? (d_depth * q.w0) * dz_b : (R)0.0;
                    for (int k = 0; k < 3; ++k) {
                        D[10 + sl[k]] += b[k] * d_zf_x;
                        D[15 + sl[k]] += b[k] * d_zb_x;
                        for (int j = 0; j < 3; ++j) {
                            D[30 + 3 * sl[k] + j] += b[k] * (dn[j] * q.w1);
                            D[45 + 3 * sl[k] + j] += b[k] * (dn[j] * q.w0);
                        }
                    }
                    /* distortion (backward_interval_tiled.wgsl:602-650) */
                    R enter = q.z_f * q.rsc, exit_d = q.z_b * q.rsc, dt_d = exit_d - enter;
                    R u = iv->dens * dt_d;
                    R en = FN(safe_exp)(-u);
                    R ad = FN(rmax)((R)1.0 - en, 0);
                    R m0 = t_prev * ad;
                    R m1 = t_prev * (enter * ad + FN(safe_div)((R)1.0 - ((R)1.0 + u) * en, iv->dens));
                    R old_z = s[5] - m0, old_w = s[6] - m1;
                    s[5] = old_z; s[6] = old_w;
                    R d_x = dl_dist ? dl_dist[5 * pix] : 0, d_y = dl_dist ? dl_dist[5 * pix + 1] : 0;
                    R d_z = s[7], d_w = s[8];
                    R d_m0 = d_z + (R)2.0 * old_w * d_y;
                    s[7] = d_z + (R)2.0 * m1 * d_x;
                    s[8] = d_w + (R)2.0 * m0 * d_y;
                    R d_u = (d_m0 * t_prev) * en;
                    D[63] += FN(safe_clip)(d_u * dt_d, MIN_VAL_R, MAX_VAL_R);
                }
            }
        }
    }
    free(st);
    free(ivs);
}

/* d(ndc z)/d(v) of vertex v with 1/w = iw: (vp row z - ndc z vp row w) / w */
static V3 FN(ivl_dz)(const CAM* c, R nz, R iw, R g) {
    V3 rz = FN(mk)(c->vp[2], c->vp[6], c->vp[10]), rw = FN(mk)(c->vp[3], c->vp[7], c->vp[11]);
    return FN(scl)(FN(scl)(FN(sub)(rz, FN(scl)(rw, nz)), iw), g);
}

/* Per-tet chain back (interval_chain_back.wgsl:160-545) of d_ivl
 * [M][IVL_NF] to the scene: d_verts [N*3], d_dens [M], d_base [M*3],
 * d_grads [M*3], d_vn [N*3] (accumulated). */
void FN(orc_interval_chain)(const R* verts, const uint32_t* idx, const R* grads, const float* cam_pos,
                            const float* vp, const float* c2w_intr, uint32_t w, uint32_t h, uint32_t M,
                            const R* ivf, const R* d_ivl, R* d_verts, R* d_dens, R* d_base, R* d_grads, R* d_vn) {
    CAM c;
    FN(cam_from_flat)(&c, cam_pos, vp, c2w_intr, w, h, 12, 0.0f);
    V3 cam = FN(mk)(c.cam[0], c.cam[1], c.cam[2]);
    for (uint32_t t = 0; t < M; ++t) {
        IVL iv, d;
        FN(ivl_unpack)(ivf + (size_t)IVL_NF * t, &iv);
        if (iv.nsil == 0) continue;
        FN(ivl_unpack)(d_ivl + (size_t)IVL_NF * t, &d);
        V3 v[4];
        for (int i = 0; i < 4; ++i) v[i] = FN(load3)(verts, idx[4 * t + i]);
        V3 g = FN(load3)(grads, t);
        R p[4][2], nz[4], off[4], iw[4];
        FN(ivl_project)(&c, v, g, p, nz, off, iw);
        V3 dv[4], dn[4];
        for (int i = 0; i < 4; ++i) { dv[i] = FN(mk)(0, 0, 0); dn[i] = FN(mk)(0, 0, 0); }
        V3 dg = FN(mk)(0, 0, 0);
        R dbs = (d.base.x + d.base.y) + d.base.z;
        dg = FN(add)(dg, FN(scl)(FN(sub)(cam, v[0]), dbs));
        dv[0] = FN(sub)(dv[0], FN(scl)(g, dbs));
#define CLIP3(x) FN(safe_clip3)((x), MIN_VAL_R, MAX_VAL_R)
        int sv[4];
        if (iv.inner < 4) {
            const int* f = FN(FACES)[FN(OPPF)[iv.inner]];
            sv[0] = f[0]; sv[1] = f[1]; sv[2] = f[2]; sv[3] = f[0];
        } else {
            const int* A = FN(TE)[iv.ea];
            const int* Bv = FN(TE)[iv.eb];
            sv[0] = A[0]; sv[1] = Bv[0]; sv[2] = A[1]; sv[3] = Bv[1];
        }
        for (int i = 0; i < 4; ++i) {
            int vi = sv[i];
            R dz = d.zf[i] + d.zb[i];
            dv[vi] = FN(add)(dv[vi], CLIP3(FN(ivl_dz)(&c, nz[vi], iw[vi], dz)));
            R doff = d.of[i] + d.ob[i];
            dv[vi] = FN(add)(dv[vi], CLIP3(FN(scl)(g, doff)));
            dg = FN(add)(dg, CLIP3(FN(scl)(FN(sub)(v[vi], cam), doff)));
            dn[vi] = FN(add)(dn[vi], FN(add)(FN(mk)(d.nf[i][0], d.nf[i][1], d.nf[i][2]),
                                             FN(mk)(d.nb[i][0], d.nb[i][1], d.nb[i][2])));
        }
        V3 cnf = FN(mk)(d.nf[4][0], d.nf[4][1], d.nf[4][2]), cnb = FN(mk)(d.nb[4][0], d.nb[4][1], d.nb[4][2]);
        if (iv.inner < 4) {
            const int* f = FN(FACES)[FN(OPPF)[iv.inner]];
            int ce = f[3];
            R z_face = iv.bw[0] * nz[f[0]] + iv.bw[1] * nz[f[1]] + iv.bw[2] * nz[f[2]];
            int ff = z_face <= nz[ce];
            R dzf = ff ? d.zf[4] : d.zb[4], dzc = ff ? d.zb[4] : d.zf[4];
            R dof = ff ? d.of[4] : d.ob[4], doc = ff ? d.ob[4] : d.of[4];
            for (int i = 0; i < 3; ++i)
                dv[f[i]] = FN(add)(dv[f[i]], CLIP3(FN(ivl_dz)(&c, nz[f[i]], iw[f[i]], iv.bw[i] * dzf)));
            dv[ce] = FN(add)(dv[ce], CLIP3(FN(ivl_dz)(&c, nz[ce], iw[ce], dzc)));
            for (int i = 0; i < 3; ++i) {
                R doi = iv.bw[i] * dof;
                dv[f[i]] = FN(add)(dv[f[i]], CLIP3(FN(scl)(g, doi)));
                dg = FN(add)(dg, CLIP3(FN(scl)(FN(sub)(v[f[i]], cam), doi)));
            }
            dv[ce] = FN(add)(dv[ce], CLIP3(FN(scl)(g, doc)));
            dg = FN(add)(dg, CLIP3(FN(scl)(FN(sub)(v[ce], cam), doc)));
            V3 dnf = ff ? cnf : cnb, dnc = ff ? cnb : cnf;
            for (int i = 0; i < 3; ++i) dn[f[i]] = FN(add)(dn[f[i]], FN(scl)(dnf, iv.bw[i]));
            dn[ce] = FN(add)(dn[ce], dnc);
        } else {
            const int* A = FN(TE)[iv.ea];
            const int* Bv = FN(TE)[iv.eb];
            R ta = iv.ta, tb = iv.tb;
            R z_a = FN(mixr)(nz[A[0]], nz[A[1]], ta), z_b = FN(mixr)(nz[Bv[0]], nz[Bv[1]], tb);
            int af = z_a <= z_b;
            R dza = af ? d.zf[4] : d.zb[4], dzb = af ? d.zb[4] : d.zf[4];
            R doa = af ? d.of[4] : d.ob[4], dob = af ? d.ob[4] : d.of[4];
            dv[A[0]] = FN(add)(dv[A[0]], CLIP3(FN(ivl_dz)(&c, nz[A[0]], iw[A[0]], ((R)1.0 - ta) * dza)));
            dv[A[1]] = FN(add)(dv[A[1]], CLIP3(FN(ivl_dz)(&c, nz[A[1]], iw[A[1]], ta * dza)));
            dv[Bv[0]] = FN(add)(dv[Bv[0]], CLIP3(FN(ivl_dz)(&c, nz[Bv[0]], iw[Bv[0]], ((R)1.0 - tb) * dzb)));
            dv[Bv[1]] = FN(add)(dv[Bv[1]], CLIP3(FN(ivl_dz)(&c, nz[Bv[1]], iw[Bv[1]], tb * dzb)));
            R oa0 = ((R)1.0 - ta) * doa, oa1 = ta * doa, ob0 = ((R)1.0 - tb) * dob, ob1 = tb * dob;
            dv[A[0]] = FN(add)(dv[A[0]], CLIP3(FN(scl)(g, oa0)));
            dv[A[1]] = FN(add)(dv[A[1]], CLIP3(FN(scl)(g, oa1)));
            dg = FN(add)(dg, CLIP3(FN(add)(FN(scl)(FN(sub)(v[A[0]], cam), oa0), FN(scl)(FN(sub)(v[A[1]], cam), oa1))));
            dv[Bv[0]] = FN(add)(dv[Bv[0]], CLIP3(FN(scl)(g, ob0)));
            dv[Bv[1]] = FN(add)(dv[Bv[1]], CLIP3(FN(scl)(g, ob1)));
            dg = FN(add)(dg, CLIP3(FN(add)(FN(scl)(FN(sub)(v[Bv[0]], cam), ob0), FN(scl)(FN(sub)(v[Bv[1]], cam), ob1))));
            V3 dna = af ? cnf : cnb, dnb = af ? cnb : cnf;
            dn[A[0]] = FN(add)(dn[A[0]], FN(scl)(dna, (R)1.0 - ta));
            dn[A[1]] = FN(add)(dn[A[1]], FN(scl)(dna, ta));
            dn[Bv[0]] = FN(add)(dn[Bv[0]], FN(scl)(dnb, (R)1.0 - tb));
            dn[Bv[1]] = FN(add)(dn[Bv[1]], FN(scl)(dnb, tb));
        }
        d_dens[t] += d.dens;
        for (int i = 0; i < 4; ++i) {
            uint32_t gi = 3 * idx[4 * t + i];
            V3 a = CLIP3(dv[i]), b = CLIP3(dn[i]);
            d_verts[gi] += a.x; d_verts[gi + 1] += a.y; d_verts[gi + 2] += a.z;
            d_vn[gi] += b.x; d_vn[gi + 1] += b.y; d_vn[gi + 2] += b.z;
        }
        V3 a = CLIP3(dg), b = CLIP3(d.base);
        d_grads[3 * t] += a.x; d_grads[3 * t + 1] += a.y; d_grads[3 * t + 2] += a.z;
        d_base[3 * t] += b.x; d_base[3 * t + 1] += b.y; d_base[3 * t + 2] += b.z;
#undef CLIP3
    }
}

/* ------------------------------------------------------------------ */
/* Per-tet error statistics for densification (forward_error)           */
/* crates/rmesh-error/src/wgsl/error_tiled_compute.wgsl:88-373,          */
/* crates/rmesh-python/src/lib.rs:1621-1711.  Same tile traversal and    */
/* forward spans as orc_forward; per covered pixel with log T >= -5.5 and */
/* a valid slab, T_val = exp(log T) * od (when exp(log T) >= 1e-6) adds   */
/* into tet_err[16 t + k]: 0 T_val, 1 T_val pe, 2 T_val pe^2, 5 T_val se, */
/* 6-8 T_val entry, 9-11 T_val exit, 12 T_val, 13-15 T_val centroid        */
/* (3, 4 stay 0); tet_count[2 t] += 1, tet_count[2 t + 1] = max(...,       */
/* i32(65535 T_val)); then log T -= od.  Outputs accumulate.              */
/* ------------------------------------------------------------------ */
void FN(orc_error)(const R* verts, const uint32_t* idx, const R* dens, const uint32_t* sorted_vals,
                   const uint32_t* ranges, const float* cam_pos, const float* vp, const float* c2w_intr, uint32_t w,
                   uint32_t h, uint32_t ts, float min_t, const R* pixel_err, const R* ssim_err, R* tet_err,
                   int32_t* tet_count) {
    CAM c;
    FN(cam_from_flat)(&c, cam_pos, vp, c2w_intr, w, h, ts, min_t);
    uint32_t tiles_x = (w + ts - 1) / ts, tiles_y = (h + ts - 1) / ts;
    uint32_t TS = ts;
    R* log_t = (R*)calloc((size_t)TS * TS, sizeof(R));
    V3 cam = FN(mk)(c.cam[0], c.cam[1], c.cam[2]);
    for (uint32_t tile = 0; tile < tiles_x * tiles_y; ++tile) {
        uint32_t tile_x = tile % tiles_x, tile_y = tile / tiles_x;
        R tile_ox = (R)(tile_x * TS), tile_oy = (R)(tile_y * TS);
        memset(log_t, 0, sizeof(R) * TS * TS);
        uint32_t rs = ranges[2 * tile], re = ranges[2 * tile + 1];
        uint32_t cursor = re;
        while (cursor > rs) {
            cursor -= 1;
            uint32_t t = sorted_vals[cursor];
            V3 v[4];
            for (int i = 0; i < 4; ++i) v[i] = FN(load3)(verts, idx[4 * t + i]);
            R proj[4][2];
            int any_behind = FN(tile_proj)(&c, v, tile_ox, tile_oy, proj);
            V3 cen = FN(scl)(FN(add)(FN(add)(FN(add)(v[0], v[1]), v[2]), v[3]), (R)0.25);
            V3 fn[4];
            for (int fi = 0; fi < 4; ++fi) {
                const int* f = FN(FACES)[fi];
                V3 n = FN(cross)(FN(sub)(v[f[2]], v[f[0]]), FN(sub)(v[f[1]], v[f[0]]));
                if (FN(dot)(n, FN(sub)(v[f[3]], v[f[0]])) < 0) n = FN(neg)(n);
                fn[fi] = n;
            }
            for (uint32_t row = 0; row < TS; ++row) {
                int32_t xl_i = 0, xr_i = -1;
                if (!any_behind) FN(row_span)(proj, row, TS, 1, &xl_i, &xr_i);
                for (int32_t col_i = xl_i; col_i <= xr_i; ++col_i) {
                    uint32_t colu = (uint32_t)col_i;
                    uint32_t pl = row * TS + colu;
                    uint32_t px = tile_x * TS + colu, py = tile_y * TS + row;
                    if (px >= w || py >= h) continue;
                    if (log_t[pl] < (R)-5.5) continue;
                    V3 rd = FN(pixel_ray)(&c, px, py);
                    V3 ro = FN(add)(cam, FN(scl)(rd, c.min_t));
                    R t_min = (R)-3.402823e38f, t_max = (R)3.402823e38f;
                    int valid = 1;
                    for (int fi = 0; fi < 4; ++fi) {
                        const int* f = FN(FACES)[fi];
                        R num = FN(dot)(fn[fi], FN(sub)(v[f[0]], ro));
                        R den = FN(dot)(fn[fi], rd);
                        if (RABS(den) < (R)1e-20f) { if (num > 0) valid = 0; continue; }
                        R tt = num / den;
                        if (den > 0) { if (tt > t_min) t_min = tt; }
                        else { if (tt < t_max) t_max = tt; }
                    }
                    if (!(valid && t_min < t_max)) continue;
                    R od = FN(clampr)(dens[t] * (t_max - t_min), (R)1e-8f, (R)88.0);
                    R trans = FN(safe_exp)(log_t[pl]);
                    if (trans < (R)1e-6f) continue;
                    R tv = trans * od;
                    V3 en = FN(add)(ro, FN(scl)(rd, t_min)), ex = FN(add)(ro, FN(scl)(rd, t_max));
                    size_t pix = (size_t)py * w + px;
                    R pe = pixel_err[pix], se = ssim_err[pix];
                    R* e = tet_err + 16 * (size_t)t;
                    e[0] += tv; e[1] += tv * pe; e[2] += tv * pe * pe; e[5] += tv * se;
                    e[6] += tv * en.x; e[7] += tv * en.y; e[8] += tv * en.z;
                    e[9] += tv * ex.x; e[10] += tv * ex.y; e[11] += tv * ex.z;
                    e[12] += tv; e[13] += tv * cen.x; e[14] += tv * cen.y; e[15] += tv * cen.z;
                    tet_count[2 * (size_t)t] += 1;
                    int32_t mx = sat_i32((double)((R)65535.0 * tv));
                    if (mx > tet_count[2 * (size_t)t + 1]) tet_count[2 * (size_t)t + 1] = mx;
                    log_t[pl] = log_t[pl] - od;
                }
            }
        }
    }
    free(log_t);
}

/* ------------------------------------------------------------------ */
/* K14: per-pixel loss gradient (+ loss value, an extension: the         */
/* reference never writes its loss buffer).                             */
/* crates/rmesh-train/src/wgsl/loss_compute.wgsl:24-72                  */
/* ------------------------------------------------------------------ */
double FN(orc_loss)(const R* rgba, const R* gt, uint32_t w, uint32_t h, uint32_t loss_type, R* dl) {
    R n_pixels = (R)(w * h);
    double loss = 0.0;
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        for (int k = 0; k < 3; ++k) {
            R d = rgba[4 * i + k] - gt[3 * i + k];
            R g;
            if (loss_type == 0) {
                R sg = d > 0 ? (R)1.0 : (d < 0 ? (R)-1.0 : (R)0.0);
                g = sg / n_pixels;
                loss += fabs((double)d);
            } else {
                g = ((R)2.0 * d) / n_pixels;
                loss += (double)d * (double)d;
            }
            dl[4 * i + k] = g;
        }
        dl[4 * i + 3] = 0;
    }
    return loss / (double)n_pixels;
}

/* ------------------------------------------------------------------ */
/* K13: reverse-order tiled adjoint                                      */
/* crates/rmesh-backward/src/wgsl/backward_tiled_compute.wgsl:154-587   */
/* Gradients accumulate (+=) into the output arrays.                    */
/*                                                                      */
/* Checkpointed mode (ck != NULL; no counterpart in the reference): the  */
/* undo restarts at every 32-tet group of the tile list from the log T   */
/* the forward stored at the group's end (or the pixel's final log T),   */
/* instead of walking the whole list back from log(1 - alpha), and tets  */
/* at or after the pixel's stop position (never composited) are skipped. */
/* That is the exact adjoint of the forward as computed.                 */
/* ------------------------------------------------------------------ */
#define BWD_TERMS 19
/* One tile; acc[(cursor - rs) * 19]: the per-tet sums of this tile
 * (density, 4 vertices, colour gradient, colour), st: TS*TS*5 and cur:
 * TS*TS scratch of the calling thread. */
static void FN(bwd_tile)(const R* verts, const uint32_t* idx, const R* colors, const R* dens, const R* grads,
                         const uint32_t* sorted_vals, const uint32_t* ranges, const CAM* cp, uint32_t w, uint32_t h,
                         uint32_t TS, uint32_t tile, const R* rgba, const R* dl, const R* ck, const R* fin_logt,
                         const uint32_t* fin_stop, R* st, int64_t* cur, R* acc) {
    const CAM c = *cp;
    uint32_t tiles_x = (w + TS - 1) / TS;
    V3 cam = FN(mk)(c.cam[0], c.cam[1], c.cam[2]);
    {
        uint32_t tile_x = tile % tiles_x, tile_y = tile / tiles_x;
        R tile_ox = (R)(tile_x * TS), tile_oy = (R)(tile_y * TS);
        for (uint32_t i = 0; i < TS * TS; ++i) {
            uint32_t row = i / TS, colu = i % TS;
            uint32_t px = tile_x * TS + colu, py = tile_y * TS + row;
            R* s = st + 5 * i;
            cur[i] = -1;
            if (px < w && py < h) {
                size_t pix = (size_t)py * w + px;
                R a = rgba[4 * pix + 3];
                s[0] = rgba[4 * pix]; s[1] = rgba[4 * pix + 1]; s[2] = rgba[4 * pix + 2];
                if (ck) {
                    s[3] = fin_logt[pix];
                    s[4] = -dl[4 * pix + 3] * FN(safe_exp)(fin_logt[pix]);
                } else {
                    s[3] = RLOG(FN(rmax)((R)1.0 - a, (R)1e-20f));
                    s[4] = -dl[4 * pix + 3] * ((R)1.0 - a);
                }
            } else {
                s[0] = s[1] = s[2] = s[3] = s[4] = 0;
            }
        }
        uint32_t rs = ranges[2 * tile], re = ranges[2 * tile + 1];
        for (uint32_t cursor = rs; cursor < re; ++cursor) {
            const uint32_t q = re - 1 - cursor; /* position in compositing order */
            uint32_t t = sorted_vals[cursor];
            uint32_t vi[4];
            V3 v[4];
            for (int i = 0; i < 4; ++i) { vi[i] = idx[4 * t + i]; v[i] = FN(load3)(verts, vi[i]); }
            R proj[4][2];
            int any_behind = FN(tile_proj)(&c, v, tile_ox, tile_oy, proj);
            R density = dens[t];
            V3 col = FN(load3)(colors, t), g = FN(load3)(grads, t);
            R acc_d = 0;
            V3 acc_v[4] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            V3 acc_g = FN(mk)(0, 0, 0), acc_c = FN(mk)(0, 0, 0);
            for (uint32_t row = 0; row < TS; ++row) {
                int32_t xl_i = 0, xr_i = -1;
                if (!any_behind) FN(row_span)(proj, row, TS, 0, &xl_i, &xr_i);
                for (int32_t col_i = xl_i; col_i <= xr_i; ++col_i) {
                    uint32_t colu = (uint32_t)col_i;
                    uint32_t pl = row * TS + colu;
                    uint32_t px = tile_x * TS + colu, py = tile_y * TS + row;
                    if (px >= w || py >= h) continue;
                    size_t pix = (size_t)py * w + px;
                    V3 rd = FN(pixel_ray)(&c, px, py);
                    R t_min = (R)-3.402823e38f, t_max = (R)3.402823e38f;
                    int valid = 1, min_face = 0, max_face = 0;
                    for (int fi = 0; fi < 4; ++fi) {
                        const int* f = FN(FACES)[fi];
                        V3 n = FN(cross)(FN(sub)(v[f[2]], v[f[0]]), FN(sub)(v[f[1]], v[f[0]]));
                        if (FN(dot)(n, FN(sub)(v[f[3]], v[f[0]])) < 0) n = FN(neg)(n);
                        R num = FN(dot)(n, FN(sub)(v[f[0]], cam));
                        R den = FN(dot)(n, rd);
                        if (RABS(den) < (R)1e-20f) { if (num > 0) valid = 0; continue; }
                        R tt = num / den;
                        if (den > 0) { if (tt > t_min) { t_min = tt; min_face = fi; } }
                        else { if (tt < t_max) { t_max = tt; max_face = fi; } }
                    }
                    if (!valid || t_min >= t_max) continue;
                    R base_offset = FN(dot)(g, FN(sub)(cam, v[0]));
                    V3 bc = FN(mk)(col.x + base_offset, col.y + base_offset, col.z + base_offset);
                    R dc_dt = FN(dot)(g, rd);
                    R a0 = dc_dt * t_min, a1 = dc_dt * t_max;
                    V3 csr = FN(mk)(bc.x + a0, bc.y + a0, bc.z + a0);
                    V3 cer = FN(mk)(bc.x + a1, bc.y + a1, bc.z + a1);
                    V3 cs = FN(mk)(FN(rmax)(csr.x, 0), FN(rmax)(csr.y, 0), FN(rmax)(csr.z, 0));
                    V3 ce = FN(mk)(FN(rmax)(cer.x, 0), FN(rmax)(cer.y, 0), FN(rmax)(cer.z, 0));
                    R dist = t_max - t_min;
                    R od = FN(clampr)(density * dist, (R)1e-8f, (R)88.0);
                    R alpha_t = FN(safe_exp)(-od);
                    R phi_v = FN(phi)(od);
                    R w0 = phi_v - alpha_t, w1 = (R)1.0 - phi_v;
                    V3 cp = FN(safe_clip3)(FN(mk)(ce.x * w0 + cs.x * w1, ce.y * w0 + cs.y * w1, ce.z * w0 + cs.z * w1),
                                           MIN_VAL_R, MAX_VAL_R);
                    R* s = st + 5 * pl;
                    if (ck) {
                        const uint32_t stop = fin_stop[pix];
                        if (q >= stop) continue; /* never composited */
                        const int64_t j = q / CK_GROUP;
                        if (j != cur[pl]) {
                            const uint32_t end = (uint32_t)(j + 1) * CK_GROUP;
                            s[3] = end < stop ? ck[FN(ck_slot)(rs, tile, (uint32_t)j + 1) * TS * TS + pl] : fin_logt[pix];
                            cur[pl] = j;
                        }
                    }
                    R log_t = s[3], d_log_t = s[4];
                    R prev_log_t = FN(rmin)(log_t + od, (R)0.0);
                    R t_prev = FN(safe_exp)(prev_log_t);
                    s[0] = s[0] - cp.x * t_prev; s[1] = s[1] - cp.y * t_prev; s[2] = s[2] - cp.z * t_prev;
                    s[3] = prev_log_t;
                    if (t_prev < (R)1e-6f) continue;
                    V3 dcol = FN(mk)(dl[4 * pix], dl[4 * pix + 1], dl[4 * pix + 2]);
                    V3 dcp = FN(safe_clip3)(FN(scl)(dcol, t_prev), MIN_VAL_R, MAX_VAL_R);
                    R d_od_state = -d_log_t;
                    s[4] = FN(safe_clip)(d_log_t + FN(dot)(dcol, cp) * t_prev, MIN_VAL_R, MAX_VAL_R);
                    R dphi = FN(dphi_dx)(od);
                    R dw0 = dphi + alpha_t, dw1 = -dphi;
                    V3 dce_i = FN(scl)(dcp, w0), dcs_i = FN(scl)(dcp, w1);
                    R d_od_int = FN(dot)(dcp, FN(mk)(ce.x * dw0 + cs.x * dw1, ce.y * dw0 + cs.y * dw1, ce.z * dw0 + cs.z * dw1));
                    R d_od = FN(safe_clip)(d_od_state + d_od_int, MIN_VAL_R, MAX_VAL_R);
                    R od_raw = density * dist;
                    int active = (od_raw >= (R)1e-8f) && (od_raw <= (R)88.0);
                    R d_od_raw = active ? d_od : (R)0.0;
                    R d_dens_l = FN(safe_clip)(d_od_raw * dist, MIN_VAL_R, MAX_VAL_R);
                    R d_dist = FN(safe_clip)(d_od_raw * density, MIN_VAL_R, MAX_VAL_R);
                    R d_tmin = -d_dist, d_tmax = d_dist;
                    V3 dcs = FN(mk)(csr.x > 0 ? dcs_i.x : 0, csr.y > 0 ? dcs_i.y : 0, csr.z > 0 ? dcs_i.z : 0);
                    V3 dce = FN(mk)(cer.x > 0 ? dce_i.x : 0, cer.y > 0 ? dce_i.y : 0, cer.z > 0 ? dce_i.z : 0);
                    V3 dbc = FN(add)(dcs, dce);
                    R sum_s = (dcs.x + dcs.y) + dcs.z, sum_e = (dce.x + dce.y) + dce.z;
                    R d_dcdt = sum_s * t_min + sum_e * t_max;
                    d_tmin = d_tmin + sum_s * dc_dt;
                    d_tmax = d_tmax + sum_e * dc_dt;
                    R dbo = (dbc.x + dbc.y) + dbc.z;
                    V3 dg_l = FN(scl)(FN(sub)(cam, v[0]), dbo);
                    V3 dv0b = FN(scl)(FN(neg)(g), dbo);
                    dg_l = FN(add)(dg_l, FN(scl)(rd, d_dcdt));
                    acc_c = FN(add)(acc_c, FN(safe_clip3)(dbc, MIN_VAL_R, MAX_VAL_R));
                    V3 dvl[4] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                    for (int side = 0; side < 2; ++side) {
                        const int* f = FN(FACES)[side == 0 ? min_face : max_face];
                        R tt = side == 0 ? t_min : t_max;
                        R dt = side == 0 ? d_tmin : d_tmax;
                        V3 va = v[f[0]], vb = v[f[1]], vc = v[f[2]];
                        V3 n = FN(cross)(FN(sub)(vc, va), FN(sub)(vb, va));
                        R den = FN(dot)(n, rd);
                        R inv_den = FN(safe_div)((R)1.0, den);
                        V3 hit = FN(add)(cam, FN(scl)(rd, tt));
                        V3 ah = FN(sub)(va, hit);
                        V3 dta = FN(scl)(FN(add)(FN(cross)(ah, FN(sub)(vb, vc)), n), inv_den);
                        V3 dtb = FN(scl)(FN(cross)(ah, FN(sub)(vc, va)), inv_den);
                        V3 dtc = FN(scl)(FN(cross)(ah, FN(sub)(va, vb)), inv_den);
                        dvl[f[0]] = FN(add)(dvl[f[0]], FN(safe_clip3)(FN(scl)(dta, dt), MIN_VAL_R, MAX_VAL_R));
                        dvl[f[1]] = FN(add)(dvl[f[1]], FN(safe_clip3)(FN(scl)(dtb, dt), MIN_VAL_R, MAX_VAL_R));
                        dvl[f[2]] = FN(add)(dvl[f[2]], FN(safe_clip3)(FN(scl)(dtc, dt), MIN_VAL_R, MAX_VAL_R));
                    }
                    dvl[0] = FN(add)(dvl[0], dv0b);
                    acc_d = acc_d + FN(safe_clip)(d_dens_l, MIN_VAL_R, MAX_VAL_R);
                    for (int k = 0; k < 4; ++k) acc_v[k] = FN(add)(acc_v[k], FN(safe_clip3)(dvl[k], MIN_VAL_R, MAX_VAL_R));
                    acc_g = FN(add)(acc_g, FN(safe_clip3)(dg_l, MIN_VAL_R, MAX_VAL_R));
                }
            }
            R* o = acc + (size_t)(cursor - rs) * BWD_TERMS;
            o[0] = acc_d;
            for (int k = 0; k < 4; ++k) { o[1 + 3 * k] = acc_v[k].x; o[2 + 3 * k] = acc_v[k].y; o[3 + 3 * k] = acc_v[k].z; }
            o[13] = acc_g.x; o[14] = acc_g.y; o[15] = acc_g.z;
            o[16] = acc_c.x; o[17] = acc_c.y; o[18] = acc_c.z;
        }
    }
}

/* Tiles run in parallel in chunks; each tile's per-tet sums are then added
 * into the gradients serially in tile and list order, so the result is
 * that of the one-thread loop bit for bit, whatever the thread count. */
void FN(orc_backward)(const R* verts, const uint32_t* idx, const R* colors, const R* dens, const R* grads,
                      const uint32_t* sorted_vals, const uint32_t* ranges,
                      const float* cam_pos, const float* vp, const float* c2w_intr,
                      uint32_t w, uint32_t h, uint32_t ts,
                      const R* rgba, const R* dl, uint32_t tile_begin, uint32_t tile_end, uint32_t tile_step,
                      R* d_verts, R* d_dens, R* d_grads, R* d_colors,
                      const R* ck, const R* fin_logt, const uint32_t* fin_stop) {
    CAM c;
    FN(cam_from_flat)(&c, cam_pos, vp, c2w_intr, w, h, ts, 0.0f);
    const uint32_t TS = ts, step = tile_step ? tile_step : 1;
    const int64_t ntiles = tile_end > tile_begin ? ((int64_t)tile_end - tile_begin + step - 1) / step : 0;
    const uint64_t kChunk = 1u << 20; /* per-tet records held at once */
    uint64_t* off = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(ntiles + 1));
    for (int64_t k0 = 0; k0 < ntiles;) {
        int64_t k1 = k0;
        off[0] = 0;
        while (k1 < ntiles && (k1 == k0 || off[k1 - k0] < kChunk)) {
            const uint32_t tile = tile_begin + (uint32_t)k1 * step;
            const uint32_t rs = ranges[2 * tile], re = ranges[2 * tile + 1];
            off[k1 - k0 + 1] = off[k1 - k0] + (re > rs ? re - rs : 0);
            ++k1;
        }
        R* acc = (R*)malloc(sizeof(R) * BWD_TERMS * (size_t)(off[k1 - k0] + 1));
#pragma omp parallel if (k1 - k0 >= 32)
        {
            R* st = (R*)malloc(sizeof(R) * TS * TS * 5);
            int64_t* cur = (int64_t*)malloc(sizeof(int64_t) * TS * TS);
#pragma omp for schedule(dynamic, 1)
            for (int64_t k = k0; k < k1; ++k)
                FN(bwd_tile)(verts, idx, colors, dens, grads, sorted_vals, ranges, &c, w, h, TS,
                             tile_begin + (uint32_t)k * step, rgba, dl, ck, fin_logt, fin_stop, st, cur,
                             acc + (size_t)off[k - k0] * BWD_TERMS);
            free(st);
            free(cur);
        }
        for (int64_t k = k0; k < k1; ++k) {
            const uint32_t tile = tile_begin + (uint32_t)k * step;
            const uint32_t rs = ranges[2 * tile], re = ranges[2 * tile + 1];
            for (uint32_t cursor = rs; cursor < re; ++cursor) {
                const uint32_t t = sorted_vals[cursor];
                const R* o = acc + (size_t)(off[k - k0] + (cursor - rs)) * BWD_TERMS;
                d_dens[t] += o[0];
                for (int i = 0; i < 4; ++i) {
                    const uint32_t vi = idx[4 * t + i];
                    d_verts[3 * vi] += o[1 + 3 * i]; d_verts[3 * vi + 1] += o[2 + 3 * i]; d_verts[3 * vi + 2] += o[3 + 3 * i];
                }
                d_grads[3 * t] += o[13]; d_grads[3 * t + 1] += o[14]; d_grads[3 * t + 2] += o[15];
                d_colors[3 * t] += o[16]; d_colors[3 * t + 1] += o[17]; d_colors[3 * t + 2] += o[18];
            }
        }
        free(acc);
        k0 = k1;
    }
    free(off);
}

/* ------------------------------------------------------------------ */
/* SH colour evaluation — crates/rmesh-util/src/wgsl/sh_eval_compute.wgsl:22-117 */
/* sh: [M][3][nc] fp32, nc = (deg+1)^2.  Colours = softplus(Σ Y_k sh + 0.5 + g·(v0−centroid)). */
/* ------------------------------------------------------------------ */
#define SH_C0 ((R)0.28209479177387814)
#define SH_C1 ((R)0.4886025119029199)
#define SH_C2_0 ((R)1.0925484305920792)
#define SH_C2_1 ((R)-1.0925484305920792)
#define SH_C2_2 ((R)0.31539156525252005)
#define SH_C2_3 ((R)-1.0925484305920792)
#define SH_C2_4 ((R)0.5462742152960396)
#define SH_C3_0 ((R)-0.5900435899266435)
#define SH_C3_1 ((R)2.890611442640554)
#define SH_C3_2 ((R)-0.4570457994644658)
#define SH_C3_3 ((R)0.3731763325901154)
#define SH_C3_4 ((R)-0.4570457994644658)
#define SH_C3_5 ((R)1.445305721320277)
#define SH_C3_6 ((R)-0.5900435899266435)

/* basis values Y_k(dir) with the signs folded in exactly as the shader applies them */
static void FN(sh_basis)(V3 d, uint32_t deg, R* Y) {
    R x = d.x, y = d.y, z = d.z;
    Y[0] = SH_C0;
    if (deg >= 1) { Y[1] = -(SH_C1 * y); Y[2] = SH_C1 * z; Y[3] = -(SH_C1 * x); }
    if (deg >= 2) {
        R xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
        Y[4] = SH_C2_0 * xy; Y[5] = SH_C2_1 * yz; Y[6] = SH_C2_2 * (((R)2.0 * zz - xx) - yy);
        Y[7] = SH_C2_3 * xz; Y[8] = SH_C2_4 * (xx - yy);
    }
    if (deg >= 3) {
        R xx = x * x, yy = y * y, zz = z * z;
        Y[9] = (SH_C3_0 * y) * ((R)3.0 * xx - yy);
        Y[10] = ((SH_C3_1 * x) * y) * z;
        Y[11] = (SH_C3_2 * y) * (((R)4.0 * zz - xx) - yy);
        Y[12] = (SH_C3_3 * z) * (((R)2.0 * zz - (R)3.0 * xx) - (R)3.0 * yy);
        Y[13] = (SH_C3_4 * x) * (((R)4.0 * zz - xx) - yy);
        Y[14] = (SH_C3_5 * z) * (xx - yy);
        Y[15] = (SH_C3_6 * x) * (xx - (R)3.0 * yy);
    }
}

/* d Y_k / d dir (3 components each) */
static void FN(sh_basis_grad)(V3 d, uint32_t deg, R dY[16][3]) {
    R x = d.x, y = d.y, z = d.z;
    for (int k = 0; k < 16; ++k) dY[k][0] = dY[k][1] = dY[k][2] = 0;
    if (deg >= 1) { dY[1][1] = -SH_C1; dY[2][2] = SH_C1; dY[3][0] = -SH_C1; }
    if (deg >= 2) {
        dY[4][0] = SH_C2_0 * y; dY[4][1] = SH_C2_0 * x;
        dY[5][1] = SH_C2_1 * z; dY[5][2] = SH_C2_1 * y;
        dY[6][0] = SH_C2_2 * (-2 * x); dY[6][1] = SH_C2_2 * (-2 * y); dY[6][2] = SH_C2_2 * (4 * z);
        dY[7][0] = SH_C2_3 * z; dY[7][2] = SH_C2_3 * x;
        dY[8][0] = SH_C2_4 * (2 * x); dY[8][1] = SH_C2_4 * (-2 * y);
    }
    if (deg >= 3) {
        R xx = x * x, yy = y * y, zz = z * z;
        dY[9][0] = SH_C3_0 * (6 * x * y); dY[9][1] = SH_C3_0 * (3 * xx - 3 * yy);
        dY[10][0] = SH_C3_1 * y * z; dY[10][1] = SH_C3_1 * x * z; dY[10][2] = SH_C3_1 * x * y;
        dY[11][0] = SH_C3_2 * (-2 * x * y); dY[11][1] = SH_C3_2 * (4 * zz - xx - 3 * yy); dY[11][2] = SH_C3_2 * (8 * y * z);
        dY[12][0] = SH_C3_3 * (-6 * x * z); dY[12][1] = SH_C3_3 * (-6 * y * z); dY[12][2] = SH_C3_3 * (6 * zz - 3 * xx - 3 * yy);
        dY[13][0] = SH_C3_4 * (4 * zz - 3 * xx - yy); dY[13][1] = SH_C3_4 * (-2 * x * y); dY[13][2] = SH_C3_4 * (8 * x * z);
        dY[14][0] = SH_C3_5 * (2 * x * z); dY[14][1] = SH_C3_5 * (-2 * y * z); dY[14][2] = SH_C3_5 * (xx - yy);
        dY[15][0] = SH_C3_6 * (3 * xx - 3 * yy); dY[15][1] = SH_C3_6 * (-6 * x * y);
    }
}

static void FN(sh_dir)(const R* verts, const uint32_t* idx, uint32_t t, const float* cam_pos,
                       V3 v[4], V3* centroid, V3* raw, R* len, V3* dir) {
    for (int i = 0; i < 4; ++i) v[i] = FN(load3)(verts, idx[4 * t + i]);
    *centroid = FN(scl)(FN(add)(FN(add)(FN(add)(v[0], v[1]), v[2]), v[3]), (R)0.25);
    *raw = FN(sub)(*centroid, FN(mk)((R)cam_pos[0], (R)cam_pos[1], (R)cam_pos[2]));
    *len = RSQRT(FN(dot)(*raw, *raw));
    *dir = FN(mk)(0, 0, 0);
    if (*len > 0) *dir = FN(mk)(raw->x / *len, raw->y / *len, raw->z / *len);
}

void FN(orc_sh_colors)(const R* verts, const uint32_t* idx, const R* sh, const R* grads, uint32_t M,
                       uint32_t deg, const float* cam_pos, R* colors) {
    uint32_t nc = (deg + 1) * (deg + 1);
#pragma omp parallel for schedule(static) if (M >= 4096)
    for (int64_t ti = 0; ti < (int64_t)M; ++ti) {
        const uint32_t t = (uint32_t)ti;
        V3 v[4], centroid, raw, dir;
        R len;
        FN(sh_dir)(verts, idx, t, cam_pos, v, &centroid, &raw, &len, &dir);
        R Y[16];
        FN(sh_basis)(dir, deg, Y);
        V3 g = FN(load3)(grads, t);
        R offset = FN(dot)(g, FN(sub)(v[0], centroid));
        for (int ch = 0; ch < 3; ++ch) {
            const R* s = sh + ((size_t)t * 3 + ch) * nc;
            R val = Y[0] * s[0];
            for (uint32_t k = 1; k < nc; ++k) val = val + Y[k] * s[k];
            colors[3 * t + ch] = FN(softplus)((val + (R)0.5) + offset);
        }
    }
}

/* Adjoint of orc_sh_colors (new: the reference has no SH adjoint).
 * Accumulates into d_sh [M*3*nc], d_grads [M*3], d_verts [N*3]. */
void FN(orc_sh_backward)(const R* verts, const uint32_t* idx, const R* sh, const R* grads, uint32_t M,
                         uint32_t deg, const float* cam_pos, const R* d_colors,
                         R* d_sh, R* d_grads, R* d_verts) {
    uint32_t nc = (deg + 1) * (deg + 1);
    /* per tet in parallel; the vertex terms are kept per tet and added in tet
     * order afterwards (the one-thread result, bit for bit) */
    V3* dvt = (V3*)malloc(sizeof(V3) * 2 * (size_t)(M ? M : 1));
#pragma omp parallel for schedule(static) if (M >= 4096)
    for (int64_t ti = 0; ti < (int64_t)M; ++ti) {
        const uint32_t t = (uint32_t)ti;
        V3 v[4], centroid, raw, dir;
        R len;
        FN(sh_dir)(verts, idx, t, cam_pos, v, &centroid, &raw, &len, &dir);
        R Y[16], dY[16][3];
        FN(sh_basis)(dir, deg, Y);
        FN(sh_basis_grad)(dir, deg, dY);
        V3 g = FN(load3)(grads, t);
        R offset = FN(dot)(g, FN(sub)(v[0], centroid));
        R s_sum = 0;
        V3 d_dir = FN(mk)(0, 0, 0);
        for (int ch = 0; ch < 3; ++ch) {
            const R* s = sh + ((size_t)t * 3 + ch) * nc;
            R val = Y[0] * s[0];
            for (uint32_t k = 1; k < nc; ++k) val = val + Y[k] * s[k];
            R pre = (val + (R)0.5) + offset;
            R d_pre = d_colors[3 * t + ch] * FN(dsoftplus)(pre);
            s_sum += d_pre;
            R* ds = d_sh + ((size_t)t * 3 + ch) * nc;
            for (uint32_t k = 0; k < nc; ++k) {
                ds[k] += d_pre * Y[k];
                d_dir.x += d_pre * s[k] * dY[k][0];
                d_dir.y += d_pre * s[k] * dY[k][1];
                d_dir.z += d_pre * s[k] * dY[k][2];
            }
        }
        V3 off = FN(sub)(v[0], centroid);
        d_grads[3 * t] += s_sum * off.x; d_grads[3 * t + 1] += s_sum * off.y; d_grads[3 * t + 2] += s_sum * off.z;
        /* offset = g·(v0 − centroid): d v0 += 0.75 s g, d v_{1,2,3} −= 0.25 s g */
        V3 dcen = FN(scl)(g, -s_sum);
        /* dir = raw/len: d raw = (d_dir − dir (dir·d_dir)) / len */
        if (len > 0) {
            R dd = FN(dot)(dir, d_dir);
            V3 draw = FN(scl)(FN(sub)(d_dir, FN(scl)(dir, dd)), (R)1.0 / len);
            dcen = FN(add)(dcen, draw);
        }
        dvt[2 * t] = FN(scl)(dcen, (R)0.25);
        dvt[2 * t + 1] = FN(scl)(g, s_sum);
    }
    for (uint32_t t = 0; t < M; ++t) {
        for (int i = 0; i < 4; ++i) {
            uint32_t vi = idx[4 * t + i];
            V3 d = dvt[2 * t];
            if (i == 0) d = FN(add)(d, dvt[2 * t + 1]);
            d_verts[3 * vi] += d.x; d_verts[3 * vi + 1] += d.y; d_verts[3 * vi + 2] += d.z;
        }
    }
    free(dvt);
}

#undef V3
#undef CAM
#undef FN
#undef CAT
#undef CAT2
