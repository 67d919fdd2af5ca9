/*
 * rmesh oracle — CPU restatement of half-potato/rmeshvk's tiled train step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (rmeshvk_amd/, rmesh_wgpu/)
 * links or calls this library; only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it, as the checker.
 *
 * Built with -O2 -ffp-contract=off (no FMA contraction, IEEE division and
 * sqrt) so the binning stages reproduce the reference's FP32 arithmetic bit
 * for bit.  See oracle_impl.h for the per-kernel functions and citations.
 *
 * Parity pinning (which check fixes each function) is listed in oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static inline uint32_t f32_bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline uint32_t umin32(uint32_t a, uint32_t b) { return a < b ? a : b; }
static inline uint32_t uclamp32(uint32_t v, uint32_t lo, uint32_t hi) { uint32_t m = v > lo ? v : lo; return m < hi ? m : hi; }
static inline int32_t imax32(int32_t a, int32_t b) { return a > b ? a : b; }
static inline int32_t imin32(int32_t a, int32_t b) { return a < b ? a : b; }
/* WGSL f32→u32 / f32→i32: truncate toward zero, saturate, NaN → 0 */
static inline uint32_t sat_u32(double x) {
    if (!(x > 0.0)) return 0u;
    if (x >= 4294967296.0) return 0xFFFFFFFFu;
    return (uint32_t)x;
}
static inline int32_t sat_i32(double x) {
    if (x != x) return 0;
    if (x <= -2147483648.0) return INT32_MIN;
    if (x >= 2147483648.0) return INT32_MAX;
    return (int32_t)x;
}

/* ---- float instantiation ---- */
#define R float
#define SFX _f32
#define RFMIN fminf
#define RFMAX fmaxf
#define RSQRT sqrtf
#define RABS fabsf
#define REXP expf
#define RLOG logf
#define RCEIL ceilf
#define RFLOOR floorf
#include "oracle_impl.h"
#undef R
#undef SFX
#undef RFMIN
#undef RFMAX
#undef RSQRT
#undef RABS
#undef REXP
#undef RLOG
#undef RCEIL
#undef RFLOOR

/* ---- double instantiation ---- */
#define R double
#define SFX _f64
#define RFMIN fmin
#define RFMAX fmax
#define RSQRT sqrt
#define RABS fabs
#define REXP exp
#define RLOG log
#define RCEIL ceil
#define RFLOOR floor
#include "oracle_impl.h"

/* ------------------------------------------------------------------ */
/* Sort: stable LSD radix sort on the low `sorting_bits` of the 64-bit   */
/* key, 8-bit digits.  Semantics of rmesh-sort's DeviceRadixSort        */
/* (crates/rmesh-sort/src/drs.rs:23-26,180-279): bits above              */
/* sorting_bits are ignored, equal masked keys keep input order.        */
/* ------------------------------------------------------------------ */
uint32_t orc_sorting_bits(uint32_t num_tiles) {
    uint32_t tile_bits = 1;
    if (num_tiles > 1) { uint32_t x = num_tiles - 1; tile_bits = 0; while (x) { tile_bits++; x >>= 1; } }
    return (32 + tile_bits + 7) & ~7u;
}

void orc_sort(uint64_t* keys, uint32_t* vals, uint64_t n, uint32_t sorting_bits) {
    if (n == 0) return;
    uint64_t* k2 = (uint64_t*)malloc(n * sizeof(uint64_t));
    uint32_t* v2 = (uint32_t*)malloc(n * sizeof(uint32_t));
    uint64_t* ks = keys; uint32_t* vs = vals; uint64_t* kd = k2; uint32_t* vd = v2;
    for (uint32_t shift = 0; shift < sorting_bits; shift += 8) {
        uint64_t cnt[257];
        memset(cnt, 0, sizeof(cnt));
        for (uint64_t i = 0; i < n; ++i) cnt[((ks[i] >> shift) & 0xFF) + 1]++;
        for (int d = 0; d < 256; ++d) cnt[d + 1] += cnt[d];
        for (uint64_t i = 0; i < n; ++i) {
            uint64_t pos = cnt[(ks[i] >> shift) & 0xFF]++;
            kd[pos] = ks[i]; vd[pos] = vs[i];
        }
        uint64_t* tk = ks; ks = kd; kd = tk;
        uint32_t* tv = vs; vs = vd; vd = tv;
    }
    if (ks != keys) { memcpy(keys, ks, n * sizeof(uint64_t)); memcpy(vals, vs, n * sizeof(uint32_t)); }
    free(k2); free(v2);
}

/* ------------------------------------------------------------------ */
/* K11: per-tile [start, end) — crates/rmesh-tile/src/wgsl/tile_ranges_compute.wgsl:10-45 */
/* `ranges` must be zeroed by the caller (the reference clears it).     */
/* A tile whose pairs are followed only by sentinels keeps end = 0, as  */
/* in the reference.                                                    */
/* ------------------------------------------------------------------ */
void orc_ranges(const uint64_t* keys, uint64_t n, uint32_t* ranges) {
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t tile = (uint32_t)(keys[i] >> 32);
        if (tile == 0xFFFFFFFFu) continue;
        if (i == 0) {
            ranges[2 * (uint64_t)tile] = (uint32_t)i;
        } else {
            uint32_t prev = (uint32_t)(keys[i - 1] >> 32);
            if (tile != prev) {
                ranges[2 * (uint64_t)tile] = (uint32_t)i;
                if (prev != 0xFFFFFFFFu) ranges[2 * (uint64_t)prev + 1] = (uint32_t)i;
            }
        }
        if (i == n - 1) ranges[2 * (uint64_t)tile + 1] = (uint32_t)(i + 1);
    }
}

/* ------------------------------------------------------------------ */
/* K15: Adam — crates/rmesh-train/src/wgsl/adam_compute.wgsl:49-77      */
/* ------------------------------------------------------------------ */
static inline float sclip(float v, float lo, float hi) { return fmaxf(fminf(v, hi), lo); }
static inline float sdiv(float a, float b) {
    if (fabsf(b) < 1.0754944e-20f) return sclip(a / 1.0754944e-20f, -1e20f, 1e20f);
    return sclip(a / b, -1e20f, 1e20f);
}
void orc_adam(float* params, const float* grads, float* m, float* v, uint64_t n,
              uint32_t step, float lr, float beta1, float beta2, float eps) {
    float step_f = (float)step;
    float b1p = expf(step_f * logf(beta1));
    float b2p = expf(step_f * logf(beta2));
    for (uint64_t i = 0; i < n; ++i) {
        float gr = grads[i];
        float g = sclip(gr != gr ? 0.0f : gr, -1e20f, 1e20f);
        float mn = beta1 * m[i] + (1.0f - beta1) * g;
        m[i] = mn;
        float vn = beta2 * v[i] + ((1.0f - beta2) * g) * g;
        v[i] = vn;
        float mh = sdiv(mn, 1.0f - b1p);
        float vh = sdiv(vn, 1.0f - b2p);
        float sq = vh < 1.0754944e-20f ? 0.0f : sqrtf(vh);
        params[i] = params[i] - lr * sdiv(mh, sq + eps);
    }
}

/* Circumsphere (cx, cy, cz, r²) per tet, f64 with the reference's
 * degenerate-denominator rule — crates/rmesh-data/src/lib.rs:601-710,
 * crates/rmesh-util/src/test_util.rs:133-164. */
void orc_circumspheres(const float* verts, const uint32_t* idx, uint64_t M, float* out) {
    for (uint64_t t = 0; t < M; ++t) {
        double p[4][3];
        for (int i = 0; i < 4; ++i)
            for (int k = 0; k < 3; ++k) p[i][k] = verts[3 * (uint64_t)idx[4 * t + i] + k];
        double a[3], b[3], c[3];
        for (int k = 0; k < 3; ++k) { a[k] = p[1][k] - p[0][k]; b[k] = p[2][k] - p[0][k]; c[k] = p[3][k] - p[0][k]; }
        double aa = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
        double bb = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
        double cc = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
        double bc[3] = {b[1] * c[2] - b[2] * c[1], b[2] * c[0] - b[0] * c[2], b[0] * c[1] - b[1] * c[0]};
        double ca[3] = {c[1] * a[2] - c[2] * a[1], c[2] * a[0] - c[0] * a[2], c[0] * a[1] - c[1] * a[0]};
        double ab[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
        double denom = 2.0 * (a[0] * bc[0] + a[1] * bc[1] + a[2] * bc[2]);
        if (fabs(denom) < 1e-12) denom = 1.0;
        double r[3];
        for (int k = 0; k < 3; ++k) r[k] = (aa * bc[k] + bb * ca[k] + cc * ab[k]) / denom;
        out[4 * t + 0] = (float)(p[0][0] + r[0]);
        out[4 * t + 1] = (float)(p[0][1] + r[1]);
        out[4 * t + 2] = (float)(p[0][2] + r[2]);
        out[4 * t + 3] = (float)(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    }
}

/* Host threads used by the tile-parallel loops (0: the OpenMP default). */
#include <omp.h>
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : omp_get_num_procs()); }
int orc_get_threads(void) { return omp_get_max_threads(); }
