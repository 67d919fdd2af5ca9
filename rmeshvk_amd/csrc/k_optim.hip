// Spherical-harmonics colours and their adjoint, fused multi-group Adam and
// the per-view loss reduction.
//
// Reference:
//   crates/rmesh-util/src/wgsl/sh_eval_compute.wgsl:22-117 -> k_sh_colors (fp32 SH,
//       channel-major [M][3][nc], colour = softplus(sum Y_k sh_k + 0.5 + g.(v0 - centroid)))
//   crates/rmesh-train/src/wgsl/adam_compute.wgsl:49-77     -> adam_update
// The SH adjoint has no counterpart in the reference (it never differentiates
// SH); it is derived from the forward above and checked by finite differences.
//
// MI355X design: the SH block of a tet is read once per step (not once per
// view): k_sh_colors loops over all views of the batch in one thread, and the
// SH adjoint (four lanes per tet, each over a quarter of the views)
// accumulates d_sh in registers and applies Adam to the tet's coefficients in
// the same pass, so d_sh never touches HBM in a training step.  The other
// parameter groups are updated by one launch.
#include "rmesh_common.hpp"
#include "rmesh_kernels.hpp"

namespace rmesh {

#define SH_C0 0.28209479177387814f
#define SH_C1 0.4886025119029199f
#define SH_C2_0 1.0925484305920792f
#define SH_C2_1 -1.0925484305920792f
#define SH_C2_2 0.31539156525252005f
#define SH_C2_3 -1.0925484305920792f
#define SH_C2_4 0.5462742152960396f
#define SH_C3_0 -0.5900435899266435f
#define SH_C3_1 2.890611442640554f
#define SH_C3_2 -0.4570457994644658f
#define SH_C3_3 0.3731763325901154f
#define SH_C3_4 -0.4570457994644658f
#define SH_C3_5 1.445305721320277f
#define SH_C3_6 -0.5900435899266435f

// Basis with the shader's signs folded in (sh_eval_compute.wgsl:66-103).
template <int DEG>
__device__ __forceinline__ void sh_basis(V3 d, float* Y) {
    float x = d.x, y = d.y, z = d.z;
    Y[0] = SH_C0;
    if (DEG >= 1) {
        Y[1] = -(SH_C1 * y);
        Y[2] = SH_C1 * z;
        Y[3] = -(SH_C1 * x);
    }
    if (DEG >= 2) {
        float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
        Y[4] = SH_C2_0 * xy;
        Y[5] = SH_C2_1 * yz;
        Y[6] = SH_C2_2 * ((2.0f * zz - xx) - yy);
        Y[7] = SH_C2_3 * xz;
        Y[8] = SH_C2_4 * (xx - yy);
    }
    if (DEG >= 3) {
        float xx = x * x, yy = y * y, zz = z * z;
        Y[9] = (SH_C3_0 * y) * (3.0f * xx - yy);
        Y[10] = ((SH_C3_1 * x) * y) * z;
        Y[11] = (SH_C3_2 * y) * ((4.0f * zz - xx) - yy);
        Y[12] = (SH_C3_3 * z) * ((2.0f * zz - 3.0f * xx) - 3.0f * yy);
        Y[13] = (SH_C3_4 * x) * ((4.0f * zz - xx) - yy);
        Y[14] = (SH_C3_5 * z) * (xx - yy);
        Y[15] = (SH_C3_6 * x) * (xx - 3.0f * yy);
    }
}

// d/d(dir) of sum_k w_k Y_k(dir), i.e. the basis Jacobian contracted with w.
template <int DEG>
__device__ __forceinline__ V3 sh_basis_vjp(V3 d, const float* w) {
    float x = d.x, y = d.y, z = d.z;
    float gx = 0.f, gy = 0.f, gz = 0.f;
    if (DEG >= 1) {
        gy += -SH_C1 * w[1];
        gz += SH_C1 * w[2];
        gx += -SH_C1 * w[3];
    }
    if (DEG >= 2) {
        gx += SH_C2_0 * y * w[4];
        gy += SH_C2_0 * x * w[4];
        gy += SH_C2_1 * z * w[5];
        gz += SH_C2_1 * y * w[5];
        gx += SH_C2_2 * (-2.f * x) * w[6];
        gy += SH_C2_2 * (-2.f * y) * w[6];
        gz += SH_C2_2 * (4.f * z) * w[6];
        gx += SH_C2_3 * z * w[7];
        gz += SH_C2_3 * x * w[7];
        gx += SH_C2_4 * (2.f * x) * w[8];
        gy += SH_C2_4 * (-2.f * y) * w[8];
    }
    if (DEG >= 3) {
        float xx = x * x, yy = y * y, zz = z * z;
        gx += SH_C3_0 * (6.f * x * y) * w[9];
        gy += SH_C3_0 * (3.f * xx - 3.f * yy) * w[9];
        gx += SH_C3_1 * y * z * w[10];
        gy += SH_C3_1 * x * z * w[10];
        gz += SH_C3_1 * x * y * w[10];
        gx += SH_C3_2 * (-2.f * x * y) * w[11];
        gy += SH_C3_2 * (4.f * zz - xx - 3.f * yy) * w[11];
        gz += SH_C3_2 * (8.f * y * z) * w[11];
        gx += SH_C3_3 * (-6.f * x * z) * w[12];
        gy += SH_C3_3 * (-6.f * y * z) * w[12];
        gz += SH_C3_3 * (6.f * zz - 3.f * xx - 3.f * yy) * w[12];
        gx += SH_C3_4 * (4.f * zz - 3.f * xx - yy) * w[13];
        gy += SH_C3_4 * (-2.f * x * y) * w[13];
        gz += SH_C3_4 * (8.f * x * z) * w[13];
        gx += SH_C3_5 * (2.f * x * z) * w[14];
        gy += SH_C3_5 * (-2.f * y * z) * w[14];
        gz += SH_C3_5 * (xx - yy) * w[14];
        gx += SH_C3_6 * (3.f * xx - 3.f * yy) * w[15];
        gy += SH_C3_6 * (-6.f * x * y) * w[15];
    }
    return mk3(gx, gy, gz);
}

__device__ __forceinline__ void tet_geom(const float4* __restrict__ rec, uint32_t t, V3 v[4], V3& g, V3& centroid) {
    float4 r0 = rec[4ull * t], r1 = rec[4ull * t + 1], r2 = rec[4ull * t + 2], r3 = rec[4ull * t + 3];
    v[0] = mk3(r0.x, r0.y, r0.z);
    v[1] = mk3(r1.x, r1.y, r1.z);
    v[2] = mk3(r2.x, r2.y, r2.z);
    v[3] = mk3(r3.x, r3.y, r3.z);
    g = mk3(r1.w, r2.w, r3.w);
    centroid = scl3(add3(add3(add3(v[0], v[1]), v[2]), v[3]), 0.25f);
}

__device__ __forceinline__ V3 view_dir(V3 centroid, const ViewParams& vw, float& len) {
    V3 raw = sub3(centroid, mk3(vw.cam[0], vw.cam[1], vw.cam[2]));
    len = sqrtf(dot3(raw, raw));
    V3 d = mk3(0.f, 0.f, 0.f);
    if (len > 0.0f) d = mk3(raw.x / len, raw.y / len, raw.z / len);
    return d;
}

// SH colours of the tets [t0, t0 + n) (one SH degree; their coefficients
// start at sh) for every view of each tet's scene.
template <int DEG>
__global__ __launch_bounds__(256) void k_sh_colors(const float4* __restrict__ rec, const float* __restrict__ sh, uint32_t t0,
                                                   uint32_t n, const SceneRange* __restrict__ scenes, uint32_t n_scenes,
                                                   const ViewParams* __restrict__ views, float4* __restrict__ colors,
                                                   bool half) {
    constexpr int NC = (DEG + 1) * (DEG + 1);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t t = t0 + i;
    const SceneRange& sc = scenes[scene_of_tet(scenes, n_scenes, t)];
    V3 v[4], g, cen;
    tet_geom(rec, t, v, g, cen);
    float c[3 * NC];
    const float* s = sh + (uint64_t)i * 3 * NC;
#pragma unroll
    for (int i = 0; i < 3 * NC; ++i) c[i] = s[i];
    // the viewer path reads the coefficients as packed f16 (project_compute.wgsl:71-75,
    // 267-318): the same values, rounded to nearest even
    if (half) {
#pragma unroll
        for (int i = 0; i < 3 * NC; ++i) c[i] = (float)(_Float16)c[i];
    }
    float offset = dot3(g, sub3(v[0], cen));
    for (uint32_t vv = sc.view_base; vv < sc.view_base + sc.V; ++vv) {
        float len;
        V3 d = view_dir(cen, views[vv], len);
        float Y[NC];
        sh_basis<DEG>(d, Y);
        float rgb[3];
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            float val = Y[0] * c[ch * NC];
#pragma unroll
            for (int k = 1; k < NC; ++k) val = val + Y[k] * c[ch * NC + k];
            rgb[ch] = softplus_f((val + 0.5f) + offset);
        }
        colors[views[vv].color_off + t] = make_float4(rgb[0], rgb[1], rgb[2], 0.0f);
    }
}

__device__ __forceinline__ float adam_update(float& p, float gr, float& m, float& v, float lr, float b1, float b2,
                                             float eps, float b1p, float b2p) {
    // adam_compute.wgsl:49-77
    float gg = safe_clip(gr != gr ? 0.0f : gr, kMinVal, kMaxVal);
    float mn = b1 * m + (1.0f - b1) * gg;
    m = mn;
    float vn = b2 * v + ((1.0f - b2) * gg) * gg;
    v = vn;
    float mh = safe_div(mn, 1.0f - b1p);
    float vh = safe_div(vn, 1.0f - b2p);
    float sq = vh < kTiny ? 0.0f : sqrtf(vh);
    p = p - lr * safe_div(mh, sq + eps);
    return p;
}

// SH adjoint (+ optional fused Adam on the SH coefficients): four lanes per
// tet, 32 tets per 128-lane workgroup; lane j of a tet takes views j, j + 4,
// ... for all three channels, so the view direction, the basis and the
// direction adjoint (one vjp of the channel-summed weights: the vjp is linear)
// are formed once per (tet, view).  The workgroup's coefficients (and Adam's
// m, v) are one contiguous block in HBM; the coefficients are moved to LDS
// with coalesced dword accesses; the four lanes' partial d_sh are summed with
// two quad_perm butterflies, staged in LDS rows (padded, see SR) and applied
// (Adam, m and v read and written in HBM directly) or stored element-parallel.
// The direction/offset adjoint is added to the tet's staging record of vertex
// slots and colour gradient (one coalesced float4 per lane), so it reaches the
// vertices through the same fixed-order gather as the compositing adjoint's
// terms: this kernel runs after that adjoint and before the gather.
constexpr int kShTets = 32;
constexpr int kShLanes = 4;

template <int CTRL>
__device__ __forceinline__ float sh_qperm(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
// Sum over the quad, left in every lane: quad_perm [1,0,3,2] then [2,3,0,1].
__device__ __forceinline__ float quad_sum(float v) {
    v = v + sh_qperm<0xB1>(v);
    return v + sh_qperm<0x4E>(v);
}

template <int DEG, bool ADAM>
__global__ __launch_bounds__(kShLanes * kShTets) void k_sh_backward(ShBwdArgs a) {
    constexpr int NC = (DEG + 1) * (DEG + 1);
    // LDS row stride: whole float4s, and 16 tets' rows (3 SR floats apart) on
    // distinct banks for ds_read_b128
    constexpr int SR = (NC + 3) / 4 * 4 + 4;
    constexpr int NT = kShLanes * kShTets;
    // only the coefficients are staged: Adam's m, v are read from HBM in the
    // element-parallel update, which keeps LDS (15 KB) off the occupancy limit
    __shared__ __align__(16) float s_io[1][3 * kShTets * SR];
    __shared__ float s_g[3 * kShTets * SR];  // summed d_sh rows
    const uint32_t lt = threadIdx.x / kShLanes, j = threadIdx.x % kShLanes;
    const uint32_t i0 = blockIdx.x * kShTets;  // first tet of the workgroup within the class
    const uint32_t t = a.t0 + i0 + lt;         // global tet id
    const uint32_t nflt = min((uint32_t)kShTets, a.M - i0) * 3u * NC;  // floats of this workgroup
    float* io_g[3] = {a.sh + (uint64_t)i0 * 3 * NC, ADAM ? a.m_sh + (uint64_t)i0 * 3 * NC : nullptr,
                      ADAM ? a.v_sh + (uint64_t)i0 * 3 * NC : nullptr};
    for (uint32_t i = threadIdx.x; i < nflt; i += NT) s_io[0][(i / NC) * SR + i % NC] = io_g[0][i];
    __syncthreads();
    const bool live = i0 + lt < a.M;
    const int row0 = (int)(lt * 3) * SR;
    V3 v[4], g = mk3(0.f, 0.f, 0.f), cen, off = mk3(0.f, 0.f, 0.f);
    float dsh[3][NC];
    float s_all = 0.0f;
    V3 dcen = mk3(0.f, 0.f, 0.f);
#pragma unroll
    for (int ch = 0; ch < 3; ++ch)
#pragma unroll
        for (int k = 0; k < NC; ++k) dsh[ch][k] = 0.0f;
    if (live) {
        tet_geom(a.rec, t, v, g, cen);
        off = sub3(v[0], cen);
        const float offset = dot3(g, off);
        const SceneRange& sc = a.scenes[scene_of_tet(a.scenes, a.n_scenes, t)];
        const uint32_t vb = sc.view_base, V = sc.V;
        // colour gradients one view ahead (their load latency overlaps a view's work)
        float dn[3] = {0.f, 0.f, 0.f};
        if (j < V) {
            const float* dcp = a.d_colors + ((uint64_t)(a.views[vb + j].color_off + t)) * 3;
            dn[0] = dcp[0];
            dn[1] = dcp[1];
            dn[2] = dcp[2];
        }
        for (uint32_t vv = j; vv < V; vv += kShLanes) {
            const float dc[3] = {dn[0], dn[1], dn[2]};
            if (vv + kShLanes < V) {
                const float* dcp = a.d_colors + ((uint64_t)(a.views[vb + vv + kShLanes].color_off + t)) * 3;
                dn[0] = dcp[0];
                dn[1] = dcp[1];
                dn[2] = dcp[2];
            }
            if (dc[0] == 0.0f && dc[1] == 0.0f && dc[2] == 0.0f) continue;
            float len;
            const V3 d = view_dir(cen, a.views[vb + vv], len);
            float Y[NC], w[NC];
            sh_basis<DEG>(d, Y);
#pragma unroll
            for (int k = 0; k < NC; ++k) w[k] = 0.0f;
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) {
                if (dc[ch] == 0.0f) continue;
                float c[NC];  // the channel's coefficients, from the LDS row
                if constexpr (NC % 4 == 0) {
                    const float4* cr = reinterpret_cast<const float4*>(&s_io[0][row0 + ch * SR]);
#pragma unroll
                    for (int q = 0; q < NC / 4; ++q) {
                        const float4 x = cr[q];
                        c[4 * q] = x.x;
                        c[4 * q + 1] = x.y;
                        c[4 * q + 2] = x.z;
                        c[4 * q + 3] = x.w;
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < NC; ++k) c[k] = s_io[0][row0 + ch * SR + k];
                }
                float val = Y[0] * c[0];
#pragma unroll
                for (int k = 1; k < NC; ++k) val = val + Y[k] * c[k];
                const float d_pre = dc[ch] * dsoftplus_f((val + 0.5f) + offset);
                s_all += d_pre;
#pragma unroll
                for (int k = 0; k < NC; ++k) {
                    dsh[ch][k] += d_pre * Y[k];
                    w[k] += d_pre * c[k];
                }
            }
            if (len > 0.0f) {
                const V3 d_dir = sh_basis_vjp<DEG>(d, w);
                const float dd = dot3(d, d_dir);
                dcen = add3(dcen, scl3(sub3(d_dir, scl3(d, dd)), 1.0f / len));
            }
        }
    }
    // sums over the tet's four lanes (every lane of the wave takes part)
    s_all = quad_sum(s_all);
    dcen = mk3(quad_sum(dcen.x), quad_sum(dcen.y), quad_sum(dcen.z));
#pragma unroll
    for (int ch = 0; ch < 3; ++ch)
#pragma unroll
        for (int k = 0; k < NC; ++k) dsh[ch][k] = quad_sum(dsh[ch][k]);
    if (live) {
        // offset = g . (v0 - centroid): its adjoint goes to the colour gradient
        // and (with the centroid's) to the four vertex slots of the tet's
        // staging record, which the vertex gather sums afterwards.  Lane j
        // adds to the record's float4 j: (density, slot 0), (slot 1, slot 2.x),
        // (slot 2.yz, slot 3.xy), (slot 3.z, colour gradient); no atomics.
        const V3 dgr = scl3(off, s_all);
        dcen = add3(dcen, scl3(g, -s_all));
        const V3 dq = scl3(dcen, 0.25f);           // slots 1..3
        const V3 d0 = add3(dq, scl3(g, s_all));   // slot 0 (v0 enters the offset too)
        float4* rp = reinterpret_cast<float4*>(a.tet_g + 16ull * t) + j;
        float4 x = *rp;
        if (j == 0) {
            x.y += d0.x;
            x.z += d0.y;
            x.w += d0.z;
        } else if (j == 1) {
            x.x += dq.x;
            x.y += dq.y;
            x.z += dq.z;
            x.w += dq.x;
        } else if (j == 2) {
            x.x += dq.y;
            x.y += dq.z;
            x.z += dq.x;
            x.w += dq.y;
        } else {
            x.x += dq.z;
            x.y += dgr.x;
            x.z += dgr.y;
            x.w += dgr.z;
        }
        *rp = x;
    }
    // lane j < 3 stages channel j's row
    if (live && j < 3) {
#pragma unroll
        for (int ch = 0; ch < 3; ++ch)
            if (ch == (int)j)
#pragma unroll
                for (int k = 0; k < NC; ++k) s_g[row0 + ch * SR + k] = dsh[ch][k];
    }
    __syncthreads();
    if constexpr (ADAM) {
        if (*a.st.skip) return;  // uniform: no barrier follows on this path
        const float stepf = (float)*a.st.step;
        const float b1p = expf(stepf * logf(a.st.beta1));
        const float b2p = expf(stepf * logf(a.st.beta2));
        for (uint32_t i = threadIdx.x; i < nflt; i += NT) {
            const uint32_t r = (i / NC) * SR + i % NC;
            float p = s_io[0][r], m = io_g[1][i], v2 = io_g[2][i];
            adam_update(p, s_g[r], m, v2, a.lr, a.st.beta1, a.st.beta2, a.st.eps, b1p, b2p);
            io_g[0][i] = p;
            io_g[1][i] = m;
            io_g[2][i] = v2;
        }
    } else {
        float* o = a.d_sh + (uint64_t)i0 * 3 * NC;
        for (uint32_t i = threadIdx.x; i < nflt; i += NT) o[i] = s_g[(i / NC) * SR + i % NC];
    }
}

__global__ __launch_bounds__(256) void k_adam(AdamArgs a) {
    if (*a.st.skip) return;
    float stepf = (float)*a.st.step;
    float b1p = expf(stepf * logf(a.st.beta1));
    float b2p = expf(stepf * logf(a.st.beta2));
    uint64_t total = 0;
    for (int k = 0; k < a.ngroups; ++k) total += a.grp[k].n;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        uint64_t j = i;
        int k = 0;
        while (j >= a.grp[k].n) {
            j -= a.grp[k].n;
            ++k;
        }
        const AdamGroup& G = a.grp[k];
        float p = G.p[j], m = G.m[j], v = G.v[j];
        adam_update(p, G.g[j], m, v, G.lr, a.st.beta1, a.st.beta2, a.st.eps, b1p, b2p);
        G.p[j] = p;
        G.m[j] = m;
        G.v[j] = v;
    }
}

__global__ void k_step_inc(uint32_t* step, const uint32_t* skip) {
    if (threadIdx.x == 0 && !*skip) *step += 1u;
}

// Per-view loss: sum of per-tile partials in tile order (deterministic).
__global__ __launch_bounds__(256) void k_reduce_loss(const float* __restrict__ partial, const ViewParams* __restrict__ views,
                                                     float* __restrict__ out) {
    __shared__ float red[4];
    const ViewParams& vw = views[blockIdx.x];
    const uint32_t nt = vw.tiles_x * vw.tiles_y;
    float s = 0.0f;
    for (uint32_t i = threadIdx.x; i < nt; i += blockDim.x) s += partial[(uint64_t)vw.tile_base + i];
    s = wave_sum(s);
    if (__lane_id() == 0) red[threadIdx.x / 64] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = ((red[0] + red[1]) + (red[2] + red[3])) * (1.0f / (float)(vw.width * vw.height));
}

void launch_sh_colors(const float4* rec, const float* sh, uint32_t deg, uint32_t t0, uint32_t n, const SceneRange* scenes,
                      uint32_t n_scenes, const ViewParams* views, float4* colors, bool half, hipStream_t s) {
    if (!n) return;
    dim3 gr((n + 255) / 256), bl(256);
    switch (deg) {
        case 0: hipLaunchKernelGGL(k_sh_colors<0>, gr, bl, 0, s, rec, sh, t0, n, scenes, n_scenes, views, colors, half); break;
        case 1: hipLaunchKernelGGL(k_sh_colors<1>, gr, bl, 0, s, rec, sh, t0, n, scenes, n_scenes, views, colors, half); break;
        case 2: hipLaunchKernelGGL(k_sh_colors<2>, gr, bl, 0, s, rec, sh, t0, n, scenes, n_scenes, views, colors, half); break;
        default: hipLaunchKernelGGL(k_sh_colors<3>, gr, bl, 0, s, rec, sh, t0, n, scenes, n_scenes, views, colors, half); break;
    }
}

template <int DEG>
static void shb(const ShBwdArgs& a, hipStream_t s) {
    dim3 gr((a.M + kShTets - 1) / kShTets), bl(kShLanes * kShTets);
    if (a.adam) hipLaunchKernelGGL((k_sh_backward<DEG, true>), gr, bl, 0, s, a);
    else hipLaunchKernelGGL((k_sh_backward<DEG, false>), gr, bl, 0, s, a);
}

void launch_sh_backward(const ShBwdArgs& a, hipStream_t s) {
    if (!a.M) return;
    switch (a.deg) {
        case 0: shb<0>(a, s); break;
        case 1: shb<1>(a, s); break;
        case 2: shb<2>(a, s); break;
        default: shb<3>(a, s); break;
    }
}

void launch_adam(const AdamArgs& a, hipStream_t s) {
    uint64_t total = 0;
    for (int k = 0; k < a.ngroups; ++k) total += a.grp[k].n;
    if (!total) return;
    uint64_t blocks = (total + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_adam, dim3((uint32_t)blocks), dim3(256), 0, s, a);
}

void launch_step_inc(uint32_t* step, const uint32_t* skip, hipStream_t s) {
    hipLaunchKernelGGL(k_step_inc, dim3(1), dim3(64), 0, s, step, skip);
}

void launch_reduce_loss(const float* partial, const ViewParams* views, uint32_t V, float* out, hipStream_t s) {
    if (V) hipLaunchKernelGGL(k_reduce_loss, dim3(V), dim3(256), 0, s, partial, views, out);
}

}  // namespace rmesh
