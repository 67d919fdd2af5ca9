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
// MI355X design: one thread per tet loops over all views of the batch so the
// 192-byte SH block of a tet is read once per step (not once per view); the
// SH adjoint accumulates d_sh over views in registers and applies Adam to the
// tet's 48 coefficients in the same pass, so d_sh never touches HBM in a
// training step.  The other parameter groups are updated by one launch.
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

template <int DEG>
__global__ __launch_bounds__(256) void k_sh_colors(const float4* __restrict__ rec, const float* __restrict__ sh, uint32_t M,
                                                   const ViewParams* __restrict__ views, uint32_t V, float4* __restrict__ colors) {
    constexpr int NC = (DEG + 1) * (DEG + 1);
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M) return;
    V3 v[4], g, cen;
    tet_geom(rec, t, v, g, cen);
    float c[3 * NC];
    const float* s = sh + (uint64_t)t * 3 * NC;
#pragma unroll
    for (int i = 0; i < 3 * NC; ++i) c[i] = s[i];
    float offset = dot3(g, sub3(v[0], cen));
    for (uint32_t vv = 0; vv < V; ++vv) {
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
        colors[(uint64_t)vv * M + t] = make_float4(rgb[0], rgb[1], rgb[2], 0.0f);
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

// SH adjoint (+ optional fused Adam on the SH coefficients): one lane per
// (tet, colour channel), 64 tets per 192-lane workgroup.  The workgroup's
// coefficients (and Adam's m, v) are one contiguous block in HBM: they are
// moved between HBM and LDS with coalesced dword accesses, and each lane works
// on its channel's row in LDS (rows padded to NC + 1 floats: conflict-free).
// The three channel partials of a tet's direction/offset adjoint are summed in
// a fixed order through LDS before one lane scatters them to the vertices.
constexpr int kShTets = 64;

template <int DEG, bool ADAM>
__global__ __launch_bounds__(3 * kShTets) void k_sh_backward(ShBwdArgs a) {
    constexpr int NC = (DEG + 1) * (DEG + 1);
    constexpr int SR = NC + 1;                      // LDS row stride
    constexpr int NIO = ADAM ? 3 : 1;               // sh (+ m, v)
    __shared__ float s_io[NIO][3 * kShTets * SR];
    __shared__ float red[kShTets][3][4];  // per channel: d(centroid).xyz, d(offset)
    const uint32_t lt = threadIdx.x / 3, ch = threadIdx.x % 3;
    const uint32_t t0 = blockIdx.x * kShTets;
    const uint32_t t = t0 + lt;
    const uint32_t nflt = min((uint32_t)kShTets, a.M - t0) * 3u * NC;  // floats of this workgroup
    float* io_g[3] = {a.sh + (uint64_t)t0 * 3 * NC, ADAM ? a.m_sh + (uint64_t)t0 * 3 * NC : nullptr,
                      ADAM ? a.v_sh + (uint64_t)t0 * 3 * NC : nullptr};
#pragma unroll
    for (int q = 0; q < NIO; ++q)
        for (uint32_t i = threadIdx.x; i < nflt; i += 3 * kShTets) s_io[q][(i / NC) * SR + i % NC] = io_g[q][i];
    __syncthreads();
    const bool live = t < a.M;
    const int row = (int)(lt * 3 + ch) * SR;
    V3 v[4], g, cen, off;
    float c[NC], dsh[NC];
    if (live) {
        tet_geom(a.rec, t, v, g, cen);
        off = sub3(v[0], cen);
        const float offset = dot3(g, off);
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            c[k] = s_io[0][row + k];
            dsh[k] = 0.0f;
        }
        float s_all = 0.0f;
        V3 dcen = mk3(0.f, 0.f, 0.f);
        for (uint32_t vv = 0; vv < a.V; ++vv) {
            const float dc = a.d_colors[((uint64_t)vv * a.M + t) * 3 + ch];
            if (dc == 0.0f) continue;
            float len;
            const V3 d = view_dir(cen, a.views[vv], len);
            float Y[NC];
            sh_basis<DEG>(d, Y);
            float val = Y[0] * c[0];
#pragma unroll
            for (int k = 1; k < NC; ++k) val = val + Y[k] * c[k];
            const float d_pre = dc * dsoftplus_f((val + 0.5f) + offset);
            s_all += d_pre;
            float w[NC];
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                dsh[k] += d_pre * Y[k];
                w[k] = d_pre * c[k];
            }
            if (len > 0.0f) {
                const V3 d_dir = sh_basis_vjp<DEG>(d, w);
                const float dd = dot3(d, d_dir);
                dcen = add3(dcen, scl3(sub3(d_dir, scl3(d, dd)), 1.0f / len));
            }
        }
        red[lt][ch][0] = dcen.x;
        red[lt][ch][1] = dcen.y;
        red[lt][ch][2] = dcen.z;
        red[lt][ch][3] = s_all;
    }
    __syncthreads();
    if (live && ch == 0) {
        const float s_all = (red[lt][0][3] + red[lt][1][3]) + red[lt][2][3];
        V3 dcen = mk3((red[lt][0][0] + red[lt][1][0]) + red[lt][2][0], (red[lt][0][1] + red[lt][1][1]) + red[lt][2][1],
                      (red[lt][0][2] + red[lt][1][2]) + red[lt][2][2]);
        // offset = g . (v0 - centroid)
        a.d_grads[3ull * t] += s_all * off.x;
        a.d_grads[3ull * t + 1] += s_all * off.y;
        a.d_grads[3ull * t + 2] += s_all * off.z;
        dcen = add3(dcen, scl3(g, -s_all));
        const uint4 q = reinterpret_cast<const uint4*>(a.idx)[t];
        const uint32_t vi[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            V3 dvv = scl3(dcen, 0.25f);
            if (i == 0) dvv = add3(dvv, scl3(g, s_all));
            if (dvv.x != 0.0f) unsafeAtomicAdd(a.d_verts + 3ull * vi[i], dvv.x);
            if (dvv.y != 0.0f) unsafeAtomicAdd(a.d_verts + 3ull * vi[i] + 1, dvv.y);
            if (dvv.z != 0.0f) unsafeAtomicAdd(a.d_verts + 3ull * vi[i] + 2, dvv.z);
        }
    }
    if (ADAM) {
        if (*a.st.skip) return;  // uniform: no barrier follows on this path
        if (live) {
            const float stepf = (float)*a.st.step;
            const float b1p = expf(stepf * logf(a.st.beta1));
            const float b2p = expf(stepf * logf(a.st.beta2));
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                float p = c[k], m = s_io[NIO > 1 ? 1 : 0][row + k], v2 = s_io[NIO > 2 ? 2 : 0][row + k];
                adam_update(p, dsh[k], m, v2, a.lr, a.st.beta1, a.st.beta2, a.st.eps, b1p, b2p);
                s_io[0][row + k] = p;
                s_io[NIO > 1 ? 1 : 0][row + k] = m;
                s_io[NIO > 2 ? 2 : 0][row + k] = v2;
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NIO; ++q)
            for (uint32_t i = threadIdx.x; i < nflt; i += 3 * kShTets) io_g[q][i] = s_io[q][(i / NC) * SR + i % NC];
    } else {
        if (live) {
#pragma unroll
            for (int k = 0; k < NC; ++k) s_io[0][row + k] = dsh[k];
        }
        __syncthreads();
        float* o = a.d_sh + (uint64_t)t0 * 3 * NC;
        for (uint32_t i = threadIdx.x; i < nflt; i += 3 * kShTets) o[i] = s_io[0][(i / NC) * SR + i % NC];
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
__global__ __launch_bounds__(256) void k_reduce_loss(const float* __restrict__ partial, uint32_t tiles_per_view,
                                                     float* __restrict__ out, float inv_npx) {
    __shared__ float red[4];
    uint32_t v = blockIdx.x;
    float s = 0.0f;
    for (uint32_t i = threadIdx.x; i < tiles_per_view; i += blockDim.x) s += partial[(uint64_t)v * tiles_per_view + i];
    s = wave_sum(s);
    if (__lane_id() == 0) red[threadIdx.x / 64] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[v] = ((red[0] + red[1]) + (red[2] + red[3])) * inv_npx;
}

void launch_sh_colors(const float4* rec, const float* sh, uint32_t deg, uint32_t M, const ViewParams* views, uint32_t V,
                      float4* colors, hipStream_t s) {
    if (!M) return;
    dim3 gr((M + 255) / 256), bl(256);
    switch (deg) {
        case 0: hipLaunchKernelGGL(k_sh_colors<0>, gr, bl, 0, s, rec, sh, M, views, V, colors); break;
        case 1: hipLaunchKernelGGL(k_sh_colors<1>, gr, bl, 0, s, rec, sh, M, views, V, colors); break;
        case 2: hipLaunchKernelGGL(k_sh_colors<2>, gr, bl, 0, s, rec, sh, M, views, V, colors); break;
        default: hipLaunchKernelGGL(k_sh_colors<3>, gr, bl, 0, s, rec, sh, M, views, V, colors); break;
    }
}

template <int DEG>
static void shb(const ShBwdArgs& a, hipStream_t s) {
    dim3 gr((a.M + kShTets - 1) / kShTets), bl(3 * kShTets);
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

void launch_reduce_loss(const float* partial, uint32_t tiles_per_view, uint32_t V, float* out, float inv_npx,
                        hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_loss, dim3(V), dim3(256), 0, s, partial, tiles_per_view, out, inv_npx);
}

}  // namespace rmesh
