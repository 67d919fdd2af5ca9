// Scene files: .rmesh (gzip or raw) and binary PLY, parsed into the
// renderer's flat arrays (crates/rmesh-data/src/lib.rs:140-394, 597-853).
//
// Host code: a file image is decoded once, in parallel over tets (f16 ->
// f32, the PCA decompression of the SH coefficients, the float64
// circumspheres), into arrays the caller copies out.  Every read is bounds
// checked; a truncated or inconsistent file is an error, not a crash.
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/rmesh_hip.h"
#include "rmesh_error.hpp"

struct rmesh_scene_file {
    rmesh_scene_info info;
    std::vector<float> vertices, densities, color_grads, circumdata, sh;
    std::vector<uint32_t> indices;
};

namespace {

rmesh_status bad(const std::string& msg) {
    rmesh::set_last_error(msg);
    return RMESH_ERR_INVALID;
}

// Runs fn(begin, end) over [0, n) in contiguous blocks on up to 16 threads.
template <class F>
void parallel_for(size_t n, F fn) {
    const size_t hw = std::max<size_t>(1, std::thread::hardware_concurrency());
    const size_t nt = std::min<size_t>({hw, 16, std::max<size_t>(1, n / 4096)});
    if (nt <= 1) {
        fn(size_t(0), n);
        return;
    }
    std::vector<std::thread> th;
    const size_t step = (n + nt - 1) / nt;
    for (size_t t = 0; t < nt; ++t) {
        const size_t b = t * step, e = std::min(n, b + step);
        if (b < e) th.emplace_back([=] { fn(b, e); });
    }
    for (auto& x : th) x.join();
}

// IEEE binary16 -> binary32 (exact: every half is a float).
float half_to_float(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t ex = (h >> 10) & 0x1Fu, man = h & 0x3FFu;
    uint32_t bits;
    if (ex == 0x1Fu) {
        bits = sign | 0x7F800000u | (man << 13);  // inf / NaN
    } else if (ex != 0) {
        bits = sign | ((ex + 112u) << 23) | (man << 13);
    } else if (man == 0) {
        bits = sign;  // +-0
    } else {
        // subnormal half: man * 2^-24, normalised
        int e = -1;
        uint32_t m = man;
        do {
            m <<= 1;
            ++e;
        } while (!(m & 0x400u));
        bits = sign | ((uint32_t)(112 - e) << 23) | ((m & 0x3FFu) << 13);
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

struct Reader {
    const uint8_t* p;
    size_t n, off = 0;
    bool need(size_t bytes) const { return off <= n && bytes <= n - off; }
    uint32_t u32() {
        uint32_t v;
        std::memcpy(&v, p + off, 4);
        off += 4;
        return v;
    }
    float f32() {
        float v;
        std::memcpy(&v, p + off, 4);
        off += 4;
        return v;
    }
    void align4() { off = (off + 3) & ~size_t(3); }
};

void f16_block(const uint8_t* src, float* dst, size_t count) {
    parallel_for(count, [&](size_t b, size_t e) {
        for (size_t i = b; i < e; ++i) {
            uint16_t h;
            std::memcpy(&h, src + 2 * i, 2);
            dst[i] = half_to_float(h);
        }
    });
}

bool gunzip(const uint8_t* data, size_t n, std::vector<uint8_t>& out, std::string& err) {
    z_stream zs;
    std::memset(&zs, 0, sizeof(zs));
    if (inflateInit2(&zs, 16 + MAX_WBITS) != Z_OK) {
        err = "zlib init failed";
        return false;
    }
    out.resize(std::max<size_t>(n * 4, 1 << 16));
    zs.next_in = const_cast<Bytef*>(data);
    size_t in_left = n;
    size_t produced = 0;
    int rc = Z_OK;
    while (rc != Z_STREAM_END) {
        if (produced == out.size()) out.resize(out.size() * 2);
        const uInt in_chunk = (uInt)std::min<size_t>(in_left, 1u << 30);
        const uInt out_chunk = (uInt)std::min<size_t>(out.size() - produced, 1u << 30);
        zs.avail_in = in_chunk;
        zs.next_out = out.data() + produced;
        zs.avail_out = out_chunk;
        rc = inflate(&zs, Z_NO_FLUSH);
        in_left -= in_chunk - zs.avail_in;
        produced += out_chunk - zs.avail_out;
        if (rc != Z_OK && rc != Z_STREAM_END && !(rc == Z_BUF_ERROR && zs.avail_out == 0)) {
            err = std::string("gzip data is corrupt or truncated (zlib: ") + (zs.msg ? zs.msg : "no message") + ")";
            inflateEnd(&zs);
            return false;
        }
        if (rc != Z_STREAM_END && in_left == 0 && zs.avail_out != 0) {
            err = "gzip data is truncated";
            inflateEnd(&zs);
            return false;
        }
    }
    inflateEnd(&zs);
    out.resize(produced);
    return true;
}

// compute_circumspheres_parallel (lib.rs:597-690): float64 circumcentre and
// radius^2; a degenerate tet (|2 a . (b x c)| < 1e-10) takes its centroid with
// the smallest vertex distance (all edges < 1e-7) or r^2 = 1e20.
struct D3 {
    double x, y, z;
};
inline D3 dsub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 dadd(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 dscl(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 dcross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double dlen(D3 a) { return std::sqrt(ddot(a, a)); }

void circumspheres(const float* v, const uint32_t* idx, size_t M, float* out) {
    parallel_for(M, [&](size_t b, size_t e) {
        for (size_t i = b; i < e; ++i) {
            D3 p[4];
            for (int k = 0; k < 4; ++k) {
                const size_t j = idx[4 * i + k];
                p[k] = {(double)v[3 * j], (double)v[3 * j + 1], (double)v[3 * j + 2]};
            }
            const D3 a = dsub(p[1], p[0]), bb_ = dsub(p[2], p[0]), c = dsub(p[3], p[0]);
            const double aa = ddot(a, a), bb = ddot(bb_, bb_), cc = ddot(c, c);
            const D3 cbc = dcross(bb_, c), cca = dcross(c, a), cab = dcross(a, bb_);
            const double denom = 2.0 * ddot(a, cbc);
            float* o = out + 4 * i;
            if (std::fabs(denom) < 1e-10) {
                const double md = std::max(std::max(std::max(dlen(dsub(p[1], p[0])), dlen(dsub(p[2], p[0]))),
                                                    std::max(dlen(dsub(p[3], p[0])), dlen(dsub(p[2], p[1])))),
                                           std::max(dlen(dsub(p[3], p[1])), dlen(dsub(p[3], p[2]))));
                const D3 cen = dscl(dadd(dadd(dadd(p[0], p[1]), p[2]), p[3]), 0.25);
                o[0] = (float)cen.x;
                o[1] = (float)cen.y;
                o[2] = (float)cen.z;
                if (md < 1e-7) {
                    const double r = std::min(std::min(dlen(dsub(p[0], cen)), dlen(dsub(p[1], cen))),
                                              std::min(dlen(dsub(p[2], cen)), dlen(dsub(p[3], cen))));
                    o[3] = (float)(r * r);
                } else {
                    o[3] = 1e20f;
                }
            } else {
                const D3 num = dadd(dadd(dscl(cbc, aa), dscl(cca, bb)), dscl(cab, cc));
                const D3 rel = {num.x / denom, num.y / denom, num.z / denom};
                const D3 cen = dadd(p[0], rel);
                o[0] = (float)cen.x;
                o[1] = (float)cen.y;
                o[2] = (float)cen.z;
                o[3] = (float)ddot(rel, rel);
            }
        }
    });
}

rmesh_status check_indices(const rmesh_scene_file& s) {
    const uint64_t N = s.info.n_vertices;
    for (size_t i = 0; i < s.indices.size(); ++i)
        if (s.indices[i] >= N)
            return bad("tet " + std::to_string(i / 4) + " references vertex " + std::to_string(s.indices[i]) +
                       " of " + std::to_string(N));
    return RMESH_OK;
}

// parse_rmesh (lib.rs:243-394)
rmesh_status parse_rmesh(const uint8_t* data, size_t n, rmesh_scene_file& s) {
    Reader r{data, n};
    if (!r.need(48)) return bad(".rmesh: header truncated");
    const uint32_t N = r.u32(), M = r.u32(), deg = r.u32(), k = r.u32();
    if (deg > 3) return bad(".rmesh: SH degree " + std::to_string(deg) + " (0-3 supported)");
    for (int i = 0; i < 7; ++i) s.info.start_pose[i] = r.f32();
    (void)r.f32();
    s.info.n_vertices = N;
    s.info.n_tets = M;
    s.info.sh_degree = deg;
    s.info.k_components = k;
    const size_t nc = (size_t)(deg + 1) * (deg + 1), dims = 3 * nc;
    const size_t ek = std::min<size_t>(k, dims);  // PCA yields at most `dims` components
    if (!r.need(12ull * N + 16ull * M + M)) return bad(".rmesh: vertex / index / density data truncated");
    s.vertices.resize(3ull * N);
    std::memcpy(s.vertices.data(), data + r.off, 12ull * N);
    r.off += 12ull * N;
    s.indices.resize(4ull * M);
    std::memcpy(s.indices.data(), data + r.off, 16ull * M);
    r.off += 16ull * M;
    // densities: u8 log-encoded, exp((v - 100) / 20)
    float lut[256];
    for (int i = 0; i < 256; ++i) lut[i] = std::exp(((float)i - 100.0f) / 20.0f);
    s.densities.resize(M);
    const uint8_t* du = data + r.off;
    parallel_for(M, [&](size_t b, size_t e) {
        for (size_t i = b; i < e; ++i) s.densities[i] = lut[du[i]];
    });
    r.off += M;
    r.align4();
    if (!r.need(2 * (dims + ek * dims + (size_t)M * ek))) return bad(".rmesh: SH data truncated");
    std::vector<float> mean(dims), basis(ek * dims), weights((size_t)M * ek);
    f16_block(data + r.off, mean.data(), dims);
    r.off += 2 * dims;
    f16_block(data + r.off, basis.data(), ek * dims);
    r.off += 2 * ek * dims;
    f16_block(data + r.off, weights.data(), (size_t)M * ek);
    r.off += 2 * (size_t)M * ek;
    r.align4();
    // colour gradients: absent when the tagged sections start here, zero-filled when truncated
    s.color_grads.assign(3ull * M, 0.0f);
    const bool tagged_here = r.need(4) && [&] {
        uint32_t m;
        std::memcpy(&m, data + r.off, 4);
        return m == 0x524D5458u;
    }();
    if (!tagged_here) {
        const size_t want = 3ull * M, avail = r.off <= n ? (n - r.off) / 2 : 0;
        const size_t got = std::min(want, avail);
        f16_block(data + r.off, s.color_grads.data(), got);
        r.off += 2 * got;
    }
    // SH: mean + sum_k weight_k basis_k, each element summed in k order
    // (lib.rs:330-346); the loop over features is innermost so it vectorises
    s.sh.resize((size_t)M * dims);
    parallel_for(M, [&](size_t b, size_t e) {
        for (size_t t = b; t < e; ++t) {
            const float* w = weights.data() + t * ek;
            float* o = s.sh.data() + t * dims;
            for (size_t f = 0; f < dims; ++f) o[f] = mean[f];
            for (size_t q = 0; q < ek; ++q) {
                const float wq = w[q];
                const float* bq = basis.data() + q * dims;
                for (size_t f = 0; f < dims; ++f) o[f] += wq * bq[f];
            }
        }
    });
    rmesh_status st = check_indices(s);
    if (st != RMESH_OK) return st;
    s.circumdata.resize(4ull * M);
    circumspheres(s.vertices.data(), s.indices.data(), M, s.circumdata.data());
    // tagged extension sections (lib.rs:396-470): counted, their contents are not used here
    r.align4();
    s.info.n_tagged = 0;
    if (r.need(8)) {
        uint32_t m;
        std::memcpy(&m, data + r.off, 4);
        if (m == 0x524D5458u) {
            r.off += 4;
            s.info.n_tagged = r.u32();
        }
    }
    return RMESH_OK;
}

// load_ply (lib.rs:711-853): binary little-endian, vertices (x, y, z) then
// tetrahedra (uchar 4, int indices[4], s, grd_xyz, sh_i_{r,g,b} interleaved);
// the SH leave planar per tet ([channel][coefficient]).
rmesh_status parse_ply(const uint8_t* data, size_t n, rmesh_scene_file& s) {
    static const char marker[] = "end_header\n";
    const uint8_t* end = std::search(data, data + n, marker, marker + 11);
    if (end == data + n) return bad("PLY: missing end_header");
    const size_t hdr = (size_t)(end - data) + 11;
    std::string header((const char*)data, (size_t)(end - data));
    uint64_t N = 0, M = 0;
    uint32_t num_sh = 0;
    bool le = false, in_tet = false;
    size_t pos = 0;
    while (pos <= header.size()) {
        size_t nl = header.find('\n', pos);
        if (nl == std::string::npos) nl = header.size();
        std::string line = header.substr(pos, nl - pos);
        pos = nl + 1;
        while (!line.empty() && (line.back() == '\r' || line.back() == ' ' || line.back() == '\t')) line.pop_back();
        size_t lead = line.find_first_not_of(" \t");
        line = lead == std::string::npos ? std::string() : line.substr(lead);
        auto starts = [&](const char* p) { return line.rfind(p, 0) == 0; };
        if (line == "format binary_little_endian 1.0") {
            le = true;
        } else if (starts("element vertex ")) {
            N = std::strtoull(line.c_str() + 15, nullptr, 10);
            in_tet = false;
        } else if (starts("element tetrahedron ")) {
            M = std::strtoull(line.c_str() + 20, nullptr, 10);
            in_tet = true;
        } else if (starts("element ")) {
            in_tet = false;
        } else if (in_tet && starts("property float sh_") && line.size() >= 2 &&
                   line.compare(line.size() - 2, 2, "_r") == 0) {
            ++num_sh;
        }
    }
    if (!le) return bad("PLY: only binary_little_endian is supported");
    if (N == 0 || M == 0) return bad("PLY: empty");
    // counts and indices are 32-bit in the scene (and keep the size products below from wrapping)
    if (N > UINT32_MAX || M > UINT32_MAX) return bad("PLY: element count above 2^32 - 1");
    const uint32_t deg = (uint32_t)std::sqrt((double)num_sh) - 1;
    const uint32_t nc = (deg + 1) * (deg + 1);
    if (num_sh == 0 || nc != num_sh) return bad("PLY: SH count " + std::to_string(num_sh) + " is not a perfect square");
    if (deg > 3) return bad("PLY: SH degree " + std::to_string(deg) + " (0-3 supported)");
    s.info.n_vertices = N;
    s.info.n_tets = M;
    s.info.sh_degree = deg;
    s.info.k_components = 0;
    s.info.n_tagged = 0;
    for (float& x : s.info.start_pose) x = 0.0f;
    const size_t stride = 1 + 16 + 4 * (1 + 3 + 3 * (size_t)nc);
    if (hdr > n || 12ull * N > n - hdr || stride * M > n - hdr - 12ull * N) return bad("PLY: data truncated");
    const uint8_t* vp = data + hdr;
    s.vertices.resize(3ull * N);
    std::memcpy(s.vertices.data(), vp, 12ull * N);
    const uint8_t* tp = vp + 12ull * N;
    s.indices.resize(4ull * M);
    s.densities.resize(M);
    s.color_grads.resize(3ull * M);
    s.sh.resize((size_t)M * 3 * nc);
    parallel_for(M, [&](size_t b, size_t e) {
        for (size_t t = b; t < e; ++t) {
            const uint8_t* rec = tp + t * stride;
            for (int k = 0; k < 4; ++k) {
                int32_t v;
                std::memcpy(&v, rec + 1 + 4 * k, 4);
                s.indices[4 * t + k] = (uint32_t)v;
            }
            std::memcpy(&s.densities[t], rec + 17, 4);
            std::memcpy(&s.color_grads[3 * t], rec + 21, 12);
            const uint8_t* sh = rec + 33;
            for (uint32_t c = 0; c < nc; ++c)
                for (int ch = 0; ch < 3; ++ch)
                    std::memcpy(&s.sh[t * 3 * nc + (size_t)ch * nc + c], sh + 4 * (3 * c + ch), 4);
        }
    });
    rmesh_status st = check_indices(s);
    if (st != RMESH_OK) return st;
    s.circumdata.resize(4ull * M);
    circumspheres(s.vertices.data(), s.indices.data(), M, s.circumdata.data());
    return RMESH_OK;
}

}  // namespace

extern "C" rmesh_status rmesh_scene_parse(const uint8_t* data, uint64_t nbytes, int32_t format,
                                          rmesh_scene_file** out, rmesh_scene_info* info) {
    if (!data || !out || !info) return bad("null argument");
    *out = nullptr;
    auto* s = new (std::nothrow) rmesh_scene_file();
    if (!s) return bad("out of host memory");
    std::memset(&s->info, 0, sizeof(s->info));
    rmesh_status st;
    try {
        if (format == RMESH_FILE_RMESH || format == RMESH_FILE_RMESH_RAW) {
            if (format == RMESH_FILE_RMESH && nbytes >= 2 && data[0] == 0x1f && data[1] == 0x8b) {
                std::vector<uint8_t> raw;
                std::string err;
                if (!gunzip(data, nbytes, raw, err)) {
                    delete s;
                    return bad(".rmesh: " + err);
                }
                st = parse_rmesh(raw.data(), raw.size(), *s);
            } else {
                st = parse_rmesh(data, nbytes, *s);
            }
        } else if (format == RMESH_FILE_PLY) {
            st = parse_ply(data, nbytes, *s);
        } else {
            delete s;
            return bad("unknown scene file format " + std::to_string(format));
        }
    } catch (const std::bad_alloc&) {
        st = bad("out of host memory");
    }
    if (st != RMESH_OK) {
        delete s;
        return st;
    }
    *info = s->info;
    *out = s;
    return RMESH_OK;
}

extern "C" rmesh_status rmesh_scene_copy(const rmesh_scene_file* s, float* vertices, uint32_t* indices,
                                         float* densities, float* color_grads, float* circumdata, float* sh_coeffs) {
    if (!s) return bad("null argument");
    auto put = [](void* dst, const void* src, size_t bytes) {
        if (dst && bytes) std::memcpy(dst, src, bytes);
    };
    put(vertices, s->vertices.data(), 4 * s->vertices.size());
    put(indices, s->indices.data(), 4 * s->indices.size());
    put(densities, s->densities.data(), 4 * s->densities.size());
    put(color_grads, s->color_grads.data(), 4 * s->color_grads.size());
    put(circumdata, s->circumdata.data(), 4 * s->circumdata.size());
    put(sh_coeffs, s->sh.data(), 4 * s->sh.size());
    return RMESH_OK;
}

extern "C" void rmesh_scene_free(rmesh_scene_file* s) { delete s; }
