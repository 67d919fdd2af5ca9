// rmesh_sort_pairs_noalloc: the pipeline's radix sort over caller-owned device
// buffers with every scratch buffer (the ping-pong key/value copies, the
// per-block digit histograms and the pair counter) in one file-scope device
// array, so the call allocates nothing: usable under hipGraph capture, in
// serving loops and in harnesses that forbid hipMalloc.
//
// Each device holds its own copy of `scratch` (kNoAllocMaxPairs pairs,
// 14 MB) until the process ends.  Its address on a device is taken once, with
// that device current, by hipGetSymbolAddress; kernels only ever receive that
// address (the array's name in host code is a host-side stand-in).  Every call
// on a device shares the device's array: calls that overlap on one GPU (two
// streams, two threads) race, and nothing from a previous call is relied on
// (the counter is rewritten by a kernel at the start of every call).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/rmesh_hip.h"
#include "rmesh_common.hpp"
#include "rmesh_error.hpp"
#include "rmesh_kernels.hpp"

namespace {

constexpr uint32_t kNoAllocMaxPairs = 1u << 20;
constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
// layout: keys [max] u64 | vals [max] u32 | histograms (256 per sort block + 256 digit totals) | counter
constexpr size_t kOffVals = align256(8ull * kNoAllocMaxPairs);
constexpr size_t kOffHist = kOffVals + align256(4ull * kNoAllocMaxPairs);
constexpr size_t kHistWords = 256ull * 2048 + 256;  // the sort's grid never exceeds 2048 blocks
constexpr size_t kOffCount = kOffHist + align256(4 * kHistWords);
constexpr size_t kScratchBytes = kOffCount + 256;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_scratch[kMaxDevices] = {};

// Address of `scratch` on device dev (current); taken once per device.
hipError_t scratch_ptr(int dev, unsigned char** out) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

__global__ __launch_bounds__(256) void k_copy_pairs(const uint64_t* __restrict__ ks, const uint32_t* __restrict__ vs,
                                                    uint64_t* __restrict__ kd, uint32_t* __restrict__ vd, uint32_t n) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        kd[i] = ks[i];
        vd[i] = vs[i];
    }
}

rmesh_status fail(rmesh_status code, const std::string& msg) {
    rmesh::set_last_error(msg);
    return code;
}

}  // namespace

extern "C" rmesh_status rmesh_sort_pairs_noalloc(uint64_t* keys, uint32_t* vals, uint64_t n, uint32_t bits,
                                                 int32_t device_id, void* hip_stream) {
    if ((!keys || !vals) && n) return fail(RMESH_ERR_INVALID, "null argument");
    if (bits % 8 != 0 || bits > 64) return fail(RMESH_ERR_INVALID, "sorting_bits must be a multiple of 8 up to 64");
    if (n > kNoAllocMaxPairs)
        return fail(RMESH_ERR_INVALID, "n exceeds the no-alloc scratch capacity (" +
                                           std::to_string(kNoAllocMaxPairs) + " pairs); use rmesh_sort_pairs");
    int prev = -1;
    if (hipGetDevice(&prev) != hipSuccess) return fail(RMESH_ERR_DEVICE, "no HIP device");
    const int dev = device_id < 0 ? prev : device_id;
    if (dev < 0 || dev >= kMaxDevices) return fail(RMESH_ERR_INVALID, "device id out of range");
    if (n == 0 || bits == 0) return RMESH_OK;
    if (dev != prev && hipSetDevice(dev) != hipSuccess) return fail(RMESH_ERR_DEVICE, "hipSetDevice failed");
    rmesh_status st = RMESH_OK;
    unsigned char* base = nullptr;
    hipError_t e = scratch_ptr(dev, &base);
    if (e != hipSuccess) {
        st = fail(RMESH_ERR_DEVICE, std::string("hipGetSymbolAddress: ") + hipGetErrorString(e));
    } else {
        hipStream_t s = static_cast<hipStream_t>(hip_stream);
        uint64_t* kb = reinterpret_cast<uint64_t*>(base);
        uint32_t* vb = reinterpret_cast<uint32_t*>(base + kOffVals);
        uint32_t* hist = reinterpret_cast<uint32_t*>(base + kOffHist);
        uint32_t* count = reinterpret_cast<uint32_t*>(base + kOffCount);
        const uint32_t cap = (uint32_t)n;
        static_assert(256ull * 2048 + 256 == kHistWords, "histogram words of the largest sort grid");
        if (rmesh::radix_hist_words(cap) > kHistWords) {
            if (dev != prev) (void)hipSetDevice(prev);
            return fail(RMESH_ERR_INVALID, "sort grid exceeds the no-alloc histogram area");
        }
        rmesh::launch_fill_u32(count, cap, 1, s);
        const bool in_b = rmesh::launch_radix_sort(keys, vals, kb, vb, count, cap, bits, hist, s);
        if (in_b) {
            const uint32_t blocks = (cap + 255) / 256 < 2048 ? (cap + 255) / 256 : 2048;
            hipLaunchKernelGGL(k_copy_pairs, dim3(blocks), dim3(256), 0, s, kb, vb, keys, vals, cap);
        }
        e = hipGetLastError();
        if (e != hipSuccess) st = fail(RMESH_ERR_DEVICE, std::string("sort launch: ") + hipGetErrorString(e));
    }
    if (dev != prev) (void)hipSetDevice(prev);
    return st;
}
