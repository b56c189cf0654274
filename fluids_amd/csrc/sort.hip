// sph_sort_pairs: the library's stable key/value radix sort on device arrays
// the caller owns, in place of the reference's clsort (opencl/algorithms/
// clsort.h:12-35, a clogs::Radixsort over cl_mem keys and values), with the
// kernels of grid.hip. sph_sort_pairs takes its ping-pong arrays and digit
// histograms from the stream-ordered allocator; sph_sort_pairs_noalloc takes
// them from `scratch`, a fixed array in the library's device image, and
// allocates nothing.
#include <mutex>
#include <string>

#include "kernels.h"
#include "solver.h"
#include "sph.h"

namespace sph {

// Device memory one sort of n pairs needs beyond its inputs: the other half
// of the key and value ping-pong and the digit histograms, each 256-byte
// aligned.
static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
static size_t sort_scratch_bytes(int n) {
    return 2 * align256((size_t)n * 4) + align256(sort_hist_words(n) * 4);
}

// Scratch of the no-allocation sort, sized for SPH_SORT_NOALLOC_MAX pairs
// (the key and value ping-pong halves and the digit histograms: 2048-key
// tiles, at most 1024 digit counters per tile, plus the digit totals). Every
// GPU the library runs on holds one copy until the process ends; all calls on
// a GPU share it. In host code `scratch` only names the symbol: its address on
// the current device comes from hipGetSymbolAddress (scratch_ptr).
constexpr size_t NOALLOC_BYTES =
    2 * ((size_t)SPH_SORT_NOALLOC_MAX * 4) + 4 * ((size_t)1024 * (SPH_SORT_NOALLOC_MAX / 2048) + 1024);
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

// Device address of `scratch` on the current device, looked up once per
// device with that device current.
static void* scratch_ptr() {
    constexpr int MAX_DEVICES = 64;
    static std::mutex mu;
    static void* ptr[MAX_DEVICES] = {};
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= MAX_DEVICES) throw Error(SPH_ERR_HIP, "device index out of range");
    std::lock_guard<std::mutex> lk(mu);
    if (!ptr[dev]) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!p) throw Error(SPH_ERR_HIP, "hipGetSymbolAddress gave no address for the sort scratch");
        ptr[dev] = p;
    }
    return ptr[dev];
}

static void check_args(const uint32_t* keys, const int32_t* values, int n, int& key_bits) {
    if (n < 0) throw Error(SPH_ERR_INVALID, "sort: negative size");
    if (key_bits == 0) key_bits = 32;
    if (key_bits < 1 || key_bits > 32) throw Error(SPH_ERR_INVALID, "sort: key_bits must be 0 or 1..32");
    if (n > 0 && (!keys || !values)) throw Error(SPH_ERR_INVALID, "sort: keys and values are required");
    if (n > MAX_SCENE_PARTICLES) throw Error(SPH_ERR_RANGE, "sort: at most 2^27 pairs");
}

// The sort of (keys, values) in place, with the other ping-pong halves and
// the histograms in `work` (sort_scratch_bytes(n) bytes). With an odd pass
// count the result lands in the work arrays and is copied back.
static void sort_in_place(uint32_t* keys, int32_t* values, int n, int key_bits, void* work, hipStream_t s) {
    if (n == 0) return;
    unsigned char* w = static_cast<unsigned char*>(work);
    uint32_t* kw = reinterpret_cast<uint32_t*>(w);
    int32_t* vw = reinterpret_cast<int32_t*>(w + align256((size_t)n * 4));
    uint32_t* hist = reinterpret_cast<uint32_t*>(w + 2 * align256((size_t)n * 4));
    if (sort_passes(key_bits) % 2 == 0) {
        launch_radix_sort(keys, kw, values, vw, hist, n, key_bits, s, nullptr, nullptr, 0, values);
    } else {
        launch_radix_sort(kw, keys, vw, values, hist, n, key_bits, s, nullptr, nullptr, 0, values);
        hip_check(hipMemcpyAsync(keys, kw, (size_t)n * 4, hipMemcpyDeviceToDevice, s), "copy");
        hip_check(hipMemcpyAsync(values, vw, (size_t)n * 4, hipMemcpyDeviceToDevice, s), "copy");
    }
    hip_check(hipGetLastError(), "sort launch");
}

// Entry point bodies (the extern "C" wrappers in sph_abi.cpp convert
// exceptions to status codes).
void sort_pairs(uint32_t* keys, int32_t* values, int n, int key_bits, void* stream) {
    check_args(keys, values, n, key_bits);
    if (n == 0) return;
    hipStream_t s = static_cast<hipStream_t>(stream);
    void* work = nullptr;
    hip_check(hipMallocAsync(&work, sort_scratch_bytes(n), s), "hipMallocAsync");
    try {
        sort_in_place(keys, values, n, key_bits, work, s);
    } catch (...) {
        (void)hipFreeAsync(work, s);
        throw;
    }
    hip_check(hipFreeAsync(work, s), "hipFreeAsync");
}

// The same sort with its work arrays in `scratch`: no allocation. The
// histograms are written in full by every pass, so nothing in the scratch has
// to start at zero; calls that overlap on one GPU share the scratch and race.
void sort_pairs_noalloc(uint32_t* keys, int32_t* values, int n, int key_bits, void* stream) {
    check_args(keys, values, n, key_bits);
    if (n > SPH_SORT_NOALLOC_MAX)
        throw Error(SPH_ERR_RANGE, "sph_sort_pairs_noalloc: " + std::to_string(n) +
                                       " pairs exceed its scratch of SPH_SORT_NOALLOC_MAX = " +
                                       std::to_string(SPH_SORT_NOALLOC_MAX) + " pairs");
    if (sort_scratch_bytes(n) > NOALLOC_BYTES) throw Error(SPH_ERR_RANGE, "sort scratch too small");
    void* work = scratch_ptr();
    sort_in_place(keys, values, n, key_bits, work, static_cast<hipStream_t>(stream));
}

// Non-zero when the hint decreases somewhere.
__global__ __launch_bounds__(256) void k_check_non_decreasing(const uint32_t* __restrict__ a, int n,
                                                              int* __restrict__ bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i + 1 < n && a[i + 1] < a[i]) atomicOr(bad, 1);
}

void sort_near(const uint32_t* hint, const uint32_t* keys, uint32_t* sorted_keys, int32_t* perm, int n,
               void* stream) {
    if (n < 0) throw Error(SPH_ERR_INVALID, "sort: negative size");
    if (n > MAX_SCENE_PARTICLES) throw Error(SPH_ERR_RANGE, "sort: at most 2^27 pairs");
    if (n == 0) return;
    if (!hint || !keys || !sorted_keys || !perm) throw Error(SPH_ERR_INVALID, "sort: NULL array");
    if (sorted_keys == keys) throw Error(SPH_ERR_INVALID, "sph_sort_near: sorted_keys must not be keys");
    hipStream_t s = static_cast<hipStream_t>(stream);
    int* bad = nullptr;
    hip_check(hipMallocAsync(reinterpret_cast<void**>(&bad), 256 + resort_bytes(n), s), "hipMallocAsync");
    struct Free {
        void* p;
        hipStream_t s;
        ~Free() { (void)hipFreeAsync(p, s); }
    } release{bad, s};
    const ResortWork w = resort_carve(reinterpret_cast<unsigned char*>(bad) + 256, n);
    hip_check(hipMemsetAsync(bad, 0, 4, s), "memset");
    resort_clear(w, n, s);
    hipLaunchKernelGGL(k_check_non_decreasing, dim3((n + 255) / 256), dim3(256), 0, s, hint, n, bad);
    hip_check(hipGetLastError(), "check launch");
    int h_bad = 0;
    hip_check(hipMemcpyAsync(&h_bad, bad, 4, hipMemcpyDeviceToHost, s), "download");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (h_bad) throw Error(SPH_ERR_INVALID, "sph_sort_near: the hint is not non-decreasing");
    if (sorted_keys != hint)
        hip_check(hipMemcpyAsync(sorted_keys, hint, (size_t)n * 4, hipMemcpyDeviceToDevice, s), "copy");
    launch_resort_samples(sorted_keys, n, w, s);
    launch_resort(n, keys, sorted_keys, perm, w, s);
    hip_check(hipGetLastError(), "sort launch");
}

}  // namespace sph
