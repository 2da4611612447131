// msa_nw_plain_noalloc: the plain kernel on caller-owned device memory, with
// its workspace in a fixed device array instead of memory the library
// allocates.
//
// msa_nw_plain (nw_plan.hip) grows a per-stream workspace with hipMalloc /
// hipFree: a SeqInfo per sequence and one strip-boundary line per workgroup.
// Callers that cannot allocate (code under hipGraph capture, tight serving
// loops, harnesses that forbid hipMalloc) call this entry point instead: the
// same two launches (seq_info_kernel, nw_plain_kernel) with the workspace
// carved out of `scratch`, a file-scope device array of kScratchBytes that
// every GPU the library runs on holds until the process ends.  Its device
// address is looked up once per device with hipGetSymbolAddress; host code
// uses that address, never the array's name.  Nothing in the workspace needs
// to start at zero: seq_info_kernel writes every SeqInfo the plain kernel
// reads, and every strip writes its bottom row before the strip below reads
// it.  Every call on a device shares that device's array, so calls that
// overlap on one device (on different streams, or from different threads)
// race on it.
#include "msa_internal.hpp"
#include "nw_planner.hpp"

#include <algorithm>
#include <mutex>

namespace msa {

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;   // the whole array, per device
constexpr size_t kInfoBytes = size_t(2) << 20;       // SeqInfo region: at most 131,072 sequences
constexpr size_t kLineBytes = kScratchBytes - kInfoBytes;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

std::mutex g_scratch_lock;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's copy of `scratch` (looked up once).
int scratch_address(int dev, uint8_t** out)
{
    if (dev < 0 || dev >= kMaxDevices) return fail(MSA_ERR_INVALID_ARGUMENT, "device %d out of range", dev);
    std::lock_guard<std::mutex> guard(g_scratch_lock);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        MSA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(MSA_ERR_RUNTIME, "hipGetSymbolAddress gave no address for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    *out = static_cast<uint8_t*>(g_scratch_ptr[dev]);
    return MSA_OK;
}

}  // namespace

}  // namespace msa

using msa::fail;

extern "C" int msa_nw_plain_noalloc(const uint16_t* d_blocks, const int64_t* d_offset, const int32_t* d_size,
                                    int32_t nseq, int32_t max_blocks, const int32_t* d_pairs, int64_t npairs,
                                    const int32_t* table625, int32_t penalty, int32_t* d_out, msa_stream stream)
{
    if (npairs == 0) return MSA_OK;
    if (npairs < 0 || nseq < 1 || max_blocks < 0 || !d_blocks || !d_offset || !d_size || !d_out || !table625)
        return fail(MSA_ERR_INVALID_ARGUMENT, "invalid arguments to msa_nw_plain_noalloc");
    if (!d_pairs && npairs != (int64_t)nseq * (nseq - 1) / 2)
        return fail(MSA_ERR_INVALID_ARGUMENT, "all-pairs mode needs npairs = nseq*(nseq-1)/2");
    if ((size_t)nseq * sizeof(msa::SeqInfo) > msa::kInfoBytes)
        return fail(MSA_ERR_TOO_LARGE, "msa_nw_plain_noalloc takes at most %zu sequences, got %d",
                    msa::kInfoBytes / sizeof(msa::SeqInfo), nseq);
    // One strip-boundary line per workgroup, as long as the longest sequence.
    const int64_t stride = ((int64_t)3 * max_blocks + 63) & ~int64_t(63);
    const int64_t line_bytes = std::max<int64_t>(stride, 1) * (int64_t)sizeof(int32_t);
    const int64_t max_grid = (int64_t)msa::kLineBytes / line_bytes;
    if (max_grid < 1)
        return fail(MSA_ERR_TOO_LARGE, "msa_nw_plain_noalloc takes sequences of at most %lld blocks, got %d",
                    (long long)(msa::kLineBytes / sizeof(int32_t) / 3), max_blocks);
    msa::Sxx sxx;
    bool symmetric = true;
    MSA_TRY(msa::make_sxx(table625, penalty, sxx, symmetric));
    int dev = 0;
    MSA_HIP(hipGetDevice(&dev));
    msa::DeviceShape shape;
    MSA_TRY(msa::device_shape(shape));
    uint8_t* base = nullptr;
    MSA_TRY(msa::scratch_address(dev, &base));
    // Fewer workgroups than msa_nw_plain when the lines would not fit: the
    // kernel strides over the pairs, so any grid scores them all.
    const int grid = (int)std::min<int64_t>(msa::plain_grid(npairs, shape.compute_units), max_grid);
    msa::SeqInfo* info = reinterpret_cast<msa::SeqInfo*>(base);
    int32_t* lines = reinterpret_cast<int32_t*>(base + msa::kInfoBytes);
    const hipStream_t s = (hipStream_t)stream;
    MSA_TRY(msa::seq_info_launch(d_blocks, d_offset, d_size, nseq, info, s));
    return msa::plain_launch(d_blocks, info, d_pairs, nullptr, npairs, sxx, penalty, d_out, lines,
                             (int32_t)std::max<int64_t>(stride, 1), grid, s);
}
