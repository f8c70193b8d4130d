// Sorted score rows with no allocation in the call: rkr_sorted_scores_noalloc.
//
// rkr_sorted_scores (ComputeScoreTableBatch::ComputeSortItemsBatch,
// ComputeScoreTableBatch.hpp:79-92) allocates a score block and a sort scratch
// row block per call and copies the rows out through pinned host buffers.
// This variant serves callers that cannot allocate (code under HIP graph
// capture, tight serving loops): the caller passes the output rows in device
// memory, the score GEMM writes them there, and the merge passes of the row
// sort use `scratch`, a file-scope device array of kNoAllocScratchBytes that
// each GPU holds from module load until the process ends.  Host code never
// names the array: it uses the address hipGetSymbolAddress returns for the
// current device, taken once per device when a context is created there.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "rkr_internal.hpp"

namespace rkr {
namespace {

constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20;  // 8M floats: the sort's scratch rows
static __device__ __align__(256) float scratch[kNoAllocScratchBytes / sizeof(float)];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float *g_scratch_ptr[kMaxDevices] = {};      // per device: the device address of scratch
hipError_t g_scratch_err[kMaxDevices] = {};  // per device: the lookup's result once tried
bool g_scratch_tried[kMaxDevices] = {};

}  // namespace

// Looks up scratch's device address on `device` (which must be current), once
// per device; later calls return the recorded result.
hipError_t noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_tried[device]) {
        void *p = nullptr;
        const hipError_t r = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch_tried[device] = true;
        g_scratch_err[device] = r;
        g_scratch_ptr[device] = r == hipSuccess ? static_cast<float *>(p) : nullptr;
        if (r != hipSuccess) (void)hipGetLastError();
    }
    return g_scratch_err[device];
}

static float *noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    return g_scratch_ptr[device];
}

}  // namespace rkr

struct rkr_ctx : public rkr::Ctx {};

extern "C" int rkr_sorted_scores_noalloc(rkr_ctx *c, int64_t user_begin, int64_t n_rows, float *out_dev) {
    using namespace rkr;
    if (!c) return RKR_ERR_INVALID;
    if (n_rows > 0 && !out_dev) return set_err(c, RKR_ERR_INVALID, "sorted_scores_noalloc: null output");
    if (!c->users) return set_err(c, RKR_ERR_STATE, "sorted_scores_noalloc: no data");
    if (user_begin < 0 || n_rows < 0 || user_begin + n_rows > c->n_user)
        return set_err(c, RKR_ERR_INVALID, "sorted_scores_noalloc: row range out of bounds");
    const int64_t I = c->n_item;
    const int64_t cap_rows = (int64_t)(kNoAllocScratchBytes / (sizeof(float) * (size_t)std::max<int64_t>(I, 1)));
    if (cap_rows < 1)
        return set_err(c, RKR_ERR_INVALID,
                       "sorted_scores_noalloc: a row of " + std::to_string(I) + " scores exceeds the " +
                           std::to_string(kNoAllocScratchBytes >> 20) + " MiB scratch (at most " +
                           std::to_string(kNoAllocScratchBytes / sizeof(float)) + " items)");
    if (n_rows == 0) return RKR_OK;
    RKR_HIP(c, hipSetDevice(c->device));
    float *tmp = noalloc_scratch(c->device);
    if (!tmp) return set_err(c, RKR_ERR_STATE, "sorted_scores_noalloc: scratch address unavailable on this device");
    // rows go straight into the caller's output; each chunk's merge passes use
    // the scratch rows, so chunks follow each other on the context's stream
    for (int64_t b0 = 0; b0 < n_rows; b0 += cap_rows) {
        const int64_t m = std::min<int64_t>(cap_rows, n_rows - b0);
        float *rows = out_dev + b0 * I;
        RKR_HIP(c, launch_scores_store(c->stream, c->users + (user_begin + b0) * c->ld, nullptr, m, c->ld, c->items,
                                       I, c->ld, c->dim, rows, I, true));
        RKR_HIP(c, launch_sort_rows(c->stream, rows, I, m, I, rows, tmp, I, c->n_cu));
    }
    return RKR_OK;
}
