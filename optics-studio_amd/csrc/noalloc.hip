// The batched complex64 propagation without device allocation
// (oc_propagate_batch_noalloc_dev_c64): for callers that cannot allocate
// (code under stream capture, tight serving loops, harnesses that forbid
// hipMalloc).  The intermediates, kernel spectra, shared transmissions,
// flags, descriptors, twiddles and axis/mask tables of the call come from
// one scratch array in device memory, reserved when the library's code
// object loads (one copy per device, held until the process ends), instead
// of the context's cached buffers.  Same kernels, same layouts and the same
// tables as oc_propagate_batch_dev_c64, so the results are the same bits.
#include "runtime.hpp"

namespace {

constexpr size_t kScratchBytes = (size_t)256 << 20;  // 256 MiB per device
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_lookup_mu;
void* g_scratch_ptr[kMaxDevices] = {};  // this device's address of `scratch`
// no-allocation calls on one device share its array: one at a time, each
// waiting for its kernels before the next starts
std::mutex g_device_mu[kMaxDevices];

// Enters the no-allocation mode for one call; leaving waits for the call's
// queued work (the host copies of its tables are released then).
struct ArenaMode {
    oc_context* ctx;
    explicit ArenaMode(oc_context* c) : ctx(c) {
        ctx->arena = true;
        ctx->arena_used = 0;
        ctx->arena_keep.clear();
        ctx->arena_tables.clear();
    }
    ~ArenaMode() {
        (void)hipStreamSynchronize(ctx->stream);
        ctx->arena = false;
        ctx->arena_used = 0;
        ctx->arena_keep.clear();
        ctx->arena_tables.clear();
    }
};

}  // namespace

namespace ocr {

// Looks up this device's address of the scratch array once per device (the
// context's device is current) and hands it to the context.
int noalloc_attach(oc_context* ctx) {
    if (ctx->device < 0 || ctx->device >= kMaxDevices) {
        ctx->scratch = nullptr;
        ctx->scratch_cap = 0;
        return OC_OK;  // the no-allocation entry point then refuses
    }
    std::lock_guard<std::mutex> lk(g_lookup_mu);
    if (!g_scratch_ptr[ctx->device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(ctx, OC_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[ctx->device] = ptr;
    }
    ctx->scratch = g_scratch_ptr[ctx->device];
    ctx->scratch_cap = kScratchBytes;
    return OC_OK;
}

}  // namespace ocr

using namespace ocr;

extern "C" int oc_propagate_batch_noalloc_dev_c64(oc_context* ctx, oc_prop_desc* d, size_t count) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx || (!d && count)) return OC_ERR_INVALID_ARGUMENT;
    if (!ctx->scratch) return fail(ctx, OC_ERR_DEVICE, "no scratch array on this device");
    (void)hipSetDevice(ctx->device);
    std::vector<Prob> v;
    int st = probs_from_descs(d, count, v);
    if (st) return st;
    const size_t need = arena_need(v, false);
    if (need > ctx->scratch_cap) {
        for (size_t i = 0; i < count; ++i) d[i].status = OC_ERR_INVALID_ARGUMENT;
        return fail(ctx, OC_ERR_INVALID_ARGUMENT,
                    "the batch needs up to " + std::to_string((need + (1 << 20) - 1) >> 20) +
                        " MiB of scratch; the no-allocation form holds " + std::to_string(ctx->scratch_cap >> 20) +
                        " MiB (use oc_propagate_batch_dev_c64 or split the batch)");
    }
    std::lock_guard<std::mutex> dev_lock(g_device_mu[ctx->device]);
    {
        ArenaMode am(ctx);
        st = run_batch(ctx, false, v);
    }
    for (size_t i = 0; i < count; ++i) d[i].status = v[i].status;
    return st;
}
