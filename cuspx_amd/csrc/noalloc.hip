// noalloc.hip -- cuspx_sign_device_noalloc: device-resident batched signing
// that allocates nothing.  The batch's workspace (plans, FORS roots, layer
// roots, chain capture, T_len parking and work counters) lives in a
// file-scope device array of kScratchBytes, one copy per GPU, held until the
// process ends.  Its address on each device is taken once, with that device
// current, on the entry point's first call there; host code uses only that
// address (the array's name is not a device pointer in host code).
//
// Every call on one GPU shares that GPU's array: calls that overlap on one
// GPU (other streams, other threads) race.  The counters the kernels need at
// zero are cleared by the batch's own prep kernel, so the array's contents
// from the previous call do not matter.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "../../include/cuspx.h"
#include "engine.hpp"

namespace {

constexpr size_t kScratchBytes = 128ull << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

std::mutex g_mu;
void* g_scratch_ptr[kMaxDevices];

// the scratch array's device address on the current device (taken once)
int scratch_ptr(void** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return spx::kErrHip;
    if (dev < 0 || dev >= kMaxDevices) return spx::kErrArg;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
            return spx::set_error(spx::kErrHip, "hipGetSymbolAddress of the scratch array failed");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return spx::kOk;
}

}  // namespace

extern "C" {

int cuspx_sign_device_noalloc(int pid, uint32_t num, const uint8_t* sk, uint32_t sk_stride,
                              const uint8_t* msg, const uint64_t* moff, const uint32_t* mlen,
                              const uint8_t* optrand, uint8_t* sig, uint64_t sig_stride,
                              void* stream) {
    const spx::SetOps* o = spx::get_set(pid);
    if (!o) return spx::kErrParam;
    if (num == 0) return spx::kOk;
    if (!sk || !msg || !moff || !mlen || !optrand || !sig)
        return spx::set_error(spx::kErrArg, "null argument");
    if (o->workspace(num, spx::kOpSign) > kScratchBytes)
        return spx::set_error(spx::kErrBudget,
                              "batch needs more workspace than the 128 MiB scratch array");
    void* ws = nullptr;
    const int rc = scratch_ptr(&ws);
    if (rc) return rc;
    spx::DevBatch b{};
    b.n = num;
    b.sk = sk;
    b.sk_stride = sk_stride;
    b.msg = msg;
    b.moff = moff;
    b.mlen = mlen;
    b.optrand = optrand;
    b.sig = sig;
    b.sig_stride = sig_stride;
    b.ws = ws;
    return o->sign(b, (hipStream_t)stream);
}

}  // extern "C"
