// engine.hpp -- host-side types shared by the per-set launchers and the
// C-ABI layer.  No torch types: plain pointers, sizes and HIP streams.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace spx {

enum { kOpSign = 0, kOpVerify = 1, kOpKeygen = 2 };

// status codes of every C-ABI entry point
enum {
    kOk = 0,
    kErrVerify = -1,   // verification failed (api.h semantics)
    kErrParam = -2,    // unknown parameter-set id
    kErrHip = -3,      // HIP runtime / launch error
    kErrArg = -4,      // bad argument (null pointer, size)
    kErrBudget = -5,   // request exceeds the device-memory budget
};

// One batch on the device.  Every pointer is device memory.
struct DevBatch {
    uint32_t n;
    const uint8_t* sk;
    uint32_t sk_stride;
    const uint8_t* pk;
    uint32_t pk_stride;
    const uint8_t* seed;
    uint32_t seed_stride;
    const uint8_t* msg;      // msg + moff[i]: 4-byte aligned, readable to a 4-byte boundary
    const uint64_t* moff;
    const uint32_t* mlen;
    uint32_t mlen_bound;     // no message is longer (0: not known)
    const uint8_t* optrand;  // [n][param n]
    const uint32_t* key_idx; // optional key row per request (sk / pk tables)
    const uint32_t* key_mid; // optional precomputed midstate per key row (8 words)
    uint8_t* sig;
    uint64_t sig_stride;
    int32_t* status;
    uint8_t* pk_out;
    uint8_t* sk_out;
    void* ws;                // workspace of workspace(n, op) bytes
};

struct SetOps {
    const char* name;
    int pid;
    int hash, n, full_h, d, a, k, len;
    int sig_bytes, pk_bytes, sk_bytes;
    size_t (*workspace)(uint32_t n, int op);
    int (*sign)(const DevBatch&, hipStream_t);
    int (*verify)(const DevBatch&, hipStream_t);
    int (*keygen)(const DevBatch&, hipStream_t);
    // mid[i] = SHA-256 midstate of public key row i (zeros for SHAKE sets)
    int (*midstates)(const uint8_t* pk, uint32_t pk_stride, uint32_t* mid, uint32_t n,
                     hipStream_t);
    // lanes per WOTS leaf k_ht (sign / keygen) or lanes per request k_vht
    // (verify) would use for a batch of n requests
    int (*split)(int op, uint32_t n);
};

const SetOps* get_set(int pid);

// returned by FamilyOps::sign when a batch does not fit the fused path
// (nothing was launched)
constexpr int kNotFused = 1;

// The fast sets of one hash family (128f, 192f, 256f) signed as one mixed
// batch by one persistent kernel (family_impl.hpp).
struct FamilyOps {
    int pid[3];
    // Signs b[0..2] (one batch per set, in the order of pid; n may be 0).
    int (*sign)(const DevBatch* b, hipStream_t s);
};

// sets the text cuspx_last_error() reports on this thread; returns code
int set_error(int code, const char* msg);

// lanes chosen by the tuning table: per WOTS leaf for a k_ht launch over
// `trees` subtrees (op kOpSign / kOpKeygen), per request for a k_vht launch
// over `trees` requests (kOpVerify); 0 = no entry
int tuned_split(int pid, int op, uint32_t trees);
int num_sets();

}  // namespace spx
