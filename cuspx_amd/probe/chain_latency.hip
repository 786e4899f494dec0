// Chain-step latency probe: how long one lane takes per WOTS chain step (one
// F = one SHA-256 compression, or one Keccak-f[1600] permutation) as the
// number of waves per SIMD grows, and whether interleaving K independent
// chains per lane (instruction-level parallelism) shortens a step.  Uses the
// library's own F (cuspx_amd/csrc/sha2_dev.hpp, shake_dev.hpp).  Each
// workgroup is 256 lanes; 256 workgroups put one wave on every SIMD of a
// 256-CU device.
//
//     ./cuspx_amd/chain_latency
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "sha2_dev.hpp"
#include "shake_dev.hpp"

using namespace spx;

template <class H, int K>
__global__ __launch_bounds__(256) void k_chain(uint32_t* out, int iters) {
    const uint32_t t = threadIdx.x + blockIdx.x * 256u;
    typename H::Key key;
    if constexpr (H::hash == kSha2) {
        for (int i = 0; i < 8; i++) key.mid[i] = t * 0x9e3779b9u + i;
        for (int i = 0; i < 4; i++) key.sk[i] = t + i;
    } else {
        for (int i = 0; i < 4; i++) {
            key.pub[i] = t + 3 * i;
            key.sk[i] = t + i;
        }
    }
    Adr a[K];
    uint32_t v[K][4];
    for (int k = 0; k < K; k++) {
        a[k] = adr_make(0, t, 0);
        adr_chain(a[k], k);
        for (int i = 0; i < 4; i++) v[k][i] = t ^ (k * 77 + i);
    }
#pragma unroll 1
    for (int s = 0; s < iters; s++) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            adr_hash(a[k], s & 15);
            H::F(key, a[k], v[k], v[k]);
        }
    }
    uint32_t x = 0;
    for (int k = 0; k < K; k++)
        for (int i = 0; i < 4; i++) x ^= v[k][i];
    out[t] = x;
}

#define CHECK(x)                                                                     \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) {                                                      \
            fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                  \
            return 1;                                                                \
        }                                                                            \
    } while (0)

int main() {
    const int kMaxBlocks = 4096;
    uint32_t* out = nullptr;
    CHECK(hipMalloc(&out, (size_t)kMaxBlocks * 256 * sizeof(uint32_t)));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    struct K {
        const char* name;
        void (*fn)(uint32_t*, int);
        int chains, iters;
    } ks[] = {
        {"SHA-256 F, 1 chain/lane", k_chain<Sha2<16>, 1>, 1, 400},
        {"SHA-256 F, 2 chains/lane", k_chain<Sha2<16>, 2>, 2, 400},
        {"SHA-256 F, 4 chains/lane", k_chain<Sha2<16>, 4>, 4, 400},
        {"Keccak F, 1 chain/lane", k_chain<Shake<16>, 1>, 1, 100},
        {"Keccak F, 2 chains/lane", k_chain<Shake<16>, 2>, 2, 100},
    };
    printf("%-26s %7s %10s %18s %12s\n", "kernel", "blocks", "ms", "us per step/chain",
           "G steps/s");
    for (const K& k : ks) {
        for (int blocks : {256, 1024, 4096}) {
            float ms = 0;
            for (int rep = 0; rep < 2; rep++) {
                CHECK(hipEventRecord(e0, 0));
                hipLaunchKernelGGL(k.fn, dim3(blocks), dim3(256), 0, 0, out, k.iters);
                CHECK(hipGetLastError());
                CHECK(hipEventRecord(e1, 0));
                CHECK(hipEventSynchronize(e1));
                CHECK(hipEventElapsedTime(&ms, e0, e1));
            }
            const double steps = (double)blocks * 256 * k.iters * k.chains;
            printf("%-26s %7d %10.3f %18.3f %12.2f\n", k.name, blocks, ms,
                   ms * 1e3 / k.iters / k.chains, steps / (ms * 1e-3) / 1e9);
        }
    }
    CHECK(hipFree(out));
    return 0;
}
