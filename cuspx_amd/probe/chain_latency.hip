// Chain-step latency probe: how long one lane takes per WOTS chain step (one
// F = one SHA-256 compression, or one Keccak-f[1600] permutation) as the
// number of waves per SIMD grows, and whether interleaving K independent
// chains per lane (instruction-level parallelism) shortens a step.  Uses the
// library's own F (cuspx_amd/csrc/sha2_dev.hpp, shake_dev.hpp).  Then the
// latency of a serial multi-block hash on one lane against a lane pair.  Each
// workgroup is 256 lanes; 256 workgroups put one wave on every SIMD of a
// 256-CU device.
//
//     ./cuspx_amd/chain_latency
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "spx_kernels.hpp"

using namespace spx;

// A serial hash of 140 payload words held in LDS (T_len of SHA2-128f: 10
// compressions) on one lane (thash_sha2_rd) or one lane pair
// (thash_sha2_pair), repeated with each result folded into the next payload;
// one such lane (pair) per 64-lane workgroup
template <bool PAIR>
__global__ __launch_bounds__(64) void k_tlen(uint32_t* out, int iters) {
    constexpr int M = 140;
    __shared__ uint32_t pk[M];
    for (int i = threadIdx.x; i < M; i += 64) pk[i] = (uint32_t)i * 0x01000193u ^ blockIdx.x;
    __syncthreads();
    uint32_t mid[8];
    for (int i = 0; i < 8; i++) mid[i] = blockIdx.x * 0x9e3779b9u + i;
    const Adr a = adr_make(0, blockIdx.x, kWotsPk);
    uint32_t node[4] = {0, 0, 0, 0};
    const int half = threadIdx.x & 1;
    if (threadIdx.x < (PAIR ? 2 : 1)) {
#pragma unroll 1
        for (int s = 0; s < iters; s++) {
            const uint32_t f = node[0];
            if constexpr (PAIR) {
                thash_sha2_pair<4, M>(mid, a, [&](int j) {
                    const int i = j + half;
                    return pk[i < 0 ? 0 : i < M ? i : M - 1] ^ f;
                }, node);
            } else {
                thash_sha2_rd<4, M>(mid, a, [&](int i) { return pk[i] ^ f; }, node);
            }
        }
    }
    out[threadIdx.x + blockIdx.x * 64u] = node[0] ^ node[1] ^ node[2] ^ node[3];
}

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
    // serial T_len: one lane or lane pair per workgroup, 256 workgroups
    printf("\n%-34s %10s %22s\n", "serial hash (140 words, 10 cmp)", "ms", "us per compression");
    struct T {
        const char* name;
        void (*fn)(uint32_t*, int);
    } ts[] = {{"T_len, one lane", k_tlen<false>}, {"T_len, lane pair", k_tlen<true>}};
    for (const T& t : ts) {
        const int iters = 40;
        float ms = 0;
        for (int rep = 0; rep < 2; rep++) {
            CHECK(hipEventRecord(e0, 0));
            hipLaunchKernelGGL(t.fn, dim3(256), dim3(64), 0, 0, out, iters);
            CHECK(hipGetLastError());
            CHECK(hipEventRecord(e1, 0));
            CHECK(hipEventSynchronize(e1));
            CHECK(hipEventElapsedTime(&ms, e0, e1));
        }
        printf("%-34s %10.3f %22.3f\n", t.name, ms, ms * 1e3 / iters / 10);
    }
    CHECK(hipFree(out));
    return 0;
}
