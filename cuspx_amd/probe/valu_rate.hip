// VALU issue-rate probe for gfx950.
//
// Each kernel repeats one block of independent VALU instructions on fixed
// registers (v32-v55, declared clobbered) in 8192 workgroups of 256 lanes
// (32 waves per SIMD over the launch) and reports wave-instructions per SIMD
// per ns.  The hash kernels are VALU-issue bound, so these rates set their
// ceiling: which instruction forms a compression is written in decides how
// close it can come to the chip's 32-bit lane rate.  Three groups:
//   * one instruction form per kernel (8 independent destinations);
//   * v_bitop3_b32 with its three sources in chosen VGPR banks (v mod 4);
//   * mixes of 4-cycle and 2-cycle forms in SHA-256-like proportions.
// No memory is touched inside the loop; each lane writes one word at the end.
//
//     ./cuspx_amd/valu_rate
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#define CLOBBER                                                                                 \
    "v32", "v33", "v34", "v35", "v36", "v37", "v38", "v39", "v40", "v41", "v42", "v43", "v44",  \
        "v45", "v46", "v47", "v48", "v49", "v50", "v51", "v52", "v53", "v54", "v55", "vcc"

#define KERNEL(NAME, BODY)                                                                 \
    __global__ __launch_bounds__(256) void NAME(uint32_t* out, uint32_t iters) {           \
        const uint32_t t = threadIdx.x + blockIdx.x * 256u;                                \
        asm volatile(                                                                      \
            "v_mov_b32 v32, %0\n v_mov_b32 v33, %0\n v_mov_b32 v34, %0\n v_mov_b32 v35, %0\n" \
            " v_mov_b32 v36, %0\n v_mov_b32 v40, %0\n v_mov_b32 v41, %0"                    \
            :                                                                              \
            : "v"(t)                                                                       \
            : CLOBBER);                                                                    \
        for (uint32_t i = 0; i < iters; i++) {                                             \
            _Pragma("unroll") for (int r = 0; r < 8; r++) asm volatile(BODY ::: CLOBBER);  \
        }                                                                                  \
        uint32_t x;                                                                        \
        asm volatile("v_mov_b32 %0, v41" : "=v"(x)::CLOBBER);                              \
        out[t] = x;                                                                        \
    }

// 8 independent destinations v41..v55 (odd banks and even banks mixed)
#define X8(OP, SRC)                                                                      \
    OP " v41, " SRC "\n " OP " v42, " SRC "\n " OP " v43, " SRC "\n " OP " v45, " SRC     \
       "\n " OP " v46, " SRC "\n " OP " v47, " SRC "\n " OP " v49, " SRC "\n " OP " v50, " SRC

KERNEL(k_xor, X8("v_xor_b32_e32", "v32, v33"))
KERNEL(k_xor64, X8("v_xor_b32_e64", "v32, v33"))
KERNEL(k_add, X8("v_add_u32_e32", "v32, v33"))
KERNEL(k_sub, X8("v_sub_u32_e32", "v32, v33"))
KERNEL(k_and, X8("v_and_b32_e32", "v32, v33"))
KERNEL(k_or, X8("v_or_b32_e32", "v32, v33"))
KERNEL(k_lshr, X8("v_lshrrev_b32_e32", "7, v33"))
KERNEL(k_mov, X8("v_mov_b32", "v32"))
KERNEL(k_bitop3, X8("v_bitop3_b32", "v32, v33, v34 bitop3:0x96"))
KERNEL(k_add3, X8("v_add3_u32", "v32, v33, v34"))
KERNEL(k_align, X8("v_alignbit_b32", "v32, v32, 7"))
KERNEL(k_alignbyte, X8("v_alignbyte_b32", "v32, v33, 1"))
KERNEL(k_perm, X8("v_perm_b32", "v32, v33, v34"))
KERNEL(k_or3, X8("v_or3_b32", "v32, v33, v34"))
KERNEL(k_bfi, X8("v_bfi_b32", "v32, v33, v34"))
KERNEL(k_lshl_or, X8("v_lshl_or_b32", "v32, 3, v33"))
KERNEL(k_lshl_add, X8("v_lshl_add_u32", "v32, 3, v33"))
KERNEL(k_add_co, X8("v_add_co_u32_e32", "vcc, v32, v33"))

KERNEL(k_b3_012, X8("v_bitop3_b32", "v32, v33, v34 bitop3:0x96"))
KERNEL(k_b3_001, X8("v_bitop3_b32", "v32, v36, v33 bitop3:0x96"))
KERNEL(k_b3_000, X8("v_bitop3_b32", "v32, v36, v40 bitop3:0x96"))
KERNEL(k_b3_same, X8("v_bitop3_b32", "v32, v33, v33 bitop3:0x96"))

#define S_(D) "v_alignbit_b32 " D ", v32, v32, 7\n "
#define A_(D) "v_add3_u32 " D ", v32, v33, v34\n "
#define F_(D) "v_bitop3_b32 " D ", v32, v33, v34 bitop3:0x96\n "
#define V_(D) "v_add_u32_e32 " D ", v32, v33\n "
KERNEL(k_mix_sf, S_("v41") F_("v42") S_("v43") F_("v45") S_("v46") F_("v47") S_("v49")
                     F_("v50"))
KERNEL(k_mix_ssff, S_("v41") S_("v42") F_("v43") F_("v45") S_("v46") S_("v47") F_("v49")
                       F_("v50"))
KERNEL(k_mix_sfff, S_("v41") F_("v42") F_("v43") F_("v45") S_("v46") F_("v47") F_("v49")
                       F_("v50"))
// 7 four-cycle : 5 two-cycle, the proportion of the SHA-256 F loop (717 : 506)
KERNEL(k_mix_sha, S_("v41") F_("v42") S_("v43") A_("v45") F_("v46") S_("v47") V_("v49")
                      A_("v50") S_("v51") F_("v53") S_("v54") V_("v55"))
// 1 four-cycle : 2 two-cycle, the proportion of a Keccak-f round (58 : 122)
KERNEL(k_mix_keccak, S_("v41") F_("v42") F_("v43") S_("v45") F_("v46") F_("v47") S_("v49")
                         F_("v50") F_("v51") S_("v53") F_("v54") F_("v55"))
// 3 four-cycle : 7 two-cycle, a bit-interleaved Keccak-f round (52 : 122)
KERNEL(k_mix_keccak_il, S_("v41") F_("v42") F_("v43") S_("v45") F_("v46") F_("v47")
                            S_("v49") F_("v50") F_("v51") F_("v53"))

#define CHECK(x)                                                                     \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) {                                                      \
            fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                  \
            return 1;                                                                \
        }                                                                            \
    } while (0)

int main() {
    const int kBlocks = 8192;
    const uint32_t kIters = 1000;
    uint32_t* out = nullptr;
    CHECK(hipMalloc(&out, (size_t)kBlocks * 256 * sizeof(uint32_t)));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    int cus = 0, clk_khz = 0;
    CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
    CHECK(hipDeviceGetAttribute(&clk_khz, hipDeviceAttributeClockRate, 0));
    const double simds = 4.0 * cus;
    struct K {
        const char* name;
        void (*fn)(uint32_t*, uint32_t);
        int insts;  // instructions per block
    } ks[] = {
        {"v_xor_b32_e32", k_xor, 8},          {"v_xor_b32_e64", k_xor64, 8},
        {"v_add_u32_e32", k_add, 8},          {"v_sub_u32_e32", k_sub, 8},
        {"v_and_b32_e32", k_and, 8},          {"v_or_b32_e32", k_or, 8},
        {"v_lshrrev_b32_e32", k_lshr, 8},     {"v_mov_b32", k_mov, 8},
        {"v_bitop3_b32", k_bitop3, 8},        {"v_add3_u32", k_add3, 8},
        {"v_alignbit_b32 (rotate)", k_align, 8}, {"v_alignbyte_b32", k_alignbyte, 8},
        {"v_perm_b32", k_perm, 8},            {"v_or3_b32", k_or3, 8},
        {"v_bfi_b32", k_bfi, 8},              {"v_lshl_or_b32", k_lshl_or, 8},
        {"v_lshl_add_u32", k_lshl_add, 8},    {"v_add_co_u32_e32", k_add_co, 8},
        {"bitop3 src banks 0,1,2", k_b3_012, 8}, {"bitop3 src banks 0,0,1", k_b3_001, 8},
        {"bitop3 src banks 0,0,0", k_b3_000, 8}, {"bitop3 src v33 twice", k_b3_same, 8},
        {"mix S F S F ..", k_mix_sf, 8},      {"mix S S F F ..", k_mix_ssff, 8},
        {"mix S F F F ..", k_mix_sfff, 8},    {"mix 7 S : 5 F (SHA-256)", k_mix_sha, 12},
        {"mix 1 S : 2 F (Keccak)", k_mix_keccak, 12},
        {"mix 3 S : 7 F (Keccak, interleaved)", k_mix_keccak_il, 10},
    };
    printf("device: %d CUs, %.0f MHz; S = 4-cycle form, F = 2-cycle form\n", cus,
           clk_khz / 1000.0);
    printf("%-26s %9s %20s %17s %15s\n", "block", "ms", "wave-instr/SIMD/ns", "cycles/instr",
           "lane-instr T/s");
    for (const K& k : ks) {
        float ms = 0;
        for (int rep = 0; rep < 2; rep++) {  // first launch warms clocks and caches
            CHECK(hipEventRecord(e0, 0));
            hipLaunchKernelGGL(k.fn, dim3(kBlocks), dim3(256), 0, 0, out, kIters);
            CHECK(hipGetLastError());
            CHECK(hipEventRecord(e1, 0));
            CHECK(hipEventSynchronize(e1));
            CHECK(hipEventElapsedTime(&ms, e0, e1));
        }
        const double winstr = (double)kBlocks * 4 * kIters * 8 * k.insts;
        const double rate = winstr / simds / (ms * 1e6);
        printf("%-26s %9.3f %20.4f %17.2f %15.2f\n", k.name, ms, rate,
               clk_khz / 1e6 / rate, winstr * 64 / (ms * 1e-3) / 1e12);
    }
    CHECK(hipFree(out));
    return 0;
}
