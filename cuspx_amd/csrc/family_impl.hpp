// family_impl.hpp -- a mixed signing batch of the fast parameter sets of one
// hash family (128f / 192f / 256f) as ONE persistent kernel.
//
// Each set's batch keeps its own inputs, outputs and workspace; the prep
// kernels of all sets run first, then k_sign_family pulls tiles from one
// device counter: every set's hypertree tiles (the longest first: 256f, 192f,
// 128f) and then every set's FORS tiles, which fill the tail.  A mixed batch
// so balances across sets inside one launch instead of running one
// persistent kernel per set on forked streams.  Then T_k (batches over 256
// requests) and the WOTS gather per set.  Family TUs instantiate it for
// SHA-256 and SHAKE256.  The kernel runs one lane per WOTS leaf; it takes a
// batch whose sets the split rule runs at one or two lanes per leaf (the two
// measured equal at those sizes: SHAKE256-128f at 4096 requests 39.6 ms
// either way).  Smaller batches, where the rule spreads leaves over four or
// more lanes, keep the per-set path.
#pragma once
#include "set_impl.hpp"

namespace spx {

struct FamilyTiles {
    uint32_t ht[3];    // hypertree tiles per set
    uint32_t fors[3];  // FORS tiles per set
    uint32_t* ctr;     // tile counter (zeroed by a prep kernel)
};

template <class P0, int J0, class P1, int J1, class P2, int J2>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(P0::hash == kSha2 ? 4 : 3)))
k_sign_family(BatchArgs A0, BatchArgs A1, BatchArgs A2, FamilyTiles T) {
    static_assert(P0::hash == P1::hash && P1::hash == P2::hash, "one hash family");
    static_assert(HtShape<P0, 1>::WG == 256 && HtShape<P1, 1>::WG == 256 &&
                      HtShape<P2, 1>::WG == 256,
                  "hypertree tiles of 256 lanes");
    constexpr int W = max_int(max_int(max_int(HtShape<P0, 1>::lds_words, fors_lds_words<P0>()),
                                      max_int(HtShape<P1, 1>::lds_words, fors_lds_words<P1>())),
                              max_int(HtShape<P2, 1>::lds_words, fors_lds_words<P2>()));
    __shared__ uint32_t tile;
    __shared__ uint32_t smem[W];  // one buffer for every tile kind of every set
#pragma unroll 1
    while (true) {
        if (threadIdx.x == 0) tile = atomicAdd(T.ctr, 1u);
        __syncthreads();
        uint32_t t = tile;
        __syncthreads();  // every lane has the tile (and left the previous one)
        if (t < T.ht[2]) {
            ht_tile<P2, 0, 1>(A2, t, smem);
            continue;
        }
        t -= T.ht[2];
        if (t < T.ht[1]) {
            ht_tile<P1, 0, 1>(A1, t, smem);
            continue;
        }
        t -= T.ht[1];
        if (t < T.ht[0]) {
            ht_tile<P0, 0, 1>(A0, t, smem);
            continue;
        }
        t -= T.ht[0];
        if (t < T.fors[2]) {
            fors_tile<P2, J2>(A2, t, smem);
            continue;
        }
        t -= T.fors[2];
        if (t < T.fors[1]) {
            fors_tile<P1, J1>(A1, t, smem);
            continue;
        }
        t -= T.fors[1];
        if (t < T.fors[0]) {
            fors_tile<P0, J0>(A0, t, smem);
            continue;
        }
        break;
    }
}

template <class P0, int J0, class P1, int J1, class P2, int J2>
struct FamilyImpl {
    static int pid_[3];

    template <class P>
    static void args(const DevBatch& b, BatchArgs& A, int pid) {
        A = BatchArgs{};
        A.sk = b.sk;
        A.sk_stride = b.sk_stride;
        A.msg = b.msg;
        A.moff = b.moff;
        A.mlen = b.mlen;
        A.optrand = b.optrand;
        A.key_idx = b.key_idx;
        A.key_mid = b.key_mid;
        A.sig = b.sig;
        A.sig_stride = b.sig_stride;
        A.n_req = b.n;
        Layout<P>::carve(pid, A, (uint8_t*)b.ws, b.n, kOpSign, 1);
    }
    template <class P, int J>
    static uint32_t fors_tiles(uint32_t n) {
        return cdiv((uint64_t)n * P::k, 256u >> (P::a - J));
    }
    template <class P>
    static void tail(const BatchArgs& A, uint32_t n, hipStream_t s) {
        if (!n) return;
        if (!A.fors_count) {
            hipLaunchKernelGGL((k_fors_pk<P>), dim3(cdiv(2ull * n, 64)), dim3(64), 0, s, A);
            prof_mark(s, "k_fors_pk");
        }
        hipLaunchKernelGGL((k_wots_gather<P>), dim3(cdiv((uint64_t)n * P::d * P::len, 256)),
                           dim3(256), 0, s, A);
        prof_mark(s, "k_wots_gather");
    }

    static int sign(const DevBatch* b, hipStream_t s) {
        const uint32_t n0 = b[0].n, n1 = b[1].n, n2 = b[2].n;
        if (n0 + n1 + n2 == 0) return kOk;
        if ((n0 && Layout<P0>::split(pid_[0], kOpSign, n0 * P0::d) > 2) ||
            (n1 && Layout<P1>::split(pid_[1], kOpSign, n1 * P1::d) > 2) ||
            (n2 && Layout<P2>::split(pid_[2], kOpSign, n2 * P2::d) > 2))
            return kNotFused;
        BatchArgs A0, A1, A2;
        args<P0>(b[0], A0, pid_[0]);
        args<P1>(b[1], A1, pid_[1]);
        args<P2>(b[2], A2, pid_[2]);
        FamilyTiles T{};
        T.ht[0] = n0 ? Layout<P0>::template ht_blocks<1>(n0 * P0::d) : 0;
        T.ht[1] = n1 ? Layout<P1>::template ht_blocks<1>(n1 * P1::d) : 0;
        T.ht[2] = n2 ? Layout<P2>::template ht_blocks<1>(n2 * P2::d) : 0;
        T.fors[0] = fors_tiles<P0, J0>(n0);
        T.fors[1] = fors_tiles<P1, J1>(n1);
        T.fors[2] = fors_tiles<P2, J2>(n2);
        // the first non-empty set's tile counter, zeroed by its prep kernel
        T.ctr = n0 ? A0.tile_ctr : n1 ? A1.tile_ctr : A2.tile_ctr;
        (void)hipGetLastError();  // report only errors of these launches
        prof_begin(s, "family");
        // prep_grid sets fields of the BatchArgs: call it before the launch
        // that copies them
        const uint32_t g0 = n0 ? prep_grid<P0>(A0, n0, b[0].mlen_bound) : 0;
        const uint32_t g1 = n1 ? prep_grid<P1>(A1, n1, b[1].mlen_bound) : 0;
        const uint32_t g2 = n2 ? prep_grid<P2>(A2, n2, b[2].mlen_bound) : 0;
        if (n0) hipLaunchKernelGGL((k_prep<P0>), dim3(g0), dim3(kPrepWg), 0, s, A0);
        if (n1) hipLaunchKernelGGL((k_prep<P1>), dim3(g1), dim3(kPrepWg), 0, s, A1);
        if (n2) hipLaunchKernelGGL((k_prep<P2>), dim3(g2), dim3(kPrepWg), 0, s, A2);
        prof_mark(s, "k_prep");
        static const uint32_t resident = [] {
            int per_cu = 0, cus = 0, dev = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &per_cu,
                    reinterpret_cast<const void*>(&k_sign_family<P0, J0, P1, J1, P2, J2>), 256,
                    0) != hipSuccess ||
                hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                    hipSuccess)
                return 256u;
            return (uint32_t)((per_cu > 0 ? per_cu : 1) * (cus > 0 ? cus : 1));
        }();
        uint64_t tiles = 0;
        for (int i = 0; i < 3; i++) tiles += (uint64_t)T.ht[i] + T.fors[i];
        const uint32_t grid = tiles < resident ? (uint32_t)tiles : resident;
        hipLaunchKernelGGL((k_sign_family<P0, J0, P1, J1, P2, J2>), dim3(grid), dim3(256), 0, s,
                           A0, A1, A2, T);
        prof_mark(s, "k_sign_family");
        tail<P0>(A0, n0, s);
        tail<P1>(A1, n1, s);
        tail<P2>(A2, n2, s);
        return hipGetLastError() == hipSuccess ? 0 : kErrHip;
    }

    static FamilyOps ops(int p0, int p1, int p2) {
        pid_[0] = p0;
        pid_[1] = p1;
        pid_[2] = p2;
        FamilyOps o{};
        o.pid[0] = p0;
        o.pid[1] = p1;
        o.pid[2] = p2;
        o.sign = &sign;
        return o;
    }
};

template <class P0, int J0, class P1, int J1, class P2, int J2>
int FamilyImpl<P0, J0, P1, J1, P2, J2>::pid_[3] = {-1, -1, -1};

}  // namespace spx

#define SPX_REGISTER_FAMILY(SYM, P0, J0, PID0, P1, J1, PID1, P2, J2, PID2)              \
    extern "C" const spx::FamilyOps* SYM(void) {                                        \
        static const spx::FamilyOps o =                                                 \
            spx::FamilyImpl<P0, J0, P1, J1, P2, J2>::ops(PID0, PID1, PID2);             \
        return &o;                                                                      \
    }
