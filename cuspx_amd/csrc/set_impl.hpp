// set_impl.hpp -- host-side launch sequences for one parameter set.
// Each set_<name>.hip instantiates SetImpl<P, J, JB> and registers its SetOps.
#pragma once
#include "engine.hpp"
#include "profile.hpp"
#include "spx_kernels.hpp"

namespace spx {

static inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// lanes that fill the chip: 1024 SIMDs x 64 lanes x 4 waves
static constexpr uint64_t kFillLanes = 1024ull * 64 * 4;

// workspace carve-up shared by sign / verify / keygen
template <class P>
struct Layout {
    static constexpr uint64_t al(uint64_t x) { return (x + 255) & ~uint64_t(255); }

    // lanes per WOTS leaf for a k_ht launch over `trees` subtrees: a tuned
    // value when the tuning table has one, else double S while the launch
    // would still fill less than half the chip (S = 1 for large batches).
    // Subtrees too wide for one workgroup are split into parts, so every
    // set takes S up to 8.
    static constexpr int max_split() { return 16; }
    static int split(int pid, int op, uint32_t trees) {
        const int t = tuned_split(pid, op, trees);
        if (t == 1 || t == 2 || t == 4 || t == 8 || t == 16) return t;
        const uint64_t leaves = (uint64_t)trees << P::th;
        int S = 1;
        while (S < max_split() && leaves * S * 2 <= kFillLanes) S *= 2;
        return S;
    }
    // Verify kernel width: G lanes per request on average (k_vht runs
    // 256 / G requests per workgroup and deals all their chains over its 256
    // lanes).  A tuned value when the table has one, else G = 16 doubled
    // while the launch would fill under an eighth of the chip: wider groups
    // shorten each layer's chain phase down to the longest chain, narrower
    // ones keep more requests per lane.
    // G = 1 selects speculative verification (k_vspec + k_vht_spec): every
    // chain from every start up front, then T_len and the climb per layer;
    // 16x the chain work for half the dependent path, so batches up to
    // kSpecMax requests only (the workspace keeps candidates for that many).
    static constexpr int vmin() { return 16; }
    static constexpr int vmax() { return P::len > 64 ? 128 : 64; }
    static constexpr uint32_t kSpecMax = 256;
    static constexpr uint32_t kSpecDefault = 64;  // built-in choice without a table
    // sign batches up to this size run T_k inside the FORS tiles
    static constexpr uint32_t kForsPkInline = 256;
    static bool vvalid(int g, uint32_t n) {
        if (g == 1) return n <= kSpecMax;
        return g >= vmin() && g <= vmax() && (g & (g - 1)) == 0;
    }
    // candidate chain ends of min(n, kSpecMax) requests: monotone in n, so a
    // workspace sized for a batch also holds every smaller batch's layout
    static uint64_t spec_bytes(uint32_t n) {
        return (uint64_t)(n < kSpecMax ? n : kSpecMax) * P::d * P::len * 16 * P::n;
    }
    static int vsplit(int pid, uint32_t n) {
        const int t = tuned_split(pid, kOpVerify, n);
        if (vvalid(t, n)) return t;
        if (n <= kSpecDefault) return 1;
        int G = vmin();
        while (G < vmax() && (uint64_t)n * G * 8 <= kFillLanes) G *= 2;
        return G;
    }
    // most parts any split gives one subtree (1 when no split subtree exists)
    static constexpr int max_parts() { return HtShape<P, 16>::PARTS; }
    template <int S>
    static uint32_t ht_blocks(uint32_t trees) {
        return HtShape<P, S>::PARTS > 1 ? trees * HtShape<P, S>::PARTS
                                        : cdiv(trees, HtShape<P, S>::TPW);
    }
    template <int S>
    static uint64_t tl_cols(uint32_t trees) {
        return HtShape<P, S>::park ? (uint64_t)ht_blocks<S>(trees) * HtShape<P, S>::NL : 0;
    }
    static uint64_t tl_cols(uint32_t trees, int S) {
        switch (S) {
            case 1: return tl_cols<1>(trees);
            case 2: return tl_cols<2>(trees);
            case 4: return tl_cols<4>(trees);
            case 8: return tl_cols<8>(trees);
            default: return tl_cols<16>(trees);
        }
    }
    static uint64_t bytes(int pid, uint32_t n, int op) {
        uint64_t b = al((uint64_t)n * sizeof(Plan));
        b += al((uint64_t)n * P::k * P::n);   // fors roots
        b += al((uint64_t)n * P::n);          // fors pk
        if (op == kOpSign) b += al((uint64_t)n * 4);  // fors root counters
        if (op == kOpVerify) b += al(spec_bytes(n));              // speculative candidates
        if (op == kOpSign) {
            b += al((uint64_t)n * P::d * P::n);                    // layer roots
            b += al((uint64_t)n * P::d * P::len * 16 * P::n);      // chain capture
            b += al(4);                                            // fused-kernel tile counter
        }
        if (op != kOpVerify) {
            // sized for every split, so a workspace stays valid whichever S runs
            const uint32_t trees = op == kOpSign ? n * P::d : n;
            uint64_t cols = 0;
            for (int S = 1; S <= 16; S *= 2) cols = tl_cols(trees, S) > cols ? tl_cols(trees, S) : cols;
            b += al(cols * P::len * P::nw * 4);
            if (max_parts() > 1) {
                b += al((uint64_t)trees * max_parts() * P::nw * 4);  // part roots
                b += al((uint64_t)trees * 4);                       // part counters
            }
        }
        (void)pid;
        return b + 256;
    }
    static void carve(int pid, BatchArgs& A, uint8_t* ws, uint32_t n, int op, int S) {
        uint8_t* p = (uint8_t*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
        A.plan = (Plan*)p;
        p += al((uint64_t)n * sizeof(Plan));
        A.fors_roots = p;
        p += al((uint64_t)n * P::k * P::n);
        A.fors_pk = p;
        p += al((uint64_t)n * P::n);
        A.fors_count = nullptr;
        if (op == kOpSign) {
            // T_k in the FORS tiles for small batches only (see fors_root_done)
            if (n <= kForsPkInline) A.fors_count = (uint32_t*)p;
            p += al((uint64_t)n * 4);
        }
        A.roots = nullptr;
        A.cap = nullptr;
        A.tl_buf = nullptr;
        A.tl_lanes = 0;
        A.ht_parts = nullptr;
        A.ht_count = nullptr;
        A.tile_ctr = nullptr;
        if (op == kOpVerify) {
            A.cap = p;
            p += al(spec_bytes(n));
        }
        if (op == kOpSign) {
            A.roots = p;
            p += al((uint64_t)n * P::d * P::n);
            A.cap = p;
            p += al((uint64_t)n * P::d * P::len * 16 * P::n);
            A.tile_ctr = (uint32_t*)p;
            p += al(4);
        }
        if (op != kOpVerify) {
            const uint32_t trees = op == kOpSign ? n * P::d : n;
            A.tl_buf = (uint32_t*)p;
            A.tl_lanes = (uint32_t)tl_cols(trees, S);
            uint64_t cols = 0;
            for (int s2 = 1; s2 <= 16; s2 *= 2) cols = tl_cols(trees, s2) > cols ? tl_cols(trees, s2) : cols;
            p += al(cols * P::len * P::nw * 4);
            if (max_parts() > 1) {
                A.ht_parts = (uint32_t*)p;
                p += al((uint64_t)trees * max_parts() * P::nw * 4);
                A.ht_count = (uint32_t*)p;
                p += al((uint64_t)trees * 4);
            }
        }
        (void)pid;
    }
};

// J: FORS leaves per lane (2^(a-J) lanes per tree) in the tiles of split
// S >= 4 (small batches: short tiles keep FORS off the critical path; one
// SHA2-192f request signs in 0.41 ms at J against 0.70 at JB) and in the
// unfused k_fors; JB: the same in the fused launch at S = 1 and 2 (larger
// batches, where longer tiles spend less on tile set-up and barriers:
// SHA2-192f at 1024 requests, S = 2, 7.07 -> 6.63 ms)
template <class P, int J, int JB>
struct SetImpl {
    using Lay = Layout<P>;
    template <int S>
    static constexpr int fors_j() { return S <= 2 ? JB : J; }

    static int pid_;
    static const char* set_name_;
    static size_t workspace(uint32_t n, int op) { return (size_t)Lay::bytes(pid_, n, op); }

    template <int MODE, int S>
    static void launch_ht(const BatchArgs& A, uint32_t trees, hipStream_t s) {
        if constexpr (HtShape<P, S>::WG <= 1024) {
            hipLaunchKernelGGL((k_ht<P, MODE, S>), dim3(Lay::template ht_blocks<S>(trees)),
                               dim3(HtShape<P, S>::WG), 0, s, A);
        }
    }
    // FORS and hypertree tiles in one persistent launch (when the split's
    // hypertree tiles are 256-lane workgroups); false = not applicable
    template <int S>
    static bool launch_fused(const BatchArgs& A, uint32_t n, hipStream_t s) {
        if constexpr (HtShape<P, S>::WG == 256) {
            constexpr int FJ = fors_j<S>();
            static const uint32_t resident = [] {
                int per_cu = 0, cus = 0, dev = 0;
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                        &per_cu, reinterpret_cast<const void*>(&k_sign_fused<P, FJ, S>), 256, 0) !=
                        hipSuccess ||
                    hipGetDevice(&dev) != hipSuccess ||
                    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                        hipSuccess)
                    return 256u;
                return (uint32_t)((per_cu > 0 ? per_cu : 1) * (cus > 0 ? cus : 1));
            }();
            constexpr uint32_t fors_tpw = 256u >> (P::a - FJ);
            const uint32_t n_ht = Lay::template ht_blocks<S>(n * P::d);
            const uint32_t n_fors = cdiv((uint64_t)n * P::k, fors_tpw);
            const uint32_t grid = n_ht + n_fors < resident ? n_ht + n_fors : resident;
            hipLaunchKernelGGL((k_sign_fused<P, FJ, S>), dim3(grid), dim3(256), 0, s, A, n_ht, n_fors);
            return true;
        }
        return false;
    }
    static bool launch_fused(const BatchArgs& A, uint32_t n, int S, hipStream_t s) {
        switch (S) {
            case 1: return launch_fused<1>(A, n, s);
            case 2: return launch_fused<2>(A, n, s);
            case 4: return launch_fused<4>(A, n, s);
            case 8: return launch_fused<8>(A, n, s);
            default: return launch_fused<16>(A, n, s);
        }
    }

    template <int MODE>
    static void launch_ht(const BatchArgs& A, uint32_t trees, int S, hipStream_t s) {
        switch (S) {
            case 1: launch_ht<MODE, 1>(A, trees, s); break;
            case 2: launch_ht<MODE, 2>(A, trees, s); break;
            case 4: launch_ht<MODE, 4>(A, trees, s); break;
            case 8: launch_ht<MODE, 8>(A, trees, s); break;
            default: launch_ht<MODE, 16>(A, trees, s); break;
        }
    }

    template <int G>
    static void launch_vht(const BatchArgs& A, uint32_t n, hipStream_t s) {
        hipLaunchKernelGGL((k_vht<P, G>), dim3(cdiv(n, 256 / G)), dim3(256), 0, s, A);
    }
    static int sign(const DevBatch& b, hipStream_t s) {
        BatchArgs A{};
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
        const uint32_t n = b.n;
        const int S = Lay::split(pid_, kOpSign, n * P::d);
        Lay::carve(pid_, A, (uint8_t*)b.ws, b.n, kOpSign, S);
        constexpr uint32_t fors_tpw = 256u >> (P::a - J);
        (void)hipGetLastError();  // report only errors of these launches
        prof_begin(s, set_name_);
        const uint32_t prep_grid_wgs = prep_grid<P>(A, n, b.mlen_bound);  // sets fields of A
        hipLaunchKernelGGL((k_prep<P>), dim3(prep_grid_wgs), dim3(kPrepWg), 0, s, A);
        prof_mark(s, "k_prep");
        // one persistent launch for FORS + hypertree tiles: FORS then overlaps
        // the hypertree instead of running before it
        // (small batches: T_k runs on the lane that stores a request's last
        // FORS root, inside the FORS tiles)
        if (launch_fused(A, n, S, s)) {
            prof_mark(s, "k_sign_fused");
            if (!A.fors_count) {
                hipLaunchKernelGGL((k_fors_pk<P>), dim3(cdiv(2ull * n, 64)), dim3(64), 0, s, A);
                prof_mark(s, "k_fors_pk");
            }
        } else {
            hipLaunchKernelGGL((k_fors<P, J>), dim3(cdiv((uint64_t)n * P::k, fors_tpw)), dim3(256),
                               0, s, A);
            prof_mark(s, "k_fors");
            if (!A.fors_count) {
                hipLaunchKernelGGL((k_fors_pk<P>), dim3(cdiv(2ull * n, 64)), dim3(64), 0, s, A);
                prof_mark(s, "k_fors_pk");
            }
            launch_ht<0>(A, n * P::d, S, s);
            prof_mark(s, "k_ht");
        }
        hipLaunchKernelGGL((k_wots_gather<P>), dim3(cdiv((uint64_t)n * P::d * P::len, 256)),
                           dim3(256), 0, s, A);
        prof_mark(s, "k_wots_gather");
        return hipGetLastError() == hipSuccess ? 0 : kErrHip;
    }

    static int verify(const DevBatch& b, hipStream_t s) {
        BatchArgs A{};
        A.pk = b.pk;
        A.pk_stride = b.pk_stride;
        A.key_idx = b.key_idx;
        A.key_mid = b.key_mid;
        A.msg = b.msg;
        A.moff = b.moff;
        A.mlen = b.mlen;
        A.sig = b.sig;
        A.sig_stride = b.sig_stride;
        A.status = b.status;
        A.n_req = b.n;
        Lay::carve(pid_, A, (uint8_t*)b.ws, b.n, kOpVerify, 1);
        const uint32_t n = b.n;
        const int G = Lay::vsplit(pid_, n);
        (void)hipGetLastError();  // report only errors of these launches
        prof_begin(s, set_name_);
        const uint32_t prep_grid_wgs = prep_grid<P>(A, n, b.mlen_bound);  // sets fields of A
        hipLaunchKernelGGL((k_vprep<P>), dim3(prep_grid_wgs), dim3(kPrepWg), 0, s, A);
        prof_mark(s, "k_vprep");
        hipLaunchKernelGGL((k_vfors<P>), dim3(cdiv((uint64_t)n * P::k, 64)), dim3(64), 0, s, A);
        prof_mark(s, "k_vfors");
        hipLaunchKernelGGL((k_fors_pk<P>), dim3(cdiv(2ull * n, 64)), dim3(64), 0, s, A);
        prof_mark(s, "k_fors_pk");
        if (G == 1) {
            hipLaunchKernelGGL((k_vspec<P>), dim3(cdiv((uint64_t)n * P::d * P::len * 8, 256)),
                               dim3(256), 0, s, A);
            prof_mark(s, "k_vspec");
            hipLaunchKernelGGL((k_vht_spec<P>), dim3(n), dim3(64), 0, s, A);
        } else if (G == 16) launch_vht<16>(A, n, s);
        else if (G == 32) launch_vht<32>(A, n, s);
        else if (G == 64) launch_vht<64>(A, n, s);
        else if constexpr (Lay::vmax() >= 128) launch_vht<128>(A, n, s);
        prof_mark(s, "k_vht");
        return hipGetLastError() == hipSuccess ? 0 : kErrHip;
    }

    static int keygen(const DevBatch& b, hipStream_t s) {
        BatchArgs A{};
        A.seed = b.seed;
        A.seed_stride = b.seed_stride;
        A.pk_out = b.pk_out;
        A.sk_out = b.sk_out;
        A.n_req = b.n;
        const uint32_t n = b.n;
        const int S = Lay::split(pid_, kOpKeygen, n);
        Lay::carve(pid_, A, (uint8_t*)b.ws, b.n, kOpKeygen, S);
        (void)hipGetLastError();  // report only errors of these launches
        prof_begin(s, set_name_);
        hipLaunchKernelGGL((k_kprep<P>), dim3(cdiv(n, 64)), dim3(64), 0, s, A);
        prof_mark(s, "k_kprep");
        launch_ht<1>(A, n, S, s);
        prof_mark(s, "k_ht_keygen");
        return hipGetLastError() == hipSuccess ? 0 : kErrHip;
    }

    static int midstates(const uint8_t* pk, uint32_t pk_stride, uint32_t* mid, uint32_t n,
                         hipStream_t s) {
        if (n == 0) return 0;
        (void)hipGetLastError();
        hipLaunchKernelGGL((k_midstate<P>), dim3(cdiv(n, 64)), dim3(64), 0, s, pk, pk_stride, mid,
                           n);
        return hipGetLastError() == hipSuccess ? 0 : kErrHip;
    }

    static int split(int op, uint32_t n) {
        if (op == kOpVerify) return Lay::vsplit(pid_, n);
        return Lay::split(pid_, op, op == kOpSign ? n * P::d : n);
    }

    static SetOps ops(const char* name, int pid) {
        pid_ = pid;
        set_name_ = name;
        SetOps o{};
        o.name = name;
        o.pid = pid;
        o.hash = P::hash;
        o.n = P::n;
        o.full_h = P::full_h;
        o.d = P::d;
        o.a = P::a;
        o.k = P::k;
        o.len = P::len;
        o.sig_bytes = P::sig_bytes;
        o.pk_bytes = P::pk_bytes;
        o.sk_bytes = P::sk_bytes;
        o.workspace = &workspace;
        o.sign = &sign;
        o.verify = &verify;
        o.keygen = &keygen;
        o.midstates = &midstates;
        o.split = &split;
        return o;
    }
};

template <class P, int J, int JB>
int SetImpl<P, J, JB>::pid_ = -1;
template <class P, int J, int JB>
const char* SetImpl<P, J, JB>::set_name_ = "";

}  // namespace spx

#define SPX_REGISTER_SET(SYM, NAME, PID, PARAMS, J, JB)                    \
    extern "C" const spx::SetOps* SYM(void) {                              \
        static const spx::SetOps o = spx::SetImpl<PARAMS, J, JB>::ops(NAME, PID); \
        return &o;                                                         \
    }
