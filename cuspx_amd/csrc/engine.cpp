// engine.cpp -- C ABI of libcuspx.so: the api.h surface of the reference
// (crypto_sign_* / face_mdp_* / randombytes), runtime-parameterised batch
// entry points with per-item status, device-resident entry points, and the
// kernel-timing hooks.  All device work goes through the HIP kernels of the
// set_*.hip translation units; there is no host fallback for any hash.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <dlfcn.h>

#include <atomic>
#include <condition_variable>
#include <fstream>
#include <functional>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cuspx.h"
#include "drbg.hpp"
#include "engine.hpp"
#include "host_pool.hpp"
#include "profile.hpp"

extern "C" {
const spx::SetOps* cuspx_set_sha2_128s(void);
const spx::SetOps* cuspx_set_sha2_128f(void);
const spx::SetOps* cuspx_set_sha2_192s(void);
const spx::SetOps* cuspx_set_sha2_192f(void);
const spx::SetOps* cuspx_set_sha2_256s(void);
const spx::SetOps* cuspx_set_sha2_256f(void);
const spx::SetOps* cuspx_set_shake256_128s(void);
const spx::SetOps* cuspx_set_shake256_128f(void);
const spx::SetOps* cuspx_set_shake256_192s(void);
const spx::SetOps* cuspx_set_shake256_192f(void);
const spx::SetOps* cuspx_set_shake256_256s(void);
const spx::SetOps* cuspx_set_shake256_256f(void);
const spx::FamilyOps* cuspx_family_sha2_f(void);
const spx::FamilyOps* cuspx_family_shake256_f(void);
}

namespace spx {

static const SetOps* (*const kSetFns[12])(void) = {
    cuspx_set_sha2_128s,     cuspx_set_sha2_128f,     cuspx_set_sha2_192s,
    cuspx_set_sha2_192f,     cuspx_set_sha2_256s,     cuspx_set_sha2_256f,
    cuspx_set_shake256_128s, cuspx_set_shake256_128f, cuspx_set_shake256_192s,
    cuspx_set_shake256_192f, cuspx_set_shake256_256s, cuspx_set_shake256_256f};

int num_sets() { return 12; }

const SetOps* get_set(int pid) {
    if (pid < 0 || pid >= 12) return nullptr;
    return kSetFns[pid]();
}

// ------------------------------------------------------------ tuning table
//
// Text file of `<set> <sign|keygen|verify> <min_units> <lanes>` lines written
// by the autotuner (python -m cuspx_amd.autotune); read once from
// $CUSPX_TUNING or tuning.txt next to libcuspx.so.  sign / keygen: lanes per
// WOTS leaf of k_ht for a launch over T subtrees; verify: lanes per request
// of k_vht for a batch of T requests.  The entry with the largest
// min_units <= T applies.

struct Tuning {
    std::mutex mu;
    bool loaded = false;
    std::map<std::pair<int, int>, std::map<uint32_t, int>> split;
    std::map<std::pair<int, int>, int> forced;  // cuspx_set_split overrides

    void load() {
        std::string path;
        if (const char* env = getenv("CUSPX_TUNING")) {
            path = env;
        } else {
            Dl_info info;
            if (dladdr((void*)&tuned_split, &info) && info.dli_fname) {
                path = info.dli_fname;
                path = path.substr(0, path.find_last_of('/') + 1) + "tuning.txt";
            }
        }
        std::ifstream f(path);
        std::string line;
        while (std::getline(f, line)) {
            if (line.empty() || line[0] == '#') continue;
            std::istringstream ls(line);
            std::string name, op;
            uint32_t min_trees = 0;
            int S = 0;
            if (!(ls >> name >> op >> min_trees >> S)) continue;
            int pid = -1;
            for (int i = 0; i < 12; i++)
                if (name == get_set(i)->name) pid = i;
            const int o = op == "sign" ? kOpSign
                          : op == "keygen" ? kOpKeygen
                          : op == "verify" ? kOpVerify : -1;
            if (pid < 0 || o < 0) continue;
            split[{pid, o}][min_trees] = S;
        }
        loaded = true;
    }
};

static Tuning& tuning() {
    static Tuning t;
    return t;
}

int tuned_split(int pid, int op, uint32_t trees) {
    Tuning& t = tuning();
    std::lock_guard<std::mutex> lk(t.mu);
    if (!t.loaded) t.load();
    auto f = t.forced.find({pid, op});
    if (f != t.forced.end()) return f->second;
    auto it = t.split.find({pid, op});
    if (it == t.split.end()) return 0;
    auto e = it->second.upper_bound(trees);
    if (e == it->second.begin()) return 0;
    return std::prev(e)->second;
}

// ------------------------------------------------------------ error text

static thread_local char g_err[256] = "";

static int fail(int code, const char* fmt, const char* what = "") {
    snprintf(g_err, sizeof(g_err), fmt, what);
    return code;
}

int set_error(int code, const char* msg) { return fail(code, "%s", msg); }

#define HIP_OK(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) return fail(kErrHip, "HIP error: %s", hipGetErrorString(e_)); \
    } while (0)

// ------------------------------------------------------------ the device

// Large host-buffer transfers move in up to kChunks pieces: one per pinned
// staging copy and one per PCIe copy, so the host copy of one piece overlaps
// the DMA of the next.
constexpr uint32_t kChunks = 32;

// One stream of the host-buffer calls with its own device arena and pinned
// staging (both grow-only), and the events that pace its chunked copies.
struct Lane {
    hipStream_t stream = nullptr;
    void* dbuf = nullptr;
    size_t dcap = 0;
    uint8_t* hbuf = nullptr;
    size_t hcap = 0;
    hipEvent_t ev[kChunks] = {};
    hipEvent_t kdone = nullptr;  // after the slice's kernels, before its output copy

    int ensure(size_t dbytes, size_t hbytes) {
        if (!stream) HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        if (!kdone) {
            for (auto& e : ev) HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            HIP_OK(hipEventCreateWithFlags(&kdone, hipEventDisableTiming));
        }
        if (dbytes > dcap) {
            if (dbuf) HIP_OK(hipFree(dbuf));
            dbuf = nullptr;
            dcap = 0;
            HIP_OK(hipMalloc(&dbuf, dbytes));
            dcap = dbytes;
        }
        if (hbytes > hcap) {
            if (hbuf) HIP_OK(hipHostFree(hbuf));
            hbuf = nullptr;
            hcap = 0;
            HIP_OK(hipHostMalloc((void**)&hbuf, hbytes, hipHostMallocDefault));
            hcap = hbytes;
        }
        return kOk;
    }
};

// Lane 0 serves the single-stream host-buffer calls; the multi-stream ones
// (face_ms_mdp_*) cut a batch into up to kLanes slices on lanes 0.., so one
// slice's PCIe copies and host staging overlap another slice's kernels.
constexpr int kLanes = 4;

struct Device {
    std::mutex mu;
    bool ready = false;
    int dev = 0;
    Lane lane[kLanes];
    size_t budget = 0;  // bytes one host-buffer call may use on the device

    int init() {
        if (ready) return kOk;
        HIP_OK(hipGetDevice(&dev));
        HIP_OK(hipStreamCreateWithFlags(&lane[0].stream, hipStreamNonBlocking));
        size_t fr = 0, tot = 0;
        HIP_OK(hipMemGetInfo(&fr, &tot));
        const char* env = getenv("CUSPX_HBM_BUDGET_MB");
        budget = env ? (size_t)atoll(env) << 20 : fr / 2;
        ready = true;
        return kOk;
    }
};

static Device& device() {
    static Device d;
    return d;
}

// ------------------------------------------------------------ settings

static int g_default_pid = -1;
static int g_debug_mode = 1;  // the reference ships with DEBUG_MODE (all_option.h:74)

static int default_pid() {
    if (g_default_pid < 0) {
        g_default_pid = 1;  // sha2-128f
        const char* env = getenv("CUSPX_PARAM");
        if (env) {
            for (int i = 0; i < 12; i++)
                if (strcmp(env, get_set(i)->name) == 0) g_default_pid = i;
        }
    }
    return g_default_pid;
}

static inline uint64_t al16(uint64_t x) { return (x + 15) & ~uint64_t(15); }

// ------------------------------------------------------- profiling hooks

struct Prof {
    std::mutex mu;
    bool on = false;
    struct Mark {
        const char* name;
        hipEvent_t ev;
    };
    std::vector<const char*> call_set;  // parameter set of each call
    std::vector<std::vector<Mark>> calls;
    std::map<hipStream_t, size_t> open;
    std::map<std::string, std::pair<double, int>> acc;
    bool by_set = false;  // key the totals by "set:kernel"
};
static Prof& prof() {
    static Prof p;
    return p;
}

void prof_begin(hipStream_t s, const char* set) {
    Prof& p = prof();
    if (!p.on) return;
    std::lock_guard<std::mutex> lk(p.mu);
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    (void)hipEventRecord(e, s);
    p.calls.push_back({{"", e}});
    p.call_set.push_back(set);
    p.open[s] = p.calls.size() - 1;
}

void prof_mark(hipStream_t s, const char* kernel) {
    Prof& p = prof();
    if (!p.on) return;
    std::lock_guard<std::mutex> lk(p.mu);
    auto it = p.open.find(s);
    if (it == p.open.end()) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    (void)hipEventRecord(e, s);
    p.calls[it->second].push_back({kernel, e});
}

// ------------------------------------------------ host-buffer batch calls

struct MsgPack {
    std::vector<uint64_t> off;  // device offsets, 16-byte aligned
    uint64_t bytes = 0;
    uint32_t longest = 0;
};

// Items [a, b) of `item` bytes from src (stride ss) to dst (stride ds).
static void copy_rows(uint8_t* dst, uint64_t ds, const uint8_t* src, uint64_t ss, uint64_t item,
                      uint32_t a, uint32_t b) {
    if (ds == item && ss == item) {
        memcpy(dst + a * item, src + a * item, (uint64_t)(b - a) * item);
        return;
    }
    for (uint32_t i = a; i < b; i++) memcpy(dst + i * ds, src + i * ss, item);
}

// How a transfer of num items of `item` bytes is cut: pieces of at least
// 1 MiB, at most kChunks of them; pieces(...).first = count, .second = items
// per piece.
static std::pair<uint32_t, uint32_t> pieces(uint32_t num, uint64_t item) {
    const uint64_t total = item * num;
    uint64_t c = total >> 20;
    c = c < 1 ? 1 : c > kChunks ? kChunks : c;
    const uint32_t per = (uint32_t)((num + c - 1) / c);
    return {(num + per - 1) / per, per};
}

// num items of `item` bytes from src (stride ss) to dst (stride ds), split
// over the host pool.
static void copy_items(uint8_t* dst, uint64_t ds, const uint8_t* src, uint64_t ss, uint64_t item,
                       uint32_t num) {
    const auto pc = pieces(num, item);
    HostPool::get().run(pc.first, [&](uint32_t c) {
        const uint32_t a = c * pc.second, b = a + pc.second < num ? a + pc.second : num;
        copy_rows(dst, ds, src, ss, item, a, b);
    });
}

static MsgPack pack_msgs(uint32_t num, const uint32_t* mlen) {
    MsgPack mp;
    mp.off.resize(num);
    uint64_t o = 0;
    for (uint32_t i = 0; i < num; i++) {
        mp.off[i] = o;
        o += al16(mlen[i]);
        mp.longest = mlen[i] > mp.longest ? mlen[i] : mp.longest;
    }
    mp.bytes = o + 16;
    return mp;
}

// One host-buffer slice in flight on a lane: its inputs are staged and
// copied, its kernels and the output copy enqueued.  finish_job waits for the
// lane and moves the outputs into the caller's buffers.
struct Job {
    Lane* lane = nullptr;
    const SetOps* o = nullptr;
    int op = -1;
    uint32_t num = 0;
    uint8_t* out = nullptr;  // sign: signatures; verify: statuses; keygen: pk
    uint64_t out_stride = 0;
    uint8_t* out2 = nullptr;  // keygen: sk
    uint32_t pieces = 0, per = 0;  // sign: output copies, items per copy
    hipEvent_t after = nullptr;    // chained slices: the previous slice's kernels
};

static int finish_job(Job& j) {
    if (!j.lane) return kOk;
    Lane& L = *j.lane;
    j.lane = nullptr;
    const SetOps* o = j.o;
    if (j.op == kOpSign) {
        // each piece is copied out as soon as its event says its DMA is done,
        // while the later pieces are still in flight
        std::atomic<int> err{kOk};
        HostPool::get().run(j.pieces, [&](uint32_t c) {
            if (hipEventSynchronize(L.ev[c]) != hipSuccess) {
                err = kErrHip;
                return;
            }
            const uint32_t a = c * j.per, b = a + j.per < j.num ? a + j.per : j.num;
            copy_rows(j.out, j.out_stride, L.hbuf, o->sig_bytes, o->sig_bytes, a, b);
        });
        HIP_OK(hipStreamSynchronize(L.stream));
        if (err != kOk) return fail(err, "output copy failed%s");
        return kOk;
    }
    HIP_OK(hipStreamSynchronize(L.stream));
    if (j.op == kOpVerify) {
        memcpy(j.out, L.hbuf, 4ull * j.num);
    } else {
        memcpy(j.out, L.hbuf, (uint64_t)j.num * o->pk_bytes);
        memcpy(j.out2, L.hbuf + al16((uint64_t)j.num * o->pk_bytes),
               (uint64_t)j.num * o->sk_bytes);
    }
    return kOk;
}

// Runs a host-buffer batch as `slices` consecutive slices on lanes 0..slices-1
// (one slice: the whole batch on lane 0).  Slice c is staged and enqueued
// while the slices before it already copy and compute.  Unchained, the
// slices' kernels run concurrently on their streams (each verify launch takes
// its own workgroups); chained, slice c's kernels start when slice c - 1's
// have finished, so each slice's output copy overlaps the next slice's
// kernels (signing, whose 17 KB-per-item output only exists once its launch
// ends).  Every lane is finished at the end, in order, so a slice's outputs
// are copied out while later slices still run.  enq(lane, first item, items,
// device bytes allowed, job) enqueues one slice.
template <class Enq>
static int run_slices(uint32_t num, int slices, Enq enq, bool chain = false) {
    if (num == 0) return kOk;
    Device& D = device();
    std::lock_guard<std::mutex> lk(D.mu);
    int rc = D.init();
    if (rc) return rc;
    slices = slices < 1 ? 1 : slices > kLanes ? kLanes : slices;
    const uint32_t per = (num + slices - 1) / slices;
    Job jobs[kLanes];
    int err = kOk;
    for (int c = 0; c < slices && err == kOk; c++) {
        const uint32_t a = c * per, b = a + per < num ? a + per : num;
        if (a >= b) break;
        if (chain && c > 0) jobs[c].after = D.lane[c - 1].kdone;
        err = enq(D.lane[c], a, b - a, D.budget / slices, jobs[c]);
    }
    for (int c = 0; c < slices; c++) {
        rc = finish_job(jobs[c]);
        if (rc && err == kOk) err = rc;
    }
    return err;
}

// sign `num` items on lane L: sk_i = sk + i*sk_stride, message i at msg + moff[i]
static int sign_enqueue(const SetOps* o, Lane& L, size_t budget, uint32_t num,
                        const uint8_t* sk, uint64_t sk_stride, const uint8_t* msg,
                        const uint64_t* moff, const uint32_t* mlen, const uint8_t* optrand,
                        uint8_t* sig, uint64_t sig_stride, Job& job) {
    MsgPack mp = pack_msgs(num, mlen);
    const uint64_t n = o->n;
    const uint64_t in_bytes = al16((uint64_t)num * o->sk_bytes) + mp.bytes + al16(8ull * num) +
                              al16(4ull * num) + al16((uint64_t)num * n);
    const uint64_t out_bytes = al16((uint64_t)num * o->sig_bytes);
    const uint64_t ws = o->workspace(num, kOpSign);
    if (in_bytes + out_bytes + ws + 512 > budget)
        return fail(kErrBudget, "sign batch exceeds the device-memory budget%s");
    int rc = L.ensure(in_bytes + out_bytes + ws + 512, in_bytes > out_bytes ? in_bytes : out_bytes);
    if (rc) return rc;
    uint8_t* h = L.hbuf;
    uint8_t* p = h;
    copy_items(p, o->sk_bytes, sk, sk_stride, o->sk_bytes, num);
    p += al16((uint64_t)num * o->sk_bytes);
    uint8_t* pm = p;
    memset(pm, 0, mp.bytes);
    for (uint32_t i = 0; i < num; i++)  // msg may be null when every message is empty
        if (mlen[i]) memcpy(pm + mp.off[i], msg + moff[i], mlen[i]);
    p += mp.bytes;
    memcpy(p, mp.off.data(), 8ull * num);
    p += al16(8ull * num);
    memcpy(p, mlen, 4ull * num);
    p += al16(4ull * num);
    memcpy(p, optrand, (uint64_t)num * n);
    uint8_t* d = (uint8_t*)L.dbuf;
    HIP_OK(hipMemcpyAsync(d, h, in_bytes, hipMemcpyHostToDevice, L.stream));
    DevBatch b{};
    b.n = num;
    b.sk = d;
    b.sk_stride = o->sk_bytes;
    uint8_t* q = d + al16((uint64_t)num * o->sk_bytes);
    b.msg = q;
    q += mp.bytes;
    b.moff = (const uint64_t*)q;
    q += al16(8ull * num);
    b.mlen = (const uint32_t*)q;
    b.mlen_bound = mp.longest > 0 ? mp.longest : 1;
    q += al16(4ull * num);
    b.optrand = q;
    uint8_t* dsig = d + in_bytes;
    b.sig = dsig;
    b.sig_stride = o->sig_bytes;
    b.ws = dsig + out_bytes;
    if (job.after) HIP_OK(hipStreamWaitEvent(L.stream, job.after, 0));
    rc = o->sign(b, L.stream);
    if (rc) return fail(rc, "kernel launch failed%s");
    HIP_OK(hipEventRecord(L.kdone, L.stream));
    const uint64_t sb = o->sig_bytes;
    const auto pc = pieces(num, sb);
    for (uint32_t c = 0; c < pc.first; c++) {
        const uint64_t a = (uint64_t)c * pc.second;
        const uint64_t cnt = a + pc.second < num ? pc.second : num - a;
        HIP_OK(hipMemcpyAsync(h + a * sb, dsig + a * sb, cnt * sb, hipMemcpyDeviceToHost,
                              L.stream));
        HIP_OK(hipEventRecord(L.ev[c], L.stream));
    }
    job.lane = &L;
    job.o = o;
    job.op = kOpSign;
    job.num = num;
    job.out = sig;
    job.out_stride = sig_stride;
    job.pieces = pc.first;
    job.per = pc.second;
    return kOk;
}

static int sign_host(const SetOps* o, uint32_t num, const uint8_t* sk, uint64_t sk_stride,
                     const uint8_t* msg, const uint64_t* moff, const uint32_t* mlen,
                     const uint8_t* optrand, uint8_t* sig, uint64_t sig_stride, int slices = 1) {
    if (num == 0) return kOk;
    if (!sk || !sig || !mlen || !moff || !optrand) return fail(kErrArg, "null argument%s");
    return run_slices(num, slices, [&](Lane& L, uint32_t a, uint32_t cnt, size_t budget,
                                       Job& job) {
        return sign_enqueue(o, L, budget, cnt, sk + a * sk_stride, sk_stride, msg, moff + a,
                            mlen + a, optrand + (uint64_t)a * o->n, sig + a * sig_stride,
                            sig_stride, job);
    }, /*chain=*/true);
}

static int verify_enqueue(const SetOps* o, Lane& L, size_t budget,
                          uint32_t num, const uint8_t* pk, uint64_t pk_stride, const uint8_t* sig,
                          uint64_t sig_stride, const uint8_t* msg, const uint64_t* moff,
                          const uint32_t* mlen, int32_t* status, Job& job) {
    MsgPack mp = pack_msgs(num, mlen);
    const uint64_t in_bytes = al16((uint64_t)num * o->pk_bytes) +
                              al16((uint64_t)num * o->sig_bytes) + mp.bytes + al16(8ull * num) +
                              al16(4ull * num);
    const uint64_t out_bytes = al16(4ull * num);
    const uint64_t ws = o->workspace(num, kOpVerify);
    if (in_bytes + out_bytes + ws + 512 > budget)
        return fail(kErrBudget, "verify batch exceeds the device-memory budget%s");
    int rc = L.ensure(in_bytes + out_bytes + ws + 512, in_bytes);
    if (rc) return rc;
    uint8_t* h = L.hbuf;
    uint8_t* d = (uint8_t*)L.dbuf;
    uint8_t* p = h;
    copy_items(p, o->pk_bytes, pk, pk_stride, o->pk_bytes, num);
    const uint64_t pk_region = al16((uint64_t)num * o->pk_bytes);
    p += pk_region;
    // the signatures (the bulk of the input): each piece goes to the device
    // as soon as it is staged, while the pool stages the next ones
    const uint64_t sb = o->sig_bytes;
    const auto pc = pieces(num, sb);
    std::atomic<int> err{kOk};
    HostPool::get().run(pc.first, [&](uint32_t c) {
        const uint32_t a = c * pc.second, e = a + pc.second < num ? a + pc.second : num;
        copy_rows(p, sb, sig, sig_stride, sb, a, e);
        if (hipMemcpyAsync(d + pk_region + a * sb, p + a * sb, (uint64_t)(e - a) * sb,
                           hipMemcpyHostToDevice, L.stream) != hipSuccess)
            err = kErrHip;
    });
    if (err != kOk) return fail(err, "input copy failed%s");
    const uint64_t sig_region = al16((uint64_t)num * sb);
    p += sig_region;
    memset(p, 0, mp.bytes);
    for (uint32_t i = 0; i < num; i++)
        if (mlen[i]) memcpy(p + mp.off[i], msg + moff[i], mlen[i]);
    p += mp.bytes;
    memcpy(p, mp.off.data(), 8ull * num);
    p += al16(8ull * num);
    memcpy(p, mlen, 4ull * num);
    const uint64_t tail = pk_region + sig_region;
    HIP_OK(hipMemcpyAsync(d, h, pk_region, hipMemcpyHostToDevice, L.stream));
    HIP_OK(hipMemcpyAsync(d + tail, h + tail, in_bytes - tail, hipMemcpyHostToDevice, L.stream));
    DevBatch b{};
    b.n = num;
    b.pk = d;
    b.pk_stride = o->pk_bytes;
    uint8_t* q = d + al16((uint64_t)num * o->pk_bytes);
    b.sig = q;
    b.sig_stride = o->sig_bytes;
    q += al16((uint64_t)num * o->sig_bytes);
    b.msg = q;
    q += mp.bytes;
    b.moff = (const uint64_t*)q;
    q += al16(8ull * num);
    b.mlen = (const uint32_t*)q;
    b.mlen_bound = mp.longest > 0 ? mp.longest : 1;
    int32_t* dst = (int32_t*)(d + in_bytes);
    b.status = dst;
    b.ws = (uint8_t*)dst + out_bytes;
    rc = o->verify(b, L.stream);
    if (rc) return fail(rc, "kernel launch failed%s");
    HIP_OK(hipMemcpyAsync(h, dst, 4ull * num, hipMemcpyDeviceToHost, L.stream));
    job.lane = &L;
    job.o = o;
    job.op = kOpVerify;
    job.num = num;
    job.out = reinterpret_cast<uint8_t*>(status);
    return kOk;
}

// status[i] per item; kErrVerify if any item failed
static int verify_host(const SetOps* o, uint32_t num, const uint8_t* pk, uint64_t pk_stride,
                       const uint8_t* sig, uint64_t sig_stride, const uint8_t* msg,
                       const uint64_t* moff, const uint32_t* mlen, int32_t* status,
                       int slices = 1) {
    if (num == 0) return kOk;
    if (!pk || !sig || !mlen || !moff || !status) return fail(kErrArg, "null argument%s");
    const int rc = run_slices(num, slices, [&](Lane& L, uint32_t a, uint32_t cnt, size_t budget,
                                               Job& job) {
        return verify_enqueue(o, L, budget, cnt, pk + a * pk_stride, pk_stride,
                              sig + a * sig_stride, sig_stride, msg, moff + a, mlen + a,
                              status + a, job);
    });
    if (rc) return rc;
    for (uint32_t i = 0; i < num; i++)
        if (status[i] != 0) return kErrVerify;
    return kOk;
}

static int keygen_enqueue(const SetOps* o, Lane& L, size_t budget,
                          uint32_t num, const uint8_t* seeds, uint64_t seed_stride, uint8_t* pk,
                          uint8_t* sk, Job& job) {
    const uint64_t sb = 3ull * o->n;
    const uint64_t in_bytes = al16(num * sb);
    const uint64_t out_bytes = al16((uint64_t)num * o->pk_bytes) + al16((uint64_t)num * o->sk_bytes);
    const uint64_t ws = o->workspace(num, kOpKeygen);
    if (in_bytes + out_bytes + ws + 512 > budget)
        return fail(kErrBudget, "keygen batch exceeds the device-memory budget%s");
    int rc = L.ensure(in_bytes + out_bytes + ws + 512, in_bytes > out_bytes ? in_bytes : out_bytes);
    if (rc) return rc;
    for (uint32_t i = 0; i < num; i++) memcpy(L.hbuf + i * sb, seeds + i * seed_stride, sb);
    uint8_t* d = (uint8_t*)L.dbuf;
    HIP_OK(hipMemcpyAsync(d, L.hbuf, num * sb, hipMemcpyHostToDevice, L.stream));
    DevBatch b{};
    b.n = num;
    b.seed = d;
    b.seed_stride = (uint32_t)sb;
    b.pk_out = d + in_bytes;
    b.sk_out = b.pk_out + al16((uint64_t)num * o->pk_bytes);
    b.ws = d + in_bytes + out_bytes;
    rc = o->keygen(b, L.stream);
    if (rc) return fail(rc, "kernel launch failed%s");
    HIP_OK(hipMemcpyAsync(L.hbuf, b.pk_out, out_bytes, hipMemcpyDeviceToHost, L.stream));
    job.lane = &L;
    job.o = o;
    job.op = kOpKeygen;
    job.num = num;
    job.out = pk;
    job.out2 = sk;
    return kOk;
}

static int keygen_host(const SetOps* o, uint32_t num, const uint8_t* seeds, uint64_t seed_stride,
                       uint8_t* pk, uint8_t* sk, int slices = 1) {
    if (num == 0) return kOk;
    if (!seeds || !pk || !sk) return fail(kErrArg, "null argument%s");
    return run_slices(num, slices, [&](Lane& L, uint32_t a, uint32_t cnt, size_t budget,
                                       Job& job) {
        return keygen_enqueue(o, L, budget, cnt, seeds + a * seed_stride, seed_stride,
                              pk + (uint64_t)a * o->pk_bytes, sk + (uint64_t)a * o->sk_bytes, job);
    });
}

// slices of a multi-stream call: at most 2, of at least 2048 items (sign,
// keygen) or 4096 (verify).  Measured on SHA2-128f host buffers, best of 3,
// slices 1 / 2 / 3 / 4: chained signing slices at 4096 items 17.43 / 17.19 /
// 17.17 / 17.09 ms and at 8192 33.80 / 32.23 / 32.98 / 32.84 ms (each slice's
// output copy hides behind the next slice's kernels, smaller launches lose
// at their tails); concurrent verify slices at 4096 3.60 / 5.02 / 5.39 /
// 5.16 ms and at 8192 6.54 / 5.82 / 8.24 / 8.16 ms (a verify launch is bound
// by its 22-layer dependent path, so it gains from sharing the CUs only when
// each slice still fills them).
static int ms_slices(int op, uint32_t num) {
    const uint32_t s = num / (op == kOpVerify ? 4096u : 2048u);
    return s < 1 ? 1 : s > 2 ? 2 : (int)s;
}

// draw the per-item optrand the way crypto_sign_signature does
// (sign.cu:1950-1954): n DRBG bytes, replaced by 0xad in debug mode
static void draw_optrand(uint8_t* out, uint32_t num, int n) {
    for (uint32_t i = 0; i < num; i++) {
        drbg_bytes(out + (uint64_t)i * n, n);
        if (g_debug_mode) memset(out + (uint64_t)i * n, 0xad, n);
    }
}

// draw keygen seeds the way crypto_sign_keypair does (sign.cu:342-352)
static void draw_seed(uint8_t* out, int n) {
    drbg_bytes(out, 3ull * n);
    if (g_debug_mode)
        for (int i = 0; i < 3 * n; i++) out[i] = (uint8_t)i;
}

// Inputs of the batched calls, which the reference makes on the device
// (dev_crypto_sign_keypair sign.cu:425-436, dev_crypto_sign_signature
// sign.cu:2022-2027): in debug mode seeds 0..3n-1 and optrand 0xad with no
// draw at all, so the host DRBG stream is left untouched; otherwise the
// reference draws from a clock-seeded curand state per thread, and these
// draw 3n / n host DRBG bytes per item instead (reproducible under
// randombytes_init, and the same bytes as that many crypto_sign_keypair /
// crypto_sign_signature calls).
static void mdp_seeds(uint8_t* out, uint32_t num, int n) {
    const uint64_t sb = 3ull * n;
    for (uint32_t i = 0; i < num; i++) {
        if (g_debug_mode)
            for (uint64_t j = 0; j < sb; j++) out[i * sb + j] = (uint8_t)j;
        else
            drbg_bytes(out + i * sb, sb);
    }
}

static void mdp_optrand(uint8_t* out, uint32_t num, int n) {
    if (g_debug_mode) {
        memset(out, 0xad, (uint64_t)num * n);
        return;
    }
    for (uint32_t i = 0; i < num; i++) drbg_bytes(out + (uint64_t)i * n, n);
}

}  // namespace spx

using namespace spx;

// =====================================================================
// C ABI
// =====================================================================

extern "C" {

int cuspx_num_params(void) { return 12; }

int cuspx_param_id(const char* name) {
    if (!name) return kErrArg;
    for (int i = 0; i < 12; i++)
        if (strcmp(name, get_set(i)->name) == 0) return i;
    return kErrParam;
}

int cuspx_get_param_info(int pid, cuspx_param_info* out) {
    const SetOps* o = get_set(pid);
    if (!o) return kErrParam;
    if (!out) return kErrArg;
    memset(out, 0, sizeof(*out));
    out->pid = pid;
    out->hash = o->hash;
    out->n = o->n;
    out->full_h = o->full_h;
    out->d = o->d;
    out->a = o->a;
    out->k = o->k;
    out->len = o->len;
    out->sig_bytes = o->sig_bytes;
    out->pk_bytes = o->pk_bytes;
    out->sk_bytes = o->sk_bytes;
    out->seed_bytes = 3 * o->n;
    snprintf(out->name, sizeof(out->name), "%s", o->name);
    return kOk;
}

int cuspx_set_split(int pid, int op, int lanes_per_leaf) {
    if (!get_set(pid)) return kErrParam;
    const SetOps* o = get_set(pid);
    if (op == kOpVerify) {
        const int g = lanes_per_leaf;
        const int hi = o->len > 64 ? 128 : 64;
        if (g != 0 && g != 1 && !(g >= 16 && g <= hi && (g & (g - 1)) == 0)) return kErrArg;
    } else if ((op != kOpSign && op != kOpKeygen) ||
               (lanes_per_leaf != 0 && lanes_per_leaf != 1 && lanes_per_leaf != 2 &&
                lanes_per_leaf != 4 && lanes_per_leaf != 8 && lanes_per_leaf != 16)) {
        return kErrArg;
    }
    Tuning& t = tuning();
    std::lock_guard<std::mutex> lk(t.mu);
    if (lanes_per_leaf == 0) t.forced.erase({pid, op});
    else t.forced[{pid, op}] = lanes_per_leaf;
    return kOk;
}

int cuspx_get_split(int pid, int op, uint32_t num) {
    const SetOps* o = get_set(pid);
    if (!o) return kErrParam;
    return o->split(op, num);
}

int cuspx_set_default_param(int pid) {
    if (!get_set(pid)) return kErrParam;
    g_default_pid = pid;
    return kOk;
}
int cuspx_get_default_param(void) { return default_pid(); }
void cuspx_set_debug_mode(int on) { g_debug_mode = on ? 1 : 0; }
int cuspx_get_debug_mode(void) { return g_debug_mode; }
const char* cuspx_last_error(void) { return g_err; }

void randombytes_init(unsigned char* entropy_input, unsigned char* personalization_string,
                      int security_strength) {
    (void)security_strength;
    drbg_init(entropy_input, personalization_string);
}

int randombytes(unsigned char* x, unsigned long long xlen) {
    drbg_bytes(x, xlen);
    return 0;
}

unsigned long long crypto_sign_secretkeybytes(void) { return get_set(default_pid())->sk_bytes; }
unsigned long long crypto_sign_publickeybytes(void) { return get_set(default_pid())->pk_bytes; }
unsigned long long crypto_sign_bytes(void) { return get_set(default_pid())->sig_bytes; }
unsigned long long crypto_sign_seedbytes(void) { return 3ull * get_set(default_pid())->n; }

int crypto_sign_seed_keypair(unsigned char* pk, unsigned char* sk, const unsigned char* seed) {
    return keygen_host(get_set(default_pid()), 1, seed, 0, pk, sk);
}

int crypto_sign_keypair(unsigned char* pk, unsigned char* sk) {
    const SetOps* o = get_set(default_pid());
    uint8_t seed[96];
    draw_seed(seed, o->n);
    return keygen_host(o, 1, seed, 0, pk, sk);
}

int crypto_sign_signature(uint8_t* sig, size_t* siglen, const uint8_t* m, size_t mlen,
                          const uint8_t* sk) {
    const SetOps* o = get_set(default_pid());
    if (mlen > UINT32_MAX) return fail(kErrArg, "message longer than 2^32 - 1 bytes%s");
    uint8_t opt[32];
    draw_optrand(opt, 1, o->n);
    const uint64_t off = 0;
    const uint32_t ml = (uint32_t)mlen;
    int rc = sign_host(o, 1, sk, 0, m, &off, &ml, opt, sig, 0);
    if (rc == kOk && siglen) *siglen = o->sig_bytes;
    return rc;
}

int crypto_sign_verify(const uint8_t* sig, size_t siglen, const uint8_t* m, size_t mlen,
                       const uint8_t* pk) {
    const SetOps* o = get_set(default_pid());
    if (siglen != (size_t)o->sig_bytes) return kErrVerify;  // sign.cu:3065-3067
    if (mlen > UINT32_MAX) return fail(kErrArg, "message longer than 2^32 - 1 bytes%s");
    const uint64_t off = 0;
    const uint32_t ml = (uint32_t)mlen;
    int32_t st = -1;
    int rc = verify_host(o, 1, pk, 0, sig, 0, m, &off, &ml, &st);
    return rc == kOk ? kOk : (rc == kErrVerify ? kErrVerify : rc);
}

int crypto_sign(unsigned char* sm, unsigned long long* smlen, const unsigned char* m,
                unsigned long long mlen, const unsigned char* sk) {
    const SetOps* o = get_set(default_pid());
    std::vector<uint8_t> mc(m, m + mlen);  // m may alias sm + SPX_BYTES
    size_t siglen = 0;
    int rc = crypto_sign_signature(sm, &siglen, mc.data(), mlen, sk);
    if (rc) return rc;
    memmove(sm + o->sig_bytes, mc.data(), mlen);
    *smlen = siglen + mlen;
    return kOk;
}

int crypto_sign_open(unsigned char* m, unsigned long long* mlen, const unsigned char* sm,
                     unsigned long long smlen, const unsigned char* pk) {
    const SetOps* o = get_set(default_pid());
    if (smlen < (unsigned long long)o->sig_bytes) {  // sign.cu:6061-6065
        memset(m, 0, smlen);
        *mlen = 0;
        return kErrVerify;
    }
    const unsigned long long ml = smlen - o->sig_bytes;
    int rc = crypto_sign_verify(sm, o->sig_bytes, sm + o->sig_bytes, ml, pk);
    if (rc) {
        memset(m, 0, smlen);
        *mlen = 0;
        return rc == kErrVerify ? kErrVerify : rc;
    }
    memmove(m, sm + o->sig_bytes, ml);
    *mlen = ml;
    return kOk;
}

}  // extern "C"

// face_mdp_* / face_ms_mdp_* bodies; slices > 1 pipelines the batch over
// several streams (ms_slices)
static int mdp_keypair(unsigned char* pk, unsigned char* sk, unsigned int num, bool ms) {
    const SetOps* o = get_set(default_pid());
    if (num && (!pk || !sk)) return fail(kErrArg, "null argument%s");
    std::vector<uint8_t> seeds(3ull * o->n * num);
    mdp_seeds(seeds.data(), num, o->n);
    return keygen_host(o, num, seeds.data(), 3ull * o->n, pk, sk,
                       ms ? ms_slices(kOpKeygen, num) : 1);
}

// messages at the fixed SPX_MLEN = 32 byte stride (sign.cu:3500-3502), so
// mlen may not exceed 32; sm records of SPX_BYTES + 32 bytes hold sig || m_i
static int mdp_sign(unsigned char* sm, unsigned long long* smlen, const unsigned char* m,
                    unsigned long long mlen, const unsigned char* sk, unsigned int num, bool ms) {
    const SetOps* o = get_set(default_pid());
    if (mlen > 32) return fail(kErrArg, "mlen exceeds the 32-byte message stride (SPX_MLEN)%s");
    if (num && (!sm || !smlen || !sk || (mlen && !m))) return fail(kErrArg, "null argument%s");
    const uint64_t stride = (uint64_t)o->sig_bytes + 32;  // SPX_SM_BYTES (params.h:150)
    std::vector<uint64_t> off(num);
    std::vector<uint32_t> ml(num, (uint32_t)mlen);
    for (unsigned i = 0; i < num; i++) off[i] = mlen ? 32ull * i : 0;
    std::vector<uint8_t> opt((uint64_t)num * o->n);
    mdp_optrand(opt.data(), num, o->n);
    static const uint8_t kNoMsg[1] = {0};
    int rc = sign_host(o, num, sk, o->sk_bytes, m ? m : kNoMsg, off.data(), ml.data(), opt.data(),
                       sm, stride, ms ? ms_slices(kOpSign, num) : 1);
    if (rc) return rc;
    for (unsigned i = 0; i < num && mlen; i++)
        memcpy(sm + i * stride + o->sig_bytes, m + 32ull * i, mlen);
    if (smlen) *smlen = stride;
    return kOk;
}

// records at stride SPX_BYTES + 32 with smlen - SPX_BYTES message bytes
// (at most 32); output slots of 32 bytes
static int mdp_open(unsigned char* m, unsigned long long* mlen, const unsigned char* sm,
                    unsigned long long smlen, const unsigned char* pk, unsigned int num, bool ms) {
    const SetOps* o = get_set(default_pid());
    const uint64_t stride = (uint64_t)o->sig_bytes + 32;
    if (num && (!m || !mlen || !sm || !pk)) return fail(kErrArg, "null argument%s");
    if (smlen < (unsigned long long)o->sig_bytes) {
        memset(m, 0, 32ull * num);
        *mlen = 0;
        return kErrVerify;
    }
    if (smlen - o->sig_bytes > 32)
        return fail(kErrArg, "smlen exceeds the SPX_BYTES + 32 record stride%s");
    const uint32_t ml = (uint32_t)(smlen - o->sig_bytes);
    std::vector<uint64_t> off(num);
    std::vector<uint32_t> mls(num, ml);
    for (unsigned i = 0; i < num; i++) off[i] = i * stride + o->sig_bytes;
    std::vector<int32_t> st(num, -1);
    int rc = verify_host(o, num, pk, o->pk_bytes, sm, stride, sm, off.data(), mls.data(), st.data(),
                         ms ? ms_slices(kOpVerify, num) : 1);
    if (rc != kOk && rc != kErrVerify) return rc;
    for (unsigned i = 0; i < num; i++) {
        if (st[i] == 0) memcpy(m + 32ull * i, sm + i * stride + o->sig_bytes, ml);
        else memset(m + 32ull * i, 0, 32);
    }
    *mlen = ml;
    return rc;
}

extern "C" {

int face_mdp_crypto_sign_keypair(unsigned char* pk, unsigned char* sk, unsigned int num) {
    return mdp_keypair(pk, sk, num, false);
}

int face_mdp_crypto_sign(unsigned char* sm, unsigned long long* smlen, const unsigned char* m,
                         unsigned long long mlen, const unsigned char* sk, unsigned int num) {
    return mdp_sign(sm, smlen, m, mlen, sk, num, false);
}

int face_mdp_crypto_sign_open(unsigned char* m, unsigned long long* mlen, const unsigned char* sm,
                              unsigned long long smlen, const unsigned char* pk, unsigned int num) {
    return mdp_open(m, mlen, sm, smlen, pk, num, false);
}

int face_ms_mdp_crypto_sign_keypair(unsigned char* pk, unsigned char* sk, unsigned int num) {
    return mdp_keypair(pk, sk, num, true);
}

int face_ms_mdp_crypto_sign(unsigned char* sm, unsigned long long* smlen, unsigned char* m,
                            unsigned long long mlen, unsigned char* sk, unsigned int num) {
    return mdp_sign(sm, smlen, m, mlen, sk, num, true);
}

int face_ms_mdp_crypto_sign_open(unsigned char* m, unsigned long long* mlen, unsigned char* sm,
                                 unsigned long long smlen, unsigned char* pk, unsigned int num) {
    return mdp_open(m, mlen, sm, smlen, pk, num, true);
}

int spx_batch_keygen(int pid, uint32_t num, const uint8_t* seeds, uint8_t* pk, uint8_t* sk) {
    const SetOps* o = get_set(pid);
    if (!o) return kErrParam;
    return keygen_host(o, num, seeds, 3ull * o->n, pk, sk);
}

int spx_batch_sign(int pid, uint32_t num, const uint8_t* sk, const uint8_t* msg,
                   const uint64_t* moff, const uint32_t* mlen, const uint8_t* optrand,
                   uint8_t* sig) {
    const SetOps* o = get_set(pid);
    if (!o) return kErrParam;
    std::vector<uint8_t> opt;
    if (!optrand) {
        opt.resize((uint64_t)num * o->n);
        draw_optrand(opt.data(), num, o->n);
        optrand = opt.data();
    }
    return sign_host(o, num, sk, o->sk_bytes, msg, moff, mlen, optrand, sig, o->sig_bytes);
}

int spx_batch_verify(int pid, uint32_t num, const uint8_t* pk, const uint8_t* sig,
                     const uint8_t* msg, const uint64_t* moff, const uint32_t* mlen,
                     int32_t* status) {
    const SetOps* o = get_set(pid);
    if (!o) return kErrParam;
    return verify_host(o, num, pk, o->pk_bytes, sig, o->sig_bytes, msg, moff, mlen, status);
}

size_t cuspx_workspace_bytes(int pid, uint32_t num, int op) {
    const SetOps* o = get_set(pid);
    if (!o || op < 0 || op > 2) return 0;
    return o->workspace(num, op);
}

static hipStream_t pick_stream(void* stream) {
    if (stream) return (hipStream_t)stream;
    Device& D = device();
    std::lock_guard<std::mutex> lk(D.mu);
    D.init();
    return D.lane[0].stream;
}

int cuspx_sign_device(int pid, uint32_t num, const uint8_t* sk, uint32_t sk_stride,
                      const uint8_t* msg, const uint64_t* moff, const uint32_t* mlen,
                      const uint8_t* optrand, uint8_t* sig, uint64_t sig_stride, void* ws,
                      void* stream) {
    const SetOps* o = get_set(pid);
    if (!o) return kErrParam;
    if (num == 0) return kOk;
    if (!sk || !msg || !moff || !mlen || !optrand || !sig || !ws) return fail(kErrArg, "null argument%s");
    DevBatch b{};
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
    return o->sign(b, pick_stream(stream));
}

int cuspx_verify_device(int pid, uint32_t num, const uint8_t* pk, uint32_t pk_stride,
                        const uint8_t* sig, uint64_t sig_stride, const uint8_t* msg,
                        const uint64_t* moff, const uint32_t* mlen, int32_t* status, void* ws,
                        void* stream) {
    const SetOps* o = get_set(pid);
    if (!o) return kErrParam;
    if (num == 0) return kOk;
    if (!pk || !sig || !msg || !moff || !mlen || !status || !ws) return fail(kErrArg, "null argument%s");
    DevBatch b{};
    b.n = num;
    b.pk = pk;
    b.pk_stride = pk_stride;
    b.sig = const_cast<uint8_t*>(sig);
    b.sig_stride = sig_stride;
    b.msg = msg;
    b.moff = moff;
    b.mlen = mlen;
    b.status = status;
    b.ws = ws;
    return o->verify(b, pick_stream(stream));
}

int cuspx_keygen_device(int pid, uint32_t num, const uint8_t* seed, uint32_t seed_stride,
                        uint8_t* pk, uint8_t* sk, void* ws, void* stream) {
    const SetOps* o = get_set(pid);
    if (!o) return kErrParam;
    if (num == 0) return kOk;
    if (!seed || !pk || !sk || !ws) return fail(kErrArg, "null argument%s");
    DevBatch b{};
    b.n = num;
    b.seed = seed;
    b.seed_stride = seed_stride;
    b.pk_out = pk;
    b.sk_out = sk;
    b.ws = ws;
    return o->keygen(b, pick_stream(stream));
}

struct cuspx_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

}  // extern "C"

// capture fn(stream) (plus forked streams it records on) into a graph
template <class Fn>
static int capture(Fn fn, cuspx_graph** out) {
    if (!out) return kErrArg;
    hipStream_t cs;
    HIP_OK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    (void)hipGetLastError();
    Prof& p = prof();
    bool was_on;
    {
        std::lock_guard<std::mutex> lk(p.mu);
        was_on = p.on;
        p.on = false;  // no timing events inside a graph
    }
    int rc = kOk;
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        rc = fail(kErrHip, "hipStreamBeginCapture failed%s");
    } else {
        rc = fn(cs);
        hipError_t e = hipStreamEndCapture(cs, &g);
        if (rc == kOk && e != hipSuccess) rc = fail(kErrHip, "hipStreamEndCapture: %s", hipGetErrorString(e));
    }
    {
        std::lock_guard<std::mutex> lk(p.mu);
        p.on = was_on;
    }
    (void)hipStreamDestroy(cs);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    cuspx_graph* cg = new cuspx_graph();
    cg->graph = g;
    if (hipGraphInstantiate(&cg->exec, g, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(g);
        delete cg;
        return fail(kErrHip, "hipGraphInstantiate failed%s");
    }
    *out = cg;
    return kOk;
}

extern "C" {

int cuspx_graph_sign(int pid, uint32_t num, const uint8_t* sk, uint32_t sk_stride,
                     const uint8_t* msg, const uint64_t* moff, const uint32_t* mlen,
                     const uint8_t* optrand, uint8_t* sig, uint64_t sig_stride, void* ws,
                     cuspx_graph** out) {
    const SetOps* o = get_set(pid);
    if (!o) return kErrParam;
    if (!num || !sk || !msg || !moff || !mlen || !optrand || !sig || !ws) return kErrArg;
    DevBatch b{};
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
    return capture([&](hipStream_t s) { return o->sign(b, s); }, out);
}

int cuspx_graph_verify(int pid, uint32_t num, const uint8_t* pk, uint32_t pk_stride,
                       const uint8_t* sig, uint64_t sig_stride, const uint8_t* msg,
                       const uint64_t* moff, const uint32_t* mlen, int32_t* status, void* ws,
                       cuspx_graph** out) {
    const SetOps* o = get_set(pid);
    if (!o) return kErrParam;
    if (!num || !pk || !sig || !msg || !moff || !mlen || !status || !ws) return kErrArg;
    DevBatch b{};
    b.n = num;
    b.pk = pk;
    b.pk_stride = pk_stride;
    b.sig = const_cast<uint8_t*>(sig);
    b.sig_stride = sig_stride;
    b.msg = msg;
    b.moff = moff;
    b.mlen = mlen;
    b.status = status;
    b.ws = ws;
    return capture([&](hipStream_t s) { return o->verify(b, s); }, out);
}

}  // extern "C"

static DevBatch group_batch(const cuspx_sign_group& gr) {
    DevBatch b{};
    b.n = gr.num;
    b.sk = gr.sk;
    b.sk_stride = gr.sk_stride;
    b.msg = gr.msg;
    b.moff = gr.moff;
    b.mlen = gr.mlen;
    b.optrand = gr.optrand;
    b.sig = gr.sig;
    b.sig_stride = gr.sig_stride;
    b.ws = gr.ws;
    return b;
}

static int check_groups(const cuspx_sign_group* groups, int ngroups) {
    if (!groups || ngroups <= 0 || ngroups > 64) return fail(kErrArg, "bad group list%s");
    for (int i = 0; i < ngroups; i++) {
        const cuspx_sign_group& g = groups[i];
        if (!get_set(g.pid)) return kErrParam;
        if (g.num && (!g.sk || !g.msg || !g.moff || !g.mlen || !g.optrand || !g.sig || !g.ws))
            return fail(kErrArg, "null argument in a group%s");
    }
    return kOk;
}

// A mixed batch whose non-empty groups are distinct fast sets of one hash
// family runs as one persistent kernel (FamilyOps); returns kNotFused if the
// groups do not form such a batch or the split rule would run a group at
// more than two lanes per WOTS leaf (nothing launched then).
static int sign_family(const cuspx_sign_group* groups, int ngroups, hipStream_t s) {
    const FamilyOps* fams[2] = {cuspx_family_sha2_f(), cuspx_family_shake256_f()};
    for (const FamilyOps* f : fams) {
        DevBatch b[3] = {};
        bool ok = true, any = false;
        for (int i = 0; i < ngroups && ok; i++) {
            if (!groups[i].num) continue;
            int slot = -1;
            for (int k = 0; k < 3; k++)
                if (f->pid[k] == groups[i].pid) slot = k;
            ok = slot >= 0 && b[slot].n == 0;
            if (ok) b[slot] = group_batch(groups[i]);
            any = true;
        }
        if (ok && any) return f->sign(b, s);
    }
    return kNotFused;
}

extern "C" {

int cuspx_sign_mixed_device(const cuspx_sign_group* groups, int ngroups, void* stream) {
    int rc = check_groups(groups, ngroups);
    if (rc) return rc;
    hipStream_t s = pick_stream(stream);
    rc = sign_family(groups, ngroups, s);
    if (rc != kNotFused) return rc;
    for (int i = 0; i < ngroups; i++) {  // one set after another on the stream
        if (!groups[i].num) continue;
        rc = get_set(groups[i].pid)->sign(group_batch(groups[i]), s);
        if (rc) return rc;
    }
    return kOk;
}

int cuspx_graph_sign_mixed(const cuspx_sign_group* groups, int ngroups, cuspx_graph** out) {
    int rc0 = check_groups(groups, ngroups);
    if (rc0) return rc0;
    std::vector<hipStream_t> fs(ngroups, nullptr);
    std::vector<hipEvent_t> ev(ngroups + 1, nullptr);
    for (int i = 0; i < ngroups; i++) HIP_OK(hipStreamCreateWithFlags(&fs[i], hipStreamNonBlocking));
    for (auto& e : ev) HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    int rc = capture(
        [&](hipStream_t s) {
            const int r = sign_family(groups, ngroups, s);
            if (r != kNotFused) return r;
            // other mixes: every group waits on the capture stream, runs its
            // own pipeline on its own stream, and the capture stream joins
            // all of them
            if (hipEventRecord(ev[ngroups], s) != hipSuccess) return (int)kErrHip;
            for (int i = 0; i < ngroups; i++) {
                const cuspx_sign_group& gr = groups[i];
                if (hipStreamWaitEvent(fs[i], ev[ngroups], 0) != hipSuccess) return (int)kErrHip;
                if (gr.num) {
                    const int r2 = get_set(gr.pid)->sign(group_batch(gr), fs[i]);
                    if (r2) return r2;
                }
                if (hipEventRecord(ev[i], fs[i]) != hipSuccess) return (int)kErrHip;
                if (hipStreamWaitEvent(s, ev[i], 0) != hipSuccess) return (int)kErrHip;
            }
            return (int)kOk;
        },
        out);
    for (auto& e : ev) (void)hipEventDestroy(e);
    for (auto& f : fs) (void)hipStreamDestroy(f);
    return rc;
}

int cuspx_graph_launch(cuspx_graph* g, void* stream) {
    if (!g) return kErrArg;
    HIP_OK(hipGraphLaunch(g->exec, pick_stream(stream)));
    return kOk;
}

int cuspx_graph_destroy(cuspx_graph* g) {
    if (!g) return kErrArg;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return kOk;
}

int cuspx_dev_alloc(void** ptr, size_t bytes) {
    if (!ptr) return kErrArg;
    HIP_OK(hipMalloc(ptr, bytes ? bytes : 1));
    return kOk;
}
int cuspx_dev_free(void* ptr) {
    HIP_OK(hipFree(ptr));
    return kOk;
}
int cuspx_memcpy_htod(void* dst, const void* src, size_t bytes) {
    HIP_OK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return kOk;
}
int cuspx_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
    HIP_OK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return kOk;
}
int cuspx_memcpy_dtod(void* dst, const void* src, size_t bytes) {
    HIP_OK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return kOk;
}
int cuspx_memset_device(void* dst, int value, size_t bytes) {
    HIP_OK(hipMemset(dst, value, bytes));
    return kOk;
}
int cuspx_stream_create(void** stream) {
    if (!stream) return kErrArg;
    hipStream_t s;
    HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return kOk;
}
int cuspx_stream_destroy(void* stream) {
    HIP_OK(hipStreamDestroy((hipStream_t)stream));
    return kOk;
}
int cuspx_stream_sync(void* stream) {
    HIP_OK(hipStreamSynchronize(pick_stream(stream)));
    return kOk;
}
int cuspx_event_create(void** event) {
    if (!event) return kErrArg;
    hipEvent_t e;
    HIP_OK(hipEventCreate(&e));
    *event = e;
    return kOk;
}
int cuspx_event_destroy(void* event) {
    HIP_OK(hipEventDestroy((hipEvent_t)event));
    return kOk;
}
int cuspx_event_record(void* event, void* stream) {
    HIP_OK(hipEventRecord((hipEvent_t)event, pick_stream(stream)));
    return kOk;
}
int cuspx_event_elapsed_ms(void* start, void* stop, float* ms) {
    HIP_OK(hipEventSynchronize((hipEvent_t)stop));
    HIP_OK(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return kOk;
}
int cuspx_device_info(int* compute_units, int* clock_khz, size_t* hbm_bytes, char* name,
                      int name_cap) {
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    hipDeviceProp_t pr;
    HIP_OK(hipGetDeviceProperties(&pr, dev));
    if (compute_units) *compute_units = pr.multiProcessorCount;
    if (clock_khz) *clock_khz = pr.clockRate;
    if (hbm_bytes) *hbm_bytes = pr.totalGlobalMem;
    if (name && name_cap > 0) {
        if (pr.name[0]) snprintf(name, name_cap, "%s (%s)", pr.name, pr.gcnArchName);
        else snprintf(name, name_cap, "%s", pr.gcnArchName);
    }
    return kOk;
}

void cuspx_profile_enable(int on) {
    Prof& p = prof();
    std::lock_guard<std::mutex> lk(p.mu);
    p.on = on != 0;
    if (on) p.by_set = on == 2;
}

void cuspx_profile_reset(void) {
    Prof& p = prof();
    std::lock_guard<std::mutex> lk(p.mu);
    for (auto& c : p.calls)
        for (auto& m : c) (void)hipEventDestroy(m.ev);
    p.calls.clear();
    p.call_set.clear();
    p.open.clear();
    p.acc.clear();
}

int cuspx_profile_read(char* names, int names_cap, float* ms, int* count, int cap) {
    Prof& p = prof();
    std::lock_guard<std::mutex> lk(p.mu);
    for (size_t ci = 0; ci < p.calls.size(); ci++) {
        auto& c = p.calls[ci];
        for (size_t i = 0; i < c.size(); i++) (void)hipEventSynchronize(c[i].ev);
        for (size_t i = 1; i < c.size(); i++) {
            float t = 0.f;
            if (hipEventElapsedTime(&t, c[i - 1].ev, c[i].ev) != hipSuccess) continue;
            auto& a = p.acc[p.by_set ? std::string(p.call_set[ci]) + ":" + c[i].name
                                     : std::string(c[i].name)];
            a.first += t;
            a.second += 1;
        }
        for (auto& m : c) (void)hipEventDestroy(m.ev);
    }
    p.calls.clear();
    p.call_set.clear();
    p.open.clear();
    int k = 0;
    std::string all;
    for (auto& kv : p.acc) {
        if (k >= cap) break;
        ms[k] = (float)kv.second.first;
        count[k] = kv.second.second;
        all += kv.first;
        all += '\n';
        k++;
    }
    if (names && names_cap > 0) snprintf(names, names_cap, "%s", all.c_str());
    (void)hipGetLastError();  // events of destroyed streams must not leak an error
    return k;
}

}  // extern "C"
