// server.cpp -- continuous-batching SPHINCS+ service on one GPU.
//
// * Keys stay resident in HBM: one table per parameter set holds sk / pk
//   rows and each key's SHA-256 pub_seed midstate (computed once on the GPU
//   when the key is added).
// * Requests (sign or verify, any parameter set, message 0..max_msg bytes)
//   queue per (set, op, message-size class: <= 2 KiB, <= 16 KiB, larger) so
//   short messages never wait for the serial hash of a long one in the same
//   launch; batches of the longest class may be held to `long_slots` slots
//   at once, so their long serial hashes leave slots to short requests.  A
//   scheduler thread keeps up to `slots` batches in flight, each on its own
//   stream with its own pinned staging and device arena.  When a slot frees,
//   the oldest queue is packed into the next batch: immediately while fewer
//   than half the slots are busy or when the batch is full, otherwise after
//   at most `batch_window_us` -- so requests that arrive while other
//   batches run are admitted as slots free up and overlap on the device.
// * Admission: the device-memory budget is split over the slots; a batch
//   takes requests while its inputs, outputs and workspace fit one slot's
//   share, and a request that could never fit is refused at submission.
// * Latency is measured from submission to the host seeing the batch's
//   completion event.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/cuspx.h"
#include "drbg.hpp"
#include "engine.hpp"

namespace spx {
namespace {

using Clock = std::chrono::steady_clock;

inline uint64_t al16(uint64_t x) { return (x + 15) & ~uint64_t(15); }
inline uint64_t al256(uint64_t x) { return (x + 255) & ~uint64_t(255); }

struct Request {
    uint64_t ticket;
    int op;  // 0 sign, 1 verify
    int pid;
    uint32_t row;
    std::vector<uint8_t> msg;
    std::vector<uint8_t> aux;  // sign: optrand (n); verify: signature
    std::vector<uint8_t> sig;  // sign result
    int32_t status = 0;
    int cls = 0;            // message-size class (queue key)
    bool done = false;
    bool cancelled = false;
    bool launched = false;  // taken from its queue into a batch
    Clock::time_point t_submit, t_launch, t_done;
};
using ReqP = std::shared_ptr<Request>;

// Key ids are (generation << 20) | (set << 16) | row: a removed key's row is
// reused with the next generation, so a stale id cannot reach the new key.
constexpr int32_t kRowBits = 16, kSetBits = 4, kGenMask = (1 << 11) - 1;
inline int32_t make_key_id(int pid, uint32_t row, uint32_t gen) {
    return (int32_t)(((gen & kGenMask) << (kRowBits + kSetBits)) | ((uint32_t)pid << kRowBits) |
                     row);
}

struct KeyTable {
    uint8_t* d_sk = nullptr;
    uint8_t* d_pk = nullptr;
    uint32_t* d_mid = nullptr;
    enum : uint8_t { kFree = 0, kLive = 1, kRemoved = 2 };
    std::vector<uint8_t> state;    // per row
    std::vector<bool> has_sk;      // rows added without a secret key verify only
    std::vector<uint32_t> refs;    // requests queued or running on the row
    std::vector<uint32_t> gen;     // generation of the row's current key
    std::vector<uint32_t> free_rows;  // removed rows whose requests have drained
    uint32_t count = 0, cap = 0;   // rows ever used, capacity

    bool live(uint32_t row, uint32_t g) const {
        return row < count && state[row] == kLive && (gen[row] & kGenMask) == g;
    }
    // a request on `row` finished (or left its queue)
    void release(uint32_t row) {
        if (row < count && refs[row] && --refs[row] == 0 && state[row] == kRemoved) {
            state[row] = kFree;
            free_rows.push_back(row);
        }
    }
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    uint8_t* h = nullptr;  // pinned staging (inputs, then outputs)
    uint8_t* d = nullptr;  // device arena: inputs, outputs, workspace
    uint64_t cap = 0;      // device arena bytes
    uint64_t hcap = 0;     // staging bytes
    bool busy = false;
    int pid = -1, op = -1, cls = 0;
    uint64_t out_off = 0;  // offset of the outputs in h / d
    std::vector<ReqP> batch;
};

}  // namespace

struct Server {
    spx_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::thread th;
    bool stop = false;
    int err = 0;
    uint64_t next_ticket = 1;
    std::map<int, KeyTable> keys;
    std::map<std::tuple<int, int, int>, std::deque<ReqP>> queues;  // (set, op, size class)
    std::map<uint64_t, ReqP> live;
    std::vector<Slot> slots;
    hipStream_t setup = nullptr;
    uint64_t pending_bytes = 0;
    // statistics
    std::vector<float> lat_us[2];
    std::map<std::pair<int, int>, std::vector<float>> lat_set;
    std::vector<float> wait_us, exec_us;  // sign: submit->launch, launch->done
    uint64_t batches = 0, batched_reqs = 0;

    // bytes one request of (pid, op) needs in a slot besides its message
    uint64_t item_bytes(const SetOps* o, int op) const {
        const uint64_t io = op == 0 ? (uint64_t)o->n + o->sig_bytes + 8 + 4 + 4 + 48
                                    : (uint64_t)o->sig_bytes + 4 + 8 + 4 + 4 + 48;
        const uint64_t ws1 = o->workspace(1, op == 0 ? kOpSign : kOpVerify);
        return io + ws1;
    }

    int launch(Slot& sl, int pid, int op, std::vector<ReqP>& reqs);
    void harvest(Slot& sl);
    void loop();
};

int Server::launch(Slot& sl, int pid, int op, std::vector<ReqP>& reqs) {
    const SetOps* o = get_set(pid);
    KeyTable& kt = keys[pid];
    const uint32_t num = (uint32_t)reqs.size();
    // host staging layout: msgs | moff | mlen | key rows | optrand or sigs
    std::vector<uint64_t> off(num);
    uint64_t mb = 0;
    for (uint32_t i = 0; i < num; i++) {
        off[i] = mb;
        mb += al16(reqs[i]->msg.size());
    }
    mb += 16;
    const uint64_t auxb = op == 0 ? (uint64_t)o->n : (uint64_t)o->sig_bytes;
    const uint64_t o_moff = al256(mb), o_mlen = o_moff + al256(8ull * num),
                   o_row = o_mlen + al256(4ull * num), o_aux = o_row + al256(4ull * num),
                   in_end = o_aux + al256(auxb * num);
    const uint64_t outb = op == 0 ? (uint64_t)o->sig_bytes * num : 4ull * num;
    const uint64_t o_out = al256(in_end), o_ws = al256(o_out + outb);
    const uint64_t ws = o->workspace(num, op == 0 ? kOpSign : kOpVerify);
    if (o_ws + ws > sl.cap || o_ws > sl.hcap) return kErrBudget;
    uint8_t* h = sl.h;
    memset(h, 0, mb);
    for (uint32_t i = 0; i < num; i++) {
        if (!reqs[i]->msg.empty()) memcpy(h + off[i], reqs[i]->msg.data(), reqs[i]->msg.size());
        ((uint64_t*)(h + o_moff))[i] = off[i];
        ((uint32_t*)(h + o_mlen))[i] = (uint32_t)reqs[i]->msg.size();
        ((uint32_t*)(h + o_row))[i] = reqs[i]->row;
        memcpy(h + o_aux + auxb * i, reqs[i]->aux.data(), auxb);
    }
    if (hipMemcpyAsync(sl.d, h, in_end, hipMemcpyHostToDevice, sl.stream) != hipSuccess)
        return kErrHip;
    DevBatch b{};
    b.n = num;
    b.msg = sl.d;
    b.moff = (const uint64_t*)(sl.d + o_moff);
    b.mlen = (const uint32_t*)(sl.d + o_mlen);
    b.mlen_bound = 1;  // the longest message (0 would mean "not known")
    for (uint32_t i = 0; i < num; i++)
        b.mlen_bound = std::max<uint32_t>(b.mlen_bound, (uint32_t)reqs[i]->msg.size());
    b.key_idx = (const uint32_t*)(sl.d + o_row);
    b.key_mid = kt.d_mid;
    b.ws = sl.d + o_ws;
    int rc;
    if (op == 0) {
        b.sk = kt.d_sk;
        b.sk_stride = o->sk_bytes;
        b.optrand = sl.d + o_aux;
        b.sig = sl.d + o_out;
        b.sig_stride = o->sig_bytes;
        rc = o->sign(b, sl.stream);
    } else {
        b.pk = kt.d_pk;
        b.pk_stride = o->pk_bytes;
        b.sig = sl.d + o_aux;
        b.sig_stride = o->sig_bytes;
        b.status = (int32_t*)(sl.d + o_out);
        rc = o->verify(b, sl.stream);
    }
    if (rc) return rc;
    if (hipMemcpyAsync(h + o_out, sl.d + o_out, outb, hipMemcpyDeviceToHost, sl.stream) !=
        hipSuccess)
        return kErrHip;
    if (hipEventRecord(sl.ev, sl.stream) != hipSuccess) return kErrHip;
    const auto tl = Clock::now();
    for (auto& r : reqs) r->t_launch = tl;
    sl.busy = true;
    sl.pid = pid;
    sl.op = op;
    sl.out_off = o_out;
    sl.batch.swap(reqs);
    batches++;
    batched_reqs += num;
    return kOk;
}

void Server::harvest(Slot& sl) {
    const SetOps* o = get_set(sl.pid);
    const auto now = Clock::now();
    KeyTable& kt = keys[sl.pid];
    for (size_t i = 0; i < sl.batch.size(); i++) {
        Request& r = *sl.batch[i];
        kt.release(r.row);
        if (sl.op == 0) {
            r.sig.assign(sl.h + sl.out_off + (uint64_t)o->sig_bytes * i,
                         sl.h + sl.out_off + (uint64_t)o->sig_bytes * (i + 1));
            r.status = 0;
        } else {
            r.status = ((const int32_t*)(sl.h + sl.out_off))[i];
        }
        r.done = true;
        r.t_done = now;
        if (!r.cancelled) {
            const float us = std::chrono::duration<float, std::micro>(now - r.t_submit).count();
            if (lat_us[sl.op].size() < (1u << 24)) {
                lat_us[sl.op].push_back(us);
                lat_set[{sl.pid, sl.op}].push_back(us);
                if (sl.op == 0) {
                    wait_us.push_back(
                        std::chrono::duration<float, std::micro>(r.t_launch - r.t_submit).count());
                    exec_us.push_back(
                        std::chrono::duration<float, std::micro>(now - r.t_launch).count());
                }
            }
        } else {
            live.erase(r.ticket);
        }
    }
    sl.batch.clear();
    sl.busy = false;
}

void Server::loop() {
    std::unique_lock<std::mutex> lk(mu);
    const auto window = std::chrono::microseconds(cfg.batch_window_us);
    while (!stop) {
        bool progressed = false;
        int inflight = 0;
        for (auto& sl : slots) {
            if (sl.busy && hipEventQuery(sl.ev) == hipSuccess) {
                harvest(sl);
                progressed = true;
            }
            inflight += sl.busy;
        }
        if (progressed) cv_done.notify_all();
        Clock::time_point wake = Clock::now() + std::chrono::milliseconds(5);
        int long_busy = 0;
        for (auto& sl : slots) long_busy += sl.busy && sl.cls == 2;
        for (auto& sl : slots) {
            if (sl.busy) continue;
            // the queue whose head request is oldest (the long-message class
            // only while it holds fewer than long_slots slots)
            std::deque<ReqP>* best = nullptr;
            std::tuple<int, int, int> bk;
            for (auto& kv : queues) {
                if (kv.second.empty()) continue;
                if (std::get<2>(kv.first) == 2 && long_busy >= (int)cfg.long_slots) continue;
                if (!best || kv.second.front()->t_submit < best->front()->t_submit) {
                    best = &kv.second;
                    bk = kv.first;
                }
            }
            if (!best) break;
            const int b_pid = std::get<0>(bk), b_op = std::get<1>(bk);
            const SetOps* o = get_set(b_pid);
            const uint64_t per = item_bytes(o, b_op);
            // how many fit the batch limit and this slot's share of the budget
            uint32_t take = 0;
            uint64_t bytes = 4096, hbytes = 4096;
            const uint64_t io = (uint64_t)o->sig_bytes + o->n + 1024;
            for (auto& r : *best) {
                const uint64_t need = per + al16(r->msg.size()) + 1024;
                const uint64_t hneed = io + al16(r->msg.size());
                if (take >= cfg.max_batch || bytes + need > slots[0].cap ||
                    hbytes + hneed > slots[0].hcap)
                    break;
                bytes += need;
                hbytes += hneed;
                take++;
            }
            // launch now while at least half the slots are idle (a wait
            // there only adds latency: requests spread over up to 18
            // queues rarely meet in one window at moderate rates); once
            // most slots are busy, hold a batch until it is full or its
            // oldest request has waited the batch window
            const auto age = Clock::now() - best->front()->t_submit;
            const bool full = take >= cfg.max_batch || take < best->size();
            const bool idle = 2 * inflight < (int)slots.size();
            if (!(idle || full || age >= window)) {
                wake = std::min(wake, best->front()->t_submit + window);
                continue;
            }
            std::vector<ReqP> reqs;
            reqs.reserve(take);
            for (uint32_t i = 0; i < take; i++) {
                ReqP r = best->front();
                best->pop_front();
                pending_bytes -= r->msg.size() + r->aux.size();
                r->launched = true;
                reqs.push_back(r);
            }
            if (reqs.empty()) continue;
            sl.cls = std::get<2>(bk);
            const int rc = launch(sl, b_pid, b_op, reqs);
            if (rc) {  // fail the batch, keep serving
                for (auto& r : reqs) {
                    r->status = rc;
                    r->done = true;
                    r->t_done = Clock::now();
                    keys[b_pid].release(r->row);
                }
                cv_done.notify_all();
                err = rc;
            } else {
                inflight++;
                long_busy += sl.cls == 2;
                progressed = true;
            }
        }
        if (!progressed) {
            // completions are polled: wait briefly unless new work arrives
            const auto poll = Clock::now() + std::chrono::microseconds(inflight ? 50 : 2000);
            cv_work.wait_until(lk, std::min(wake, poll));
        }
    }
}

}  // namespace spx

using namespace spx;

extern "C" {

int spx_server_create(const spx_server_config* cfg, spx_server** out) {
    if (!out) return kErrArg;
    Server* s = new Server();
    if (cfg) s->cfg = *cfg;
    if (!s->cfg.max_batch) s->cfg.max_batch = 4096;
    if (!s->cfg.slots) s->cfg.slots = 4;
    if (!s->cfg.long_slots || s->cfg.long_slots > s->cfg.slots) s->cfg.long_slots = s->cfg.slots;
    if (!s->cfg.max_msg_bytes) s->cfg.max_msg_bytes = 65536;
    if (!s->cfg.max_keys_per_set) s->cfg.max_keys_per_set = 1024;
    if (s->cfg.max_keys_per_set > (1u << kRowBits)) {  // rows are 16 bits of a key id
        delete s;
        return kErrArg;
    }
    if (!s->cfg.hbm_budget_bytes) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            delete s;
            return kErrHip;
        }
        s->cfg.hbm_budget_bytes = std::min<uint64_t>(fr / 4, 16ull << 30);
    }
    if (hipStreamCreateWithFlags(&s->setup, hipStreamNonBlocking) != hipSuccess) {
        delete s;
        return kErrHip;
    }
    const uint64_t share = s->cfg.hbm_budget_bytes / s->cfg.slots;
    // staging holds one batch of messages plus the larger of its signatures
    const uint64_t hshare = std::min<uint64_t>(
        share, (uint64_t)s->cfg.max_batch * (s->cfg.max_msg_bytes + 50 * 1024) + (1 << 20));
    s->slots.resize(s->cfg.slots);
    for (auto& sl : s->slots) {
        sl.cap = share;
        sl.hcap = hshare;
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess ||
            hipMalloc((void**)&sl.d, share) != hipSuccess ||
            hipHostMalloc((void**)&sl.h, hshare, hipHostMallocDefault) != hipSuccess) {
            spx_server_destroy((spx_server*)s);
            return kErrHip;
        }
    }
    s->th = std::thread([s] { s->loop(); });
    *out = (spx_server*)s;
    return kOk;
}

int spx_server_add_key(spx_server* sv, int pid, const uint8_t* sk, const uint8_t* pk,
                       int32_t* key_id) {
    Server* s = (Server*)sv;
    const SetOps* o = get_set(pid);
    if (!s || !key_id || !pk) return kErrArg;
    if (!o) return kErrParam;
    std::lock_guard<std::mutex> lk(s->mu);
    KeyTable& kt = s->keys[pid];
    if (kt.cap == 0) {
        kt.cap = s->cfg.max_keys_per_set;
        if (hipMalloc((void**)&kt.d_sk, (uint64_t)kt.cap * o->sk_bytes) != hipSuccess ||
            hipMalloc((void**)&kt.d_pk, (uint64_t)kt.cap * o->pk_bytes) != hipSuccess ||
            hipMalloc((void**)&kt.d_mid, (uint64_t)kt.cap * 32) != hipSuccess)
            return kErrHip;
    }
    if (kt.state.size() < kt.cap) {
        kt.state.resize(kt.cap, KeyTable::kFree);
        kt.has_sk.resize(kt.cap, false);
        kt.refs.resize(kt.cap, 0);
        kt.gen.resize(kt.cap, 0);
    }
    // a removed row whose requests have drained, else a new one
    const bool reuse = !kt.free_rows.empty();
    if (!reuse && kt.count >= kt.cap) return kErrBudget;
    const uint32_t row = reuse ? kt.free_rows.back() : kt.count;
    std::vector<uint8_t> skb(o->sk_bytes, 0);
    if (sk) memcpy(skb.data(), sk, o->sk_bytes);
    if (hipMemcpy(kt.d_sk + (uint64_t)row * o->sk_bytes, skb.data(), o->sk_bytes,
                  hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(kt.d_pk + (uint64_t)row * o->pk_bytes, pk, o->pk_bytes,
                  hipMemcpyHostToDevice) != hipSuccess)
        return kErrHip;
    int rc = o->midstates(kt.d_pk + (uint64_t)row * o->pk_bytes, o->pk_bytes, kt.d_mid + row * 8,
                          1, s->setup);
    if (rc || hipStreamSynchronize(s->setup) != hipSuccess) return kErrHip;
    if (reuse) {
        kt.free_rows.pop_back();
        kt.gen[row]++;
    } else {
        kt.count++;
    }
    kt.has_sk[row] = sk != nullptr;
    kt.state[row] = KeyTable::kLive;
    kt.refs[row] = 0;
    *key_id = make_key_id(pid, row, kt.gen[row]);
    return kOk;
}

int spx_server_remove_key(spx_server* sv, int32_t key_id) {
    Server* s = (Server*)sv;
    if (!s || key_id < 0) return kErrArg;
    const int pid = (key_id >> kRowBits) & ((1 << kSetBits) - 1);
    const uint32_t row = (uint32_t)key_id & ((1u << kRowBits) - 1);
    const uint32_t g = (uint32_t)key_id >> (kRowBits + kSetBits);
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->keys.find(pid);
    if (it == s->keys.end() || !it->second.live(row, g)) return kErrArg;
    KeyTable& kt = it->second;
    kt.state[row] = KeyTable::kRemoved;  // no new requests; queued / running ones finish
    if (kt.refs[row] == 0) {
        kt.state[row] = KeyTable::kFree;
        kt.free_rows.push_back(row);
    }
    return kOk;
}

int spx_server_submit(spx_server* sv, int op, int32_t key_id, const uint8_t* msg, uint32_t mlen,
                      const uint8_t* aux, uint64_t* ticket) {
    Server* s = (Server*)sv;
    if (!s || !ticket || (op != 0 && op != 1) || (mlen && !msg)) return kErrArg;
    if (key_id < 0) return kErrParam;
    const int pid = (key_id >> kRowBits) & ((1 << kSetBits) - 1);
    const uint32_t row = (uint32_t)key_id & ((1u << kRowBits) - 1);
    const uint32_t g = (uint32_t)key_id >> (kRowBits + kSetBits);
    const SetOps* o = get_set(pid);
    if (!o) return kErrParam;
    if (op == 1 && !aux) return kErrArg;
    if (mlen > s->cfg.max_msg_bytes) return kErrBudget;
    auto r = std::make_shared<Request>();
    r->op = op;
    r->pid = pid;
    r->row = row;
    r->msg.assign(msg, msg + mlen);
    if (op == 0) {
        r->aux.resize(o->n);
        if (aux) {
            memcpy(r->aux.data(), aux, o->n);
        } else {  // as crypto_sign_signature: n DRBG bytes, 0xad in debug mode
            drbg_bytes(r->aux.data(), o->n);
            if (cuspx_get_debug_mode()) memset(r->aux.data(), 0xad, o->n);
        }
    } else {
        r->aux.assign(aux, aux + o->sig_bytes);
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->keys.find(pid);
        if (it == s->keys.end() || !it->second.live(row, g)) return kErrArg;  // unknown / removed
        if (op == 0 && !it->second.has_sk[row]) return kErrArg;  // verify-only key
        // one request must fit a slot; the queue holds at most one budget of bytes
        const uint64_t need = s->item_bytes(o, op) + al16(mlen) + 8192;
        const uint64_t hneed = (uint64_t)o->sig_bytes + o->n + al16(mlen) + 8192;
        if (need > s->slots[0].cap || hneed > s->slots[0].hcap ||
            s->pending_bytes + mlen + r->aux.size() > s->cfg.hbm_budget_bytes)
            return kErrBudget;
        r->ticket = s->next_ticket++;
        r->t_submit = Clock::now();
        s->pending_bytes += mlen + r->aux.size();
        const int cls = mlen <= 2048 ? 0 : (mlen <= 16384 ? 1 : 2);
        r->cls = cls;
        s->queues[std::make_tuple(pid, op, cls)].push_back(r);
        s->live[r->ticket] = r;
        it->second.refs[row]++;  // the row stays until this request is done
        *ticket = r->ticket;
    }
    s->cv_work.notify_one();
    return kOk;
}

int spx_server_wait(spx_server* sv, uint64_t ticket, int timeout_ms, uint8_t* sig_out,
                    int32_t* status) {
    Server* s = (Server*)sv;
    if (!s) return kErrArg;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->live.find(ticket);
    if (it == s->live.end()) return kErrArg;
    ReqP r = it->second;
    const auto until = Clock::now() + std::chrono::milliseconds(timeout_ms < 0 ? 1 << 30 : timeout_ms);
    while (!r->done) {
        if (s->cv_done.wait_until(lk, until) == std::cv_status::timeout && !r->done) return 1;
    }
    if (status) *status = r->status;
    if (sig_out && r->op == 0 && !r->sig.empty()) memcpy(sig_out, r->sig.data(), r->sig.size());
    s->live.erase(ticket);
    return r->status < -1 ? r->status : kOk;
}

int spx_server_cancel(spx_server* sv, uint64_t ticket) {
    Server* s = (Server*)sv;
    if (!s) return kErrArg;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->live.find(ticket);
    if (it == s->live.end()) return kErrArg;
    ReqP r = it->second;
    r->cancelled = true;
    if (!r->launched && !r->done) {
        // still queued: leave the queue now (its bytes count no more) and
        // complete with -4, waking any thread waiting on it
        auto q = s->queues.find(std::make_tuple(r->pid, r->op, r->cls));
        if (q != s->queues.end()) {
            auto& dq = q->second;
            for (auto x = dq.begin(); x != dq.end(); ++x)
                if (x->get() == r.get()) {
                    dq.erase(x);
                    break;
                }
        }
        s->pending_bytes -= r->msg.size() + r->aux.size();
        s->keys[r->pid].release(r->row);
        r->status = kErrArg;
        r->done = true;
        r->t_done = Clock::now();
        s->live.erase(it);
        lk.unlock();
        s->cv_done.notify_all();
        return kOk;
    }
    // launched (the GPU work is not stopped; the result is dropped) or done
    if (r->done) s->live.erase(it);
    return 1;
}

static float pct(std::vector<float> v, double q) {
    if (v.empty()) return 0.f;
    const size_t k = std::min(v.size() - 1, (size_t)(q * (v.size() - 1) + 0.5));
    std::nth_element(v.begin(), v.begin() + k, v.end());
    return v[k];
}

int spx_server_stats(spx_server* sv, spx_server_statistics* st) {
    Server* s = (Server*)sv;
    if (!s || !st) return kErrArg;
    std::lock_guard<std::mutex> lk(s->mu);
    memset(st, 0, sizeof(*st));
    st->signs = s->lat_us[0].size();
    st->verifies = s->lat_us[1].size();
    st->batches = s->batches;
    st->mean_batch = s->batches ? (double)s->batched_reqs / s->batches : 0.0;
    st->sign_p50_us = pct(s->lat_us[0], 0.50);
    st->sign_p99_us = pct(s->lat_us[0], 0.99);
    st->verify_p50_us = pct(s->lat_us[1], 0.50);
    st->verify_p99_us = pct(s->lat_us[1], 0.99);
    uint64_t pend = 0;
    for (auto& kv : s->queues) pend += kv.second.size();
    st->pending = pend;
    st->hbm_bytes = s->cfg.hbm_budget_bytes;
    st->last_error = s->err;
    return kOk;
}

int spx_server_reset_stats(spx_server* sv) {
    Server* s = (Server*)sv;
    if (!s) return kErrArg;
    std::lock_guard<std::mutex> lk(s->mu);
    s->lat_us[0].clear();
    s->lat_us[1].clear();
    s->lat_set.clear();
    s->wait_us.clear();
    s->exec_us.clear();
    s->batches = s->batched_reqs = 0;
    return kOk;
}

int spx_server_latency(spx_server* sv, int pid, int op, float* p50_us, float* p99_us,
                       uint64_t* count) {
    Server* s = (Server*)sv;
    if (!s || !p50_us || !p99_us || !count) return kErrArg;
    std::lock_guard<std::mutex> lk(s->mu);
    if (pid < 0) {  // pid -1: sign queue wait (op 0) / sign execution (op 1)
        const std::vector<float>& x = op == 0 ? s->wait_us : s->exec_us;
        *p50_us = pct(x, 0.50);
        *p99_us = pct(x, 0.99);
        *count = x.size();
        return kOk;
    }
    auto it = s->lat_set.find({pid, op});
    if (it == s->lat_set.end()) {
        *p50_us = *p99_us = 0.f;
        *count = 0;
        return kOk;
    }
    *p50_us = pct(it->second, 0.50);
    *p99_us = pct(it->second, 0.99);
    *count = it->second.size();
    return kOk;
}

int spx_server_destroy(spx_server* sv) {
    Server* s = (Server*)sv;
    if (!s) return kErrArg;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->th.joinable()) s->th.join();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        for (auto& kv : s->live)
            if (!kv.second->done) {
                kv.second->status = kErrArg;
                kv.second->done = true;
            }
    }
    s->cv_done.notify_all();
    for (auto& sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.d) (void)hipFree(sl.d);
        if (sl.h) (void)hipHostFree(sl.h);
        if (sl.ev) (void)hipEventDestroy(sl.ev);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    for (auto& kv : s->keys) {
        (void)hipFree(kv.second.d_sk);
        (void)hipFree(kv.second.d_pk);
        (void)hipFree(kv.second.d_mid);
    }
    if (s->setup) (void)hipStreamDestroy(s->setup);
    delete s;
    return kOk;
}

}  // extern "C"
