/*
 * cuspx.h -- C ABI of the MI355X SPHINCS+ engine (libcuspx.so).
 *
 * Plain pointers and sizes only.  Every entry point returns 0 on success or
 * a negative status:
 *   -1  verification failed (api.h semantics)     -2  unknown parameter set
 *   -3  HIP runtime / launch error                 -4  bad argument
 *   -5  request exceeds the device-memory budget
 *
 * Parameter sets are chosen at run time (the reference fixes one per build
 * with params.h macros).  Ids: 0..5 = sha2-{128s,128f,192s,192f,256s,256f},
 * 6..11 = shake256-{128s,...,256f}.  The api.h names below act on the
 * process-wide default set (sha2-128f unless CUSPX_PARAM names another set
 * or cuspx_set_default_param() is called).
 *
 * DEBUG_MODE: the reference is shipped with DEBUG_MODE defined
 * (all_option.h:74): keygen seeds become 0..3n-1 and optrand 0xad..ad.  The
 * single-item calls draw their DRBG bytes first and then replace them
 * (sign.cu:342-349, 1950-1954); the batched face_mdp_* / face_ms_mdp_* calls
 * draw nothing (sign.cu:425-436, 2022-2027).  The library starts in that
 * mode; cuspx_set_debug_mode(0) selects DRBG-derived seeds and optrand.
 *
 * Threads (the reference's calls are not reentrant): the host-buffer calls
 * (crypto_sign*, face_mdp_*, face_ms_mdp_*, spx_batch_*) may be made from
 * several threads; on one device they take turns behind a lock.  The DRBG
 * is one shared, locked stream, so draws made by concurrent threads
 * interleave in no fixed order.  The device-resident calls use only the
 * caller's buffers, workspace and stream; calls that overlap need
 * workspaces of their own.  The default parameter set and debug mode are
 * process-wide.
 */
#ifndef CUSPX_H
#define CUSPX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- set-up */

typedef struct {
    int pid, hash, n, full_h, d, a, k, len;
    int sig_bytes, pk_bytes, sk_bytes, seed_bytes;
    char name[32];
} cuspx_param_info;

int cuspx_num_params(void);
int cuspx_param_id(const char *name);            /* "sha2-128f", ... ; -2 if unknown */
int cuspx_get_param_info(int pid, cuspx_param_info *out);
int cuspx_set_default_param(int pid);
/* Work split of set pid at every batch size.  op 0 (sign) / 2 (keygen): lanes
 * per WOTS leaf of the hypertree kernel (1, 2, 4, 8, 16).  op 1 (verify): mean
 * lanes per request of the verify kernel, which runs 256 / lanes requests per
 * workgroup (16, 32, 64; also 128 when len = 67), or 1 for speculative
 * verification (every WOTS chain from every start computed up front, then
 * T_len and the subtree climb per layer; batches of up to 256 requests, larger
 * ones fall back to the table / built-in choice).  0 restores the tuning
 * table / built-in choice.  cuspx_get_split reports the choice for a batch
 * size. */
int cuspx_set_split(int pid, int op, int lanes_per_leaf);
int cuspx_get_split(int pid, int op, uint32_t num);
int cuspx_get_default_param(void);
void cuspx_set_debug_mode(int on);
int cuspx_get_debug_mode(void);
const char *cuspx_last_error(void);

/* ---------------------------------------------- api.h surface (rng.h, api.h) */

/* replaces rng.h:48 randombytes_init / rng.cu:139 */
void randombytes_init(unsigned char *entropy_input, unsigned char *personalization_string,
                      int security_strength);
/* replaces rng.h:53 randombytes / rng.cu:156 */
int randombytes(unsigned char *x, unsigned long long xlen);

/* replaces api.h:19-34 size getters / sign.cu:124-147 */
unsigned long long crypto_sign_secretkeybytes(void);
unsigned long long crypto_sign_publickeybytes(void);
unsigned long long crypto_sign_bytes(void);
unsigned long long crypto_sign_seedbytes(void);
/* replaces api.h:41 / sign.cu:154 */
int crypto_sign_seed_keypair(unsigned char *pk, unsigned char *sk, const unsigned char *seed);
/* replaces api.h:48 / sign.cu:342 (draws 3n DRBG bytes) */
int crypto_sign_keypair(unsigned char *pk, unsigned char *sk);
/* replaces api.h:74 / sign.cu:1924 (draws n DRBG bytes of optrand); messages
 * of 2^32 bytes or more are refused (-4), here and in the calls below */
int crypto_sign_signature(uint8_t *sig, size_t *siglen, const uint8_t *m, size_t mlen,
                          const uint8_t *sk);
/* replaces api.h:80 / sign.cu:3048 */
int crypto_sign_verify(const uint8_t *sig, size_t siglen, const uint8_t *m, size_t mlen,
                       const uint8_t *pk);
/* replaces api.h:95 / sign.cu:3302 */
int crypto_sign(unsigned char *sm, unsigned long long *smlen, const unsigned char *m,
                unsigned long long mlen, const unsigned char *sk);
/* replaces api.h:136 / sign.cu:6058 */
int crypto_sign_open(unsigned char *m, unsigned long long *mlen, const unsigned char *sm,
                     unsigned long long smlen, const unsigned char *pk);

/* ------------------------------------- batched data-parallel surface (api.h) */

/* replaces api.h:59 face_mdp_crypto_sign_keypair / sign.cu:1208.  num keys;
 * pk[num][2n], sk[num][4n].  Seeds: 0..3n-1 in debug mode with no DRBG draw
 * (as the reference), else 3n DRBG bytes per key in order (the reference
 * draws curand).  face_mdp_crypto_sign's optrand likewise: 0xad..ad with no
 * draw in debug mode, else n DRBG bytes per item in order. */
int face_mdp_crypto_sign_keypair(unsigned char *pk, unsigned char *sk, unsigned int num);
/* replaces api.h:107 face_mdp_crypto_sign / sign.cu:5098.  Message i is
 * m + 32*i (SPX_MLEN stride) of mlen <= 32 bytes (-4 otherwise); sm record i
 * at stride SPX_BYTES+32 holds sig || m_i; *smlen = SPX_BYTES+32. */
int face_mdp_crypto_sign(unsigned char *sm, unsigned long long *smlen, const unsigned char *m,
                         unsigned long long mlen, const unsigned char *sk, unsigned int num);
/* replaces api.h:142 face_mdp_crypto_sign_open / sign.cu:6856.  Records at
 * stride SPX_BYTES+32, message length smlen-SPX_BYTES (at most 32, else -4),
 * output slots of 32 bytes.  Returns 0 only if every item verifies (the
 * reference reports item 0 only, sign.cu:6154); failed items get a zeroed
 * slot. */
int face_mdp_crypto_sign_open(unsigned char *m, unsigned long long *mlen, const unsigned char *sm,
                              unsigned long long smlen, const unsigned char *pk, unsigned int num);
/* Multi-stream forms: replace api.h:61 face_ms_mdp_crypto_sign_keypair /
 * sign.cu:1400, api.h:112 face_ms_mdp_crypto_sign / sign.cu:5305 and
 * api.h:151 face_ms_mdp_crypto_sign_open / sign.cu:7080 (the reference's
 * non-const pointers).  Same arguments, outputs and DRBG draws as the
 * face_mdp_* calls above; the batch is cut into up to 2 slices on separate
 * streams (slices of >= 2048 items, >= 4096 for verification; smaller
 * batches run as one slice).  Signing slices run back to back on the GPU, so
 * one slice's output copy overlaps the next slice's kernels; verification
 * and keygen slices run concurrently. */
int face_ms_mdp_crypto_sign_keypair(unsigned char *pk, unsigned char *sk, unsigned int num);
int face_ms_mdp_crypto_sign(unsigned char *sm, unsigned long long *smlen, unsigned char *m,
                            unsigned long long mlen, unsigned char *sk, unsigned int num);
int face_ms_mdp_crypto_sign_open(unsigned char *m, unsigned long long *mlen, unsigned char *sm,
                                 unsigned long long smlen, unsigned char *pk, unsigned int num);

/* ------------------------- runtime-parameterised batches with per-item status */

/* seeds[num][3n] -> pk[num][2n], sk[num][4n] */
int spx_batch_keygen(int pid, uint32_t num, const uint8_t *seeds, uint8_t *pk, uint8_t *sk);
/* message i = msg + moff[i], mlen[i] bytes (any alignment, any length);
 * optrand[num][n] or NULL (then drawn like crypto_sign_signature, item by
 * item); sig[num][SPX_BYTES] */
int spx_batch_sign(int pid, uint32_t num, const uint8_t *sk, const uint8_t *msg,
                   const uint64_t *moff, const uint32_t *mlen, const uint8_t *optrand,
                   uint8_t *sig);
/* status[i] = 0 valid / -1 invalid; returns 0 if all valid, -1 otherwise */
int spx_batch_verify(int pid, uint32_t num, const uint8_t *pk, const uint8_t *sig,
                     const uint8_t *msg, const uint64_t *moff, const uint32_t *mlen,
                     int32_t *status);

/* ---------------------------- device-resident batches (inputs already in HBM)
 * All pointers are device memory on the current HIP device; stream is a
 * hipStream_t (NULL = the library's stream).  msg + moff[i] must be 4-byte
 * aligned and readable up to the next 4-byte boundary.  ws is a workspace
 * of cuspx_workspace_bytes(pid, num, op) bytes (op 0 sign, 1 verify,
 * 2 keygen); it holds the batch's plans, intermediate nodes and work
 * counters, so launches that may overlap need workspaces of their own (a
 * graph keeps its workspace for every replay).  Calls are asynchronous on
 * the stream. */
size_t cuspx_workspace_bytes(int pid, uint32_t num, int op);
int cuspx_sign_device(int pid, uint32_t num, const uint8_t *sk, uint32_t sk_stride,
                      const uint8_t *msg, const uint64_t *moff, const uint32_t *mlen,
                      const uint8_t *optrand, uint8_t *sig, uint64_t sig_stride, void *ws,
                      void *stream);
int cuspx_verify_device(int pid, uint32_t num, const uint8_t *pk, uint32_t pk_stride,
                        const uint8_t *sig, uint64_t sig_stride, const uint8_t *msg,
                        const uint64_t *moff, const uint32_t *mlen, int32_t *status, void *ws,
                        void *stream);
int cuspx_keygen_device(int pid, uint32_t num, const uint8_t *seed, uint32_t seed_stride,
                        uint8_t *pk, uint8_t *sk, void *ws, void *stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc): cuspx_sign_device without the workspace argument.  The
 * workspace is a 128 MiB array the library holds on each GPU until the
 * process ends; the call allocates nothing.  A batch whose workspace
 * (cuspx_workspace_bytes(pid, num, 0)) exceeds 128 MiB is refused with -5
 * before any launch (up to 451 sha2-128f, 114 sha2-256f, 64 sha2-128s
 * requests).  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (other streams, other threads) race and must be
 * serialised by the caller.  stream NULL = HIP's null stream. */
int cuspx_sign_device_noalloc(int pid, uint32_t num, const uint8_t *sk, uint32_t sk_stride,
                              const uint8_t *msg, const uint64_t *moff, const uint32_t *mlen,
                              const uint8_t *optrand, uint8_t *sig, uint64_t sig_stride,
                              void *stream);

/* ------------------------------------------------------------ hipGraph replay
 * Capture the launch sequence of one device-resident batch (the arguments of
 * cuspx_sign_device / cuspx_verify_device) into an instantiated hipGraph and
 * replay it: every replay reads the same device buffers, so the caller
 * refreshes their contents between replays.  No memset nodes are recorded. */
typedef struct cuspx_graph cuspx_graph;
int cuspx_graph_sign(int pid, uint32_t num, const uint8_t *sk, uint32_t sk_stride,
                     const uint8_t *msg, const uint64_t *moff, const uint32_t *mlen,
                     const uint8_t *optrand, uint8_t *sig, uint64_t sig_stride, void *ws,
                     cuspx_graph **out);
int cuspx_graph_verify(int pid, uint32_t num, const uint8_t *pk, uint32_t pk_stride,
                       const uint8_t *sig, uint64_t sig_stride, const uint8_t *msg,
                       const uint64_t *moff, const uint32_t *mlen, int32_t *status, void *ws,
                       cuspx_graph **out);
/* A mixed batch: `groups` single-set sign batches (device-resident, as in
 * cuspx_sign_device; each with its own workspace).  When the non-empty
 * groups are distinct fast sets of one hash family (sha2-128f/192f/256f or
 * shake256-128f/192f/256f) and each is large enough to run at one or two
 * lanes per WOTS leaf, they are signed by ONE persistent kernel whose tile queue spans every set
 * (after each set's prep kernel; T_k and the WOTS gather per set after it).
 * Other mixes run each group's own pipeline on a forked stream.  Not in the
 * reference (its batches are single-set, sign.cu:5098). */
typedef struct {
    int pid;
    uint32_t num;
    const uint8_t *sk;
    uint32_t sk_stride;
    const uint8_t *msg;
    const uint64_t *moff;
    const uint32_t *mlen;
    const uint8_t *optrand;
    uint8_t *sig;
    uint64_t sig_stride;
    void *ws;
} cuspx_sign_group;
int cuspx_graph_sign_mixed(const cuspx_sign_group *groups, int ngroups, cuspx_graph **out);
/* the same mixed batch launched directly on `stream` (the persistent kernel
 * when it applies, else the groups one after another) */
int cuspx_sign_mixed_device(const cuspx_sign_group *groups, int ngroups, void *stream);
int cuspx_graph_launch(cuspx_graph *g, void *stream);
int cuspx_graph_destroy(cuspx_graph *g);

/* ------------------------------------------------------------ serving
 * Continuous-batching service: keys resident in HBM (sk / pk rows and the
 * SHA-256 pub_seed midstate per key), sign / verify requests of any
 * parameter set and message length queued per (set, op), up to `slots`
 * batches in flight on separate streams, batches formed as slots free up and
 * bounded by the device-memory budget.  Not in the reference (it has no
 * serving path); sign results equal crypto_sign_signature for the same key,
 * message and optrand. */
typedef struct {
    uint64_t hbm_budget_bytes;  /* device memory for batches (0: min(free/4, 16 GiB)) */
    uint32_t max_batch;         /* requests per launch (0: 4096) */
    uint32_t slots;             /* batches in flight (0: 4) */
    uint32_t max_msg_bytes;     /* largest message accepted (0: 65536) */
    uint32_t batch_window_us;   /* longest wait for a partial batch while at least half
                                   the slots are busy (fewer: it launches at once) */
    uint32_t max_keys_per_set;  /* resident keys per parameter set (0: 1024, at most 65536) */
    uint32_t long_slots;        /* most slots running batches of messages over 16 KiB
                                   at once (0: all slots); fewer trades p99 and peak
                                   rate for p50, since short requests then never
                                   wait behind long serial message hashes */
} spx_server_config;

typedef struct {
    uint64_t signs, verifies, batches, pending;
    double mean_batch;
    float sign_p50_us, sign_p99_us, verify_p50_us, verify_p99_us;
    uint64_t hbm_bytes;
    int32_t last_error;
} spx_server_statistics;

typedef struct spx_server spx_server;
int spx_server_create(const spx_server_config *cfg, spx_server **out);
/* sk may be NULL for a verify-only key (sign requests on it are refused with
 * -4).  key_id is opaque (set, row and a generation).  At most
 * max_keys_per_set keys per set are resident at once (-5 beyond). */
int spx_server_add_key(spx_server *s, int pid, const uint8_t *sk, const uint8_t *pk,
                       int32_t *key_id);
/* New requests on the key are refused (-4) from now on; requests already
 * queued or running on it complete normally, and its row is reused by a
 * later spx_server_add_key once they have (under a new key_id).  -4 for an
 * unknown or already removed key.  The generation is 11 bits: a key_id kept
 * past 2048 reuses of its row may name that row's current key again. */
int spx_server_remove_key(spx_server *s, int32_t key_id);
/* op 0 = sign (aux = optrand[n] or NULL to draw it like crypto_sign_signature),
 * op 1 = verify (aux = signature).  -5 if the request can never fit one
 * slot's share of the budget, or the queue already holds a budget of bytes */
int spx_server_submit(spx_server *s, int op, int32_t key_id, const uint8_t *msg, uint32_t mlen,
                      const uint8_t *aux, uint64_t *ticket);
/* 0 done (status: 0 ok / -1 invalid signature), 1 timeout (the request stays
 * live: wait again or cancel), <0 error (-4: unknown or cancelled ticket) */
int spx_server_wait(spx_server *s, uint64_t ticket, int timeout_ms, uint8_t *sig_out,
                    int32_t *status);
/* 0: the request was still queued; it leaves the queue (freeing its bytes)
 * and completes with -4, waking any thread waiting on it.  1: it was already
 * launched (the GPU work is not stopped; the result is discarded) or done.
 * -4: unknown ticket. */
int spx_server_cancel(spx_server *s, uint64_t ticket);
int spx_server_stats(spx_server *s, spx_server_statistics *st);
int spx_server_reset_stats(spx_server *s);
/* latency percentiles of one (parameter set, op) since the last reset; with
 * pid = -1, op 0 reports the sign queue wait (submit to launch) and op 1 the
 * sign execution (launch to completion seen by the host) */
int spx_server_latency(spx_server *s, int pid, int op, float *p50_us, float *p99_us,
                       uint64_t *count);
int spx_server_destroy(spx_server *s);

/* ------------------------------------------------------ device plumbing
 * Minimal device-memory, stream and event helpers over the library's own
 * HIP runtime, so callers that hold no HIP runtime of their own (Python via
 * ctypes) can keep batches resident in HBM and time them. */
int cuspx_dev_alloc(void **ptr, size_t bytes);
int cuspx_dev_free(void *ptr);
int cuspx_memcpy_htod(void *dst, const void *src, size_t bytes);
int cuspx_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int cuspx_memcpy_dtod(void *dst, const void *src, size_t bytes);
int cuspx_memset_device(void *dst, int value, size_t bytes);
int cuspx_stream_create(void **stream);
int cuspx_stream_destroy(void *stream);
int cuspx_stream_sync(void *stream);
int cuspx_event_create(void **event);
int cuspx_event_destroy(void *event);
int cuspx_event_record(void *event, void *stream);
int cuspx_event_elapsed_ms(void *start, void *stop, float *ms);
/* compute units, peak engine clock (kHz), HBM bytes, device name */
int cuspx_device_info(int *compute_units, int *clock_khz, size_t *hbm_bytes, char *name,
                      int name_cap);

/* ------------------------------------------------------ kernel timing hooks */
/* When enabled (on = 1; on = 2 keys the totals by "set:kernel"), every
 * launcher records a HIP event after each kernel on its stream.
 * cuspx_profile_read synchronises, then reports per kernel name the summed
 * milliseconds and launch counts since the last reset. */
void cuspx_profile_enable(int on);
int cuspx_profile_read(char *names, int names_cap, float *ms, int *count, int cap);
void cuspx_profile_reset(void);

#ifdef __cplusplus
}
#endif
#endif
