/*
 * perset.c -- libspx_<set>.so: one shared object per parameter set.
 *
 * The reference builds one library per parameter set, the set chosen at
 * compile time by -DSHA256 / -DSHAKE256 and -DSPX_<size> (Makefile:109-155),
 * so its callers keep api.h unchanged and pick the set by the library they
 * link.  libcuspx.so serves all twelve sets behind a run-time default; this
 * file, compiled once per set with -DCUSPX_SET_NAME="<set>", makes a library
 * that
 *   - makes <set> libcuspx.so's default parameter set when it is loaded, and
 *   - defines every api.h / rng.h function a caller of the reference links
 *     against, each calling libcuspx.so's definition of the same name,
 * so `-lspx_sha2_128f` alone satisfies such a caller's link (also under
 * --as-needed).  libcuspx.so's definitions are looked up in libcuspx.so
 * itself, so a caller that also links libcuspx.so (after this library:
 * listed before it, --as-needed drops this one, whose names libcuspx.so
 * already defines) reaches the same code either way.  A library whose
 * lookups fail stops the process at load with a message: there is no other
 * path to fall back to.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <stdio.h>
#include <stdlib.h>

#include "cuspx.h"

#ifndef CUSPX_SET_NAME
#error "compile with -DCUSPX_SET_NAME=\"<parameter set>\""
#endif

static void (*p_randombytes_init)(unsigned char*, unsigned char*, int);
static int (*p_randombytes)(unsigned char*, unsigned long long);
static unsigned long long (*p_secretkeybytes)(void);
static unsigned long long (*p_publickeybytes)(void);
static unsigned long long (*p_bytes)(void);
static unsigned long long (*p_seedbytes)(void);
static int (*p_seed_keypair)(unsigned char*, unsigned char*, const unsigned char*);
static int (*p_keypair)(unsigned char*, unsigned char*);
static int (*p_signature)(uint8_t*, size_t*, const uint8_t*, size_t, const uint8_t*);
static int (*p_verify)(const uint8_t*, size_t, const uint8_t*, size_t, const uint8_t*);
static int (*p_sign)(unsigned char*, unsigned long long*, const unsigned char*,
                     unsigned long long, const unsigned char*);
static int (*p_open)(unsigned char*, unsigned long long*, const unsigned char*,
                     unsigned long long, const unsigned char*);
static int (*p_mdp_keypair)(unsigned char*, unsigned char*, unsigned int);
static int (*p_mdp_sign)(unsigned char*, unsigned long long*, const unsigned char*,
                         unsigned long long, const unsigned char*, unsigned int);
static int (*p_mdp_open)(unsigned char*, unsigned long long*, const unsigned char*,
                         unsigned long long, const unsigned char*, unsigned int);
static int (*p_ms_keypair)(unsigned char*, unsigned char*, unsigned int);
static int (*p_ms_sign)(unsigned char*, unsigned long long*, unsigned char*,
                        unsigned long long, unsigned char*, unsigned int);
static int (*p_ms_open)(unsigned char*, unsigned long long*, unsigned char*,
                        unsigned long long, unsigned char*, unsigned int);

static void* need(void* lib, const char* name) {
    void* f = dlsym(lib, name);
    if (!f) {
        fprintf(stderr, "libspx_%s: %s not found in libcuspx.so\n", CUSPX_SET_NAME, name);
        abort();
    }
    return f;
}

__attribute__((constructor)) static void spx_select_set(void) {
    /* libcuspx.so is the object that defines cuspx_param_id (this library
     * does not); open it by the path it was loaded from */
    Dl_info info;
    void* lib = NULL;
    if (dladdr((void*)&cuspx_param_id, &info) && info.dli_fname)
        lib = dlopen(info.dli_fname, RTLD_LAZY | RTLD_NOLOAD);
    if (!lib) {
        fprintf(stderr, "libspx_%s: libcuspx.so is not loaded\n", CUSPX_SET_NAME);
        abort();
    }
    *(void**)&p_randombytes_init = need(lib, "randombytes_init");
    *(void**)&p_randombytes = need(lib, "randombytes");
    *(void**)&p_secretkeybytes = need(lib, "crypto_sign_secretkeybytes");
    *(void**)&p_publickeybytes = need(lib, "crypto_sign_publickeybytes");
    *(void**)&p_bytes = need(lib, "crypto_sign_bytes");
    *(void**)&p_seedbytes = need(lib, "crypto_sign_seedbytes");
    *(void**)&p_seed_keypair = need(lib, "crypto_sign_seed_keypair");
    *(void**)&p_keypair = need(lib, "crypto_sign_keypair");
    *(void**)&p_signature = need(lib, "crypto_sign_signature");
    *(void**)&p_verify = need(lib, "crypto_sign_verify");
    *(void**)&p_sign = need(lib, "crypto_sign");
    *(void**)&p_open = need(lib, "crypto_sign_open");
    *(void**)&p_mdp_keypair = need(lib, "face_mdp_crypto_sign_keypair");
    *(void**)&p_mdp_sign = need(lib, "face_mdp_crypto_sign");
    *(void**)&p_mdp_open = need(lib, "face_mdp_crypto_sign_open");
    *(void**)&p_ms_keypair = need(lib, "face_ms_mdp_crypto_sign_keypair");
    *(void**)&p_ms_sign = need(lib, "face_ms_mdp_crypto_sign");
    *(void**)&p_ms_open = need(lib, "face_ms_mdp_crypto_sign_open");
    const int pid = cuspx_param_id(CUSPX_SET_NAME);
    if (pid < 0 || cuspx_set_default_param(pid) != 0) {
        fprintf(stderr, "libspx_%s: libcuspx.so does not know this parameter set\n",
                CUSPX_SET_NAME);
        abort();
    }
}

void randombytes_init(unsigned char* entropy_input, unsigned char* personalization_string,
                      int security_strength) {
    p_randombytes_init(entropy_input, personalization_string, security_strength);
}
int randombytes(unsigned char* x, unsigned long long xlen) { return p_randombytes(x, xlen); }

unsigned long long crypto_sign_secretkeybytes(void) { return p_secretkeybytes(); }
unsigned long long crypto_sign_publickeybytes(void) { return p_publickeybytes(); }
unsigned long long crypto_sign_bytes(void) { return p_bytes(); }
unsigned long long crypto_sign_seedbytes(void) { return p_seedbytes(); }

int crypto_sign_seed_keypair(unsigned char* pk, unsigned char* sk, const unsigned char* seed) {
    return p_seed_keypair(pk, sk, seed);
}
int crypto_sign_keypair(unsigned char* pk, unsigned char* sk) { return p_keypair(pk, sk); }
int crypto_sign_signature(uint8_t* sig, size_t* siglen, const uint8_t* m, size_t mlen,
                          const uint8_t* sk) {
    return p_signature(sig, siglen, m, mlen, sk);
}
int crypto_sign_verify(const uint8_t* sig, size_t siglen, const uint8_t* m, size_t mlen,
                       const uint8_t* pk) {
    return p_verify(sig, siglen, m, mlen, pk);
}
int crypto_sign(unsigned char* sm, unsigned long long* smlen, const unsigned char* m,
                unsigned long long mlen, const unsigned char* sk) {
    return p_sign(sm, smlen, m, mlen, sk);
}
int crypto_sign_open(unsigned char* m, unsigned long long* mlen, const unsigned char* sm,
                     unsigned long long smlen, const unsigned char* pk) {
    return p_open(m, mlen, sm, smlen, pk);
}

int face_mdp_crypto_sign_keypair(unsigned char* pk, unsigned char* sk, unsigned int num) {
    return p_mdp_keypair(pk, sk, num);
}
int face_mdp_crypto_sign(unsigned char* sm, unsigned long long* smlen, const unsigned char* m,
                         unsigned long long mlen, const unsigned char* sk, unsigned int num) {
    return p_mdp_sign(sm, smlen, m, mlen, sk, num);
}
int face_mdp_crypto_sign_open(unsigned char* m, unsigned long long* mlen, const unsigned char* sm,
                              unsigned long long smlen, const unsigned char* pk, unsigned int num) {
    return p_mdp_open(m, mlen, sm, smlen, pk, num);
}
int face_ms_mdp_crypto_sign_keypair(unsigned char* pk, unsigned char* sk, unsigned int num) {
    return p_ms_keypair(pk, sk, num);
}
int face_ms_mdp_crypto_sign(unsigned char* sm, unsigned long long* smlen, unsigned char* m,
                            unsigned long long mlen, unsigned char* sk, unsigned int num) {
    return p_ms_sign(sm, smlen, m, mlen, sk, num);
}
int face_ms_mdp_crypto_sign_open(unsigned char* m, unsigned long long* mlen, unsigned char* sm,
                                 unsigned long long smlen, unsigned char* pk, unsigned int num) {
    return p_ms_open(m, mlen, sm, smlen, pk, num);
}
