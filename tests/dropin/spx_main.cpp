// A caller written the way the reference's test/spx.cu drives the api.h
// surface (keypair, crypto_sign, crypto_sign_open, in-place open, bit flips),
// compiled against include/cuspx_api.h alone with the reference's set flags:
//
//   g++ -DSHA256 -DSPX_128F -I include spx_main.cpp -L cuspx_amd -lcuspx
//
// or, with -DPLAIN_API_H, against the plain declarations and size macros
// only (cuspx.h + cuspx_params.h, as a caller keeping its api.h and
// params.h would), linked to the per-set library, which alone picks the set:
//
//   g++ -DSHA256 -DSPX_128F -DPLAIN_API_H -I include spx_main.cpp -L cuspx_amd -lspx_sha2_128f
//
// The DRBG is seeded with the NIST KAT entropy 00..2f; the library runs in
// the reference's shipped DEBUG_MODE (keygen seed 0..3n-1, optrand 0xad).
// Prints "pk", "sk", "m" and "sm" lines in hex for the test to compare
// with the CPU oracle, then "OK" and exit status 0 if every check passed.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#ifdef PLAIN_API_H
#include "cuspx.h"
#include "cuspx_params.h"
#else
#include "cuspx_api.h"
#endif

static void put_hex(const char* tag, const unsigned char* b, size_t n) {
    printf("%s ", tag);
    for (size_t i = 0; i < n; i++) printf("%02x", b[i]);
    printf("\n");
}

static int failures = 0;
static void expect(bool ok, const char* what) {
    if (!ok) {
        printf("  X %s\n", what);
        failures++;
    }
}

int main() {
    setbuf(stdout, NULL);
    unsigned char entropy[48];
    for (int i = 0; i < 48; i++) entropy[i] = (unsigned char)i;
    randombytes_init(entropy, NULL, 256);

    std::vector<unsigned char> pk(SPX_PK_BYTES), sk(SPX_SK_BYTES), m(SPX_MLEN);
    std::vector<unsigned char> sm(SPX_BYTES + SPX_MLEN), mout(SPX_BYTES + SPX_MLEN);
    unsigned long long smlen = 0, mlen = 0;

    expect(crypto_sign_bytes() == SPX_BYTES, "crypto_sign_bytes != SPX_BYTES");
    expect(crypto_sign_publickeybytes() == CRYPTO_PUBLICKEYBYTES, "public key size");
    expect(crypto_sign_secretkeybytes() == CRYPTO_SECRETKEYBYTES, "secret key size");
    expect(crypto_sign_seedbytes() == CRYPTO_SEEDBYTES, "seed size");

    randombytes(m.data(), SPX_MLEN);
    expect(crypto_sign_keypair(pk.data(), sk.data()) == 0, "crypto_sign_keypair failed");
    expect(crypto_sign(sm.data(), &smlen, m.data(), SPX_MLEN, sk.data()) == 0,
           "crypto_sign failed");
    expect(smlen == SPX_BYTES + SPX_MLEN, "smlen != SPX_BYTES + SPX_MLEN");
    put_hex("pk", pk.data(), SPX_PK_BYTES);
    put_hex("sk", sk.data(), SPX_SK_BYTES);
    put_hex("m", m.data(), SPX_MLEN);
    put_hex("sm", sm.data(), (size_t)smlen);

    expect(crypto_sign_open(mout.data(), &mlen, sm.data(), smlen, pk.data()) == 0,
           "verification failed");
    expect(mlen == SPX_MLEN, "recovered mlen");
    expect(memcmp(m.data(), mout.data(), SPX_MLEN) == 0, "recovered message");

    // detached signature and verification
    std::vector<unsigned char> sig(SPX_BYTES);
    size_t siglen = 0;
    expect(crypto_sign_signature(sig.data(), &siglen, m.data(), SPX_MLEN, sk.data()) == 0,
           "crypto_sign_signature failed");
    expect(siglen == SPX_BYTES && memcmp(sig.data(), sm.data(), SPX_BYTES) == 0,
           "detached signature differs from crypto_sign's");
    expect(crypto_sign_verify(sig.data(), siglen, m.data(), SPX_MLEN, pk.data()) == 0,
           "crypto_sign_verify failed");

    // in place: the message is recovered into the signed-message buffer
    std::vector<unsigned char> inplace(sm);
    expect(crypto_sign_open(inplace.data(), &mlen, inplace.data(), smlen, pk.data()) == 0,
           "in-place verification failed");
    expect(mlen == SPX_MLEN && memcmp(inplace.data(), m.data(), SPX_MLEN) == 0,
           "in-place recovered message");

    // a flipped message bit invalidates; the output is cleared
    sm[smlen - 1] ^= 1;
    expect(crypto_sign_open(mout.data(), &mlen, sm.data(), smlen, pk.data()) != 0,
           "flipping a message bit did not invalidate the signature");
    expect(mlen == 0, "mlen not cleared on failure");
    sm[smlen - 1] ^= 1;

    // one flipped bit in every n-byte hash of the signature invalidates it
    int missed = 0;
    for (unsigned long long j = 0; j < smlen - SPX_MLEN; j += SPX_N) {
        sm[j] ^= 1;
        if (crypto_sign_open(mout.data(), &mlen, sm.data(), smlen, pk.data()) == 0) missed++;
        sm[j] ^= 1;
    }
    expect(missed == 0, "a flipped signature hash did not invalidate the signature");

    // shorter than a signature: rejected
    expect(crypto_sign_open(mout.data(), &mlen, sm.data(), SPX_BYTES - 1, pk.data()) != 0,
           "smlen < SPX_BYTES accepted");

    if (failures) {
        printf("FAILED %d\n", failures);
        return 1;
    }
    printf("OK %s\n", CUSPX_PARAM_NAME);
    return 0;
}
