// drbg.hpp -- host AES-256-CTR DRBG (rng.cu:116-241 semantics).
#pragma once
#include <stdint.h>

namespace spx {
void aes256_ecb(const uint8_t key[32], const uint8_t in[16], uint8_t out[16]);
void drbg_init(const uint8_t entropy[48], const uint8_t* pers);
void drbg_bytes(uint8_t* x, uint64_t xlen);
// true: use the round-table AES even where the CPU has the AES instructions
// (both give the same bytes; the native test runs each)
void aes_force_portable(bool on);
}  // namespace spx
