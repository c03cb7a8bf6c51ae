// SHA-256 (FIPS 180-4) for the state digest (Sources/StrouhalCore/Sim.swift:441-444
// hashes the raw DDF buffer with CryptoKit; this is the host-side equivalent).
#pragma once

#include <stddef.h>
#include <stdint.h>

namespace strh {

struct Sha256 {
    uint32_t h[8];
    uint8_t buf[64];
    uint64_t total;   // bytes absorbed
    size_t fill;      // bytes pending in buf

    Sha256();
    void update(const void* data, size_t len);
    void finish(uint8_t out[32]);
};

// True when the x86 SHA extensions drive the block function.
bool sha256_uses_shani();

}  // namespace strh
