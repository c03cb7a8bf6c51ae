// SHA-256 per FIPS 180-4 §6.2, with the block function on the x86 SHA
// extensions when the host CPU has them (selected once at run time).
#include "sha256.h"

#include <string.h>

#if defined(__x86_64__)
#include <immintrin.h>
#endif

namespace strh {

static const uint32_t K256[64] = {
    0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u, 0x923f82a4u, 0xab1c5ed5u,
    0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u, 0x72be5d74u, 0x80deb1feu, 0x9bdc06a7u, 0xc19bf174u,
    0xe49b69c1u, 0xefbe4786u, 0x0fc19dc6u, 0x240ca1ccu, 0x2de92c6fu, 0x4a7484aau, 0x5cb0a9dcu, 0x76f988dau,
    0x983e5152u, 0xa831c66du, 0xb00327c8u, 0xbf597fc7u, 0xc6e00bf3u, 0xd5a79147u, 0x06ca6351u, 0x14292967u,
    0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu, 0x53380d13u, 0x650a7354u, 0x766a0abbu, 0x81c2c92eu, 0x92722c85u,
    0xa2bfe8a1u, 0xa81a664bu, 0xc24b8b70u, 0xc76c51a3u, 0xd192e819u, 0xd6990624u, 0xf40e3585u, 0x106aa070u,
    0x19a4c116u, 0x1e376c08u, 0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au, 0x5b9cca4fu, 0x682e6ff3u,
    0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u, 0x90befffau, 0xa4506cebu, 0xbef9a3f7u, 0xc67178f2u};

static inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

// FIPS 180-4 §6.2.2, one 512-bit block at a time.
static void blocks_portable(uint32_t h[8], const uint8_t* p, size_t nblocks) {
    uint32_t w[64];
    for (; nblocks; nblocks--, p += 64) {
        for (int t = 0; t < 16; t++)
            w[t] = ((uint32_t)p[4 * t] << 24) | ((uint32_t)p[4 * t + 1] << 16) |
                   ((uint32_t)p[4 * t + 2] << 8) | (uint32_t)p[4 * t + 3];
        for (int t = 16; t < 64; t++) {
            const uint32_t s0 = rotr(w[t - 15], 7) ^ rotr(w[t - 15], 18) ^ (w[t - 15] >> 3);
            const uint32_t s1 = rotr(w[t - 2], 17) ^ rotr(w[t - 2], 19) ^ (w[t - 2] >> 10);
            w[t] = w[t - 16] + s0 + w[t - 7] + s1;
        }
        uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
        for (int t = 0; t < 64; t++) {
            const uint32_t S1 = rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25);
            const uint32_t ch = (e & f) ^ (~e & g);
            const uint32_t t1 = hh + S1 + ch + K256[t] + w[t];
            const uint32_t S0 = rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22);
            const uint32_t maj = (a & b) ^ (a & c) ^ (b & c);
            const uint32_t t2 = S0 + maj;
            hh = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
        }
        h[0] += a; h[1] += b; h[2] += c; h[3] += d;
        h[4] += e; h[5] += f; h[6] += g; h[7] += hh;
    }
}

#if defined(__x86_64__)
// The same compression on SHA-NI: the eight working variables live as
// (A,B,E,F) and (C,D,G,H) in two registers; sha256rnds2 does two rounds,
// sha256msg1/msg2 extend the message schedule four words at a time.
__attribute__((target("sha,sse4.1,ssse3")))
static void blocks_shani(uint32_t h[8], const uint8_t* p, size_t nblocks) {
    const __m128i bswap = _mm_set_epi64x(0x0c0d0e0f08090a0bLL, 0x0405060700010203LL);
    __m128i t = _mm_loadu_si128((const __m128i*)&h[0]);      // A B C D
    __m128i s1 = _mm_loadu_si128((const __m128i*)&h[4]);     // E F G H
    t = _mm_shuffle_epi32(t, 0xB1);
    s1 = _mm_shuffle_epi32(s1, 0x1B);
    __m128i s0 = _mm_alignr_epi8(t, s1, 8);                   // A B E F
    s1 = _mm_blend_epi16(s1, t, 0xF0);                        // C D G H
    for (; nblocks; nblocks--, p += 64) {
        const __m128i save0 = s0, save1 = s1;
        __m128i m[4];
        for (int k = 0; k < 16; k++) {
            __m128i& cur = m[k & 3];
            if (k < 4) {
                cur = _mm_shuffle_epi8(_mm_loadu_si128((const __m128i*)(p + 16 * k)), bswap);
            } else {
                __m128i x = _mm_sha256msg1_epu32(m[k & 3], m[(k + 1) & 3]);
                x = _mm_add_epi32(x, _mm_alignr_epi8(m[(k + 3) & 3], m[(k + 2) & 3], 4));
                cur = _mm_sha256msg2_epu32(x, m[(k + 3) & 3]);
            }
            __m128i wk = _mm_add_epi32(cur, _mm_loadu_si128((const __m128i*)&K256[4 * k]));
            s1 = _mm_sha256rnds2_epu32(s1, s0, wk);
            wk = _mm_shuffle_epi32(wk, 0x0E);
            s0 = _mm_sha256rnds2_epu32(s0, s1, wk);
        }
        s0 = _mm_add_epi32(s0, save0);
        s1 = _mm_add_epi32(s1, save1);
    }
    t = _mm_shuffle_epi32(s0, 0x1B);                          // F E B A
    s1 = _mm_shuffle_epi32(s1, 0xB1);                         // D C H G
    s0 = _mm_blend_epi16(t, s1, 0xF0);                        // D C B A
    s1 = _mm_alignr_epi8(s1, t, 8);                           // H G F E
    _mm_storeu_si128((__m128i*)&h[0], s0);
    _mm_storeu_si128((__m128i*)&h[4], s1);
}

static bool detect_shani() {
    __builtin_cpu_init();
    return __builtin_cpu_supports("sha") && __builtin_cpu_supports("sse4.1");
}
#endif

static bool g_force_portable = false;

bool sha256_uses_shani() {
#if defined(__x86_64__)
    static const bool has = detect_shani();
    return has && !g_force_portable;
#else
    return false;
#endif
}

static void blocks(uint32_t h[8], const uint8_t* p, size_t nblocks) {
#if defined(__x86_64__)
    if (sha256_uses_shani()) { blocks_shani(h, p, nblocks); return; }
#endif
    blocks_portable(h, p, nblocks);
}

Sha256::Sha256() : total(0), fill(0) {
    static const uint32_t iv[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au,
                                   0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};
    memcpy(h, iv, sizeof(iv));
}

void Sha256::update(const void* data, size_t len) {
    const uint8_t* p = (const uint8_t*)data;
    total += len;
    if (fill) {
        const size_t take = len < 64 - fill ? len : 64 - fill;
        memcpy(buf + fill, p, take);
        fill += take; p += take; len -= take;
        if (fill < 64) return;
        blocks(h, buf, 1);
        fill = 0;
    }
    const size_t nb = len / 64;
    if (nb) { blocks(h, p, nb); p += nb * 64; len -= nb * 64; }
    if (len) { memcpy(buf, p, len); fill = len; }
}

// Padding, FIPS 180-4 §5.1.1: 0x80, zeros to 56 mod 64, 64-bit big-endian length.
void Sha256::finish(uint8_t out[32]) {
    const uint64_t bits = total * 8;
    uint8_t pad[72];
    memset(pad, 0, sizeof(pad));
    pad[0] = 0x80;
    const size_t padlen = (fill < 56) ? (56 - fill) : (120 - fill);
    for (int i = 0; i < 8; i++) pad[padlen + i] = (uint8_t)(bits >> (56 - 8 * i));
    update(pad, padlen + 8);
    for (int i = 0; i < 8; i++) {
        out[4 * i] = (uint8_t)(h[i] >> 24);
        out[4 * i + 1] = (uint8_t)(h[i] >> 16);
        out[4 * i + 2] = (uint8_t)(h[i] >> 8);
        out[4 * i + 3] = (uint8_t)h[i];
    }
}

}  // namespace strh

// Host-only test hooks (no device involvement): hash a buffer, choose the path.
extern "C" __attribute__((visibility("default")))
int strh_sha256(const void* data, size_t len, uint8_t out[32]) {
    strh::Sha256 s;
    s.update(data, len);
    s.finish(out);
    return 0;
}

extern "C" __attribute__((visibility("default")))
int strh_sha256_force_portable(int enable) {
    strh::g_force_portable = enable != 0;
    return strh::sha256_uses_shani() ? 1 : 0;
}
