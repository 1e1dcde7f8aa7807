// 64-bit content fingerprint of host byte ranges: tells apart scenes, material tables, envmaps
// and checkpoint payloads. Not cryptographic; it catches a snapshot resumed into another
// scene and bytes changed in storage.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <cstring>

namespace rgpu {

inline uint64_t fp_mix(uint64_t h) {  // the splitmix64 finaliser
    h ^= h >> 30;
    h *= 0xbf58476d1ce4e5b9ull;
    h ^= h >> 27;
    h *= 0x94d049bb133111ebull;
    h ^= h >> 31;
    return h;
}

// Four independent multiply-rotate lanes over 32-byte blocks (so the loop runs near memory
// speed), folded with the length and the tail at the end.
inline uint64_t fingerprint64(const void* data, size_t bytes, uint64_t seed = 0) {
    const uint8_t* p = static_cast<const uint8_t*>(data);
    const uint64_t k = 0x9e3779b97f4a7c15ull;
    uint64_t a = seed ^ 0x243f6a8885a308d3ull, b = seed ^ 0x13198a2e03707344ull;
    uint64_t c = seed ^ 0xa4093822299f31d0ull, d = seed ^ 0x082efa98ec4e6c89ull;
    size_t i = 0;
    for (; i + 32 <= bytes; i += 32) {
        uint64_t w[4];
        std::memcpy(w, p + i, 32);
        a = ((a ^ w[0]) * k);
        a = (a << 29) | (a >> 35);
        b = ((b ^ w[1]) * k);
        b = (b << 29) | (b >> 35);
        c = ((c ^ w[2]) * k);
        c = (c << 29) | (c >> 35);
        d = ((d ^ w[3]) * k);
        d = (d << 29) | (d >> 35);
    }
    uint64_t h = fp_mix(a) ^ fp_mix(b + 1) ^ fp_mix(c + 2) ^ fp_mix(d + 3) ^ fp_mix(bytes + k);
    for (; i < bytes; ++i) h = fp_mix(h ^ (uint64_t)p[i] ^ ((uint64_t)i << 8));
    return fp_mix(h);
}

}  // namespace rgpu
