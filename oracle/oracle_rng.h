// TEST INFRASTRUCTURE ONLY — part of the CPU oracle (see oracle_math.h).
//
// PCG32 (util/rng.h:7-148) and MurmurHash64A (util/hash.h:9-80), restated from
// O'Neill's PCG paper and Appleby's MurmurHash2-64A as the reference uses them.
#pragma once

#include <cstdint>
#include <cstring>
#include <algorithm>

namespace orc {

constexpr uint64_t kPcgState = 0x853c49e6748fea9bULL;
constexpr uint64_t kPcgStream = 0xda3e39cb94b95bdbULL;
constexpr uint64_t kPcgMult = 0x5851f42d4c957f2dULL;

// util/rng.h:15-22 — Stafford's "Mix13" 64-bit finaliser.
inline uint64_t mix_bits(uint64_t v) {
    v ^= (v >> 31);
    v *= 0x7fb5d329728ea185ULL;
    v ^= (v >> 27);
    v *= 0x81dadef4bc2dd44dULL;
    v ^= (v >> 33);
    return v;
}

struct Pcg32 {
    uint64_t state = kPcgState, inc = kPcgStream;

    uint32_t next_u32() {  // rng.h:80-86 (XSH-RR output of the old state)
        uint64_t old = state;
        state = old * kPcgMult + inc;
        uint32_t xorshifted = uint32_t(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = uint32_t(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((~rot + 1u) & 31));
    }
    void set_sequence(uint64_t seq, uint64_t offset) {  // rng.h:117-123
        state = 0u;
        inc = (seq << 1u) | 1u;
        next_u32();
        state += offset;
        next_u32();
    }
    void set_sequence(uint64_t seq) { set_sequence(seq, mix_bits(seq)); }
    float next_float() {  // rng.h:125-128
        return std::min<float>(float(0x1.fffffep-1), float(next_u32()) * 0x1p-32f);
    }
    void advance(int64_t idelta) {  // rng.h:135-148 (Brown's log-time jump)
        uint64_t cur_mult = kPcgMult, cur_plus = inc, acc_mult = 1u, acc_plus = 0u;
        uint64_t delta = uint64_t(idelta);
        while (delta > 0) {
            if (delta & 1) {
                acc_mult *= cur_mult;
                acc_plus = acc_plus * cur_mult + cur_plus;
            }
            cur_plus = (cur_mult + 1) * cur_plus;
            cur_mult *= cur_mult;
            delta /= 2;
        }
        state = acc_mult * state + acc_plus;
    }
};

// util/hash.h:9-54 — MurmurHash2, 64-bit version A.
inline uint64_t murmur64a(const unsigned char *key, size_t len, uint64_t seed) {
    const uint64_t m = 0xc6a4a7935bd1e995ULL;
    const int r = 47;
    uint64_t h = seed ^ (len * m);
    size_t nblocks = len / 8;
    for (size_t i = 0; i < nblocks; ++i) {
        uint64_t k;
        std::memcpy(&k, key + 8 * i, 8);
        k *= m;
        k ^= k >> r;
        k *= m;
        h ^= k;
        h *= m;
    }
    const unsigned char *tail = key + 8 * nblocks;
    switch (len & 7) {
        case 7: h ^= uint64_t(tail[6]) << 48; [[fallthrough]];
        case 6: h ^= uint64_t(tail[5]) << 40; [[fallthrough]];
        case 5: h ^= uint64_t(tail[4]) << 32; [[fallthrough]];
        case 4: h ^= uint64_t(tail[3]) << 24; [[fallthrough]];
        case 3: h ^= uint64_t(tail[2]) << 16; [[fallthrough]];
        case 2: h ^= uint64_t(tail[1]) << 8; [[fallthrough]];
        case 1: h ^= uint64_t(tail[0]); h *= m;
    }
    h ^= h >> r;
    h *= m;
    h ^= h >> r;
    return h;
}

}  // namespace orc
