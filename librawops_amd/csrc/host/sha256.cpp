// SHA-256 compression and padding as specified by FIPS 180-4 sections 4.2.2,
// 5.1.1 and 6.2.2 (round constants: the first 32 bits of the fractional parts
// of the cube roots of the first 64 primes).
#include "sha256.hpp"

#include <cstring>

namespace rawengine::detail {
namespace {

constexpr std::uint32_t kRound[64] = {
    0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u, 0x923f82a4u, 0xab1c5ed5u,
    0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u, 0x72be5d74u, 0x80deb1feu, 0x9bdc06a7u, 0xc19bf174u,
    0xe49b69c1u, 0xefbe4786u, 0x0fc19dc6u, 0x240ca1ccu, 0x2de92c6fu, 0x4a7484aau, 0x5cb0a9dcu, 0x76f988dau,
    0x983e5152u, 0xa831c66du, 0xb00327c8u, 0xbf597fc7u, 0xc6e00bf3u, 0xd5a79147u, 0x06ca6351u, 0x14292967u,
    0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu, 0x53380d13u, 0x650a7354u, 0x766a0abbu, 0x81c2c92eu, 0x92722c85u,
    0xa2bfe8a1u, 0xa81a664bu, 0xc24b8b70u, 0xc76c51a3u, 0xd192e819u, 0xd6990624u, 0xf40e3585u, 0x106aa070u,
    0x19a4c116u, 0x1e376c08u, 0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au, 0x5b9cca4fu, 0x682e6ff3u,
    0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u, 0x90befffau, 0xa4506cebu, 0xbef9a3f7u, 0xc67178f2u};

inline std::uint32_t rotr(std::uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

}  // namespace

void Sha256::compress(const std::uint8_t* block) {
    std::uint32_t w[64];
    for (int t = 0; t < 16; ++t)
        w[t] = std::uint32_t(block[4 * t]) << 24 | std::uint32_t(block[4 * t + 1]) << 16 |
               std::uint32_t(block[4 * t + 2]) << 8 | std::uint32_t(block[4 * t + 3]);
    for (int t = 16; t < 64; ++t) {
        const std::uint32_t s0 = rotr(w[t - 15], 7) ^ rotr(w[t - 15], 18) ^ (w[t - 15] >> 3);
        const std::uint32_t s1 = rotr(w[t - 2], 17) ^ rotr(w[t - 2], 19) ^ (w[t - 2] >> 10);
        w[t] = w[t - 16] + s0 + w[t - 7] + s1;
    }
    std::uint32_t v[8];
    std::memcpy(v, h_, sizeof v);
    for (int t = 0; t < 64; ++t) {
        const std::uint32_t e = v[4], a = v[0];
        const std::uint32_t t1 = v[7] + (rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25)) + ((e & v[5]) ^ (~e & v[6])) +
                                 kRound[t] + w[t];
        const std::uint32_t t2 = (rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22)) + ((a & v[1]) ^ (a & v[2]) ^ (v[1] & v[2]));
        for (int i = 7; i > 0; --i) v[i] = v[i - 1];
        v[4] += t1;
        v[0] = t1 + t2;
    }
    for (int i = 0; i < 8; ++i) h_[i] += v[i];
}

void Sha256::update(const void* data, std::size_t size) {
    auto* p = static_cast<const std::uint8_t*>(data);
    total_ += size;
    if (fill_) {
        const std::size_t take = size < 64 - fill_ ? size : 64 - fill_;
        std::memcpy(pending_ + fill_, p, take);
        fill_ += take; p += take; size -= take;
        if (fill_ < 64) return;
        compress(pending_);
        fill_ = 0;
    }
    for (; size >= 64; p += 64, size -= 64) compress(p);
    std::memcpy(pending_, p, size);
    fill_ = size;
}

std::array<std::uint8_t, 32> Sha256::finish() {
    const std::uint64_t bits = total_ * 8;
    pending_[fill_++] = 0x80;
    if (fill_ > 56) {
        std::memset(pending_ + fill_, 0, 64 - fill_);
        compress(pending_);
        fill_ = 0;
    }
    std::memset(pending_ + fill_, 0, 56 - fill_);
    for (int i = 0; i < 8; ++i) pending_[56 + i] = std::uint8_t(bits >> (56 - 8 * i));
    compress(pending_);
    std::array<std::uint8_t, 32> out{};
    for (int i = 0; i < 8; ++i)
        for (int b = 0; b < 4; ++b) out[4 * i + b] = std::uint8_t(h_[i] >> (24 - 8 * b));
    return out;
}

}  // namespace rawengine::detail
