// SHA-256 (FIPS 180-4) for source fingerprints and tile-cache signatures.
// Host code: a digest is a sequential chain over 64-byte blocks, so there is
// nothing to spread over the GPU; the render path never waits on it.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>

namespace rawengine::detail {

class Sha256 {
public:
    void update(const void* data, std::size_t size);
    std::array<std::uint8_t, 32> finish();  // once per object

private:
    void compress(const std::uint8_t* block);
    std::uint32_t h_[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au,
                           0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};
    std::uint8_t pending_[64] = {};
    std::size_t fill_ = 0;
    std::uint64_t total_ = 0;
};

}  // namespace rawengine::detail
