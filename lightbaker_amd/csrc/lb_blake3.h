// BLAKE3, unkeyed, 32-byte output: the incremental hasher the tile dependencies stream into.
#pragma once

#include <cstddef>
#include <cstdint>

namespace lbk {

class Blake3 {
public:
    Blake3();
    void reset();
    void update(const void *data, size_t n);
    void finalize(uint8_t out[32]) const;

private:
    void push_block();
    void push_chunk();
    uint32_t cv_[8];
    uint64_t chunk_;
    uint32_t blocks_;
    uint8_t buf_[64];
    uint32_t buf_len_;
    uint32_t stack_[54][8];  // a subtree a level: 2^54 chunks
    int depth_;
};

}  // namespace lbk

extern "C" int lb_blake3(const void *data, size_t n, uint8_t out[32]);
