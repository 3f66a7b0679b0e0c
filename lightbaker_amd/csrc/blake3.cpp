// BLAKE3 (the hash the reference keys its cache with: deps.rs:60-121, bake.rs:461-477), written
// from the published specification (O'Connor, Aumasson, Neves, Wilcox-O'Hearn, "BLAKE3: one
// function, fast everywhere", 2020): the 7-round compression function over 64-byte blocks,
// 1 KiB chunks chained by their counters, and a binary tree of parent nodes built with a stack of
// chaining values; 32-byte output, unkeyed. Scalar and portable: the host hashes a few MB a bake.
#include "lb_blake3.h"

#include <cstring>

namespace lbk {
namespace {

constexpr uint32_t IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                            0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
constexpr uint8_t SCHEDULE[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
enum : uint32_t { CHUNK_START = 1, CHUNK_END = 2, PARENT = 4, ROOT = 8 };

inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

inline void mix(uint32_t *s, int a, int b, int c, int d, uint32_t x, uint32_t y) {
    s[a] = s[a] + s[b] + x;
    s[d] = rotr(s[d] ^ s[a], 16);
    s[c] = s[c] + s[d];
    s[b] = rotr(s[b] ^ s[c], 12);
    s[a] = s[a] + s[b] + y;
    s[d] = rotr(s[d] ^ s[a], 8);
    s[c] = s[c] + s[d];
    s[b] = rotr(s[b] ^ s[c], 7);
}

// the first 8 words of the compression's output: the chaining value
void compress(const uint32_t cv[8], const uint8_t block[64], uint64_t counter, uint32_t len,
              uint32_t flags, uint32_t out[8]) {
    uint32_t m[16];
    for (int i = 0; i < 16; i++)
        m[i] = (uint32_t)block[4 * i] | (uint32_t)block[4 * i + 1] << 8 |
               (uint32_t)block[4 * i + 2] << 16 | (uint32_t)block[4 * i + 3] << 24;
    uint32_t s[16] = {cv[0], cv[1], cv[2], cv[3], cv[4], cv[5], cv[6], cv[7],
                      IV[0], IV[1], IV[2], IV[3], (uint32_t)counter, (uint32_t)(counter >> 32),
                      len, flags};
    for (int r = 0; r < 7; r++) {
        mix(s, 0, 4, 8, 12, m[0], m[1]);
        mix(s, 1, 5, 9, 13, m[2], m[3]);
        mix(s, 2, 6, 10, 14, m[4], m[5]);
        mix(s, 3, 7, 11, 15, m[6], m[7]);
        mix(s, 0, 5, 10, 15, m[8], m[9]);
        mix(s, 1, 6, 11, 12, m[10], m[11]);
        mix(s, 2, 7, 8, 13, m[12], m[13]);
        mix(s, 3, 4, 9, 14, m[14], m[15]);
        if (r < 6) {
            uint32_t t[16];
            for (int i = 0; i < 16; i++) t[i] = m[SCHEDULE[i]];
            std::memcpy(m, t, sizeof m);
        }
    }
    for (int i = 0; i < 8; i++) out[i] = s[i] ^ s[i + 8];
}

void parent_cv(const uint32_t left[8], const uint32_t right[8], uint32_t flags, uint32_t out[8]) {
    uint8_t block[64];
    for (int i = 0; i < 8; i++)
        for (int k = 0; k < 4; k++) {
            block[4 * i + k] = (uint8_t)(left[i] >> (8 * k));
            block[32 + 4 * i + k] = (uint8_t)(right[i] >> (8 * k));
        }
    compress(IV, block, 0, 64, PARENT | flags, out);
}

}  // namespace

Blake3::Blake3() { reset(); }

void Blake3::reset() {
    std::memcpy(cv_, IV, sizeof cv_);
    chunk_ = 0;
    blocks_ = 0;
    buf_len_ = 0;
    depth_ = 0;
}

// The current chunk's block in buf_ is full and more input follows: compress it into the chunk.
void Blake3::push_block() {
    uint32_t flags = blocks_ == 0 ? CHUNK_START : 0;
    compress(cv_, buf_, chunk_, 64, flags, cv_);
    blocks_++;
    buf_len_ = 0;
}

// The current chunk (16 blocks, its last in buf_) is complete and more input follows.
void Blake3::push_chunk() {
    uint32_t cv[8];
    compress(cv_, buf_, chunk_, 64, (blocks_ == 0 ? CHUNK_START : 0) | CHUNK_END, cv);
    chunk_++;
    // merge completed subtrees: one per trailing zero of the chunk count
    for (uint64_t total = chunk_; (total & 1) == 0; total >>= 1) parent_cv(stack_[--depth_], cv, 0, cv);
    std::memcpy(stack_[depth_++], cv, sizeof cv);
    std::memcpy(cv_, IV, sizeof cv_);
    blocks_ = 0;
    buf_len_ = 0;
}

void Blake3::update(const void *data, size_t n) {
    const uint8_t *p = static_cast<const uint8_t *>(data);
    while (n) {
        if (buf_len_ == 64) {
            // a full block waits: it is the chunk's last when the chunk holds 16
            if (blocks_ == 15)
                push_chunk();
            else
                push_block();
        }
        size_t take = 64 - buf_len_;
        if (take > n) take = n;
        std::memcpy(buf_ + buf_len_, p, take);
        buf_len_ += (uint32_t)take;
        p += take;
        n -= take;
    }
}

void Blake3::finalize(uint8_t out[32]) const {
    // the last chunk's last block, then the stack folded from the top; ROOT on the final one
    uint8_t block[64] = {};
    std::memcpy(block, buf_, buf_len_);
    uint32_t flags = (blocks_ == 0 ? CHUNK_START : 0) | CHUNK_END;
    uint32_t cv[8];
    if (depth_ == 0) {
        compress(cv_, block, chunk_, buf_len_, flags | ROOT, cv);
    } else {
        compress(cv_, block, chunk_, buf_len_, flags, cv);
        for (int d = depth_ - 1; d >= 0; d--) parent_cv(stack_[d], cv, d == 0 ? ROOT : 0, cv);
    }
    for (int i = 0; i < 8; i++)
        for (int k = 0; k < 4; k++) out[4 * i + k] = (uint8_t)(cv[i] >> (8 * k));
}

}  // namespace lbk

extern "C" int lb_blake3(const void *data, size_t n, uint8_t out[32]) {
    if ((n && !data) || !out) return 1;
    lbk::Blake3 h;
    h.update(data, n);
    h.finalize(out);
    return 0;
}
