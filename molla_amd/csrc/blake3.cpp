// BLAKE3 hash mode (host).  Algorithm as published by its authors (2020);
// the constants below are the ones that publication fixes: the IV is the
// SHA-256 / BLAKE2s initial state and the message schedule is one fixed
// permutation applied between rounds.
#include "blake3.hpp"

#include <algorithm>
#include <cstring>
#include <thread>

namespace mpm {
namespace {

constexpr uint32_t IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                            0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
constexpr int PERM[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
constexpr uint32_t CHUNK_START = 1, CHUNK_END = 2, PARENT = 4, ROOT = 8;
constexpr size_t BLOCK = 64, CHUNK = 1024;

inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

inline void g(uint32_t* v, int a, int b, int c, int d, uint32_t x, uint32_t y) {
    v[a] = v[a] + v[b] + x;
    v[d] = rotr(v[d] ^ v[a], 16);
    v[c] = v[c] + v[d];
    v[b] = rotr(v[b] ^ v[c], 12);
    v[a] = v[a] + v[b] + y;
    v[d] = rotr(v[d] ^ v[a], 8);
    v[c] = v[c] + v[d];
    v[b] = rotr(v[b] ^ v[c], 7);
}

// Compression: 16-word state (cv, IV[0..4), counter, block length, flags),
// seven rounds of column then diagonal G steps, message permuted between
// rounds; the new chaining value is the xor of the state's two halves.
void compress(const uint32_t cv[8], const uint32_t block[16], uint64_t counter, uint32_t block_len,
              uint32_t flags, uint32_t out[8]) {
    uint32_t v[16] = {cv[0], cv[1], cv[2], cv[3], cv[4], cv[5], cv[6], cv[7],
                      IV[0], IV[1], IV[2], IV[3], uint32_t(counter), uint32_t(counter >> 32), block_len, flags};
    uint32_t m[16], t[16];
    std::memcpy(m, block, sizeof(m));
    for (int r = 0; r < 7; ++r) {
        g(v, 0, 4, 8, 12, m[0], m[1]);
        g(v, 1, 5, 9, 13, m[2], m[3]);
        g(v, 2, 6, 10, 14, m[4], m[5]);
        g(v, 3, 7, 11, 15, m[6], m[7]);
        g(v, 0, 5, 10, 15, m[8], m[9]);
        g(v, 1, 6, 11, 12, m[10], m[11]);
        g(v, 2, 7, 8, 13, m[12], m[13]);
        g(v, 3, 4, 9, 14, m[14], m[15]);
        if (r < 6) {
            for (int i = 0; i < 16; ++i) t[i] = m[PERM[i]];
            std::memcpy(m, t, sizeof(m));
        }
    }
    for (int i = 0; i < 8; ++i) out[i] = v[i] ^ v[i + 8];
}

// The input is the virtual concatenation head || body.
struct Input {
    const uint8_t* head;
    size_t n_head;
    const uint8_t* body;
    size_t n_body;
    size_t size() const { return n_head + n_body; }
    // Little-endian words of the block at byte offset `off` holding `len`
    // bytes, zero-padded to 64.
    void block_words(size_t off, size_t len, uint32_t w[16]) const {
        uint8_t b[BLOCK] = {0};
        size_t k = 0;
        if (off < n_head) {
            const size_t h = std::min(len, n_head - off);
            std::memcpy(b, head + off, h);
            k = h;
        }
        if (k < len) std::memcpy(b + k, body + (off + k - n_head), len - k);
        for (int i = 0; i < 16; ++i)
            w[i] = uint32_t(b[4 * i]) | uint32_t(b[4 * i + 1]) << 8 | uint32_t(b[4 * i + 2]) << 16 |
                   uint32_t(b[4 * i + 3]) << 24;
    }
};

// A node whose final compression is still open, so the root can add ROOT.
struct Node {
    uint32_t cv[8];
    uint32_t block[16];
    uint64_t counter;
    uint32_t len;
    uint32_t flags;
    void chaining(uint32_t out[8]) const { compress(cv, block, counter, len, flags, out); }
};

// Chunk `index` covering bytes [off, off + len), len in [0, 1024].
Node chunk_node(const Input& in, size_t off, size_t len, uint64_t index) {
    Node n;
    std::memcpy(n.cv, IV, sizeof(IV));
    const size_t nblocks = len == 0 ? 1 : (len + BLOCK - 1) / BLOCK;
    for (size_t b = 0; b + 1 < nblocks; ++b) {
        uint32_t w[16];
        in.block_words(off + b * BLOCK, BLOCK, w);
        compress(n.cv, w, index, BLOCK, b == 0 ? CHUNK_START : 0, n.cv);
    }
    const size_t last = len - (nblocks - 1) * BLOCK;
    in.block_words(off + (nblocks - 1) * BLOCK, last, n.block);
    n.counter = index;
    n.len = uint32_t(last);
    n.flags = CHUNK_END | (nblocks == 1 ? CHUNK_START : 0);
    return n;
}

Node parent_node(const uint32_t l[8], const uint32_t r[8]) {
    Node n;
    std::memcpy(n.cv, IV, sizeof(IV));
    std::memcpy(n.block, l, 32);
    std::memcpy(n.block + 8, r, 32);
    n.counter = 0;
    n.len = BLOCK;
    n.flags = PARENT;
    return n;
}

// Open node of the subtree over bytes [off, off + len) whose first chunk has
// index off / 1024.  `spawn` levels of recursion hash the left half on a new
// thread.
Node subtree(const Input& in, size_t off, size_t len, int spawn) {
    if (len <= CHUNK) return chunk_node(in, off, len, off / CHUNK);
    size_t left = CHUNK;  // largest power-of-two chunk count leaving >= 1 byte on the right
    while (2 * left < len) left *= 2;
    uint32_t lcv[8], rcv[8];
    if (spawn > 0 && len >= (size_t(1) << 20)) {
        std::thread t([&] { subtree(in, off, left, spawn - 1).chaining(lcv); });
        subtree(in, off + left, len - left, spawn - 1).chaining(rcv);
        t.join();
    } else {
        subtree(in, off, left, 0).chaining(lcv);
        subtree(in, off + left, len - left, 0).chaining(rcv);
    }
    return parent_node(lcv, rcv);
}

}  // namespace

void blake3_hash2(const void* head, size_t n_head, const void* body, size_t n_body, uint8_t out[BLAKE3_OUT_LEN]) {
    const Input in{static_cast<const uint8_t*>(head), n_head, static_cast<const uint8_t*>(body), n_body};
    const Node root = subtree(in, 0, in.size(), 4);
    uint32_t w[8];
    compress(root.cv, root.block, root.counter, root.len, root.flags | ROOT, w);
    for (int i = 0; i < 8; ++i)
        for (int k = 0; k < 4; ++k) out[4 * i + k] = uint8_t(w[i] >> (8 * k));
}

}  // namespace mpm
