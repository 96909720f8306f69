// BLAKE3 (unkeyed hash mode, 32-byte output), host code.
//
// The soil checkpoint digests its payload with the `blake3` crate 1.8.5
// (Cargo.lock:1467-1470, mpm_checkpoint.rs:87-94).  The crate is not part of
// this build, so this is a restatement of the published algorithm
// (O'Connor, Aumasson, Neves, Wilcox-O'Hearn, "BLAKE3: one function, fast
// everywhere", 2020): 64-byte blocks, 1 KiB chunks, a binary Merkle tree of
// chunk chaining values whose left subtree always holds the largest power of
// two of chunks that leaves at least one byte to the right, and a 7-round
// compression function built on the BLAKE2s G function.
#pragma once

#include <cstddef>
#include <cstdint>

namespace mpm {

constexpr size_t BLAKE3_OUT_LEN = 32;

// Digest of the concatenation head[0, n_head) || body[0, n_body).  The lattice
// checkpoint digests its version word and then the payload; the plain one
// passes n_head = 0.  Inputs of several MiB are hashed by up to 16 host
// threads, one per subtree (the tree shape, and so the digest, is fixed by
// the input length alone).
void blake3_hash2(const void* head, size_t n_head, const void* body, size_t n_body, uint8_t out[BLAKE3_OUT_LEN]);

inline void blake3_hash(const void* data, size_t len, uint8_t out[BLAKE3_OUT_LEN]) {
    blake3_hash2(nullptr, 0, data, len, out);
}

}  // namespace mpm
