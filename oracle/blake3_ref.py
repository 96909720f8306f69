"""BLAKE3 hash mode in pure Python -- TEST INFRASTRUCTURE ONLY.

The soil checkpoint digest is ``blake3::Hasher`` from the ``blake3`` crate
1.8.5 (Cargo.lock:1467-1470, used at mpm_checkpoint.rs:87-94), which is not
installed here.  This restates the published algorithm (BLAKE3 paper, 2020)
in its incremental form: chunks are finished as input arrives and their
chaining values are merged on a stack, one merge per trailing zero bit of
the finished-chunk count; the product library (blake3.cpp) uses the
recursive split form instead, so agreement of the two checks the tree shape.

Pinned by the published known answers in tests/golden/blake3_vectors.json
(input bytes i mod 251) and by "abc".
"""
from __future__ import annotations

import struct

IV = (0x6A09E667, 0xBB67AE85, 0x3C6EF372, 0xA54FF53A, 0x510E527F, 0x9B05688C, 0x1F83D9AB, 0x5BE0CD19)
PERM = (2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8)
CHUNK_START, CHUNK_END, PARENT, ROOT = 1, 2, 4, 8
M32 = 0xFFFFFFFF


def _rotr(x: int, n: int) -> int:
    return ((x >> n) | (x << (32 - n))) & M32


def _g(v, a, b, c, d, x, y):
    v[a] = (v[a] + v[b] + x) & M32
    v[d] = _rotr(v[d] ^ v[a], 16)
    v[c] = (v[c] + v[d]) & M32
    v[b] = _rotr(v[b] ^ v[c], 12)
    v[a] = (v[a] + v[b] + y) & M32
    v[d] = _rotr(v[d] ^ v[a], 8)
    v[c] = (v[c] + v[d]) & M32
    v[b] = _rotr(v[b] ^ v[c], 7)


def compress(cv, block_words, counter: int, block_len: int, flags: int):
    v = list(cv) + list(IV[:4]) + [counter & M32, (counter >> 32) & M32, block_len, flags]
    m = list(block_words)
    for r in range(7):
        _g(v, 0, 4, 8, 12, m[0], m[1])
        _g(v, 1, 5, 9, 13, m[2], m[3])
        _g(v, 2, 6, 10, 14, m[4], m[5])
        _g(v, 3, 7, 11, 15, m[6], m[7])
        _g(v, 0, 5, 10, 15, m[8], m[9])
        _g(v, 1, 6, 11, 12, m[10], m[11])
        _g(v, 2, 7, 8, 13, m[12], m[13])
        _g(v, 3, 4, 9, 14, m[14], m[15])
        m = [m[PERM[i]] for i in range(16)]
    return [v[i] ^ v[i + 8] for i in range(8)]


def _words(block: bytes):
    return list(struct.unpack("<16I", block.ljust(64, b"\0")))


class _Chunk:
    def __init__(self, counter: int):
        self.cv = list(IV)
        self.counter = counter
        self.buf = b""
        self.blocks = 0

    def length(self) -> int:
        return 64 * self.blocks + len(self.buf)

    def update(self, data: bytes) -> bytes:
        """Absorb up to the chunk's end; returns the unused tail."""
        while data and self.length() < 1024:
            if len(self.buf) == 64:
                flags = CHUNK_START if self.blocks == 0 else 0
                self.cv = compress(self.cv, _words(self.buf), self.counter, 64, flags)
                self.blocks += 1
                self.buf = b""
            take = min(64 - len(self.buf), len(data))
            self.buf += data[:take]
            data = data[take:]
        return data

    def output(self):
        flags = CHUNK_END | (CHUNK_START if self.blocks == 0 else 0)
        return (self.cv, _words(self.buf), self.counter, len(self.buf), flags)


def _parent(left, right):
    return (list(IV), left + right, 0, 64, PARENT)


def blake3(data: bytes) -> bytes:
    """32-byte BLAKE3 hash of ``data``."""
    stack = []
    chunk = _Chunk(0)
    data = bytes(data)
    while True:
        data = chunk.update(data)
        if not data:
            break
        # the chunk is full and more input follows: it is not the root
        cv = compress(*chunk.output())
        total = chunk.counter + 1
        while total & 1 == 0:
            cv = compress(*_parent(stack.pop(), cv))
            total >>= 1
        stack.append(cv)
        chunk = _Chunk(chunk.counter + 1)
    node = chunk.output()
    while stack:
        node = _parent(stack.pop(), compress(*node))
    cv, block, counter, length, flags = node
    return struct.pack("<8I", *compress(cv, block, counter, length, flags | ROOT))
