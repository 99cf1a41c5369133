"""FNV-1a-32 digest (native/main.rs:152-163) against textbook vectors, in the
library and in the oracle."""

import numpy as np

import oracle as orc
import rainbow_fmm as rf


def test_textbook_vectors(golden):
    for v in golden["fnv1a"]:
        data = bytes.fromhex(v["hex"])
        assert rf.fnv1a(data) == v["digest"]
        assert orc.fnv1a(data) == v["digest"]


def test_more_vectors():
    # "foo" from the FNV reference suite; 48 zero bytes is the digest of one
    # empty Field (the reference survey's value).
    assert rf.fnv1a(b"foo") == 0xA9F37ED7 == orc.fnv1a(b"foo")
    assert rf.fnv1a(bytes(48)) == 0xC655FF85 == orc.fnv1a(bytes(48))


def test_word_order_matches_harness():
    # The harness hashes u32 words' little-endian bytes in buffer order, which
    # is the buffer's byte order on a little-endian host.
    words = np.array([1, 0x01020304, 0xFFFFFFFF], dtype=np.uint32)
    h = 2166136261
    for w in words:
        for b in int(w).to_bytes(4, "little"):
            h = ((h ^ b) * 16777619) & 0xFFFFFFFF
    assert rf.fnv1a(words) == h == orc.fnv1a(words)
