"""Randomness of key generation and encryption (fhea_chacha20_block, FHEA_SEED_RANDOM).

The secure source is ChaCha20 (RFC 8439) keyed from the OS; its block function is
pinned by the RFC's known-answer vector (section 2.3.2) and by the all-zero-key
vector of appendix A.1.  Seeded calls stay reproducible; unseeded calls differ
from call to call and still give working keys (decryption invariant)."""
import ctypes as C

import numpy as np
import pytest

from anns_fhe_amd import _lib


def _block(key, counter, nonce):
    lib = _lib.load()
    k = (C.c_uint32 * 8)(*key)
    n = (C.c_uint32 * 3)(*nonce)
    out = (C.c_uint32 * 16)()
    assert lib.fhea_chacha20_block(k, counter, n, out) == 0
    return list(out)


def _words_le(hexstr):
    b = bytes.fromhex(hexstr)
    return [int.from_bytes(b[4 * i:4 * i + 4], "little") for i in range(len(b) // 4)]


def test_chacha20_block_rfc8439_vector():
    key = _words_le("000102030405060708090a0b0c0d0e0f101112131415161718191a1b1c1d1e1f")
    nonce = _words_le("000000090000004a00000000")
    got = _block(key, 1, nonce)
    assert got == [0xe4e7f110, 0x15593bd1, 0x1fdd0f50, 0xc47120a3, 0xc7f4d1c7, 0x0368c033, 0x9aaa2204, 0x4e6cd4c3,
                   0x466482d2, 0x09aa9f07, 0x05d7c214, 0xa2028bd9, 0xd19c12b5, 0xb94e16de, 0xe883d0cb, 0x4e3c50a2]


def test_chacha20_block_zero_key_vector():
    got = _block([0] * 8, 0, [0, 0, 0])
    want = _words_le("76b8e0ada0f13d90405d6ae55386bd28bdd219b8a08ded1aa836efcc8b770dc7"
                     "da41597c5157488d7724e03fb8d84a376a43b8f41518a11cc387b669b2ee6586")
    assert got == want


def test_seeded_reproducible_unseeded_fresh():
    from anns_fhe_amd.client import CkksClient
    cl = CkksClient(1024, 3, 1, 3)
    assert np.array_equal(cl.keygen_secret(5), cl.keygen_secret(5))
    a, b = cl.keygen_secret(), cl.keygen_secret()
    assert not np.array_equal(a, b)
    for s in (a, b):
        assert set(np.unique(s)) <= {-1, 0, 1}
        # a ternary draw: each value near a third of the coefficients
        assert all(abs(np.mean(s == v) - 1 / 3) < 0.08 for v in (-1, 0, 1))
    pk1, pk2 = cl.keygen_public(a), cl.keygen_public(a)
    assert not np.array_equal(pk1, pk2)
    z = np.linspace(-1, 1, 512)
    for pk in (pk1, pk2):
        ct = cl.encrypt_values(pk, z, 3, 2.0 ** 40)
        assert np.max(np.abs(cl.decrypt_values(a, ct, 3, 2.0 ** 40, 512) - z)) < 1e-6
    with pytest.raises(ValueError):
        cl.keygen_secret(-1)
