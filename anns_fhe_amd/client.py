"""CKKS client over the libfhea C-ABI (host only, no GPU).

Plays the role OpenFHE plays for the reference's client
(software_golden_model/fhe_core/fhe_context_manager.cpp:81-128 key generation,
:457-512 encrypt / decrypt; integration_tools/generate_interactive_test_vectors.cpp:161-224):
secret, public, relinearisation and rotation keys, encoding of real/complex
vectors into plaintext polynomials (canonical embedding, slot j at
zeta^(5^j mod 2N)), public-key encryption, decryption and decoding.

Ciphertexts and plaintexts are numpy uint64 arrays, limb-major, evaluation
domain, exactly the words the engine's arenas hold: a ciphertext at `limbs`
limbs is c0 (limbs*N words) followed by c1.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check


def _p64(a):
    return a.ctypes.data_as(_lib.u64p)


def _seed(seed):
    """None -> FHEA_SEED_RANDOM (secure generator); an integer -> reproducible stream."""
    if seed is None:
        return _lib.SEED_RANDOM
    if not 0 <= int(seed) < _lib.SEED_RANDOM:
        raise ValueError("seed must be None or an integer in [0, 2^64 - 1)")
    return int(seed)


def _pf(a):
    return a.ctypes.data_as(_lib.f64p)


class CkksClient:
    def __init__(self, N, L, K, dnum, q_bits=50, q_primes=None, p_primes=None, ctx=None, psi=None):
        """Build from a parameter block, or from a device context (`ctx`, same tables)."""
        self.lib = _lib.load()
        h = C.c_void_p()
        if ctx is not None:
            check(self.lib.fhea_client_from_ctx(ctx.h, C.byref(h)), "fhea_client_from_ctx")
        else:
            prm = _lib.fhea_params()
            prm.N, prm.L, prm.K, prm.dnum, prm.q_bits = N, L, K, dnum, q_bits
            self._qa = np.array(q_primes, np.uint64) if q_primes is not None else None
            self._pa = np.array(p_primes, np.uint64) if p_primes is not None else None
            prm.q_primes = _p64(self._qa) if self._qa is not None else C.cast(None, _lib.u64p)
            prm.p_primes = _p64(self._pa) if self._pa is not None else C.cast(None, _lib.u64p)
            self._psi = np.array(psi, np.uint64) if psi is not None else None
            prm.psi = _p64(self._psi) if self._psi is not None else C.cast(None, _lib.u64p)
            check(self.lib.fhea_client_create(C.byref(prm), C.byref(h)), "fhea_client_create")
        self.h = h
        info = (C.c_uint32 * 5)()
        check(self.lib.fhea_client_info(self.h, info), "fhea_client_info")
        self.N, self.L, self.K, self.dnum, self.alpha = list(info)
        pr = np.zeros(self.L + self.K, np.uint64)
        check(self.lib.fhea_client_primes(self.h, _p64(pr)), "fhea_client_primes")
        self.primes = [int(x) for x in pr]
        self.q = self.primes[: self.L]
        self.p = self.primes[self.L:]

    def close(self):
        if getattr(self, "h", None):
            self.lib.fhea_client_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def table(self, which):
        """0 psi, 1 N^-1 (per prime); 2 twiddles, 3 inverse twiddles ((L+K) x N)."""
        T = self.L + self.K
        n = T if which in (0, 1) else T * self.N
        out = np.zeros(n, np.uint64)
        check(self.lib.fhea_client_table(self.h, which, _p64(out), n), "fhea_client_table")
        return out if which in (0, 1) else out.reshape(T, self.N)

    # ---- keys.  seed None (default): the secure generator (ChaCha20 keyed by the OS);
    # an integer: a reproducible std::mt19937_64 stream for tests, not secure.
    def keygen_secret(self, seed=None):
        s = np.zeros(self.N, np.int8)
        check(self.lib.fhea_client_keygen_secret(self.h, _seed(seed), s.ctypes.data_as(_lib.i8p)), "keygen_secret")
        return s

    def keygen_public(self, s, seed=None):
        pk = np.zeros(2 * self.L * self.N, np.uint64)
        check(self.lib.fhea_keygen_public(self.h, s.ctypes.data_as(_lib.i8p), _seed(seed), _p64(pk)),
              "keygen_public")
        return pk

    def key_size_words(self, key_limbs):
        return 2 * self.dnum * (key_limbs + self.K) * self.N

    def relin_key(self, s, key_limbs, seed=None):
        k = np.zeros(self.key_size_words(key_limbs), np.uint64)
        check(self.lib.fhea_client_relin_key(self.h, s.ctypes.data_as(_lib.i8p), key_limbs, _seed(seed), _p64(k)),
              "relin_key")
        return k

    def rotation_key(self, s, galois, key_limbs, seed=None):
        k = np.zeros(self.key_size_words(key_limbs), np.uint64)
        check(self.lib.fhea_client_rotation_key(self.h, s.ctypes.data_as(_lib.i8p), galois, key_limbs, _seed(seed),
                                                _p64(k)), "rotation_key")
        return k

    def galois_for_step(self, step):
        half = self.N // 2
        return pow(5, step % half, 2 * self.N)

    # ---- encoding
    def encode(self, values, limbs, scale):
        v = np.asarray(values)
        n = v.size
        re = np.ascontiguousarray(v.real, dtype=np.float64).reshape(-1)
        im = np.ascontiguousarray(v.imag, dtype=np.float64).reshape(-1) if np.iscomplexobj(v) else None
        pt = np.zeros(limbs * self.N, np.uint64)
        check(self.lib.fhea_encode(self.h, _pf(re), _pf(im) if im is not None else C.cast(None, _lib.f64p),
                                   n, limbs, scale, _p64(pt)), "fhea_encode")
        return pt

    def encode_batch(self, values, limbs, scale, threads=0):
        """values: (count, n) real array -> (count, limbs*N) plaintexts, encoded on host threads."""
        v = np.ascontiguousarray(values, dtype=np.float64)
        count, n = v.shape
        pt = np.zeros((count, limbs * self.N), np.uint64)
        check(self.lib.fhea_encode_batch(self.h, _pf(v), C.cast(None, _lib.f64p), count, n, limbs, scale,
                                         _p64(pt), threads), "fhea_encode_batch")
        return pt

    def decode(self, pt, limbs, scale, n, complex_out=False):
        pt = np.ascontiguousarray(pt, dtype=np.uint64)
        re = np.zeros(n, np.float64)
        im = np.zeros(n, np.float64) if complex_out else None
        check(self.lib.fhea_decode(self.h, _p64(pt), limbs, scale, n, _pf(re),
                                   _pf(im) if im is not None else C.cast(None, _lib.f64p)), "fhea_decode")
        return re + 1j * im if complex_out else re

    # ---- encryption
    def encrypt(self, pk, pt, limbs, seed=None):
        pt = np.ascontiguousarray(pt, dtype=np.uint64)
        ct = np.zeros(2 * limbs * self.N, np.uint64)
        check(self.lib.fhea_encrypt(self.h, _p64(pk), _p64(pt), limbs, _seed(seed), _p64(ct)), "fhea_encrypt")
        return ct

    def decrypt(self, s, ct, limbs):
        ct = np.ascontiguousarray(ct, dtype=np.uint64)
        pt = np.zeros(limbs * self.N, np.uint64)
        check(self.lib.fhea_decrypt(self.h, s.ctypes.data_as(_lib.i8p), _p64(ct), limbs, _p64(pt)), "fhea_decrypt")
        return pt

    def set_host_ntt(self, mode):
        """0: FP64 CPU transforms where q < 2^50 and the CPU has AVX2+FMA (default);
        1: 64-bit integer transforms.  The words are the same."""
        check(self.lib.fhea_client_set_host_ntt(self.h, mode), "fhea_client_set_host_ntt")

    def secret_eval(self, s, limbs=None):
        """The secret in the evaluation domain on the first `limbs` Q primes (for
        decrypt_values_batch; computed once per secret)."""
        limbs = limbs or self.L
        out = np.zeros(limbs * self.N, np.uint64)
        check(self.lib.fhea_secret_eval(self.h, s.ctypes.data_as(_lib.i8p), limbs, _p64(out)), "fhea_secret_eval")
        return out

    def decrypt_values_batch(self, s_eval, cts, limbs, scale, n, threads=1, complex_out=False):
        """Decrypts and decodes `cts` (count x 2*limbs*N words, or one flat array of
        them) with s_eval = secret_eval(s): (count, n) slot values, equal to
        decrypt_values on each ciphertext."""
        cts = np.ascontiguousarray(cts, dtype=np.uint64).reshape(-1)
        W = 2 * limbs * self.N
        if cts.size % W or s_eval.size < limbs * self.N:
            raise ValueError("ciphertext or secret size does not match limbs")
        count = cts.size // W
        re = np.zeros((count, n), np.float64)
        im = np.zeros((count, n), np.float64) if complex_out else None
        check(self.lib.fhea_decrypt_decode_batch(self.h, _p64(np.ascontiguousarray(s_eval, np.uint64)), _p64(cts),
                                                 count, limbs, scale, n, _pf(re),
                                                 _pf(im) if im is not None else C.cast(None, _lib.f64p), threads),
              "fhea_decrypt_decode_batch")
        return re + 1j * im if complex_out else re

    def encrypt_values(self, pk, values, limbs, scale, seed=None):
        return self.encrypt(pk, self.encode(values, limbs, scale), limbs, seed)

    def decrypt_values(self, s, ct, limbs, scale, n, complex_out=False):
        return self.decode(self.decrypt(s, ct, limbs), limbs, scale, n, complex_out)
