"""The generated evaluators and the per-item kernel bodies compiled for the host
(libmct_host.so, csrc/host_emu.cpp): a CPU twin of the HIP pipeline, used as a parallel CPU
checker of large GPU runs and to bisect a mismatch between the kernels and the oracle without
a GPU (the reference's "gen-host" implementation, tests/cuda_impls.cpp:9-15). It is never a
fallback of the library: libmct.so has no CPU path.

  HostGenerator().generate(seed, chunks, threads=8) -> (terrain u8 [n, 384, 16, 16], biomes u8 [n, 96, 4, 4])
"""
import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libmct_host.so')
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError('libmct_host.so is not built (run build.py): %s' % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        P, I, I64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
        L.mcte_create.restype = P
        L.mcte_destroy.argtypes = [P]
        L.mcte_last_error.restype = ctypes.c_char_p
        L.mcte_generate.argtypes = [P, I64, P, I, P, P, P, P, P]
        L.mcte_biome_query.argtypes = [P, I64, P, I, P]
        _lib = L
    return _lib


class HostGenerator:
    def _run(self, seed, xz, t, b):
        L = lib()
        h = L.mcte_create()
        try:
            st = np.zeros(4, np.int64)
            rc = L.mcte_generate(h, int(seed), xz.ctypes.data, len(xz), t.ctypes.data, b.ctypes.data, None, None, st.ctypes.data)
            if rc != 0:
                raise RuntimeError('host emulation: ' + L.mcte_last_error().decode())
        finally:
            L.mcte_destroy(h)

    def generate(self, seed, chunks, threads=8):
        """TERRAIN of the chunks under seed, chunks split over `threads` host threads (one
        emulation handle each; each chunk is generated with its own 3 x 3 biome neighbourhood)."""
        xz = np.ascontiguousarray(np.asarray(chunks, np.int32).reshape(-1, 2))
        n = len(xz)
        t = np.zeros((n, 384, 16, 16), np.uint8)
        b = np.zeros((n, 96, 4, 4), np.uint8)
        parts = [p for p in np.array_split(np.arange(n), max(1, min(threads, n))) if len(p)]
        errors = []

        def work(p):
            try:
                lo, hi = int(p[0]), int(p[-1]) + 1
                tt, bb = t[lo:hi], b[lo:hi]
                self._run(seed, np.ascontiguousarray(xz[lo:hi]), tt, bb)
            except Exception as e:   # reported after the join
                errors.append(e)
        ths = [threading.Thread(target=work, args=(p,)) for p in parts]
        for th in ths:
            th.start()
        for th in ths:
            th.join()
        if errors:
            raise errors[0]
        return t, b

    def biome_query(self, seed, xyz):
        p = np.ascontiguousarray(np.asarray(xyz, np.int32).reshape(-1, 3))
        out = np.zeros(len(p), np.int32)
        L = lib()
        h = L.mcte_create()
        try:
            if L.mcte_biome_query(h, int(seed), p.ctypes.data, len(p), out.ctypes.data) != 0:
                raise RuntimeError('host emulation: ' + L.mcte_last_error().decode())
        finally:
            L.mcte_destroy(h)
        return out
