"""The generated evaluators and the per-item kernel bodies compiled for the host
(libmct_host.so, csrc/host_emu.cpp): a CPU twin of the HIP pipeline, used as a parallel CPU
checker of large GPU runs and to bisect a mismatch between the kernels and the oracle without
a GPU (the reference's "gen-host" implementation, tests/cuda_impls.cpp:9-15). It is never a
fallback of the library: libmct.so has no CPU path.

  HostGenerator().generate(seed, chunks, threads=8) -> (terrain u8 [n, 384, 16, 16], biomes u8 [n, 96, 4, 4])
  op_counts(seed, chunks) -> algorithmic operations per stage and per stage unit (libmct_count.so)

Operation counts: the code generator gives every scope of the generated evaluators and every
noise octave a static count of its source operations by class (FP64, FP32, INT32, memory
reads; csrc/gen/opcounts.h), and the hand-written stage bodies carry the same hooks. The
counting build adds a scope's counts each time the scope runs, so a stage's total is the sum
over scopes of static count x invocations, with the invocations of the actual chunks.
"""
import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libmct_host.so')
COUNT_PATH = os.path.join(HERE, 'libmct_count.so')
_lib = None
_count_lib = None
OP_CLASSES = ('f64', 'f32', 'i32', 'mem')


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError('libmct_host.so is not built (run build.py): %s' % LIB_PATH)
        _lib = _bind(ctypes.CDLL(LIB_PATH))
    return _lib


def _bind(L):
    P, I, I64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
    L.mcte_create.restype = P
    L.mcte_destroy.argtypes = [P]
    L.mcte_last_error.restype = ctypes.c_char_p
    L.mcte_generate.argtypes = [P, I64, P, I, P, P, P, P, P]
    L.mcte_biome_query.argtypes = [P, I64, P, I, P]
    L.mcte_stage_ops.argtypes = [I, ctypes.POINTER(ctypes.c_char_p), P]
    L.mcte_check_divisions.argtypes = [I, P]
    L.mcte_check_minmax.argtypes = [I, P]
    L.mcte_check_dense.argtypes = [I, I, I, I, P]
    L.mcte_set_replay_all.argtypes = [P, I]
    return L


def count_lib():
    global _count_lib
    if _count_lib is None:
        if not os.path.exists(COUNT_PATH):
            raise RuntimeError('libmct_count.so is not built (run build.py): %s' % COUNT_PATH)
        L = _bind(ctypes.CDLL(COUNT_PATH))
        if L.mcte_counts_ops() != 1:
            raise RuntimeError('%s was built without operation counting' % COUNT_PATH)
        _count_lib = L
    return _count_lib


# What each stage runs over: its operations per launch are ops per unit x units per launch.
STAGE_UNIT = {
    'biome_columns': 'biome_chunks', 'biome_quarts': 'biome_chunks', 'biome_ties': 'biome_chunks', 'biome_chain': 'biome_chunks',
    'dedupe_columns': 'chunks', 'columns': 'columns', 'corners': 'columns', 'dedupe_grid': 'chunks', 'fts': 'grid_keys',
    'grid_scatter+aquifer_skip': 'chunks', 'aquifer_cells': 'chunks', 'fill': 'chunks', 'heights': 'chunks',
    'surface_columns': 'chunks', 'surface1': 'chunks', 'surface2+ore_veins': 'chunks', 'surface_check+exact': 'chunks',
}


def stage_units(seeds, chunks):
    """Units of one batch: chunks, distinct biome chunks (3 x 3 neighbourhoods), distinct
    corner columns (5 x 5 quart columns per chunk) and distinct surface-level grid keys
    (27 x 19 quarts around each chunk, the 11 x 11 prefill window on the volume path, and the
    4 chunk corners on the volume path), each per seed, as the batch's dedupe finds them.
    Quart coordinates stay within +-2^20 for chunks within the world border."""
    xz = np.asarray(chunks, np.int64).reshape(-1, 2)
    s = np.broadcast_to(np.asarray(seeds, np.int64), (len(xz),)).astype(np.int64)
    cx, cz = xz[:, 0], xz[:, 1]

    _, sidx = np.unique(s, return_inverse=True)

    def distinct(a, b, flag=None, lead=1):
        # one int64 key per position: seed index, x and z offset to 21 bits each, the flag
        si = np.repeat(sidx.astype(np.int64), a.size // len(xz))
        key = (si << 43) | ((a.ravel() + (1 << 20)) << 22) | ((b.ravel() + (1 << 20)) << 1)
        if flag is not None:
            key |= flag.ravel()
        return len(np.unique(key))
    d = np.arange(-1, 2)
    bx = (cx[:, None, None] + d[None, None, :]).repeat(3, 1)
    bz = (cz[:, None, None] + d[None, :, None]).repeat(3, 2)
    k = np.arange(25)
    qx = cx[:, None] * 4 + k % 5
    qz = cz[:, None] * 4 + k // 5
    e = np.arange(27 * 19)
    gx = 4 * cx[:, None] - 16 + e % 27
    gz = 4 * cz[:, None] - 8 + e // 27
    vol = (gx >= 4 * cx[:, None] - 4) & (gx <= 4 * cx[:, None] + 6) & (gz >= 4 * cz[:, None] - 4) & (gz <= 4 * cz[:, None] + 6)
    j = np.arange(4)
    gx = np.concatenate([gx, 4 * cx[:, None] + 4 * (j & 1)], 1)
    gz = np.concatenate([gz, 4 * cz[:, None] + 4 * (j >> 1)], 1)
    vol = np.concatenate([vol, np.ones((len(xz), 4), bool)], 1)
    return {'chunks': len(xz),
            'biome_chunks': distinct(bx, bz),
            'columns': distinct(qx, qz),
            'grid_keys': distinct(gx, gz, vol.astype(np.int64))}


def op_counts(seed, chunks):
    """Algorithmic operations of the TERRAIN step on these chunks under one seed, per stage:
    {stage: {'unit', 'units', 'ops': {class: total}, 'per_unit': {class: ops / unit}}}, from
    one run of the counting emulation (single thread; about 0.1 s per chunk)."""
    L = count_lib()
    xz = np.ascontiguousarray(np.asarray(chunks, np.int32).reshape(-1, 2))
    n = len(xz)
    h = L.mcte_create()
    try:
        t = np.zeros((n, 384, 16, 16), np.uint8)
        b = np.zeros((n, 96, 4, 4), np.uint8)
        st = np.zeros(4, np.int64)
        if L.mcte_generate(h, int(seed), xz.ctypes.data, n, t.ctypes.data, b.ctypes.data, None, None, st.ctypes.data) != 0:
            raise RuntimeError('host emulation: ' + L.mcte_last_error().decode())
    finally:
        L.mcte_destroy(h)
    units = stage_units(np.full(n, seed, np.int64), xz)
    out = {}
    name = ctypes.c_char_p()
    ops = np.zeros(4, np.uint64)
    for i in range(L.mcte_stage_ops(0, None, None)):
        L.mcte_stage_ops(i, ctypes.byref(name), ops.ctypes.data)
        stage = name.value.decode()
        unit = STAGE_UNIT[stage]
        tot = {c: int(v) for c, v in zip(OP_CLASSES, ops)}
        out[stage] = {'unit': unit, 'units': units[unit], 'ops': tot, 'per_unit': {c: v / units[unit] for c, v in tot.items()}}
    return out


DIVISORS = (25.0, 1.5, 2.5, 3.0, 10.0)


def check_divisions(threads=8):
    """The device's rewrites of divisions and products, checked bit for bit on the CPU with the
    same source (libmct_host.so, mcte_check_divisions): mismatch counts of
      'squeeze'   t / 24 (squeeze_div24) over every float t in [-1, 1];
      'aquifer'   x / d (div_known) for d in DIVISORS over every half-integer |x| <= 2^19;
      'raw'       the uncorrected x * RN(1/d) over the same x (must mismatch: sensitivity);
      'gradient'  fma(g, v, c) against g * v + c for g in {-1, 0, 1} (noise dots)."""
    out = np.zeros(12, np.int64)
    if lib().mcte_check_divisions(int(threads), out.ctypes.data) != 0:
        raise RuntimeError('mcte_check_divisions failed')
    return {'squeeze': int(out[0]), 'aquifer': dict(zip(DIVISORS, out[1:6].tolist())),
            'raw': dict(zip(DIVISORS, out[6:11].tolist())), 'gradient': int(out[11])}


def check_minmax(threads=8):
    """Java Math.min / max (float) and Mth.clamp as the device computes them (select forms in
    jmath.h), against the JDK's own branching definitions on the CPU with the same source
    (libmct_host.so, mcte_check_minmax): mismatch counts over 'special' value pairs / triples,
    'random' bit-pattern pairs and 'clamp' of every float for four bound pairs."""
    out = np.zeros(3, np.int64)
    if lib().mcte_check_minmax(int(threads), out.ctypes.data) != 0:
        raise RuntimeError('mcte_check_minmax failed')
    return {'special': int(out[0]), 'random': int(out[1]), 'clamp': int(out[2])}


def check_dense(batches=40, n=64, nseeds=3, spread=40):
    """The dense dedupe's key index on random batches (mcte_check_dense): counts of keys indexed
    outside the dense space, not given back by dense_key, and out of order; batches checked."""
    out = np.zeros(4, np.int64)
    if lib().mcte_check_dense(int(batches), int(n), int(nseeds), int(spread), out.ctypes.data) != 0:
        raise RuntimeError('mcte_check_dense failed')
    return {'outside': int(out[0]), 'roundtrip': int(out[1]), 'order': int(out[2]), 'batches': int(out[3])}


class HostGenerator:
    def __init__(self, replay_all=False):
        """replay_all: the exact surface pass walks every column of a chunk it redoes again (the
        plain sequential order) instead of only the columns whose steep test changed."""
        self.replay_all = replay_all
        self.exact_redone = self.exact_columns = self.exact_changed = 0

    def _run(self, seed, xz, t, b):
        L = lib()
        h = L.mcte_create()
        try:
            L.mcte_set_replay_all(h, 1 if self.replay_all else 0)
            st = np.zeros(4, np.int64)
            rc = L.mcte_generate(h, int(seed), xz.ctypes.data, len(xz), t.ctypes.data, b.ctypes.data, None, None, st.ctypes.data)
            if rc != 0:
                raise RuntimeError('host emulation: ' + L.mcte_last_error().decode())
            return st[:3].copy()
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
        redone = []

        def work(p):
            try:
                lo, hi = int(p[0]), int(p[-1]) + 1
                tt, bb = t[lo:hi], b[lo:hi]
                redone.append(self._run(seed, np.ascontiguousarray(xz[lo:hi]), tt, bb))
            except Exception as e:   # reported after the join
                errors.append(e)
        ths = [threading.Thread(target=work, args=(p,)) for p in parts]
        for th in ths:
            th.start()
        for th in ths:
            th.join()
        if errors:
            raise errors[0]
        # chunks the exact surface pass redid, their columns walked again, the chunks it changed
        self.exact_redone, self.exact_columns, self.exact_changed = (int(v) for v in np.sum(redone, axis=0)) if redone else (0, 0, 0)
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
