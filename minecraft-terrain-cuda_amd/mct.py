"""Python binding of libmct.so (the C ABI in include/mct.h), shaped like the reference's
generator class FullTerrainGenerator (cuda/include/mccuda_cuda/terrain_gpu.h:37-55).

Arrays are numpy; the library is loaded from this directory and must have been built
(`python build.py`). Nothing here computes terrain: every call goes to the HIP library,
and a missing or unloadable library is an error.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libmct.so')
IR_PATH = os.path.join(HERE, 'data', 'synthetic_overworld_ir.json')
TERRAIN_PATH = os.path.join(HERE, 'data', 'synthetic_overworld_terrain.json')

BLOCKS_PER_CHUNK = 16 * 384 * 16
QUARTS_PER_CHUNK = 4 * 96 * 4

_lib = None


class MctError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__('mct error %d: %s' % (code, msg))
        self.code = code


TUNE_REPLAY_ALL = 1   # mct_tuning.flags: the exact surface pass walks every column of a redone chunk


class Tuning(ctypes.Structure):
    _fields_ = [('workgroup', ctypes.c_int32 * 24), ('run_len', ctypes.c_int32), ('flags', ctypes.c_int32), ('ms', ctypes.c_double * 24)]

    def as_dict(self):
        n = kernel_group_count()
        names = [lib().mct_kernel_name(i).decode() for i in range(n)]
        d = {'workgroup': {names[i]: int(self.workgroup[i]) for i in range(n) if self.workgroup[i]}, 'run_len': int(self.run_len),
             'ms': {names[i]: round(float(self.ms[i]), 4) for i in range(n) if self.ms[i] > 0}}
        if self.flags & TUNE_REPLAY_ALL:
            d['surface_replay_all'] = True
        return d

    @classmethod
    def from_dict(cls, d):
        t = cls()
        n = kernel_group_count()
        names = [lib().mct_kernel_name(i).decode() for i in range(n)]
        for k, v in d.get('workgroup', {}).items():
            t.workgroup[names.index(k)] = int(v)
        t.run_len = int(d.get('run_len', 8))
        t.flags = TUNE_REPLAY_ALL if d.get('surface_replay_all') else 0
        return t


def kernel_group_count():
    n = 0
    while lib().mct_kernel_name(n):
        n += 1
    return n


class Timing(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ('h2d_ms', 'biome_ms', 'noise_ms', 'aquifer_ms', 'fill_ms', 'surface_ms', 'd2h_ms', 'wall_ms')]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MctError(-1, 'libmct.so is not built (run build.py): %s' % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        P, I, I64, SZ = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_size_t
        L.mct_last_error.restype = ctypes.c_char_p
        L.mct_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p, I, I, I, ctypes.POINTER(P)]
        L.mct_destroy.argtypes = [P]
        L.mct_set_seed.argtypes = [P, I64]
        L.mct_generate.argtypes = [P, P, I, P, P, P, P]
        L.mct_generate_multi.argtypes = [P, P, P, I, P, P, P, P]
        L.mct_generate_device.argtypes = [P, P, P, I, P, P, P]
        L.mct_biome_tie_stats.argtypes = [P, P]
        L.mct_generate_core.argtypes = [P, P, P, I, P, P, P]
        L.mct_autotune.argtypes = [P, P, P, I, P]
        L.mct_get_tuning.argtypes = [P, P]
        L.mct_set_tuning.argtypes = [P, P]
        L.mct_generate_core_device.argtypes = [P, P, P, I, P, P, P]
        L.mct_last_batch_done_ms.argtypes = [P, P, I, ctypes.POINTER(I)]
        L.mct_alloc_pinned.argtypes = [SZ]
        L.mct_alloc_pinned.restype = P
        L.mct_free_pinned.argtypes = [P]
        L.mct_alloc_device.argtypes = [SZ]
        L.mct_alloc_device.restype = P
        L.mct_free_device.argtypes = [P]
        L.mct_device_to_host.argtypes = [P, P, SZ]
        L.mct_chunk_digests.argtypes = [P, I64, I64, P]
        L.mct_chunk_digests_noalloc.argtypes = [P, I64, I64, P]
        L.mct_topdown.argtypes = [P, I, I, P]
        L.mct_info.argtypes = [P]
        L.mct_profile.argtypes = [P, I]
        L.mct_kernel_times.argtypes = [P, P, P, I, ctypes.POINTER(I)]
        L.mct_set_launch_mode.argtypes = [P, I]
        L.mct_kernel_name.argtypes = [I]
        L.mct_kernel_name.restype = ctypes.c_char_p
        L.mct_biome_query.argtypes = [P, P, P, I, P]
        L.mct_zoom_selftest.argtypes = [I64, I64, P]
        L.mct_server_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p, I, I, SZ, ctypes.POINTER(P)]
        L.mct_server_submit.argtypes = [P, I64, P, I, P, P, ctypes.POINTER(I64)]
        L.mct_server_poll.argtypes = [P, P, P, I, ctypes.POINTER(I)]
        L.mct_server_drain.argtypes = [P]
        L.mct_server_stats.argtypes = [P, P]
        L.mct_server_destroy.argtypes = [P]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise MctError(rc, lib().mct_last_error().decode())


def _ptr(a):
    return None if a is None else a.ctypes.data


def _host_out(a, n, per_chunk, what):
    """A caller's host output array for n chunks of per_chunk bytes: it must be a writeable,
    C-contiguous uint8 array of at least that size (the library writes through its pointer)."""
    if not isinstance(a, np.ndarray) or a.dtype != np.uint8 or not a.flags['C_CONTIGUOUS'] or not a.flags['WRITEABLE']:
        raise ValueError('%s must be a writeable C-contiguous uint8 numpy array' % what)
    if a.nbytes < n * per_chunk:
        raise ValueError('%s holds %d bytes, %d chunks need %d' % (what, a.nbytes, n, n * per_chunk))
    return a


def _xz(chunks):
    a = np.ascontiguousarray(np.asarray(chunks, dtype=np.int32).reshape(-1, 2))
    return a


class FullTerrainGenerator:
    """generate(chunks) -> (terrain u8[n, 384, 16, 16], biomes u8[n, 96, 4, 4])."""

    def __init__(self, max_batch_chunks=256, pipeline_slots=2, seed_capacity=64, ir_path=IR_PATH, terrain_path=TERRAIN_PATH):
        h = ctypes.c_void_p()
        _check(lib().mct_create(ir_path.encode(), terrain_path.encode(), max_batch_chunks, pipeline_slots, seed_capacity, ctypes.byref(h)))
        self.h = h
        self.max_batch = max_batch_chunks

    def close(self):
        if self.h:
            lib().mct_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_seed(self, seed):
        _check(lib().mct_set_seed(self.h, int(seed)))

    def generate(self, chunks, biomes=True, aquifer=False, timing=None, out=None):
        xz = _xz(chunks)
        n = len(xz)
        t = _host_out(out, n, BLOCKS_PER_CHUNK, 'out') if out is not None else np.empty((n, 384, 16, 16), np.uint8)
        b = np.empty((n, 96, 4, 4), np.uint8) if biomes else None
        a = np.empty((n, 384, 16, 16), np.uint8) if aquifer else None
        _check(lib().mct_generate(self.h, _ptr(xz), n, _ptr(t), _ptr(b), _ptr(a), ctypes.byref(timing) if timing is not None else None))
        return (t, b, a) if aquifer else (t, b)

    def generate_multi(self, seeds, chunks, biomes=True, aquifer=False, timing=None):
        xz = _xz(chunks)
        n = len(xz)
        s = np.ascontiguousarray(np.asarray(seeds, dtype=np.int64))
        if len(s) != n:
            raise ValueError('one seed per chunk')
        t = np.empty((n, 384, 16, 16), np.uint8)
        b = np.empty((n, 96, 4, 4), np.uint8) if biomes else None
        a = np.empty((n, 384, 16, 16), np.uint8) if aquifer else None
        _check(lib().mct_generate_multi(self.h, _ptr(s), _ptr(xz), n, _ptr(t), _ptr(b), _ptr(a), ctypes.byref(timing) if timing is not None else None))
        return (t, b, a) if aquifer else (t, b)

    def generate_device(self, seeds, chunks, d_terrain, d_biomes=None, timing=None):
        """Outputs stay in device memory (pointers from alloc_device)."""
        xz = _xz(chunks)
        s = np.ascontiguousarray(np.asarray(seeds, dtype=np.int64))
        if len(s) != len(xz):
            raise ValueError('one seed per chunk')
        _check(lib().mct_generate_device(self.h, _ptr(s), _ptr(xz), len(xz), d_terrain, d_biomes, ctypes.byref(timing) if timing is not None else None))

    def autotune(self, seeds, chunks):
        """Times the launch shapes on the sample and keeps the fastest; returns the tuning."""
        xz = _xz(chunks)
        s = np.ascontiguousarray(np.broadcast_to(np.asarray(seeds, dtype=np.int64), (len(xz),)))
        t = Tuning()
        _check(lib().mct_autotune(self.h, _ptr(s), _ptr(xz), len(xz), ctypes.byref(t)))
        return t.as_dict()

    def tuning(self):
        t = Tuning()
        _check(lib().mct_get_tuning(self.h, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, d):
        _check(lib().mct_set_tuning(self.h, ctypes.byref(Tuning.from_dict(d))))

    def generate_core(self, seeds, chunks, density=True, fill=True, timing=None):
        """Core configuration (TerrainGenerator): final density float32 (n, 384, 16, 16) and the
        core fill uint8 (0 air, 1 stone, 2 water, 3 lava), either may be skipped (None)."""
        xz = _xz(chunks)
        n = len(xz)
        s = np.ascontiguousarray(np.broadcast_to(np.asarray(seeds, dtype=np.int64), (n,)))
        d = np.empty((n, 384, 16, 16), np.float32) if density else None
        f = np.empty((n, 384, 16, 16), np.uint8) if fill else None
        _check(lib().mct_generate_core(self.h, _ptr(s), _ptr(xz), n, _ptr(d), _ptr(f), ctypes.byref(timing) if timing is not None else None))
        return d, f

    def generate_core_device(self, seeds, chunks, d_density=None, d_fill=None, timing=None):
        xz = _xz(chunks)
        s = np.ascontiguousarray(np.broadcast_to(np.asarray(seeds, dtype=np.int64), (len(xz),)))
        _check(lib().mct_generate_core_device(self.h, _ptr(s), _ptr(xz), len(xz), d_density, d_fill, ctypes.byref(timing) if timing is not None else None))

    def biome_query(self, seeds, xyz):
        """Biome stored by the TERRAIN step for each block (x, y, z) under its seed."""
        p = np.ascontiguousarray(np.asarray(xyz, dtype=np.int32).reshape(-1, 3))
        s = np.ascontiguousarray(np.broadcast_to(np.asarray(seeds, dtype=np.int64), (len(p),)))
        out = np.empty(len(p), np.uint8)
        _check(lib().mct_biome_query(self.h, _ptr(s), _ptr(p), len(p), _ptr(out)))
        return out

    def profile(self, enable=True):
        _check(lib().mct_profile(self.h, 1 if enable else 0))

    def launch_mode(self, graphs=True):
        """graphs=True: each batch replays a captured HIP graph; False: launch by launch."""
        _check(lib().mct_set_launch_mode(self.h, 1 if graphs else 0))

    def kernel_times(self):
        ms = np.zeros(64, np.float64)
        cnt = np.zeros(64, np.int64)
        n = ctypes.c_int()
        _check(lib().mct_kernel_times(self.h, _ptr(ms), _ptr(cnt), 64, ctypes.byref(n)))
        return {lib().mct_kernel_name(i).decode(): (float(ms[i]), int(cnt[i])) for i in range(n.value)}

    def biome_tie_stats(self):
        out = np.zeros(4, np.int64)
        _check(lib().mct_biome_tie_stats(self.h, _ptr(out)))
        return {'quarts': int(out[0]), 'overrides': int(out[1]), 'biome_changes': int(out[2]), 'exact_surface_redo': int(out[3])}

    def last_batch_done_ms(self):
        n = ctypes.c_int()
        buf = np.zeros(1 << 16, np.float64)
        _check(lib().mct_last_batch_done_ms(self.h, _ptr(buf), len(buf), ctypes.byref(n)))
        return buf[:min(n.value, len(buf))].copy()


def zoom_selftest(world_seed, count):
    """(checked, mismatches, near_ties) of the column-cached zoom against the plain FP64 zoom;
    near ties are picks with a second candidate within 8e-6 of the minimal distance (the
    reference's FP64 fallbacks, terrain_gpu.h:56-59)."""
    out = np.zeros(3, np.int64)
    _check(lib().mct_zoom_selftest(int(world_seed), int(count), _ptr(out)))
    return tuple(int(v) for v in out)


def alloc_device(nbytes):
    p = lib().mct_alloc_device(nbytes)
    if not p:
        raise MctError(3, lib().mct_last_error().decode())
    return p


def free_device(p):
    lib().mct_free_device(p)


def device_to_host(p, nbytes):
    out = np.empty(nbytes, np.uint8)
    _check(lib().mct_device_to_host(_ptr(out), p, nbytes))
    return out


def chunk_digests(p, count, bytes_per_chunk):
    """Per-chunk 64-bit digests (uint64 array) of count chunks of device memory at p."""
    out = np.zeros(int(count), np.uint64)
    _check(lib().mct_chunk_digests(p, int(count), int(bytes_per_chunk), _ptr(out)))
    return out


def chunk_digests_noalloc(p, count, bytes_per_chunk):
    """chunk_digests through the library's fixed device buffer (no device allocation; at most
    DIGESTS_MAX chunks per call)."""
    out = np.zeros(int(count), np.uint64)
    _check(lib().mct_chunk_digests_noalloc(p, int(count), int(bytes_per_chunk), _ptr(out)))
    return out


DIGESTS_MAX = 65536


def digest_host(chunks):
    """The same digest computed on the host (numpy), for arrays of whole chunks (n, ...) of any
    element type (the digest is of each chunk's bytes; a multiple of 8 per chunk)."""
    a = np.ascontiguousarray(chunks).reshape(len(chunks), -1).view(np.uint8)
    w = a.view('<u8')
    with np.errstate(over='ignore'):
        def mix(z):
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            return z ^ (z >> np.uint64(31))
        i = np.arange(1, w.shape[1] + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
        h = mix(w + i[None, :]).sum(axis=1, dtype=np.uint64)
        return mix(h ^ np.uint64(a.shape[1]))


def topdown(terrain, count=None):
    """Top-down records (uint8 [n, 256, 8], columns in z * 16 + x order) of n chunks of terrain:
    a host array [n, 384, 16, 16], or a device pointer with count."""
    if isinstance(terrain, np.ndarray):
        t = np.ascontiguousarray(terrain, dtype=np.uint8).reshape(-1, BLOCKS_PER_CHUNK)
        n, ptr, dev = len(t), _ptr(t), 0
    else:
        n, ptr, dev = int(count), terrain, 1
    out = np.zeros((n, 256, 8), np.uint8)
    _check(lib().mct_topdown(ptr, n, dev, _ptr(out)))
    return out


def info():
    out = np.zeros(6, np.int32)
    _check(lib().mct_info(_ptr(out)))
    return dict(zip(('blocks_per_chunk', 'quarts_per_chunk', 'min_y', 'height', 'palette', 'biomes'), (int(v) for v in out)))


class Server:
    """Serving loop (include/mct.h, mct_server_*): submit region requests from any thread,
    poll completions with their latency. Output buffers must stay alive until the request
    completes; pinned buffers (PinnedBuffer) make the copy back asynchronous."""

    def __init__(self, max_batch_chunks=1024, seed_capacity=64, hbm_budget_bytes=32 << 30, ir_path=IR_PATH, terrain_path=TERRAIN_PATH):
        h = ctypes.c_void_p()
        _check(lib().mct_server_create(ir_path.encode(), terrain_path.encode(), max_batch_chunks, seed_capacity, hbm_budget_bytes, ctypes.byref(h)))
        self.h = h
        self._keep = {}

    def submit(self, seed, chunks, terrain, biomes=None):
        xz = _xz(chunks)
        _host_out(terrain, len(xz), BLOCKS_PER_CHUNK, 'terrain')
        if biomes is not None:
            _host_out(biomes, len(xz), QUARTS_PER_CHUNK, 'biomes')
        rid = ctypes.c_int64()
        _check(lib().mct_server_submit(self.h, int(seed), _ptr(xz), len(xz), _ptr(terrain), _ptr(biomes), ctypes.byref(rid)))
        self._keep[rid.value] = (xz, terrain, biomes)
        return rid.value

    def poll(self, cap=4096):
        ids = np.zeros(cap, np.int64)
        lat = np.zeros(cap, np.float64)
        n = ctypes.c_int()
        _check(lib().mct_server_poll(self.h, _ptr(ids), _ptr(lat), cap, ctypes.byref(n)))
        out = list(zip(ids[:n.value].tolist(), lat[:n.value].tolist()))
        for rid, _ in out:
            self._keep.pop(rid, None)
        return out

    def drain(self):
        _check(lib().mct_server_drain(self.h))

    def stats(self):
        out = np.zeros(8, np.int64)
        _check(lib().mct_server_stats(self.h, _ptr(out)))
        keys = ('submitted', 'completed', 'batches', 'chunks', 'arena_chunks', 'max_arena_used', 'mixed_batches', 'seed_capacity')
        return dict(zip(keys, (int(v) for v in out)))

    def close(self):
        if self.h:
            lib().mct_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PinnedBuffer:
    """Page-locked host memory (mct_alloc_pinned) viewed as a numpy array."""

    def __init__(self, shape, dtype=np.uint8):
        n = int(np.prod(shape)) * np.dtype(dtype).itemsize
        self.p = lib().mct_alloc_pinned(max(n, 1))
        if not self.p:
            raise MctError(3, 'pinned allocation failed')
        self.array = np.ctypeslib.as_array(ctypes.cast(self.p, ctypes.POINTER(ctypes.c_uint8)), shape=(max(n, 1),))[:n].view(dtype).reshape(shape)

    def free(self):
        if self.p:
            lib().mct_free_pinned(self.p)
            self.p = None
