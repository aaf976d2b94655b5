"""Python handle on the oracle (oracle/liboracle.so): the CPU restatement of the reference's
TERRAIN semantics. Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's
CPU-baseline leg use it as the checker; the library under minecraft-terrain-cuda_amd/ never does.

Build with `make -C oracle`. The reference cpu/ port compiled from its own sources
(`make -C oracle ref`, output oracle/_ref/libmcref.so) is wrapped by `Reference` for the
cross-checks that pin the oracle.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'liboracle.so')
REF_PATH = os.path.join(HERE, '_ref', 'libmcref.so')
DATA = os.path.join(os.path.dirname(HERE), 'minecraft-terrain-cuda_amd', 'data')
IR_PATH = os.path.join(DATA, 'synthetic_overworld_ir.json')
TERRAIN_PATH = os.path.join(DATA, 'synthetic_overworld_terrain.json')

HEIGHT = 384
PER_CHUNK = 16 * 16 * HEIGHT
QUARTS = 16 * (HEIGHT // 4)

P, I, I64, U64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_uint64
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f'{LIB_PATH} is not built: run make -C oracle')
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            'orc_last_error': (ctypes.c_char_p, []),
            'orc_create': (P, [ctypes.c_char_p, ctypes.c_char_p]),
            'orc_destroy': (None, [P]),
            'orc_generate': (I, [P, I64, P, I, I, P, P, P, P]),
            'orc_sample': (I, [P, I64, I, P, P]),
            'orc_point': (I, [P, I64, I, P, I, P]),
            'orc_root': (I, [P, ctypes.c_char_p]),
            'orc_node_count': (I, [P]),
            'orc_tree': (I, [P, P, I, I, P, P, P, P]),
            'orc_leaf_biome': (I, [P, I]),
            'orc_quantize': (I64, [ctypes.c_float]),
            'orc_zoom': (None, [I64, P, I, P]),
            'orc_zoom_seed': (I64, [I64]),
            'orc_md5': (None, [ctypes.c_char_p, P]),
            'orc_sha256': (None, [P, ctypes.c_size_t, P]),
            'orc_xoro': (None, [I64, I, P]),
            'orc_xoro_ints': (None, [I64, I, I, P]),
            'orc_xoro_doubles': (None, [I64, I, P]),
            'orc_lcg_ints': (None, [I64, I, I, P]),
            'orc_lcg_doubles': (None, [I64, I, P]),
            'orc_lcg_next': (None, [I64, I, I, P]),
            'orc_mix13': (U64, [U64]),
            'orc_block_seed': (I64, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
            'orc_positional': (U64, [I64, ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
            'orc_to_int': (ctypes.c_int32, [ctypes.c_double]),
            'orc_floor': (ctypes.c_int32, [ctypes.c_double]),
            'orc_fmin': (ctypes.c_float, [ctypes.c_float, ctypes.c_float]),
            'orc_fmax': (ctypes.c_float, [ctypes.c_float, ctypes.c_float]),
            'orc_wrap': (ctypes.c_double, [ctypes.c_double]),
            'orc_stream_sum': (ctypes.c_double, [P, I]),
            'orc_octave': (ctypes.c_float, [P] + [ctypes.c_double] * 7 + [I]),
            'orc_biome_temperature': (ctypes.c_float, [P, I64, I, I, I, I]),
            'orc_octaves': (I, [P, I64, I, P, P]),
            'orc_clay': (None, [P, I64, P]),
            'orc_noise_layers': (I, [P, I64, ctypes.c_char_p, P, P, I]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def core_fill(density, sea_level=None, min_y=None):
    """Vanilla doFill with aquifers disabled, over (n, height, 16, 16) densities (y, z, x): 1 stone
    where density > 0, else 3 lava below min(-54, sea level) for y < -54, 2 water below sea
    level, 0 air (cpu/include/mccuda/volume_eval.h:76-85). Sea level and min y default to the
    IR's."""
    if sea_level is None or min_y is None:
        import json
        ir = json.load(open(IR_PATH))
        sea_level = ir['sea_level'] if sea_level is None else sea_level
        min_y = ir['min_y'] if min_y is None else min_y
    y = (np.arange(density.shape[1]) + min_y)[None, :, None, None]
    lava_level = min(sea_level, -54)
    fluid = np.where(y < lava_level, np.where(y < -54, 3, 0), np.where(y < sea_level, 2, 0))
    return np.where(density.astype(np.float64) > 0.0, 1, np.broadcast_to(fluid, density.shape)).astype(np.uint8)


def palette_flags():
    """(air, fluid) boolean arrays over palette indices: air is minecraft:air itself, fluid any
    water or lava state (cpu/src/terrain_model.cpp:120-127)."""
    import json
    pal = json.load(open(TERRAIN_PATH))['palette']
    air = np.array([p == 'minecraft:air' for p in pal] + [False] * (256 - len(pal)))
    fluid = np.array([p.startswith('minecraft:water') or p.startswith('minecraft:lava') for p in pal] + [False] * (256 - len(pal)))
    return air, fluid


def topdown_records(terrain):
    """Top-down column records of terrain [n, height, 16, 16] (y, z, x) as [n, 256, 8] with
    columns in z * 16 + x order (bench/bench_main.cpp:132-166): from the top down past air
    (stopping at y index 0) -> top; then down past fluid (stopping at 0) -> floor; bytes top,
    floor, min(255, top_y - floor_y), 0, int16 minY + top_y, int16 minY + floor_y."""
    import json
    min_y = json.load(open(IR_PATH))['min_y']
    air, fluid = palette_flags()
    t = np.asarray(terrain).reshape(len(terrain), terrain.shape[1], 256).astype(np.int64)
    h = t.shape[1]
    ys = np.arange(h)[None, :, None]
    stop_top = ~air[t] | (ys == 0)
    top_y = h - 1 - np.argmax(stop_top[:, ::-1, :], axis=1)                 # highest y that stops the air scan
    stop_floor = (~fluid[t] | (ys == 0)) & (ys <= top_y[:, None, :])
    floor_y = h - 1 - np.argmax(stop_floor[:, ::-1, :], axis=1)
    n = len(t)
    out = np.zeros((n, 256, 8), np.uint8)
    cols = np.arange(256)[None, :]
    out[:, :, 0] = t[np.arange(n)[:, None], top_y, cols]
    out[:, :, 1] = t[np.arange(n)[:, None], floor_y, cols]
    out[:, :, 2] = np.minimum(255, top_y - floor_y)
    out[:, :, 4:6] = (top_y + min_y).astype('<i2')[..., None].view(np.uint8).reshape(n, 256, 2)
    out[:, :, 6:8] = (floor_y + min_y).astype('<i2')[..., None].view(np.uint8).reshape(n, 256, 2)
    return out


class Oracle:
    """The oracle over one IR + terrain model (the synthetic overworld by default)."""

    def __init__(self, ir_path=IR_PATH, terrain_path=TERRAIN_PATH):
        L = lib()
        self.h = L.orc_create(ir_path.encode(), terrain_path.encode())
        if not self.h:
            raise RuntimeError('oracle: ' + L.orc_last_error().decode())

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError('oracle: ' + lib().orc_last_error().decode())

    def generate(self, seed, chunks, threads=8, density=False):
        """Full TERRAIN for chunk coordinates [(cx, cz), ...] under `seed`.

        Returns (terrain, biomes, aquifer[, density]): terrain/aquifer uint16 (n, 384, 16, 16) as
        (y, z, x); biomes uint8 (n, 96, 4, 4) as (qy, qz, qx); density float32 like terrain.
        """
        n = len(chunks)
        xz = np.ascontiguousarray(np.array(chunks, dtype=np.int32).reshape(n, 2))
        t = np.zeros((n, HEIGHT, 16, 16), np.uint16)
        a = np.zeros((n, HEIGHT, 16, 16), np.uint16)
        b = np.zeros((n, HEIGHT // 4, 4, 4), np.uint8)
        d = np.zeros((n, HEIGHT, 16, 16), np.float32) if density else None
        self._check(lib().orc_generate(self.h, int(seed), xz.ctypes.data, n, int(threads), b.ctypes.data, a.ctypes.data, t.ctypes.data, _ptr(d)))
        return (t, b, a, d) if density else (t, b, a)

    def generate_core(self, seed, chunks, threads=8):
        """Core configuration: (final density f32, core fill u8), (n, 384, 16, 16) as (y, z, x).
        The density is the TERRAIN step's (the same final-density volume sample of the chunk)."""
        d = self.generate(seed, chunks, threads=threads, density=True)[3]
        return d, core_fill(d)

    def biome_query(self, seeds, xyz):
        """The biome the TERRAIN step stores for each block: BiomeManager's zoom picks the quart
        (BiomeManager.getBiome), quart y clamps to the world height (Climate / noiseBiome), and
        the answer is that quart's entry in its chunk's biome array."""
        p = np.ascontiguousarray(np.asarray(xyz, dtype=np.int32).reshape(-1, 3))
        s = np.broadcast_to(np.asarray(seeds, dtype=np.int64), (len(p),))
        out = np.zeros(len(p), np.uint8)
        cache = {}
        for i in range(len(p)):
            seed = int(s[i])
            q = np.zeros((1, 3), np.int32)
            lib().orc_zoom(lib().orc_zoom_seed(seed), p[i:i + 1].ctypes.data, 1, q.ctypes.data)
            qx, qy, qz = (int(v) for v in q[0])
            qy = min(max(qy, -16), -16 + HEIGHT // 4 - 1) + 16
            key = (seed, qx >> 2, qz >> 2)
            if key not in cache:
                cache[key] = self.generate(seed, [key[1:]], threads=1)[1][0]
            out[i] = cache[key][qy, qz & 3, qx & 3]
        return out

    def root(self, name):
        r = lib().orc_root(self.h, name.encode())
        if r < 0:
            raise KeyError(name)
        return r

    def node_count(self):
        return lib().orc_node_count(self.h)

    def sample(self, seed, node, vol):
        """Volume path of IR node `node` over vol = (sx, sy, sz, x0, y0, z0, dx, dy, dz)."""
        v = np.array(vol, dtype=np.int32)
        out = np.zeros(int(v[0]) * int(v[1]) * int(v[2]), np.float32)
        self._check(lib().orc_sample(self.h, int(seed), int(node), v.ctypes.data, out.ctypes.data))
        return out

    def point(self, seed, node, xyz):
        p = np.ascontiguousarray(np.array(xyz, dtype=np.int32).reshape(-1, 3))
        out = np.zeros(len(p), np.float32)
        self._check(lib().orc_point(self.h, int(seed), int(node), p.ctypes.data, len(p), out.ctypes.data))
        return out

    def tree(self, targets, chain=True):
        t = np.ascontiguousarray(np.asarray(targets, dtype=np.int64).reshape(-1, 7))
        n = len(t)
        leaf, brute = np.zeros(n, np.int32), np.zeros(n, np.int32)
        dist, cdist = np.zeros(n, np.int64), np.zeros(n, np.int64)
        self._check(lib().orc_tree(self.h, t.ctypes.data, n, 1 if chain else 0, leaf.ctypes.data, brute.ctypes.data, dist.ctypes.data, cdist.ctypes.data))
        return leaf, brute, dist, cdist

    def leaf_biome(self, leaf):
        return lib().orc_leaf_biome(self.h, int(leaf))

    def biome_temperature(self, seed, biome, x, y, z):
        return lib().orc_biome_temperature(self.h, int(seed), int(biome), int(x), int(y), int(z))

    def octave(self, seed, k):
        offs = np.zeros(4, np.float64)
        perm = np.zeros(256, np.uint8)
        count = lib().orc_octaves(self.h, int(seed), int(k), offs.ctypes.data, perm.ctypes.data)
        return count, offs, perm

    def noise_layers(self, seed, key):
        """(params (n, 6) float64: frequency, amplitude, ox, oy, oz, smear; perms (n, 256) u8)."""
        params, perms = np.zeros((256, 6), np.float64), np.zeros((256, 256), np.uint8)
        n = lib().orc_noise_layers(self.h, int(seed), key.encode(), params.ctypes.data, perms.ctypes.data, 256)
        if n < 0:
            raise KeyError(key)
        return params[:n].copy(), perms[:n].copy()

    def clay(self, seed):
        out = np.zeros(192, np.int32)
        lib().orc_clay(self.h, int(seed), out.ctypes.data)
        return out


class Reference:
    """The reference cpu/ port compiled from its sources (oracle/_ref/libmcref.so)."""

    def __init__(self, ir_path=IR_PATH, terrain_path=TERRAIN_PATH):
        if not os.path.exists(REF_PATH):
            raise RuntimeError(f'{REF_PATH} is not built: run make -C oracle ref')
        L = ctypes.CDLL(REF_PATH)
        L.ref_last_error.restype = ctypes.c_char_p
        L.ref_create.restype = P
        L.ref_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
        L.ref_destroy.argtypes = [P]
        L.ref_generate_terrain.argtypes = [P, I64, P, I, P, P, P, P]
        L.ref_generate_region.argtypes = [P, I64, I, I, I, I, I, P, P]
        L.ref_sample_id.argtypes = [P, I64, I, P, P]
        L.ref_point_id.argtypes = [P, I64, I, P, I, P]
        L.ref_tree_search.argtypes = [P, P, P, I, P]
        L.ref_zoom.argtypes = [I64, P, I, P]
        L.ref_seed_misc.argtypes = [P, I64, P, P, P]
        L.ref_seed_layers.argtypes = [P, I64, P, P]
        L.ref_noise_ref.argtypes = [P, I64, ctypes.c_char_p, P, P]
        L.ref_chunk_biomes.argtypes = [P, I64, I, I, P]
        L.ref_generate_core.argtypes = [P, I64, P, I, P, P]
        L.ref_digest_terrain.argtypes = [P, P, P, I64, I, I, P, P]
        L.ref_query_biomes.argtypes = [P, P, P, I64, I, I, P]
        self.L = L
        self.h = L.ref_create(ir_path.encode(), terrain_path.encode())
        if not self.h:
            raise RuntimeError('reference: ' + L.ref_last_error().decode())

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError('reference: ' + self.L.ref_last_error().decode())

    def generate(self, seed, chunks, density=False):
        n = len(chunks)
        xz = np.ascontiguousarray(np.array(chunks, dtype=np.int32).reshape(n, 2))
        t = np.zeros((n, HEIGHT, 16, 16), np.uint16)
        a = np.zeros((n, HEIGHT, 16, 16), np.uint16)
        b = np.zeros((n, HEIGHT // 4, 4, 4), np.uint8)
        d = np.zeros((n, HEIGHT, 16, 16), np.float32) if density else None
        self._check(self.L.ref_generate_terrain(self.h, int(seed), xz.ctypes.data, n, b.ctypes.data, a.ctypes.data, t.ctypes.data, _ptr(d)))
        return (t, b, a, d) if density else (t, b, a)

    def generate_core(self, seed, chunks):
        """Core configuration (cpu_generate_chunk): (density f32, fill u8), (n, 384, 16, 16)."""
        n = len(chunks)
        xz = np.ascontiguousarray(np.array(chunks, dtype=np.int32).reshape(n, 2))
        d = np.zeros((n, HEIGHT, 16, 16), np.float32)
        f = np.zeros((n, HEIGHT, 16, 16), np.uint8)
        self._check(self.L.ref_generate_core(self.h, int(seed), xz.ctypes.data, n, d.ctypes.data, f.ctypes.data))
        return d, f

    def digest_terrain(self, seeds, xz, threads=8, run=16):
        """Per-chunk 64-bit digests (terrain u8, biomes) of the full TERRAIN step of chunks
        xz [n, 2] under seeds [n], in the form of the library's device digest."""
        seeds = np.ascontiguousarray(seeds, np.int64)
        xz = np.ascontiguousarray(xz, np.int32).reshape(-1, 2)
        n = len(xz)
        if len(seeds) != n:
            raise ValueError('one seed per chunk')
        dt, db = np.zeros(n, np.uint64), np.zeros(n, np.uint64)
        self._check(self.L.ref_digest_terrain(self.h, seeds.ctypes.data, xz.ctypes.data, n, threads, run, dt.ctypes.data, db.ctypes.data))
        return dt, db

    def query_biomes(self, seeds, xyz, threads=8, run=64):
        """Answers of biome-only block queries: the biome stored for each block's zoomed quart."""
        seeds = np.ascontiguousarray(seeds, np.int64)
        xyz = np.ascontiguousarray(xyz, np.int32).reshape(-1, 3)
        n = len(xyz)
        if len(seeds) != n:
            raise ValueError('one seed per query')
        out = np.zeros(n, np.uint8)
        self._check(self.L.ref_query_biomes(self.h, seeds.ctypes.data, xyz.ctypes.data, n, threads, run, out.ctypes.data))
        return out

    def region(self, seed, x0, z0, w, d, threads=8):
        t = np.zeros((w * d, HEIGHT, 16, 16), np.uint16)
        b = np.zeros((w * d, HEIGHT // 4, 4, 4), np.uint8)
        self._check(self.L.ref_generate_region(self.h, int(seed), x0, z0, w, d, threads, b.ctypes.data, t.ctypes.data))
        return t, b

    def sample(self, seed, node, vol):
        v = np.array(vol, dtype=np.int32)
        out = np.zeros(int(v[0]) * int(v[1]) * int(v[2]), np.float32)
        self._check(self.L.ref_sample_id(self.h, int(seed), int(node), v.ctypes.data, out.ctypes.data))
        return out

    def noise_layers(self, seed, key):
        n = self.L.ref_seed_layers(self.h, int(seed), None, None)
        if n < 0:
            self._check(1)
        params, perms = np.zeros((n, 6), np.float64), np.zeros((n, 256), np.uint8)
        self.L.ref_seed_layers(self.h, int(seed), params.ctypes.data, perms.ctypes.data)
        first, count = ctypes.c_int(), ctypes.c_int()
        self._check(self.L.ref_noise_ref(self.h, int(seed), key.encode(), ctypes.byref(first), ctypes.byref(count)))
        return params[first.value:first.value + count.value].copy(), perms[first.value:first.value + count.value].copy()

    def seed_misc(self, seed):
        zs = ctypes.c_int64()
        pos = np.zeros(6, np.uint64)
        clay = np.zeros(192, np.int32)
        self._check(self.L.ref_seed_misc(self.h, int(seed), ctypes.byref(zs), pos.ctypes.data, clay.ctypes.data))
        return zs.value, pos, clay

    def tree_search(self, targets, candidates):
        t = np.ascontiguousarray(np.asarray(targets, dtype=np.int64).reshape(-1, 7))
        c = np.ascontiguousarray(np.asarray(candidates, dtype=np.int32))
        out = np.zeros(len(t), np.int32)
        self._check(self.L.ref_tree_search(self.h, t.ctypes.data, c.ctypes.data, len(t), out.ctypes.data))
        return out

    def zoom(self, zoom_seed, xyz):
        p = np.ascontiguousarray(np.array(xyz, dtype=np.int32).reshape(-1, 3))
        out = np.zeros_like(p)
        self.L.ref_zoom(int(zoom_seed), p.ctypes.data, len(p), out.ctypes.data)
        return out

    def chunk_biomes(self, seed, cx, cz):
        out = np.zeros(QUARTS, np.uint8)
        self._check(self.L.ref_chunk_biomes(self.h, int(seed), int(cx), int(cz), out.ctypes.data))
        return out

    def point(self, seed, node, xyz):
        p = np.ascontiguousarray(np.array(xyz, dtype=np.int32).reshape(-1, 3))
        out = np.zeros(len(p), np.float32)
        self._check(self.L.ref_point_id(self.h, int(seed), int(node), p.ctypes.data, len(p), out.ctypes.data))
        return out
