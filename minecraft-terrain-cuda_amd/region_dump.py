"""Region outputs in the formats the reference's consumers read (bench/bench_main.cpp:132-187):

  dump directory   terrain.u8  each chunk's blocks in the canonical layout, biomes.u8 each
                   chunk's quart biomes, both in row-major chunk order ((cz - z0) * region +
                   (cx - x0)); meta.json {seed, x0, z0, region, height, min_y, chunk_bytes,
                   biome_bytes} (what the in-game recorder in client/ loads)
  MCTD file        "MCTD", int32 version 1, region, x0, z0, min_y, then 8 bytes per block
                   column, row-major over the region's blocks ((rz + z) * side + rx + x): top
                   state, floor state, min(255, top_y - floor_y), 0, int16 top y, int16 floor y
                   (the top-down view the videos are drawn from)

The column records are computed on the GPU (mct_topdown), so a top-down view of a region needs
8 bytes per column on the host instead of the whole terrain.

  python region_dump.py --seed 123456789 --x0 0 --z0 0 --region 32 --dump out/ --topdown out/view.mctd
"""
import argparse
import json
import os
import struct

import numpy as np

import mct

MAGIC = b'MCTD'
VERSION = 1


def region_order(x0, z0, region):
    """Chunks of the region in row-major order: index (cz - z0) * region + (cx - x0)."""
    return [(x0 + i % region, z0 + i // region) for i in range(region * region)]


def _positions(order, x0, z0, region):
    pos = np.empty(len(order), np.int64)
    for i, (cx, cz) in enumerate(order):
        k = (cz - z0) * region + (cx - x0)
        if not (0 <= cx - x0 < region and 0 <= cz - z0 < region):
            raise ValueError('chunk (%d, %d) is outside the region' % (cx, cz))
        pos[i] = k
    if len(order) != region * region or len(set(pos.tolist())) != len(order):
        raise ValueError('the chunk list must cover the region once')
    return pos


def write_dump(path, seed, x0, z0, region, order, terrain, biomes, height=384, min_y=-64):
    """terrain [n, 384, 16, 16] / biomes [n, 96, 4, 4] of the chunks `order` (any order covering
    the region once) -> path/terrain.u8, path/biomes.u8, path/meta.json."""
    pos = _positions(order, x0, z0, region)
    inv = np.empty_like(pos)
    inv[pos] = np.arange(len(pos))
    os.makedirs(path, exist_ok=True)
    t = np.ascontiguousarray(terrain, dtype=np.uint8).reshape(len(order), -1)
    b = np.ascontiguousarray(biomes, dtype=np.uint8).reshape(len(order), -1)
    t[inv].tofile(os.path.join(path, 'terrain.u8'))
    b[inv].tofile(os.path.join(path, 'biomes.u8'))
    meta = '{"seed": %d, "x0": %d, "z0": %d, "region": %d, "height": %d, "min_y": %d, "chunk_bytes": %d, "biome_bytes": %d}\n' % (
        seed, x0, z0, region, height, min_y, t.shape[1], b.shape[1])
    with open(os.path.join(path, 'meta.json'), 'w') as f:
        f.write(meta)


def read_dump(path):
    """(meta dict, terrain [region^2, height, 16, 16], biomes [region^2, height / 4, 4, 4]) in
    row-major chunk order."""
    meta = json.load(open(os.path.join(path, 'meta.json')))
    n = meta['region'] ** 2
    t = np.fromfile(os.path.join(path, 'terrain.u8'), np.uint8).reshape(n, meta['height'], 16, 16)
    b = np.fromfile(os.path.join(path, 'biomes.u8'), np.uint8).reshape(n, meta['height'] // 4, 4, 4)
    return meta, t, b


def topdown_raster(records, order, x0, z0, region):
    """Per-chunk column records [n, 256, 8] (z * 16 + x) -> the region raster [side, side, 8]."""
    side = region * 16
    out = np.zeros((side, side, 8), np.uint8)
    r = np.asarray(records, np.uint8).reshape(len(order), 16, 16, 8)
    for i, (cx, cz) in enumerate(order):
        rx, rz = (cx - x0) * 16, (cz - z0) * 16
        out[rz:rz + 16, rx:rx + 16] = r[i]
    return out


def write_topdown(path, x0, z0, region, raster, min_y=-64):
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(path, 'wb') as f:
        f.write(MAGIC)
        f.write(struct.pack('<5i', VERSION, region, x0, z0, min_y))
        f.write(np.ascontiguousarray(raster, np.uint8).tobytes())


def read_topdown(path):
    """(header dict, records [side, side, 8] uint8)."""
    with open(path, 'rb') as f:
        data = f.read()
    if data[:4] != MAGIC:
        raise ValueError('not an MCTD file: ' + path)
    version, region, x0, z0, min_y = struct.unpack('<5i', data[4:24])
    side = region * 16
    rec = np.frombuffer(data, np.uint8, side * side * 8, 24).reshape(side, side, 8)
    return {'version': version, 'region': region, 'x0': x0, 'z0': z0, 'min_y': min_y}, rec


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument('--seed', type=int, default=123456789)
    ap.add_argument('--x0', type=int, default=0)
    ap.add_argument('--z0', type=int, default=0)
    ap.add_argument('--region', type=int, default=32)
    ap.add_argument('--dump', help='directory for terrain.u8 / biomes.u8 / meta.json')
    ap.add_argument('--topdown', help='MCTD file')
    ap.add_argument('--batch', type=int, default=4096)
    a = ap.parse_args()
    if not a.dump and not a.topdown:
        ap.error('nothing to write (--dump and/or --topdown)')
    order = region_order(a.x0, a.z0, a.region)
    n = len(order)
    g = mct.FullTerrainGenerator(max_batch_chunks=min(a.batch, n), pipeline_slots=2)
    seeds = np.full(n, a.seed, np.int64)
    if a.dump:
        t, b = g.generate_multi(seeds, order)
        write_dump(a.dump, a.seed, a.x0, a.z0, a.region, order, t, b)
        rec = mct.topdown(t) if a.topdown else None
    else:
        d_t = mct.alloc_device(n * mct.BLOCKS_PER_CHUNK)
        g.generate_device(seeds, order, d_t)
        rec = mct.topdown(d_t, n)
        mct.free_device(d_t)
    if a.topdown:
        write_topdown(a.topdown, a.x0, a.z0, a.region, topdown_raster(rec, order, a.x0, a.z0, a.region))
    g.close()


if __name__ == '__main__':
    main()
