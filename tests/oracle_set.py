"""Writes a small `mccuda-oracle/1` set (with its terrain part) from the oracle, in the layout
the reference's Java oracle writes from the game (oracle/src/main/java/mccuda/oracle/
ChunkDump.java:58-246, TerrainDump.java:100-172): test data for the set reader and the
set checker, which otherwise need the game's own dumps."""
import hashlib
import json
import os

import numpy as np

CORE_PALETTE = ['minecraft:air', 'minecraft:stone', 'minecraft:water', 'minecraft:lava']


def file_name(name):
    return name.replace(':', '_').replace('/', '.') + '.f32'


def write_set(directory, orc, seeds, named=(), terrain=True, min_y=-64, height=384):
    """seeds: [(seed, [(cx, cz), ...]), ...]; named: IR root names dumped over each seed's first
    chunk."""
    import oracle
    ir = json.load(open(oracle.IR_PATH))
    tm = json.load(open(oracle.TERRAIN_PATH))
    hashes, thashes, mseeds = {}, {}, []
    for d in ('density', 'fill', 'nodes', 'terrain'):
        os.makedirs(os.path.join(directory, d), exist_ok=True)

    def put(rel, data, table):
        with open(os.path.join(directory, rel), 'wb') as f:
            f.write(data)
        table[rel] = hashlib.sha256(data).hexdigest()

    for i, (seed, chunks) in enumerate(seeds):
        t, b, a, d = orc.generate(seed, chunks, density=True)
        put('density/s%d.f32' % i, d.astype('<f4').tobytes(), hashes)
        put('fill/s%d.u8' % i, oracle.core_fill(d).tobytes(), hashes)
        dumps = []
        if named:
            dump = 's%d_c0' % i
            os.makedirs(os.path.join(directory, 'nodes', dump), exist_ok=True)
            cx, cz = chunks[0]
            for name in named:
                v = orc.sample(seed, ir['roots'][name], [16, height, 16, cx * 16, min_y, cz * 16, 1, 1, 1])
                v = v.reshape(16, 16, height).transpose(2, 0, 1)      # volume index y + (x + z * 16) * h -> (y, z, x)
                put('nodes/%s/%s' % (dump, file_name(name)), np.ascontiguousarray(v).astype('<f4').tobytes(), hashes)
            dumps.append(dump)
        if terrain:
            put('terrain/biomes_s%d.u8' % i, b.tobytes(), thashes)
            put('terrain/aquifer_s%d.u16' % i, a.astype('<u2').tobytes(), thashes)
            put('terrain/terrain_s%d.u16' % i, t.astype('<u2').tobytes(), thashes)
        mseeds.append({'index': i, 'seed': str(seed), 'quick': i == 0, 'chunks': [[cx, cz, 'c%d' % k] for k, (cx, cz) in enumerate(chunks)],
                       'node_dumps': dumps})
    manifest = {'format': 'mccuda-oracle/1', 'minecraft': 'synthetic', 'quick': False, 'min_y': min_y, 'height': height,
                'layout': 'index = ((y - min_y) * 16 + z) * 16 + x; density float32 LE, fill uint8 palette index', 'palette': CORE_PALETTE,
                'named_functions': [{'name': n, 'file': file_name(n)} for n in named], 'seeds': mseeds, 'sha256': hashes}
    json.dump(manifest, open(os.path.join(directory, 'manifest.json'), 'w'), indent=1)
    if terrain:
        json.dump({'format': 'mccuda-terrain-oracle/1', 'minecraft': 'synthetic', 'palette': tm['palette'], 'biome_names': tm['biome_names'],
                   'sha256': thashes}, open(os.path.join(directory, 'manifest_terrain.json'), 'w'), indent=1)
