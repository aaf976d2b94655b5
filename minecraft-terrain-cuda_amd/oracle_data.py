"""Reader of the reference's oracle dump sets (format `mccuda-oracle/1` with the terrain part
`mccuda-terrain-oracle/1`), the files its Java oracle writes from the game itself
(oracle/src/main/java/mccuda/oracle/ChunkDump.java:58-246, TerrainDump.java:100-172; reader
tests/oracle_data.cpp:27-63). A directory holds:

  manifest.json                      format, min_y, height, palette of the core fill, named
                                     functions [{name, file}], seeds [{index, seed (string),
                                     quick, chunks [[cx, cz, tag]], node_dumps}], sha256 {path: hex}
  density/s<i>.f32                   final density per block, float32 LE, chunks back to back
  fill/s<i>.u8                       core fill (aquifers off), indices into the manifest palette
  nodes/s<i>_c<k>/<file>             named density functions over chunk k of seed i, float32
  manifest_terrain.json              palette, biome_names, sha256 (the TERRAIN part)
  terrain/biomes_s<i>.u8             quart biomes, (qy * 4 + qz) * 4 + qx
  terrain/aquifer_s<i>.u16           blocks after doFill, palette indices, uint16 LE
  terrain/terrain_s<i>.u16           blocks after buildSurface, palette indices, uint16 LE

Block layout everywhere: ((y - min_y) * 16 + z) * 16 + x. Arrays come back shaped
(chunks, height, 16, 16) / (chunks, height / 4, 4, 4).
"""
import hashlib
import json
import os
from dataclasses import dataclass, field

import numpy as np

FORMAT = 'mccuda-oracle/1'
TERRAIN_FORMAT = 'mccuda-terrain-oracle/1'


@dataclass
class OracleChunk:
    cx: int
    cz: int
    tag: str


@dataclass
class OracleSeed:
    index: int
    seed: int
    quick: bool
    chunks: list = field(default_factory=list)
    node_dumps: list = field(default_factory=list)

    def xz(self):
        return [(c.cx, c.cz) for c in self.chunks]


class OracleError(RuntimeError):
    pass


class OracleSet:
    def __init__(self, directory):
        self.dir = directory
        path = os.path.join(directory, 'manifest.json')
        if not os.path.exists(path):
            raise OracleError('no oracle manifest in %s' % directory)
        m = json.load(open(path))
        if m.get('format') != FORMAT:
            raise OracleError('unexpected oracle format %r (want %s)' % (m.get('format'), FORMAT))
        self.manifest = m
        self.min_y, self.height = int(m['min_y']), int(m['height'])
        self.palette = list(m.get('palette', []))
        self.named = [(n['name'], n['file']) for n in m.get('named_functions', [])]
        self.seeds = [OracleSeed(int(s['index']), int(s['seed']), bool(s['quick']), [OracleChunk(int(c[0]), int(c[1]), str(c[2])) for c in s['chunks']],
                                 list(s.get('node_dumps', []))) for s in m['seeds']]
        tpath = os.path.join(directory, 'manifest_terrain.json')
        self.terrain_manifest = None
        if os.path.exists(tpath):
            t = json.load(open(tpath))
            if t.get('format') != TERRAIN_FORMAT:
                raise OracleError('unexpected terrain oracle format %r (want %s)' % (t.get('format'), TERRAIN_FORMAT))
            self.terrain_manifest = t

    def chunk_blocks(self):
        return 16 * 16 * self.height

    def _read(self, rel, dtype, count, shape):
        path = os.path.join(self.dir, rel)
        if not os.path.exists(path):
            raise OracleError('cannot open ' + path)
        a = np.fromfile(path, dtype=np.dtype(dtype).newbyteorder('<'))
        if a.size != count:
            raise OracleError('%s: %d values, want %d' % (rel, a.size, count))
        return a.reshape(shape)

    def density(self, s):
        n = len(s.chunks)
        return self._read('density/s%d.f32' % s.index, np.float32, n * self.chunk_blocks(), (n, self.height, 16, 16))

    def fill(self, s):
        n = len(s.chunks)
        return self._read('fill/s%d.u8' % s.index, np.uint8, n * self.chunk_blocks(), (n, self.height, 16, 16))

    def node(self, dump, file):
        return self._read('nodes/%s/%s' % (dump, file), np.float32, self.chunk_blocks(), (self.height, 16, 16))

    def terrain(self, s):
        """(biomes u8, aquifer u16, terrain u16) of seed s, or OracleError without the terrain part."""
        if self.terrain_manifest is None:
            raise OracleError('no terrain manifest in ' + self.dir)
        n = len(s.chunks)
        b = self._read('terrain/biomes_s%d.u8' % s.index, np.uint8, n * 16 * (self.height // 4), (n, self.height // 4, 4, 4))
        a = self._read('terrain/aquifer_s%d.u16' % s.index, np.uint16, n * self.chunk_blocks(), (n, self.height, 16, 16))
        t = self._read('terrain/terrain_s%d.u16' % s.index, np.uint16, n * self.chunk_blocks(), (n, self.height, 16, 16))
        return b, a, t

    def verify_hashes(self):
        """Files whose SHA-256 differs from the manifests' (missing files included)."""
        bad = []
        for m in (self.manifest, self.terrain_manifest or {}):
            for rel, want in m.get('sha256', {}).items():
                path = os.path.join(self.dir, rel)
                if not os.path.exists(path) or hashlib.sha256(open(path, 'rb').read()).hexdigest() != want:
                    bad.append(rel)
        return bad


def load_oracle_set(directory):
    return OracleSet(directory)
