"""CRV photon lookup tables, format version 6: reader, writer and the in-memory table.

Follows the reference's response/lookup.py:1-180 (the format Offline CRVResponse MakeCrvPhotons::LoadLookupTable
reads, written by lookupgen/mu2e/writer.py:22-91). Binary layout, little-endian, as the C++ structs are written:

  LookupConstants       3 int32 (version major, minor, reflector), 4 bytes of struct padding, 10 float64 (half
                        thickness, half width, half length, fibre separation, hole radius x, hole radius y, fibre
                        radius, corner radius, Birks constant, fibre decay time) = 96 bytes;
  LookupCerenkov        two maps, each uint64 n then n pairs of float64 (beta, photons per mm): scintillator, fibre;
  LookupBinDefinitions  7 vectors, each uint64 n then n float64: x, y, z, beta, theta, phi, r bin edges;
  bins                  the three tables in order, scintillation in the scintillator [x, y, z], Cerenkov in the
                        scintillator [x, y, z, beta], Cerenkov in the fibre [beta, theta, phi, r, z] (last index
                        fastest); each bin is uint32 bin number, float32 arrival probability at SiPM 0, uint64 n + n
                        uint8 delay weights (index = delay in ns), uint64 n + n uint8 fibre-emission weights.

In memory the weights of a table are uint8 rows padded with zeros to the table's longest vector, with their int32
row sums (a delay index i is drawn with probability w_i / sum w). The released tables are Mu2e-internal: tests use
tables written by `write` from synthetic content. Loads are not checked against a pinned content hash (the
reference's library pinning needs its hepsimlib package); `provenance` records the file, its size and its sha256.
"""
import hashlib
import os
import struct
from dataclasses import dataclass, field

import numpy as np

DEFAULT_DIR = "/cvmfs/mu2e.opensciencegrid.org/DataFiles/CRVConditions/v6_0"
TABLES = ("scint", "ckscint", "ckfiber")
EDGES = ("x", "y", "z", "beta", "theta", "phi", "r")
CONSTANTS = ("half_thickness", "half_width", "half_length", "fiber_separation", "hole_radius_x", "hole_radius_y",
             "fiber_radius", "corner_radius", "birks", "fiber_decay_time")


def lookup_dir():
    return os.environ.get("CRVFS_LOOKUP_DIR", DEFAULT_DIR)


@dataclass
class Table:
    prob: np.ndarray        # [N] float32 arrival probability at SiPM 0
    delays: np.ndarray      # [N, Dmax] uint8 weights
    delay_sum: np.ndarray   # [N] int32
    emis: np.ndarray        # [N, Emax] uint8 weights
    emis_sum: np.ndarray    # [N] int32


@dataclass
class LookupTable:
    name: str
    version: tuple
    reflector: int
    half_thickness: float
    half_width: float
    half_length: float
    fiber_separation: float
    hole_radius_x: float
    hole_radius_y: float
    fiber_radius: float
    corner_radius: float
    birks: float
    fiber_decay_time: float
    ck_scint: np.ndarray    # [n, 2] (beta, photons per mm), sorted by beta
    ck_fiber: np.ndarray
    edges: dict             # x, y, z, beta, theta, phi, r -> float64 arrays
    tables: dict            # scint, ckscint, ckfiber -> Table
    provenance: dict = field(default=None, compare=False)


def n_bins(edges):
    """Bin counts of the three tables for a dict of edges."""
    nb = {k: len(v) - 1 for k, v in edges.items()}
    s = nb["x"] * nb["y"] * nb["z"]
    return {"scint": s, "ckscint": s * nb["beta"],
            "ckfiber": nb["beta"] * nb["theta"] * nb["phi"] * nb["r"] * nb["z"]}


def table_from_weights(prob, delays, emis):
    """A Table from per-bin probabilities and weight rows (lists of uint8 vectors or padded 2-D arrays)."""
    def pad(rows):
        if isinstance(rows, np.ndarray) and rows.ndim == 2:
            return np.ascontiguousarray(rows, np.uint8)
        m = max([len(r) for r in rows], default=0)
        out = np.zeros((len(rows), max(m, 1)), np.uint8)
        for i, r in enumerate(rows):
            out[i, :len(r)] = r
        return out
    d, e = pad(delays), pad(emis)
    return Table(np.asarray(prob, np.float32), d, d.sum(1, dtype=np.int32), e, e.sum(1, dtype=np.int32))


def _read_bins(buf, o, n):
    """n bin records from offset o: (Table, new offset). Weight vectors keep their stored lengths up to the padding."""
    prob = np.empty(n, np.float32)
    d_off, d_len, e_off, e_len = (np.empty(n, np.int64) for _ in range(4))
    unpack = struct.Struct("<If").unpack_from
    size_t = struct.Struct("<Q").unpack_from
    for i in range(n):
        num, p = unpack(buf, o)
        if num != i:
            raise ValueError(f"corrupt lookup table: bin {num} at position {i}")
        prob[i] = p
        nd, = size_t(buf, o + 8)
        d_off[i], d_len[i] = o + 16, nd
        o += 16 + nd
        ne, = size_t(buf, o)
        e_off[i], e_len[i] = o + 8, ne
        o += 8 + ne
    if o > len(buf):
        raise ValueError("corrupt lookup table: truncated bin records")
    arr = np.frombuffer(buf, np.uint8)

    def gather(off, ln):
        m = max(int(ln.max()) if n else 0, 1)
        out = np.zeros((n, m), np.uint8)
        col = np.arange(m)[None, :]
        ok = col < ln[:, None]
        out[ok] = arr[(off[:, None] + col)[ok]]
        return out

    d, e = gather(d_off, d_len), gather(e_off, e_len)
    return Table(prob, d, d.sum(1, dtype=np.int32), e, e.sum(1, dtype=np.int32)), o


def from_bytes(buf, name="LookupTable"):
    """Decode a version-6 table (bytes)."""
    if len(buf) < 96:
        raise ValueError(f"{name}: {len(buf)} bytes, shorter than the table header")
    v1, v2, refl = struct.unpack_from("<3i", buf, 0)
    consts = struct.unpack_from("<10d", buf, 16)
    if v1 != 6:
        raise ValueError(f"{name}: lookup table version {v1}.{v2}, expected 6.x")
    o = 96

    def vec(width):
        nonlocal o
        if o + 8 > len(buf):
            raise ValueError(f"{name}: truncated header")
        n, = struct.unpack_from("<Q", buf, o)
        o += 8
        if o + 8 * width * n > len(buf):
            raise ValueError(f"{name}: truncated header")
        a = np.frombuffer(buf, "<f8", width * n, o).astype(np.float64)
        o += 8 * width * n
        return a

    ck_s, ck_f = vec(2).reshape(-1, 2), vec(2).reshape(-1, 2)
    edges = {k: vec(1) for k in EDGES}
    if any(len(v) < 2 for v in edges.values()):
        raise ValueError(f"{name}: every bin-edge vector needs at least two edges")
    tables = {}
    for t, n in n_bins(edges).items():
        tables[t], o = _read_bins(buf, o, n)
    if o != len(buf):
        raise ValueError(f"{name}: {len(buf) - o} trailing bytes")
    return LookupTable(name, (v1, v2), refl, *(float(x) for x in consts), ck_s, ck_f, edges, tables)


def load(name, directory=None):
    """Read LookupTable_<name> (e.g. '6000_0') from `directory` (default: $CRVFS_LOOKUP_DIR or the CVMFS path);
    a name starting with 'LookupTable' is taken as the file name."""
    base = name if name.startswith("LookupTable") else f"LookupTable_{name}"
    path = os.path.join(directory or lookup_dir(), base)
    with open(path, "rb") as fh:
        buf = fh.read()
    L = from_bytes(buf, base)
    L.provenance = {"name": f"crv.lookup.{base.replace('LookupTable_', '')}", "file": base, "size": len(buf),
                    "sha256": hashlib.sha256(buf).hexdigest(), "status": "unpinned"}
    return L


def to_bytes(L):
    """The version-6 file of a LookupTable; each bin's weight vectors are written up to their last nonzero entry."""
    out = bytearray(struct.pack("<3i4x", int(L.version[0]), int(L.version[1]), int(L.reflector)))
    out += struct.pack("<10d", *(float(getattr(L, k)) for k in CONSTANTS))
    for ck in (L.ck_scint, L.ck_fiber):
        ck = np.asarray(ck, "<f8").reshape(-1, 2)
        out += struct.pack("<Q", len(ck)) + ck.tobytes()
    for k in EDGES:
        e = np.asarray(L.edges[k], "<f8")
        out += struct.pack("<Q", len(e)) + e.tobytes()
    nb = n_bins(L.edges)
    for t in TABLES:
        T = L.tables[t]
        if len(T.prob) != nb[t]:
            raise ValueError(f"table {t}: {len(T.prob)} bins, the edges give {nb[t]}")
        for i in range(nb[t]):
            out += struct.pack("<If", i, float(T.prob[i]))
            for w in (T.delays[i], T.emis[i]):
                nz = np.flatnonzero(w)
                m = nz[-1] + 1 if len(nz) else 0
                out += struct.pack("<Q", m) + np.asarray(w[:m], np.uint8).tobytes()
    return bytes(out)


def write(L, path):
    with open(path, "wb") as fh:
        fh.write(to_bytes(L))
