"""Stage-boundary tables in the reference's dump layouts, as parquet or npz (uproot and ROOT are absent here).

Inputs: the CrvStep table of the `CrvStepTableDumper` layout (TFileService directory dump/, tree `steps`; columns
in reference crv/crvfastsim/README.md:30-40 and crv/art/CrvFastSim/src/CrvStepTableDumper_module.cc:123-134): run,
subrun, event, nstep, bar, sp, edep (MeV), t0, t1 (ns), x0 y0 z0 x1 y1 z1 (mm), px0 py0 pz0, p1 (MeV/c), len (mm).
`read_steps` returns it as a pandas DataFrame that the stage functions take as their step table.

Outputs: the `CrvRecoTableDumper` layout (directory crvdump/; reference crv/crvfastsim/reco/reftables.py:1-14 and
crv/art/CrvFastSim/src/CrvRecoTableDumper_module.cc:97-185) for the trees this chain produces:

* events   run, subrun, event, spill, evlen, pbtime, ndigi, npulse, ncluster
* digis    run, subrun, event, idx, bar, sipm, starttdc, nzs, roc, feb, febch, adc[]
* pulses   run, subrun, event, idx, bar, sipm, pe, peph, height, beta, chi2, ped, time, letime, seq, flags,
           pedfromdb, wfidx[]
* clusters run, subrun, event, idx, sectortype, start, end, pe, slope, layers, pulses[], poscalc, avgtime, x, y, z,
           sidepe0, sidepe1, sidetime0, sidetime1, sidehits0, sidehits1
* calib    run, channel, pedestal, pulsearea, pulseheight, timeoffset, status, roc, feb, febch

`idx` is the row's index in its event's collection; a cluster's `pulses` are the `idx` of its member pulses in the
pulse tree; `flags` bit 0 is the failed fit (the bit the reference decodes, reco/cli.py:47); `layers` holds the
sorted distinct layers of the cluster's member pulses (reference reco/coinc.py:391, read back as a jagged int column by
CrvCoincidenceClustersFromTable_module.cc:69); `wfidx` holds the pulse's reco-waveform number in the event (this chain
keeps no art digi collection to point into); spill, evlen and pbtime are 0 (no onspill folding). Column types are the
dumper's branch types (CrvRecoTableDumper_module.cc:59-92; its reader, CrvRecoTable.hh:33-37, keys rows by UInt_t
run, subrun, event): unsigned run, subrun, event; int counters and indices; float PE, height, slope, positions and
calibration constants; double times; short ADC words.

Storage: a path ending in `.npz` is one archive with keys `<tree>/<column>`, a jagged column stored flat with its
per-row counts under `<tree>/n<column>` (the counted-array convention the reference's table reader accepts,
CrvRecoTable.hh:61-79); any other path is a directory of `<tree>.parquet` files with jagged columns as lists.
"""
import os

import numpy as np

STEP_COLUMNS = ("run", "subrun", "event", "nstep", "bar", "sp", "edep", "t0", "t1", "x0", "y0", "z0", "x1", "y1", "z1",
                "px0", "py0", "pz0", "p1", "len")
STEP_REQUIRED = ("subrun", "event", "bar", "edep", "t0", "t1", "z0", "z1")
TREES = {
    "events": ("run", "subrun", "event", "spill", "evlen", "pbtime", "ndigi", "npulse", "ncluster"),
    "digis": ("run", "subrun", "event", "idx", "bar", "sipm", "starttdc", "nzs", "roc", "feb", "febch", "adc"),
    "pulses": ("run", "subrun", "event", "idx", "bar", "sipm", "pe", "peph", "height", "beta", "chi2", "ped", "time",
               "letime", "seq", "flags", "pedfromdb", "wfidx"),
    "clusters": ("run", "subrun", "event", "idx", "sectortype", "start", "end", "pe", "slope", "layers", "pulses",
                 "poscalc", "avgtime", "x", "y", "z", "sidepe0", "sidepe1", "sidetime0", "sidetime1", "sidehits0",
                 "sidehits1"),
    "calib": ("run", "channel", "pedestal", "pulsearea", "pulseheight", "timeoffset", "status", "roc", "feb", "febch"),
}
JAGGED = {"digis": ("adc",), "pulses": ("wfidx",), "clusters": ("layers", "pulses")}
# the dumper's branch types (reference CrvRecoTableDumper_module.cc:59-92; jagged columns: the element type)
_U, _I, _F, _D = np.uint32, np.int32, np.float32, np.float64
_IDS = {"run": _U, "subrun": _U, "event": _U}
_CHAN = {"idx": _I, "bar": _I, "sipm": _I}
DTYPES = {
    "events": {**_IDS, "spill": _I, "evlen": _D, "pbtime": _D, "ndigi": _I, "npulse": _I, "ncluster": _I},
    "digis": {**_IDS, **_CHAN, "starttdc": _I, "nzs": _I, "roc": _I, "feb": _I, "febch": _I, "adc": np.int16},
    "pulses": {**_IDS, **_CHAN, **{k: _F for k in ("pe", "peph", "height", "beta", "chi2", "ped")}, "time": _D,
               "letime": _D, "seq": _I, "flags": _I, "pedfromdb": _I, "wfidx": _I},
    "clusters": {**_IDS, "idx": _I, "sectortype": _I, "start": _D, "end": _D, "pe": _F, "slope": _F, "layers": _I,
                 "pulses": _I, "poscalc": _I, "avgtime": _D, **{k: _F for k in ("x", "y", "z", "sidepe0", "sidepe1")},
                 "sidetime0": _D, "sidetime1": _D, "sidehits0": _I, "sidehits1": _I},
    "calib": {"run": _U, "channel": _I, **{k: _F for k in ("pedestal", "pulsearea", "pulseheight", "timeoffset")},
              "status": _I, "roc": _I, "feb": _I, "febch": _I},
}


def _typed(tree, cols):
    """A tree's columns in the dumper's types."""
    dt = DTYPES[tree]
    return {c: Jagged(v.values.astype(dt[c]), v.counts) if isinstance(v, Jagged) else np.asarray(v).astype(dt[c])
            for c, v in cols.items()}


class Jagged:
    """A jagged column: flat values and per-row counts."""

    def __init__(self, values, counts):
        self.values = np.asarray(values)
        self.counts = np.asarray(counts, np.int64)
        if self.counts.sum() != len(self.values):
            raise ValueError("jagged column: counts do not add up to the values")

    def __len__(self):
        return len(self.counts)

    def offsets(self):
        return np.concatenate([[0], np.cumsum(self.counts)])

    def row(self, i):
        o = self.offsets()
        return self.values[o[i]:o[i + 1]]


def _key_parts(evkey):
    evkey = np.asarray(evkey, np.int64)
    return evkey >> 20, evkey & ((1 << 20) - 1)


def _rank_in_event(ev):
    """Index of each row within its event (rows grouped by event)."""
    ev = np.asarray(ev)
    if len(ev) == 0:
        return np.zeros(0, np.int64)
    start = np.flatnonzero(np.r_[True, ev[1:] != ev[:-1]])
    return np.arange(len(ev)) - np.repeat(start, np.diff(np.r_[start, len(ev)]))


# ---- steps ----------------------------------------------------------------------------------------------------------

def read_steps(path):
    """The CrvStep table of a dump file (parquet file or directory holding steps.parquet, or npz with keys steps/<col>
    or <col>), sorted by (run, subrun, event) with the file's row order kept within an event."""
    import pandas as pd
    if str(path).endswith(".npz"):
        with np.load(path, allow_pickle=False) as z:
            pre = "steps/" if any(k.startswith("steps/") for k in z.files) else ""
            df = pd.DataFrame({k[len(pre):]: z[k] for k in z.files if k.startswith(pre) and "/" not in k[len(pre):]})
    else:
        f = os.path.join(path, "steps.parquet") if os.path.isdir(path) else path
        df = pd.read_parquet(f)
    missing = [c for c in STEP_REQUIRED if c not in df.columns]
    if missing:
        raise ValueError(f"step table lacks columns {missing} (CrvStepTableDumper layout)")
    if "run" not in df.columns:
        df["run"] = 0
    return df.sort_values(["run", "subrun", "event"], kind="stable").reset_index(drop=True)


def steps_frame(ev, counter, edep, z, t, run=0):
    """A synthetic step table (event key, counter, edep, z along the counter, t) in the dump layout: a point step
    (x0 = x1, ...) at the counter, its event key split into subrun and event; other kinematics 0."""
    import pandas as pd
    sub, evt = _key_parts(ev)
    n = len(ev)
    z = np.asarray(z, float)
    t = np.asarray(t, float)
    nstep = np.bincount(_group_ids(ev), minlength=1)[_group_ids(ev)] if n else np.zeros(0, np.int64)
    return pd.DataFrame({"run": np.full(n, run, np.int64), "subrun": sub, "event": evt, "nstep": nstep,
                         "bar": np.asarray(counter, np.int64), "sp": np.zeros(n, np.int64),
                         "edep": np.asarray(edep, float), "t0": t, "t1": t, "x0": np.zeros(n), "y0": np.zeros(n),
                         "z0": z, "x1": np.zeros(n), "y1": np.zeros(n), "z1": z, "px0": np.zeros(n),
                         "py0": np.zeros(n), "pz0": np.zeros(n), "p1": np.zeros(n), "len": np.zeros(n)})


def _group_ids(ev):
    ev = np.asarray(ev)
    if len(ev) == 0:
        return np.zeros(0, np.int64)
    return np.cumsum(np.r_[True, ev[1:] != ev[:-1]]) - 1


def write_steps(path, df):
    """Write a step table in the dump layout (parquet file, or npz with keys steps/<col>)."""
    cols = {c: np.asarray(df[c]) for c in STEP_COLUMNS if c in df}
    if str(path).endswith(".npz"):
        np.savez(path, **{f"steps/{c}": v for c, v in cols.items()})
    else:
        import pandas as pd
        pd.DataFrame(cols).to_parquet(path, index=False)


# ---- reco tables ----------------------------------------------------------------------------------------------------

def pulses_tree(P, run=0):
    """crvdump/pulses from a pulse table (ev, chan, wf, seq, pe, peph, time, height, beta, chi2, letime, ped,
    failed, pedfromdb; rows grouped by event in collection order)."""
    ev = np.asarray(P["ev"], np.int64)
    sub, evt = _key_parts(ev)
    chan = np.asarray(P["chan"], np.int64)
    n = len(ev)
    wf = np.asarray(P["wf"], np.int64)
    wf_in_event = wf - np.repeat(_first_in_event(ev, wf), np.diff(np.r_[_starts(ev), n])) if n else wf
    out = {"run": np.full(n, run, np.int64), "subrun": sub, "event": evt, "idx": _rank_in_event(ev),
            "bar": chan // 4, "sipm": chan % 4,
            **{k: np.asarray(P[k], np.float64) for k in ("pe", "peph", "height", "beta", "chi2", "ped", "time",
                                                          "letime")},
            "seq": np.asarray(P["seq"], np.int64), "flags": np.asarray(P["failed"], np.int64) & 1,
            "pedfromdb": np.asarray(P["pedfromdb"], bool), "wfidx": Jagged(wf_in_event, np.ones(n, np.int64))}
    return _typed("pulses", out)


def _starts(ev):
    ev = np.asarray(ev)
    return np.flatnonzero(np.r_[True, ev[1:] != ev[:-1]]) if len(ev) else np.zeros(0, np.int64)


def _first_in_event(ev, v):
    return np.asarray(v)[_starts(ev)]


def clusters_tree(C, layer_of_bar, pulse_bar, pulse_ev, run=0):
    """crvdump/clusters from a cluster table (cl_*, mem_cl, mem_idx, mem_event); member pulses are given by their
    index in their event's pulse collection, whose bars (pulse_bar, with events pulse_ev) give the layer count."""
    ev = np.asarray(C["cl_event"], np.int64)
    sub, evt = _key_parts(ev)
    nc = len(ev)
    mem_cl = np.asarray(C["mem_cl"], np.int64)
    mem_idx = np.asarray(C["mem_idx"], np.int64)
    counts = np.bincount(mem_cl, minlength=nc)[:nc] if nc else np.zeros(0, np.int64)
    # layers: the sorted distinct layers of each cluster's member pulses
    pev = np.asarray(pulse_ev, np.int64)  # sorted by event
    rows = np.searchsorted(pev, np.repeat(ev, counts), side="left") + mem_idx
    lay = np.asarray(layer_of_bar)[np.asarray(pulse_bar, np.int64)[rows]] if len(rows) else rows
    layers = Jagged(np.zeros(0, np.int64), np.zeros(nc, np.int64))
    if len(lay):
        pair = np.unique(mem_cl * 8 + np.asarray(lay, np.int64))
        layers = Jagged(pair % 8, np.bincount(pair // 8, minlength=nc)[:nc])
    pos = np.asarray(C["cl_pos"], float).reshape(nc, 3)
    spe = np.asarray(C["cl_sidepe"], float).reshape(nc, 2)
    st = np.asarray(C["cl_sidetime"], float).reshape(nc, 2)
    sh = np.asarray(C["cl_sidehits"], np.int64).reshape(nc, 2)
    return _typed("clusters", {
            "run": np.full(nc, run, np.int64), "subrun": sub, "event": evt, "idx": _rank_in_event(ev),
            "sectortype": np.asarray(C["cl_type"], np.int64), "start": np.asarray(C["cl_start"], float),
            "end": np.asarray(C["cl_end"], float), "pe": np.asarray(C["cl_pe"], float),
            "slope": np.asarray(C["cl_slope"], float), "layers": layers, "pulses": Jagged(mem_idx, counts),
            "poscalc": np.asarray(C["cl_poscalc"], bool), "avgtime": np.asarray(C["cl_avgtime"], float),
            "x": pos[:, 0], "y": pos[:, 1], "z": pos[:, 2], "sidepe0": spe[:, 0], "sidepe1": spe[:, 1],
            "sidetime0": st[:, 0], "sidetime1": st[:, 1], "sidehits0": sh[:, 0], "sidehits1": sh[:, 1]})


def digis_tree(D, cond, run=0):
    """crvdump/digis from a digi table (chan = SiPM label = event key * 32768 + channel, starttdc, adc [n, 12])."""
    from .detector import CHANNELS_PER_FEB
    lab = np.asarray(D["chan"], np.int64)
    ev, ch = lab >> 15, lab & 32767
    sub, evt = _key_parts(ev)
    adc = np.asarray(D["adc"])
    n = len(lab)
    return _typed("digis", {
        "run": np.full(n, run, np.int64), "subrun": sub, "event": evt, "idx": _rank_in_event(ev),
        "bar": ch // 4, "sipm": ch % 4, "starttdc": np.asarray(D["starttdc"], np.int64),
        "nzs": np.zeros(n, bool), "roc": np.asarray(cond.roc, np.int64)[ch],
        "feb": np.asarray(cond.feb, np.int64)[ch], "febch": ch % CHANNELS_PER_FEB,
        "adc": Jagged(adc.reshape(-1).astype(np.int16), np.full(n, adc.shape[1] if adc.ndim == 2 else 0))})


def calib_tree(cond, run=0):
    """crvdump/calib from the conditions (one row per channel)."""
    from .detector import CHANNELS_PER_FEB
    n = len(cond.pedestal)
    ch = np.arange(n, dtype=np.int64)
    return _typed("calib", {
        "run": np.full(n, run, np.int64), "channel": ch, "pedestal": np.asarray(cond.pedestal, float),
        "pulsearea": np.asarray(cond.pulsearea, float), "pulseheight": np.asarray(cond.pulseheight, float),
        "timeoffset": np.asarray(cond.timeoffset, float), "status": np.asarray(cond.status, np.int64),
        "roc": np.asarray(cond.roc, np.int64), "feb": np.asarray(cond.feb, np.int64),
        "febch": ch % CHANNELS_PER_FEB})


def events_tree(event_keys, pulses, clusters, digis=None, run=0):
    """crvdump/events: one row per event key with its digi, pulse and cluster counts."""
    ev = np.unique(np.asarray(event_keys, np.int64))
    sub, evt = _key_parts(ev)

    def count(tree):
        if tree is None:
            return np.zeros(len(ev), np.int64)
        k = (np.asarray(tree["subrun"], np.int64) << 20) | np.asarray(tree["event"], np.int64)
        return np.searchsorted(k, ev, side="right") - np.searchsorted(k, ev, side="left")

    n = len(ev)
    return _typed("events", {
        "run": np.full(n, run, np.int64), "subrun": sub, "event": evt, "spill": np.zeros(n, np.int64),
        "evlen": np.zeros(n, float), "pbtime": np.zeros(n, float), "ndigi": count(digis),
        "npulse": count(pulses), "ncluster": count(clusters)})


def write_reco(path, trees):
    """Write crvdump trees ({name: {column: array or Jagged}}) to `path` (npz archive or parquet directory)."""
    for name, cols in trees.items():
        if name not in TREES:
            raise ValueError(f"unknown crvdump tree {name}")
        if tuple(cols) != TREES[name]:
            raise ValueError(f"{name}: columns {tuple(cols)} differ from the dumper's {TREES[name]}")
    if str(path).endswith(".npz"):
        out = {}
        for name, cols in trees.items():
            for c, v in cols.items():
                if isinstance(v, Jagged):
                    out[f"{name}/{c}"] = v.values
                    out[f"{name}/n{c}"] = v.counts
                else:
                    out[f"{name}/{c}"] = np.asarray(v)
        np.savez(path, **out)
        return
    import pyarrow as pa
    import pyarrow.parquet as pq
    os.makedirs(path, exist_ok=True)
    for name, cols in trees.items():
        arrs = {}
        for c, v in cols.items():
            if isinstance(v, Jagged):
                arrs[c] = pa.ListArray.from_arrays(pa.array(v.offsets().astype(np.int32)), pa.array(v.values))
            else:
                arrs[c] = pa.array(np.asarray(v))
        pq.write_table(pa.table(arrs), os.path.join(path, f"{name}.parquet"))


def read_reco(path, tree):
    """One crvdump tree as {column: numpy array or Jagged}."""
    if tree not in TREES:
        raise ValueError(f"unknown crvdump tree {tree}")
    if str(path).endswith(".npz"):
        with np.load(path, allow_pickle=False) as z:
            out = {}
            for c in TREES[tree]:
                if c in JAGGED.get(tree, ()):
                    out[c] = Jagged(z[f"{tree}/{c}"], z[f"{tree}/n{c}"])
                else:
                    out[c] = z[f"{tree}/{c}"]
            return out
    import pyarrow.parquet as pq
    t = pq.read_table(os.path.join(path, f"{tree}.parquet"))
    out = {}
    for c in TREES[tree]:
        col = t.column(c).combine_chunks()
        if c in JAGGED.get(tree, ()):
            out[c] = Jagged(col.values.to_numpy(zero_copy_only=False), np.diff(col.offsets.to_numpy()))
        else:
            out[c] = col.to_numpy(zero_copy_only=False)
    return out


def reco_trees(out, cond, geom, digis=None, run=0):
    """The crvdump trees of a chain run: `out` is run_chain's result ({"pulses", "clusters"}), digis optional."""
    P = pulses_tree(out["pulses"], run)
    C = clusters_tree(out["clusters"], geom.layer, P["bar"], np.asarray(out["pulses"]["ev"], np.int64), run)
    trees = {"pulses": P, "clusters": C}
    if digis is not None:
        trees["digis"] = digis_tree(digis, cond, run)
    keys = np.asarray(out["pulses"]["ev"], np.int64)
    trees["events"] = events_tree(keys, P, C, trees.get("digis"), run)
    trees["calib"] = calib_tree(cond, run)
    return {k: trees[k] for k in ("events", "digis", "pulses", "clusters", "calib") if k in trees}
