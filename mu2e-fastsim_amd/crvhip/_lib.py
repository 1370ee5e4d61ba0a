"""ctypes binding of libcrvhip.so (include/crvhip.h) and the device context of the CRV chain.

PyTorch provides device memory (torch.empty(..., device="cuda")) and streams; every computation runs in the HIP
kernels of libcrvhip.so. A missing library is an error: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcrvhip.so")

CRV_OK, CRV_E_ARG, CRV_E_HIP, CRV_E_CAPACITY, CRV_E_STATE, CRV_E_OVERFLOW = 0, -1, -2, -3, -4, -5

P = C.c_void_p
i32, i64, u32, u64, f64 = C.c_int32, C.c_int64, C.c_uint32, C.c_uint64, C.c_double


class Params(C.Structure):
    _fields_ = [("seed", u64), ("evkey_hi", u32), ("precision", i32)] + \
        [(k, f64) for k in ("pe_per_mev", "ref_distance", "w_short", "l_short", "w_long", "l_long", "half_length",
                            "inv_speed", "tau", "photon_window_hi")] + \
        [(k, i32) for k in ("n_pixels_x", "n_pixels_y", "pixel_lrint", "_pad0")] + \
        [(k, f64) for k in ("overvoltage", "time_constant", "capacitance", "p1", "p2", "cross_talk",
                            "single_pe_reference_charge", "single_pe_max_time", "single_pe_precision", "min_voltage",
                            "noise")] + \
        [(k, i32) for k in ("noise_on", "samples_zs", "window_samples", "feb_key_roc")] + \
        [(k, f64) for k in ("adc_conversion", "window_lo", "window_hi", "feb_phase_width")] + \
        [(k, i32) for k in ("pedestal", "min_adc", "max_adc", "_pad1")] + \
        [(k, f64) for k in ("min_adc_difference", "default_beta", "min_beta", "max_beta", "max_time_difference",
                            "min_pulse_height_ratio", "max_pulse_height_ratio", "le_time_factor",
                            "pedestal_undershoot_threshold", "time_offset_scale", "limit_tolerance")] + \
        [(k, i32) for k in ("samples_before", "samples_after", "use_time_offset_db", "ignore_channels",
                            "lm_iterations", "_pad2")] + \
        [(k, f64) for k in ("fiber_signal_speed", "max_time_interval", "time_offset")] + \
        [(k, i32) for k in ("compensate_mask", "ignore_secondary_pulses", "big_cluster_threshold",
                            "add_opposite_side_pulses")]


class Steps(C.Structure):
    _fields_ = [("event", P), ("counter", P), ("edep", P), ("z", P), ("t", P), ("n", i64)]


class Charges(C.Structure):
    _fields_ = [("label", P), ("time", P), ("charge", P), ("pe", P), ("photon", P), ("capacity", i64), ("n", i64),
                ("n_photons", i64), ("capped", i64)]


class PhotonList(C.Structure):
    _fields_ = [("label", P), ("time", P), ("photon", P), ("j", P), ("rank", P), ("n", i64)]


class Photons(C.Structure):
    _fields_ = [("label", P), ("time", P), ("photon", P), ("j", P), ("rank", P), ("capacity", i64), ("n", i64),
                ("capped", i64)]


class LookupTableC(C.Structure):
    _fields_ = [(k, f64) for k in ("half_thickness", "half_width", "half_length", "fiber_separation", "hole_radius_x",
                                   "hole_radius_y", "fiber_radius", "corner_radius", "inv_speed", "atten_length")] + \
        [("edges", P * 7), ("n_edges", i32 * 7), ("ck", P * 2), ("n_ck", i32 * 2), ("prob", P * 3), ("delays", P * 3),
         ("dmax", i32 * 3)]


class LookupBars(C.Structure):
    _fields_ = [("nbars", i32), ("table", P), ("reflector", P), ("yield_", P), ("speed_ratio", P),
                ("atten_ratio", P), ("center", P), ("local_to_world", P)]


class LookupSteps(C.Structure):
    _fields_ = [("event", P), ("bar", P), ("t0", P), ("t1", P), ("pos0", P), ("pos1", P), ("edep", P), ("len", P),
                ("p0", P), ("p1", P), ("mass", P), ("charge", P), ("n", i64)]


class SipmNoise(C.Structure):
    _fields_ = [(k, f64) for k in ("thermal_rate", "trap0_prob", "trap1_prob", "trap0_lifetime", "trap1_lifetime",
                                   "t_start", "t_end")] + [("thermal_labels", P), ("n_thermal", i64)]


class Digis(C.Structure):
    _fields_ = [("label", P), ("start", P), ("tdc0", P), ("starttdc", P), ("adc", P), ("record", P), ("sample", P),
                ("capacity", i64), ("n", i64), ("record_adc", P), ("record_label", P), ("record_tick", P),
                ("record_capacity", i64), ("n_records", i64), ("overlong", i64)]


class Waveforms(C.Structure):
    _fields_ = [("ev", P), ("chan", P), ("start", P), ("nsamp", P), ("off", P), ("adc", P), ("capacity", i64),
                ("n", i64), ("sample_capacity", i64), ("n_samples", i64)]


PULSE_F64 = ("pe", "peph", "time", "height", "beta", "chi2", "letime", "ped", "fit_a", "fit_mu", "fit_b")
PULSE_U8 = ("failed", "zerondf", "pedfromdb")


class Pulses(C.Structure):
    _fields_ = [("wf", P), ("seq", P), ("ev", P), ("chan", P)] + [(k, P) for k in PULSE_F64] + \
        [(k, P) for k in PULSE_U8] + [("capacity", i64), ("n", i64)]


CLUSTER_F64 = ("cl_pe", "cl_start", "cl_end", "cl_slope", "cl_avgtime", "cl_pos", "cl_sidepe", "cl_sidetime")


class Clusters(C.Structure):
    _fields_ = [("cl_event", P), ("cl_type", P), ("cl_poscalc", P), ("cl_sidehits", P)] + \
        [(k, P) for k in CLUSTER_F64] + [("cl_mem_off", P), ("capacity", i64), ("n", i64), ("mem_idx", P),
                                         ("mem_capacity", i64), ("n_mem", i64)]


TIMING_COUNTS = ("n_steps", "n_hit_counters", "n_photons", "n_charges", "n_digis", "n_records", "n_waveforms",
                 "n_pulses", "n_clusters", "n_members")


class Timing(C.Structure):
    _fields_ = [(k, f64) for k in ("charges_ms", "digis_ms", "waveforms_ms", "pulses_ms", "clusters_ms", "total_ms",
                                   "k_sipm_ms", "k_digi_ms", "k_pulse_ms", "k_coinc_ms")] + \
        [(k, i64) for k in TIMING_COUNTS]


class BatchOpts(C.Structure):
    _fields_ = [("grid", i32), ("tile", i32), ("pulse_slots", i32), ("sort_tasks", i32), ("log_budget", i64),
                ("record_adc", P), ("record_label", P), ("record_tick", P), ("n_records", i64), ("n_waveforms", i64),
                ("overlong", i32), ("capped", i32), ("pulse_slots_used", i32), ("log_passes", i32),
                ("pipeline", i32), ("_pad", i32)]


EXPORTS = ("crv_ctx_create", "crv_ctx_destroy", "crv_set_params", "crv_last_error", "crv_upload_conditions",
           "crv_upload_template", "crv_upload_photon_map", "crv_upload_geometry", "crv_charges_run", "crv_digis_run",
           "crv_reco_waveforms_run", "crv_pulses_run", "crv_clusters_run", "crv_chain_run", "crv_chain_fetch",
           "crv_chain_records", "crv_chain_run_batched", "crv_ctx_workspace", "crv_ctx_trim",
           "crv_pulses_run_noalloc", "crv_graph_create", "crv_graph_launch", "crv_graph_fetch", "crv_graph_stats",
           "crv_graph_destroy", "crv_sipm_run", "crv_photons_run", "crv_lookup_upload", "crv_lookup_photons_run")

_lib = None


def load(path=LIB_PATH):
    """The shared library with argument types declared. Raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libcrvhip.so not built ({path}); run __graft_entry__.build()")
    lib = C.CDLL(path)
    lib.crv_ctx_create.argtypes = [C.c_int, C.POINTER(P)]
    lib.crv_ctx_destroy.argtypes = [P]
    lib.crv_set_params.argtypes = [P, C.POINTER(Params)]
    lib.crv_last_error.argtypes = [P, C.c_char_p, i64]
    lib.crv_upload_conditions.argtypes = [P, i32] + [P] * 8
    lib.crv_upload_template.argtypes = [P, P, i32]
    lib.crv_upload_photon_map.argtypes = [P, P, i32, i32, P]
    lib.crv_upload_geometry.argtypes = [P, i32, P, P, P, P, P, P, i32] + [P] * 8
    lib.crv_charges_run.argtypes = [P, C.POINTER(Steps), C.POINTER(Charges), P]
    lib.crv_sipm_run.argtypes = [P, C.POINTER(PhotonList), C.POINTER(SipmNoise), C.POINTER(Charges), P]
    lib.crv_photons_run.argtypes = [P, C.POINTER(Steps), P, f64, C.POINTER(Photons), P]
    lib.crv_lookup_upload.argtypes = [P, C.POINTER(LookupTableC), i32, C.POINTER(LookupBars)]
    lib.crv_lookup_photons_run.argtypes = [P, C.POINTER(LookupSteps), C.POINTER(Photons), P]
    lib.crv_digis_run.argtypes = [P, P, P, P, i64, C.POINTER(Digis), P]
    lib.crv_reco_waveforms_run.argtypes = [P, P, P, P, i64, C.POINTER(Waveforms), P]
    lib.crv_pulses_run.argtypes = [P] + [P] * 6 + [i64] + [P] * 4 + [C.POINTER(Pulses), P]
    lib.crv_pulses_run_noalloc.argtypes = [P] + [P] * 6 + [i64] + [P] * 4 + [C.POINTER(Pulses), P]
    lib.crv_clusters_run.argtypes = [P] + [P] * 6 + [i64, C.POINTER(Clusters), P]
    lib.crv_chain_run.argtypes = [P, C.POINTER(Steps), C.POINTER(Pulses), C.POINTER(Clusters), C.POINTER(Timing), P]
    lib.crv_chain_fetch.argtypes = [P, C.POINTER(Pulses), C.POINTER(Clusters), P]
    lib.crv_chain_records.argtypes = [P, C.POINTER(P), C.POINTER(P), C.POINTER(i64)]
    lib.crv_chain_run_batched.argtypes = [P, C.POINTER(Steps), C.POINTER(BatchOpts), C.POINTER(Pulses),
                                          C.POINTER(Clusters), C.POINTER(Timing), P]
    lib.crv_graph_create.argtypes = [P, C.POINTER(Steps), i64, C.POINTER(BatchOpts), C.c_double, C.POINTER(P)]
    lib.crv_graph_launch.argtypes = [P, C.POINTER(Steps), C.POINTER(Pulses), C.POINTER(Clusters), C.POINTER(Timing),
                                     P]
    lib.crv_graph_fetch.argtypes = [P, C.POINTER(Pulses), C.POINTER(Clusters), P]
    lib.crv_graph_stats.argtypes = [P, C.POINTER(i64), C.POINTER(i64), C.POINTER(i64)]
    lib.crv_graph_destroy.argtypes = [P]
    lib.crv_ctx_workspace.argtypes = [P, C.POINTER(C.c_uint64)]
    lib.crv_ctx_trim.argtypes = [P, C.POINTER(C.c_uint64)]
    for name in EXPORTS:
        getattr(lib, name).restype = C.c_int
    _lib = lib
    return lib


class CrvError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libcrvhip error {code}: {msg}")
        self.code = code


def params_from(cfg, seed, evkey_hi=0, precision=64):
    """crv_params of a ChainConfig (reference dataclass fields, see crvhip.config)."""
    p = Params()
    p.seed, p.evkey_hi, p.precision = int(seed) & (2 ** 64 - 1), int(evkey_hi), int(precision)
    o = cfg.optics
    for k in ("pe_per_mev", "ref_distance", "w_short", "l_short", "w_long", "l_long", "half_length", "inv_speed", "tau"):
        setattr(p, k, float(getattr(o, k)))
    p.photon_window_hi = float(cfg.window[1])
    s = cfg.sipm
    from .contract import validate_sipm, validate_window
    validate_sipm(s)
    validate_window(cfg.window)
    p.n_pixels_x, p.n_pixels_y, p.pixel_lrint = s.n_pixels_x, s.n_pixels_y, int(s.pixel_id == "lrint")
    for k in ("overvoltage", "time_constant", "capacitance", "p1", "p2", "cross_talk"):
        setattr(p, k, float(getattr(s, k)))
    w = cfg.waveform
    for k in ("single_pe_reference_charge", "single_pe_max_time", "single_pe_precision", "min_voltage", "noise",
              "adc_conversion"):
        setattr(p, k, float(getattr(w, k)))
    p.noise_on = int(bool(cfg.noise))
    p.samples_zs, p.window_samples = int(w.samples_zs), int(cfg.window_samples)
    p.feb_key_roc = int(cfg.feb_phase_key == "roc_feb")
    p.window_lo, p.window_hi, p.feb_phase_width = float(cfg.window[0]), float(cfg.window[1]), float(cfg.feb_phase_width)
    p.pedestal, p.min_adc, p.max_adc = int(w.pedestal), int(w.min_adc), int(w.max_adc)
    r = cfg.pulse_reco
    for k in ("min_adc_difference", "default_beta", "min_beta", "max_beta", "max_time_difference",
              "min_pulse_height_ratio", "max_pulse_height_ratio", "le_time_factor", "pedestal_undershoot_threshold",
              "time_offset_scale", "limit_tolerance"):
        setattr(p, k, float(getattr(r, k)))
    p.samples_before, p.samples_after = int(r.samples_before), int(r.samples_after)
    p.use_time_offset_db, p.ignore_channels = int(r.use_time_offset_db), int(r.ignore_channels)
    p.lm_iterations = int(r.lm_iterations)
    f = cfg.finder
    p.fiber_signal_speed, p.max_time_interval, p.time_offset = (float(f.fiber_signal_speed),
                                                               float(f.max_time_interval), float(f.time_offset))
    p.compensate_mask = int(sum(1 << b for b in f.compensate_status_bits))
    p.ignore_secondary_pulses = int(f.ignore_secondary_pulses)
    p.big_cluster_threshold = int(f.big_cluster_threshold)
    p.add_opposite_side_pulses = int(f.add_opposite_side_pulses)
    return p


def ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


class Context:
    """A libcrvhip context on one device with the conditions, template, photon map and finder geometry uploaded."""

    def __init__(self, cfg, cond, geom, tmpl, device=0):
        import torch
        self.torch = torch
        self.lib = load()
        self.device = device
        self.h = C.c_void_p()
        rc = self.lib.crv_ctx_create(int(device), C.byref(self.h))
        if rc:
            buf = C.create_string_buffer(512)
            self.lib.crv_last_error(None, buf, 512)
            raise CrvError(rc, buf.value.decode())
        self.cfg = cfg
        self._pmap = None
        self.set_params(cfg, seed=0)
        self.upload_conditions(cond, cfg)
        self.upload_template(tmpl)
        self.upload_photon_map(cfg)
        if geom is not None:
            self.upload_geometry(geom, cfg.finder)

    def close(self):
        if self.h:
            self.lib.crv_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        if rc:
            buf = C.create_string_buffer(1024)
            self.lib.crv_last_error(self.h, buf, 1024)
            raise CrvError(rc, buf.value.decode())

    def workspace_bytes(self):
        b = C.c_uint64(0)
        self.check(self.lib.crv_ctx_workspace(self.h, C.byref(b)))
        return int(b.value)

    def trim(self):
        """Return the context's workspace to the device; returns the bytes freed."""
        b = C.c_uint64(0)
        self.check(self.lib.crv_ctx_trim(self.h, C.byref(b)))
        return int(b.value)

    def set_params(self, cfg, seed, evkey_hi=0, precision=64):
        self.params = params_from(cfg, seed, evkey_hi, precision)
        self.check(self.lib.crv_set_params(self.h, C.byref(self.params)))

    @staticmethod
    def _h(a, dt):
        return np.ascontiguousarray(np.asarray(a), dtype=dt)

    def upload_conditions(self, cond, cfg):
        from .detector import effective_status
        status = self._h(effective_status(cond), np.int32)
        n = len(status)
        dev = self._h((np.asarray(cond.deviation, float) * cfg.response.photon_yield_variation_scale + 1.0)
                      * cfg.response.photon_yield_scale, np.float64)
        arrs = [status, self._h(cond.roc, np.int32), self._h(cond.feb, np.int32), dev,
                self._h(cond.pedestal, np.float64), self._h(cond.pulsearea, np.float64),
                self._h(cond.pulseheight, np.float64), self._h(cond.timeoffset, np.float64)]
        self.check(self.lib.crv_upload_conditions(self.h, n, *[a.ctypes.data_as(P) for a in arrs]))
        self._cond_arrays = arrs

    def upload_template(self, tmpl):
        t = self._h(tmpl, np.float64)
        self.check(self.lib.crv_upload_template(self.h, t.ctypes.data_as(P), len(t)))

    def use_photon_map(self, cfg, pmap=None):
        """Make `pmap` (or cfg.photon_map when None) the context's photon map; uploads only on a change."""
        pm = cfg.photon_map if pmap is None else pmap
        if self._pmap is not pm:
            a = np.asarray(pm, float)
            if a.ndim != 2 or a.shape[0] < 1 or a.shape[1] < 1 or not np.isfinite(a).all() or (a < 0).any() \
                    or a.sum() <= 0:
                raise ValueError("pmap: the photon map must be a 2-d array of finite, non-negative weights")
            self.upload_photon_map(cfg, pm)

    def upload_photon_map(self, cfg, pmap=None):
        pm = self._h(cfg.photon_map if pmap is None else pmap, np.float64)
        self._pmap = cfg.photon_map if pmap is None else pmap
        dead = np.zeros((41, 41), np.uint8)
        for x, y in cfg.sipm.inactive:
            dead[x, y] = 1
        self.check(self.lib.crv_upload_photon_map(self.h, pm.ctypes.data_as(P), pm.shape[0], pm.shape[1],
                                                  dead.ctypes.data_as(P)))

    def upload_geometry(self, G, fcfg):
        nb = len(G.sector)
        ns = len(G.sector_names)
        cuts = np.zeros((ns, 10))
        has = np.zeros(ns, np.uint8)
        fields = ("pe_threshold", "max_dt_adjacent", "max_dt", "min_slope", "max_slope", "max_slope_diff",
                  "coincidence_layers", "min_cluster_pes", "initial_max_distance", "final_max_distance")
        for i, name in enumerate(G.sector_names):
            if name in fcfg.cuts:
                has[i] = 1
                cuts[i] = [float(getattr(fcfg.cuts[name], f)) for f in fields]
        arrs = [self._h(G.sector, np.int32), self._h(G.layer, np.int32), self._h(G.counter_in_type, np.int32),
                self._h(G.x, np.float64), self._h(G.y, np.float64), self._h(G.pos, np.float64)]
        sarrs = [self._h(G.sector_type, np.int32), self._h(G.length_dir, np.int32), self._h(G.pos_side0, np.float64),
                 self._h(G.pos_side1, np.float64), self._h(G.sipms_side0, np.uint8), self._h(G.sipms_side1, np.uint8),
                 has, self._h(cuts, np.float64)]
        self.check(self.lib.crv_upload_geometry(self.h, nb, *[a.ctypes.data_as(P) for a in arrs], ns,
                                                *[a.ctypes.data_as(P) for a in sarrs]))


class Graph:
    """A captured FP32 batched chain (crv_graph_create) on a context, for calls of at most max_steps steps."""

    def __init__(self, ctx, sample, max_steps, opts, margin=0.25):
        self.ctx = ctx
        self.h = C.c_void_p()
        ctx.check(ctx.lib.crv_graph_create(ctx.h, C.byref(sample), int(max_steps), C.byref(opts), float(margin),
                                           C.byref(self.h)))

    def launch(self, steps, pulses, clusters, timing=None, stream=None):
        """crv_graph_launch; returns its status (0 or CRV_E_CAPACITY with the sizes in the outputs' n)."""
        rc = self.ctx.lib.crv_graph_launch(self.h, C.byref(steps), C.byref(pulses), C.byref(clusters),
                                           C.byref(timing) if timing is not None else None, stream)
        if rc not in (CRV_OK, CRV_E_CAPACITY):
            self.ctx.check(rc)
        return rc

    def fetch(self, pulses, clusters, stream=None):
        self.ctx.check(self.ctx.lib.crv_graph_fetch(self.h, C.byref(pulses), C.byref(clusters), stream))

    def stats(self):
        la, fb = i64(0), i64(0)
        caps = (i64 * 7)()
        self.ctx.check(self.ctx.lib.crv_graph_stats(self.h, C.byref(la), C.byref(fb), caps))
        names = ("steps", "log_chunks", "records", "runs", "pulses", "pulse_slots", "last_fell_back")
        return {"launches": la.value, "fallbacks": fb.value, "capacities": dict(zip(names, list(caps)))}

    def close(self):
        if self.h:
            self.ctx.lib.crv_graph_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

