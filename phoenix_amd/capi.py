"""ctypes binding of libphx.so (include/phx.h).

The raymarch runs only through the HIP library: loading fails loudly when
libphx.so is missing, and a render fails with the library's own error when no
GPU is present.  There is no CPU path here.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PHX_LIB", os.path.join(HERE, "libphx.so"))

MAX_FREQS = 32
MAX_LEVELS = 16
FP32, FP64, MIXED = 32, 64, 96
TABLE_FP32, TABLE_FP64 = 1, 2

c_double, c_int, c_int64, c_char = ctypes.c_double, ctypes.c_int32, ctypes.c_int64, ctypes.c_char
P_dbl, P_int = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32)


class FreqOut(ctypes.Structure):
    _fields_ = [("IQUV", c_double * 4), ("tau", c_double), ("tauF", c_double)]


class Info(ctypes.Structure):
    _fields_ = [("device", c_char * 256), ("cus", c_int), ("blocks", c_int), ("block", c_int),
                ("vgprs", c_int), ("sgprs", c_int), ("scratch_bytes", c_int), ("lds_bytes", c_int),
                ("waves_per_cu", c_int), ("upload_s", c_double), ("kernel_s", c_double),
                ("decode_s", c_double), ("steps", c_int64), ("nplasma", c_int64), ("bad", c_int64),
                ("rays", c_int64), ("fp64_steps", c_int64), ("launches", c_int), ("reruns", c_int)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["device"] = self.device.decode(errors="replace")
        return d


class Model(ctypes.Structure):
    _fields_ = [
        ("inclination", c_double), ("fov_x", c_double), ("fov_y", c_double), ("r_cam", c_double),
        ("t_init", c_double), ("pa", c_double), ("phi_cam", c_double),
        ("img_width", c_int), ("img_height", c_int), ("nfreqs", c_int), ("pad0", c_int),
        ("frequencies", c_double * MAX_FREQS),
        ("mass", c_double), ("distance", c_double),
        ("stepsize", c_double), ("horizon_margin", c_double),
        ("max_steps", c_int), ("max_order", c_int),
        ("m_unit", c_double), ("r_low", c_double), ("r_high", c_double), ("beta_c", c_double),
        ("sigma_cut", c_double), ("thetae_min", c_double), ("thetae_max", c_double), ("rt_outer", c_double),
        ("nspin", c_int), ("bpol", c_int), ("nlevels", c_int), ("sfc", c_int),
        ("nxlone", c_int * 3), ("nx", c_int * 3), ("ng", c_int * 3),
        ("nleafs", c_int), ("ndim", c_int), ("nwini", c_int), ("nws", c_int), ("cells", c_int),
        ("xprobmin", c_double * 3), ("xprobmax", c_double * 3), ("hslope_grid", c_double),
        ("a", c_double), ("hslope", c_double), ("gam", c_double), ("Ne", c_double), ("Bu", c_double),
        ("rhigh", c_double), ("rlow", c_double), ("ibc", c_double), ("r_out", c_double), ("cut_in", c_double),
        ("cos_pa", c_double), ("sin_pa", c_double),
        ("L", c_double), ("T_unit", c_double), ("rho_unit", c_double), ("U_unit", c_double),
        ("B_unit", c_double), ("Ne_unit", c_double),
        ("freq", c_double * MAX_FREQS), ("hconst", c_double * MAX_FREQS), ("jscale", c_double * MAX_FREQS),
        ("lvl_w", (c_double * 3) * MAX_LEVELS), ("x_lo", c_double * 4), ("x_hi", c_double * 4),
        ("level_dims", (c_int * 3) * MAX_LEVELS),
        ("lb", P_dbl), ("dxc", P_dbl), ("prim", P_dbl),
        ("level_map", P_int * MAX_LEVELS), ("leaf_level", P_int), ("leaf_index", P_int),
    ]


class Config(ctypes.Structure):
    _fields_ = [("outdir", c_char * 1024), ("image", c_char * 16), ("block_size", c_int),
                ("dump_file", c_char * 1024), ("grid_file", c_char * 1024), ("gpu_decode", c_int),
                ("show_header", c_int)]


class Tune(ctypes.Structure):
    _fields_ = [("queue", c_int), ("blocks_per_cu", c_int), ("order", c_int), ("fp64_wave", c_int),
                ("refill", c_int), ("escape", c_int), ("block", c_int), ("fetch", c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerOpts(ctypes.Structure):
    _fields_ = [("precision", c_int), ("quantum_jobs", c_int), ("hbm_budget", c_int64), ("max_active", c_int),
                ("order", c_int)]


class ServeStats(ctypes.Structure):
    _fields_ = [("submitted", c_int64), ("completed", c_int64), ("failed", c_int64), ("active", c_int64),
                ("waiting", c_int64), ("rays", c_int64), ("launches", c_int64), ("peak_bytes", c_int64),
                ("busy_s", c_double), ("wall_s", c_double), ("p50_s", c_double), ("p99_s", c_double),
                ("mean_s", c_double), ("max_s", c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# device::Progress (src/device/kernel.cuh:32): jobs done in the chunk, user pointer
PROGRESS = ctypes.CFUNCTYPE(None, ctypes.c_int32, ctypes.c_void_p)


class PhxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"phx error {code}: {msg}")
        self.code = code


_lib = None

# every entry point of include/phx.h, with its ctypes signature
SIGNATURES = {
    "phx_last_error": (ctypes.c_char_p, []),
    "phx_version": (ctypes.c_char_p, []),
    "phx_load_config": (c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p), c_int, ctypes.POINTER(Config),
                                ctypes.POINTER(Model)]),
    "phx_load_grid": (c_int, [ctypes.c_char_p, ctypes.POINTER(Model)]),
    "phx_load_dump": (c_int, [ctypes.POINTER(Model), ctypes.c_char_p]),
    "phx_model_derive": (c_int, [ctypes.POINTER(Model)]),
    "phx_free_model": (None, [ctypes.POINTER(Model)]),
    "phx_schedule": (c_int, [ctypes.POINTER(Model), c_int, P_int, c_int]),
    "phx_compute_spec": (c_int, [ctypes.POINTER(Model), ctypes.POINTER(FreqOut), c_int, P_dbl]),
    "phx_compute_tau": (c_int, [ctypes.POINTER(Model), ctypes.POINTER(FreqOut), c_int, P_dbl]),
    "phx_write_image_text": (c_int, [ctypes.POINTER(Model), ctypes.POINTER(FreqOut), c_int, ctypes.c_char_p]),
    "phx_write_spectrum": (c_int, [ctypes.POINTER(Model), ctypes.POINTER(FreqOut), ctypes.c_char_p]),
    "phx_write_image_hdf5": (c_int, [ctypes.POINTER(Model), ctypes.POINTER(FreqOut), c_int, ctypes.c_char_p]),
    "phx_write_output": (c_int, [ctypes.POINTER(Model), ctypes.POINTER(Config), ctypes.POINTER(FreqOut),
                                 ctypes.c_char_p, c_int]),
    "phx_set_device": (c_int, [c_int]),
    "phx_dev_warmup": (c_int, []),
    "phx_dev_render": (c_int, [ctypes.POINTER(Model), ctypes.c_char_p, c_int, P_int, c_int, ctypes.POINTER(FreqOut),
                               ctypes.POINTER(Info), ctypes.c_void_p, ctypes.c_void_p]),
    "phx_snapshot_open": (c_int, [ctypes.POINTER(Model), ctypes.c_char_p, c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "phx_snapshot_close": (c_int, [ctypes.c_void_p]),
    "phx_snapshot_bytes": (c_int64, [ctypes.c_void_p]),
    "phx_snapshot_render": (c_int, [ctypes.c_void_p, ctypes.POINTER(Model), c_int, P_int, c_int,
                                    ctypes.POINTER(FreqOut), ctypes.POINTER(Info)]),
    "phx_snapshot_render_batch": (c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.POINTER(Model)), c_int, c_int,
                                          ctypes.POINTER(ctypes.POINTER(FreqOut)), ctypes.POINTER(Info)]),
    "phx_autotune": (c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.POINTER(Model)), c_int, c_int, c_int,
                             ctypes.POINTER(Tune), ctypes.POINTER(c_double)]),
    "phx_snapshot_set_tune": (c_int, [ctypes.c_void_p, ctypes.POINTER(Tune)]),
    "phx_server_start": (c_int, [ctypes.POINTER(ctypes.c_void_p), c_int, ctypes.POINTER(ServerOpts),
                                 ctypes.POINTER(ctypes.c_void_p)]),
    "phx_server_submit": (c_int, [ctypes.c_void_p, c_int, ctypes.POINTER(Model), ctypes.POINTER(FreqOut),
                                  ctypes.POINTER(c_int64)]),
    "phx_server_wait": (c_int, [ctypes.c_void_p, c_int64, c_double, P_int, ctypes.POINTER(c_double)]),
    "phx_server_stats": (c_int, [ctypes.c_void_p, ctypes.POINTER(ServeStats)]),
    "phx_server_reset_stats": (c_int, [ctypes.c_void_p]),
    "phx_server_stop": (c_int, [ctypes.c_void_p]),
    "phx_snapshot_result": (c_int, [ctypes.c_void_p, ctypes.POINTER(FreqOut), c_int64]),
    "phx_snapshot_flux": (c_int, [ctypes.c_void_p, P_dbl, c_int]),
    "phx_server_submit_spec": (c_int, [ctypes.c_void_p, c_int, ctypes.POINTER(Model), ctypes.POINTER(FreqOut), P_dbl,
                                       ctypes.POINTER(c_int64)]),
    "phx_snapshot_render_noalloc": (c_int, [ctypes.c_void_p, ctypes.POINTER(Model), c_int, ctypes.c_void_p,
                                            ctypes.POINTER(Info)]),
    "phx_snapshot_prims": (c_int, [ctypes.c_void_p, c_int, P_dbl, c_int64]),
    "phx_graph_create": (c_int, [ctypes.c_void_p, ctypes.POINTER(Model), c_int, ctypes.c_void_p,
                                 ctypes.POINTER(ctypes.c_void_p)]),
    "phx_graph_launch": (c_int, [ctypes.c_void_p, ctypes.POINTER(Model), ctypes.POINTER(Info)]),
    "phx_graph_destroy": (c_int, [ctypes.c_void_p]),
}


def lib():
    """The loaded libphx.so; raises if it was not built (there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m phoenix_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise PhxError(rc, lib().phx_last_error().decode(errors="replace"))
    return rc


def _res_array(n):
    arr = np.zeros((n, 6), dtype=np.float64)
    return arr, arr.ctypes.data_as(ctypes.POINTER(FreqOut))


class Run:
    """config.toml + grid + dump loaded by the library's host layer (phx_load_*)."""

    def __init__(self, config_path: str, overrides=()):
        self.model = Model()
        self.config = Config()
        ov = [o.encode() for o in overrides]
        arr = (ctypes.c_char_p * max(1, len(ov)))(*ov)
        check(lib().phx_load_config(config_path.encode(), arr, len(ov), ctypes.byref(self.config),
                                    ctypes.byref(self.model)))
        check(lib().phx_load_grid(self.config.grid_file, ctypes.byref(self.model)))
        check(lib().phx_load_dump(ctypes.byref(self.model), self.config.dump_file))
        self._loaded = True

    @property
    def dump(self) -> str:
        return self.config.dump_file.decode()

    def close(self):
        if getattr(self, "_loaded", False):
            lib().phx_free_model(ctypes.byref(self.model))
            self._loaded = False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def request(self, **kw) -> Model:
        """A copy of the model with user fields replaced and per-request scalars re-derived."""
        m = Model()
        ctypes.pointer(m)[0] = self.model
        for k, v in kw.items():
            if k == "freqs":
                m.nfreqs = len(v)
                for i, f in enumerate(v):
                    m.frequencies[i] = f
            else:
                setattr(m, k, v)
        check(lib().phx_model_derive(ctypes.byref(m)))
        return m

    def schedule(self, morton: bool = True, model: Model | None = None) -> np.ndarray:
        m = model or self.model
        n = m.img_width * m.img_height
        jobs = np.zeros(n, dtype=np.int32)
        check(lib().phx_schedule(ctypes.byref(m), int(morton), jobs.ctypes.data_as(P_int), n))
        return jobs

    def dev_render(self, precision: int = FP32, model: Model | None = None, jobs=None, progress=None):
        """phx_dev_render: the drop-in one-shot device path.  Returns (res [W*H, nf, 6], info).
        progress: called with the jobs done after each chunk (device::Progress)."""
        m = model or self.model
        if jobs is None:
            jobs = self.schedule(True, m)
        jobs = np.ascontiguousarray(jobs, dtype=np.int32)
        n = m.img_width * m.img_height * m.nfreqs
        res, rp = _res_array(n)
        info = Info()
        cb = PROGRESS(lambda done, user: progress(done)) if progress else None
        check(lib().phx_dev_render(ctypes.byref(m), self.config.dump_file, precision, jobs.ctypes.data_as(P_int),
                                   jobs.size, rp, ctypes.byref(info), ctypes.cast(cb, ctypes.c_void_p) if cb else None,
                                   None))
        return res.reshape(m.img_width * m.img_height, m.nfreqs, 6), info.as_dict()

    def write_output(self, res: np.ndarray, outdir: str | None = None, image: str | None = None,
                     block_size: int | None = None, model: Model | None = None) -> str:
        """phx_write_output: the reference's write_output (HDF5 and/or text image, spectrum).
        Returns the messages it would print."""
        m = model or self.model
        cfg = Config()
        ctypes.pointer(cfg)[0] = self.config
        if outdir is not None:
            cfg.outdir = outdir.encode()
        if image is not None:
            cfg.image = image.encode()
        if block_size is not None:
            cfg.block_size = block_size
        r = np.ascontiguousarray(res.reshape(-1, 6), dtype=np.float64)
        log = ctypes.create_string_buffer(8192)
        check(lib().phx_write_output(ctypes.byref(m), ctypes.byref(cfg), r.ctypes.data_as(ctypes.POINTER(FreqOut)),
                                     log, len(log)))
        return log.value.decode()

    def flux(self, res: np.ndarray, model: Model | None = None) -> np.ndarray:
        m = model or self.model
        r = np.ascontiguousarray(res.reshape(-1, 6))
        out = np.zeros((m.nfreqs, 4))
        for f in range(m.nfreqs):
            S = np.zeros(4)
            check(lib().phx_compute_spec(ctypes.byref(m), r.ctypes.data_as(ctypes.POINTER(FreqOut)), f,
                                         S.ctypes.data_as(P_dbl)))
            out[f] = S
        return out


class Snapshot:
    """A snapshot resident in HBM (phx_snapshot_*)."""

    def __init__(self, run: Run, tables: int = TABLE_FP32 | TABLE_FP64):
        self.run = run
        self.h = ctypes.c_void_p()
        check(lib().phx_snapshot_open(ctypes.byref(run.model), run.config.dump_file, tables, ctypes.byref(self.h)))
        self._tune = {"queue": -1, "blocks_per_cu": 0, "order": 0, "fp64_wave": 1, "refill": 0, "escape": 0,
                      "block": 0, "fetch": 0}

    def close(self):
        if self.h:
            check(lib().phx_snapshot_close(self.h))
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def bytes(self) -> int:
        return int(lib().phx_snapshot_bytes(self.h))

    def render(self, request: Model | None = None, precision: int = FP32, jobs=None, to_host: bool = True,
               plain: bool = False, tune: dict | None = None):
        """One render on the resident snapshot.  plain: the reference kernel's way, each lane marching
        its pixel with the samples processed in place, the wave taking new pixels when all its
        lanes are done and every ray marched to r_out (queue 0, refill 64, escape 1).  tune: launch
        settings for this render only."""
        if plain or tune:
            saved = dict(self._tune)
            self.set_tune(**dict(saved, **(dict(queue=0, refill=64, escape=1) if plain else {}), **(tune or {})))
            try:
                return self.render(request, precision, jobs, to_host)
            finally:
                self.set_tune(**saved)
        m = request or self.run.model
        n = m.img_width * m.img_height * m.nfreqs
        jp, nj = None, 0
        if jobs is not None:
            jobs = np.ascontiguousarray(jobs, dtype=np.int32)
            jp, nj = jobs.ctypes.data_as(P_int), jobs.size
        info = Info()
        if to_host:
            res, rp = _res_array(n)
        else:
            res, rp = None, None
        check(lib().phx_snapshot_render(self.h, ctypes.byref(m), precision, jp, nj, rp, ctypes.byref(info)))
        if res is not None:
            res = res.reshape(m.img_width * m.img_height, m.nfreqs, 6)
        return res, info.as_dict()

    def render_batch(self, requests, precision: int = FP32, to_host: bool = True):
        """phx_snapshot_render_batch: one launch over every pixel of every request.
        Returns ([res_i [W*H, nf, 6] or None], info)."""
        n = len(requests)
        rq = (ctypes.POINTER(Model) * n)(*[ctypes.pointer(m) for m in requests])
        outs, ptrs = [], (ctypes.POINTER(FreqOut) * n)()
        for i, m in enumerate(requests):
            if to_host:
                arr, rp = _res_array(m.img_width * m.img_height * m.nfreqs)
                outs.append(arr)
                ptrs[i] = rp
            else:
                outs.append(None)
        info = Info()
        check(lib().phx_snapshot_render_batch(self.h, rq, n, precision, ptrs if to_host else None,
                                              ctypes.byref(info)))
        res = [None if a is None else a.reshape(m.img_width * m.img_height, m.nfreqs, 6)
               for a, m in zip(outs, requests)]
        return res, info.as_dict()

    def autotune(self, requests, precision: int = FP32, sample_stride: int = 8):
        """phx_autotune: returns (tune dict, sampled rays*freq/s); the choice stays with the snapshot."""
        n = len(requests)
        rq = (ctypes.POINTER(Model) * n)(*[ctypes.pointer(m) for m in requests])
        t, rate = Tune(), c_double()
        check(lib().phx_autotune(self.h, rq, n, precision, sample_stride, ctypes.byref(t), ctypes.byref(rate)))
        self._tune = t.as_dict()
        return t.as_dict(), rate.value

    def set_tune(self, queue: int = -1, blocks_per_cu: int = 0, order: int = 0, fp64_wave: int = 1, refill: int = 0,
                 escape: int = 0, block: int = 0, fetch: int = 0):
        t = Tune(queue, blocks_per_cu, order, fp64_wave, refill, escape, block, fetch)
        check(lib().phx_snapshot_set_tune(self.h, ctypes.byref(t)))
        self._tune = t.as_dict()

    def flux(self, nfreqs: int) -> np.ndarray:
        """phx_snapshot_flux: [nfreqs, 4] Jy of the last render, reduced on the device."""
        spec = np.zeros((nfreqs, 4))
        check(lib().phx_snapshot_flux(self.h, spec.ctypes.data_as(P_dbl), nfreqs))
        return spec

    def render_noalloc(self, request: Model, precision: int, d_res: int) -> dict:
        """phx_snapshot_render_noalloc into the device buffer at address d_res; returns the info."""
        info = Info()
        check(lib().phx_snapshot_render_noalloc(self.h, ctypes.byref(request), precision, ctypes.c_void_p(d_res),
                                                ctypes.byref(info)))
        return info.as_dict()

    def graph(self, request: Model, precision: int, d_res: int) -> "Graph":
        """phx_graph_create: the allocation-free render of request captured as a HIP graph, results
        into the device buffer at address d_res."""
        return Graph(self, request, precision, d_res)

    def prims(self, table: int) -> np.ndarray:
        m = self.run.model
        n = m.nleafs * m.cells * 8
        out = np.zeros(n)
        check(lib().phx_snapshot_prims(self.h, table, out.ctypes.data_as(P_dbl), n))
        return out.reshape(m.nleafs, m.cells, 8)


class Graph:
    """phx_graph_*: a captured render replayed for parameter-sweep points of one camera."""

    def __init__(self, snap: Snapshot, request: Model, precision: int, d_res: int):
        self.snap = snap
        self.h = ctypes.c_void_p()
        check(lib().phx_graph_create(snap.h, ctypes.byref(request), precision, ctypes.c_void_p(d_res),
                                     ctypes.byref(self.h)))

    def launch(self, request: Model | None = None) -> dict:
        """phx_graph_launch: replays the graph (for request, a new sweep point, when given); returns the info."""
        info = Info()
        check(lib().phx_graph_launch(self.h, ctypes.byref(request) if request is not None else None,
                                     ctypes.byref(info)))
        return info.as_dict()

    def close(self):
        if self.h:
            check(lib().phx_graph_destroy(self.h))
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """phx_server_*: continuous-batching render service over resident snapshots."""

    def __init__(self, snapshots, precision: int = FP32, quantum_jobs: int = 0, hbm_budget: int = 0,
                 max_active: int = 0, order: int = 0):
        arr = (ctypes.c_void_p * len(snapshots))(*[sn.h.value for sn in snapshots])
        opts = ServerOpts(precision, quantum_jobs, hbm_budget, max_active, order)
        self.h = ctypes.c_void_p()
        check(lib().phx_server_start(arr, len(snapshots), ctypes.byref(opts), ctypes.byref(self.h)))
        self._keep = {}

    def submit(self, snap: int, request: Model, image: bool = True, spec: bool = False) -> int:
        """Queues a request; returns its ticket.  image: the result comes back to the host; spec: the
        flux per frequency, reduced on the device (phx_server_submit_spec).  The arrays stay alive
        until wait()."""
        n = request.img_width * request.img_height * request.nfreqs
        arr, rp = _res_array(n) if image else (None, None)
        sp = np.zeros((request.nfreqs, 4)) if spec else None
        req = Model()
        ctypes.pointer(req)[0] = request
        t = c_int64()
        check(lib().phx_server_submit_spec(self.h, snap, ctypes.byref(req), rp,
                                           sp.ctypes.data_as(P_dbl) if spec else None, ctypes.byref(t)))
        self._keep[t.value] = (arr, req, sp)
        return t.value

    def wait(self, ticket: int, timeout: float = -1.0):
        """(result [W*H, nf, 6] or None, latency_s) once done, or (None, 0.0) on timeout."""
        res, _, lat = self.wait_spec(ticket, timeout)
        return res, lat

    def wait_spec(self, ticket: int, timeout: float = -1.0):
        """(result or None, spectrum [nf, 4] or None, latency_s) once done, or (None, None, 0.0)."""
        done, lat = c_int(), c_double()
        check(lib().phx_server_wait(self.h, ticket, timeout, ctypes.byref(done), ctypes.byref(lat)))
        if not done.value:
            return None, None, 0.0
        arr, req, sp = self._keep.pop(ticket)
        res = None if arr is None else arr.reshape(req.img_width * req.img_height, req.nfreqs, 6)
        return res, sp, lat.value

    def stats(self) -> dict:
        st = ServeStats()
        check(lib().phx_server_stats(self.h, ctypes.byref(st)))
        return st.as_dict()

    def reset_stats(self):
        check(lib().phx_server_reset_stats(self.h))

    def stop(self):
        if self.h:
            check(lib().phx_server_stop(self.h))
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass
