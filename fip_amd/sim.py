"""ctypes binding of libfiphip (include/fip.h). No fallback: if the HIP library is
missing or a call fails, this raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .workloads import BC_SIDES, BC_TYPES, Case

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libfiphip.so")

FIELDS = {"h": 0, "qx": 1, "qy": 2, "hmax": 3, "bed": 4, "kst": 5}
EV_NONE, EV_FINISHED, EV_DUMP = 0, 1, 2
EXPORTS = ["fip_last_error", "fip_create", "fip_destroy", "fip_bc_generate", "fip_bc_source",
           "fip_bc_launch",
           "fip_bc_apply", "fip_bc_params", "fip_control_timestep", "fip_compute_step",
           "fip_floodplain", "fip_set_run", "fip_simulate", "fip_prepare", "fip_read_samples", "fip_get_time",
           "fip_set_dt", "fip_download", "fip_upload", "fip_tiles", "fip_get_masks",
           "fip_get_cfl", "fip_set_launch", "fip_tune", "fip_set_timing", "fip_get_stats",
           "fip_sync", "fip_check_rcp"]
CLOSURE_EXPORTS = ["fip_restrict", "fip_closure_create", "fip_closure_destroy", "fip_closure_get_params",
                   "fip_closure_set_params", "fip_closure_add_sample", "fip_closure_num_samples",
                   "fip_closure_train", "fip_closure_adam_step", "fip_closure_eval", "fip_closure_attach",
                   "fip_closure_attach_ensemble", "fip_closure_detach", "fip_closure_status",
                   "fip_closure_forward", "fip_closure_forward_noalloc", "fip_closure_backward", "fip_closure_fused_outputs",
                   "fip_closure_set_kernel", "fip_closure_set_activation", "fip_closure_get_activation",
                   "fip_mfma_probe"]
CL_PARAMS = 4 * (9 * 32 * 32 + 32)
ENS_EXPORTS = ["fip_ens_create", "fip_ens_destroy", "fip_ens_set_bcs", "fip_ens_set_run",
               "fip_ens_simulate", "fip_ens_prepare", "fip_ens_get_time", "fip_ens_download", "fip_ens_set_launch",
               "fip_ens_tune",
               "fip_ens_set_timing", "fip_ens_get_stats"]


class FipDesc(C.Structure):
    _fields_ = [("W", C.c_int), ("H", C.c_int), ("dx", C.c_float), ("dt", C.c_float),
                ("no_data", C.c_float), ("kst", C.c_float), ("bed", C.c_void_p),
                ("h", C.c_void_p), ("qx", C.c_void_p), ("qy", C.c_void_p),
                ("kst_var", C.c_void_p), ("variable_dt", C.c_int), ("staggered", C.c_int),
                ("mode", C.c_int), ("variant", C.c_int), ("tile_rows", C.c_int),
                ("device", C.c_int)]


class FipBC(C.Structure):
    _fields_ = [("type", C.c_int), ("side", C.c_int), ("from_", C.c_int), ("to", C.c_int),
                ("z", C.c_float), ("q", C.c_float), ("zt", C.c_void_p), ("zv", C.c_void_p),
                ("zn", C.c_int), ("qt", C.c_void_p), ("qv", C.c_void_p), ("qn", C.c_int)]


class FipStats(C.Structure):
    _fields_ = [("steps", C.c_longlong), ("tiles_total", C.c_longlong),
                ("tiles_computed", C.c_longlong), ("stencil_launches", C.c_longlong),
                ("stencil_ms", C.c_double), ("timed_launches", C.c_longlong),
                ("kst_classes", C.c_int), ("dense_span", C.c_int)]


class FipTune(C.Structure):
    _fields_ = [("tile_rows", C.c_int), ("waves_per_block", C.c_int), ("blocks_per_cu", C.c_int),
                ("stencil_ms", C.c_double), ("tried", C.c_int)]


class FipError(RuntimeError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FipError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.fip_last_error.restype = C.c_char_p
        vp = C.c_void_p
        L.fip_create.argtypes = [C.POINTER(FipDesc), C.POINTER(vp)]
        L.fip_destroy.argtypes = [vp]
        L.fip_bc_generate.argtypes = [vp, C.POINTER(FipBC), C.c_int]
        L.fip_bc_source.argtypes = [C.POINTER(FipBC), C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]
        L.fip_bc_launch.argtypes = [vp, C.POINTER(C.c_float)]
        L.fip_bc_apply.argtypes = [vp]
        L.fip_bc_params.argtypes = [vp, C.POINTER(C.c_float), C.c_int]
        for nm in ("fip_control_timestep", "fip_compute_step", "fip_floodplain", "fip_sync"):
            getattr(L, nm).argtypes = [vp]
        ip = C.POINTER(C.c_int)
        L.fip_set_run.argtypes = [vp, C.c_double, C.c_int, ip, ip, ip, C.c_int,
                                  C.POINTER(C.c_double), C.c_int]
        L.fip_simulate.argtypes = [vp, C.c_longlong, ip, C.POINTER(C.c_longlong)]
        L.fip_prepare.argtypes = [vp, C.c_longlong]
        L.fip_read_samples.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_float), C.c_int, ip]
        L.fip_get_time.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_float), ip]
        L.fip_set_dt.argtypes = [vp, C.c_float]
        L.fip_download.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_float)]
        L.fip_upload.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_float)]
        L.fip_tiles.argtypes = [vp, ip, ip, ip, ip]
        L.fip_get_masks.argtypes = [vp, C.POINTER(C.c_uint64), C.c_int]
        L.fip_get_cfl.argtypes = [vp, C.POINTER(C.c_float)]
        L.fip_set_timing.argtypes = [vp, C.c_int]
        L.fip_get_stats.argtypes = [vp, C.POINTER(FipStats)]
        L.fip_ens_create.argtypes = [C.POINTER(FipDesc), C.c_int, C.POINTER(vp)]
        L.fip_ens_destroy.argtypes = [vp]
        L.fip_ens_set_bcs.argtypes = [vp, C.c_int, C.POINTER(FipBC), C.c_int]
        L.fip_ens_set_run.argtypes = [vp, C.POINTER(C.c_double)]
        L.fip_ens_simulate.argtypes = [vp, C.c_longlong, ip, C.POINTER(C.c_longlong)]
        L.fip_ens_prepare.argtypes = [vp, C.c_longlong]
        L.fip_ens_get_time.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_float), ip]
        L.fip_ens_download.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float)]
        L.fip_ens_set_launch.argtypes = [vp, C.c_int, C.c_int]
        L.fip_ens_tune.argtypes = [C.POINTER(FipDesc), C.c_int, C.POINTER(C.POINTER(FipBC)),
                                   ip, C.c_int, C.POINTER(FipTune)]
        L.fip_set_launch.argtypes = [vp, C.c_int, C.c_int]
        fp, dp = C.POINTER(C.c_float), C.POINTER(C.c_double)
        L.fip_restrict.argtypes = [vp, vp, C.c_int]
        L.fip_closure_create.argtypes = [C.c_int, C.c_ulonglong, C.c_int, C.c_int, C.c_int, C.POINTER(vp)]
        L.fip_closure_destroy.argtypes = [vp]
        L.fip_closure_get_params.argtypes = [vp, fp]
        L.fip_closure_set_params.argtypes = [vp, fp]
        L.fip_closure_add_sample.argtypes = [vp, vp, vp, C.c_int, fp, C.c_float]
        L.fip_closure_num_samples.argtypes = [vp, ip]
        L.fip_closure_train.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int,
                                        C.c_ulonglong, fp]
        L.fip_closure_adam_step.argtypes = [vp, fp, C.c_float]
        L.fip_closure_eval.argtypes = [vp, C.c_int, C.c_int, dp, dp]
        L.fip_closure_attach.argtypes = [vp, vp, fp, C.c_float, C.c_float, C.c_double]
        L.fip_closure_attach_ensemble.argtypes = [vp, vp, fp, fp, C.c_float, C.c_double]
        L.fip_closure_detach.argtypes = [vp]
        L.fip_closure_status.argtypes = [vp, vp, vp, C.c_int, ip, C.POINTER(C.c_uint), dp]
        L.fip_closure_forward.argtypes = [vp, fp, C.c_int, C.c_int, C.c_int, fp]
        L.fip_closure_forward_noalloc.argtypes = [vp, vp, C.c_int, C.c_int, C.c_int, vp]
        L.fip_closure_backward.argtypes = [vp, fp, C.c_int, C.c_int, C.c_int, fp, fp,
                                           C.POINTER(C.c_ubyte), C.c_float, fp, dp]
        L.fip_closure_fused_outputs.argtypes = [vp, vp, fp, C.c_float, fp]
        L.fip_closure_set_kernel.argtypes = [vp, C.c_int]
        L.fip_closure_set_activation.argtypes = [vp, C.c_int]
        L.fip_closure_get_activation.argtypes = [vp, C.POINTER(C.c_int)]
        L.fip_mfma_probe.argtypes = [C.c_int, C.POINTER(C.c_ushort), C.POINTER(C.c_ushort), fp]
        L.fip_check_rcp.argtypes = [C.c_int, C.c_float, C.c_float, C.POINTER(C.c_ulonglong), fp]
        L.fip_tune.argtypes = [C.POINTER(FipDesc), C.POINTER(FipBC), C.c_int, C.c_int,
                               C.POINTER(FipTune)]
        L.fip_ens_set_timing.argtypes = [vp, C.c_int]
        L.fip_ens_get_stats.argtypes = [vp, C.POINTER(FipStats)]
        _lib = L
    return _lib


def _check(rc: int):
    if rc != 0:
        raise FipError(f"libfiphip error {rc}: {lib().fip_last_error().decode()}")


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _desc(case: Case, keep: list, mode: str, variant: str, tile_rows: int, device: int,
          staggered: bool) -> FipDesc:
    def arr(a):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=np.float32)
        keep.append(a)
        return a

    bed, h, qx, qy, kv = (arr(case.bed), arr(case.h), arr(case.qx), arr(case.qy),
                          arr(case.kst_var))
    return FipDesc(case.W, case.H, case.dx, case.dt,
                   case.no_data if case.no_data is not None else -9999.0, case.kst,
                   _ptr(bed), _ptr(h), _ptr(qx), _ptr(qy), _ptr(kv),
                   1 if case.variable_dt else 0, 1 if staggered else 0,
                   1 if mode == "masked" else 0, 1 if variant == "fast" else 0, tile_rows, device)


def _bc_array(case: Case, bcs, keep: list):
    arr = (FipBC * max(1, len(bcs)))()
    for i, bc in enumerate(bcs):
        frm, to = case.bc_range(bc)
        b = arr[i]
        b.type, b.side, b.from_, b.to = BC_TYPES[bc.type], BC_SIDES[bc.side], frm, to
        b.z = bc.z if bc.z is not None else 0.0
        b.q = bc.q if bc.q is not None else 0.0
        for nm, ser in (("z", bc.z_series), ("q", bc.q_series)):
            if ser is not None:
                t = np.asarray(ser[0], dtype=np.float32)
                v = np.asarray(ser[1], dtype=np.float32)
                keep += [t, v]
                setattr(b, nm + "t", _ptr(t))
                setattr(b, nm + "v", _ptr(v))
                setattr(b, nm + "n", len(t))
    return arr


class Ensemble:
    """Many scenarios advanced together on one GPU (include/fip_ensemble.h)."""

    def __init__(self, cases, mode: str = "masked", variant: str = "exact", tile_rows: int = 32,
                 device: int = 0, staggered: bool = True):
        L = lib()
        self.cases = list(cases)
        keep = []
        descs = (FipDesc * len(self.cases))(*[_desc(c, keep, mode, variant, tile_rows, device,
                                                     staggered) for c in self.cases])
        e = C.c_void_p()
        _check(L.fip_ens_create(descs, len(self.cases), C.byref(e)))
        self.e = e
        for i, c in enumerate(self.cases):
            k2 = []
            arr = _bc_array(c, c.bcs, k2)
            _check(L.fip_ens_set_bcs(self.e, i, arr, len(c.bcs)))

    def close(self):
        if getattr(self, "e", None):
            lib().fip_ens_destroy(self.e)
            self.e = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_run(self, targets):
        t = (C.c_double * len(self.cases))(*[float(x) for x in targets])
        _check(lib().fip_ens_set_run(self.e, t))

    def simulate(self, nsteps):
        act, done = C.c_int(), C.c_longlong()
        _check(lib().fip_ens_simulate(self.e, int(nsteps), C.byref(act), C.byref(done)))
        return act.value, done.value

    def prepare(self, nsteps):
        """Capture the graphs of the next simulate(nsteps) call (runs nothing)."""
        _check(lib().fip_ens_prepare(self.e, int(nsteps)))

    def time(self, s):
        t, dt, it = C.c_double(), C.c_float(), C.c_int()
        _check(lib().fip_ens_get_time(self.e, s, C.byref(t), C.byref(dt), C.byref(it)))
        return t.value, dt.value, it.value

    def field(self, s, name, side=0):
        c = self.cases[s]
        a = np.empty((c.H, c.W), dtype=np.float32)
        _check(lib().fip_ens_download(self.e, s, FIELDS[name], side,
                                      a.ctypes.data_as(C.POINTER(C.c_float))))
        return a

    def set_launch(self, waves_per_block=0, blocks_per_cu=-1):
        _check(lib().fip_ens_set_launch(self.e, int(waves_per_block), int(blocks_per_cu)))

    def set_timing(self, on: bool):
        _check(lib().fip_ens_set_timing(self.e, 1 if on else 0))

    def stats(self):
        st = FipStats()
        _check(lib().fip_ens_get_stats(self.e, C.byref(st)))
        return {k: getattr(st, k) for k, _ in FipStats._fields_}


class Sim:
    """One simulation on one GPU, created from a workload Case."""

    def __init__(self, case: Case, mode: str = "dense", variant: str = "exact",
                 tile_rows: int = 64, device: int = 0, staggered: bool = True,
                 generate_bcs: bool = True):
        L = lib()
        self.case = case
        self.W, self.H = case.W, case.H
        keep = []

        def arr(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float32)
            keep.append(a)
            return a

        bed, h, qx, qy, kv = (arr(case.bed), arr(case.h), arr(case.qx), arr(case.qy),
                              arr(case.kst_var))
        d = FipDesc(case.W, case.H, case.dx, case.dt,
                    case.no_data if case.no_data is not None else -9999.0, case.kst,
                    _ptr(bed), _ptr(h), _ptr(qx), _ptr(qy), _ptr(kv),
                    1 if case.variable_dt else 0, 1 if staggered else 0,
                    1 if mode == "masked" else 0, 1 if variant == "fast" else 0, tile_rows, device)
        s = C.c_void_p()
        _check(L.fip_create(C.byref(d), C.byref(s)))
        self.s = s
        if generate_bcs:
            self.bc_generate(case.bcs)

    def close(self):
        if getattr(self, "s", None):
            lib().fip_destroy(self.s)
            self.s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bc_generate(self, bcs):
        arr = (FipBC * max(1, len(bcs)))()
        self._series = []
        for i, bc in enumerate(bcs):
            frm, to = self.case.bc_range(bc)
            b = arr[i]
            b.type, b.side, b.from_, b.to = BC_TYPES[bc.type], BC_SIDES[bc.side], frm, to
            b.z = bc.z if bc.z is not None else 0.0
            b.q = bc.q if bc.q is not None else 0.0
            for nm, ser in (("z", bc.z_series), ("q", bc.q_series)):
                if ser is not None:
                    t = np.asarray(ser[0], dtype=np.float32)
                    v = np.asarray(ser[1], dtype=np.float32)
                    self._series += [t, v]
                    setattr(b, nm + "t", _ptr(t))
                    setattr(b, nm + "v", _ptr(v))
                    setattr(b, nm + "n", len(t))
        _check(lib().fip_bc_generate(self.s, arr, len(bcs)))
        self.nbc = len(bcs)

    def bc_launch(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        _check(lib().fip_bc_launch(self.s, p.ctypes.data_as(C.POINTER(C.c_float))))

    def bc_apply(self):
        _check(lib().fip_bc_apply(self.s))

    def bc_params(self):
        p = np.zeros(2 * max(1, self.nbc), dtype=np.float32)
        _check(lib().fip_bc_params(self.s, p.ctypes.data_as(C.POINTER(C.c_float)), self.nbc))
        return p[:2 * self.nbc].reshape(-1, 2)

    def control(self):
        _check(lib().fip_control_timestep(self.s))

    def compute_step(self):
        _check(lib().fip_compute_step(self.s))

    def floodplain(self):
        _check(lib().fip_floodplain(self.s))

    def step_reference_order(self):
        """One simulate() iteration through the fine-grained seam (src/main.cpp:100-135)."""
        t, dt, it = self.time()
        if self.case.variable_dt:
            self.control()
        if it == 0:
            self.bc_apply()
        self.compute_step()
        self.bc_apply()
        if (it + 1) % 100 == 0:
            self.floodplain()

    def set_run(self, target, sampling_interval=0, points=(), save_times=()):
        n = len(points)
        xs = (C.c_int * max(1, n))(*[p[0] for p in points])
        ys = (C.c_int * max(1, n))(*[p[1] for p in points])
        ts = (C.c_int * max(1, n))(*[p[2] for p in points])
        sv = (C.c_double * max(1, len(save_times)))(*save_times)
        _check(lib().fip_set_run(self.s, float(target), int(sampling_interval), xs, ys, ts, n, sv,
                                 len(save_times)))
        self.npoints = n

    def simulate(self, nsteps):
        ev, done = C.c_int(), C.c_longlong()
        _check(lib().fip_simulate(self.s, int(nsteps), C.byref(ev), C.byref(done)))
        return ev.value, done.value

    def prepare(self, nsteps):
        """Capture the graphs that simulate(nsteps) launches next (nothing runs)."""
        _check(lib().fip_prepare(self.s, int(nsteps)))

    def read_samples(self, max_rows=256):
        t = np.zeros(max_rows, dtype=np.float64)
        v = np.zeros(max_rows * max(1, self.npoints), dtype=np.float32)
        n = C.c_int()
        _check(lib().fip_read_samples(self.s, t.ctypes.data_as(C.POINTER(C.c_double)),
                                      v.ctypes.data_as(C.POINTER(C.c_float)), max_rows, C.byref(n)))
        return t[:n.value], v[:n.value * self.npoints].reshape(n.value, self.npoints)

    def time(self):
        t, dt, it = C.c_double(), C.c_float(), C.c_int()
        _check(lib().fip_get_time(self.s, C.byref(t), C.byref(dt), C.byref(it)))
        return t.value, dt.value, it.value

    def set_dt(self, dt):
        _check(lib().fip_set_dt(self.s, dt))

    def field(self, name, side=0):
        out = np.empty((self.H, self.W), dtype=np.float32)
        _check(lib().fip_download(self.s, FIELDS[name], side, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def set_field(self, name, side, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        _check(lib().fip_upload(self.s, FIELDS[name], side, a.ctypes.data_as(C.POINTER(C.c_float))))

    def state(self, side=0):
        return {k: self.field(k, side) for k in ("h", "qx", "qy")}

    def tiles(self):
        a = [C.c_int() for _ in range(4)]
        _check(lib().fip_tiles(self.s, *[C.byref(x) for x in a]))
        return tuple(x.value for x in a)

    def masks(self):
        tx, ty, _, _ = self.tiles()
        out = np.empty(tx * ty, dtype=np.uint64)
        _check(lib().fip_get_masks(self.s, out.ctypes.data_as(C.POINTER(C.c_uint64)), len(out)))
        return out

    def cfl(self):
        c = C.c_float()
        _check(lib().fip_get_cfl(self.s, C.byref(c)))
        return c.value

    def set_launch(self, waves_per_block=0, blocks_per_cu=-1):
        _check(lib().fip_set_launch(self.s, int(waves_per_block), int(blocks_per_cu)))

    def set_timing(self, on: bool):
        _check(lib().fip_set_timing(self.s, 1 if on else 0))

    def stats(self):
        st = FipStats()
        _check(lib().fip_get_stats(self.s, C.byref(st)))
        return {k: getattr(st, k) for k, _ in FipStats._fields_}

    def sync(self):
        _check(lib().fip_sync(self.s))


def _tune_dict(r: FipTune) -> dict:
    return {k: getattr(r, k) for k, _ in FipTune._fields_}


def tune(case: Case, mode: str = "dense", variant: str = "fast", steps: int = 20,
         staggered: bool = True) -> dict:
    """fip_tune: the fastest tile rows / launch shape for this case."""
    keep = []
    d = _desc(case, keep, mode, variant, 0, 0, staggered)
    arr = _bc_array(case, case.bcs, keep)
    r = FipTune()
    _check(lib().fip_tune(C.byref(d), arr, len(case.bcs), int(steps), C.byref(r)))
    return _tune_dict(r)


def tune_ensemble(cases, mode: str = "masked", variant: str = "fast", steps: int = 20,
                  staggered: bool = True) -> dict:
    """fip_ens_tune: the fastest tile rows / launch shape for an ensemble."""
    keep = []
    descs = (FipDesc * len(cases))(*[_desc(c, keep, mode, variant, 0, 0, staggered) for c in cases])
    arrs = [_bc_array(c, c.bcs, keep) for c in cases]
    ptrs = (C.POINTER(FipBC) * len(cases))(*[C.cast(a, C.POINTER(FipBC)) for a in arrs])
    nbcs = (C.c_int * len(cases))(*[len(c.bcs) for c in cases])
    r = FipTune()
    _check(lib().fip_ens_tune(descs, len(cases), ptrs, nbcs, int(steps), C.byref(r)))
    return _tune_dict(r)


def _f32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def restrict(fine: "Sim", coarse: "Sim", factor: int = 4):
    """fip_restrict: the fine simulation's current state onto the coarse one."""
    _check(lib().fip_restrict(fine.s, coarse.s, int(factor)))


def check_rcp(lo, hi, device=0):
    """fip_check_rcp: (number of binary32 b in [lo, hi] where the exact variant's face-depth
    reciprocal differs from the IEEE division, the smallest such b or 0)."""
    n, f = C.c_ulonglong(), C.c_float()
    _check(lib().fip_check_rcp(int(device), float(lo), float(hi), C.byref(n), C.byref(f)))
    return n.value, f.value


def mfma_probe(A, B, device=0):
    A = np.ascontiguousarray(A, dtype=np.uint16)
    B = np.ascontiguousarray(B, dtype=np.uint16)
    D = np.zeros((16, 16), np.float32)
    _check(lib().fip_mfma_probe(device, A.ctypes.data_as(C.POINTER(C.c_ushort)),
                                B.ctypes.data_as(C.POINTER(C.c_ushort)), _f32p(D)))
    return D


ACTIVATIONS = {"gelu": 0, "hsilu": 1}  # FIP_CLOSURE_ACT_GELU, FIP_CLOSURE_ACT_HSILU


class Closure:
    """The learned subgrid closure (include/fip_closure.h). activation: "gelu" (default) or
    "hsilu" (hard SiLU, x * clamp(x / 4 + 1 / 2, 0, 1))."""

    def __init__(self, Wc, Hc, max_samples=0, seed=1, device=0, activation="gelu"):
        h = C.c_void_p()
        _check(lib().fip_closure_create(device, seed, Wc, Hc, max_samples, C.byref(h)))
        self.c = h
        self.Wc, self.Hc = Wc, Hc
        try:
            self.set_activation(activation)
        except Exception:
            self.close()
            raise

    @property
    def activation(self) -> str:
        a = C.c_int()
        _check(lib().fip_closure_get_activation(self.c, C.byref(a)))
        return {v: k for k, v in ACTIVATIONS.items()}[a.value]

    def set_activation(self, activation: str):
        """Switch the network's activation (not while attached); the parameters stay."""
        if activation not in ACTIVATIONS:
            raise ValueError(f"activation must be one of {sorted(ACTIVATIONS)}")
        _check(lib().fip_closure_set_activation(self.c, ACTIVATIONS[activation]))

    def close(self):
        if getattr(self, "c", None):
            lib().fip_closure_destroy(self.c)
            self.c = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self):
        p = np.zeros(CL_PARAMS, np.float32)
        _check(lib().fip_closure_get_params(self.c, _f32p(p)))
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == CL_PARAMS
        _check(lib().fip_closure_set_params(self.c, _f32p(p)))

    def add_sample(self, coarse: "Sim", fine: "Sim", bed_std, zref, factor=4):
        b = np.ascontiguousarray(bed_std, dtype=np.float32)
        _check(lib().fip_closure_add_sample(self.c, coarse.s, fine.s, factor, _f32p(b), float(zref)))

    def num_samples(self):
        n = C.c_int()
        _check(lib().fip_closure_num_samples(self.c, C.byref(n)))
        return n.value

    def adam_step(self, grad, lr):
        g = np.ascontiguousarray(grad, dtype=np.float32)
        assert g.size == CL_PARAMS
        _check(lib().fip_closure_adam_step(self.c, _f32p(g), float(lr)))

    def train(self, steps, batch=32, patch=128, lr=1e-3, first=0, n=None, seed=7):
        n = self.num_samples() - first if n is None else n
        out = np.zeros(steps, np.float32)
        _check(lib().fip_closure_train(self.c, steps, batch, patch, lr, first, n, seed, _f32p(out)))
        return out

    def eval(self, first=0, n=None):
        n = self.num_samples() - first if n is None else n
        a, b = C.c_double(), C.c_double()
        _check(lib().fip_closure_eval(self.c, first, n, C.byref(a), C.byref(b)))
        return a.value, b.value

    def attach(self, coarse: "Sim", bed_std, zref, sat=2.65, mass_limit=1e-4):
        b = np.ascontiguousarray(bed_std, dtype=np.float32)
        _check(lib().fip_closure_attach(self.c, coarse.s, _f32p(b), float(zref), float(sat),
                                        float(mass_limit)))

    def attach_ensemble(self, ens: "Ensemble", bed_std, zref, sat=2.65, mass_limit=1e-4):
        b = np.ascontiguousarray(bed_std, dtype=np.float32)
        z = np.ascontiguousarray(zref, dtype=np.float32)
        _check(lib().fip_closure_attach_ensemble(self.c, ens.e, _f32p(b), _f32p(z), float(sat),
                                                 float(mass_limit)))

    @staticmethod
    def detach(coarse: "Sim"):
        _check(lib().fip_closure_detach(coarse.s))

    def status(self, coarse: "Sim" = None, ens: "Ensemble" = None, s=0):
        fb, ns, mr = C.c_int(), C.c_uint(), C.c_double()
        _check(lib().fip_closure_status(self.c, coarse.s if coarse else None, ens.e if ens else None, s,
                                        C.byref(fb), C.byref(ns), C.byref(mr)))
        return {"fallback": fb.value, "saturated_faces": ns.value, "max_rel_mass_error": mr.value}

    def forward(self, feat):
        f = np.ascontiguousarray(feat, dtype=np.float32)
        N, H, W, _ = f.shape
        out = np.zeros((N, H, W, 2), np.float32)
        _check(lib().fip_closure_forward(self.c, _f32p(f), N, H, W, _f32p(out)))
        return out

    def forward_noalloc(self, feat_dev, N, H, W, out_dev):
        """fip_closure_forward_noalloc on device pointers (integers): feat [N][H][W][6] and
        out [N][H][W][2] fp32 in device memory."""
        _check(lib().fip_closure_forward_noalloc(self.c, C.c_void_p(feat_dev), N, H, W, C.c_void_p(out_dev)))

    def set_kernel(self, form: int):
        """Launch form of the fused inference kernel: 2 row march (default), 1 persistent
        tiles, 0 one work-group per tile."""
        _check(lib().fip_closure_set_kernel(self.c, int(form)))

    def fused_outputs(self, coarse: "Sim", bed_std, zref):
        b = np.ascontiguousarray(bed_std, dtype=np.float32)
        out = np.zeros((coarse.H, coarse.W, 2), np.float32)
        _check(lib().fip_closure_fused_outputs(self.c, coarse.s, _f32p(b), float(zref), _f32p(out)))
        return out

    def backward(self, feat, qp, qt, mask, bound=0.1):
        f = np.ascontiguousarray(feat, dtype=np.float32)
        N, H, W, _ = f.shape
        qp = np.ascontiguousarray(qp, dtype=np.float32)
        qt = np.ascontiguousarray(qt, dtype=np.float32)
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        g = np.zeros(CL_PARAMS, np.float32)
        loss = C.c_double()
        _check(lib().fip_closure_backward(self.c, _f32p(f), N, H, W, _f32p(qp), _f32p(qt),
                                          m.ctypes.data_as(C.POINTER(C.c_ubyte)), float(bound), _f32p(g),
                                          C.byref(loss)))
        return loss.value, g
