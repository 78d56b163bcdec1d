"""Compile generated boundary-kernel sources with hipRTC for gfx950 (no GPU needed)."""
import ctypes as C
import os

from fip_amd.sim import FipBC, lib
from fip_amd.workloads import BC_SIDES, BC_TYPES

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "fip_amd", "csrc")
HEADERS = ["fip_ctl.h", "fip_detmath.h", "fip_sched.h", "fip_step_common.h", "fip_bc_common.h"]


def bc_source(case, staggered=True):
    arr = (FipBC * max(1, len(case.bcs)))()
    keep = []
    import numpy as np
    for i, bc in enumerate(case.bcs):
        frm, to = case.bc_range(bc)
        b = arr[i]
        b.type, b.side, b.from_, b.to = BC_TYPES[bc.type], BC_SIDES[bc.side], frm, to
        b.z = bc.z or 0.0
        b.q = bc.q or 0.0
        for nm, ser in (("z", bc.z_series), ("q", bc.q_series)):
            if ser is not None:
                t = np.asarray(ser[0], np.float32)
                v = np.asarray(ser[1], np.float32)
                keep += [t, v]
                setattr(b, nm + "t", t.ctypes.data_as(C.c_void_p))
                setattr(b, nm + "v", v.ctypes.data_as(C.c_void_p))
                setattr(b, nm + "n", len(t))
    L = lib()
    L.fip_bc_source.argtypes = [C.POINTER(FipBC), C.c_int, C.c_int, C.c_int, C.c_int, C.c_char_p,
                                C.c_size_t, C.POINTER(C.c_size_t)]
    n = C.c_size_t()
    rc = L.fip_bc_source(arr, len(case.bcs), case.W, case.H, 1 if staggered else 0, None, 0,
                         C.byref(n))
    assert rc == 0, L.fip_last_error()
    buf = C.create_string_buffer(n.value + 1)
    L.fip_bc_source(arr, len(case.bcs), case.W, case.H, 1 if staggered else 0, buf, n.value + 1,
                    C.byref(n))
    return buf.value.decode()


def rtc_compile(src, arch="gfx950"):
    R = C.CDLL("libhiprtc.so")
    prog = C.c_void_p()
    hs = [open(os.path.join(CSRC, h)).read().encode() for h in HEADERS]
    hsrc = (C.c_char_p * len(hs))(*hs)
    hnames = (C.c_char_p * len(hs))(*[h.encode() for h in HEADERS])
    assert R.hiprtcCreateProgram(C.byref(prog), src.encode(), b"bc.hip", len(hs), hsrc, hnames) == 0
    opts = [f"--offload-arch={arch}", "-O3", "-std=c++17", "-ffp-contract=off", "-Xclang",
            "-target-feature", "-Xclang", "-packed-fp32-ops"]
    copts = (C.c_char_p * len(opts))(*[o.encode() for o in opts])
    rc = R.hiprtcCompileProgram(prog, len(opts), copts)
    n = C.c_size_t()
    R.hiprtcGetProgramLogSize(prog, C.byref(n))
    log = C.create_string_buffer(n.value + 1)
    R.hiprtcGetProgramLog(prog, log)
    code = C.c_size_t()
    R.hiprtcGetCodeSize(prog, C.byref(code))
    R.hiprtcDestroyProgram(C.byref(prog))
    return rc, log.value.decode(errors="replace"), code.value
