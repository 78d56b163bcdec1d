"""The C ABI library: loads, exports every declared entry point, fails loudly without a
GPU, and generates boundary kernels that hipRTC compiles for gfx950. CPU only."""
import ctypes as C
import os
import re

import pytest

from conftest import gpu_available
from fip_amd.workloads import BCSpec, Case, small_case
from rtc_check import bc_source, rtc_compile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            names |= set(re.findall(r"\b(fip_[a-z_0-9]+)\s*\(", txt))
    return sorted(names)


def test_exports_every_declared_symbol():
    from fip_amd.sim import LIB_PATH
    lib = C.CDLL(LIB_PATH)
    names = declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_exports():
    from fip_amd.sim import CLOSURE_EXPORTS, ENS_EXPORTS, EXPORTS
    assert set(EXPORTS) | set(ENS_EXPORTS) | set(CLOSURE_EXPORTS) == set(declared())


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU error path")
def test_create_fails_loudly_without_gpu():
    from fip_amd.sim import FipError, Sim
    with pytest.raises(FipError, match="hipSetDevice"):
        Sim(small_case(16, 16))


def _case_with(bcs, W=50, H=40):
    c = small_case(W, H, wet=True)
    c.bcs = bcs
    return c


BC_LISTS = {
    "reference_cfg0": None,
    "all_types_all_sides": [BCSpec("z", "left", 3, 9, z=12.0), BCSpec("q", "right", 0, 20, z=11.0, q=5.0),
                            BCSpec("open", "top"), BCSpec("close", "bottom"),
                            BCSpec("open", "left", 20, 39), BCSpec("close", "right", 21, 39)],
    "series": [BCSpec("q", "left", 2, 30, z_series=([0.0, 100.0], [10.0, 12.0]),
                      q_series=([-60.0, 1e4], [0.0, 50.0])),
               BCSpec("z", "bottom", 5, 10, z_series=([0.0], [3.0]))],
    "overlapping_same_side": [BCSpec("z", "left", 0, 20, z=10.0), BCSpec("close", "left", 10, 39)],
    "empty": [],
}


@pytest.mark.parametrize("staggered", [True, False])
@pytest.mark.parametrize("name", list(BC_LISTS))
def test_generated_bc_kernel_compiles(name, staggered):
    case = small_case(512, 512, wet=False) if name == "reference_cfg0" else _case_with(BC_LISTS[name])
    src = bc_source(case, staggered)
    rc, log, size = rtc_compile(src)
    assert rc == 0, log
    assert size > 0
    if name == "overlapping_same_side":
        assert src.count("// stage") == 2  # the overlap orders the two BCs


def test_bench_refuses_several_gpus():
    """bench.py measures one GPU: --gpus N with N > 1 exits with a message, before any
    GPU work."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""
