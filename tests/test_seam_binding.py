"""The reference-side binding fip_amd/seam/fip_hip_seam.cpp compiles against the reference's
own seam headers (src/kernels.h, src/bc_kernel_gen.h, src/simdata.h, src/boundary_condition.h,
CUDA-free) and links against libfiphip: it defines exactly the functions of the seam, and
what it leaves undefined is either exported by libfiphip or a SimData global of the
reference's simdata.cpp. CPU only; needs the reference tree."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src"
SEAM = os.path.join(ROOT, "fip_amd", "seam")
LIB = os.path.join(ROOT, "fip_amd", "libfiphip.so")

pytestmark = pytest.mark.skipif(not os.path.isfile(os.path.join(REF, "kernels.h")),
                                reason="the reference tree is absent")


def symbols(path, *flags):
    out = subprocess.run(["nm", "-C", *flags, path], check=True, capture_output=True, text=True).stdout
    return {line.split(None, 2)[-1] for line in out.splitlines() if line.strip()}


def test_seam_binding_compiles_and_links(tmp_path):
    obj = tmp_path / "fip_hip_seam.o"
    so = tmp_path / "libseam.so"
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fPIC", "-c",
                    os.path.join(SEAM, "fip_hip_seam.cpp"), "-I", os.path.join(ROOT, "include"),
                    "-I", SEAM, "-I", REF, "-o", str(obj)], check=True)
    subprocess.run(["g++", "-shared", "-o", str(so), str(obj), "-L", os.path.dirname(LIB), "-lfiphip"],
                   check=True)
    defined = symbols(str(obj), "--defined-only", "--extern-only")
    for want in ("launchComputeKernel()", "launchFloodPlainKernel()", "BoundaryConditionKernel::generate()",
                 "BoundaryConditionKernel::launch(void*, std::vector<JitParam, std::allocator<JitParam> > const&)",
                 "initDeviceBuffers(float const*, float const*, float const*, float const*, float const*)"):
        assert want in defined, (want, sorted(defined))
    exported = symbols(LIB, "-D", "--defined-only")
    undefined = symbols(str(so), "-u")
    fip = {s for s in undefined if s.startswith("fip_")}
    assert fip and fip <= exported, fip - exported
    ours = {s for s in undefined if not s.startswith(("fip_", "std::", "operator", "__", "_ITM", "_Unwind"))
            and "@" not in s and "GLIBC" not in s}
    assert ours and all(s.startswith("SimData::") for s in ours), sorted(ours)
