"""fip_amd: the fip shallow-water time step on MI355X (gfx950).

libfiphip.so (built by fip_amd/build.py) holds the HIP kernels and the C ABI of
include/fip.h; fip_amd.sim is a ctypes binding of it; fip_amd.workloads generates the
synthetic flood scenarios; bin/fip is the drop-in `fip parameters.yml` CLI.
"""
