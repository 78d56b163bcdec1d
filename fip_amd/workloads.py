"""Synthetic flood scenarios for the fip time step, generated from a frozen recipe.

PRNG: splitmix64 (Steele, Lea & Flood 2014); the stream for (config c, scenario s)
starts from seed c * 2**32 + s. A uniform draw in [0,1) takes the top 24 bits of one
64-bit output. Values are computed in binary64 and stored as float32 rasters.

Bed for cell (x, y): slope * dx * (W-1-x) + sum_k A_k exp(-r^2 / (2 sigma_k^2)) - channel,
where each Gaussian hill is evaluated inside its window |x-cx|,|y-cy| <= ceil(4 sigma)
(outside it the hill contributes 0).

Rasters are written as raw float32 after a 36-byte header so that the `[path, 36]`
byte-offset form of parameters.yml (src/simdata.cpp:143-150) is exercised.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

MASK64 = (1 << 64) - 1
HEADER_BYTES = 36

BC_TYPES = {"z": 0, "q": 1, "close": 2, "open": 3}
BC_SIDES = {"left": 0, "right": 1, "top": 2, "bottom": 3}


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next() >> 40) / float(1 << 24))


def seed_for(config: int, scenario: int = 0) -> int:
    return config * (1 << 32) + scenario


@dataclass
class BCSpec:
    type: str
    side: str
    frm: int | None = None
    to: int | None = None
    z: float | None = None
    q: float | None = None
    z_series: tuple[list[float], list[float]] | None = None  # (times s, values)
    q_series: tuple[list[float], list[float]] | None = None


@dataclass
class Case:
    name: str
    W: int
    H: int
    dx: float
    dt: float
    duration: float
    bed: np.ndarray
    kst: float = 30.0
    kst_var: np.ndarray | None = None
    h: np.ndarray | None = None
    qx: np.ndarray | None = None
    qy: np.ndarray | None = None
    no_data: float | None = None
    variable_dt: bool = True
    bcs: list[BCSpec] = field(default_factory=list)

    def bc_range(self, bc: BCSpec) -> tuple[int, int]:
        frm = 0 if bc.frm is None else bc.frm
        to = bc.to
        if to is None:
            to = self.H - 1 if bc.side in ("left", "right") else self.W - 1
        return frm, to


def add_hills(bed: np.ndarray, rng: SplitMix64, n: int, amp: tuple[float, float],
              sigma: tuple[float, float]) -> None:
    H, W = bed.shape
    for _ in range(n):
        a = rng.uniform(*amp)
        s = rng.uniform(*sigma)
        cx = rng.uniform(0.0, W)
        cy = rng.uniform(0.0, H)
        r = int(math.ceil(4.0 * s))
        x0, x1 = max(0, int(cx) - r), min(W, int(cx) + r + 1)
        y0, y1 = max(0, int(cy) - r), min(H, int(cy) + r + 1)
        if x0 >= x1 or y0 >= y1:
            continue
        xs = np.arange(x0, x1, dtype=np.float64) - cx
        ys = np.arange(y0, y1, dtype=np.float64) - cy
        gx = np.exp(-(xs * xs) / (2.0 * s * s))
        gy = np.exp(-(ys * ys) / (2.0 * s * s))
        bed[y0:y1, x0:x1] += a * np.outer(gy, gx)


def terrain_plane_hills(W: int, H: int, dx: float, slope: float, nhills: int,
                        rng: SplitMix64, amp=(0.0, 5.0), sigma=(10.0, 60.0)) -> np.ndarray:
    x = np.arange(W, dtype=np.float64)
    bed = np.broadcast_to(slope * dx * (W - 1 - x), (H, W)).copy()
    add_hills(bed, rng, nhills, amp, sigma)
    return bed


def config0(W: int = 512, H: int = 512, nhills: int = 64, duration: float = 3600.0,
            scenario: int = 0) -> Case:
    """Plain single domain: 512^2 at dx=3, dry start, left z inflow into a channel.

    Channel: 20 cells wide along y = H/2 (rows H/2-10 .. H/2+9), 3 m deep.
    BCs: one left type-z entry per row H/2-20 .. H/2+20 with z = bed(0,row) + 2 m,
    right open, top and bottom closed. Sizes other than 512 scale the rows.
    """
    rng = SplitMix64(seed_for(0, scenario))
    dx = 3.0
    bed = terrain_plane_hills(W, H, dx, 0.002, nhills, rng)
    c = H // 2
    bed[c - 10:c + 10, :] -= 3.0
    bed32 = bed.astype(np.float32)
    bcs = []
    for row in range(max(0, c - 20), min(H, c + 21)):
        bcs.append(BCSpec("z", "left", row, row, z=float(np.float32(bed32[row, 0]) + np.float32(2.0))))
    bcs.append(BCSpec("open", "right"))
    bcs.append(BCSpec("close", "top"))
    bcs.append(BCSpec("close", "bottom"))
    return Case(name=f"cfg0_{W}x{H}", W=W, H=H, dx=dx, dt=0.1, duration=duration, bed=bed32,
                kst=30.0, bcs=bcs)


def _convex_hull(pts: np.ndarray) -> np.ndarray:
    """Andrew's monotone chain; returns the hull counter-clockwise."""
    p = sorted(map(tuple, pts))

    def cross(o, a, b):
        return (a[0] - o[0]) * (b[1] - o[1]) - (a[1] - o[1]) * (b[0] - o[0])

    lower, upper = [], []
    for q in p:
        while len(lower) >= 2 and cross(lower[-2], lower[-1], q) <= 0:
            lower.pop()
        lower.append(q)
    for q in reversed(p):
        while len(upper) >= 2 and cross(upper[-2], upper[-1], q) <= 0:
            upper.pop()
        upper.append(q)
    return np.array(lower[:-1] + upper[:-1], dtype=np.float64)


def inside_convex(hull: np.ndarray, W: int, H: int) -> np.ndarray:
    inside = np.ones((H, W), dtype=bool)
    x = np.arange(W, dtype=np.float64)[None, :]
    n = len(hull)
    for row0 in range(0, H, 512):
        y = np.arange(row0, min(H, row0 + 512), dtype=np.float64)[:, None]
        blk = np.ones((y.shape[0], W), dtype=bool)
        for k in range(n):
            ax, ay = hull[k]
            bx, by = hull[(k + 1) % n]
            blk &= ((bx - ax) * (y - ay) - (by - ay) * (x - ax)) >= -1e-9
        inside[row0:row0 + y.shape[0]] = blk
    return inside


def config1(W: int = 5667, H: int = 6500, nhills: int = 2048, duration: float = 12 * 3600.0,
            scenario: int = 0, amp_sine: float = 400.0, wavelength: float = 3000.0,
            width: int = 30, nsites: int = 512) -> Case:
    """Plain large domain at the real shape 5667 x 6500, dx = 3 m.

    Meandering channel y = H/2 + amp*sin(2 pi x / wavelength), `width` cells, 4 m deep;
    0.001 slope; kSt from a Voronoi patchwork of `nsites` sites with values in
    {15,25,35,45,60}; no_data = -1 outside a seeded convex polygon that contains the
    channel's inflow (x=0) and outflow (x=W-1) sections. Left BC type q spans the
    channel at x=0 with a triangular hydrograph 0 -> 600 m^3/s (at 4 h) -> 0 (at 12 h)
    that starts rising 60 s before t=0, and a constant z = channel bed + 2.5 m.
    """
    from scipy.spatial import cKDTree

    rng = SplitMix64(seed_for(1, scenario))
    dx = 3.0
    bed = terrain_plane_hills(W, H, dx, 0.001, nhills, rng)
    x = np.arange(W, dtype=np.float64)
    yc = H / 2.0 + amp_sine * np.sin(2.0 * math.pi * x / wavelength)
    ys = np.arange(H, dtype=np.float64)[:, None]
    chan = np.abs(ys - yc[None, :]) < width / 2.0
    bed[chan] -= 4.0
    # Voronoi roughness
    sites = np.array([[rng.uniform(0, W), rng.uniform(0, H)] for _ in range(nsites)])
    classes = np.array([int(rng.uniform(0, 5)) for _ in range(nsites)])
    values = np.array([15.0, 25.0, 35.0, 45.0, 60.0], dtype=np.float32)
    tree = cKDTree(sites)
    kst = np.empty((H, W), dtype=np.float32)
    gx = np.arange(W, dtype=np.float64)
    for row0 in range(0, H, 256):
        rows = np.arange(row0, min(H, row0 + 256), dtype=np.float64)
        pts = np.stack(np.meshgrid(gx, rows), axis=-1).reshape(-1, 2)
        _, idx = tree.query(pts, workers=-1)
        kst[row0:row0 + len(rows)] = values[classes[idx]].reshape(len(rows), W)
    # convex no_data polygon
    c0, c1 = yc[0], yc[-1]
    pts = [(0.0, c0 - width), (0.0, c0 + width), (W - 1.0, c1 - width), (W - 1.0, c1 + width)]
    for k in range(16):
        t = k / 16.0
        inset = rng.uniform(0.0, 0.15)
        if k % 4 == 0:
            pts.append((t * W, inset * H))
        elif k % 4 == 1:
            pts.append(((1 - inset) * W, t * H))
        elif k % 4 == 2:
            pts.append(((1 - t) * W, (1 - inset) * H))
        else:
            pts.append((inset * W, (1 - t) * H))
    hull = _convex_hull(np.array(pts))
    inside = inside_convex(hull, W, H)
    bed32 = bed.astype(np.float32)
    bed32[~inside] = np.float32(-1.0)
    rows = np.nonzero(chan[:, 0])[0]
    frm, to = int(rows.min()), int(rows.max())
    zc = float(np.float32(bed32[frm:to + 1, 0].min()) + np.float32(2.5))
    qt = [-60.0, 4 * 3600.0, 12 * 3600.0]
    qv = [0.0, 600.0, 0.0]
    bcs = [BCSpec("q", "left", frm, to, z_series=([0.0, duration], [zc, zc]), q_series=(qt, qv)),
           BCSpec("open", "right")]
    return Case(name=f"cfg1_{W}x{H}", W=W, H=H, dx=dx, dt=0.1, duration=duration, bed=bed32,
                kst_var=kst, no_data=-1.0, bcs=bcs)


def level_state(case: Case, frac: float = 0.30) -> Case:
    """A flat water level that wets `frac` of the valid cells (h = max(0, level - bed)):
    the late, partly flooded state of a run, as an initial state. Returns the case."""
    valid = case.bed != np.float32(case.no_data if case.no_data is not None else -9999.0)
    level = np.float32(np.quantile(case.bed[valid], frac))
    case.h = np.where(valid, np.maximum(np.float32(0), level - case.bed), np.float32(0)).astype(np.float32)
    return case


CONFIG2_SIZES = ((256, 256), (512, 384), (1024, 1024), (1536, 1024), (2048, 1536))


def config2(scenario: int, sizes=CONFIG2_SIZES, duration: float = 6 * 3600.0) -> Case:
    """One scenario of the batched ensemble: a config-0 style domain of a drawn size.

    Size drawn uniformly from `sizes` (W, H), dx = 3 m; terrain as config 0 (slope 0.002,
    64 hills per 512^2 of area, a 20-cell channel 3 m deep along y = H/2); uniform
    kSt ~ U(20, 45); left type q over the channel with a triangular hydrograph rising
    from 0 (60 s before t=0) to a peak ~ U(50, 800) m^3/s at t_peak ~ U(0.5, 3) h and
    back to 0 at 6 h, z constant at the channel bed + 2.5 m; right open; top and bottom
    closed.
    """
    rng = SplitMix64(seed_for(2, scenario))
    W, H = sizes[min(len(sizes) - 1, int(rng.uniform(0, len(sizes))))]
    dx = 3.0
    nhills = max(1, int(round(64 * W * H / (512.0 * 512.0))))
    bed = terrain_plane_hills(W, H, dx, 0.002, nhills, rng)
    c = H // 2
    bed[c - 10:c + 10, :] -= 3.0
    kst = float(np.float32(rng.uniform(20.0, 45.0)))
    peak = rng.uniform(50.0, 800.0)
    tpeak = rng.uniform(0.5, 3.0) * 3600.0
    bed32 = bed.astype(np.float32)
    frm, to = c - 10, c + 9
    zc = float(np.float32(bed32[frm:to + 1, 0].min()) + np.float32(2.5))
    qt = [-60.0, tpeak, 6 * 3600.0]
    qv = [0.0, peak, 0.0]
    bcs = [BCSpec("q", "left", frm, to, z_series=([0.0, duration], [zc, zc]), q_series=(qt, qv)),
           BCSpec("open", "right"), BCSpec("close", "top"), BCSpec("close", "bottom")]
    return Case(name=f"cfg2_{scenario}_{W}x{H}", W=W, H=H, dx=dx, dt=0.1, duration=duration,
                bed=bed32, kst=kst, bcs=bcs)


def small_case(W: int, H: int, scenario: int = 0, wet: bool = True, kst_var: bool = False,
               no_data: bool = False, bc_kinds=("z", "open", "close", "close")) -> Case:
    """A small config-0 style case for parity tests, optionally with an initial lake."""
    rng = SplitMix64(seed_for(9, scenario))
    dx = 3.0
    bed = terrain_plane_hills(W, H, dx, 0.002, max(2, (W * H) // 4096), rng,
                              sigma=(2.0, max(3.0, min(W, H) / 6.0)))
    c = H // 2
    half = max(1, H // 25)
    bed[max(0, c - half):c + half, :] -= 3.0
    bed32 = bed.astype(np.float32)
    h = None
    if wet:
        level = np.float32(np.percentile(bed32, 40))
        h = np.maximum(np.float32(0), level - bed32).astype(np.float32)
    kv = None
    if kst_var == "many":  # about 1000 distinct values: kSt is read as floats
        jj, ii = np.mgrid[0:H, 0:W]
        kv = (20.0 + ((ii * 7919 + jj * 104729) % 1000) * 0.025).astype(np.float32)
    elif kst_var:  # five values: kSt is read as a one-byte class per cell
        kv = np.empty((H, W), dtype=np.float32)
        for j in range(H):
            for i in range(W):
                kv[j, i] = (15.0, 25.0, 35.0, 45.0, 60.0)[((i // 7) * 3 + (j // 5)) % 5]
    nd = None
    if no_data:
        nd = -1.0
        yy, xx = np.mgrid[0:H, 0:W]
        cut = (xx > 0.75 * W) & (yy > 0.7 * H)
        bed32[cut] = np.float32(-1.0)
        if h is not None:
            h[cut] = 0
    bcs = []
    zl, zr, zt, zb = bc_kinds
    lo, hi = max(0, c - 2 * half), min(H - 1, c + 2 * half)
    if zl == "z":
        for row in range(lo, hi + 1):
            bcs.append(BCSpec("z", "left", row, row, z=float(bed32[row, 0] + np.float32(2.0))))
    elif zl == "q":
        zc = float(bed32[lo:hi + 1, 0].min() + np.float32(2.0))
        bcs.append(BCSpec("q", "left", lo, hi, z_series=([0.0, 1e6], [zc, zc]),
                          q_series=([-60.0, 600.0, 3600.0], [0.0, 40.0, 0.0])))
    elif zl:
        bcs.append(BCSpec(zl, "left"))
    if zr:
        bcs.append(BCSpec(zr, "right"))
    if zt:
        bcs.append(BCSpec(zt, "top"))
    if zb:
        bcs.append(BCSpec(zb, "bottom"))
    return Case(name=f"small_{W}x{H}_{scenario}", W=W, H=H, dx=dx, dt=0.1, duration=600.0,
                bed=bed32, kst=30.0, kst_var=kv, h=h, no_data=nd, bcs=bcs)


def _write_raster(path: str, a: np.ndarray) -> None:
    with open(path, "wb") as f:
        f.write(b"FIPRAW32".ljust(HEADER_BYTES, b"\0"))
        f.write(np.ascontiguousarray(a, dtype=np.float32).tobytes())


def _fmt(x: float) -> str:
    return repr(float(x))


def write_case(case: Case, directory: str, sampling=None, sampling_interval: int | None = None,
               save_state=None, extra: dict | None = None) -> str:
    """Writes rasters, CSV series and parameters.yml; returns the yml path."""
    os.makedirs(directory, exist_ok=True)
    d = os.path.abspath(directory)
    lines = [f"name: {os.path.join(d, case.name)}"]
    _write_raster(os.path.join(d, "terrain.bin"), case.bed)
    lines.append(f"terrain: [{os.path.join(d, 'terrain.bin')}, {HEADER_BYTES}]")
    lines += [f"W: {case.W}", f"H: {case.H}", f"dx: {_fmt(case.dx)}", f"dt: {_fmt(case.dt)}",
              f"duration: {_fmt(case.duration)}"]
    if case.kst_var is not None:
        _write_raster(os.path.join(d, "kst.bin"), case.kst_var)
        lines.append(f"kSt_var: [{os.path.join(d, 'kst.bin')}, {HEADER_BYTES}]")
    else:
        lines.append(f"kSt: {_fmt(case.kst)}")
    if case.h is not None:
        z = (case.h.astype(np.float32) + case.bed.astype(np.float32)).astype(np.float32)
        _write_raster(os.path.join(d, "z.bin"), z)
        lines.append(f"z: [{os.path.join(d, 'z.bin')}, {HEADER_BYTES}]")
    for nm, a in (("qx", case.qx), ("qy", case.qy)):
        if a is not None:
            _write_raster(os.path.join(d, f"{nm}.bin"), a)
            lines.append(f"{nm}: [{os.path.join(d, nm + '.bin')}, {HEADER_BYTES}]")
    if case.no_data is not None:
        lines.append(f"no_data: {_fmt(case.no_data)}")
    lines.append(f"variable_dt: {'true' if case.variable_dt else 'false'}")
    if extra:
        for k, v in extra.items():
            lines.append(f"{k}: {v}")
    if save_state:
        lines.append("save_state:")
        for t in save_state:
            lines.append(f"  - {t}")
    if sampling:
        lines.append(f"sampling_interval: {sampling_interval or 60}")
        lines.append("sampling:")
        for sp in sampling:
            lines.append(f"  - {{ x: {sp['x']}, y: {sp['y']}, label: {sp['label']}, type: {sp['type']} }}")
    if case.bcs:
        lines.append("boundary_conditions:")
        for k, bc in enumerate(case.bcs):
            lines.append(f"  - side: {bc.side}")
            lines.append(f"    type: {bc.type}")
            if bc.frm is not None:
                lines.append(f"    from: {bc.frm}")
            if bc.to is not None:
                lines.append(f"    to: {bc.to}")
            for nm, val, ser in (("z", bc.z, bc.z_series), ("q", bc.q, bc.q_series)):
                if ser is not None:
                    p = os.path.join(d, f"bc{k}_{nm}.csv")
                    with open(p, "w") as f:
                        f.write("time;value\n")
                        for t, v in zip(*ser):
                            f.write(f"{_fmt(t)};{_fmt(v)}\n")
                    lines.append(f"    {nm}: [{p}, 0, 1]")
                elif val is not None:
                    lines.append(f"    {nm}: {_fmt(val)}")
    path = os.path.join(d, "parameters.yml")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path
