"""Seeded geometry and initial conditions: building masks, noise fields.

The solid mask follows the reference's ``InitFlags(int[])`` contract
(FvmSolverGpu.cs:972-987, InitShaderColl.compute:345-364): an ``int32`` array of
length X*Y*Z indexed ``z*X*Y + y*X + x`` with 1 = solid.  Buildings are
half-open integer boxes ``[x0, x1) x [0, h) x [z0, z1)`` standing on the ground
(y is up), rasterised like ``CSInitBoxFlag`` (InitShaderColl.compute:302-343).

Random numbers come from SplitMix64 (Steele, Lea & Flood, 2014): state
``s += 0x9E3779B97F4A7C15``, then the two xor-shift-multiply rounds; an
integer in ``[lo, hi]`` is ``lo + next() % (hi - lo + 1)`` and a float in
[0, 1) is ``(next() >> 40) * 2**-24``.  The same stream in the C library's
``uws_init_noise`` gives the same velocity noise bit for bit.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

MASK64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15


class SplitMix64:
    def __init__(self, seed: int):
        self.s = seed & MASK64

    def next(self) -> int:
        self.s = (self.s + GOLDEN) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi]."""
        return lo + self.next() % (hi - lo + 1)

    def uniform(self) -> float:
        return (self.next() >> 40) * (1.0 / 16777216.0)


def splitmix64_stream(seed: int, n: int) -> np.ndarray:
    """The first n outputs of SplitMix64(seed), vectorised."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed & MASK64) + k * np.uint64(GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def noise_velocity(flags: np.ndarray, seed: int, amp: float) -> np.ndarray:
    """u = U(-amp, amp) per component in fluid cells, 0 in solid cells, as
    ``uws_init_noise``: one draw per component of every cell in cell order."""
    n = flags.size
    z = splitmix64_stream(seed, 3 * n)
    u01 = (z >> np.uint64(40)).astype(np.float32) * np.float32(1.0 / 16777216.0)
    r = (np.float32(2.0) * u01 - np.float32(1.0)) * np.float32(amp)
    u = r.reshape(n, 3)
    u[flags.reshape(-1) != 0] = 0.0
    return np.ascontiguousarray(u, dtype=np.float32)


@dataclass(frozen=True)
class Box:
    """Half-open building box: x in [x0, x1), y in [0, h), z in [z0, z1)."""

    x0: int
    x1: int
    z0: int
    z1: int
    h: int


def rasterise(grid, boxes, angle_deg: float = 0.0, centre=None) -> np.ndarray:
    """int32 mask [z][y][x] flattened; with angle != 0 every footprint is
    rotated by -angle about `centre` ((x, z), default the domain centre): the
    inflow direction rotated into the solver frame, whose inflow is always at
    X0.  A cell is solid when its centre lies inside a rotated footprint."""
    X, Y, Z = grid
    m = np.zeros((Z, Y, X), dtype=np.int32)
    if angle_deg % 360.0 == 0.0:
        for b in boxes:
            x0, x1 = max(b.x0, 0), min(b.x1, X)
            z0, z1 = max(b.z0, 0), min(b.z1, Z)
            if x0 < x1 and z0 < z1:
                m[z0:z1, 0:min(b.h, Y), x0:x1] = 1
        return m.reshape(-1)
    th = math.radians(angle_deg)
    c, s = math.cos(th), math.sin(th)
    cx, cz = (X / 2.0, Z / 2.0) if centre is None else (float(centre[0]), float(centre[1]))
    xc = np.arange(X) + 0.5
    zc = np.arange(Z) + 0.5
    XX, ZZ = np.meshgrid(xc, zc)  # [z][x]
    # cell centre mapped back into the un-rotated layout frame
    bx = c * (XX - cx) + s * (ZZ - cz) + cx
    bz = -s * (XX - cx) + c * (ZZ - cz) + cz
    for b in boxes:
        inside = (bx >= b.x0) & (bx < b.x1) & (bz >= b.z0) & (bz < b.z1)
        col = np.zeros((Z, X), dtype=np.int32)
        col[inside] = 1
        m[:, 0:min(b.h, Y), :] |= col[:, None, :]
    return m.reshape(-1)


def lattice_layout(grid, n_buildings: int, lattice, seed: int, foot=(4, 12), height=(4, 24),
                   region=((0.15, 0.85), (0.1, 0.9))) -> list:
    """Buildings on a jittered nx x nz lattice, at most one per lattice cell.

    The lattice spans the x fraction ``region[0]`` and z fraction ``region[1]``
    of the domain; ``n_buildings`` lattice cells are drawn without
    replacement, each building's footprint sides are uniform in ``foot``
    (clamped so it fits its lattice cell with a one-cell gap) and its height is
    uniform in ``height``; the footprint is placed uniformly inside the cell.
    Neighbouring buildings therefore never overlap or touch.
    """
    X, Y, Z = grid
    nx, nz = lattice
    if n_buildings > nx * nz:
        raise ValueError("more buildings than lattice cells")
    rng = SplitMix64(seed)
    xa, xb = int(region[0][0] * X), int(region[0][1] * X)
    za, zb = int(region[1][0] * Z), int(region[1][1] * Z)
    cells = list(range(nx * nz))
    # Fisher-Yates with the seeded stream; the first n cells get a building
    for i in range(len(cells) - 1, 0, -1):
        j = rng.randint(0, i)
        cells[i], cells[j] = cells[j], cells[i]
    boxes = []
    for cell in sorted(cells[:n_buildings]):
        i, k = cell % nx, cell // nx
        cx0 = xa + (xb - xa) * i // nx
        cx1 = xa + (xb - xa) * (i + 1) // nx
        cz0 = za + (zb - za) * k // nz
        cz1 = za + (zb - za) * (k + 1) // nz
        wmax_x = max(1, min(foot[1], cx1 - cx0 - 1))
        wmax_z = max(1, min(foot[1], cz1 - cz0 - 1))
        wx = rng.randint(min(foot[0], wmax_x), wmax_x)
        wz = rng.randint(min(foot[0], wmax_z), wmax_z)
        h = rng.randint(height[0], height[1])
        x0 = cx0 + rng.randint(0, (cx1 - cx0 - 1) - wx)
        z0 = cz0 + rng.randint(0, (cz1 - cz0 - 1) - wz)
        boxes.append(Box(x0, x0 + wx, z0, z0 + wz, h))
    return boxes


def read_flag_file(path, ncells: int) -> np.ndarray:
    """A mask written by the reference's FlagExporter (FlagExporter.cs:13-34):
    the flat index z*X*Y + y*X + x of every solid cell, one per line.  Returns
    the int32 flag array InitFlags takes (1 solid, 0 fluid)."""
    flags = np.zeros(int(ncells), dtype=np.int32)
    with open(path) as fh:
        for line in fh:
            line = line.strip()
            if not line:
                continue
            i = int(line)
            if i < 0 or i >= ncells:
                raise ValueError(f"solid index {i} outside a grid of {ncells} cells")
            flags[i] = 1
    return flags


def write_flag_file(path, flags) -> None:
    """The FlagExporter format: one line per non-zero flag with its flat index."""
    idx = np.flatnonzero(np.asarray(flags).reshape(-1))
    with open(path, "w") as fh:
        fh.write("".join(f"{int(i)}\n" for i in idx))
