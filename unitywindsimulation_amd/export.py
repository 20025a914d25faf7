"""Flow-field export in the layout of the reference's netCDF exporter
(``Assets/Code/IO/NcFlowFieldExporter.cs:128-271``).

The reference writes a netCDF-3 classic file (``nc_create`` with NC_CLOBBER)
with dimensions ``time`` (unlimited), ``velX velY velZ``, ``velVec`` (4),
``presX presY presZ``; global attributes ``gridLayoutType`` (text) and ``dt``
(float); float coordinate variables at (i + 1/2) dx (``velVec`` holds
0 1 2 3); and the record variables ``time(time)``,
``velField(time, velZ, velY, velX, velVec)`` and
``presField(time, velZ..)`` = ``presField(time, presZ, presY, presX)``,
appended one frame at a time by a writer thread fed from a queue
(``NcFlowFieldExporter.cs:274-392``).  No netCDF library is available here,
so this module writes the classic format itself (big-endian header, fixed
variables, then interleaved records; the record count in the header is
updated after every frame).  Frames are enqueued from a solver
(``enqueue_field``) and written by a background thread, as in the reference.
"""
from __future__ import annotations

import queue
import struct
import threading
from pathlib import Path

import numpy as np

NC_DIMENSION, NC_VARIABLE, NC_ATTRIBUTE = 0x0A, 0x0B, 0x0C
NC_CHAR, NC_INT, NC_FLOAT = 2, 4, 5


def _pad4(n: int) -> int:
    return (n + 3) & ~3


def _name(s: str) -> bytes:
    b = s.encode()
    return struct.pack(">i", len(b)) + b + b"\0" * (_pad4(len(b)) - len(b))


class NcFlowFieldExporter:
    """Streams frames of (time, u, p) into a netCDF-3 classic file."""

    def __init__(self, path, cf, save_vel: bool = True, save_pres: bool = True, threaded: bool = True):
        cf.Init()
        self.path = Path(path)
        self.X, self.Y, self.Z = cf.gridRes
        self.save_vel, self.save_pres = save_vel, save_pres
        self.num_steps_saved = 0
        dx = float(cf.dx)
        grid_type = {0: "Collocated", 1: "CollNonUniform", 2: "Staggered"}[int(cf.gridType)]
        X, Y, Z = self.X, self.Y, self.Z
        dims = [("time", 0), ("velX", X), ("velY", Y), ("velZ", Z), ("velVec", 4), ("presX", X), ("presY", Y),
                ("presZ", Z)]
        did = {n: i for i, (n, _) in enumerate(dims)}
        coords = {
            "velX": (np.arange(X, dtype=np.float32) + np.float32(0.5)) * np.float32(dx),
            "velY": (np.arange(Y, dtype=np.float32) + np.float32(0.5)) * np.float32(dx),
            "velZ": (np.arange(Z, dtype=np.float32) + np.float32(0.5)) * np.float32(dx),
            "velVec": np.arange(4, dtype=np.float32),
            "presX": (np.arange(X, dtype=np.float32) + np.float32(0.5)) * np.float32(dx),
            "presY": (np.arange(Y, dtype=np.float32) + np.float32(0.5)) * np.float32(dx),
            "presZ": (np.arange(Z, dtype=np.float32) + np.float32(0.5)) * np.float32(dx),
        }
        # variables in definition order: (name, dim names, record?)
        self._vars = [("time", ["time"], True)] + [(k, [k], False) for k in coords]
        self._vars += [("velField", ["time", "velZ", "velY", "velX", "velVec"], True),
                       ("presField", ["time", "presZ", "presY", "presX"], True)]
        dimlen = dict(dims)

        def vsize(names):
            n = 4
            for d in names:
                if d != "time":
                    n *= dimlen[d]
            return _pad4(n)

        def header(begins):
            h = b"CDF\x01" + struct.pack(">i", self.num_steps_saved)
            h += struct.pack(">ii", NC_DIMENSION, len(dims))
            for n, ln in dims:
                h += _name(n) + struct.pack(">i", ln)
            gat = grid_type.encode()
            h += struct.pack(">ii", NC_ATTRIBUTE, 2)
            h += _name("gridLayoutType") + struct.pack(">ii", NC_CHAR, len(gat)) + gat + b"\0" * (_pad4(len(gat)) - len(gat))
            h += _name("dt") + struct.pack(">ii", NC_FLOAT, 1) + struct.pack(">f", float(cf.dt))
            h += struct.pack(">ii", NC_VARIABLE, len(self._vars))
            for (n, dn, _), b in zip(self._vars, begins):
                h += _name(n) + struct.pack(">i", len(dn)) + b"".join(struct.pack(">i", did[d]) for d in dn)
                h += struct.pack(">ii", 0, 0)  # no variable attributes
                h += struct.pack(">iii", NC_FLOAT, vsize(dn), b)
            return h

        hlen = len(header([0] * len(self._vars)))
        begins, off = [], hlen
        for n, dn, rec in self._vars:  # fixed variables first
            if not rec:
                begins.append(off)
                off += vsize(dn)
            else:
                begins.append(None)
        self._recsize = 0
        for i, (n, dn, rec) in enumerate(self._vars):
            if rec:
                begins[i] = off + self._recsize
                self._recsize += vsize(dn)
        self._begins = {n: b for (n, _, _), b in zip(self._vars, begins)}
        self._rec_start = off
        self._header = header
        self._begin_list = begins
        self._fh = open(self.path, "wb+")
        self._fh.write(header(begins))
        for n, _, rec in self._vars:
            if not rec:
                self._fh.write(coords[n].astype(">f4").tobytes())
        self._fh.flush()
        self._q: queue.Queue = queue.Queue()
        self._thread = None
        if threaded:
            self._thread = threading.Thread(target=self._writer_loop, daemon=True)
            self._thread.start()

    # ---- writing ----
    def write_frame(self, t: float, vel=None, pres=None) -> None:
        """Append one record: vel [z][y][x][4] (or [..][3], w = 0 is added), pres [z][y][x]."""
        r = self.num_steps_saved
        base = self._rec_start + r * self._recsize
        rec = bytearray(self._recsize)
        o = 0
        for n, dn, isrec in self._vars:
            if not isrec:
                continue
            if n == "time":
                data = np.asarray([t], dtype=">f4")
            elif n == "velField":
                if vel is None or not self.save_vel:
                    data = np.zeros(self.Z * self.Y * self.X * 4, dtype=">f4")
                else:
                    v = np.asarray(vel, dtype=np.float32).reshape(self.Z * self.Y * self.X, -1)
                    if v.shape[1] == 3:
                        v = np.concatenate([v, np.zeros((v.shape[0], 1), np.float32)], axis=1)
                    data = v.astype(">f4").reshape(-1)
            else:
                if pres is None or not self.save_pres:
                    data = np.zeros(self.Z * self.Y * self.X, dtype=">f4")
                else:
                    data = np.asarray(pres, dtype=np.float32).astype(">f4").reshape(-1)
            b = data.tobytes()
            rec[o:o + len(b)] = b
            o += _pad4(len(b))
        self._fh.seek(base)
        self._fh.write(bytes(rec))
        self.num_steps_saved = r + 1
        self._fh.seek(4)
        self._fh.write(struct.pack(">i", self.num_steps_saved))
        self._fh.flush()

    def enqueue_field(self, solver) -> None:
        """EnqueueField (NcFlowFieldExporter.cs:295-336): read u and p of the
        solver now and hand the frame to the writer thread."""
        cf = solver.cf
        vel = solver.velocity(4) if self.save_vel else None
        pres = solver.pressure() if self.save_pres else None
        item = (float(np.float32(cf.currPhyTime)), vel, pres)
        if self._thread is None:
            self.write_frame(*item)
        else:
            self._q.put(item)

    def _writer_loop(self):
        while True:
            item = self._q.get()
            if item is None:
                self._q.task_done()
                return
            self.write_frame(*item)
            self._q.task_done()

    def close(self) -> None:
        """CloseNcFile: wait for the queue to drain, then close the file."""
        if self._thread is not None:
            self._q.put(None)
            self._thread.join()
            self._thread = None
        if self._fh:
            self._fh.close()
            self._fh = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
