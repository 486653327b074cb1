"""MCS acquisition geometry from navigation and SU files, as the reference sets it up
on the host before the inversion (get_mcs_data.f90:89-140 read_nav_shot_mcs,
:142-165 get_geometry_mcs, :376-426 get_bathymetry_model_mcs, :429-519
get_pos_channels).  Arithmetic in REAL(4) like the reference; positions map to
1-based grid indices by 1 + ceiling(x / dmodel).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .lib import akima, su_header_word

f32 = np.float32


def grid(x) -> np.ndarray:
    """1 + ceiling(x / dmodel) for REAL(4) positions already divided by dmodel."""
    return 1 + np.ceil(np.asarray(x, np.float32)).astype(np.int64)


@dataclass
class MCSGeometry:
    dmodel: float
    added_space: float
    shot_id: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    pos_shot: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))       # metres along the model
    pos_shot_grid: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    pos_bat: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))        # water depth at the shot
    pos_bat_grid: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    dshots: float = 0.0
    length_model: float = 0.0
    nxmodel: int = 0
    pos_rec: np.ndarray = field(default_factory=lambda: np.zeros((0, 0), np.float32))  # (nrec, nshots)
    pos_rec_grid: np.ndarray = field(default_factory=lambda: np.zeros((0, 0), np.int64))

    @classmethod
    def from_navigation(cls, path, nshots, dmodel, added_space):
        """read_nav_shot_mcs: lines `shotID xs bat`; shot x = added_space + |xs|, water depth |bat|."""
        rows = []
        with open(path) as fh:
            for line in fh:
                parts = line.split()
                if len(parts) >= 3:
                    rows.append((int(float(parts[0])), f32(parts[1]), f32(parts[2])))
        if len(rows) < nshots:
            raise ValueError(f"navigation holds {len(rows)} shots, {nshots} wanted")
        rows = rows[:nshots]
        d, add = f32(dmodel), f32(added_space)
        g = cls(float(d), float(add))
        g.shot_id = np.array([r[0] for r in rows], np.int64)
        xs = np.array([r[1] for r in rows], np.float32)
        g.pos_shot = (add + np.abs(xs)).astype(np.float32)
        g.pos_shot_grid = grid(g.pos_shot / d)
        g.pos_bat = np.abs(np.array([r[2] for r in rows], np.float32))
        g.pos_bat_grid = grid(g.pos_bat / d)
        g.dshots = float(np.abs(xs[-1] - xs[0]) / f32(nshots - 1)) if nshots > 1 else 0.0
        g.length_model = float(xs[-1] + f32(2.0) * add)
        g.nxmodel = int(grid(f32(g.length_model) / d))
        return g

    def receivers_nominal(self, nrec, near_offset, drec):
        """get_pos_channels with offset_header = 0: channel k at shot x - near_offset - (k-1) drec."""
        k = np.arange(nrec, dtype=np.float32)[:, None]
        self.pos_rec = (self.pos_shot[None, :] - f32(near_offset) - k * f32(drec)).astype(np.float32)
        self.pos_rec_grid = grid(self.pos_rec / f32(self.dmodel))
        return self

    def receivers_from_su(self, su_files, nrec, byte_offset, offset_unit, big_endian=True):
        """get_pos_channels with offset_header = 1: the signed 4-byte offset at 1-based byte byte_offset
        of each trace header, scaled by offset_unit (> 0: multiply, < 0: divide by |offset_unit|)."""
        pos = np.zeros((nrec, len(su_files)), np.float32)
        for s, path in enumerate(su_files):
            for r in range(nrec):
                off = su_header_word(path, r, byte_offset, 4, big_endian)
                if off == 0:
                    raise ValueError(f"{path}: offset 0 in the header of trace {r + 1}")
                o = f32(abs(off)) * f32(offset_unit) if offset_unit > 0 else f32(abs(off)) / f32(abs(offset_unit))
                pos[r, s] = self.pos_shot[s] - o
        self.pos_rec = pos
        self.pos_rec_grid = grid(pos / f32(self.dmodel))
        return self

    def source_receiver_indices(self, shot_depth, streamer_depth):
        """get_geometry_mcs: 1-based grid indices (nxSou, nySou, nxRec (nrec, nshots), nyRec)."""
        d = f32(self.dmodel)
        ny_s = int(grid(f32(shot_depth) / d))
        ny_r = int(grid(f32(streamer_depth) / d))
        return self.pos_shot_grid.copy(), ny_s, self.pos_rec_grid.copy(), ny_r

    def bathymetry(self, nxmodel, length_model):
        """get_bathymetry_model_mcs: water depth under every model column by Akima interpolation of the
        shots' depths (extended flat to x = 0 and x = length_model); returns (metres, grid rows)."""
        n = len(self.pos_shot)
        X = np.zeros(n + 2, np.float32)
        Y = np.zeros(n + 2, np.float32)
        X[1:n + 1] = self.pos_shot
        Y[1:n + 1] = self.pos_bat
        X[0], X[-1] = 0.0, f32(length_model)
        Y[0], Y[-1] = Y[1], Y[n]
        xmodel = np.arange(nxmodel, dtype=np.float32) * f32(self.dmodel)
        bat = akima(X, Y, xmodel)
        return bat, grid(bat / f32(self.dmodel))
