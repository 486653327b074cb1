"""Synthetic acquisitions for the parity case and the throughput workload.

The reference ships no data (its README points at datasets that are not in
the tree), so these generators are this package's own, seeded.  Grid
conventions follow the reference: model arrays are (ny, nx) with ny rows of
depth, positions are 1-based, a depth or offset in metres maps to index
1 + ceil(x / h) (get_mcs_data.f90:114,150,154), and the source wavelet is the
reference's Ricker (get_mcs_data.f90:331-362).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


def grid_index(x_m: float, h: float) -> int:
    """Metres -> 1-based grid index, 1 + ceiling(x/h) (get_mcs_data.f90:114)."""
    return 1 + int(math.ceil(np.float32(x_m) / np.float32(h)))


def ricker(f0: float, nt: int, dt: float) -> np.ndarray:
    """Ricker wavelet as Ricker_mcs builds it (get_mcs_data.f90:339-362), in REAL(4):
    S0 = (1 - 2 (pi f0 (t - t0))^2) exp(-(pi f0 (t - t0))^2), t0 = 1.5/f0, zeroed for t >= 10/f0."""
    f32 = np.float32
    pi = f32(3.14159265)
    f0 = f32(f0)
    t0 = f32(1.5) / f0
    tmp = f32(10.0) / f0
    t = np.arange(nt, dtype=np.float32) * f32(dt)
    a = (pi * f0) * (t - t0)
    a2 = a * a
    s = (f32(1.0) - f32(2.0) * a2) * np.exp(f32(-1.0) * a2)
    s[t >= tmp] = 0.0
    return s.astype(np.float32)


def _smooth(a, sigma):
    from scipy.ndimage import gaussian_filter
    return gaussian_filter(a, sigma, mode="nearest")


@dataclass
class Acquisition:
    """A model plus shots.  Shot positions are 1-based window coordinates."""
    name: str
    ny: int
    nx: int
    h: float
    dt: float
    nt: int
    vp_true: np.ndarray
    vp_init: np.ndarray
    water_rows: np.ndarray            # per column: rows 1..water_rows[i] are not inverted
    shots: list = field(default_factory=list)   # dicts: x0, w, src_x, src_y, rec_x, rec_y, f0
    wavelets: list = field(default_factory=list)  # per shot (nt, nsrc) float32
    free_surface: int = 0
    dpml: int = 20
    store_snap: int = 10
    notes: dict = field(default_factory=dict)

    def cell_updates(self, sweeps: int = 2) -> int:
        """Credited cell-updates: padded window cells x (nt-1) steps x sweeps, summed over shots."""
        dn = self.dpml + 6
        return sum((s["w"] + 2 * dn) * (self.ny + 2 * dn) for s in self.shots) * (self.nt - 1) * sweeps


def single_shot_400x150(free_surface: int = 0) -> Acquisition:
    """Single-shot parity case: nx=400, nz=150, h=25 m, dt=2 ms, nt=2000.

    Water (1500 m/s) down to 2 km, then a linear 1800 -> 4500 m/s gradient to
    the bottom; a +8 % Gaussian anomaly (sigma 500 m) at the model centre,
    applied to the whole model (the centre lies in the water; recorded
    choice).  Initial model: the true model smoothed with sigma 500 m, water
    reset to 1500 m/s.  One 8 Hz Ricker at 10 m depth, x = 2.5 km; 300
    receivers at 10 m depth every 25 m starting at x = 1250 m (recorded
    choice, centred spread)."""
    ny, nx, h, dt, nt = 150, 400, 25.0, 0.002, 2000
    z = np.arange(ny) * h
    x = np.arange(nx) * h
    zb = z[-1]
    v1 = np.where(z < 2000.0, 1500.0, 1800.0 + (z - 2000.0) / (zb - 2000.0) * 2700.0)
    vp = np.repeat(v1[:, None], nx, axis=1)
    xc, zc = x[-1] / 2, z[-1] / 2
    r2 = (x[None, :] - xc) ** 2 + (z[:, None] - zc) ** 2
    vt = vp * (1.0 + 0.08 * np.exp(-r2 / (2 * 500.0 ** 2)))
    water = z < 2000.0
    vi = _smooth(vt, 500.0 / h)
    vi[water, :] = 1500.0
    wr = np.full(nx, grid_index(2000.0, h), np.int32)     # rows 1..81 (calculate_misfit_gradient.f90:345-351)
    sy = grid_index(10.0, h)
    shot = dict(x0=0, w=nx, src_x=[grid_index(2500.0, h)], src_y=[sy],
                rec_x=[grid_index(1250.0 + 25.0 * i, h) for i in range(300)], rec_y=[sy] * 300, f0=8.0)
    acq = Acquisition("single_shot_400x150", ny, nx, h, dt, nt, vt.astype(np.float32), vi.astype(np.float32), wr,
                      [shot], [ricker(8.0, nt, dt)[:, None]], free_surface=free_surface)
    acq.notes = {"anomaly_centre_m": (float(xc), float(zc)), "receiver_start_m": 1250.0}
    return acq


def marine_layered(name: str, ny: int, nx: int, h: float, dt: float, nt: int, nshots: int, wmin: int, wmax: int,
                   seed: int = 0) -> Acquisition:
    """Layered marine model on an (ny, nx) grid of spacing h with OBS-style shots.

    Water (1500 m/s) over a smooth random seafloor at 1.5-3 km, then 5 layers
    at 2000, 2800, 3800, 5000, 6500 m/s with smooth random interfaces and
    smooth +-5 % perturbations; smoothing lengths are fixed in metres, so the
    model is the same geology on any grid.  Each shot: Ricker with
    f0 ~ U(6,10) Hz at 10 m depth, aperture window W ~ U[wmin, wmax] cells at a
    random position, source at a random point of the central half of its
    window, one receiver per window column at the seafloor cell."""
    rng = np.random.default_rng(seed)
    k = 25.0 / h                                              # smoothing lengths given in 25 m cells

    def smooth_curve(lo, hi, corr):
        c = _smooth(rng.standard_normal(nx), corr * k)
        c = (c - c.min()) / max(c.max() - c.min(), 1e-12)
        return lo + (hi - lo) * c

    seafloor = smooth_curve(1500.0, 3000.0, 150.0)            # metres
    vel = [2000.0, 2800.0, 3800.0, 5000.0, 6500.0]
    z = np.arange(ny)[:, None] * h
    vp = np.full((ny, nx), 1500.0)
    base = seafloor.copy()
    thick = [900.0, 1300.0, 1700.0, 2200.0]
    bounds = [seafloor]
    for th in thick:
        base = base + th * (0.8 + 0.4 * smooth_curve(0.0, 1.0, 100.0))
        bounds.append(base.copy())
    for li, v in enumerate(vel):
        top = bounds[li][None, :]
        vp = np.where(z >= top, v, vp)
    pert = _smooth(rng.standard_normal((ny, nx)), 8.0 * k)
    pert = 0.05 * pert / np.abs(pert).max()
    sub = z >= seafloor[None, :]
    vp = np.where(sub, vp * (1.0 + pert), vp)
    vtrue = vp.astype(np.float32)
    vinit = _smooth(vp, 20.0 * k)
    vinit = np.where(sub, vinit, 1500.0).astype(np.float32)
    wr = np.array([grid_index(seafloor[i], h) for i in range(nx)], np.int32)

    shots, wavelets = [], []
    sy = grid_index(10.0, h)
    for _ in range(nshots):
        w = int(rng.integers(wmin, wmax + 1))
        x0 = int(rng.integers(0, nx - w + 1))
        f0 = float(rng.uniform(6.0, 10.0))
        sx = int(round(rng.uniform(0.25, 0.75) * w)) + 1
        rx = np.arange(1, w + 1)
        ry = wr[x0:x0 + w]
        shots.append(dict(x0=x0, w=w, src_x=[sx], src_y=[sy], rec_x=rx.tolist(), rec_y=ry.tolist(), f0=f0))
        wavelets.append(ricker(f0, nt, dt)[:, None])
    acq = Acquisition(name, ny, nx, h, dt, nt, vtrue, vinit, wr, shots, wavelets, free_surface=0)
    acq.notes = {"seed": seed, "vmax": float(vtrue.max()), "courant": float(vtrue.max()) * dt / h}
    return acq


def marine_64shot_2000x400(seed: int = 0, nshots: int = 64, nt: int = 5000) -> Acquisition:
    """Batched-shot workload: nx=2000, nz=400, h=25 m, dt=2 ms, nt=5000, 64 shots
    (marine_layered; Courant number <= 6825*0.002/25 = 0.546), windows of
    600-1600 cells."""
    return marine_layered("marine_64shot_2000x400", 400, 2000, 25.0, 0.002, nt, nshots, 600, 1600, seed)


def checkpoint_6000x1000(seed: int = 0, nshots: int = 16, nt: int = 15000) -> Acquisition:
    """Checkpointed-adjoint workload: nx=6000, nz=1000, h=10 m, dt=0.8 ms,
    nt=15000 (Courant number 0.546 as above).  The model is not specified
    beyond the grid; recorded choice: the layered marine model on this grid,
    windows of 1500-3000 cells (15-30 km)."""
    return marine_layered("checkpoint_6000x1000", 1000, 6000, 10.0, 0.0008, nt, nshots, 1500, 3000, seed)


def precision_4000x800(nshots: int = 32, nt: int = 10000, window: int = 1200) -> Acquisition:
    """Precision workload: the single-shot model (single_shot_400x150) on a finer,
    wider grid -- nx=4000, nz=800, h=12.5 m, dt=1 ms, nt=10000.  Water to 2 km,
    then 1800 -> 4500 m/s to the bottom, +8 % Gaussian anomaly (sigma 500 m)
    at the model centre; initial model smoothed with sigma 500 m, water kept.
    Recorded choices for the 32 shots: evenly spaced windows of `window`
    columns (15 km), an 8 Hz Ricker at 10 m depth in the window centre and
    one receiver per window column at 10 m depth."""
    ny, nx, h, dt = 800, 4000, 12.5, 0.001
    z = np.arange(ny) * h
    x = np.arange(nx) * h
    zb = z[-1]
    v1 = np.where(z < 2000.0, 1500.0, 1800.0 + (z - 2000.0) / (zb - 2000.0) * 2700.0)
    vp = np.repeat(v1[:, None], nx, axis=1)
    xc, zc = x[-1] / 2, z[-1] / 2
    r2 = (x[None, :] - xc) ** 2 + (z[:, None] - zc) ** 2
    vt = vp * (1.0 + 0.08 * np.exp(-r2 / (2 * 500.0 ** 2)))
    water = z < 2000.0
    vi = _smooth(vt, 500.0 / h)
    vi[water, :] = 1500.0
    wr = np.full(nx, grid_index(2000.0, h), np.int32)
    sy = grid_index(10.0, h)
    shots, wavelets = [], []
    starts = np.linspace(0, nx - window, nshots).round().astype(int) if nshots > 1 else [0]
    for x0 in starts:
        shots.append(dict(x0=int(x0), w=window, src_x=[window // 2 + 1], src_y=[sy],
                          rec_x=list(range(1, window + 1)), rec_y=[sy] * window, f0=8.0))
        wavelets.append(ricker(8.0, nt, dt)[:, None])
    acq = Acquisition("precision_4000x800", ny, nx, h, dt, nt, vt.astype(np.float32), vi.astype(np.float32), wr,
                      shots, wavelets, free_surface=0, store_snap=4)
    return acq


def small_case(seed: int = 0, ny: int = 40, nx: int = 60, nt: int = 200):
    """Tiny layered model with a few sources and receivers, for quick parity checks."""
    rng = np.random.default_rng(seed)
    c = np.full((ny, nx), 1500.0, np.float32)
    c[ny * 3 // 8:, :] = 2500.0
    c[ny * 5 // 8:, :] = 3200.0
    c = (c * (1 + 0.05 * rng.standard_normal((ny, nx)))).astype(np.float32)
    w = ricker(10.0, nt, 0.002)
    S = np.stack([w, 0.5 * w], 1).astype(np.float32)
    nxs = np.array([nx // 3, 2 * nx // 3])
    nys = np.array([2, 3])
    nxr = np.arange(5, nx - 4, 8)
    nyr = np.full(len(nxr), 2)
    nyr[len(nyr) // 2] = ny - 10
    return c, S, nxs, nys, nxr, nyr
