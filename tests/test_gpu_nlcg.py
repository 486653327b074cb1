"""NLCG on the GPU: FP64 Polak-Ribiere beta (drop-in Beta_calculous and the
device-resident form), search direction, initial step scale and the FP32 model
update against numpy, and a short inversion that lowers the misfit.

Tolerances: FP64 sums in a fixed order vs numpy's pairwise sums, 1e-12
relative; the drop-in returns REAL(4), so within one FP32 rounding of the FP64
value; against the reference's REAL(4)-accumulated Beta_calculous 2e-6
relative; the model update (expf, sqrtf in FP32) within 4 FP32 ulps of numpy's
float32 evaluation of the same expression."""
import ctypes as C

import numpy as np
import pytest

import fwi_amd
import refbind as rb
from fwi_amd import synth
from fwi_amd.lib import beta_calculous
from fwi_amd.nlcg import NLCG

pytestmark = pytest.mark.gpu


def _pr(g, go):
    g, go = g.astype(np.float64), go.astype(np.float64)
    return float((g * (g - go)).sum() / (go * go).sum())


def test_beta_dropin():
    rng = np.random.default_rng(1)
    g = np.asfortranarray(rng.standard_normal((150, 400)).astype(np.float32))
    go = np.asfortranarray((0.8 * g + 0.2 * rng.standard_normal((150, 400))).astype(np.float32))
    b = beta_calculous(g, go)
    b64 = _pr(g, go)
    assert b == np.float32(b64) or abs(b - b64) <= 1.5 * np.spacing(np.float32(b64))
    if rb.ref_available("fs1_r4"):
        r = C.c_float(0)
        p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
        rb.ref("fs1_r4").beta_calculous_(p(g), p(go), C.byref(C.c_int(400)), C.byref(C.c_int(150)), C.byref(r))
        assert b == pytest.approx(r.value, rel=2e-6)


def _small_problem(nt=800):
    """Crosswell transmission: sources in column 6, receivers down column 115,
    a +10 % Gaussian anomaly between them; the initial model is the 2000 m/s
    background.  The anomaly delays the direct arrivals by about a fifth of a
    period, well inside the range where the L2 misfit is convex."""
    ny, nx, h, dt = 60, 120, 25.0, 0.002
    yy, xx = np.mgrid[0:ny, 0:nx]
    vt = 2000.0 * (1 + 0.10 * np.exp(-(((xx - 60) / 8.0) ** 2 + ((yy - 30) / 8.0) ** 2)))
    vi = np.full((ny, nx), 2000.0)
    shots, wav = [], []
    for sy in (10, 30, 50):
        shots.append(dict(x0=0, w=nx, src_x=[6], src_y=[sy], rec_x=[115] * ny, rec_y=list(range(1, ny + 1))))
        wav.append(synth.ricker(10.0, nt, dt)[:, None])
    ctx = fwi_amd.Context(ny, nx, np.float32(h), np.float32(dt), nt, 10, dpml=20, free_surface=0)
    ctx.set_shots(shots)
    ctx.set_sources(wav)
    ctx.set_model(vt.astype(np.float32))
    ctx.forward(False)
    ctx.observed_from_synthetic()
    ctx.set_model(vi.astype(np.float32))
    water = np.zeros(nx, np.int32)
    return ctx, vt.astype(np.float32), vi.astype(np.float32), water


def test_device_direction_scale_and_update():
    ctx, vt, vi, water = _small_problem(nt=400)
    ctx.set_base_model()
    ctx.gradient(1.0, water)
    g1 = ctx.get_gradient()
    assert ctx.nlcg_direction(1, first=True) == 0.0
    s1 = ctx.get_search()
    assert np.array_equal(s1, -g1)
    lam0 = (np.float32(1000.0) * (vi * vi)).astype(np.float32)
    ratio = np.abs(s1 / np.log(lam0.astype(np.float64))).max()
    assert ctx.search_scale() == pytest.approx(ratio, rel=1e-12)
    step = 1.0 / 100.0 / ratio
    ctx.apply_step(step, commit=False)
    vp = ctx.get_model()
    lam = (np.exp(s1.astype(np.float32) * np.float32(step)) * lam0).astype(np.float32)
    expect = np.sqrt(lam / np.float32(1000.0)).astype(np.float32)
    ulp = np.abs(vp - expect) / np.spacing(expect)
    assert ulp.max() <= 4
    assert np.abs(vp - vi).max() > 0
    # second gradient at the trial model: Polak-Ribiere direction
    ctx.gradient(1.0, water)
    g2 = ctx.get_gradient()
    beta = ctx.nlcg_direction(1, first=False)
    assert beta == pytest.approx(_pr(g2, g1), rel=1e-12)
    s2 = ctx.get_search()
    expect_s = -g2 + beta * s1 if beta > 0 else -g2
    assert np.allclose(s2, expect_s, rtol=1e-12, atol=1e-12 * np.abs(expect_s).max())
    # steepest descent ignores beta
    ctx.gradient(1.0, water)
    assert ctx.nlcg_direction(0, first=False) == 0.0
    assert np.array_equal(ctx.get_search(), -ctx.get_gradient())


def test_nlcg_iterations_lower_misfit():
    ctx, vt, vi, water = _small_problem()
    opt = NLCG(ctx, water_rows=water)
    recs = [opt.iterate() for _ in range(3)]
    ctx.gradient(1.0, water)
    final = ctx.misfit()
    print("\n" + "\n".join(f"it {r['it']}: misfit {r['misfit']:.6e} beta {r['beta']:.3f} step {r['step']:.3e} "
                           f"trials {r['trials']}" for r in recs) + f"\nfinal misfit {final:.6e}")
    assert all(r["accepted"] for r in recs)
    assert final < 0.8 * recs[0]["misfit"]
    err0 = np.sqrt(((vi - vt) ** 2).mean())
    err = np.sqrt(((ctx.get_model() - vt) ** 2).mean())
    print(f"model RMS error {err0:.3f} -> {err:.3f} m/s")
    assert err < err0
