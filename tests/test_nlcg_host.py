"""Host side of the NLCG iteration: the quadratic line-search fit (polyfit.f90)
and the Polak-Ribiere beta of the reference (Beta_calculous.f90).

polyfit.f90 includes mpif.h, so it is not in the reference build (oracle/_ref);
it is pinned here by closed forms: a quadratic through three points is
reproduced exactly, and the returned step is the grid point nearest to the
parabola's vertex (or the interval end when the vertex lies outside).  The
reference's compiled Beta_calculous (REAL(4) accumulation) is compared with
the FP64 Polak-Ribiere formula the GPU uses."""
import ctypes as C

import numpy as np
import pytest

import refbind as rb
from fwi_amd.nlcg import polyfit


@pytest.mark.parametrize("loopind", [2, 3, 5])
@pytest.mark.parametrize("seed", range(6))
def test_polyfit_vertex(loopind, seed):
    rng = np.random.default_rng(seed)
    nut = float(np.float32(10 ** rng.uniform(-4, 1)))
    a = rng.uniform(0.5, 3.0)
    x0 = rng.uniform(-0.5, 1.5) * nut
    c = rng.uniform(1.0, 5.0)
    f = lambda x: a * (x - x0) ** 2 / nut ** 2 + c  # noqa: E731
    lo = 0.0 if loopind == 2 else 0.25 * nut
    pts = [lo, 0.5 * nut, nut]
    step, expected = polyfit(nut, *[f(float(np.float32(p))) for p in pts], loopind)
    grid = np.linspace(lo, nut, 2001)
    h = grid[1] - grid[0]
    target = min(max(x0, lo), nut)
    assert abs(step - target) <= 0.5 * h + 1e-6 * nut
    assert expected == pytest.approx(f(step), rel=1e-9, abs=1e-12)


def test_polyfit_below_two_trials_leaves_step():
    assert polyfit(1.0, 1.0, 0.5, 0.7, 1) == (None, None)


def test_polyfit_last_minimum_wins():
    # a constant misfit: every grid point attains the minimum; the reference's
    # loop keeps overwriting EndStep, so the last point (nut) wins
    step, _ = polyfit(2.0, 3.0, 3.0, 3.0, 2)
    assert step == 2.0


@pytest.mark.skipif(not rb.ref_available("fs1_r4"), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("variant", ["fs1_r4", "fs1_r8"])
def test_reference_beta_matches_fp64_formula(variant):
    rng = np.random.default_rng(3)
    ny, nx = 37, 53
    dt = np.float64 if variant.endswith("r8") else np.float32
    g = np.asfortranarray(rng.standard_normal((ny, nx)).astype(dt))
    go = np.asfortranarray((0.7 * g + 0.3 * rng.standard_normal((ny, nx))).astype(dt))
    b = (C.c_double if dt == np.float64 else C.c_float)(0)
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    rb.ref(variant).beta_calculous_(p(g), p(go), C.byref(C.c_int(nx)), C.byref(C.c_int(ny)), C.byref(b))
    g64, go64 = g.astype(np.float64), go.astype(np.float64)
    beta64 = float((g64 * (g64 - go64)).sum() / (go64 * go64).sum())
    tol = 1e-13 if dt == np.float64 else 2e-6
    assert b.value == pytest.approx(beta64, rel=tol)


def test_get_f1f2_bands():
    """Band edges of the frequency loop (auxiliar.f90:223-261): growing and sliding bands."""
    from fwi_amd.multiscale import get_f1f2
    assert [get_f1f2(i, 3.0, 2.0, 4, 1) for i in (2, 3, 4)] == [(3.0, 5.0), (3.0, 7.0), (3.0, 9.0)]
    assert [get_f1f2(i, 3.0, 2.0, 4, 2) for i in (2, 3, 4)] == [(3.0, 5.0), (5.0, 7.0), (7.0, 9.0)]
    with pytest.raises(ValueError):
        get_f1f2(2, 3.0, 2.0, 4, 3)
    with pytest.raises(ValueError):
        get_f1f2(1, 3.0, 2.0, 4, 1)
