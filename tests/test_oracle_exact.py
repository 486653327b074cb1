"""The oracle's exact discrete adjoint pair (no counterpart in the reference).

Pinned by mathematics only:
* its forward is the reference scheme: bit-identical to the C++ restatement
  of solver_acoustic_SS (itself bit-identical to the reference build), free
  surface on and off, thin and standard frames, receivers in the frame,
  a duplicate source cell;
* dot-product test in FP64: <F s, r> = <s, F^T r> for random N(0,1) source
  traces s and receiver traces r, to 1e-12 relative;
* the velocity gradient of J = dt/2 sum (d - d_obs)^2 matches central finite
  differences along a smooth (sigma 300 m) relative perturbation, with the
  error falling as eps^2 from eps = 1e-3 to 1e-5 (FP64)."""
import numpy as np
import pytest
from scipy.ndimage import gaussian_filter

import refbind as rb
from fwi_amd import synth


def _acq(nx, ny, dpml):
    nxs, nys = np.array([10, 30, 10]), np.array([5, 20, 5])      # first and last share a cell
    nxr = np.array([3, 15, 40, -2, nx + dpml + 4, 22])
    nyr = np.array([2, 28, 10, -dpml - 3, ny + 3, 1])
    return nxs, nys, nxr, nyr


@pytest.mark.parametrize("fs,dpml", [(0, 5), (0, 20), (1, 5), (1, 20)])
def test_exact_forward_is_the_reference_scheme(fs, dpml):
    rng = np.random.default_rng(fs * 10 + dpml)
    ny, nx, nt = 30, 44, 160
    c = (1800 + 900 * rng.random((ny, nx))).astype(np.float32)
    nxs, nys, nxr, nyr = _acq(nx, ny, dpml)
    S = rng.standard_normal((nt, 3)).astype(np.float32)
    d, dt = np.float32(20.0), np.float32(0.0015)
    O = rb.oracle_solver("f32", 3, nxs, nys, nxr, nyr, S, None, d, dt, c, 1, 0, nt, dpml=dpml, free_surface=fs)
    R, _ = rb.oracle_exact_forward("f32", nxs, nys, nxr, nyr, S, d, dt, c, nt, dpml=dpml, free_surface=fs)
    assert np.abs(O.rec).max() > 0
    assert np.array_equal(O.rec, R)


@pytest.mark.parametrize("fs,dpml", [(0, 5), (0, 20), (1, 5), (1, 20)])
def test_exact_adjoint_dot_product(fs, dpml):
    rng = np.random.default_rng(100 + fs * 10 + dpml)
    ny, nx, nt = 30, 44, 160
    c = 1800 + 900 * rng.random((ny, nx))
    nxs, nys, nxr, nyr = _acq(nx, ny, dpml)
    s = rng.standard_normal((nt, 3))
    r = rng.standard_normal((nt, len(nxr)))
    F, _ = rb.oracle_exact_forward("f64", nxs, nys, nxr, nyr, s, 20.0, 0.0015, c, nt, dpml=dpml, free_surface=fs)
    Ft, _ = rb.oracle_exact_adjoint("f64", nxs, nys, nxr, nyr, r, 20.0, 0.0015, c, nt, dpml=dpml, free_surface=fs)
    lhs, rhs = float((F * r).sum()), float((s * Ft).sum())
    assert abs(lhs) > 1e-3
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    # the second of the two sources on a shared cell is overridden (last wins): no sensitivity
    assert np.all(Ft[:, 0] == 0) and np.abs(Ft[:, 2]).max() > 0


def _fd_case():
    ny, nx, nt, d, dt, dpml = 30, 44, 300, 20.0, 0.0015, 10
    yy, xx = np.mgrid[0:ny, 0:nx]
    vt = 2000 + 300 * np.exp(-(((xx - 22) / 5.0) ** 2 + ((yy - 15) / 5.0) ** 2))
    vi = np.full((ny, nx), 2000.0)
    nxs, nys = np.array([4, 40]), np.array([15, 8])
    nxr = np.arange(1, nx + 1, 3)
    nyr = np.full(len(nxr), 2)
    w = synth.ricker(15.0, nt, dt).astype(np.float64)
    S = np.stack([w, 0.7 * w], 1)
    return ny, nx, nt, d, dt, dpml, vt, vi, (nxs, nys, nxr, nyr), S


@pytest.mark.slow
def test_exact_gradient_matches_finite_differences():
    ny, nx, nt, d, dt, dpml, vt, vi, acq, S = _fd_case()
    obs, _ = rb.oracle_exact_forward("f64", *acq, S, d, dt, vt, nt, dpml=dpml, free_surface=0)

    def J(v):
        r, _ = rb.oracle_exact_forward("f64", *acq, S, d, dt, v, nt, dpml=dpml, free_surface=0)
        return 0.5 * dt * ((r - obs) ** 2).sum()

    rec, hist = rb.oracle_exact_forward("f64", *acq, S, d, dt, vi, nt, dpml=dpml, free_surface=0, history=True)
    _, g = rb.oracle_exact_adjoint("f64", *acq, dt * (rec - obs), d, dt, vi, nt, hist=hist, dpml=dpml,
                                   free_surface=0)
    dv = gaussian_filter(np.random.default_rng(1).standard_normal((ny, nx)), 300.0 / d / 2)
    dv /= np.abs(dv).max()
    gd = float((g * dv * vi).sum())
    errs = []
    for eps in (1e-3, 1e-4, 1e-5, 1e-6):
        fd = (J(vi * (1 + eps * dv)) - J(vi * (1 - eps * dv))) / (2 * eps)
        errs.append(abs(fd - gd) / abs(gd))
    print("\nFD relative errors at eps 1e-3..1e-6:", " ".join(f"{e:.1e}" for e in errs))
    assert errs[0] < 1e-5 and errs[1] < errs[0] / 50 and errs[2] < errs[1] / 50 and errs[3] < 1e-8
