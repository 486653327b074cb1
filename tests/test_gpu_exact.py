"""Exact discrete adjoint on the GPU (validation mode).

* The GPU transpose and gradient follow the oracle's exact_adjoint operation
  by operation: source adjoints and velocity gradients are bit-identical to
  the oracle's FP32 run (tests/test_oracle_exact.py pins the oracle by the
  dot-product test and finite differences in FP64).
* Dot-product test of the GPU pair in FP32: <F s, r> = <s, F^T r> to 1e-4
  relative (FP32 rounding over the sweep with N(0,1) data injected at every
  step; 2e-7 to 2e-5 measured; the same pair in FP64 holds to 1e-12).
* Distance of the GPU FP32 gradient from the oracle's FP64 gradient is stated
  and bounded (1e-3 relative L2, cosine >= 0.9999).
* Several shots with uneven windows in one context: the gradient is the sum of
  the per-shot oracle gradients, each on its window of the model."""
import numpy as np
import pytest

import fwi_amd
import refbind as rb
from fwi_amd import synth

pytestmark = pytest.mark.gpu


def _ctx(ny, nx, d, dt, nt, shots, wav, vp, dpml, fs):
    ctx = fwi_amd.Context(ny, nx, np.float32(d), np.float32(dt), nt, 10, dpml=dpml, free_surface=fs)
    ctx.set_shots(shots)
    ctx.set_sources(wav)
    ctx.set_model(vp.astype(np.float32))
    return ctx


def _case(seed=0, ny=30, nx=44, nt=160):
    rng = np.random.default_rng(seed)
    c = (1800 + 900 * rng.random((ny, nx))).astype(np.float32)
    nxs, nys = np.array([10, 30, 10]), np.array([5, 20, 5])
    nxr = np.array([3, 15, 40, 22, 8, 1])
    nyr = np.array([2, 28, 10, 1, 14, 30])
    return rng, c, nxs, nys, nxr, nyr


@pytest.mark.parametrize("fs,dpml", [(0, 5), (0, 20), (1, 20)])
def test_transpose_matches_oracle_and_passes_dot_product(fs, dpml):
    rng, c, nxs, nys, nxr, nyr = _case(fs * 10 + dpml)
    ny, nx = c.shape
    nt, d, dt = 160, 20.0, 0.0015
    shot = dict(x0=0, w=nx, src_x=nxs.tolist(), src_y=nys.tolist(), rec_x=nxr.tolist(), rec_y=nyr.tolist())
    s = rng.standard_normal((nt, 3)).astype(np.float32)
    r = rng.standard_normal((nt, len(nxr))).astype(np.float32)
    ctx = _ctx(ny, nx, d, dt, nt, [shot], [s], c, dpml, fs)
    ctx.forward(False)
    Fs = ctx.synthetic()[0]
    Ftr = ctx.exact_transpose([r])[0]
    O, _ = rb.oracle_exact_adjoint("f32", nxs, nys, nxr, nyr, r, np.float32(d), np.float32(dt), c, nt, dpml=dpml,
                                   free_surface=fs)
    assert np.abs(O).max() > 0
    assert np.array_equal(Ftr, O)
    lhs = float((Fs.astype(np.float64) * r).sum())
    rhs = float((s.astype(np.float64) * Ftr).sum())
    print(f"\nGPU dot-product test fs={fs} dpml={dpml}: <Fs,r>={lhs:.9e} <s,F^T r>={rhs:.9e} "
          f"rel {abs(lhs - rhs) / abs(lhs):.1e}")
    assert abs(lhs - rhs) <= 1e-4 * abs(lhs)


def _fd_model():
    ny, nx, nt, d, dt, dpml = 30, 44, 300, 20.0, 0.0015, 10
    yy, xx = np.mgrid[0:ny, 0:nx]
    vt = (2000 + 300 * np.exp(-(((xx - 22) / 5.0) ** 2 + ((yy - 15) / 5.0) ** 2))).astype(np.float32)
    vi = np.full((ny, nx), 2000.0, np.float32)
    w = synth.ricker(15.0, nt, dt)
    return ny, nx, nt, d, dt, dpml, vt, vi, w


@pytest.mark.parametrize("fs", [0, 1])
def test_exact_gradient_matches_oracle(fs):
    ny, nx, nt, d, dt, dpml, vt, vi, w = _fd_model()
    nxs, nys = np.array([4, 40]), np.array([15, 8])
    nxr = np.arange(1, nx + 1, 3)
    nyr = np.full(len(nxr), 2)
    S = np.stack([w, 0.7 * w], 1).astype(np.float32)
    shot = dict(x0=0, w=nx, src_x=nxs.tolist(), src_y=nys.tolist(), rec_x=nxr.tolist(), rec_y=nyr.tolist())
    ctx = _ctx(ny, nx, d, dt, nt, [shot], [S], vt, dpml, fs)
    ctx.forward(False)
    ctx.observed_from_synthetic()
    obs = ctx.synthetic()[0]
    ctx.set_model(vi)
    m, g = ctx.exact_gradient()
    f32 = np.float32
    rec, hist = rb.oracle_exact_forward("f32", nxs, nys, nxr, nyr, S, f32(d), f32(dt), vi, nt, dpml=dpml,
                                        free_surface=fs, history=True)
    e = (f32(dt) * (rec - obs)).astype(np.float32)
    _, go = rb.oracle_exact_adjoint("f32", nxs, nys, nxr, nyr, e, f32(d), f32(dt), vi, nt, hist=hist, dpml=dpml,
                                    free_surface=fs)
    m_o = 0.5 * float(dt) * float(((rec - obs).astype(np.float64) ** 2).sum())
    assert m == pytest.approx(m_o, rel=1e-12)
    assert np.abs(g).max() > 0
    assert np.array_equal(g, go)
    # against the FP64 oracle
    S8 = S.astype(np.float64)
    obs8, _ = rb.oracle_exact_forward("f64", nxs, nys, nxr, nyr, S8, d, dt, vt.astype(np.float64), nt, dpml=dpml,
                                      free_surface=fs)
    rec8, hist8 = rb.oracle_exact_forward("f64", nxs, nys, nxr, nyr, S8, d, dt, vi.astype(np.float64), nt,
                                          dpml=dpml, free_surface=fs, history=True)
    _, g8 = rb.oracle_exact_adjoint("f64", nxs, nys, nxr, nyr, dt * (rec8 - obs8), d, dt, vi.astype(np.float64), nt,
                                    hist=hist8, dpml=dpml, free_surface=fs)
    e_rel = float(np.linalg.norm(g - g8) / np.linalg.norm(g8))
    cos = float((g * g8).sum() / (np.linalg.norm(g) * np.linalg.norm(g8)))
    print(f"\nexact gradient fs={fs}: FP32 GPU vs FP64 oracle rel-L2 {e_rel:.2e}, cosine {cos:.9f}")
    assert e_rel < 1e-3 and cos >= 0.9999


def test_exact_gradient_sums_shot_windows():
    ny, nx, nt, d, dt, dpml, vt, vi, w = _fd_model()
    nx2 = 80
    vt2 = np.concatenate([vt, vt[:, :nx2 - nx]], 1).astype(np.float32)
    vi2 = np.full((ny, nx2), 2000.0, np.float32)
    shots = [dict(x0=0, w=44, src_x=[5], src_y=[10], rec_x=list(range(1, 45, 2)), rec_y=[2] * 22),
             dict(x0=30, w=50, src_x=[30], src_y=[3], rec_x=list(range(2, 51, 3)), rec_y=[25] * 17)]
    wav = [w[:, None].astype(np.float32), (0.5 * w)[:, None].astype(np.float32)]
    ctx = _ctx(ny, nx2, d, dt, nt, shots, wav, vt2, dpml, 0)
    ctx.forward(False)
    ctx.observed_from_synthetic()
    obs = ctx.synthetic()
    ctx.set_model(vi2)
    m, g = ctx.exact_gradient()
    f32 = np.float32
    g_o = np.zeros((ny, nx2))
    m_o = 0.0
    for s, wv, ob in zip(shots, wav, obs):
        sub = np.asfortranarray(vi2[:, s["x0"]:s["x0"] + s["w"]])
        args = (s["src_x"], s["src_y"], s["rec_x"], s["rec_y"])
        rec, hist = rb.oracle_exact_forward("f32", *args, wv, f32(d), f32(dt), sub, nt, dpml=dpml, free_surface=0,
                                            history=True)
        _, gs = rb.oracle_exact_adjoint("f32", *args, (f32(dt) * (rec - ob)).astype(np.float32), f32(d), f32(dt),
                                        sub, nt, hist=hist, dpml=dpml, free_surface=0)
        g_o[:, s["x0"]:s["x0"] + s["w"]] += gs
        m_o += 0.5 * float(dt) * float(((rec - ob).astype(np.float64) ** 2).sum())
    assert m == pytest.approx(m_o, rel=1e-12)
    assert np.allclose(g, g_o, rtol=1e-12, atol=1e-12 * np.abs(g_o).max())
