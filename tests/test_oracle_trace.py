"""Trace-conditioning oracle (oracle/traces.py, oracle_akima_f32) against the
reference's own filters_function1D/FFTPACK and INTRPL (oracle/_ref), and
against closed forms:
* the response magnitude is the Butterworth one, |H|^2 = 1/(1 + W^(2n)) with
  W = f/fc (low-pass), fs/f (high-pass) or Q (f/f0 - f0/f) (band-pass), n = 6;
* Akima reproduces linear data and the identity resampling exactly;
* a zero trace stays zero; the filter is linear."""
import numpy as np
import pytest

import refbind as rb
import traces as tr
from fwi_amd import synth

needs_ref = pytest.mark.skipif(not rb.ref_available("trace"), reason="reference build (oracle/_ref) not present")


def _data(seed, nt, ntr):
    rng = np.random.default_rng(seed)
    d = np.cumsum(rng.standard_normal((nt, ntr)), 0).astype(np.float32)
    d[:, 0] = synth.ricker(6.0, nt, 0.004)
    return d


@needs_ref
@pytest.mark.parametrize("dt_in,dt_out,nt_in,nt_out", [
    (0.004, 0.0013, 300, 900),     # finer grid
    (0.002, 0.002, 300, 300),      # identity
    (0.002, 0.0031, 300, 250),     # coarser grid running past the last input time (extrapolation)
    (0.004, 0.001, 2, 9),          # two input points
    (0.004, 0.0007, 5, 40),        # short input, both ends extrapolated in the interval logic
])
def test_akima_oracle_matches_reference(dt_in, dt_out, nt_in, nt_out):
    d = _data(1, nt_in, 4)
    R = rb.ref_akima(d, dt_in, dt_out, nt_out)
    O = tr.akima(rb.oracle(), d, dt_in, dt_out, nt_out)
    assert np.array_equal(R, O)


def test_akima_linear_and_identity():
    t = np.arange(200) * np.float32(0.004)
    lin = np.stack([np.float32(2.5) * t + np.float32(1.0), np.float32(-3.0) * t], 1).astype(np.float32)
    out = tr.akima(rb.oracle(), lin, 0.004, 0.0017, 450)
    u = np.arange(450) * np.float32(0.0017)
    assert np.allclose(out[:, 0], 2.5 * u + 1.0, rtol=1e-6, atol=1e-5)
    assert np.allclose(out[:, 1], -3.0 * u, rtol=1e-6, atol=1e-5)
    d = _data(2, 120, 3)
    assert np.array_equal(tr.akima(rb.oracle(), d, 0.004, 0.004, 120), d)


@needs_ref
@pytest.mark.parametrize("typef,f1,fc", [(3, 3.0, 5.0), (3, 5.0, 7.0), (1, 0.0, 8.0), (2, 0.0, 10.0)])
def test_filter_oracle_matches_reference(typef, f1, fc):
    d = _data(3, 200, 3)
    d[:, 2] = 0.0
    R = rb.ref_time_filter(d, 0.004, typef, f1, fc)
    O = tr.time_filter(d, 0.004, typef, f1, fc)
    assert np.array_equal(R[:, 2], d[:, 2]) and np.array_equal(O[:, 2], d[:, 2])
    for j in range(2):
        # FFT rounding in REAL(4) scales with the trace, not with the (possibly tiny) filtered output
        assert np.linalg.norm(R[:, j] - O[:, j]) <= 2e-6 * np.linalg.norm(d[:, j])


@pytest.mark.parametrize("typef", [1, 2, 3])
def test_butterworth_magnitude(typef):
    ts, dt, f1, fc = 4800, 0.004, 3.0, 5.0
    H = tr.butterworth_response(ts, dt, typef, f1, fc)
    f32 = np.float32
    step = (f32(1.0) / f32(dt)) / f32(ts)
    f = np.zeros(ts // 2 + 1, np.float32)
    acc = f32(0.0)
    for j in range(len(f)):
        f[j] = acc
        acc = f32(acc + step)
    w = f[1:].astype(np.float64)
    f0 = float(np.sqrt(f32(f1) * f32(fc)))
    Q = float(f32(f0) / (f32(fc) - f32(f1)))
    W = {1: w / fc, 2: 40.0 / w, 3: Q * (w / f0 - f0 / w)}[typef]
    mag2 = 1.0 / (1.0 + W ** 12)
    assert np.allclose(np.abs(H[1:ts // 2 + 1]) ** 2, mag2, rtol=1e-6, atol=1e-12)
    # conjugate mirror of the negative bins
    assert np.allclose(H[ts // 2 + 1:], np.conj(H[1:ts // 2][::-1]))


def test_filter_linearity():
    d = _data(4, 150, 2)
    a = tr.time_filter(d, 0.004, 3, 3.0, 5.0)
    b = tr.time_filter(np.float32(2.0) * d, 0.004, 3, 3.0, 5.0)
    assert np.allclose(b, 2.0 * a, rtol=1e-12, atol=1e-12 * np.abs(a).max())


def test_wiener_core_identities():
    """Identical data give s(w) = kk/(kk+eps) on the band: a zero-phase pulse at t = 0 of height
    sum_band kk/(kk+eps) / npad; data delayed by m samples move the pulse to t = m."""
    rng = np.random.default_rng(7)
    nt, nr, dt, f1, f2 = 400, 6, 0.004, 3.0, 9.0
    Gs = np.zeros((nt, nr), np.float32)
    Gs[20:120] = rng.standard_normal((100, nr))
    s = tr.wiener_source(Gs, Gs, nt, dt, f1, f2)
    npad = 2 * nt
    df = np.float32(1.0) / (np.float32(nt) * np.float32(dt))
    if1, if2 = 1 + int(np.floor(np.float32(f1) / df)), 1 + int(np.floor(np.float32(f2) / df))
    W = np.fft.fft(Gs.astype(np.float64), n=npad, axis=0)[if1 - 1:if2] / npad
    kk = (np.abs(W) ** 2).sum(1)
    expect0 = (kk / (kk + 0.001 * kk.max())).sum() / npad
    assert s[0] == pytest.approx(expect0, rel=1e-12)
    assert np.argmax(s) == 0
    m = 37
    Go = np.zeros_like(Gs)
    Go[m:] = Gs[:-m]
    assert np.argmax(tr.wiener_source(Gs, Go, nt, dt, f1, f2)) == m


def test_estimation_windows():
    """Water 1500 m/s down to row 9 (z = 200 m at d = 25 m): the window of a receiver at offset h
    ends half a period before the water-bottom reflection 2 sqrt((h/2)^2 + 200^2)/1500."""
    ny, nx, nt, d, dt = 30, 60, 500, 25.0, 0.002
    model = np.full((ny, nx), 1500.0, np.float32)
    model[8:, :] = 2000.0
    it2 = tr._estimation_windows(nx, ny, nt, d, dt, 20, np.array([20, 40, 60]), model, 3.0, 5.0)
    for r, x in enumerate((20, 40, 60)):
        t2 = 2 * np.sqrt((abs(x - 20) * d / 2) ** 2 + 200.0 ** 2) / 1500.0 - 0.5 / 4.0
        assert abs(it2[r] - (1 + int(t2 / dt))) <= 1
