"""Trace-conditioning oracle -- TEST INFRASTRUCTURE ONLY (tests/, smoke(), bench's
CPU leg may use it; the product never does).

butterworth_response restates the frequency response of filters.f90:81-184
(analog Butterworth of order 6 in its low-, high- or band-pass form, H = 1/B(if)
on bins 0..ts/2 with the bin frequency accumulated in REAL(4) as f = f + fe/ts,
the DC bin forced to 1 (low-pass) or 0, the negative bins mirrored conjugate).
time_filter restates time_filter.f90:1-38: each nonzero trace is placed at the
end of a 24x zero-padded record, transformed, multiplied by the response,
transformed back, and its last nt samples kept (filters.f90:47-64; FFTPACK's
forward transform carries the 1/N).  numpy FFTs in FP64 stand in for FFTPACK.
akima wraps the C++ restatement of INTRPL (akima433.f90) in liboracle.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np


def butterworth_response(ts: int, dt: float, typef: int, f1: float, fc: float, n: int = 6) -> np.ndarray:
    """H(1..ts) of filters.f90 (complex128; REAL(4) frequency accumulation)."""
    f32 = np.float32
    pi = f32(3.14159265)
    fe = f32(1.0) / f32(dt)
    fs = f32(40.0)                              # filters.f90:30 (high-pass start frequency)
    f2 = f32(fc)
    f0 = np.sqrt(f32(f1) * f2)
    Q = f0 / (f2 - f32(f1))
    half = ts // 2 + 1
    freqs = np.zeros(half, np.float32)
    f = f32(0.0)
    step = fe / f32(ts)
    for j in range(half):
        freqs[j] = f
        f = f32(f + step)
    w = freqs.astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        if typef == 1:
            z = 1j * w / float(fc)
        elif typef == 2:
            z = float(fs) / (1j * w)
        else:
            z = float(Q) * (1j * w / float(f0) + float(f0) / (1j * w))
        B = np.ones(half, np.complex128)
        for k in range(1, n // 2 + 1):
            c = float(np.cos(f32((2 * k + n - 1) * pi) / f32(2 * n)))
            B = B * (z * z - 2 * z * c + 1)
        H = np.zeros(ts, np.complex128)
        H[:half] = 1.0 / B
    H[0] = 1.0 if typef == 1 else 0.0
    for j in range(2, ts // 2 + 1):
        H[ts // 2 + j - 1] = np.conj(H[ts // 2 + 1 - (j - 1) - 1])
    return H


def time_filter(data: np.ndarray, dt: float, typef: int, f1: float, fc: float) -> np.ndarray:
    """time_filter(Data_trace(nt, Num), nt, dt, Num, typef, f1, fc), FP64 result."""
    nt, num = data.shape
    ts = 24 * nt
    H = butterworth_response(ts, dt, typef, f1, fc)
    out = np.array(data, dtype=np.float64, copy=True)
    for i in range(num):
        tr = data[:, i]
        if not np.any(tr != 0):
            continue
        tmp = np.zeros(ts)
        tmp[ts - nt:] = tr
        y = np.fft.ifft(np.fft.fft(tmp) * H)
        out[:, i] = y.real[ts - nt:]
    return out


def akima(lib, data: np.ndarray, dt_in: float, dt_out: float, nt_out: int) -> np.ndarray:
    """interpolation_akima(Data_tmp, dt_Data, nt_Data, Data_trace, NumRec, dt, nt) via the C++ oracle."""
    nt_in, ntr = data.shape
    src = np.asfortranarray(data, dtype=np.float32)
    out = np.zeros((nt_out, ntr), np.float32, order="F")
    fn = lib.oracle_akima_f32
    fn.restype = C.c_int
    err = fn(src.ctypes.data_as(C.c_void_p), C.c_float(dt_in), C.c_int(nt_in), out.ctypes.data_as(C.c_void_p),
             C.c_int(ntr), C.c_float(dt_out), C.c_int(nt_out))
    assert err == 0, err
    return out


def _estimation_windows(nx, ny, nt, dmodel, dt, nxs, nxr, model, f1, f2):
    """Per-receiver last kept sample it2 (source_estimation.f90:79-109) in REAL(4): the water bottom is the
    first row with vp > 1500 under the source-receiver midpoint, the window ends half a period before the
    water-bottom reflection."""
    f32 = np.float32
    d = f32(dmodel)
    zwb = np.full(nx, f32(ny - 1) * d, np.float32)
    for i in range(nx):
        rows = np.nonzero(model[:, i] > f32(1500.0))[0]
        if len(rows):
            zwb[i] = f32(rows[0]) * d
    fc = (f32(f1) + f32(f2)) / f32(2.0)
    it2 = np.zeros(len(nxr), np.int64)
    for r, x in enumerate(nxr):
        xmid = f32(0.5) * f32(int(x) + int(nxs)) * d
        ixm = int(xmid / d) + 1
        z = zwb[ixm - 1]
        h = f32(abs(int(x) - int(nxs))) * d * f32(0.5)
        t_wb = f32(2.0) * np.sqrt(h * h + z * z) / f32(1500.0)
        t2 = t_wb - f32(0.5) / fc
        it2[r] = max(0, min(nt, 1 + int(t2 / f32(dt))))
    return it2


def wiener_source(Gs, Go, nt, dt, f1, f2):
    """Frequency-domain core of source_estimation.f90:120-167 in FP64: 2x padded FFTs (FFTPACK's forward
    1/N), kk = sum |Gs|^2 and bb = sum conj(Gs) Go over receivers on bins if1..if2 (bin spacing taken as
    1/(nt dt) as the reference does on its 2nt-point transform), s = bb / (kk + 1e-3 max kk), inverse
    transform, real part / (2 nt)."""
    f32 = np.float32
    npad = 2 * nt
    df = f32(1.0) / (f32(nt) * f32(dt))
    if1, if2 = 1 + int(np.floor(f32(f1) / df)), 1 + int(np.floor(f32(f2) / df))
    Ws = np.fft.fft(Gs.astype(np.float64), n=npad, axis=0) / npad
    Wo = np.fft.fft(Go.astype(np.float64), n=npad, axis=0) / npad
    band = slice(if1 - 1, if2)
    kk = (np.conj(Ws[band]) * Ws[band]).sum(1)
    bb = (np.conj(Ws[band]) * Wo[band]).sum(1)
    eps = 0.001 * kk.real.max()
    ss = np.zeros(npad, np.complex128)
    ss[band] = bb / (kk + eps)
    return (np.fft.ifft(ss) * npad).real[:nt] / npad


def source_estimation(obs, f1, f2, nx, ny, nt, nr, dmodel, dt, nxs, nys, nxr, nyr, model, forward_simple):
    """source_estimation.f90:1-171 with the observed data passed in (the reference reads them with
    get_shot_data) and without its diagnostic file output.  forward_simple(nxs, nys, nxr, nyr, S, d, dt,
    model, nt) -> (nt, nr) is the modelling routine (solver_forward_simple)."""
    f32 = np.float32
    s0 = np.zeros(nt, np.float32)
    s0[0] = f32(1.0) / f32(dt)
    Gs = np.array(forward_simple(nxs, nys, nxr, nyr, s0, f32(dmodel), f32(dt), model, nt), dtype=np.float32)
    Go = np.array(obs, dtype=np.float32, copy=True)
    it2 = _estimation_windows(nx, ny, nt, dmodel, dt, nxs, nxr, model, f1, f2)
    for r in range(nr):
        Gs[it2[r]:, r] = 0.0
        Go[it2[r]:, r] = 0.0
    Gs = Gs / np.abs(Gs).max()
    Go = Go / np.abs(Go).max()
    return wiener_source(Gs, Go, nt, dt, f1, f2)
