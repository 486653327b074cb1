"""Edge cases of the batched context on the GPU, against the oracle.

* Bad shot descriptions are refused with a message (window outside the model,
  a source or receiver outside the padded grid, no shots), calls out of
  order are refused, and the context works normally afterwards.
* Ragged acquisition in one batch: a shot without receivers, a shot without
  sources, sources and receivers on the window corners and inside the CPML
  frame, windows 1, 2 and 5 columns wide.  Seismograms, adjoint sources and
  imaging kernels bit-exact per shot; the misfit to 1e-5 (FP64 sum of the
  oracle's FP32 terms).
* A batch with no receivers at all: zero misfit and a zero gradient.
* Records as short as one snapshot interval, and snapshots every 2 and 3 steps.
* The drop-in solver with store_snap 1 and 0: the reference's snapshot counter
  never fires there, so Fields and kernel are left as they are (the context
  refuses store_snap < 2 instead).
"""
import numpy as np
import pytest

import fwi_amd
import refbind as rb
from fwi_amd import synth

pytestmark = pytest.mark.gpu

D, DT = np.float32(25.0), np.float32(0.002)


def _shot(x0, w, sx, sy, rx, ry):
    return dict(x0=x0, w=w, src_x=list(sx), src_y=list(sy), rec_x=list(rx), rec_y=list(ry))


def test_shot_errors_leave_context_usable():
    ny, nx, nt, ss = 40, 60, 120, 10
    dn = 20 + 6
    ctx = fwi_amd.Context(ny, nx, D, DT, nt, ss, dpml=20, free_surface=0)
    try:
        with pytest.raises(fwi_amd.FwiError, match="set shots first"):
            ctx.set_model(np.full((ny, nx), 2000.0, np.float32))
        bad = [
            ([], "bad arguments to fwi_set_shots"),
            ([_shot(50, 20, [1], [1], [2], [2])], "window outside the model"),
            ([_shot(0, 0, [], [], [], [])], "window outside the model"),
            ([_shot(-1, 10, [1], [1], [2], [2])], "window outside the model"),
            ([_shot(0, 10, [10 + dn + 1], [1], [2], [2])], "source outside the padded grid"),
            ([_shot(0, 10, [1], [-dn], [2], [2])], "source outside the padded grid"),
            ([_shot(0, 10, [1], [1], [2], [ny + dn + 1])], "receiver outside the padded grid"),
            ([_shot(0, 10, [1], [1], [-dn], [2])], "receiver outside the padded grid"),
        ]
        for shots, message in bad:
            with pytest.raises(fwi_amd.FwiError, match=message):
                ctx.set_shots(shots)
        c, S, nxs, nys, nxr, nyr = synth.small_case(2, ny, nx, nt)
        # positions on the padded grid's outer cells are accepted
        ctx.set_shots([_shot(0, nx, nxs, nys, nxr, nyr), _shot(0, nx, [1 - dn, nx + dn], [1 - dn, ny + dn],
                                                                [1 - dn, nx + dn], [ny + dn, 1 - dn])])
        with pytest.raises(fwi_amd.FwiError, match="model not set"):
            ctx.forward(False)
        ctx.set_model(c)
        ctx.set_sources([S, np.ones((nt, 2), np.float32)])
        with pytest.raises(fwi_amd.FwiError, match="run fwi_forward first"):
            ctx.misfit()
        ctx.forward(store_snapshots=False)
        with pytest.raises(fwi_amd.FwiError, match="run fwi_forward with snapshots first"):
            ctx.adjoint()
        rec = ctx.synthetic()
    finally:
        ctx.close()
    O = rb.oracle_solver("f32", 1, nxs, nys, nxr, nyr, S, None, D, DT, c, ss, nt // ss, nt, free_surface=0)
    assert np.abs(O.rec).max() > 0
    assert np.array_equal(rec[0], O.rec)
    # sources on the frame's outer cells inject nothing (only model cells are
    # updated with the source term); receivers there sample the frame
    O1 = rb.oracle_solver("f32", 1, [1 - dn, nx + dn], [1 - dn, ny + dn], [1 - dn, nx + dn], [ny + dn, 1 - dn],
                          np.ones((nt, 2), np.float32), None, D, DT, c, ss, nt // ss, nt, free_surface=0)
    assert np.array_equal(rec[1], O1.rec)


def _batch_vs_oracle(ny, nx, nt, ss, shots, wavs, vp, obs_of):
    ctx = fwi_amd.Context(ny, nx, D, DT, nt, ss, dpml=20, free_surface=0)
    try:
        ctx.set_shots(shots)
        ctx.set_model(vp)
        ctx.set_sources(wavs)
        ctx.forward(store_snapshots=True)
        syn = ctx.synthetic()
        obs = [obs_of(i, a) for i, a in enumerate(syn)]
        ctx.set_observed(obs)
        mis = ctx.misfit()
        adj = ctx.adjoint_sources()
        ctx.adjoint()
        kers = ctx.kernels()
        ctx.reduce_gradient(1.0, None)
        grad = ctx.get_gradient()
    finally:
        ctx.close()
    mis_o, g_o = 0.0, np.zeros((ny, nx))
    for i, (s, w) in enumerate(zip(shots, wavs)):
        sub = np.asfortranarray(vp[:, s["x0"]:s["x0"] + s["w"]])
        nr = len(s["rec_x"])
        O = rb.oracle_solver("f32", 1, s["src_x"], s["src_y"], s["rec_x"], s["rec_y"], w, None, D, DT, sub, ss,
                             nt // ss, nt, free_surface=0)
        assert syn[i].shape == (nt, nr)
        if nr:
            assert np.array_equal(syn[i], O.rec), f"shot {i}: seismogram differs"
            m, adj_o = rb.oracle_misfit(O.rec, obs[i], float(DT))
            assert np.array_equal(adj[i], adj_o), f"shot {i}: adjoint source differs"
            mis_o += m
        else:
            adj_o = np.zeros((nt, 0), np.float32)
        O2 = rb.oracle_solver("f32", 2, s["src_x"], s["src_y"], s["rec_x"], s["rec_y"], w, adj_o, D, DT, sub, ss,
                              nt // ss, nt, fields=O.fields, free_surface=0)
        assert np.array_equal(kers[i], O2.kernel), f"shot {i}: kernel differs"
        g_o[:, s["x0"]:s["x0"] + s["w"]] += O2.kernel.astype(np.float64)
    assert abs(mis - mis_o) <= 1e-5 * max(abs(mis_o), 1e-30)
    assert np.array_equal(grad, g_o)      # FP64 sum in shot order, then the scale: same operations
    return syn, kers, mis


def test_ragged_acquisition_vs_oracle():
    ny, nx, nt, ss = 40, 90, 200, 10
    rng = np.random.default_rng(11)
    vp = (1800 + 1200 * rng.random((ny, nx))).astype(np.float32)
    shots = [
        _shot(0, 60, [20], [5], [], []),                                   # no receivers
        _shot(10, 70, [], [], [5, 30, 60], [3, 20, 38]),                   # no sources
        _shot(20, 70, [1, 70, -10], [1, 40, 20], [1, 70, 1, 70, 80, -4], [1, 40, 40, 1, 10, -2]),  # corners, frame
        _shot(89, 1, [1], [8], [1, 1], [12, 30]),                          # one column
        _shot(0, 2, [2], [10], [1, 2], [14, 25]),                          # two columns
        _shot(40, 5, [3], [6], [1, 5, 3], [9, 20, 39]),                    # five columns
    ]
    wavs = [synth.ricker(12.0, nt, float(DT))[:, None] * np.ones((1, len(s["src_x"])), np.float32) for s in shots]
    wavs = [np.asfortranarray(w.astype(np.float32)) for w in wavs]

    def obs_of(i, a):
        return (0.8 * a + np.float32(1e-3) * np.float32(i + 1)).astype(np.float32)
    syn, kers, mis = _batch_vs_oracle(ny, nx, nt, ss, shots, wavs, vp, obs_of)
    assert syn[0].shape == (nt, 0) and np.abs(kers[0]).max() == 0       # nothing to back-propagate
    assert np.abs(syn[1]).max() == 0 and np.abs(kers[1]).max() == 0     # no incident field
    for i in (2, 3, 4, 5):
        assert np.abs(syn[i]).max() > 0 and np.abs(kers[i]).max() > 0
    assert mis > 0


def test_no_receivers_anywhere():
    ny, nx, nt, ss = 30, 50, 100, 10
    vp = np.full((ny, nx), 2200.0, np.float32)
    shots = [_shot(0, 50, [10], [5], [], []), _shot(5, 30, [3, 20], [4, 9], [], [])]
    wavs = [synth.ricker(15.0, nt, float(DT))[:, None] * np.ones((1, len(s["src_x"])), np.float32) for s in shots]
    syn, kers, mis = _batch_vs_oracle(ny, nx, nt, ss, shots, [w.astype(np.float32) for w in wavs], vp,
                                      lambda i, a: a)
    assert mis == 0.0
    assert all(s.shape == (nt, 0) for s in syn) and all(np.abs(k).max() == 0 for k in kers)


@pytest.mark.parametrize("nt, ss", [(10, 10), (24, 2), (30, 3)])
def test_short_records_and_dense_snapshots(nt, ss):
    ny, nx = 36, 48
    rng = np.random.default_rng(nt + ss)
    vp = (1500 + 2000 * rng.random((ny, nx))).astype(np.float32)
    # receivers next to the sources and noise wavelets, so short records carry signal
    shots = [_shot(0, 48, [20, 30], [10, 25], [21, 22, 31, 29], [10, 11, 26, 24]),
             _shot(8, 30, [15], [18], [15, 16, 14], [19, 18, 17])]
    wavs = [np.asfortranarray(rng.standard_normal((nt, len(s["src_x"]))).astype(np.float32)) for s in shots]

    def obs_of(i, a):
        # traces are normalised by their maxima, so a rescaled copy would fit exactly
        return np.ascontiguousarray(np.roll(a, 1, axis=0)).astype(np.float32)
    syn, kers, _ = _batch_vs_oracle(ny, nx, nt, ss, shots, wavs, vp, obs_of)
    assert all(np.abs(s).max() > 0 for s in syn)
    # with a single snapshot interval the only correlation taken is with the
    # zero initial snapshot (slot 0), so the reference's kernel is zero
    assert all((np.abs(k).max() > 0) == (nt // ss >= 2) for k in kers)


@pytest.mark.parametrize("ss", [1, 0])
def test_dropin_store_snap_below_two(ss):
    c, S, nxs, nys, nxr, nyr = synth.small_case(4, 36, 48, 120)
    nt, nts = S.shape[0], 120
    rng = np.random.default_rng(ss)
    F0 = np.asfortranarray(rng.standard_normal((36, 48 * (nts + 1))).astype(np.float32))
    k0 = np.asfortranarray(rng.standard_normal((36, 48)).astype(np.float32))
    rec, F, _ = fwi_amd.solver_acoustic_SS(1, nxs, nys, nxr, nyr, S, None, D, DT, c, ss, nts, nt, Fields=F0.copy(order="F"),
                                           free_surface=0)
    O = rb.oracle_solver("f32", 1, nxs, nys, nxr, nyr, S, None, D, DT, c, max(ss, 1), nts, nt, fields=F0,
                         free_surface=0)
    assert np.abs(rec).max() > 0 and np.array_equal(rec, O.rec)
    assert np.array_equal(F, F0) and np.array_equal(O.fields, F0)
    adj = np.asfortranarray(rng.standard_normal((nt, len(nxr))).astype(np.float32))
    rec2, _, k = fwi_amd.solver_acoustic_SS(2, nxs, nys, nxr, nyr, S, adj, D, DT, c, ss, nts, nt,
                                            kernel=k0.copy(order="F"), Fields=F0.copy(order="F"), free_surface=0)
    assert np.array_equal(k, k0) and np.abs(rec2).max() == 0
    if rb.ref_available("fs0_r4"):
        R = rb.ref_solver("fs0_r4", 1, nxs, nys, nxr, nyr, S, None, D, DT, c, ss, nts, nt)
        assert np.array_equal(R.rec, rec) and np.abs(R.fields).max() == 0


@pytest.mark.parametrize("nt, ss, nts", [(205, 10, 20), (200, 10, 25), (123, 7, 17)])
def test_dropin_record_not_a_multiple_of_store_snap(nt, ss, nts):
    """The reference's caller rounds nt down to a multiple of store_snap, but the
    routine itself takes any nt and nt_Store: snapshots at every store_snap-th
    step into slots 1, 2, ...; the adjoint correlates slot nt_Store - k/store_snap
    at step k (slots past the stored ones hold what the caller passed)."""
    c, S, nxs, nys, nxr, nyr = synth.small_case(6, 36, 52, nt)
    rec, F, _ = fwi_amd.solver_acoustic_SS(1, nxs, nys, nxr, nyr, S, None, D, DT, c, ss, nts, nt, free_surface=0)
    O = rb.oracle_solver("f32", 1, nxs, nys, nxr, nyr, S, None, D, DT, c, ss, nts, nt, free_surface=0)
    assert np.array_equal(rec, O.rec) and np.array_equal(F, O.fields)
    adj = -O.rec[::-1].copy()
    _, _, k = fwi_amd.solver_acoustic_SS(2, nxs, nys, nxr, nyr, S, adj, D, DT, c, ss, nts, nt,
                                         Fields=O.fields.copy(order="F"), free_surface=0)
    O2 = rb.oracle_solver("f32", 2, nxs, nys, nxr, nyr, S, adj, D, DT, c, ss, nts, nt, fields=O.fields,
                          free_surface=0)
    assert np.abs(O2.kernel).max() > 0 and np.array_equal(k, O2.kernel)
