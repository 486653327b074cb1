"""Writes tests/golden/marine64_nt5000_fs0_r4.npz: the reference's own Fortran
solver (oracle/_ref, REAL(4), free surface off) on two shots of the 64-shot
marine gradient workload at full size -- nx=2000 nz=400 h=25 m dt=2 ms
nt=5000, the widest aperture window and the median one.

Per shot: observed data in the true model (iprop=3), forward with snapshots in
the initial model (iprop=1), Misfit_L2 on traces normalised as
Misfit_function does, adjoint with the imaging condition (iprop=2) -- the
same path as make_golden.gradient_path.  The full arrays are too large to
keep, so the fixture holds SHA-256 digests of the full observed and
synthetic seismograms, adjoint source and imaging kernel (Fortran-order
float32 bytes), decimated copies for diagnostics, and the misfit.

Run:  python tests/golden/make_marine_fullsize.py   (about 7 minutes, two processes)
"""
from __future__ import annotations

import multiprocessing as mp
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import refbind as rb  # noqa: E402
from fwi_amd import synth  # noqa: E402
from make_golden import gradient_path, sha  # noqa: E402

OUT = os.path.join(HERE, "marine64_nt5000_fs0_r4.npz")


def picked_shots(acq):
    order = sorted(range(len(acq.shots)), key=lambda i: (acq.shots[i]["w"], i))
    return [order[-1], order[len(order) // 2]]


def one_shot(i):
    acq = synth.marine_64shot_2000x400()
    s = acq.shots[i]
    x0, w = s["x0"], s["w"]
    vt = np.asfortranarray(acq.vp_true[:, x0:x0 + w])
    vi = np.asfortranarray(acq.vp_init[:, x0:x0 + w])
    f32 = np.float32
    obs, F, mis, adj, A = gradient_path("fs0_r4", vt, vi, acq.wavelets[i], s["src_x"], s["src_y"], s["rec_x"],
                                        s["rec_y"], f32(acq.h), f32(acq.dt), acq.nt, acq.store_snap)
    return dict(shot=i, w=w, obs_sha=sha(obs), rec_sha=sha(F.rec), adj_sha=sha(adj), kernel_sha=sha(A.kernel),
                misfit=float(mis),
                rec_dec=F.rec[::25, ::16].copy(), kernel_dec=A.kernel[::4, ::16].copy())


def main():
    if not rb.ref_available("fs0_r4"):
        raise SystemExit("oracle/_ref is not built: run oracle/build_ref.sh first")
    acq = synth.marine_64shot_2000x400()
    picks = picked_shots(acq)
    with mp.get_context("fork").Pool(len(picks)) as pool:
        res = pool.map(one_shot, picks)
    out = {}
    for k, r in enumerate(res):
        for key, v in r.items():
            out[f"s{k}_{key}"] = v
        print(f"shot {r['shot']} (W={r['w']}): misfit {r['misfit']:.9e}, kernel sha {r['kernel_sha'][:16]}")
    np.savez_compressed(OUT, nshots=len(res), **out)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
