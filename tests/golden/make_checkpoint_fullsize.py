"""Writes tests/golden/checkpoint16_nt15000_fs0_r4.npz: the reference's own
Fortran solver (oracle/_ref, REAL(4), free surface off) on one shot of the
16-shot checkpointing workload at full size -- nx=6000 nz=1000 h=10 m
dt=0.8 ms nt=15000, the narrowest of the 16 windows (1611 columns; the
reference holds its 9.7 GB of snapshots in host memory).

Same path as make_golden.gradient_path: observed data in the true model,
forward with snapshots in the initial model, Misfit_L2 on normalised traces,
adjoint with the imaging condition.  The fixture holds SHA-256 digests of the
full observed and synthetic seismograms, adjoint source and imaging kernel
(Fortran-order float32 bytes), decimated copies and the misfit.

Run:  python tests/golden/make_checkpoint_fullsize.py   (about 40 minutes, one core)
"""
from __future__ import annotations

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

OUT = os.path.join(HERE, "checkpoint16_nt15000_fs0_r4.npz")


def picked_shot(acq):
    return min(range(len(acq.shots)), key=lambda i: (acq.shots[i]["w"], i))


def main():
    if not rb.ref_available("fs0_r4"):
        raise SystemExit("oracle/_ref is not built: run oracle/build_ref.sh first")
    acq = synth.checkpoint_6000x1000(nshots=16)
    i = picked_shot(acq)
    s = acq.shots[i]
    x0, w = s["x0"], s["w"]
    vt = np.asfortranarray(acq.vp_true[:, x0:x0 + w])
    vi = np.asfortranarray(acq.vp_init[:, x0:x0 + w])
    f32 = np.float32
    obs, F, mis, adj, A = gradient_path("fs0_r4", vt, vi, acq.wavelets[i], s["src_x"], s["src_y"], s["rec_x"],
                                        s["rec_y"], f32(acq.h), f32(acq.dt), acq.nt, acq.store_snap)
    np.savez_compressed(OUT, shot=i, w=w, obs_sha=sha(obs), rec_sha=sha(F.rec), adj_sha=sha(adj),
                        kernel_sha=sha(A.kernel), misfit=float(mis), rec_dec=F.rec[::50, ::16].copy(),
                        kernel_dec=A.kernel[::8, ::16].copy())
    print(f"shot {i} (W={w}): misfit {float(mis):.9e}; wrote {OUT}")


if __name__ == "__main__":
    main()
