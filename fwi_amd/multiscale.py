"""Multiscale inversion loop on a Context: the reference's frequency-band loop
(00-main_Inversion.f90:76-88 with get_f1f2, auxiliar.f90:223-261) around the
NLCG iterations of FWI.f90.

Per band the raw observed data and source traces are band-filtered
(time_filter, `typef`) and resampled onto the context's time grid on the GPU,
as calculate_misfit_gradient.f90:193-244 does per shot, and a fresh NLCG
(FWI.f90:59-61: EndStep = 0, OldGrad = 0, var_orig = lambda) runs `iters`
accepted iterations.  A line search that needs more than 6 trials resets the
search direction and repeats the iteration; the second such failure in a band
ends the band (FWI.f90:244-260).

The loop only calls Context methods, so any object with the same methods (a
CPU restatement, for instance) runs the identical sequence of decisions.
"""
from __future__ import annotations

from .nlcg import NLCG


def get_f1f2(ifreq: int, f_init: float, step_freq: float, num_freq: int, bp_type: int):
    """Band (f1, f2) of frequency index ifreq = 2..num_freq (auxiliar.f90:223-261).

    bp_type 1: growing band [f_init, Freq(ifreq)]; 2: sliding band
    [Freq(ifreq) - step_freq, Freq(ifreq)], with Freq(i) = f_init + (i-1) step_freq.
    bp_type 3 also switches the misfit function (misfit_type 3, then 1), which
    only the L2 misfit of this library does not cover, so it is rejected."""
    if not 2 <= ifreq <= num_freq:
        raise ValueError("ifreq runs from 2 to num_freq")
    freq = [f_init + (i - 1) * step_freq for i in range(1, num_freq + 1)]
    if bp_type == 1:
        return f_init, freq[ifreq - 1]
    if bp_type == 2:
        return freq[ifreq - 1] - step_freq, freq[ifreq - 1]
    raise ValueError(f"bp_type {bp_type}: only 1 (growing band) and 2 (sliding band) use the L2 misfit")


def band_loop(ctx, raw_obs, raw_src, dt_data, bands, iters, typef=3, water_rows=None, scale=1.0, method=1,
              on_iteration=None):
    """Runs the multiscale inversion on `ctx` (shots and starting model set).

    raw_obs / raw_src: per-shot (nt_data, nrec) / (nt_data, nsrc) traces on
    dt_data; bands: [(f1, f2), ...] in order; iters: accepted NLCG iterations
    per band (nFWI).  Returns the list of iteration records (NLCG.iterate's
    dicts with the band added); on_iteration(record) is called after each."""
    history = []
    for band, (f1, f2) in enumerate(bands):
        ctx.set_observed_raw(raw_obs, dt_data, typef, f1, f2)
        ctx.set_sources_raw(raw_src, dt_data, typef, f1, f2)
        opt = NLCG(ctx, water_rows=water_rows, scale=scale, method=method)
        accepted = failures = 0
        while accepted < iters:
            rec = opt.iterate()
            rec.update(band=band, f1=f1, f2=f2)
            history.append(rec)
            if on_iteration is not None:
                on_iteration(rec)
            if rec["accepted"]:
                accepted += 1
            else:
                failures += 1
                if failures >= 2:
                    break
    return history
