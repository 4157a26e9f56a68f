"""Several outputs from one binning: the column-density-weighted fields of
Spectra.get_velocity / get_temp / get_dens_weighted_density
(spectra.py:945-1043) on device-resident particles, and the optical depth
an observer would use for an ion (get_observer_tau, spectra.py:894-943).  (Several lines' optical
depths from one binning are spectra_priv.tau with a list of lines.)

The reference makes one _Particle_Interpolate call per weighted field, each
rebuilding the sightline table; here one Plan serves every field, and all
weighted column densities go through one batched colden launch.
"""
from __future__ import annotations

from typing import Dict, Sequence

import torch

from . import fluxstatistics as FS
from . import spectra_priv as SP


def weighted_fields(plan: SP.Plan, nbins: int, kernel: int, velfac: float, dens: torch.Tensor,
                    weights: Dict[str, torch.Tensor]) -> Dict[str, torch.Tensor]:
    """Column-density-weighted mean of each per-particle field in `weights`:
    sum_pairs dens w K / sum_pairs dens K per pixel, the quotient of
    _get_mass_weight_quantity (spectra.py:957-980): pixels with no column
    density divide by 1 (and are 0).  Also returns "colden", the column
    density itself.  All fields: float32 [N] device tensors."""
    d = SP._require(dens, torch.float32, "dens")
    names = list(weights)
    arrays = [d] + [(d * SP._require(weights[n], torch.float32, n)).contiguous() for n in names]
    outs = SP.colden(plan, nbins, kernel, velfac, arrays)
    den = outs[0].clone()
    den[den == 0] = 1.0
    res = {"colden": outs[0]}
    for n, o in zip(names, outs[1:]):
        res[n] = o / den
    return res


def velocity_temperature(plan: SP.Plan, nbins: int, kernel: int, velfac: float, dens: torch.Tensor,
                         vel: torch.Tensor, temp: torch.Tensor, atime: float) -> Dict[str, torch.Tensor]:
    """get_velocity (3 components, vel * sqrt(atime) as in spectra.py:952-954)
    and get_temp (spectra.py:997-1010) of one species, from one batched
    colden launch of five weighted densities."""
    v = SP._require(vel, torch.float32, "vel").reshape(-1, 3) * float(atime) ** 0.5
    w = {"vel_x": v[:, 0].contiguous(), "vel_y": v[:, 1].contiguous(), "vel_z": v[:, 2].contiguous(),
         "temp": SP._require(temp, torch.float32, "temp")}
    r = weighted_fields(plan, nbins, kernel, velfac, dens, w)
    r["velocity"] = torch.stack([r.pop("vel_x"), r.pop("vel_y"), r.pop("vel_z")], dim=-1)
    return r


def observer_lines(maxtaus: torch.Tensor) -> torch.Tensor:
    """Line kept for each sightline by get_observer_tau (spectra.py:918-936)
    from maxtaus [nlines, nlos]: the strongest line with 0.1 < maxtau < 3,
    else the weakest line with maxtau > 0.1, else the strongest line; the
    first line of equal values.  Returns int64 [nlos] on maxtaus' device."""
    m = maxtaus
    inf = torch.tensor(float("inf"), dtype=m.dtype, device=m.device)
    band = (m > 0.1) & (m < 3)
    # torch.argmax / argmin return the first index of equal values
    strongest_in_band = torch.argmax(torch.where(band, m, -inf), dim=0)
    weakest_above = torch.argmin(torch.where(m > 0.1, m, inf), dim=0)
    strongest = torch.argmax(m, dim=0)
    return torch.where(band.any(dim=0), strongest_in_band,
                       torch.where((m > 0.1).any(dim=0), weakest_above, strongest))


def observer_tau(plan: SP.Plan, nbins: int, kernel: int, velfac: float, vel: torch.Tensor, temp: torch.Tensor,
                 lines: Sequence, dens: torch.Tensor, dvbin: float, spec_res: float):
    """get_observer_tau (spectra.py:894-943) for the lines of one ion, all of
    whose optical depths come from one tau launch over one binning: the
    maximum over pixels of each line's tau smoothed to the spectrograph's
    resolution spec_res (km/s FWHM, pixels of dvbin km/s) picks one line per
    sightline, whose unsmoothed tau is returned.  dens is the ion's float32
    [N] density, shared by every line.  Returns (tau float64 [nlos, nbins],
    chosen line index int64 [nlos]).  spec_res must be positive: the
    reference's Gaussian smoothing divides by it."""
    if not spec_res > 0:
        raise ValueError("spec_res must be positive (the smoothing kernel divides by its width)")
    if len(lines) == 0:
        raise ValueError("at least one line is required")
    taus = SP.tau(plan, nbins, kernel, velfac, vel, temp, lines, [dens] * len(lines))
    stack = torch.stack(taus)                                   # [nlines, nlos, nbins]
    maxtaus = FS.res_corr(stack, dvbin, spec_res).amax(dim=-1)  # [nlines, nlos]
    choice = observer_lines(maxtaus)
    rows = torch.arange(stack.shape[1], device=stack.device)
    return stack[choice, rows], choice
