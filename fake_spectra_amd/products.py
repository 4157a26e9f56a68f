"""Several outputs from one binning: the column-density-weighted fields of
Spectra.get_velocity / get_temp / get_dens_weighted_density
(spectra.py:945-1043) on device-resident particles.  (Several lines' optical
depths from one binning are spectra_priv.tau with a list of lines.)

The reference makes one _Particle_Interpolate call per weighted field, each
rebuilding the sightline table; here one Plan serves every field, and all
weighted column densities go through one batched colden launch.
"""
from __future__ import annotations

from typing import Dict

import torch

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
