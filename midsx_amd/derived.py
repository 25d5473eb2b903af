"""Derived quantities of a surface tally (derived_quantities.cpp:3-68):
primary fluence per energy bin and primary air kerma
AK = sum_i E_i * fluence_i * (mu_en/rho)_air(E_i).

The reference keeps every incident energy and counts, for each bin centre
E_i of the source spectrum, the primary photons with energy in
[E_i - w/2, E_i + w/2] (derived_quantities.cpp:32,39; w = the spectrum's bin
spacing).  Here the surface tally carries an IncidentEnergy histogram whose
bins are the spectrum's bins (`spectrum_disc_tally`), so each primary photon
- whose energy is the source energy it was born with - lands in its own bin.
The air coefficients are those of the AAPM TG-195 table (`Context.air_muen`).
Only the non-cosine-weighted form is provided (the reference's drivers use
that one)."""
import math

import numpy as np

from . import TALLY_DISC, TALLY_RECT


def spectrum_bins(energies_eV):
    """(lo, width, n) of histogram bins centred on evenly spaced energies."""
    e = np.asarray(energies_eV, dtype=np.float64)
    w = float(e[1] - e[0])
    return float(e[0] - w / 2), w, int(e.size)


def spectrum_disc_tally(centre, normal, radius, energies_eV):
    """A disc surface tally (surface_tally.cpp:55-69) whose IncidentEnergy
    histogram has one bin per spectrum energy."""
    lo, w, n = spectrum_bins(energies_eV)
    return dict(kind=TALLY_DISC, p0=tuple(centre), normal=tuple(normal), radius=float(radius), hist_bins=n,
                hist_lo=lo, hist_width=w)


def tally_area(t):
    """Area of a disc (pi r^2) or rectangle (|edge1 x edge2|) tally."""
    if t["kind"] == TALLY_DISC:
        return math.pi * t["radius"] ** 2
    if t["kind"] == TALLY_RECT:
        return float(np.linalg.norm(np.cross(t["p1"], t["p2"])))
    raise ValueError("fluence needs a disc or rectangle tally")


def primary_fluence(primary_hist, area):
    """Primary photons per bin and cm^2 (derived_quantities.cpp:3-45)."""
    return np.asarray(primary_hist, dtype=np.float64) / area


def primary_spectrum_air_kerma(ctx, tally_index, energies_eV):
    """Primary air kerma (eV/g) at surface tally `tally_index` of the selected
    job for a source spectrum on energies_eV (derived_quantities.cpp:48-68);
    the tally must have been set with spectrum_disc_tally(..., energies_eV)."""
    t = ctx.tallies[tally_index]
    lo, w, n = spectrum_bins(energies_eV)
    if t.get("hist_bins") != n or abs(t.get("hist_lo") - lo) > 1e-9 * w or abs(t.get("hist_width") - w) > 1e-9 * w:
        raise ValueError("the tally's energy histogram does not match the spectrum's bins")
    e = np.asarray(energies_eV, dtype=np.float64)
    phi = primary_fluence(ctx.tally(tally_index)["hist"][0], tally_area(t))
    return float(np.sum(e * phi * ctx.air_muen(e)))
