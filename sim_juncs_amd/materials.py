"""Material models of the sim_juncs junctions, in the units the engine takes.

The reference writes poles in `.geom` files as ``[f0, gamma, sigma, "drude"|"lorentz"]``
with f0 and gamma in 1/um (eV / 1.23984193) and divides them by ``um_scale`` when it builds
``meep::lorentzian_susceptibility`` (reference src/disp.cpp:537-546).  sigma is divided by
the sheet thickness for ``make_2d`` roots (src/disp.cpp:519-524,541).
"""
from __future__ import annotations

from dataclasses import dataclass, field

# hc in eV*um used to convert Rakic's eV parameters to 1/um frequencies.
HC_EV_UM = 1.23984193


@dataclass(frozen=True)
class Pole:
    f0: float        # 1/um (before / um_scale)
    gamma: float     # 1/um (before / um_scale)
    sigma: float
    drude: bool

    def scaled(self, um_scale: float, thickness: float = 1.0) -> "Pole":
        return Pole(self.f0 / um_scale, self.gamma / um_scale, self.sigma / thickness, self.drude)


@dataclass(frozen=True)
class MaterialSpec:
    eps_inf: float = 1.0
    poles: tuple = field(default_factory=tuple)

    def scaled(self, um_scale: float, thickness: float = 1.0) -> "MaterialSpec":
        return MaterialSpec(self.eps_inf, tuple(p.scaled(um_scale, thickness) for p in self.poles))

    def chi(self, f):
        """Continuous-time susceptibility at meep frequency f (poles already scaled):
        sum sigma w0^2 / (L w0^2 - w^2 - i w gamma'), w = 2 pi f, gamma' = 2 pi gamma."""
        import numpy as np

        w = 2 * np.pi * np.asarray(f, dtype=complex)
        out = np.zeros_like(w)
        for p in self.poles:
            w0 = 2 * np.pi * p.f0
            g = 2 * np.pi * p.gamma
            L = 0.0 if p.drude else 1.0
            out = out + p.sigma * w0 * w0 / (L * w0 * w0 - w * w - 1j * w * g)
        return out

    def eps(self, f):
        return self.eps_inf + self.chi(f)


VACUUM = MaterialSpec(1.0)

# Au: Rakic et al. 1998 Drude + first Lorentz pole as written in
# junctions/Au_SiO2_box/junc.geom:60 (1/um units).
AU_DRUDE = Pole(1e-10, 0.04274738474121455, 4.0314052191361974e21, True)
AU_L1 = Pole(0.3347200880680007, 0.19437961740816426, 11.3629356945, False)
# Second Rakic Lorentz pole (omega_2 = 0.830 eV, Gamma_2 = 0.345 eV, f_2 = 0.010), converted
# as the first: f0 = omega_2 / hc, gamma = Gamma_2 / hc, sigma = f_2 * omega_p^2 / omega_2^2 with
# omega_p = 9.03 eV.
AU_L2 = Pole(0.6694401761, 0.278261278, 1.183639135, False)
AU = MaterialSpec(1.0, (AU_DRUDE, AU_L1, AU_L2))
# What the reference actually simulates for Au: the Lorentz pole on junc.geom:60 lacks commas,
# so the list parser fuses it into a 2-element list and parse_susceptibilities rejects it
# (src/cgs_read.cpp:275,1906; src/disp.cpp:424-430); only the Drude pole survives.
AU_REFERENCE_AS_RUN = MaterialSpec(1.0, (AU_DRUDE,))

# SiO2: one Lorentz pole (junctions/Au_SiO2_box/junc.geom:68).
SIO2 = MaterialSpec(1.0, (Pole(9.67865314895427, 0.08065544290795199, 1.12, False),))
# Non-dispersive silica slab used by the plain 96^3 workload.
SIO2_CONST = MaterialSpec(2.13)

# Graphene as a Drude sheet (the reference's graphene entry, junctions/Au_graphene_box/
# junc_template.geom:18, is a legacy Lorentz pole with an unphysical weight).  The intraband
# sheet conductivity sigma_s(w) = (e^2 E_F / (pi hbar^2)) i / (w + i Gamma) smeared over a
# layer of thickness t gives eps(w) = 1 - w_p^2 / (w^2 + i w Gamma) with
# w_p^2 = e^2 E_F / (pi hbar^2 eps0 t).  In the engine's Drude form sigma * w0^2 = w_p^2 with a
# dummy w0 (f0 = 1e-10, as the reference writes its Au Drude pole).
GRAPHENE_EF_EV = 0.4
GRAPHENE_GAMMA_EV = 0.02


def graphene_sheet(thickness_um: float, ef_ev: float = GRAPHENE_EF_EV,
                   gamma_ev: float = GRAPHENE_GAMMA_EV) -> MaterialSpec:
    """Drude sheet of graphene smeared over `thickness_um`, poles in 1/um frequency units."""
    import math

    e, hbar, eps0, c = 1.602176634e-19, 1.054571817e-34, 8.8541878128e-12, 299792458.0
    wp2 = e * e * (ef_ev * e) / (math.pi * hbar * hbar * eps0 * (thickness_um * 1e-6))
    fp2_um = wp2 / (2 * math.pi * c) ** 2 * 1e-12  # (w_p / 2 pi c)^2 in um^-2
    f0 = 1e-10
    return MaterialSpec(1.0, (Pole(f0, gamma_ev / HC_EV_UM, fp2_um / (f0 * f0), True),))
