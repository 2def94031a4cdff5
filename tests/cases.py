"""Small cases shared by the parity tests (sizes the oracle runs in seconds)."""
import functools

from flite3d_amd import mesh
from flite3d_amd.params import InputVariables


@functools.lru_cache(maxsize=None)
def levels(n, nl, jitter=0.15, first_cell=None, wall=1):
    return tuple(mesh.bump_levels(n, nl, jitter=jitter, seed=0, first_cell=first_cell,
                                  wall_indicator=wall))


def euler(order=1, **kw):
    base = dict(reynoldsnumber=0.0, machnumber=0.5, alpha=2.0, fluxtype=4,
                hllcfluxorder=order, automaticlipping=1, cflnumber=2.0,
                multigridscheme=1, restartnumber=0)
    base.update(kw)
    return InputVariables(**base)


def rans(**kw):
    base = dict(reynoldsnumber=6.0e6, machnumber=0.7, alpha=1.0, fluxtype=4,
                hllcfluxorder=2, automaticlipping=1, hllcclipping=1.0,
                gradientmethod=1, turbulencemodel=1, wallsareisentropic=True,
                cflnumber=1.0, multigridscheme=1, restartnumber=0)
    base.update(kw)
    return InputVariables(**base)


def laminar(**kw):
    base = dict(reynoldsnumber=500.0, machnumber=0.5, alpha=2.0, fluxtype=4,
                hllcfluxorder=2, automaticlipping=1, turbulencemodel=0,
                wallsareisentropic=False, cflnumber=1.0, multigridscheme=1,
                restartnumber=0)
    base.update(kw)
    return InputVariables(**base)
