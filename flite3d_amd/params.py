"""InputVariables: defaults, <case>.inp namelist reading and post-processing.

Mirrors readInputVariables (SolverSource/mgnsg3d_def_par.F90:201-422): the
same defaults (DEF:217-379), the same `&inputVariables ivd%field = value /`
namelist syntax (DEF:384, SolverExampleInput/C13_1.inp), and the same
post-processing after the read (DEF:386-415): steady flow forces
maximumnumberofcycles=1000000 and addTimeStep=.false., HLLC zeroes the JST
factors, and the inflow state is rebuilt from Mach, alpha and beta.

Only the hot path's options are accepted with non-default values; asking for
an option outside it (unsteady flow, k-omega/SST, engines, JST/Roe, least
squares gradients) raises instead of running something else.
"""
from __future__ import annotations

import ctypes
import math
import re

# default values, DEF:217-379 (names are the ivd field names, lower case)
DEFAULTS = {
    "numberofmgiterations": 100,
    "numberofsgiterations": 0,
    "numberofcflincrements": 0,
    "writetofileinterval": 25,
    "gamma": 1.4,
    "machnumber": 0.5,
    "alpha": 0.0,
    "beta": 0.0,
    "cflnumber": 1.0,
    "reynoldsnumber": 1.0e6,
    "viscosityscheme": 1,
    "inflowtemperature": 528.0,
    "wallsareisentropic": False,
    "movingwall": False,
    "walltemperature": 528.0,
    "patchinitialization": False,
    "dissipationscheme": 2,
    "coarsegriddissipationscheme": 2,
    "secondorderdissipationfactor": 0.4,
    "fourthorderdissipationfactor": 0.2,
    "coarsegriddissipationfactor": 0.75,
    "hartenscorrectionfactor": 0.0,
    "usematrixdissipation": False,
    "usedissipationweighting": False,
    "residualsmoothingfactor": 0.0,
    "multigridscheme": 3,
    "numberofrelaxationsteps": 1,
    "numberofgridstouse": 1,
    "numberofrssteps": 0,
    "numberofpssteps": 0,
    "multigridbcrelaxation": 1.0,
    "turbulencemodel": 0,
    "numberofturbulencegrids": 1,
    "turbulentcflnumber": 1.0,
    "tripfactor": 1.0,
    "referencearea": 1.0,
    "forcecalctype": 1,
    "flowtype": 0,
    "addtimestep": True,
    "timescheme": 2,
    "physicaltimestep": 0.1,
    "maximumnumberofcycles": 10000,
    "restartnumber": 1,
    "fluxtype": 1,
    "hllcfluxorder": 2,
    "hllcclipping": 1.0,
    "automaticlipping": 0,
    "gradientmethod": 1,
    "explicit": False,
    "turbprolongationrelaxation": 1.0,
    "prolongationsmoothingfactor": 0.0,
    "prolongationmappingscheme": 1,
    "prolongationrelaxation": 1.0,
    "prandtlnumber": 0.72,
    "turbulentprandtlnumber": 0.9,
    "numberofturbulencesteps": 9999999,
    "maxturbulencevalue": 3000.0,
    "numberoftriggersteps": 0,
    "maxincrementfactor": 0.25,
    "minimumdensity": 0.1,
    "turbk": 0.1,
    "boundaryterm": 1,
    "groundplane": -9999999.9,
    "datum": [0.0, 0.0, 0.0],
    "usetimeresidual": False,
}


def _parse_value(text):
    t = text.strip().rstrip(",")
    lo = t.lower()
    if lo in (".true.", "t", ".t."):
        return True
    if lo in (".false.", "f", ".f."):
        return False
    if t.startswith(("'", '"')):
        return t.strip("'\"")
    if "," in t:
        return [_parse_value(x) for x in t.split(",") if x.strip()]
    v = lo.replace("d", "e")
    try:
        if re.fullmatch(r"[+-]?\d+", v):
            return int(v)
        return float(v)
    except ValueError:
        return t


def read_namelist(text):
    """Parse `&inputVariables ivd%name = value ... /` into a dict."""
    m = re.search(r"&\s*inputvariables(.*?)(?:^\s*/|\n\s*/|/\s*$)", text,
                  re.S | re.I | re.M)
    body = m.group(1) if m else text
    out = {}
    body = "\n".join(line.split("!")[0] for line in body.splitlines())
    for mm in re.finditer(r"ivd%(\w+)(\([^)]*\))?\s*=\s*([^=]*?)(?=\s*ivd%|\Z)",
                          body, re.S | re.I):
        name = mm.group(1).lower()
        idx = mm.group(2)
        val = _parse_value(mm.group(3).replace("\n", " "))
        if idx:
            out.setdefault(name, {})[idx] = val
        else:
            out[name] = val
    return out


class InputVariables:
    """ivd after readInputVariables (DEF:201-422)."""

    def __init__(self, **overrides):
        self.v = dict(DEFAULTS)
        for k, val in overrides.items():
            self.v[k.lower()] = val
        self._post()

    @classmethod
    def from_file(cls, path):
        with open(path) as f:
            return cls(**read_namelist(f.read()))

    def __getattr__(self, name):
        try:
            return self.__dict__["v"][name.lower()]
        except KeyError as e:
            raise AttributeError(name) from e

    def _post(self):
        v = self.v
        g, M = v["gamma"], v["machnumber"]
        # inflowField(:,2) keeps the pre-read defaults (DEF:238-245)
        inflow2 = [1.0, 1.0, 0.0, 0.0, 1.0, 1.0 * (0.5 ** 2) / g]
        if v["flowtype"] == 0:
            v["maximumnumberofcycles"] = 1000000
            v["addtimestep"] = False
        if v["fluxtype"] == 4:
            v["secondorderdissipationfactor"] = 0.0
            v["fourthorderdissipationfactor"] = 0.0
        PI = 4.0 * math.atan(1.0)
        a = v["alpha"] * PI / 180.0
        b = v["beta"] * PI / 180.0
        inflow1 = [1.0, math.cos(a) * math.cos(b), math.sin(b), math.sin(a),
                   (1.0 / ((g - 1.0) * g * M ** 2)) + 0.5, 1.0 / (g * M ** 2)]
        v["walltemperature_scaled"] = ((1.0 / ((g - 1.0) * M ** 2)) *
                                       (v["walltemperature"] /
                                        v["inflowtemperature"]))
        v["inflowfield"] = [inflow1, inflow2]

    def check_supported(self):
        v = self.v
        bad = []
        if v["flowtype"] not in (0, 1, 2):
            bad.append("flowType other than 0-2")
        if v["flowtype"] != 0 and v["explicit"]:
            bad.append("explicit time stepping")
        if v["turbulencemodel"] not in (0, 1):
            bad.append("turbulenceModel other than 0/1")
        if v["fluxtype"] not in (1, 2, 3, 4):
            bad.append("fluxType other than 1-4")
        if v["fluxtype"] != 4 and (v["dissipationscheme"] != 2 or v["coarsegriddissipationscheme"] != 2):
            bad.append("dissipationScheme / coarseGridDissipationScheme other than 2 (JST)")
        if v["fluxtype"] != 4 and v["usedissipationweighting"]:
            bad.append("useDissipationWeighting")
        if v["hllcfluxorder"] == 2 and v["gradientmethod"] != 1:
            bad.append("least-squares gradients")
        if v["viscosityscheme"] != 1:
            bad.append("viscosityScheme 2")
        if v["residualsmoothingfactor"] > 0 or v["prolongationsmoothingfactor"] > 0:
            bad.append("residual smoothing")
        if v["movingwall"] or v["patchinitialization"]:
            bad.append("moving walls / patch initialisation")
        if v["multigridscheme"] not in (0, 1, 3) and v["multigridscheme"] > 0:
            bad.append("multigridScheme other than 0, 1, 3")
        if bad:
            raise NotImplementedError("outside the implemented path: " +
                                      ", ".join(bad))

    def to_struct(self, coarse_fp32=False, rk=None):
        """flite_params; rk = stage factors (alpha_1..alpha_s, s <= 5) to
        replace the reference's 3-stage (0.6, 0.6, 1.0)."""
        v = self.v
        p = FliteParams()
        p.gamma = v["gamma"]
        p.mach_number = v["machnumber"]
        p.reynolds_number = v["reynoldsnumber"]
        p.prandtl_number = v["prandtlnumber"]
        p.turbulent_prandtl_number = v["turbulentprandtlnumber"]
        p.alpha = v["alpha"]
        p.beta = v["beta"]
        for k in range(2):
            for j in range(6):
                p.inflow_field[k][j] = v["inflowfield"][k][j]
        p.inflow_temperature = v["inflowtemperature"]
        p.walls_are_isentropic = int(bool(v["wallsareisentropic"]))
        p.turbulence_model = v["turbulencemodel"]
        p.flux_type = v["fluxtype"]
        p.hllc_flux_order = v["hllcfluxorder"]
        p.automatic_clipping = v["automaticlipping"]
        p.gradient_method = v["gradientmethod"]
        p.hllc_clipping = v["hllcclipping"]
        p.boundary_term = v["boundaryterm"]
        p.cfl_number = v["cflnumber"]
        p.turbulent_cfl_number = v["turbulentcflnumber"]
        alphas = list(rk) if rk is not None else [0.6, 0.6, 1.0]
        if not 1 <= len(alphas) <= 5:
            raise ValueError("1 to 5 Runge-Kutta stages")
        p.rk_coefficients[:] = alphas + [0.0] * (5 - len(alphas))
        p.rk_stages = len(alphas)
        p.max_increment_factor = v["maxincrementfactor"]
        p.minimum_density = v["minimumdensity"]
        p.max_turbulence_value = v["maxturbulencevalue"]
        p.turb_k = v["turbk"]
        p.second_order_dissipation_factor = v["secondorderdissipationfactor"]
        p.fourth_order_dissipation_factor = v["fourthorderdissipationfactor"]
        p.multigrid_scheme = v["multigridscheme"]
        p.number_of_relaxation_steps = v["numberofrelaxationsteps"]
        p.number_of_turbulence_grids = v["numberofturbulencegrids"]
        p.number_of_turbulence_steps = v["numberofturbulencesteps"]
        p.prolongation_relaxation = v["prolongationrelaxation"]
        p.turb_prolongation_relaxation = v["turbprolongationrelaxation"]
        p.multigrid_bc_relaxation = v["multigridbcrelaxation"]
        p.reference_area = v["referencearea"]
        p.groundplane = v["groundplane"]
        d = v["datum"]
        if isinstance(d, dict):
            d = [d.get(f"({i})", 0.0) for i in (1, 2, 3)]
        p.datum[:] = list(d)
        p.coarse_fp32 = int(bool(coarse_fp32))
        p.coarse_grid_dissipation_factor = v["coarsegriddissipationfactor"]
        p.dissipation_scheme = v["dissipationscheme"]
        p.coarse_grid_dissipation_scheme = v["coarsegriddissipationscheme"]
        p.use_matrix_dissipation = int(bool(v["usematrixdissipation"]))
        p.hartens_correction_factor = v["hartenscorrectionfactor"]
        p.flow_type = v["flowtype"]
        p.time_scheme = v["timescheme"]
        p.add_time_step = int(bool(v["addtimestep"]))
        p.physical_time_step = v["physicaltimestep"]
        return p


class FliteParams(ctypes.Structure):
    """flite_params (include/flite3d_amd.h)."""
    _fields_ = [
        ("gamma", ctypes.c_double),
        ("mach_number", ctypes.c_double),
        ("reynolds_number", ctypes.c_double),
        ("prandtl_number", ctypes.c_double),
        ("turbulent_prandtl_number", ctypes.c_double),
        ("alpha", ctypes.c_double),
        ("beta", ctypes.c_double),
        ("inflow_field", (ctypes.c_double * 6) * 2),
        ("inflow_temperature", ctypes.c_double),
        ("walls_are_isentropic", ctypes.c_int),
        ("turbulence_model", ctypes.c_int),
        ("flux_type", ctypes.c_int),
        ("hllc_flux_order", ctypes.c_int),
        ("automatic_clipping", ctypes.c_int),
        ("gradient_method", ctypes.c_int),
        ("hllc_clipping", ctypes.c_double),
        ("boundary_term", ctypes.c_int),
        ("cfl_number", ctypes.c_double),
        ("turbulent_cfl_number", ctypes.c_double),
        ("rk_coefficients", ctypes.c_double * 5),
        ("rk_stages", ctypes.c_int),
        ("max_increment_factor", ctypes.c_double),
        ("minimum_density", ctypes.c_double),
        ("max_turbulence_value", ctypes.c_double),
        ("turb_k", ctypes.c_double),
        ("second_order_dissipation_factor", ctypes.c_double),
        ("fourth_order_dissipation_factor", ctypes.c_double),
        ("multigrid_scheme", ctypes.c_int),
        ("number_of_relaxation_steps", ctypes.c_int),
        ("number_of_turbulence_grids", ctypes.c_int),
        ("number_of_turbulence_steps", ctypes.c_int),
        ("prolongation_relaxation", ctypes.c_double),
        ("turb_prolongation_relaxation", ctypes.c_double),
        ("multigrid_bc_relaxation", ctypes.c_double),
        ("reference_area", ctypes.c_double),
        ("groundplane", ctypes.c_double),
        ("datum", ctypes.c_double * 3),
        ("coarse_fp32", ctypes.c_int),
        ("coarse_grid_dissipation_factor", ctypes.c_double),
        ("dissipation_scheme", ctypes.c_int),
        ("coarse_grid_dissipation_scheme", ctypes.c_int),
        ("use_matrix_dissipation", ctypes.c_int),
        ("hartens_correction_factor", ctypes.c_double),
        ("flow_type", ctypes.c_int),
        ("time_scheme", ctypes.c_int),
        ("add_time_step", ctypes.c_int),
        ("physical_time_step", ctypes.c_double),
    ]


class FliteLevelView(ctypes.Structure):
    """flite_level_view (include/flite3d_amd.h)."""
    _fields_ = [
        ("index_base", ctypes.c_int),
        ("n_nodes", ctypes.c_int),
        ("n_sides", ctypes.c_int),
        ("n_bsides", ctypes.c_int),
        ("n_bnodes", ctypes.c_int),
        ("side_index", ctypes.POINTER(ctypes.c_int)),
        ("side_weights", ctypes.POINTER(ctypes.c_double)),
        ("bside_index", ctypes.POINTER(ctypes.c_int)),
        ("bside_weights", ctypes.POINTER(ctypes.c_double)),
        ("node_volume", ctypes.POINTER(ctypes.c_double)),
        ("coordinates", ctypes.POINTER(ctypes.c_double)),
        ("wall_distance", ctypes.POINTER(ctypes.c_double)),
        ("node_indicator_register", ctypes.POINTER(ctypes.c_int)),
        ("node_indicator_array", ctypes.POINTER(ctypes.c_int)),
        ("n_fine", ctypes.c_int),
        ("prolongation", ctypes.POINTER(ctypes.c_int)),
        ("restriction", ctypes.POINTER(ctypes.c_double)),
        ("side_gcl", ctypes.POINTER(ctypes.c_double)),
        ("bside_gcl", ctypes.POINTER(ctypes.c_double)),
        ("node_volume_p", ctypes.POINTER(ctypes.c_double)),
        ("node_volume_p2", ctypes.POINTER(ctypes.c_double)),
        ("coordinate_movement", ctypes.POINTER(ctypes.c_double)),
    ]


def level_view(level, is_fine):
    """Build a FliteLevelView over a mesh.Level (0-based); keeps refs alive."""
    import numpy as np
    keep = []

    def P(a, ct):
        a = np.ascontiguousarray(a)
        keep.append(a)
        return a.ctypes.data_as(ctypes.POINTER(ct))

    v = FliteLevelView()
    v.index_base = 0
    v.n_nodes = level.n_nodes
    v.n_sides = level.n_sides
    v.n_bsides = level.n_bsides
    v.n_bnodes = level.n_bnodes
    v.side_index = P(level.side_index.astype(np.int32), ctypes.c_int)
    v.side_weights = P(level.side_w.astype(np.float64), ctypes.c_double)
    v.bside_index = P(level.bside_index.astype(np.int32), ctypes.c_int)
    v.bside_weights = P(level.bside_w.astype(np.float64), ctypes.c_double)
    v.node_volume = P(level.volume.astype(np.float64), ctypes.c_double)
    v.coordinates = (P(level.coords.astype(np.float64), ctypes.c_double)
                     if is_fine else None)
    v.wall_distance = P(level.wall_dist.astype(np.float64), ctypes.c_double)
    v.node_indicator_register = P(level.reg.astype(np.int32), ctypes.c_int)
    v.node_indicator_array = P(level.node_ind_array.astype(np.int32),
                               ctypes.c_int)
    if level.f2c is not None:
        v.n_fine = level.f2c.shape[0]
        v.prolongation = P(level.f2c.astype(np.int32), ctypes.c_int)
        v.restriction = P(level.rweight.astype(np.float64), ctypes.c_double)
    else:
        v.n_fine = 0
    # moving mesh (flowType 2): optional attributes of the level
    for name, attr in (("side_gcl", "gcl"), ("bside_gcl", "bgcl"), ("node_volume_p", "volume_p"),
                       ("node_volume_p2", "volume_p2"), ("coordinate_movement", "coord_move")):
        a = getattr(level, attr, None)
        if a is not None:
            setattr(v, name, P(np.asarray(a, dtype=np.float64), ctypes.c_double))
    v._keep = keep
    return v


FIELD_NCOMP = {0: 6, 1: 6, 2: 6, 3: 5, 4: 1, 5: 1, 6: 1, 7: 1, 8: 1, 9: 1,
               10: 6, 11: 1, 12: 25, 13: 3, 14: 6, 15: 6}
FIELDS = dict(U=0, UPREV=1, RHS=2, DISSIPATION=3, PRESSURE=4, DT=5,
              DT_TURB=6, LAMINAR_VISCOSITY=7, VORTICITY=8, DIVERGENCE=9,
              SOURCE=10, TURB_DIFFUSION=11, GRADIENT=12, WALL_STRESS=13,
              UOLD=14, UOLD2=15)
OPS = dict(PRESSURE=0, PRESSURE_ABS=1, TIMESTEPS=2, DISSIPATION=3, VISCOUS=4,
           RHS=5, SA_TIMESTEPS=6, SA_DIFFUSION=7)
