"""A CompLaB3D run from its own input files: CompLaB.xml and geometry.dat.

Reads the tags of the coupled reaction step the way the reference's readers
do -- initialize_complab (complab_functions.hh:825-1600: domain, numerics,
chemistry, microbiology, equilibrium), the immobile / precipitation /
dissolution blocks (complab.cpp:440-574) and the diagnostics block
(complab3d_integration.hh:375-423) -- with the same defaults, and fills a
parameter block, the flow and transport set-up and the geometry.

Two things the files do not carry: the rate laws (the reference compiles them
from defineKinetics.hh / defineAbioticKinetics.hh / defineDissolutionRate into
the executable), which the caller sets on `params` (bio_law, mu_max, rxn_*,
dissol_k[phase], ...); and the kinetics families this build runs (dual Monod,
mass action, the dissolution power law of each phase) bound what those can
be.  A run the path cannot take -- a finite-difference or LB biomass solver,
FBA or surrogate reactions, more than four dissolvable phases -- is refused
with the reason.

readGeometry reads one x slice at a time through Palabos' stream operator,
which walks y and then z fastest; the reference's generator (writeDat,
complab3d_geometry.hh:260-270) writes x fastest.  `read_geometry_dat` follows
the reader by default (order="slices") and takes the generator's order on
request (order="x_fastest").
"""
import os
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import params as P
from .driver import FlowSetup, TransportSetup

_MISSING = object()
_TRUE = ("true", "1", "yes", "on")


class InputError(ValueError):
    """The input files describe a run this path cannot take, or are malformed."""


class _Doc:
    def __init__(self, root):
        self.root = root

    def text(self, *path, default=_MISSING):
        node = self.root
        for tag in path:
            node = node.find(tag) if node is not None else None
        if node is None or node.text is None or not node.text.strip():
            if default is _MISSING:
                raise InputError("<%s> is required" % "/".join(path))
            return default
        return node.text.strip()

    def has(self, *path):
        return self.text(*path, default=None) is not None

    def num(self, *path, default=_MISSING, cast=float):
        t = self.text(*path, default=None)
        if t is None:
            if default is _MISSING:
                raise InputError("<%s> is required" % "/".join(path))
            return default
        try:
            return cast(float(t.split()[0])) if cast is int else cast(t.split()[0])
        except ValueError:
            raise InputError("<%s> is not a number: %r" % ("/".join(path), t))

    def nums(self, *path, default=_MISSING):
        t = self.text(*path, default=None)
        if t is None:
            if default is _MISSING:
                raise InputError("<%s> is required" % "/".join(path))
            return default
        return [float(x) for x in t.split()]

    def flag(self, *path, default=False):
        t = self.text(*path, default=None)
        return default if t is None else t.lower() in _TRUE


def read_geometry_dat(path, nx, ny, nz, order="slices"):
    """The material field of an nx x ny x nz geometry.dat as an int8 array
    (nz, ny, nx + 2): the two ghost x columns copy the first and last slices,
    as readGeometry does (complab_functions.hh:178-207)."""
    with open(path) as f:
        vals = np.array(f.read().split(), dtype=np.int64)
    if vals.size != nx * ny * nz:
        raise InputError("%s holds %d values, the domain %dx%dx%d needs %d" % (path, vals.size, nx, ny, nz, nx * ny * nz))
    if order == "slices":                       # x slowest, z fastest
        core = vals.reshape(nx, ny, nz).transpose(2, 1, 0)
    elif order == "x_fastest":
        core = vals.reshape(nz, ny, nx)
    else:
        raise InputError("order must be 'slices' or 'x_fastest'")
    g = np.empty((nz, ny, nx + 2), dtype=np.int64)
    g[:, :, 1:-1] = core
    g[:, :, 0] = core[:, :, 0]
    g[:, :, -1] = core[:, :, -1]
    return g


def _btype(s, where):
    s = s.lower()
    if s == "dirichlet":
        return P.BC_DIRICHLET
    if s == "neumann":
        return P.BC_NEUMANN
    if s in ("closed", "noflux", "no_flux", "wall"):
        return P.BC_CLOSED
    raise InputError("%s (%s) should be Dirichlet, Neumann or closed" % (where, s))


# ---------------------------------------------------------------------------------------------
# .thm energetics files (complab3d_thermo.hh:299-508): `key value...` lines, `#` comments;
# file keys provenance, version, temperature (K), energy_units (kJ | J), gas_constant
# (kJ/mol/K), concentration_scale; blocks opened by `reaction NAME` with microbe, stoich
# (name/number pairs), dG0, atp, dGatp, chi, cmin, dGloss and the yield keys
# ---------------------------------------------------------------------------------------------
@dataclass
class ThmReaction:
    name: str
    microbe: str = ""
    species: List[str] = field(default_factory=list)
    nu: List[float] = field(default_factory=list)
    dG0: Optional[float] = None
    atp: float = 0.0
    dGatp: float = 50.0
    chi: float = 1.0
    cmin: float = 1e-9
    dGloss: float = 0.0
    yield_on: bool = False
    numbers: dict = field(default_factory=dict)        # dGana, carbon, reduction, dGdis, yieldmax


@dataclass
class ThmModel:
    provenance: str = ""
    temperature: float = 298.15
    unit_scale: float = 1.0                             # to kJ/mol
    gas_constant: float = 8.314462618e-3
    conc_scale: float = 1.0
    reactions: List[ThmReaction] = field(default_factory=list)

    def rt(self) -> float:
        return self.gas_constant * self.temperature


_THM_NUMBERS = ("dG0", "atp", "dGatp", "chi", "cmin", "dGloss", "dGana", "carbon", "reduction", "dGdis", "yieldmax")


def read_thm(path) -> ThmModel:
    """A .thm file, strictly: an unknown keyword, a malformed line, a block
    without stoich, chi <= 0, cmin <= 0 or a negative dGloss is an error."""
    M = ThmModel()
    cur = None
    with open(path) as fh:
        for ln, line in enumerate(fh, 1):
            t = line.split("#", 1)[0].split()
            if not t:
                continue
            at = " (%s line %d)" % (os.path.basename(path), ln)
            key = t[0]

            def one_number(positive=False):
                if len(t) != 2:
                    raise InputError("%s needs exactly one number%s" % (key, at))
                try:
                    v = float(t[1])
                except ValueError:
                    raise InputError("%s needs exactly one number%s" % (key, at))
                if positive and not v > 0:
                    raise InputError("%s needs one positive number%s" % (key, at))
                return v
            if key == "provenance":
                M.provenance = " ".join(t[1:])
            elif key == "version":
                pass
            elif key == "temperature":
                M.temperature = one_number(True)
            elif key == "energy_units":
                if len(t) != 2 or t[1] not in ("kJ", "kj", "J", "j"):
                    raise InputError("energy_units is kJ or J%s" % at)
                M.unit_scale = 1.0 if t[1] in ("kJ", "kj") else 1e-3
            elif key == "concentration_scale":
                M.conc_scale = one_number(True)
            elif key == "gas_constant":
                M.gas_constant = one_number(True)
            elif key == "reaction":
                if len(t) != 2:
                    raise InputError("reaction needs exactly one name%s" % at)
                if any(r.name == t[1] for r in M.reactions):
                    raise InputError("two reactions are both named '%s'%s" % (t[1], at))
                cur = ThmReaction(t[1])
                M.reactions.append(cur)
            elif cur is None:
                raise InputError("'%s' appears before any reaction block%s" % (key, at))
            elif key == "stoich":
                if len(t) < 3 or (len(t) - 1) % 2:
                    raise InputError("stoich needs name/number pairs%s" % at)
                for name, num in zip(t[1::2], t[2::2]):
                    try:
                        v = float(num)
                    except ValueError:
                        raise InputError("'%s' is not a number in the stoich line%s" % (num, at))
                    if v == 0.0:
                        raise InputError("species '%s' has a zero coefficient%s" % (name, at))
                    cur.species.append(name)
                    cur.nu.append(v)
            elif key == "microbe":
                if len(t) != 2:
                    raise InputError("microbe needs one name, index or 'all'%s" % at)
                cur.microbe = t[1]
            elif key == "yield":
                if len(t) != 2 or t[1].lower() not in set(_TRUE) | {"off", "false", "no", "0"}:
                    raise InputError("yield needs on or off%s" % at)
                cur.yield_on = t[1].lower() in _TRUE
            elif key in _THM_NUMBERS:
                v = one_number()
                if key in ("dG0", "atp", "dGatp", "chi", "cmin", "dGloss"):
                    setattr(cur, key, v)
                else:
                    cur.numbers[key] = v
            else:
                raise InputError("'%s' is not a keyword this file understands%s" % (key, at))
    if not M.reactions:
        raise InputError("%s declares no reaction blocks" % path)
    for r in M.reactions:
        if not r.species:
            raise InputError("reaction '%s' has no stoich line" % r.name)
        if r.dG0 is None:
            raise InputError("reaction '%s' has no dG0" % r.name)
        if not r.chi > 0:
            raise InputError("reaction '%s' has chi <= 0" % r.name)
        if not r.cmin > 0:
            raise InputError("reaction '%s' has cmin <= 0, so ln Q can be -inf" % r.name)
        if r.dGloss < 0:
            raise InputError("reaction '%s' has a negative dGloss" % r.name)
        r.dG0 *= M.unit_scale
        r.dGatp *= M.unit_scale
        r.dGloss *= M.unit_scale
    return M


def apply_thm(p, M: ThmModel, species: List[str], microbes: List[str], kinetics_in_use: bool = True):
    """Binds the model to the run (species and microbe names) and sets the
    parameter block's gate.  The compiled-kinetics path computes one rate
    vector per voxel, so while an organism takes its rates from it the file
    holds one reaction block, for microbe0 or all of them
    (complab3d_integration.hh:838-858).  A block naming one organism gates
    that organism alone (specOfMicrobe, complab3d_thermo.hh:640-648); this
    path carries one gate, so one block in every case."""
    if len(M.reactions) > 1:
        if kinetics_in_use:
            raise InputError("the energetics file declares more than one reaction block; the kinetics path "
                             "computes one combined rate vector per voxel and takes one gate")
        raise InputError("the energetics file declares more than one reaction block; this path carries one gate")
    r = M.reactions[0]
    p.thm_microbe = -1
    if r.microbe not in ("", "all", "*"):
        idx = microbes.index(r.microbe) if r.microbe in microbes else (
            int(r.microbe) if r.microbe.isdigit() and int(r.microbe) < len(microbes) else -1)
        if idx < 0:
            raise InputError("reaction '%s' names microbe '%s', which is not in <name_of_microbes>"
                             % (r.name, r.microbe))
        if idx != 0 and kinetics_in_use:
            raise InputError("reaction '%s' is for an organism other than microbe0, but an organism takes its "
                             "rates from the kinetics path, which gates only as microbe0 (write `microbe all`)"
                             % r.name)
        p.thm_microbe = idx
    for s in range(len(p.thm_nu)):
        p.thm_nu[s] = 0.0
    for name, v in zip(r.species, r.nu):
        if name not in species:
            raise InputError("reaction '%s' uses species '%s', which is not in <name_of_substrates>. This run has: %s"
                             % (r.name, name, " ".join(species)))
        p.thm_nu[species.index(name)] += v
    p.thm_on = 1
    p.thm_rt = M.rt()
    p.thm_dg0 = r.dG0
    p.thm_dgloss = r.dGloss
    p.thm_threshold = r.atp * r.dGatp
    p.thm_chi = r.chi
    p.thm_cmin = r.cmin
    p.thm_conc_scale = M.conc_scale


@dataclass
class Run:
    params: P.CaseParams
    geometry: np.ndarray                        # (nz, ny, nx + 2) in this build's material codes
    flow: FlowSetup
    transport: TransportSetup
    steps: int                                  # ade_max_iT
    species: List[str]
    microbes: List[str]
    dx: float                                   # metres
    diag_enabled: bool = False
    diag_interval: int = 0
    diag_tol: float = 1e-6
    conserve: List[str] = field(default_factory=list)
    materials: dict = field(default_factory=dict)   # file material number -> this build's code
    # output (complab_functions.hh:1166-1178, 1465-1474; complab3d_integration.hh:383-384)
    output_path: str = "output"                 # absolute, or as the XML names it under its folder
    vtk_interval: int = 1000                    # <IO><save_VTK_interval>
    chk_interval: int = 1000000                 # <IO><save_CHK_interval>
    mask_filename: str = "maskLattice"
    summary_csv: str = "summary.csv"
    symbolic: dict = field(default_factory=dict)   # microbe index (params.SYM_ABIOTIC: the abiotic law) -> params.SymProgram


def load(xml_path, geometry_path: Optional[str] = None, precision=P.FP64, order="slices") -> Run:
    """Reads CompLaB.xml (and the geometry file it names, under its
    <input_path>, unless `geometry_path` is given) into a Run."""
    doc = _Doc(ET.parse(xml_path).getroot())
    base = os.path.dirname(os.path.abspath(xml_path))
    biotic = doc.flag("simulation_mode", "biotic_mode", default=True)
    if doc.flag("simulation_mode", "enable_fba_glpk") or doc.flag("simulation_mode", "enable_fba_cobrapy") \
            or doc.flag("simulation_mode", "enable_surrogate"):
        raise InputError("FBA and surrogate reactions are not on this path")

    # ---- domain (complab_functions.hh:999-1003, 1195-1233)
    nx = doc.num("LB_numerics", "domain", "nx", cast=int)
    ny = doc.num("LB_numerics", "domain", "ny", cast=int)
    nz = doc.num("LB_numerics", "domain", "nz", cast=int)
    dx = doc.num("LB_numerics", "domain", "dx")
    pe = doc.num("LB_numerics", "Peclet", default=0.0)
    char_len = doc.num("LB_numerics", "domain", "characteristic_length", default=None)
    if char_len is None:
        if pe > 1e-14:
            raise InputError("<characteristic_length> must be given when Peclet > 0")
        char_len = 0.0
    unit = doc.text("LB_numerics", "domain", "unit", default="um")
    scale = {"m": 1.0, "mm": 1e-3, "um": 1e-6}.get(unit)
    if scale is None:
        raise InputError("unit (%s) must be either m, mm, or um" % unit)
    char_len /= dx
    dx_m = dx * scale
    delta_p = doc.num("LB_numerics", "delta_P", default=0.0)
    tau = doc.num("LB_numerics", "tau", default=0.8)
    it = ("LB_numerics", "iteration")
    ns_max1 = doc.num(*it, "ns_max_iT1", default=100000, cast=int)
    ns_max2 = doc.num(*it, "ns_max_iT2", default=100000, cast=int)
    ns_eps1 = doc.num(*it, "ns_converge_iT1", default=1e-8)
    ns_eps2 = doc.num(*it, "ns_converge_iT2", default=1e-6)
    ns_upd = doc.num(*it, "ns_update_interval", default=1, cast=int)
    ade_upd = doc.num(*it, "ade_update_interval", default=1, cast=int)
    steps = doc.num(*it, "ade_max_iT", default=10000000, cast=int)

    # ---- chemistry (complab_functions.hh:1006-1027, 1280-1295; complab.cpp:440-458)
    S = doc.num("chemistry", "number_of_substrates", cast=int)
    if not 1 <= S <= P.MAX_SPECIES:
        raise InputError("number_of_substrates %d outside 1..%d" % (S, P.MAX_SPECIES))
    p = P.default_params(nx + 2, ny, nz, S)
    p.precision = precision
    names, d_pore, d_bio = [], [], []
    for s in range(S):
        tag = "substrate%d" % s
        names.append(doc.text("chemistry", tag, "name_of_substrates", default=tag))
        p.c0[s] = doc.num("chemistry", tag, "initial_concentration", default=0.0)
        d_pore.append(doc.num("chemistry", tag, "substrate_diffusion_coefficients", "in_pore", default=0.0))
        d_bio.append(doc.num("chemistry", tag, "substrate_diffusion_coefficients", "in_biofilm", default=0.0))
        p.left_btype[s] = _btype(doc.text("chemistry", tag, "left_boundary_type"), tag + " left_boundary_type")
        p.right_btype[s] = _btype(doc.text("chemistry", tag, "right_boundary_type"), tag + " right_boundary_type")
        p.left_bc[s] = doc.num("chemistry", tag, "left_boundary_condition")
        p.right_bc[s] = doc.num("chemistry", tag, "right_boundary_condition")
        p.mobile[s] = 0 if doc.flag("chemistry", tag, "immobile") else 1

    # ---- materials (complab_functions.hh:1217-1233)
    pore_mats = [int(x) for x in doc.nums("LB_numerics", "domain", "material_numbers", "pore", default=[2])]
    solid_mat = doc.num("LB_numerics", "domain", "material_numbers", "solid", default=0, cast=int)
    bb_mat = doc.num("LB_numerics", "domain", "material_numbers", "bounce_back", default=1, cast=int)
    mats = {solid_mat: P.SOLID, bb_mat: P.WALL}
    for m in pore_mats:
        mats[m] = P.PORE

    # ---- microbiology (complab_functions.hh:1036-1440)
    microbes = []
    M = doc.num("microbiology", "number_of_microbes", default=0, cast=int) if biotic else 0
    if M > P.MAX_MICROBES:
        raise InputError("%d microbes, this path holds %d" % (M, P.MAX_MICROBES))
    perm = []
    rtypes = []
    for m in range(M):
        tag = "microbe%d" % m
        solver = doc.text("microbiology", tag, "solver_type").lower()
        if solver not in ("ca", "cellular automata", "cellular_automata"):
            raise InputError("%s: solver_type %s; only cellular-automaton biofilms are on this path" % (tag, solver))
        rt = doc.text("microbiology", tag, "reaction_type", default="kinetics").lower()
        if rt not in ("kinetics", "none", "symbolic"):
            raise InputError("%s: reaction_type %s is not on this path" % (tag, rt))
        rtypes.append(rt)
        microbes.append(doc.text("microbiology", tag, "name_of_microbes", default=tag))
        b0 = doc.nums("microbiology", tag, "initial_densities")
        bio_mats = doc.nums("LB_numerics", "domain", "material_numbers", tag, default=None)
        if bio_mats is None:
            raise InputError("%s has no material number: planktonic biomass is not on this path" % tag)
        if len(bio_mats) != len(b0):
            raise InputError("%s: %d material numbers for %d initial densities" % (tag, len(bio_mats), len(b0)))
        if len(set(b0)) != 1:
            raise InputError("%s: one initial density per microbe on this path" % tag)
        for bm in bio_mats:
            mats[int(bm)] = P.BIOFILM + m
        p.b0[m] = b0[0]
        p.kdecay[m] = doc.num("microbiology", tag, "decay_coefficient", default=0.0)
        ratio = doc.num("microbiology", tag, "viscosity_ratio_in_biofilm", default=None)
        if ratio is None:
            raise InputError("%s: viscosity_ratio_in_biofilm must be given for a CA biofilm" % tag)
        perm.append(1.0 / ratio if ratio > 1e-14 else -99.0)
    p.n_microbes = M

    # ---- symbolic rate laws (complab3d_integration.hh:264-300; chooseFiles,
    # complab3d_symbolic.hh:910-965): a microbe's own <expressions_file>, else
    # the shared <symbolic><expressions_file>
    symbolic = {}
    if "symbolic" in rtypes:
        if not doc.flag("symbolic", "enabled"):
            raise InputError("a microbe has reaction_type symbolic but <symbolic><enabled> is not true")
        shared = doc.text("symbolic", "expressions_file", default=None)
        from . import symbolic as SYM
        for m, rt in enumerate(rtypes):
            if rt != "symbolic":
                continue
            f = doc.text("microbiology", "microbe%d" % m, "expressions_file", default=None) or shared
            if not f:
                raise InputError("%s has reaction_type symbolic but no rate law" % microbes[m])
            try:
                law = SYM.read_sym(f if os.path.isabs(f) else os.path.join(base, f))
                symbolic[m] = SYM.compile_law(law, names, microbes[m])
            except SYM.SymError as e:
                raise InputError(str(e))
    if M > 0:
        if len(set(perm)) != 1:
            raise InputError("one biofilm viscosity ratio for every microbe on this path")
        p.perm_ratio = perm[0] if perm[0] > 0 else 0.0
        p.thrd_bfilm_frac = doc.num("microbiology", "thrd_biofilm_fraction")
        p.bmax = doc.num("microbiology", "maximum_biomass_density")
        ca = doc.text("microbiology", "CA_method", default="fraction").lower()
        if ca in ("fraction", "0", "no"):
            p.ca_method = P.CA_FRACTION
        elif ca in ("half", "1", "yes"):
            p.ca_method = P.CA_HALF
        else:
            raise InputError("CA_method (%s) should be either half or fraction" % ca)

    # ---- flow
    p.ns_tau = tau
    p.delta_p = delta_p
    p.flow_on = 1 if pe > 1e-14 else 0
    p.ns_update_interval = ns_upd
    p.ade_update_interval = ade_upd
    p.ns_max_iter = ns_max2
    p.ns_conv_rule = P.CONV_TRACER
    p.ns_conv_eps = ns_eps2
    flow = FlowSetup(pe=pe, char_length=char_len, delta_p=delta_p, ns_tau=tau, max_iter=ns_max1,
                     rule=P.CONV_TRACER, eps=ns_eps1, has_biofilm=M > 0)
    transport = TransportSetup(d_pore=d_pore, d_biofilm=d_bio, dx=dx_m)

    # ---- simulation mode (complab_functions.hh:865-990)
    p.abiotic_rate_scale = doc.num("simulation_mode", "abiotic_rate_scale", default=1.0)

    # ---- equilibrium (complab_functions.hh:1484-1582; complab.cpp:609-618)
    if doc.flag("equilibrium", "enabled"):
        comps = doc.text("equilibrium", "components", default="").split()
        nc = len(comps)
        if not 1 <= nc <= P.MAX_COMP:
            raise InputError("equilibrium needs 1..%d components" % P.MAX_COMP)
        p.eq_on = 1
        p.eq_solver = P.EQ_AAPCF
        p.eq_nc = nc
        p.eq_tol = 1e-10
        p.eq_max_iter = 200
        p.eq_anderson_depth = 4
        rows = []
        for s in range(S):
            row = doc.nums("equilibrium", "stoichiometry", "species%d" % s, default=None)
            if row is None or len(row) != nc:
                row = [0.0] * nc
            rows.append(row)
            for j in range(nc):
                p.eq_stoich[s][j] = row[j]
            p.eq_logk[s] = doc.num("equilibrium", "logK", "species%d" % s, default=0.0)
        for j, cname in enumerate(comps):
            k = names.index(cname) if cname in names else -1
            if k < 0:
                unit_rows = [s for s in range(S) if all(rows[s][i] == (1.0 if i == j else 0.0) for i in range(nc))]
                k = unit_rows[0] if unit_rows else -1
            p.eq_comp_species[j] = k

    # ---- precipitation (complab.cpp:485-503)
    if doc.flag("precipitation", "enabled"):
        p.precip_on = 1
        p.precip_species = doc.num("precipitation", "solid_substrate", default=-1, cast=int)
        p.precip_max_rho = doc.num("precipitation", "max_precipRho", default=1e30)
        p.surface_only = doc.num("precipitation", "surface_only", default=1, cast=int)
        p.precip_perm_ratio = doc.num("precipitation", "perm_ratio", default=0.0)
        p.precip_interval = doc.num("precipitation", "update_interval", default=200, cast=int)

    # ---- dissolution (complab.cpp:516-574): declared phases <phase0>, <phase1>, ...
    if doc.flag("dissolution", "enabled"):
        phases = []
        for k in range(64):
            if not doc.has("dissolution", "phase%d" % k, "material_number"):
                break
            phases.append(k)
        if not phases:
            raise InputError("<dissolution> is enabled but no <phase0> block was given")
        if len(phases) > P.MAX_PHASES:
            raise InputError("%d dissolvable phases; this path takes up to %d" % (len(phases), P.MAX_PHASES))
        p.dissol_on = 1
        p.dissol_n_phases = len(phases)
        p.dissol_reopen_frac = doc.num("dissolution", "reopen_fraction", default=0.9)
        p.dissol_surface_only = doc.num("dissolution", "surface_only", default=1, cast=int)
        p.dissol_interval = doc.num("dissolution", "update_interval", default=100, cast=int)
        if not 0 < p.dissol_reopen_frac < 1:
            raise InputError("<reopen_fraction> must be strictly between 0 and 1")
        for k in phases:
            ph = ("dissolution", "phase%d" % k)
            mat = doc.num(*ph, "material_number", cast=int)
            precipitate = doc.flag(*ph, "is_precipitate")
            if not precipitate and mats.get(mat) not in (P.SOLID, P.WALL):
                raise InputError("grain phase %d lies on material %d; on this path a grain phase lies on the "
                                 "solid or bounce-back material" % (k, mat))
            p.dissol_is_precipitate[k] = 1 if precipitate else 0
            p.dissol_material[k] = -1 if precipitate else mats[mat]
            p.dissol_species[k] = doc.num(*ph, "substrate", default=-1, cast=int)
            p.dissol_full[k] = doc.num(*ph, "full_density", default=0.0)
            p.dissol_initial[k] = 0.0 if precipitate else doc.num(*ph, "initial_fill", default=0.0)
            if not 0 <= p.dissol_species[k] < S:
                raise InputError("dissolution substrate %d is outside 0..%d" % (p.dissol_species[k], S - 1))
            if not p.dissol_full[k] > 0:
                raise InputError("<full_density> of phase %d must be positive" % k)
            if precipitate:
                p.precip_phase_id = k + 1          # the last precipitate phase declared (complab.cpp:530)

    # ---- thermodynamic gate (complab3d_integration.hh:315-326, 810-860)
    if doc.flag("thermodynamics", "enabled"):
        f = doc.text("thermodynamics", "energetics_file", default=None)
        if not f:
            raise InputError("<thermodynamics> is enabled but <energetics_file> is not set")
        kinetics_in_use = biotic and any(rt == "kinetics" for rt in rtypes)
        apply_thm(p, read_thm(f if os.path.isabs(f) else os.path.join(base, f)), names, microbes, kinetics_in_use)

    # ---- diagnostics (complab3d_integration.hh:375-423)
    run_diag = doc.flag("diagnostics", "enabled")
    conserve = []
    if run_diag:
        for k in range(32):
            e = doc.text("diagnostics", "conserve" if k == 0 else "conserve%d" % k, default=None)
            if e is None:
                break
            conserve.append(e)

    # ---- geometry
    if geometry_path is None:
        gdir = doc.text("path", "input_path", default="input")
        geometry_path = os.path.join(base, gdir, doc.text("LB_numerics", "domain", "filename"))
    raw = read_geometry_dat(geometry_path, nx, ny, nz, order=order)
    unknown = sorted(set(np.unique(raw).tolist()) - set(mats))
    if unknown:
        raise InputError("geometry holds material numbers %s that the XML does not declare" % unknown)
    geo = np.empty(raw.shape, dtype=np.int8)
    for src, dst in mats.items():
        geo[raw == src] = dst

    # the abiotic symbolic law (<symbolic><abiotic_file>): every open voxel, no
    # organism, its own block whether or not mass-action kinetics are on
    # (complab.cpp:2455-2493)
    ab = doc.text("symbolic", "abiotic_file", default=None) if doc.flag("symbolic", "enabled") else None
    if ab:
        if p.dissol_on:
            raise InputError("<symbolic><abiotic_file> together with dissolution is not combined on this path")
        from . import symbolic as SYM
        try:
            law = SYM.read_sym(ab if os.path.isabs(ab) else os.path.join(base, ab))
            symbolic[P.SYM_ABIOTIC] = SYM.compile_law(law, names, None)
        except SYM.SymError as e:
            raise InputError(str(e))

    if symbolic:
        # the symbolic processors skip fixed substrates and can read speciation
        # totals (complab3d_processors_symbolic.hh:124, 167, from
        # complab3d_metabolic.hh:829-859, 640-655); neither is carried here
        fix = doc.text("chemistry", "fix_concentration", default=None)
        if fix is not None and any(v.lower() in ("yes", "true", "1") for v in fix.split()):
            raise InputError("<chemistry><fix_concentration> holds a substrate fixed; the symbolic rate "
                             "path here does not hold substrates fixed")
        basis = doc.text("simulation_mode", "fba_concentration_basis", default="free").strip().lower()
        if basis == "total":
            raise InputError("<fba_concentration_basis> total: the symbolic rate path here reads free "
                             "concentrations only")
        if basis != "free":
            raise InputError("fba_concentration_basis \"%s\" is not understood. Use free or total" % basis)
    return Run(params=p, geometry=geo, flow=flow, transport=transport, steps=steps, species=names,
               microbes=microbes, dx=dx_m, diag_enabled=run_diag,
               diag_interval=doc.num("diagnostics", "interval", default=0, cast=int),
               diag_tol=doc.num("diagnostics", "tolerance", default=1e-6), conserve=conserve,
               materials=dict(mats),
               output_path=os.path.join(base, doc.text("path", "output_path", default="output")),
               vtk_interval=doc.num("IO", "save_VTK_interval", default=1000, cast=int),
               chk_interval=doc.num("IO", "save_CHK_interval", default=1000000, cast=int),
               mask_filename=doc.text("IO", "mask_filename", default="maskLattice"),
               summary_csv=doc.text("diagnostics", "summary_csv", default="summary.csv"),
               symbolic=symbolic)
