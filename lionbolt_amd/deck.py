"""Input decks (``lb.in``) read as the reference reads them.

Follows ReadInput.f90 (header loop, duplicate and required blocks),
InputReadModule.f90 (each block's keys and messages), ValidateInput.f90 (mesh
and material checks) and InputBuildModule.f90 (mesh, angular set, solver).
The line reader restates NittanyPhysics' ``ReadSplitSimplify(..., ic=.TRUE.)``
(the library is not vendored): text from ``#`` on is a comment, words are split
on white space and compared in lower case; names and file names keep their
case (the reference re-reads those lines with ``ReadAndSplit``).

What the transport path needs from a deck is built here: the tetrahedral mesh
(``MESH GMSH`` + ``TRANSLATE`` + ``SCALE``), the angular set of each particle
(``ANGULAR SN`` with ``PNSCATTERING L`` -> SNPrep(L)), the solver
(``SOLVER GMRES|SI``) and the energy-group edges of ``GRID``.  Cross sections
come from NittanyPhysics in the reference and are supplied by the caller here.
One-dimensional slab problems (``PROBLEM SLAB``) are parsed and validated but
have no tetrahedral mesh to sweep; ``build_mesh`` rejects them.

Errors raise ``DeckError`` with the reference's ``stophere`` message.  Where
the reference keeps a flag across blocks (``axisset`` in BeamsInputRead and
``matgiven`` in SlabInputRead are initialised in their declarations, so
Fortran keeps them between calls), each block here starts afresh; the two
differ only on decks where a later BEAM omits the AXIS an earlier one gave, or
a later SLAB omits the MATERIAL an earlier one gave.
"""
import os
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

# Chemical symbols by atomic number (PeriodicTable, declared at
# InputReadModule.f90:271-275 and supplied by NittanyPhysics).
ELEMENTS = (
    "H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se Br Kr "
    "Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb "
    "Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn Fr Ra Ac Th Pa U Np Pu Am Cm Bk Cf Es Fm Md No Lr "
    "Rf Db Sg Bh Hs Mt Ds Rg Cn Nh Fl Mc Lv Ts Og").split()


def periodic_table(symbol):
    """Atomic number of a chemical symbol, 0 when it is not one (case as written)."""
    try:
        return ELEMENTS.index(symbol) + 1
    except ValueError:
        return 0


class DeckError(ValueError):
    """An input read or validation error (the reference stops with the same text)."""


@dataclass
class SlabInput:               # InputModule.f90:34-42
    material: int = 1
    elements: Optional[int] = None
    thickness: Optional[float] = None
    structure: Optional[str] = None


@dataclass
class MeshInput:               # InputModule.f90:23-32
    fname: Optional[str] = None
    dotrans: bool = False
    doscale: bool = False
    trans: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    scale: Tuple[float, float, float] = (1.0, 1.0, 1.0)
    slabs: List[SlabInput] = field(default_factory=list)


@dataclass
class MaterialInput:           # InputModule.f90:44-53
    name: str = ""
    Z: List[int] = field(default_factory=list)
    N: List[int] = field(default_factory=list)
    dens: float = 0.0


@dataclass
class ParticleInput:           # InputModule.f90:64-90
    name: str = ""
    identity: Optional[str] = None
    gridstruct: Optional[str] = None
    fname: Optional[str] = None
    Emin: float = -1.0
    Emax: float = -1.0
    G: int = -1
    angular: Optional[str] = None
    solver: Optional[str] = None
    noscat: bool = False
    L: int = 0
    slab: bool = False


@dataclass
class BeamInput:               # InputModule.f90:100-118
    particle: Optional[int] = None
    spectrum: Optional[str] = None
    mcgroups: Optional[List[int]] = None
    E0: float = -1.0
    dE: float = -1.0
    weight: float = 1.0
    axis: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    origin: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    cutout: Optional[str] = None
    cutoutparams: Tuple[float, float, float, float] = (0.0, 0.0, 0.0, 0.0)
    angdist: Optional[str] = None
    slab: bool = False


@dataclass
class PostProcInput:           # InputModule.f90:129-138
    fluence: bool = False
    outputunc: bool = False
    energy: bool = False
    dose: bool = False
    charge: bool = False


@dataclass
class OptionsInput:            # InputModule.f90:139-147
    storlim: bool = False
    memlim: bool = False
    debug: bool = False


@dataclass
class Deck:                    # InputModule.f90:149-162
    path: str
    general: bool
    slab: bool
    L: int = 0
    mesh: MeshInput = field(default_factory=MeshInput)
    materials: List[MaterialInput] = field(default_factory=list)
    particles: List[ParticleInput] = field(default_factory=list)
    beams: List[BeamInput] = field(default_factory=list)
    postproc: PostProcInput = field(default_factory=PostProcInput)
    options: OptionsInput = field(default_factory=OptionsInput)

    @property
    def base_dir(self):
        return os.path.dirname(os.path.abspath(self.path))

    def mesh_path(self):
        """The GMSH file, relative names taken from the deck's directory (the
        reference runs in it)."""
        if self.mesh.fname is None:
            return None
        return self.mesh.fname if os.path.isabs(self.mesh.fname) else os.path.join(self.base_dir, self.mesh.fname)


class _Lines:
    """Line source with the reference's splitting rules."""

    def __init__(self, text):
        self.lines = text.splitlines()
        self.pos = 0

    def next(self, where):
        """(lower-case line, lower-case words, true-case line, true-case words);
        DeckError at the end of the file inside a block."""
        got = self.next_or_none()
        if got is None:
            raise DeckError(f"Input read error: End of file inside the {where} block.")
        return got

    def next_or_none(self):
        if self.pos >= len(self.lines):
            return None
        raw = self.lines[self.pos]
        self.pos += 1
        raw = raw.split("#", 1)[0]
        words = raw.split()
        true_line = " ".join(words)
        return true_line.lower(), [w.lower() for w in words], true_line, words


def _reals(words, n, msg):
    try:
        return tuple(float(words[1 + i].replace("d", "e").replace("D", "e")) for i in range(n))
    except (IndexError, ValueError):
        raise DeckError(msg) from None


def _num(word, kind, msg):
    try:
        return kind(word.replace("d", "e").replace("D", "e")) if kind is float else kind(word)
    except (TypeError, ValueError):
        raise DeckError(msg) from None


def _word(words, i):
    return words[i] if i < len(words) else ""


def _read_problem(src):                       # ProblemInputRead, InputReadModule.f90:4-48
    while True:
        got = src.next_or_none()
        if got is None:
            raise DeckError("Input read error: Blank input file.")
        line, words, _, _ = got
        if line:
            break
    if len(words) != 2:
        raise DeckError("Input read error: Either too many or too few words in first input line. "
                        'Use two words - "PROBLEM" followed by the desired problem type. Your input was: ' + line)
    if words[0] != "problem":
        raise DeckError('Input read error: First line must denote solve type. Use key "PROBLEM". '
                        "Key used was: " + words[0])
    if words[1] not in ("general", "slab"):
        raise DeckError("Input read error: Invalid problem type : " + line)
    return words[1] == "general", words[1] == "slab"


def _read_slab(src):                          # SlabInputRead, :127-206
    s = SlabInput()
    matgiven = False
    while True:
        line, words, _, _ = src.next("SLAB")
        key = _word(words, 0)
        if key == "structure":
            if _word(words, 1) not in ("linear", "logarithmic"):
                raise DeckError("Input read error: Unknown structure for SLAB Structure key:" + _word(words, 1))
            s.structure = words[1]
        elif key == "thickness":
            s.thickness = _num(_word(words, 1), float, "Input read error: Missing value for SLAB Thickness key.")
        elif key == "elements":
            s.elements = _num(_word(words, 1), int, "Input read error: Missing value for SLAB Elements key.")
        elif key == "material":
            s.material = _num(_word(words, 1), int, "Input read error: Missing value for SLAB Material key.")
            matgiven = True
        elif key == "end":
            if not matgiven:
                s.material = 1
            return s
        # other keys: the reference's select has no default branch and skips them


def _read_mesh(src):                          # MeshInputRead, :50-125
    m = MeshInput()
    while True:
        line, words, true_line, true_words = src.next("MESH")
        key = _word(words, 0)
        if key == "gmsh":
            if len(words) < 2:
                raise DeckError("Input read error: Missing file name for MESH GMSH key.")
            m.fname = true_line[len(true_words[0]):].strip()
        elif key == "translate":
            m.trans = _reals(words, 3, "Input read error: Unable to read MESH translation vector. Make sure this "
                                       "is given as three separate real numbers.")
            m.dotrans = True
        elif key == "scale":
            m.scale = _reals(words, 3, "Input read error: Unable to read MESH scale vector. Make sure this "
                                       "is given as three separate real numbers.")
            m.doscale = True
        elif key == "slab":
            m.slabs.append(_read_slab(src))
        elif key == "":
            continue
        elif key == "end":
            return m
        else:
            raise DeckError("Input read error: Unknown key in MESH block : " + key)


def _read_material(src, name):                # MaterialInputRead, :257-339
    mat = MaterialInput(name=name)
    while True:
        line, words, _, true_words = src.next("MATERIAL")
        key = _word(words, 0)
        if key == "density":
            mat.dens = _num(_word(words, 1), float, "Input read error: Missing value for MATERIAL Density key.")
        elif key == "":
            continue
        elif key == "end":
            return mat
        else:
            Z = periodic_table(true_words[0])
            if Z == 0:
                raise DeckError("Input read error: Unknown key in a MATERIAL block : " + " ".join(true_words))
            mat.Z.append(Z)
            mat.N.append(_num(_word(true_words, 1), int,
                              "Input read error: Missing atom count in a MATERIAL block : " + " ".join(true_words)))


def _read_materials(src):                     # MaterialsInputRead, :208-255
    mats = []
    while True:
        line, words, true_line, _ = src.next("MATERIALS")
        if line == "":
            continue
        if _word(words, 0) == "end":
            if not mats:
                raise DeckError("Input read error: No materials listed in MATERIALS block.")
            return mats
        mats.append(_read_material(src, true_line))


def _read_angular(src):                       # ReadPNScattering, :341-374
    L = 0
    while True:
        line, words, _, _ = src.next("ANGULAR")
        key = _word(words, 0)
        if key == "pnscattering":
            L = _num(_word(words, 1), int, "Input read error: Missing value for PNSCATTERING key.")
        elif key == "":
            continue
        elif key == "end":
            return L
        else:
            raise DeckError("Input read error: Unknown key in ANGULAR block : " + key)


def _read_grid(p, words, true_words):         # ParticleInputRead GRID, :482-535
    w = _word(words, 1)
    if w in ("linear", "logarithmic", "exponential"):
        p.gridstruct = w
        return
    form1 = w.count(":") == 2 and "," not in w
    form2 = ":" not in w and "," not in w
    if form1:
        a, b, c = w.split(":")
        p.gridstruct = "linear"
        p.Emin = _num(a, float, "Input read error: Unknown value for PARTICLE GRID key  : " + w)
        p.Emax = _num(b, float, "Input read error: Unknown value for PARTICLE GRID key  : " + w)
        p.G = _num(c, int, "Input read error: Unknown value for PARTICLE GRID key  : " + w)
    elif form2 and w:
        p.gridstruct = "null"
        p.fname = true_words[1]
    else:
        raise DeckError("Input read error: Unknown value for PARTICLE GRID key  : " + w)


def _read_particle(src, name):                # ParticleInputRead, :437-602
    p = ParticleInput(name=name)
    while True:
        line, words, _, true_words = src.next("PARTICLE")
        key = _word(words, 0)
        if key == "physics":
            if _word(words, 1) not in ("electrons", "photons"):
                raise DeckError("Input read error: Unknown particle in PHYSICS key : " + _word(words, 1))
            p.identity = words[1]
        elif key == "grid":
            _read_grid(p, words, true_words)
        elif key == "min":
            p.Emin = _num(_word(words, 1), float, "Input read error: Missing value for PARTICLE Min key.")
        elif key == "max":
            p.Emax = _num(_word(words, 1), float, "Input read error: Missing value for PARTICLE Max key.")
        elif key == "groups":
            p.G = _num(_word(words, 1), int, "Input read error: Missing value for PARTICLE Groups key.")
        elif key == "solver":
            if _word(words, 1) not in ("gmres", "si"):
                raise DeckError("Input read error: Unknown value for PARTICLE solver key  : " + _word(words, 1))
            p.solver = words[1].upper()
        elif key == "angular":
            if _word(words, 1) not in ("sn", "pn"):
                raise DeckError("Input read error: Unknown value for PARTICLE angular key : " + _word(words, 1))
            p.angular = words[1].upper()
        elif key == "noscatter":
            p.noscat = True
        elif key == "":
            continue
        elif key == "end":
            return p
        else:
            raise DeckError("Input read error: Unknown key in PARTICLE block : " + key)


def _read_particles(src):                     # ParticlesInputRead, :376-435
    parts = []
    while True:
        line, words, true_line, _ = src.next("PARTICLES")
        if line == "":
            continue
        if _word(words, 0) == "end":
            return parts
        p = _read_particle(src, true_line)
        if any(q.name == p.name for q in parts):
            raise DeckError("Input read error: Duplicate particle name " + p.name + ". "
                            "If you wish to transport multiple particles of the same kind, name them "
                            "differently, but give them the same PHYSICS key.")
        if p.name.lower() in ("electrons", "photons"):
            p.identity = p.name.lower()
        if p.identity is None:
            raise DeckError("Input read error: Particle " + p.name + " was given without a PHYSICS key.")
        parts.append(p)


def _read_beam(src):                          # BeamsInputRead, :604-790
    b = BeamInput()
    axisset = False
    bad_mono = "Input read error: Improper input after BEAM MONOCHROMATIC key. Please consult documentation."
    while True:
        line, words, _, true_words = src.next("BEAM")
        key = _word(words, 0)
        if key == "particle":
            b.particle = _num(_word(words, 1), int, "Input read error: Missing value for BEAM Particle key.")
        elif key == "polychromatic":
            if len(words) < 2:
                raise DeckError("Input read error: Missing file name for BEAM Polychromatic key.")
            b.spectrum = true_words[1]
        elif key == "monochromatic":
            b.spectrum = "constant"
            args = words[1:]
            nr = sum("." in a for a in args)
            ni = len(args) - nr
            if nr == 2:
                if ni != 1 or "." in args[0]:
                    raise DeckError(bad_mono)
                b.mcgroups = [_num(args[0], int, bad_mono)]
                b.E0 = _num(args[1], float, bad_mono)
                b.dE = _num(args[2], float, bad_mono)
            else:
                if nr != 0:
                    raise DeckError(bad_mono)
                b.mcgroups = [_num(a, int, bad_mono) for a in args]
        elif key == "weight":
            b.weight = _num(_word(words, 1), float, "Input read error: Missing value for BEAM Weight key.")
        elif key == "origin":
            b.origin = _reals(words, 3, "Input read error: Missing parameters for BEAM Origin key.")
        elif key == "axis":
            n = len(words) - 1
            if n == 1:
                b.axis = (_num(words[1], float, "Input read error: Missing parameters for BEAM Axis key."), 0.0, 0.0)
                b.slab = True
            elif n == 3:
                a = np.array(_reals(words, 3, "Input read error: Missing parameters for BEAM Axis key."))
                b.axis = tuple(float(x) for x in a / np.linalg.norm(a))
                b.slab = False
            else:
                raise DeckError("Input read error: Incorrect number of parameters for BEAM Axis key.")
            axisset = True
        elif key == "cutout":
            kind = _word(words, 1)
            if kind == "none":
                b.cutout = "none"
            elif kind == "circle":
                b.cutout = "circle"
                c = _reals(words[1:], 2, "Input read error: Missing Cutout Circle parameters. You must enter two.")
                b.cutoutparams = (c[0], c[1], 0.0, 0.0)
            elif kind == "rectangle":
                b.cutout = "rectangle"
                b.cutoutparams = _reals(words[1:], 4, "Input read error: Missing Cutout Rectangle parameters. "
                                                      "You must enter four.")
            elif kind == "":
                raise DeckError("Input read error: Missing all parameters for BEAM Cutout key.")
            else:
                raise DeckError("Input read error: Unknown cutout type for BEAM Cutout key: " + kind)
        elif key in ("spherical", "planar"):
            b.angdist = key
        elif key == "":
            continue
        elif key == "end":
            if not axisset:
                b.slab = True
            return b
        else:
            raise DeckError("Input read error: Unknown key in BEAM block : " + key)


def _read_postproc(src, pp):                  # PostProcInputRead, :792-822
    flags = {"fluence": "fluence", "uncollided": "outputunc", "energy": "energy", "dose": "dose", "charge": "charge"}
    while True:
        line, words, _, _ = src.next("POSTPROCESSING")
        key = _word(words, 0)
        if key in flags:
            setattr(pp, flags[key], True)
        elif key == "":
            continue
        elif key == "end":
            return
        else:
            raise DeckError("Input read error: Unknown key in POSTPROCESSING block : " + key)


def parse_deck(text, path="lb.in", check_files=True):
    """Deck from its text (ReadInput.f90:3-121); ``path`` places relative file
    names.  ``check_files=False`` skips the checks that open the mesh file."""
    src = _Lines(text)
    general, slab = _read_problem(src)
    d = Deck(path=path, general=general, slab=slab)
    done = set()
    while True:
        got = src.next_or_none()
        if got is None:
            break
        line = got[0]
        if line in ("mesh", "materials", "angular", "particles", "postprocessing") and line in done:
            raise DeckError(f"Input read error: Duplicate {line.upper()} block.")
        if line == "mesh":
            d.mesh = _read_mesh(src)
        elif line == "materials":
            d.materials = _read_materials(src)
        elif line == "angular":
            d.L = _read_angular(src)
        elif line == "particles":
            d.particles = _read_particles(src)
        elif line == "beam":
            d.beams.append(_read_beam(src))
        elif line == "postprocessing":
            _read_postproc(src, d.postproc)
        elif line == "memorylimited":
            d.options.memlim = True
        elif line == "storagelimited":
            d.options.storlim = True
        elif line == "debug":
            d.options.debug = True
        elif line == "":
            continue
        else:
            raise DeckError("Input read error: Invalid header : " + line)
        done.add(line)
    for block, msg in (("mesh", "MESH block was not found."),
                       ("materials", "MATERIALS block was not found."),
                       ("angular", "ANGULAR block was not found."),
                       ("particles", "PARTICLES block was not found."),
                       ("beam", "Must provide at least on BEAM block.")):
        if block not in done:
            raise DeckError("Input read error: " + msg)
    for p in d.particles:                     # ReadInput.f90:113-116
        p.L = d.L
        p.slab = d.slab
    validate(d, check_files=check_files)
    return d


def read_deck(path, check_files=True):
    with open(path) as f:
        return parse_deck(f.read(), path=path, check_files=check_files)


def count_physical_names(path):
    """The count on the line after ``$PhysicalNames`` (ValidateInput.f90:70-90)."""
    with open(path) as f:
        for line in f:
            if line.strip() == "$PhysicalNames":
                return int(f.readline().split()[0])
    raise DeckError("Input validation error: Could not find the number of materials in the mesh. "
                    "Are you sure you defined physical volumes?")


def validate(d, check_files=True):            # ValidateInput.f90:3-128
    if d.general:
        if d.mesh.slabs:
            raise DeckError("Input validation error: Using PROBLEM GENERAL, but slabs were specified.")
        if d.mesh.fname is None:
            raise DeckError("Input validation error: Using PROBLEM GENERAL, but mesh file was not specified.")
        if check_files and not os.path.exists(d.mesh_path()):
            raise DeckError("Input validation error: Provided mesh file does not exist: " + d.mesh.fname)
    else:
        if d.mesh.fname is not None:
            raise DeckError("Input validation error: Using PROBLEM SLAB, but mesh file was specified.")
        if not d.mesh.slabs:
            raise DeckError("Input validation error: Using PROBLEM SLAB, but no slabs were specified.")
    for m in d.materials:
        if m.dens <= 0.0:
            raise DeckError("Input validation error: Density negative, zero, or not provided for material "
                            + m.name + ".")
        if not m.Z or not m.N:
            raise DeckError("Input validation error: Atomic composition not provided for material " + m.name + ".")
    nm = len(d.materials)
    mismatch = "Input validation error: Number of materials specified does not match with number of materials in mesh."
    if d.general:
        if nm > 1 and check_files and count_physical_names(d.mesh_path()) != nm:
            raise DeckError(mismatch)
    elif any(s.material > nm for s in d.mesh.slabs):
        raise DeckError(mismatch)


# --- what the transport path builds from a deck -------------------------------------------------------------


def build_mesh(d):
    """MeshInputBuild (InputBuildModule.f90:15-55): the GMSH mesh, translated
    and then scaled (that order, whichever key came first)."""
    from .mesh import TetMesh, read_gmsh41
    if not d.general:
        raise DeckError("PROBLEM SLAB builds a one-dimensional slab mesh, which the tetrahedral sweep does not take.")
    m = read_gmsh41(d.mesh_path())
    rg = np.array(m.rg, dtype=np.float64)
    if d.mesh.dotrans:
        rg = rg + np.asarray(d.mesh.trans)       # Geometry.f90:195-219
    if d.mesh.doscale:
        rg = rg * np.asarray(d.mesh.scale)       # Geometry.f90:221-245
    return TetMesh(rg, m.node, m.mat, m.nbr)


def build_angular(d, particle):
    """ParticleInputBuildAngular (InputBuildModule.f90:77-107): SNPrep(L) for
    ANGULAR SN.  PN is a work in progress in the reference."""
    from .angular import snprep
    p = d.particles[particle] if isinstance(particle, int) else particle
    if p.angular != "SN":
        raise DeckError(f"Particle {p.name}: only ANGULAR SN builds an angular set.")
    if p.solver not in ("SI", "GMRES"):
        raise DeckError("InputBuildModule.f90: ParticleInputBuildAngular: Unknown solver given to BuildAngular.")
    return snprep(p.L)


def solver_kind(p):
    """Keyword for ``Operator.boltzmann_matinv``: SOLVER GMRES -> "gmres", SI -> "si"."""
    if p.solver not in ("SI", "GMRES"):
        raise DeckError(f"Particle {p.name}: no SOLVER given.")
    return p.solver.lower()


def energy_edges(p):
    """Group edges, highest energy first (group 1 is the highest, as the BEAM
    MONOCHROMATIC groups count them), for LINEAR and LOGARITHMIC grids.  The
    reference builds them in NittanyPhysics' EnergyGrid, which is not vendored:
    equal widths or equal ratios between MIN and MAX are the textbook grids the
    names denote."""
    if p.G < 1 or p.Emin <= 0.0 or p.Emax <= p.Emin:
        raise DeckError(f"Particle {p.name}: GRID needs GROUPS >= 1 and 0 < MIN < MAX.")
    t = np.arange(p.G + 1) / p.G
    if p.gridstruct == "linear":
        return p.Emax - (p.Emax - p.Emin) * t
    if p.gridstruct == "logarithmic":
        return p.Emax * (p.Emin / p.Emax) ** t
    raise DeckError(f"Particle {p.name}: GRID {p.gridstruct} is defined by NittanyPhysics, which is not available.")
