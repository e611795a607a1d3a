"""Input decks: the reference's twelve lb.in files (tests/golden/decks, taken
unchanged from examples/ and validation/) parse to the values written in them,
and malformed decks stop with the reference's messages
(ReadInput.f90, InputReadModule.f90, ValidateInput.f90)."""
import glob
import os

import numpy as np
import pytest

from lionbolt_amd import deck
from lionbolt_amd import mesh as M

HERE = os.path.dirname(os.path.abspath(__file__))
DECKS = os.path.join(HERE, "golden", "decks")
CUBE = os.path.join(HERE, "golden", "simple_cube.msh")

# Values read off each deck by hand: problem, L, mesh file, translate, scale,
# slabs (thickness, elements, structure, material), materials (name, density,
# Z, N), particles (name, grid, Emin, Emax, G, solver), beams (particle,
# spectrum, groups, axis, origin, cutout, angular distribution), storagelimited.
WATER = ("WATER", 1.0, [1, 8], [2, 1])
BEAM_Z = (1, "constant", [1], (0.0, 0.0, -1.0), (0.0, 0.0, 130.0), ("rectangle", (100.0, 0.0, 10.0, 10.0)), "spherical")
EXPECTED = {
    "three_beam_water_tank": ("general", 7, "three_beam.msh", None, (30.0, 30.0, 30.0), [], [WATER],
                              [("PHOTONS", "linear", 0.001, 1.0, 25, "SI"), ("ELECTRONS", "linear", 0.001, 1.0, 35, "GMRES")],
                              [BEAM_Z,
                               (1, "constant", [1], (0.0, 0.86602540, 0.5), (0.0, -99.59292144, -42.5),
                                ("rectangle", (100.0, 0.0, 10.0, 10.0)), "spherical"),
                               (1, "constant", [1], (0.0, -0.86602540, 0.5), (0.0, 99.59292144, -42.5),
                                ("rectangle", (100.0, 0.0, 10.0, 10.0)), "spherical")], True),
    "water_tank_with_aluminum": ("general", 15, "het_cube_spherical_beam.msh", (-0.5, -0.5, 0.0), (30.0, 30.0, 30.0),
                                 [], [WATER, ("ALUMINUM", 2.7, [13], [1])],
                                 [("PHOTONS", "logarithmic", 0.001, 6.0, 50, "GMRES"),
                                  ("ELECTRONS", "logarithmic", 0.001, 6.0, 50, "GMRES")], [BEAM_Z], True),
    "water_tank_1MeV": ("general", 11, "water_tank.msh", None, (30.0, 30.0, 30.0), [], [WATER],
                        [("PHOTONS", "linear", 0.001, 1.0, 25, "GMRES"),
                         ("ELECTRONS", "logarithmic", 0.001, 1.0, 35, "GMRES")], [BEAM_Z], True),
    "water_tank_6MeV": ("general", 11, "water_tank.msh", None, (30.0, 30.0, 30.0), [], [WATER],
                        [("PHOTONS", "linear", 0.001, 6.0, 25, "SI"),
                         ("ELECTRONS", "logarithmic", 0.001, 6.0, 35, "GMRES")], [BEAM_Z], True),
    "water_tank_6MV": ("general", 11, "water_tank.msh", None, (30.0, 30.0, 30.0), [], [WATER],
                       [("PHOTONS", "logarithmic", 0.001, 5.4, 50, "SI"),
                        ("ELECTRONS", "logarithmic", 0.001, 5.4, 50, "GMRES")],
                       [(1, "6MV_photons.txt", None, (0.0, 0.0, -1.0), (0.0, 0.0, 130.0),
                         ("rectangle", (100.0, 0.0, 10.0, 10.0)), "spherical")], True),
    "AlAuAl_1MeV_0deg": ("slab", 15, None, None, None,
                         [(0.01, 50, "linear", 1), (0.002, 50, "linear", 2), (0.07, 150, "linear", 1)],
                         [("ALUMINUM", 2.7, [13], [1]), ("GOLD", 19.3, [79], [1])],
                         [("ELECTRONS", "linear", 0.001, 1.0, 35, "GMRES")],
                         [(1, "constant", [1], None, None, (None, None), None)], False),
    "Al_521keV_0deg": ("slab", 15, None, None, None, [(0.0867, 200, "linear", 1)], [("ALUMINUM", 2.7, [13], [1])],
                       [("ELECTRONS", "linear", 0.001, 0.521, 75, "GMRES")],
                       [(1, "constant", [1], None, None, (None, None), None)], False),
    "Al_6MV_0deg": ("slab", 15, None, None, None, [(0.0867, 200, "linear", 1)], [("ALUMINUM", 2.7, [13], [1])],
                    [("ELECTRONS", "linear", 0.001, 5.4, 35, "GMRES")],
                    [(1, "just_some_electrons.txt", None, None, None, (None, None), None)], False),
    "Au_10MeV_0deg": ("slab", 15, None, None, None, [(0.2, 200, "linear", 1)], [("GOLD", 19.3, [79], [1])],
                      [("ELECTRONS", "logarithmic", 0.001, 10.0, 50, "GMRES")],
                      [(1, "constant", [1], None, None, (None, None), None)], False),
    "Au_1MeV_0deg": ("slab", 15, None, None, None, [(0.01, 200, "linear", 1)], [("GOLD", 19.3, [79], [1])],
                     [("ELECTRONS", "linear", 0.001, 1.0, 50, "GMRES")],
                     [(1, "constant", [1], None, None, (None, None), None)], False),
    "U_1MeV_0deg": ("slab", 15, None, None, None, [(0.02, 200, "linear", 1)], [("URANIUM", 18.9, [92], [1])],
                    [("ELECTRONS", "logarithmic", 0.001, 1.0, 35, "GMRES")],
                    [(1, "constant", [1], None, None, (None, None), None)], False),
    "Pb_10MeV_0deg": ("slab", 15, None, None, None, [(5.0, 500, "logarithmic", 1)], [("LEAD", 11.34, [82], [1])],
                      [("PHOTONS", "linear", 0.001, 10.0, 50, "SI"),
                       ("ELECTRONS", "logarithmic", 0.001, 10.0, 100, "GMRES")],
                      [(1, "constant", [1], None, None, (None, None), None)], False),
}


def test_every_reference_deck_is_covered():
    names = sorted(os.path.basename(f)[:-3] for f in glob.glob(os.path.join(DECKS, "*.in")))
    assert names == sorted(EXPECTED)


@pytest.mark.parametrize("name", sorted(EXPECTED))
def test_reference_deck_values(name):
    prob, L, fname, trans, scale, slabs, mats, parts, beams, storlim = EXPECTED[name]
    d = deck.read_deck(os.path.join(DECKS, name + ".in"), check_files=False)
    assert (d.general, d.slab) == (prob == "general", prob == "slab")
    assert d.L == L
    assert d.mesh.fname == fname
    assert d.mesh.dotrans == (trans is not None) and d.mesh.doscale == (scale is not None)
    if trans is not None:
        assert d.mesh.trans == trans
    if scale is not None:
        assert d.mesh.scale == scale
    assert [(s.thickness, s.elements, s.structure, s.material) for s in d.mesh.slabs] == slabs
    assert [(m.name, m.dens, m.Z, m.N) for m in d.materials] == mats
    assert [(p.name, p.gridstruct, p.Emin, p.Emax, p.G, p.solver) for p in d.particles] == parts
    for p in d.particles:
        assert p.angular == "SN" and p.L == L and p.slab == d.slab and p.identity == p.name.lower()
    assert len(d.beams) == len(beams)
    for b, (part, spec, groups, axis, origin, (cut, cutp), angdist) in zip(d.beams, beams):
        assert (b.particle, b.spectrum, b.mcgroups, b.cutout, b.angdist) == (part, spec, groups, cut, angdist)
        assert b.slab == (axis is None)
        if axis is not None:
            a = np.array(axis)
            np.testing.assert_allclose(b.axis, a / np.linalg.norm(a), rtol=0, atol=1e-15)
            assert abs(np.linalg.norm(b.axis) - 1.0) < 1e-15
            assert b.origin == origin and b.cutoutparams == cutp
    pp = d.postproc
    assert (pp.fluence, pp.energy, pp.dose, pp.charge, pp.outputunc) == (True, True, True, True, False)
    assert d.options.storlim == storlim


GOOD = """PROBLEM GENERAL
MESH
    GMSH {mesh}
    {extra_mesh}
END
MATERIALS
    WATER
        DENSITY 1.0
        H 2
        O 1
    END
END
ANGULAR
    PNSCATTERING 3
END
PARTICLES
    PHOTONS
        GRID 0.01:1.0:4
        ANGULAR SN
        SOLVER GMRES
    END
END
BEAM
    PARTICLE 1
    MONOCHROMATIC 1
    AXIS 0.0 0.0 -2.0
    ORIGIN 0.0 0.0 10.0
    CUTOUT CIRCLE 5.0 1.0
    PLANAR
END
"""


def good(**kw):
    kw.setdefault("mesh", CUBE)
    kw.setdefault("extra_mesh", "")
    return GOOD.format(**kw)


def test_good_deck_and_its_build(tmp_path):
    d = deck.parse_deck(good(extra_mesh="SCALE 2.0 3.0 4.0\n    TRANSLATE 1.0 -1.0 0.5"), path=str(tmp_path / "lb.in"))
    assert d.beams[0].axis == (0.0, 0.0, -1.0) and d.beams[0].cutout == "circle"
    assert d.beams[0].cutoutparams == (5.0, 1.0, 0.0, 0.0) and d.beams[0].angdist == "planar"
    m = deck.build_mesh(d)
    ref = M.read_gmsh41(CUBE)
    # translate first, then scale, whatever order the keys came in (InputBuildModule.f90:40-41)
    np.testing.assert_array_equal(m.rg, (ref.rg + [1.0, -1.0, 0.5]) * [2.0, 3.0, 4.0])
    assert np.array_equal(m.node, ref.node) and np.array_equal(m.nbr, ref.nbr) and np.array_equal(m.mat, ref.mat)
    ang = deck.build_angular(d, 0)
    assert ang.L == 3 and ang.NI == 2 * 16 and abs(ang.w.sum() - 4 * np.pi) < 1e-12
    assert deck.solver_kind(d.particles[0]) == "gmres"
    E = deck.energy_edges(d.particles[0])
    np.testing.assert_allclose(E, [1.0, 0.7525, 0.505, 0.2575, 0.01], rtol=0, atol=1e-15)


def test_energy_edges_closed_forms():
    p = deck.ParticleInput(name="x", gridstruct="logarithmic", Emin=0.001, Emax=10.0, G=4)
    np.testing.assert_allclose(deck.energy_edges(p), [10.0, 1.0, 0.1, 0.01, 0.001], rtol=1e-14)
    p.gridstruct = "exponential"
    with pytest.raises(deck.DeckError, match="NittanyPhysics"):
        deck.energy_edges(p)


def test_slab_problem_has_no_tet_mesh():
    d = deck.read_deck(os.path.join(DECKS, "Au_1MeV_0deg.in"))
    with pytest.raises(deck.DeckError, match="slab"):
        deck.build_mesh(d)


BAD = [
    ("", "Blank input file"),
    ("PROBLEM\n", "too few words"),
    ("SOLVE GENERAL\n", 'Use key "PROBLEM"'),
    ("PROBLEM CYLINDER\n", "Invalid problem type : problem cylinder"),
    (lambda: good() + "FOO\n", "Invalid header : foo"),
    (lambda: good() + "MESH\nGMSH x.msh\nEND\n", "Duplicate MESH block"),
    (lambda: good().replace("    GMSH", "    ROTATE 1\n    GMSH"), "Unknown key in MESH block : rotate"),
    (lambda: good().replace("H 2", "Xx 2"), "Unknown key in a MATERIAL block : Xx 2"),
    (lambda: good().replace("DENSITY 1.0", ""), "Density negative, zero, or not provided for material WATER"),
    (lambda: good().replace("PNSCATTERING 3", "ORDER 3"), "Unknown key in ANGULAR block : order"),
    (lambda: good().replace("GRID 0.01:1.0:4", "GRID 0.01:1.0"), "Unknown value for PARTICLE GRID key"),
    (lambda: good().replace("SOLVER GMRES", "SOLVER CG"), "Unknown value for PARTICLE solver key"),
    (lambda: good().replace("ANGULAR SN", "ANGULAR DG"), "Unknown value for PARTICLE angular key"),
    (lambda: good().replace("    PHOTONS\n", "    NEUTRONS\n"), "Particle NEUTRONS was given without a PHYSICS key"),
    (lambda: good().replace("END\nEND\nBEAM", "END\n    PHOTONS\n        SOLVER SI\n    END\nEND\nBEAM"),
     "Duplicate particle name PHOTONS"),
    (lambda: good().replace("MONOCHROMATIC 1", "MONOCHROMATIC 1 0.5"), "Improper input after BEAM MONOCHROMATIC"),
    (lambda: good().replace("AXIS 0.0 0.0 -2.0", "AXIS 0.0 1.0"), "Incorrect number of parameters for BEAM Axis"),
    (lambda: good().replace("CUTOUT CIRCLE 5.0 1.0", "CUTOUT HEXAGON"), "Unknown cutout type for BEAM Cutout key: hexagon"),
    (lambda: good().replace("    PLANAR\n", "    CONE\n"), "Unknown key in BEAM block : cone"),
    (lambda: good().split("BEAM")[0], "Must provide at least on BEAM block"),
    (lambda: good().replace("ANGULAR\n    PNSCATTERING 3\nEND\n", ""), "ANGULAR block was not found"),
    (lambda: good().replace("    GMSH", "    SLAB\n        THICKNESS 1.0\n    END\n    GMSH"),
     "Using PROBLEM GENERAL, but slabs were specified"),
    (lambda: good(mesh="no_such_file.msh"), "Provided mesh file does not exist: no_such_file.msh"),
    (lambda: good().replace("    END\nEND\nANGULAR", "    END\n    AIR\n        DENSITY 0.001\n        N 2\n    END\nEND\nANGULAR"),
     "Number of materials specified does not match"),
    (lambda: good() + "POSTPROCESSING\n    FLUENCE\n    TALLY\nEND\n", "Unknown key in POSTPROCESSING block : tally"),
    (lambda: good().rsplit("END", 1)[0], "End of file inside the BEAM block"),
]


@pytest.mark.parametrize("text,msg", BAD, ids=[m for _, m in BAD])
def test_malformed_decks_stop_with_the_reference_message(text, msg, tmp_path):
    text = text() if callable(text) else text
    with pytest.raises(deck.DeckError, match=msg.replace("(", r"\(").replace(")", r"\)")):
        deck.parse_deck(text, path=str(tmp_path / "lb.in"))


def test_comments_case_and_blank_lines(tmp_path):
    text = good().replace("PROBLEM GENERAL", "# leading comment\n\nproblem general   # trailing").replace(
        "SOLVER GMRES", "solver si  # lower case")
    d = deck.parse_deck(text, path=str(tmp_path / "lb.in"))
    assert d.particles[0].solver == "SI" and d.materials[0].name == "WATER"
    assert deck.solver_kind(d.particles[0]) == "si"


def test_slab_decks_keep_defaults():
    # SLAB without MATERIAL takes material 1; unknown SLAB keys are skipped as the reference skips them
    text = ("PROBLEM SLAB\nMESH\n SLAB\n  THICKNESS 1.0\n  ELEMENTS 10\n  COLOUR red\n END\nEND\n"
            "MATERIALS\n LEAD\n  DENSITY 11.34\n  Pb 1\n END\nEND\nANGULAR\n PNSCATTERING 1\nEND\n"
            "PARTICLES\n PHOTONS\n  GRID LINEAR\n  GROUPS 2\n  MIN 0.1\n  MAX 1.0\n  SOLVER SI\n  ANGULAR SN\n END\nEND\n"
            "BEAM\n PARTICLE 1\n MONOCHROMATIC 2 1.5 0.25\nEND\n")
    d = deck.parse_deck(text)
    assert d.mesh.slabs[0].material == 1 and d.mesh.slabs[0].elements == 10
    b = d.beams[0]
    assert b.slab and b.mcgroups == [2] and (b.E0, b.dE) == (1.5, 0.25)
    np.testing.assert_allclose(deck.energy_edges(d.particles[0]), [1.0, 0.55, 0.1], atol=1e-15)


def test_periodic_table():
    assert [deck.periodic_table(s) for s in ("H", "O", "Al", "Au", "Pb", "U", "Og")] == [1, 8, 13, 79, 82, 92, 118]
    assert deck.periodic_table("AL") == 0 and deck.periodic_table("Xx") == 0
