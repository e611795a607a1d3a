"""An input deck through the HIP path: the deck's mesh (GMSH + SCALE), its
PNSCATTERING order (11, as the reference's water-tank decks) and each
particle's SOLVER drive the solve, and the results go to the reference's HDF5
layout.  Cross sections and the source are synthetic (the reference generates
them in NittanyPhysics).  Tolerances: the GMRES particle's scalar flux within
1e-7 of the oracle (both converged to rtol 1e-10); the SI particle's within
1e-7 of the HIP GMRES solution of the same problem."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from lionbolt_amd import deck
from oracle.oracle import OracleOp

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CUBE = os.path.join(HERE, "golden", "simple_cube.msh")

DECK = f"""PROBLEM GENERAL
MESH
    GMSH {CUBE}
    SCALE 2.0 2.0 2.0
END
MATERIALS
    WATER
        DENSITY 1.0
        H 2
        O 1
    END
END
ANGULAR
    PNSCATTERING 11
END
PARTICLES
    PHOTONS
        GRID 0.001:1.0:25
        ANGULAR SN
        SOLVER SI
    END
    ELECTRONS
        GRID LOGARITHMIC
        GROUPS 35
        MIN 0.001
        MAX 1.0
        ANGULAR SN
        SOLVER GMRES
    END
END
BEAM
    PARTICLE 1
    MONOCHROMATIC 1
    AXIS 0.0 0.0 -1.0
    ORIGIN 0.0 0.0 130.0
    CUTOUT RECTANGLE 100.0 0.0 10.0 10.0
    SPHERICAL
END
"""


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_deck_drives_the_solve_and_the_results_file(tmp_path):
    from lionbolt_amd import h5out
    from lionbolt_amd.ops import Operator
    path = tmp_path / "lb.in"
    path.write_text(DECK)
    d = deck.read_deck(str(path))
    m = deck.build_mesh(d)
    assert m.NE == 12741 and abs(m.rg.max() - 1.0) < 1e-12
    L = d.L
    sigs = 0.5 * 0.7 ** np.arange(L + 1)[None, :]
    c = m.centroids()
    q_cell = np.where(np.all(np.abs(c) < 0.4, axis=1), 1.0 / (4 * np.pi), 0.0)
    psi_all, phi_all = [], []
    for p in d.particles:
        ang = deck.build_angular(d, p)
        assert ang.L == 11 and ang.NI == 288
        op = Operator(m, ang)
        op.set_physics([1.0], sigs)
        q = np.tile(np.repeat(q_cell, 4), ang.NI)
        kind = deck.solver_kind(p)
        print(f"particle {p.name}: solver {kind}, NE {m.NE}, NI {ang.NI}", flush=True)
        psi, st = op.boltzmann_matinv(q, kind=kind, space="moment" if kind == "gmres" else "psi", m=30, rtol=1e-10)
        assert st["status"] == 0 and st["true_residual"] <= 1e-9
        phi = op.fluence(psi)
        if kind == "gmres":
            orc = OracleOp(m, ang)
            orc.set_physics([1.0], sigs)
            b, ost = orc.boltzmann_matinv(q, kind="gmres", m=30, rtol=1e-10)
            assert ost["status"] == 0
            assert rel(phi, orc.fluence(b)) <= 1e-7
        else:
            g, gst = op.boltzmann_matinv(q, kind="gmres", space="moment", m=30, rtol=1e-10)
            assert gst["status"] == 0
            assert rel(phi, op.fluence(g)) <= 1e-7
        psi_all.append(psi)
        phi_all.append(phi)
    out = str(tmp_path / "results.h5")
    ang = deck.build_angular(d, d.particles[1])
    E = deck.energy_edges(d.particles[1])
    h5out.write_results(out, m, ang, particle=d.particles[1].name, psi=[psi_all[1]], phi=[phi_all[1]], energy_grid=E)
    h5dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    if os.path.exists(h5dump):
        txt = subprocess.run([h5dump, "-w", "0", "-y", "-m", "%.17g", "-d", "/electrons/fluence/energy_1", out],
                             capture_output=True, text=True, check=True).stdout
        body = txt[txt.index("DATA {") + 6:]
        vals = np.array([float(t) for t in body[:body.index("}")].replace(",", " ").split()])
        np.testing.assert_array_equal(vals, phi_all[1])
