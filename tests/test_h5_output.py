"""Results written in the reference's HDF5 layout (LionboltHDF5.f90, names of
src/IO/schema.yaml), read back with the HDF5 library's own h5dump."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from lionbolt_amd import angular, mesh as M

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)
pytestmark = pytest.mark.skipif(H5DUMP is None, reason="h5dump (HDF5 tools) not available")


def h5(*args):
    """h5dump with the given options; the file is the last argument."""
    return subprocess.run([H5DUMP, *args], capture_output=True, text=True, check=True).stdout


def values(path, obj, attr=False):
    out = h5("-w", "0", "-y", "-m", "%.17g", "-a" if attr else "-d", obj, path)
    body = out[out.index("DATA {") + 6:]
    body = body[:body.index("}")]
    if '"' in body:
        return re.findall(r'"([^"]*)"', body)
    return np.array([float(t) for t in re.findall(r"[-+]?\d+\.?\d*(?:[eE][-+]?\d+)?", body)])


def test_results_file_layout_and_values(tmp_path):
    from lionbolt_amd import h5out
    m = M.jittered_cube(2, a=3.0, seed=5, slab=(1.0, 2.0))
    ang = angular.snprep(1)
    rng = np.random.default_rng(4)
    psi = [rng.random(ang.NI * m.NENK) for _ in range(2)]
    phi = [(ang.w[:, None] * p.reshape(ang.NI, -1)).sum(0) for p in psi]
    unc = [rng.random(ang.NI * m.NENK)]
    E = np.array([1.0, 0.5, 0.1])
    path = str(tmp_path / "results.h5")
    h5out.write_results(path, m, ang, particle="PHOTONS", psi=psi, phi=phi, psi_unc=unc, energy_grid=E)
    head = h5("-H", path)
    for name in ("mesh", "material_to_spatial_dof", "material_1", "material_2", "photons", "angular_fluence",
                 "uncollided_angular_fluence", "energy_1", "energy_2", "angle_1", "fluence", "global_nodes",
                 "connectivity", "offset", "element_volumes", "angular_weights", "abscissae", "energy_grid"):
        assert f'"{name}"' in head, name
    assert values(path, "/problem_type", attr=True) == ["general"]
    assert values(path, "/mesh/number_of_elements", attr=True)[0] == m.NE
    assert values(path, "/mesh/number_of_spatial_dofs", attr=True)[0] == m.NENK
    assert values(path, "/mesh/number_of_global_nodes", attr=True)[0] == m.NK
    assert values(path, "/mesh/number_of_materials", attr=True)[0] == 2
    assert values(path, "/photons/object_type", attr=True) == ["particle"]
    assert values(path, "/photons/number_of_angular_dofs", attr=True)[0] == ang.NI
    assert values(path, "/photons/number_of_energy_dofs", attr=True)[0] == 2
    # 1-based connectivity(offset(e) + k) = global node of (e, k)
    conn = values(path, "/mesh/connectivity")
    off = values(path, "/mesh/offset")
    assert np.array_equal(conn, m.node.reshape(-1) + 1)
    assert np.array_equal(off, 4 * np.arange(m.NE + 1))
    assert np.allclose(values(path, "/mesh/global_nodes").reshape(-1, 3), m.rg, rtol=1e-12, atol=0)
    sd2 = values(path, "/mesh/material_to_spatial_dof/material_2")
    assert np.array_equal(sd2, (np.nonzero(np.repeat(m.mat, 4) == 2)[0] + 1))
    assert np.allclose(values(path, "/mesh/element_volumes"), m.volumes(), rtol=1e-12)
    assert np.allclose(values(path, "/photons/abscissae").reshape(-1, 3), ang.k, rtol=1e-12)
    a3 = values(path, "/photons/angular_fluence/energy_2/angle_3")
    assert np.allclose(a3, psi[1].reshape(ang.NI, -1)[2], rtol=1e-12)
    assert np.allclose(values(path, "/photons/fluence/energy_1"), phi[0], rtol=1e-12)
    u1 = values(path, "/photons/uncollided_angular_fluence/energy_1/angle_1")
    assert np.allclose(u1, unc[0].reshape(ang.NI, -1)[0], rtol=1e-12)


def test_bad_sizes_are_rejected(tmp_path):
    from lionbolt_amd import h5out
    m = M.jittered_cube(1)
    ang = angular.snprep(1)
    with pytest.raises(ValueError):
        h5out.write_results(str(tmp_path / "x.h5"), m, ang, psi=[np.zeros(3)])


def test_writer_library_exports_every_declared_symbol():
    from lionbolt_amd import h5out
    hdr = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "lionbolt_h5.h")
    names = re.findall(r"^(?:int|const char\*)\s+(lbh5_\w+)\(", open(hdr).read(), flags=re.M)
    assert len(names) == 8
    L = h5out.lib()
    for n in names:
        assert hasattr(L, n) and n in h5out.SIGNATURES, n
