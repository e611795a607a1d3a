"""MI355X-native transport-solve path of Lionbolt: DFEM SN sweep, scattering
moment product and the GMRES / source-iteration loop, behind the reference's
operator API (TransportOp, ScatteringOp, BoltzmannOp, SpaceAngleVector)."""
from . import angular, mesh, rng  # noqa: F401

__all__ = ["angular", "mesh", "rng", "ops", "build"]


def build():
    """Compile liblbhip.so for gfx950 in-tree."""
    import os
    import subprocess
    here = os.path.dirname(os.path.abspath(__file__))
    subprocess.run(["make", "-s", "-C", os.path.join(here, "csrc"), "-j4"], check=True)


def __getattr__(name):
    if name == "ops":
        import importlib
        return importlib.import_module(".ops", __name__)
    raise AttributeError(name)
