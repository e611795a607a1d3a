"""Planar external beam: the host mirror of the reference's ``FieldGeometry`` and
``ExternalBeam`` (src/sources/Sources.f90:13-80) over the ``lb_beam_*`` entry
points.  Only the planar angular distribution is on the transport path's
first-collision source; spherical beams and slab meshes stay out.

* ``FieldGeometry(axis, origin, cutout, cutoutparams)``: the field; ``e1``, ``e2``
  span the source plane.  The reference builds them only for rectangular cutouts
  (``ZYZBasis`` of the axis angles and the rotation ``cutoutparams(2)`` degrees,
  init_FieldGeometryGen.f90:99-107, restated here as the z-y-z Euler rotation) and
  leaves them unset otherwise, although RayTracing.f90:176-180 back-projects every
  quadrature point with them; here any orthonormal pair normal to the axis is used
  then, which gives the back-projection along the axis that RayTracing.f90:149-153
  describes.
* ``ExternalBeam(fldgeo, n)``: ``uncollided(op, attn)`` is the projected uncollided
  fluence phi0 (FluenceQuadrature + the projection of FCSSource.f90:134-176);
  ``moments(op, phi0)`` its first-collision and uncollided flux moments;
  ``Source(op, s, phi0)`` adds the first-collision source to a device space-angle
  vector (FCSSource.f90:178-215); ``AddAngularFluence(op, phi0, unc)`` adds the
  uncollided angular fluence (AddAngularFluence.f90:153-166).
"""
import ctypes
import math

import numpy as np

from . import _lib
from .ops import _dev, _torch_done


def _zyz_basis(a, b, c):
    """Columns 1 and 2 of R = Rz(a) Ry(b) Rz(c) (its column 3 is the axis)."""
    ca, sa, cb, sb, cc, sc = math.cos(a), math.sin(a), math.cos(b), math.sin(b), math.cos(c), math.sin(c)
    e1 = np.array([ca * cb * cc - sa * sc, sa * cb * cc + ca * sc, -sb * cc])
    e2 = np.array([-ca * cb * sc - sa * cc, -sa * cb * sc + ca * cc, sb * sc])
    return e1, e2


class FieldGeometry:
    CUTOUTS = {"none": _lib.CUTOUT_NONE, "circle": _lib.CUTOUT_CIRCLE, "rectangle": _lib.CUTOUT_RECTANGLE}

    def __init__(self, axis, origin, cutout="none", cutoutparams=(0.0, 0.0, 0.0, 0.0)):
        a = np.asarray(axis, dtype=np.float64)
        self.axis = a / np.linalg.norm(a)
        self.origin = np.asarray(origin, dtype=np.float64)
        self.cutout = cutout
        self.cutoutparams = np.zeros(4)
        self.cutoutparams[:len(cutoutparams)] = cutoutparams
        if cutout == "rectangle":
            self.e1, self.e2 = _zyz_basis(math.atan2(self.axis[1], self.axis[0]), math.acos(self.axis[2]),
                                          self.cutoutparams[1] * math.pi / 180.0)
        else:
            t = np.array([1.0, 0.0, 0.0]) if abs(self.axis[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
            e1 = t - self.axis.dot(t) * self.axis
            self.e1 = e1 / np.linalg.norm(e1)
            self.e2 = np.cross(self.axis, self.e1)

    def desc(self, n):
        d = _lib.BeamDesc()
        for c in range(3):
            d.axis[c], d.origin[c], d.e1[c], d.e2[c] = self.axis[c], self.origin[c], self.e1[c], self.e2[c]
        d.cutout = self.CUTOUTS[self.cutout]
        for c in range(4):
            d.params[c] = self.cutoutparams[c]
        d.n = float(n)
        return d


class ExternalBeam:
    def __init__(self, fldgeo, n=1.0):
        self.fldgeo = fldgeo
        self.n = n

    def uncollided(self, op, attn, elems=None, out=None):
        """phi0 [NENK] (cuda float64) for cross sections ``attn`` (sigma_t per material);
        ``elems=(e0, e1)`` traces only those elements (written into ``out``)."""
        import torch
        attn = np.ascontiguousarray(np.atleast_1d(attn), dtype=np.float64)
        phi0 = torch.zeros(op.NENK, dtype=torch.float64, device="cuda") if out is None else out
        p = _dev(phi0, op.NENK, "out")
        d = self.fldgeo.desc(self.n)
        if elems is None:
            _lib.check(_lib.lib().lb_beam_uncollided_dev(op.h, ctypes.byref(d), _lib.dptr(attn), p))
        else:
            _lib.check(_lib.lib().lb_beam_uncollided_range_dev(op.h, ctypes.byref(d), _lib.dptr(attn), int(elems[0]),
                                                               int(elems[1]), p))
        return phi0

    def moments(self, op, phi0):
        """(fcs, unc) [nmom][NENK]: Y_lm(axis) phi0 and the uncollided fluence's moments D g phi0."""
        import torch
        p = _dev(phi0, op.NENK, "phi0")
        fcs = torch.empty((op.nmom, op.NENK), dtype=torch.float64, device="cuda")
        unc = torch.empty_like(fcs)
        _torch_done()
        d = self.fldgeo.desc(self.n)
        _lib.check(_lib.lib().lb_beam_moments_dev(op.h, ctypes.byref(d), p, fcs.data_ptr(), unc.data_ptr()))
        return fcs, unc

    def Source(self, op, s, phi0, g=0):
        """FCSSource: s [NI][NENK] (cuda) += M diag(sigma_l(mat, g)) Y(axis) phi0."""
        fcs, _ = self.moments(op, phi0)
        op.moments_to_discrete_dev(fcs, s, g, 1, accumulate=True)

    def source_host(self, op, s):
        """FCSSource on a host vector (reference layout, NENK * NI): s + the
        first-collision source with the operator's single-group sigma_t and
        sigma_l (``lb_beam_source``); returns the new vector."""
        x = op._vec(s)
        d = self.fldgeo.desc(self.n)
        _lib.check(_lib.lib().lb_beam_source(op.h, ctypes.byref(d), _lib.dptr(x), x.size))
        return x

    def add_angular_fluence_host(self, op, attn, unc):
        """AddAngularFluence on a host vector: unc + the uncollided angular fluence
        for attenuation ``attn`` (per material) (``lb_beam_add_angular_fluence``)."""
        x = op._vec(unc)
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(attn, dtype=np.float64), (op.Nm,)))
        d = self.fldgeo.desc(self.n)
        _lib.check(_lib.lib().lb_beam_add_angular_fluence(op.h, ctypes.byref(d), _lib.dptr(a), _lib.dptr(x), x.size))
        return x

    def AddAngularFluence(self, op, phi0, unc):
        """unc [NI][NENK] (cuda) += phi0 sum_l (2l+1)/(4 pi) P_l(k_i . axis)."""
        p, u = _dev(phi0, op.NENK, "phi0"), _dev(unc, op.n, "unc")
        _torch_done()
        d = self.fldgeo.desc(self.n)
        _lib.check(_lib.lib().lb_beam_angular_fluence_dev(op.h, ctypes.byref(d), p, u))
