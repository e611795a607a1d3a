"""Discrete-ordinates sets and the angular scattering kernels.

``snprep(L)`` reproduces the reference's ``SNPrep`` product set
(``src/angular/SNPrep.f90:39-76``): ``Ng = L+1`` Gauss-Legendre polar cosines in
ascending order, ``Nc = 2 Ng`` Chebyshev azimuths ``phi_j = (2j-1) pi / Nc``,
direction ``ii = (j-1) Ng + i`` (polar index fastest), weight ``wG_i 2 pi / Nc``
(so the weights sum to 4 pi).  ``product(Np, Na)`` is the same construction with
independent polar and azimuthal counts, which the reference's public
``AngularClass`` components allow (``AngularSpace.f90:5-20``).
``level_symmetric_s8()`` is the LQ_n S8 set of Lewis & Miller, *Computational
Methods of Neutron Transport* (1984), Table 4-1, with the exact weights
49/405, 49/540 and 5/54 per octant.

``yy_kernel`` is ``YY(ip,i,l) = w_ip (2l+1)/(4 pi) P_l(k_i . k_ip)``
(``SNPrep.f90:106-133``); ``moment_matrices`` returns the real-spherical-harmonic
discrete-to-moment ``D`` and moment-to-discrete ``M`` matrices whose product
reproduces ``YY`` by the addition theorem.
"""
import math

import numpy as np


def gauss_legendre(n):
    """Gauss-Legendre nodes (ascending) and weights on [-1, 1] by Newton iteration."""
    x = np.empty(n)
    w = np.empty(n)
    for i in range(n):
        z = math.cos(math.pi * (i + 0.75) / (n + 0.5))
        for _ in range(100):
            p0, p1 = 1.0, z
            for k in range(2, n + 1):
                p0, p1 = p1, ((2 * k - 1) * z * p1 - (k - 1) * p0) / k
            if n == 1:
                p0, p1 = 1.0, z
            dp = n * (z * p1 - p0) / (z * z - 1.0)
            dz = p1 / dp
            z -= dz
            if abs(dz) < 1e-16:
                break
        p0, p1 = 1.0, z
        for k in range(2, n + 1):
            p0, p1 = p1, ((2 * k - 1) * z * p1 - (k - 1) * p0) / k
        dp = n * (z * p1 - p0) / (z * z - 1.0)
        x[i] = z
        w[i] = 2.0 / ((1.0 - z * z) * dp * dp)
    order = np.argsort(x)
    x, w = x[order], w[order]
    # enforce exact antisymmetry of the nodes, as the symmetric rule requires
    xs = 0.5 * (x - x[::-1])
    ws = 0.5 * (w + w[::-1])
    return xs, ws


class AngularSet:
    """Directions ``k`` (NI, 3), weights ``w`` (NI,), and scattering order ``L``."""

    def __init__(self, k, w, L, name=""):
        self.k = np.ascontiguousarray(k, dtype=np.float64)
        self.w = np.ascontiguousarray(w, dtype=np.float64)
        self.L = int(L)
        self.name = name

    @property
    def NI(self):
        return int(self.k.shape[0])

    @property
    def n_moments(self):
        return (self.L + 1) ** 2


def product(Np, Na, L):
    """Np Gauss-Legendre polar x Na Chebyshev azimuthal product set, polar index fastest."""
    mu, wg = gauss_legendre(Np)
    wc = 2.0 * math.pi / Na
    phi = (2.0 * np.arange(1, Na + 1) - 1.0) * math.pi / Na
    k = np.empty((Na * Np, 3))
    w = np.empty(Na * Np)
    ii = 0
    for j in range(Na):
        for i in range(Np):
            s = math.sqrt(1.0 - mu[i] ** 2)
            k[ii] = (math.cos(phi[j]) * s, math.sin(phi[j]) * s, mu[i])
            w[ii] = wg[i] * wc
            ii += 1
    return AngularSet(k, w, L, name=f"GLxC {Np}x{Na}")


def snprep(L):
    """The reference's own set: Ng = L+1, Nc = 2 Ng (``SNPrep.f90:39-76``)."""
    s = product(L + 1, 2 * (L + 1), L)
    s.name = f"SNPrep L={L}"
    return s


def level_symmetric_s8(L=0):
    """LQ_n S8 (80 directions); octants in the order of sign patterns (+,+,+), (-,+,+), ..."""
    mu = np.sqrt((1.0 + 6.0 * np.arange(4)) / 21.0)
    w_cls = {"114": 49.0 / 405.0, "123": 49.0 / 540.0, "222": 5.0 / 54.0}
    pts = []
    for a in range(4):
        for b in range(4):
            c = 6 - 3 - a - b   # 1-based indices sum to 6
            if 0 <= c < 4:
                cls = "".join(sorted(str(t + 1) for t in (a, b, c)))
                pts.append(((mu[a], mu[b], mu[c]), w_cls[cls]))
    k, w = [], []
    for sz in (1, -1):
        for sy in (1, -1):
            for sx in (1, -1):
                for (x, y, z), wt in pts:
                    k.append((sx * x, sy * y, sz * z))
                    w.append(wt * math.pi / 2.0)
    return AngularSet(np.array(k), np.array(w), L, name="LQ S8")


def legendre(L, x):
    """P_0..P_L at x by the three-term recurrence; shape (L+1, ...)."""
    x = np.asarray(x, dtype=np.float64)
    p = np.empty((L + 1,) + x.shape)
    p[0] = 1.0
    if L >= 1:
        p[1] = x
    for n in range(1, L):
        p[n + 1] = ((2 * n + 1) * x * p[n] - n * p[n - 1]) / (n + 1)
    return p


def yy_kernel(ang):
    """``YY(ip, i, l)`` as an array of shape (L+1, NI, NI) indexed [l, ip, i]."""
    kk = np.clip(ang.k @ ang.k.T, -1.0, 1.0)
    np.fill_diagonal(kk, 1.0)
    p = legendre(ang.L, kk)
    ell = np.arange(ang.L + 1)[:, None, None]
    return (2 * ell + 1) / (4 * math.pi) * p * ang.w[None, :, None]


def real_sph_harm(L, k):
    """Real orthonormal spherical harmonics Y_lm(k) for l <= L, shape ((L+1)^2, N).

    Row index ``l*l + l + m`` for ``m = -l..l``; ``Y_l^m`` uses cos(m phi) for
    m > 0 and sin(|m| phi) for m < 0, with the Condon-Shortley phase omitted, so
    that ``sum_m Y_lm(a) Y_lm(b) = (2l+1)/(4 pi) P_l(a.b)``.
    """
    k = np.asarray(k, dtype=np.float64)
    mu = np.clip(k[:, 2], -1.0, 1.0)
    phi = np.arctan2(k[:, 1], k[:, 0])
    s = np.sqrt(np.maximum(0.0, 1.0 - mu * mu))
    n = k.shape[0]
    # associated Legendre P_l^m without Condon-Shortley phase
    P = np.zeros((L + 1, L + 1, n))
    P[0, 0] = 1.0
    for m in range(1, L + 1):
        P[m, m] = (2 * m - 1) * s * P[m - 1, m - 1]
    for m in range(0, L):
        P[m + 1, m] = (2 * m + 1) * mu * P[m, m]
    for m in range(0, L + 1):
        for l in range(m + 2, L + 1):
            P[l, m] = ((2 * l - 1) * mu * P[l - 1, m] - (l + m - 1) * P[l - 2, m]) / (l - m)
    Y = np.zeros(((L + 1) ** 2, n))
    for l in range(L + 1):
        for m in range(-l, l + 1):
            am = abs(m)
            norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - am) / math.factorial(l + am))
            if m == 0:
                Y[l * l + l] = norm * P[l, 0]
            elif m > 0:
                Y[l * l + l + m] = math.sqrt(2.0) * norm * P[l, am] * np.cos(am * phi)
            else:
                Y[l * l + l + m] = math.sqrt(2.0) * norm * P[l, am] * np.sin(am * phi)
    return Y


def moment_matrices(ang):
    """``D`` ((L+1)^2, NI) = w_i Y_lm(k_i) and ``M`` (NI, (L+1)^2) = Y_lm(k_i)."""
    Y = real_sph_harm(ang.L, ang.k)
    return Y * ang.w[None, :], np.ascontiguousarray(Y.T)


def moment_degree(L):
    """Legendre order l of each moment row ``l*l + l + m``."""
    return np.concatenate([np.full(2 * l + 1, l) for l in range(L + 1)])
