"""Average-gradient (G-tilt) operator GS0 (mkg, lib/mkg.c:108-468): the
setup library's CSR operator and the oracle's cell-by-cell quadrature,
pinned by

- planes: any amplitude-weighted average of a constant gradient is that
  gradient, so every subaperture with enough illuminated area returns the
  plane's slope exactly;
- a brute-force midpoint quadrature (numpy) of the definition, the
  amplitude-weighted average over the subaperture square of the gradient of
  the bilinear OPD over cells whose four corners exist with non-zero
  amplitude, on a small ragged layout with missing neighbours and zero
  amplitudes;
- the operator applied to random OPDs equals the oracle's direct
  evaluation.
"""
import numpy as np


def _csr_apply(rowp, coli, val, opd):
    out = np.zeros(len(rowp) - 1)
    for r in range(len(rowp) - 1):
        s, e = rowp[r], rowp[r + 1]
        out[r] = np.dot(val[s:e], opd[coli[s:e]])
    return out


class _Geo:
    """A hand-made layout: subapertures on a 3x3 block with two missing,
    random amplitudes with a dark corner."""

    def __init__(self, nxsa=4, seed=3):
        rng = np.random.default_rng(seed)
        self.nxsa, self.dx = nxsa, 0.1
        cells = [(0, 0), (1, 0), (2, 0), (0, 1), (1, 1), (1, 2), (2, 2)]
        self.nsa = len(cells)
        self.ptsx = np.array([c[0] * nxsa * self.dx for c in cells]) + 0.05
        self.ptsy = np.array([c[1] * nxsa * self.dx for c in cells]) + 0.05
        self.amp = rng.uniform(0.3, 1.0, self.nsa * nxsa * nxsa)
        self.amp[2 * nxsa * nxsa: 2 * nxsa * nxsa + 3] = 0.0  # a dark corner in subaperture 2
        self.saa = np.ones(self.nsa)


def _mkg(g, saat=0.0):
    from maos_amd import _lib
    from maos_amd._lib import dptr, lptr
    nz = 2 * g.nsa * (g.nxsa + 2) ** 2
    rowp = np.zeros(2 * g.nsa + 1, dtype=np.int64)
    coli = np.zeros(nz, dtype=np.int64)
    val = np.zeros(nz)
    n = _lib.setup().mxs_mkg(lptr(rowp), lptr(coli), dptr(val), nz, dptr(g.ptsx), dptr(g.ptsy), g.nsa, g.nxsa,
                             g.dx, dptr(g.amp), dptr(g.saa), saat)
    assert n > 0
    return rowp, coli[:n], val[:n]


def _oracle_gtilt(g, opd, saat=0.0):
    import ctypes as C
    import oracle
    d = lambda a: np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.POINTER(C.c_double))
    out = np.zeros(2 * g.nsa)
    keep = [np.ascontiguousarray(x, dtype=np.float64) for x in (opd, g.amp, g.ptsx, g.ptsy, g.saa)]
    oracle.lib().orc_gtilt(d(out), d(keep[0]), d(keep[1]), d(keep[2]), d(keep[3]), g.nsa, g.nxsa, g.dx, d(keep[4]),
                           saat)
    return out


def _brute(g, opd, m=40):
    """Midpoint rule with m x m samples per cell of the common grid."""
    n = g.nxsa
    pts = {}
    for isa in range(g.nsa):
        ox = int(round((g.ptsx[isa] - g.ptsx[0]) / g.dx))
        oy = int(round((g.ptsy[isa] - g.ptsy[0]) / g.dx))
        for iy in range(n):
            for ix in range(n):
                k = isa * n * n + iy * n + ix
                pts[(ox + ix, oy + iy)] = (g.amp[k], opd[k])
    out = np.zeros(2 * g.nsa)
    t = (np.arange(m) + 0.5) / m
    U, V = np.meshgrid(t, t)
    for isa in range(g.nsa):
        ox = int(round((g.ptsx[isa] - g.ptsx[0]) / g.dx))
        oy = int(round((g.ptsy[isa] - g.ptsy[0]) / g.dx))
        sa = sx = sy = 0.0
        for cy in range(oy - 1, oy + n):
            for cx in range(ox - 1, ox + n):
                c = [pts.get((cx + i, cy + j)) for j in (0, 1) for i in (0, 1)]
                if any(v is None or abs(v[0]) <= 1e-14 for v in c):
                    continue
                X, Y = cx + U, cy + V
                inside = (X >= ox - 0.5) & (X <= ox + n - 0.5) & (Y >= oy - 0.5) & (Y <= oy + n - 0.5)
                a = [v[0] for v in c]
                p = [v[1] for v in c]
                A = a[0] * (1 - U) * (1 - V) + a[1] * U * (1 - V) + a[2] * (1 - U) * V + a[3] * U * V
                Du = (p[1] - p[0]) * (1 - V) + (p[3] - p[2]) * V
                Dv = (p[2] - p[0]) * (1 - U) + (p[3] - p[1]) * U
                w = inside / (m * m)
                sa += np.sum(w * A)
                sx += np.sum(w * A * Du)
                sy += np.sum(w * A * Dv)
        if sa > 0.01 * n * n:
            out[isa] = sx / (sa * g.dx)
            out[isa + g.nsa] = sy / (sa * g.dx)
    return out


def _plane(g, a, b, c=1e-7):
    n = g.nxsa
    x = (g.ptsx[:, None, None] + np.arange(n)[None, None, :] * g.dx).repeat(n, 1)
    y = (g.ptsy[:, None, None] + np.arange(n)[None, :, None] * g.dx).repeat(n, 2)
    return (c + a * x + b * y).reshape(-1)


def test_gtilt_plane_exact():
    g = _Geo()
    opd = _plane(g, 2.5e-6, -1.25e-6)
    rowp, coli, val = _mkg(g)
    got = _csr_apply(rowp, coli, val, opd)
    np.testing.assert_allclose(got[:g.nsa], 2.5e-6, rtol=1e-12)
    np.testing.assert_allclose(got[g.nsa:], -1.25e-6, rtol=1e-12)
    orc = _oracle_gtilt(g, opd)
    np.testing.assert_allclose(orc[:g.nsa], 2.5e-6, rtol=1e-12)
    np.testing.assert_allclose(orc[g.nsa:], -1.25e-6, rtol=1e-12)


def test_gtilt_matches_brute_force_quadrature():
    g = _Geo(nxsa=4)
    rng = np.random.default_rng(9)
    opd = rng.normal(size=g.nsa * 16) * 1e-7
    want = _brute(g, opd, m=60)
    rowp, coli, val = _mkg(g)
    got = _csr_apply(rowp, coli, val, opd)
    orc = _oracle_gtilt(g, opd)
    scale = np.abs(want).max()
    # the midpoint rule's error is O(1/m^2) per cell
    np.testing.assert_allclose(got, want, atol=2e-4 * scale)
    np.testing.assert_allclose(orc, want, atol=2e-4 * scale)
    np.testing.assert_allclose(got, orc, rtol=1e-11, atol=1e-12 * scale)


def test_gtilt_operator_equals_oracle_on_a_pupil(lgs_small):
    """On a real pupil (annulus, partial subapertures, missing neighbours at
    the edges) the CSR operator times a random OPD equals the oracle's
    direct evaluation."""
    import oracle
    cfg, p, _ = lgs_small
    rowp, coli, val = p.gs0()
    assert len(rowp) == 2 * p.nsa + 1
    rng = np.random.default_rng(2)
    opd = rng.normal(size=p.nsa * p.nxsa ** 2) * 1e-7
    got = _csr_apply(rowp, coli, val, opd)
    want = oracle.Oracle(p).gtilt(opd)
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-12 * np.abs(want).max())
    # every row reads its own grid and at most one ring of neighbours
    for r in range(2 * p.nsa):
        cols = coli[rowp[r]:rowp[r + 1]]
        assert len(cols) <= (p.nxsa + 2) ** 2


def test_gtilt_saat_leaves_rows_empty():
    g = _Geo()
    g.saa = np.array([1, 1, 0.2, 1, 1, 1, 1.0])
    rowp, coli, val = _mkg(g, saat=0.3)
    assert rowp[3] == rowp[2] and rowp[g.nsa + 3] == rowp[g.nsa + 2]
    opd = _plane(g, 1e-6, 1e-6)
    orc = np.full(2 * g.nsa, 7.0)
    from ctypes import POINTER, c_double
    import oracle
    keep = [np.ascontiguousarray(x, dtype=np.float64) for x in (opd, g.amp, g.ptsx, g.ptsy, g.saa)]
    d = lambda a: a.ctypes.data_as(POINTER(c_double))
    oracle.lib().orc_gtilt(d(orc), d(keep[0]), d(keep[1]), d(keep[2]), d(keep[3]), g.nsa, g.nxsa, g.dx, d(keep[4]),
                           0.3)
    assert orc[2] == 0.0 and orc[g.nsa + 2] == 0.0  # an empty row applied gives 0
    assert orc[0] != 7.0
