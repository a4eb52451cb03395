"""Coax line monitor on the CPU (em_lab_amd.coax, the oracle's or_line_accumulate): each piece is
checked against what the mathematics fixes, not against a retyped formula.

- geometry factor (src/Coax.cpp:42-83): equal to a direct dense solve of the same lattice
  Laplace problem, and to the charge on the inner conductor (Green's identity);
- V terms (src/Coax.cpp:136-149): the line integral of a discrete gradient field E = -grad phi
  with phi = 1 on the inner conductor's nodes and 0 on the shield's is exactly 1, whatever phi
  does in between;
- I / Icm loops (src/Coax.cpp:119-129,152-159): equal to the discrete curl of B summed over the
  enclosed faces (Stokes on the lattice), for a random B;
- the oracle's accumulation (shaders/line_accumulate.comp:30-48): one step from zero phasors gives
  stage 1 = blend * 2 s e^{-i wt} and stage 2 = blend^2 * 2 s e^{-i wt};
- the two-wave fit (src/Coax.cpp:248-288): synthetic V(t) = A e^{-j beta t} + B e^{+j beta t}
  returns A and B, and the derived reflection, VSWR and impedance.
"""
import ctypes as C
import math

import numpy as np
import pytest

from em_lab_amd import coax as X
from em_lab_amd.emfdtd import coax_geometry_factor
from oracle import oracle as O


def _direct_g(inner, outer):
    n = outer + 1
    lo, hi = (outer - inner) // 2, (outer - inner) // 2 + inner
    fixed = np.zeros((n, n), bool)
    phi = np.zeros((n, n))
    fixed[0, :] = fixed[-1, :] = fixed[:, 0] = fixed[:, -1] = True
    fixed[lo:hi + 1, lo:hi + 1] = True
    phi[lo:hi + 1, lo:hi + 1] = 1.0
    free = [(j, i) for j in range(n) for i in range(n) if not fixed[j, i]]
    idx = {p: k for k, p in enumerate(free)}
    A = np.zeros((len(free), len(free)))
    b = np.zeros(len(free))
    for (j, i), k in idx.items():
        A[k, k] = 4.0
        for jj, ii in ((j - 1, i), (j + 1, i), (j, i - 1), (j, i + 1)):
            if (jj, ii) in idx:
                A[k, idx[(jj, ii)]] = -1.0
            else:
                b[k] += phi[jj, ii]
    if free:
        for (j, i), v in zip(free, np.linalg.solve(A, b)):
            phi[j, i] = v
    g = float(np.sum(np.diff(phi, axis=0) ** 2) + np.sum(np.diff(phi, axis=1) ** 2))
    # Green's identity: G = the net flux out of the inner conductor (its charge in units of eps)
    q = 0.0
    for j in range(lo, hi + 1):
        for i in range(lo, hi + 1):
            for jj, ii in ((j - 1, i), (j + 1, i), (j, i - 1), (j, i + 1)):
                if not (lo <= jj <= hi and lo <= ii <= hi):
                    q += phi[j, i] - phi[jj, ii]
    return g, q


@pytest.mark.parametrize("inner,outer", [(1, 3), (1, 5), (1, 7), (2, 8), (3, 11)])
def test_geometry_factor_against_direct_solve(inner, outer):
    g_dir, charge = _direct_g(inner, outer)
    assert charge == pytest.approx(g_dir, rel=1e-10)
    assert coax_geometry_factor(inner, outer) == pytest.approx(g_dir, rel=1e-9)


def _field_arrays(n):
    return np.zeros((n ** 3, 4), np.float32), np.zeros((n ** 3, 4), np.float32)


def _evaluate(E, B, n, q, t, w, rot=(1.0, 0.0, 1.0, 0.0), blend=1.0, phasors=None):
    g = O.grid(n, n, n, 0.5, 0)
    ph = np.zeros((q.shape[0], 4), np.float32) if phasors is None else phasors
    now = np.zeros(q.shape[0], np.float32)
    O.lib().or_line_accumulate(C.byref(g), O.ptr(E), O.ptr(B), O.ptr(q), q.shape[0], O.ptr(t), O.ptr(w),
                               O.ptr(np.array(rot, np.float32)), blend, O.ptr(ph), O.ptr(now))
    return ph, now


@pytest.mark.parametrize("axis,sign,inner,outer", [(0, 1, 1, 5), (1, -1, 1, 5), (2, 1, 2, 8), (2, -1, 1, 7)])
def test_voltage_of_a_gradient_field(axis, sign, inner, outer):
    n = 24
    c = X.CoaxSetup(axis=axis, sign=sign, start=6 if sign > 0 else 17, center=(12.5, 12.5, 12.5), length=10,
                    inner=inner, outer=outer)
    q, t, w, layout = X.build_terms([c])
    assert layout == [(0, 9)] and q.shape[0] == 27
    rng = np.random.default_rng(5)
    u, v = (axis + 1) % 3, (axis + 2) % 3
    lo_u, hi_u = X._node_lo(12.5, inner), X._node_hi(12.5, inner)
    so_u, sh_u = X._node_lo(12.5, outer), X._node_hi(12.5, outer)
    # node potential: 1 on the inner conductor's nodes, 0 on and outside the shield, random between
    idx = np.arange(n)
    uu, vv = np.meshgrid(idx, idx, indexing="ij")
    inside_inner = (uu >= lo_u) & (uu <= hi_u) & (vv >= lo_u) & (vv <= hi_u)
    inside_shield = (uu > so_u) & (uu < sh_u) & (vv > so_u) & (vv < sh_u)
    phi2 = np.where(inside_inner, 1.0, np.where(inside_shield, rng.uniform(-2, 2, (n, n)), 0.0))
    E, B = _field_arrays(n)
    Ec = E.reshape(n, n, n, 4)  # (z, y, x, comp)
    for a_idx in range(n):
        for comp, d in ((u, (1, 0)), (v, (0, 1))):
            nxt = np.roll(phi2, -1, axis=0 if d == (1, 0) else 1)
            eu = (phi2 - nxt).astype(np.float32)  # E on the edge from node i to node i + 1
            sl = [slice(None)] * 3
            sl[2 - axis] = a_idx
            # place eu[u index, v index] at lattice (axis=a_idx, u, v)
            arr = np.moveaxis(Ec[..., comp], [2 - axis, 2 - u, 2 - v], [0, 1, 2])
            arr[a_idx] = eu
    _, now = _evaluate(E, B, n, q, t, w)
    np.testing.assert_allclose(now[0::3], 1.0, rtol=0, atol=2e-6)


def test_loop_is_enclosed_curl():
    n = 20
    rng = np.random.default_rng(11)
    E, B = _field_arrays(n)
    B[:, :3] = rng.standard_normal((n ** 3, 3)).astype(np.float32)
    for axis in range(3):
        c = X.CoaxSetup(axis=axis, sign=1, start=4, center=(10.5, 10.5, 10.5), length=6, inner=1, outer=5)
        q, t, w, _ = X.build_terms([c])
        _, now = _evaluate(E, B, n, q, t, w)
        u, v = (axis + 1) % 3, (axis + 2) % 3
        Bc = np.moveaxis(B.reshape(n, n, n, 4).astype(np.float64), [2 - axis, 2 - u, 2 - v], [0, 1, 2])
        # discrete curl along +axis at the face of node (iu, iv): the E-update backward difference
        # (dBv/du - dBu/dv) with B_v at (iu, iv) - (iu - 1, iv) and B_u at (iu, iv) - (iu, iv - 1)
        for k in range(1, c.length):
            ib = c.start + k
            for quantity, size in ((1, c.inner), (2, c.outer + 2 * c.wall)):
                u0, u1 = X._node_lo(10.5, size), X._node_hi(10.5, size)
                v0, v1 = u0, u1
                curl = 0.0
                for iu in range(u0, u1 + 1):
                    for iv in range(v0, v1 + 1):
                        curl += (Bc[ib, iu, iv, v] - Bc[ib, iu - 1, iv, v]) - (Bc[ib, iu, iv, u] - Bc[ib, iu, iv - 1, u])
                assert now[3 * (k - 1) + quantity] == pytest.approx(curl, abs=1e-4)


def test_oracle_accumulation_one_step():
    n = 8
    E, B = _field_arrays(n)
    E[:, 0] = 0.75
    B[:, 1] = -0.5
    q = np.array([[0, 2, 0, 0], [2, 1, 1, 0]], np.int32)
    t = np.array([[1, 2, 3, 0], [99, -5, 3, 0], [4, 4, 4, 4]], np.int32)  # the second term clamps
    w = np.array([0.5, 2.0, 3.0], np.float32)
    rot = (0.6, 0.8, 0.28, 0.96)
    ph, now = _evaluate(E, B, n, q, t, w, rot=rot, blend=0.25)
    s0, s1 = 0.5 * 0.75 + 2.0 * 0.75, 3.0 * -0.5
    assert now.tolist() == [np.float32(s0), np.float32(s1)]
    for k, (s, (c, sn)) in enumerate(((s0, rot[:2]), (s1, rot[2:]))):
        obs = np.array([2 * s * c, -2 * s * sn])
        np.testing.assert_allclose(ph[k, :2], 0.25 * obs, rtol=1e-6)
        np.testing.assert_allclose(ph[k, 2:], 0.0625 * obs, rtol=1e-6)


def test_two_wave_fit_recovers_waves():
    c = X.CoaxSetup(length=60, outer=5, wavelength=20.0, epsilon_r=2.25, z0_ohm=50.0, source_ohm=50.0, drive=1.0)
    dt = 0.5
    beta = X.lattice_beta(c.wavelength, c.epsilon_r, dt)
    A, Bw = 0.3 - 0.2j, 0.05 + 0.07j
    z0n = 50.0 / X.ETA0
    n = c.length - 1
    ph = np.zeros((3 * n, 4), np.float64)
    for k in range(n):
        tv, ti = k + 1.0, k + 1.5
        V = A * np.exp(-1j * beta * tv) + Bw * np.exp(1j * beta * tv)
        I = (A * np.exp(-1j * beta * ti) - Bw * np.exp(1j * beta * ti)) / z0n
        ph[3 * k, 2:] = V.real, V.imag
        ph[3 * k + 1, 2:] = I.real, I.imag
        ph[3 * k + 2, 2:] = 1e-3 * abs(A) / z0n, 0.0
    # exact in FP64 before the float32 round trip the device phasors take: compare loosely
    r = X.analyse(c, n, ph.astype(np.float32), np.zeros(3 * n, np.float32), dt)
    assert r.valid
    assert abs(r.A - A) < 1e-6 and abs(r.B - Bw) < 1e-6
    assert r.fit_residual < 1e-6
    assert r.z0_measured == pytest.approx(50.0, rel=1e-5)
    g = Bw / A * np.exp(2j * beta * c.length)
    assert abs(r.gamma_end - g) < 1e-5
    assert r.vswr == pytest.approx((1 + abs(g)) / (1 - abs(g)), rel=1e-5)
    assert abs(r.z_end - 50.0 * (1 + g) / (1 - g)) < 1e-3
    assert r.common_mode_db == pytest.approx(-60.0, abs=1e-3)
    assert r.forward_fraction == pytest.approx(abs(A) ** 2 / (2 * z0n) / (1.0 / (8 * z0n)), rel=1e-5)
    # a pulsed line or fewer than 8 stations: no fit
    assert not X.analyse(X.CoaxSetup(length=60, pulsed=True), n, ph.astype(np.float32), np.zeros(3 * n), dt).valid


def test_scene_coax_builder():
    """Scene.coax (src/ScriptEngine.cpp:1961-2127): dielectric from the geometry factor, the
    inner conductor's extent per termination, generator and load feeds, the sleeve balun."""
    from em_lab_amd.scene import Scene
    g = coax_geometry_factor(1, 5)
    for term, inner_cells in (("load", 62), ("short", 63), ("open", 61)):
        sc = Scene(size=(96, 32, 32), pml=6, wavelength=16.0)
        c, _, _ = sc.coax((10, 16, 16), axis=0, termination=term)
        assert c.epsilon_r == pytest.approx((X.ETA0 / g / 50.0) ** 2, rel=1e-6) and c.z0_ohm == pytest.approx(50.0)
        assert sc.eps[16, 16, 11] == 1.0 and sc.sigma[16, 16, 11] > 1e5  # inner conductor
        assert int((sc.sigma[16, 16, :] > 1e5).sum()) == inner_cells + 2  # + both end plates
        assert len(sc.ports) == (2 if term == "load" else 1)
    sc = Scene(size=(96, 40, 40), pml=6, wavelength=16.0)
    c, _, _ = sc.coax((10, 20, 20), axis=0, balun=True, termination="open")
    assert (c.sleeve_start, c.sleeve_end, c.sleeve_outer) == (64 - 4 - 1, 64, 7 + 2 * 3)
    with pytest.raises(ValueError):
        Scene(size=(96, 32, 32)).coax((10, 16, 16), outer=4)
