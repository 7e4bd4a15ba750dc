"""The FP64 CPU reference gradient (oracle/des_adjoint.c) against finite
differences of the oracle's forward step, and its geometric building blocks
against the oracle's volume and shape functions.

Finite differences of a quasi-static run are only as good as the forward
model's own rounding noise allows: the checks use per-element viscosities of
1e19-1e21 Pa s, where Maxwell relaxation changes the surface within a few
steps, and central differences in log space (x * exp(+-1e-3))."""
import ctypes as C

import numpy as np
import pytest

import scenarios as SC
from dynearthsol_amd import abi, model as M
from oracle import OracleState, adjoint_lib


def _setup(case, nelem=600, perturb=0.0, coh_scale=1.0, damping=None):
    m = M.config_adjoint(case, nelem_target=nelem, log_visc=(19.0, 21.0))
    if damping is not None:
        m.params.damping_option = damping
    if perturb:
        SC._perturb_stress(m, perturb, 11)
        m.elem_cohesion[:] *= coh_scale
    top = (m.bcflag & 32) != 0
    cb = np.zeros((m.nnode, 3))
    cb[top, 2] = np.random.default_rng(0).standard_normal(int(top.sum()))
    return m, cb


def _fd(o0, m, cb, nsteps, which, e, h=1e-3):
    base = m.elem_visc if which == "visc" else m.elem_cohesion
    vals = []
    for sgn in (1, -1):
        o = o0.copy()
        a = base.copy()
        a[e] = base[e] * np.exp(sgn * h)
        o.set_elem_params(**{("elem_visc" if which == "visc" else "elem_cohesion"): a})
        o.step(nsteps)
        vals.append(float(np.sum(cb * o.f["coord"])))
    return (vals[0] - vals[1]) / (2 * h) / base[e]


def test_tet_gradient_blocks():
    """G_i = d vol / d x_i equals the reference's shape-function derivative
    times the volume (geometry.cxx:400-436), and the Hessian-vector product
    matches central differences of G."""
    lib = adjoint_lib()
    m = SC.build("evp3")
    o = OracleState(m)
    rng = np.random.default_rng(1)
    for e in rng.integers(0, m.nelem, 20):
        x = np.ascontiguousarray(o.f["coord"][m.conn[e]])
        G = np.zeros((4, 3))
        lib.adj3_tet_grad(abi.dptr(x), abi.dptr(G))
        vol = o.f["volume"][e]
        shp = np.stack([o.f["shpdx"][e], o.f["shpdy"][e], o.f["shpdz"][e]], axis=1)
        assert np.allclose(G, shp * vol, rtol=1e-12, atol=1e-12 * np.abs(G).max())
        Gb = rng.standard_normal((4, 3))
        xb = np.zeros((4, 3))
        lib.adj3_tet_grad_adj(abi.dptr(x), abi.dptr(np.ascontiguousarray(Gb)), abi.dptr(xb))
        d = rng.standard_normal((4, 3))
        h = 1e-3 * np.abs(x).max()
        Gp, Gm = np.zeros((4, 3)), np.zeros((4, 3))
        lib.adj3_tet_grad(abi.dptr(np.ascontiguousarray(x + h * d)), abi.dptr(Gp))
        lib.adj3_tet_grad(abi.dptr(np.ascontiguousarray(x - h * d)), abi.dptr(Gm))
        fd = np.sum(Gb * (Gp - Gm)) / (2 * h)      # exact: G is quadratic in x
        assert abs(np.sum(xb * d) - fd) <= 1e-9 * max(abs(fd), np.abs(xb).max() * np.abs(d).sum())


@pytest.mark.parametrize("damping", [0, 1])
def test_maxwell_gradient_matches_fd(damping):
    """Damping option 1 (fields.cxx:251-256) is piecewise linear in the force,
    with kinks where a force component changes sign; a finite difference that
    straddles one differs from the one-sided derivative, so the damped case
    is held to 1e-4 and the undamped one to 1e-5."""
    m, cb = _setup("maxwell", damping=damping)
    o0 = OracleState(m)
    vb, hb = o0.copy().gradient(15, cb)
    assert not np.any(hb)
    tol = 1e-5 if damping == 0 else 1e-4
    for e in np.argsort(-np.abs(vb * m.elem_visc))[:4]:
        fd = _fd(o0, m, cb, 15, "visc", e)
        assert abs(vb[e] - fd) <= tol * abs(fd), (e, vb[e], fd)


def test_plastic_gradient_matches_fd():
    """Elasto-visco-plastic, with a deviatoric initial stress and low cohesion
    so that elements yield from the first step: viscosity and cohesion
    gradients of yielding elements."""
    m, cb = _setup("plastic", perturb=30e6, coh_scale=0.2, damping=0)
    o0 = OracleState(m)
    vb, hb = o0.copy().gradient(10, cb)
    assert np.count_nonzero(hb) > 10
    for which, g, base in (("visc", vb, m.elem_visc), ("coh", hb, m.elem_cohesion)):
        for e in np.argsort(-np.abs(g * base))[:4]:
            fd = _fd(o0, m, cb, 10, which, e)
            assert abs(g[e] - fd) <= 1e-5 * abs(fd), (which, e, g[e], fd)


def test_directional_derivative():
    """All parameters at once: the adjoint's directional derivative along a
    random log-space direction against a central difference."""
    m, cb = _setup("plastic", perturb=30e6, coh_scale=0.2, damping=1)
    o0 = OracleState(m)
    vb, hb = o0.copy().gradient(10, cb)
    rng = np.random.default_rng(7)
    dv, dc = rng.standard_normal(m.nelem), rng.standard_normal(m.nelem)
    ad = np.sum(vb * m.elem_visc * dv) + np.sum(hb * m.elem_cohesion * dc)
    h = 1e-4
    vals = []
    for sgn in (1, -1):
        o = o0.copy()
        o.set_elem_params(elem_visc=m.elem_visc * np.exp(sgn * h * dv),
                          elem_cohesion=m.elem_cohesion * np.exp(sgn * h * dc))
        o.step(10)
        vals.append(float(np.sum(cb * o.f["coord"])))
    fd = (vals[0] - vals[1]) / (2 * h)
    assert abs(ad - fd) <= 1e-4 * abs(fd), (ad, fd)


def test_unsupported_configuration_is_refused():
    m = SC.build("evp3")          # thermal expansion, adaptive dt
    o = OracleState(m)
    with pytest.raises(ValueError):
        o.gradient(1, np.zeros((m.nnode, 3)))
