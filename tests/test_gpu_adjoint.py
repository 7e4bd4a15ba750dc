"""The HIP adjoint (des_adj_forward / des_adj_backward) against the FP64 CPU
reference gradient (oracle/des_adjoint.c, itself pinned by finite
differences of the oracle), against finite differences of the HIP forward,
and through the torch.autograd binding."""
import math

import numpy as np
import pytest

import scenarios as SC
from dynearthsol_amd import model as M
from oracle import OracleState

pytestmark = pytest.mark.gpu


def _case(case, nelem=600, perturb=0.0, coh_scale=1.0, damping=None, log_visc=(19.0, 21.0)):
    m = M.config_adjoint(case, nelem_target=nelem, log_visc=log_visc)
    if damping is not None:
        m.params.damping_option = damping
    if perturb:
        SC._perturb_stress(m, perturb, 11)
        m.elem_cohesion[:] *= coh_scale
    top = (m.bcflag & 32) != 0
    cb = np.zeros((m.nnode, 3))
    cb[top, 2] = np.random.default_rng(0).standard_normal(int(top.sum()))
    return m, cb


def _gpu_grad(m, cb, nsteps, window):
    from dynearthsol_amd.lib import DesModel
    g = DesModel(m)
    g.adj_forward(nsteps, window)
    vb, hb = g.adj_backward(cb)
    return g, vb, hb


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("case,kw", [("maxwell", {}), ("maxwell", {"damping": 0}),
                                     ("plastic", {"perturb": 30e6, "coh_scale": 0.2})])
def test_gradient_matches_cpu_reference(case, kw):
    """15 steps in windows of 4 (three full windows and a partial one)."""
    m, cb = _case(case, **kw)
    print(case, kw, "elements", m.nelem, flush=True)
    g, vb, hb = _gpu_grad(m, cb, 15, 4)
    o = OracleState(m)
    vr, hr = o.gradient(15, cb)
    ev = _rel(vb, vr)
    eh = _rel(hb, hr) if np.any(hr) else float(np.max(np.abs(hb)))
    print(f"{case}: viscosity gradient {ev:.2e}, cohesion gradient {eh:.2e}")
    # both FP64; the forwards differ by libm and summation rounding, which a
    # yielding element's return amplifies
    tol = 1e-8 if case == "maxwell" else 1e-6
    assert ev < tol and eh < tol
    # the forward leaves the final state: the oracle's
    got = g.download(["coord"])["coord"]
    assert SC.rel_err(got, o.f["coord"]) < 1e-10


def test_gradient_matches_finite_differences():
    """Central differences in log space of the HIP forward, on the elements
    with the largest gradients: 1e-5 relative (maxwell, no damping)."""
    from dynearthsol_amd.lib import DesModel
    m, cb = _case("maxwell", nelem=3000, damping=0)
    g, vb, _ = _gpu_grad(m, cb, 20, 8)
    h = 1e-3
    for e in np.argsort(-np.abs(vb * m.elem_visc))[:6]:
        vals = []
        for sgn in (1, -1):
            a = m.elem_visc.copy()
            a[e] *= math.exp(sgn * h)
            gg = DesModel(m)
            gg.set_elem_params(elem_visc=a)
            gg.run(20)
            vals.append(float(np.sum(cb * gg.download(["coord"])["coord"])))
        fd = (vals[0] - vals[1]) / (2 * h) / m.elem_visc[e]
        print(f"e {e}: adjoint {vb[e]:.8e} fd {fd:.8e}", flush=True)
        assert abs(vb[e] - fd) <= 1e-5 * abs(fd)


def test_gradient_matches_finite_differences_on_random_elements():
    """64 elements drawn at random (seed 5), as the adjoint configuration's
    check prescribes, each to 1e-5 relative, over the configuration's
    horizon: 2000 steps, checkpointed every 100.

    A finite difference resolves a gradient only above the FP64 forward's
    rounding noise, about 1e-9 in d loss / d log(visc) here.  With the loss
    on the top surface alone some elements have gradients of 1e-6 of the
    largest, at that noise (measured: 2.4e-4 relative on one of the 64 after
    1000 steps).  So the check weighs the vertical displacement of every
    node (seeded weights): the same reverse sweep with a coord_bar that is
    non-zero everywhere, which gives every element a gradient of 2e-3 of the
    largest or more.  The difference: central differences at h = 0.4, 0.2,
    0.1 in log viscosity combined by two levels of Richardson extrapolation
    (error O(h^6)).  Maxwell, no damping.  Measured worst: 1.3e-6."""
    from dynearthsol_amd.lib import DesModel
    nsteps = 2000
    m, _ = _case("maxwell", nelem=2000, damping=0)
    cb = np.zeros((m.nnode, 3))
    cb[:, 2] = np.random.default_rng(3).standard_normal(m.nnode)
    g, vb, _ = _gpu_grad(m, cb, nsteps, 100)
    g.close()
    x0 = m.coord.copy()
    picks = np.random.default_rng(5).choice(m.nelem, 64, replace=False)
    ad_all = vb * m.elem_visc
    gg = DesModel(m)

    def loss(e, fac):
        a = m.elem_visc.copy()
        a[e] *= fac
        gg.reset()
        gg.set_elem_params(elem_visc=a)
        gg.run(nsteps)
        return float(np.sum(cb * (gg.download(["coord"])["coord"] - x0)))

    hs = (0.4, 0.2, 0.1)
    errs = []
    for e in picks:
        D = [(loss(e, math.exp(h)) - loss(e, math.exp(-h))) / (2 * h) for h in hs]
        R1 = [(4 * D[k + 1] - D[k]) / 3 for k in range(2)]
        fd = (16 * R1[1] - R1[0]) / 15              # d loss / d log(visc)
        ad = ad_all[e]
        errs.append(abs(ad - fd) / abs(fd))
        print(f"e {e:5d} |grad|/max {abs(ad) / np.abs(ad_all).max():.1e} adjoint {ad: .10e} fd {fd: .10e} "
              f"rel {errs[-1]:.1e}", flush=True)
    print(f"64 random elements: worst relative difference {max(errs):.2e}")
    assert max(errs) <= 1e-5


def test_plastic_gradient_matches_finite_differences():
    """Elasto-visco-plastic with a deviatoric initial stress and low cohesion,
    so that elements yield from the first step and stay on the yield
    surface: viscosity and cohesion gradients of the four elements with the
    largest of each, against central differences (h = 1e-3, log space) of
    the HIP forward, 1e-5 relative."""
    from dynearthsol_amd.lib import DesModel
    m, cb = _case("plastic", perturb=30e6, coh_scale=0.2, damping=0)
    g, vb, hb = _gpu_grad(m, cb, 10, 4)
    assert np.count_nonzero(hb) > 10
    gg = DesModel(m)
    h = 1e-3
    for which, grad, base in (("elem_visc", vb, m.elem_visc), ("elem_cohesion", hb, m.elem_cohesion)):
        for e in np.argsort(-np.abs(grad * base))[:4]:
            vals = []
            for sgn in (1, -1):
                a = base.copy()
                a[e] *= math.exp(sgn * h)
                gg.reset()
                gg.set_elem_params(**{which: a})
                gg.run(10)
                vals.append(float(np.sum(cb * gg.download(["coord"])["coord"])))
            fd = (vals[0] - vals[1]) / (2 * h) / base[e]
            print(f"{which} e {e}: adjoint {grad[e]:.8e} fd {fd:.8e}", flush=True)
            assert abs(grad[e] - fd) <= 1e-5 * abs(fd), (which, e, grad[e], fd)


def test_window_size_does_not_change_the_gradient():
    """Recomputation from checkpoints reproduces the forward exactly: the
    gradient is bitwise the same for any window."""
    m, cb = _case("plastic", perturb=30e6, coh_scale=0.2)
    _, v1, h1 = _gpu_grad(m, cb, 12, 3)
    _, v2, h2 = _gpu_grad(m, cb, 12, 12)
    _, v3, h3 = _gpu_grad(m, cb, 12, 5)
    assert np.array_equal(v1, v2) and np.array_equal(h1, h2)
    assert np.array_equal(v1, v3) and np.array_equal(h1, h3)


def test_repeated_backward_gives_the_same_gradient():
    """The forward records its last window, which the first backward uses
    without recomputing it and then overwrites: a second backward recomputes
    every window and must give the same gradient bitwise."""
    m, cb = _case("plastic", perturb=30e6, coh_scale=0.2)
    g, v1, h1 = _gpu_grad(m, cb, 11, 4)
    v2, h2 = g.adj_backward(cb)
    assert np.array_equal(v1, v2) and np.array_equal(h1, h2)


def test_taped_run_matches_checkpointed_run():
    """window=0 keeps every step's records from the forward: the same
    gradient bitwise, the same final state, and a second backward over the
    kept tape gives the same again."""
    m, cb = _case("plastic", perturb=30e6, coh_scale=0.2)
    gw, v1, h1 = _gpu_grad(m, cb, 12, 5)
    gt, v2, h2 = _gpu_grad(m, cb, 12, 0)
    assert np.array_equal(v1, v2) and np.array_equal(h1, h2)
    fw = gw.download(["coord", "vel", "stress"])
    ft = gt.download(["coord", "vel", "stress"])
    for k in fw:
        assert np.array_equal(fw[k], ft[k]), k
    v3, h3 = gt.adj_backward(cb)
    assert np.array_equal(v2, v3) and np.array_equal(h2, h3)


def test_oversized_tape_is_refused():
    """Refused before any step; a failed forward leaves nothing for a
    backward to run on (status 101, call out of order)."""
    from dynearthsol_amd.lib import DesError, DesModel
    m, cb = _case("maxwell")
    g = DesModel(m)
    g.adj_forward(3, 2)
    with pytest.raises(DesError):
        g.adj_forward(2_000_000_000, 0)
    with pytest.raises(DesError) as ex:
        g.adj_backward(cb)
    assert ex.value.code == 101


@pytest.mark.parametrize("dev", ["cuda", "cpu"])
def test_torch_autograd_binding(dev):
    """Device tensors (passed to the library as device pointers) and host
    tensors (host copies) give the C-ABI's gradient."""
    import torch
    from dynearthsol_amd.autograd import SurfaceTopography
    m, _ = _case("maxwell")
    topo = SurfaceTopography(m, nsteps=10, window=4)
    assert torch.cuda.is_available(), "torch sees no GPU next to the library"
    lv = torch.tensor(np.log10(m.elem_visc), dtype=torch.float64, device=dev, requires_grad=True)
    coh = torch.tensor(m.elem_cohesion, dtype=torch.float64, device=dev, requires_grad=True)
    uz = topo(lv, coh)
    target = uz.detach() * 0.9
    loss = ((uz - target) ** 2).sum()
    loss.backward()
    # the same gradient straight from the C-ABI
    cb = np.zeros((m.nnode, 3))
    cb[topo.top, 2] = 2 * (uz - target).detach().cpu().numpy()
    _, vb, hb = _gpu_grad(m, cb, 10, 4)
    # 10 ** log10(visc) may differ from visc in the last bit
    assert np.allclose(lv.grad.cpu().numpy(), vb * m.elem_visc * math.log(10), rtol=1e-10, atol=0)
    assert np.allclose(coh.grad.cpu().numpy(), hb, rtol=1e-10, atol=0)
    # a second evaluation restarts from the initial state
    uz2 = topo(lv.detach(), coh.detach())
    assert torch.equal(uz2, uz.detach())


def test_unsupported_configuration_is_refused():
    from dynearthsol_amd.lib import DesError, DesModel
    g = DesModel(SC.build("evp3"))
    with pytest.raises(DesError):
        g.adj_forward(2, 1)


def test_torch_binding_refuses_a_stale_graph():
    """The device keeps the checkpoints of the latest forward only: a graph
    whose forward was followed by another forward cannot be differentiated
    (it would return the later run's gradient), and says so."""
    import torch
    from dynearthsol_amd.autograd import SurfaceTopography
    m, _ = _case("maxwell")
    topo = SurfaceTopography(m, nsteps=6, window=3)
    assert torch.cuda.is_available()
    dev = "cuda"
    lv = torch.tensor(np.log10(m.elem_visc), dtype=torch.float64, device=dev, requires_grad=True)
    coh = torch.tensor(m.elem_cohesion, dtype=torch.float64, device=dev, requires_grad=True)
    first = topo(lv, coh).sum()
    second = topo(lv * 1.001, coh).sum()
    with pytest.raises(RuntimeError, match="another forward"):
        first.backward()
    second.backward()          # the latest graph still works
    assert torch.isfinite(lv.grad).all()


def test_device_pointer_entry_points_equal_host_ones():
    """des_upload_fields_device / des_download_fields_device /
    des_adj_backward_device on torch device tensors give bitwise what the
    host-pointer entry points give."""
    import torch
    from dynearthsol_amd.lib import DesModel
    m, cb = _case("maxwell")
    assert torch.cuda.is_available()
    dev = torch.device("cuda")
    g = DesModel(m)
    h = DesModel(m)
    v = torch.tensor(m.elem_visc * 1.5, dtype=torch.float64, device=dev)
    c = torch.tensor(m.elem_cohesion * 0.8, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    g.upload_device(elem_visc=v.data_ptr(), elem_cohesion=c.data_ptr())
    h.upload(elem_visc=m.elem_visc * 1.5, elem_cohesion=m.elem_cohesion * 0.8)
    g.adj_forward(8, 3)
    h.adj_forward(8, 3)
    x = torch.empty((m.nnode, 3), dtype=torch.float64, device=dev)
    s = torch.empty((m.nelem, 6), dtype=torch.float64, device=dev)
    g.download_device(coord=x.data_ptr(), stress=s.data_ptr())
    hf = h.download(["coord", "stress"])
    assert np.array_equal(x.cpu().numpy(), hf["coord"]) and np.array_equal(s.cpu().numpy(), hf["stress"])
    cbt = torch.tensor(cb, dtype=torch.float64, device=dev)
    vb = torch.empty(m.nelem, dtype=torch.float64, device=dev)
    hb = torch.empty(m.nelem, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    g.adj_backward_device(cbt.data_ptr(), vb.data_ptr(), hb.data_ptr())
    vh, hh = h.adj_backward(cb)
    assert np.array_equal(vb.cpu().numpy(), vh) and np.array_equal(hb.cpu().numpy(), hh)
