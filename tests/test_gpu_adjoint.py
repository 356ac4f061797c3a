"""The differentiable flow past a sphere and its hand-written adjoint (pic_adjoint_*), on the
GPU, against central finite differences of the GPU forward run and against the FP64 CPU
reference gradient: central finite differences of the oracle's restatement of the same run
(oracle/pic_oracle_diff.c).

Two cases on the ch3/ver2 domain at 21x21x41 nodes, 1e5 ions loaded plus 2000 injected per
step, 20 steps, checkpoints every 5:
  - the sphere case's parameters (7 km/s, -100 V, 1.5 eV) and its loss, the ion current to
    the sphere plus the wake density L2.  Its -100 V sheath fills the domain, so Te moves
    the loss by a few parts per million of what the other two do; Te is checked against the
    size of the whole gradient there, and component by component in the second case;
  - a sphere at -3 V, where the Boltzmann electrons shape the potential, with the wake
    potential L2 added to the loss: all three components are checked one by one.
The finite differences use eps = 1e-4 relative to each parameter; the tolerance is 1e-3.
"""
import numpy as np
import pytest

import pic_amd as P
from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

CASE = dict(ni=21, nj=21, nk=41, x0=(-0.1, -0.1, 0.0), xm=(0.1, 0.1, 0.4),
            sphere_c=(0.0, 0.0, 0.15), sphere_r=0.05, dt=1e-7, steps=20, checkpoint_every=5,
            mass=16 * O.AMU, charge=O.QE, n0=1e10, T_ion=1000.0, n_init=100000, n_inject=2000,
            soft_width=1.0, cg_tol=1e-9, nr_tol=1e-11, adj_rel_tol=1e-10, seed=7)
EPS = 1e-4
TOL = 1e-3


@pytest.fixture(scope="module")
def run():
    r = P.SphereAdjointRun(**CASE)
    yield r
    r.close()


@pytest.fixture(scope="module")
def mask(run):
    return run.wake_mask()


def gpu_outputs(run, theta):
    dev = torch.device("cuda", 0)
    den = torch.empty(run.nn, dtype=torch.float64, device=dev)
    phi = torch.empty(run.nn, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    cur = run.forward(theta, den.data_ptr(), phi.data_ptr())
    torch.cuda.synchronize()
    return cur, den.cpu().numpy(), phi.cpu().numpy()


def oracle_outputs(theta):
    cur, den, phi, _ = O.dsph_run(O.dsph_config(**CASE), theta)
    return cur, den, phi


class Loss:
    """L = I/I0 + mean_wake(den^2)/D0 [+ mean_wake(phi^2)/P0], normalised at theta0."""

    def __init__(self, mask, outs0, with_phi):
        cur, den, phi = outs0
        self.mask, self.with_phi = mask, with_phi
        self.I0 = cur
        self.D0 = np.mean(den[mask] ** 2)
        self.P0 = np.mean(phi[mask] ** 2)

    def __call__(self, outs):
        cur, den, phi = outs
        L = cur / self.I0 + np.mean(den[self.mask] ** 2) / self.D0
        if self.with_phi:
            L += np.mean(phi[self.mask] ** 2) / self.P0
        return L

    def adjoint_inputs(self, outs):
        cur, den, phi = outs
        m = self.mask
        dden = np.where(m, 2 * den / (m.sum() * self.D0), 0.0)
        dphi = np.where(m, 2 * phi / (m.sum() * self.P0), 0.0) if self.with_phi else None
        return 1.0 / self.I0, dden, dphi


def adjoint_gradient(run, theta, loss):
    outs = gpu_outputs(run, theta)
    dcur, dden, dphi = loss.adjoint_inputs(outs)
    dev = torch.device("cuda", 0)
    tden = torch.tensor(dden, device=dev)
    tphi = torch.tensor(dphi, device=dev) if dphi is not None else None
    torch.cuda.synchronize()
    g = run.backward(dcur, tden.data_ptr(), tphi.data_ptr() if tphi is not None else 0)
    return g, outs


def central_fd(fn, loss, theta):
    g = np.zeros(3)
    for a in range(3):
        h = EPS * abs(theta[a])
        tp, tm = theta.copy(), theta.copy()
        tp[a] += h
        tm[a] -= h
        g[a] = (loss(fn(tp)) - loss(fn(tm))) / (2 * h)
    return g


def test_forward_matches_oracle_run(run, mask):
    """The GPU run and the oracle's: the same loads, injections, pushes, soft absorption and
    fixed point of the Newton solve (the CG differs: multigrid-preconditioned on the GPU,
    Jacobi in the oracle, both converged to RMS(F) < 1e-9)."""
    theta = np.array([7000.0, -100.0, 1.5])
    cur, den, phi = gpu_outputs(run, theta)
    cur_o, den_o, phi_o = oracle_outputs(theta)
    assert cur > 0 and np.isfinite(den).all() and np.isfinite(phi).all()
    assert abs(cur - cur_o) <= 1e-6 * abs(cur_o), (cur, cur_o)
    assert np.linalg.norm(den - den_o) <= 1e-6 * np.linalg.norm(den_o)
    assert np.linalg.norm(phi - phi_o) <= 1e-8 * np.linalg.norm(phi_o)
    st = run.stats()
    assert st[2] >= CASE["steps"] + 4  # one Newton solve per step, four for the initial one


def _check(name, g_adj, g_ref, theta, per_component):
    s_adj, s_ref = g_adj * np.abs(theta), g_ref * np.abs(theta)  # d L / d ln theta
    scale = np.abs(s_ref).max()
    for a in range(3):
        bound = TOL * (abs(s_ref[a]) if per_component[a] else scale)
        assert abs(s_adj[a] - s_ref[a]) <= bound, (name, a, s_adj, s_ref)


def test_adjoint_matches_finite_differences_sphere_case(run, mask):
    theta = np.array([7000.0, -100.0, 1.5])
    loss = Loss(mask, gpu_outputs(run, theta), with_phi=False)
    g_adj, _ = adjoint_gradient(run, theta, loss)
    g_fd = central_fd(lambda t: gpu_outputs(run, t), loss, theta)
    _check("gpu fd", g_adj, g_fd, theta, (True, True, False))
    loss_o = Loss(mask, oracle_outputs(theta), with_phi=False)
    g_or = central_fd(oracle_outputs, loss_o, theta)
    _check("oracle fd", g_adj, g_or, theta, (True, True, False))


def test_adjoint_matches_finite_differences_all_parameters(run, mask):
    theta = np.array([7000.0, -3.0, 1.5])
    loss = Loss(mask, gpu_outputs(run, theta), with_phi=True)
    g_adj, _ = adjoint_gradient(run, theta, loss)
    g_fd = central_fd(lambda t: gpu_outputs(run, t), loss, theta)
    _check("gpu fd", g_adj, g_fd, theta, (True, True, True))
    loss_o = Loss(mask, oracle_outputs(theta), with_phi=True)
    g_or = central_fd(oracle_outputs, loss_o, theta)
    _check("oracle fd", g_adj, g_or, theta, (True, True, True))


def test_autograd_function_gives_the_adjoint_gradient(run, mask):
    """The torch.autograd.Function over the same C ABI: loss.backward() yields the library's
    adjoint gradient."""
    import adjoint
    theta0 = np.array([7000.0, -100.0, 1.5])
    loss = Loss(mask, gpu_outputs(run, theta0), with_phi=False)
    g_ref, _ = adjoint_gradient(run, theta0, loss)
    theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    cur, den, phi = adjoint.SphereRunFunction.apply(theta, run)
    m = torch.tensor(mask, device=den.device)
    L = adjoint.sphere_loss(cur, den, m, loss.I0, loss.D0)
    assert abs(L.item() - 2.0) < 1e-9
    L.backward()
    np.testing.assert_allclose(theta.grad.numpy(), g_ref, rtol=1e-9, atol=0)
