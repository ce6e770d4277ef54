"""The torch binding of the adjoint (sphere_amd.autograd.Trajectory):
gradients of a loss on positions sampled along a run, with respect to the
contact parameters and the initial state.

* against central differences of the same loss on the oracle with FP64
  lengths (the adjoint differentiates lengths as exact roots), on a box
  where contacts form during the run;
* against the sum of sph_adjoint runs, one per sample (the loss is
  linear in its seeds);
* on the granular column collapse (2,000 spheres, radii uniform in
  [0.008, 0.012] m, settled in a 0.3 m wide column and released onto a
  frictional floor, 500 steps sampled every 50, MSE against a trajectory
  made with k_n* = 2e7 and gamma_n* = 0.3 critical, from k_n = 1e7 and
  gamma_n = 0.2 critical): a step against the gradient lowers the loss.
"""

import numpy
import pytest
import torch

import oracle
from sphere_amd import autograd, engine, scenes

from . import cases

pytestmark = pytest.mark.gpu


def _oracle_samples(image, every, nsamples):
    o = oracle.Oracle(image)
    out = []
    for _ in range(nsamples):
        o.step(every)
        out.append(o.get("x")[:, :3].copy())
    return numpy.stack(out)


def test_trajectory_gradients_match_central_differences():
    s = cases.settling_case(300, 11)
    every, nsamples = 10, 4
    traj = autograd.Trajectory(s, every, nsamples, params=("k_n", "gamma_n", "mu"))
    rng = numpy.random.default_rng(8)
    target = torch.as_tensor(_oracle_samples(s.tobytes(), every, nsamples) +
                             1e-4 * rng.standard_normal((nsamples, 300, 3)))
    weights = torch.as_tensor(rng.uniform(0.5, 1.5, (nsamples, 1, 1)))
    theta0 = [float(s.k_n[0]), float(s.gamma_n[0]), float(s.mu_s[0])]
    theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    v0 = torch.as_tensor(s.vel[:, :3].copy(), dtype=torch.float64).requires_grad_(True)

    def loss_of(x):
        return float((weights * (torch.as_tensor(x) - target) ** 2).sum())

    X = traj(theta, v0=v0)
    loss = (weights * (X - target) ** 2).sum()
    loss.backward()
    oracle.set_exact_length(True)
    try:
        for j in range(3):
            h = 1e-6 * theta0[j]
            vals = []
            for sign in (1.0, -1.0):
                th = list(theta0)
                th[j] += sign * h
                vals.append(loss_of(_oracle_samples(traj.image(th), every, nsamples)))
            fd = (vals[0] - vals[1]) / (2 * h)
            ad = float(theta.grad[j])
            # The friction term comes from the few contacts that slide in
            # this short run; its central difference carries stick/slip
            # changes within +-h, hence the looser bound.
            tol = 1e-3 if traj.params[j] == "mu" else 1e-4
            assert abs(ad - fd) <= tol * abs(fd), (traj.params[j], ad, fd)
        d = rng.standard_normal((300, 3))
        h = 1e-6
        vals = []
        for sign in (1.0, -1.0):
            img = traj.image(theta0, v0=s.vel[:, :3] + sign * h * d)
            vals.append(loss_of(_oracle_samples(img, every, nsamples)))
        fd = (vals[0] - vals[1]) / (2 * h)
        ad = float((v0.grad.numpy() * d).sum())
        assert abs(ad - fd) <= 1e-4 * abs(fd), ("v0", ad, fd)
    finally:
        oracle.set_exact_length(False)


def test_sampled_adjoint_is_the_sum_of_single_adjoints():
    """sph_adjoint_samples with a term on every sample equals the sum of
    sph_adjoint runs, one per sample, each from the start to its sample
    (the loss is linear in the seeds); and the torch binding returns it."""
    s = cases.fixture("shear1k.bin")
    s.wmode[0] = 0
    every, nsamples = 8, 3
    n = int(s.np[0])
    rng = numpy.random.default_rng(3)
    sx = rng.standard_normal((nsamples, n, 3))
    image = s.tobytes()
    with engine.Context(image) as ctx:
        g_all = ctx.adjoint_samples(every, nsamples, sx)
    total = None
    for k in range(nsamples):
        with engine.Context(image) as ctx:
            g = ctx.adjoint(every * (k + 1), sx[k])
        total = g if total is None else {key: total[key] + g[key] for key in g}
    for key in ("x", "v", "w"):
        scale = numpy.abs(total[key]).max()
        assert numpy.allclose(g_all[key], total[key], rtol=1e-10, atol=1e-10 * scale), key
    for key in engine.ADJOINT_PARAMS:
        assert abs(g_all[key] - total[key]) <= 1e-9 * max(abs(total[key]), 1e-30), key
    traj = autograd.Trajectory(s, every, nsamples, params=("k_n", "gamma_n", "mu"))
    theta = torch.tensor([float(s.k_n[0]), float(s.gamma_n[0]), float(s.mu_s[0])],
                         dtype=torch.float64, requires_grad=True)
    x0 = torch.as_tensor(s.x[:, :3].copy()).requires_grad_(True)
    v0 = torch.as_tensor(s.vel[:, :3].copy()).requires_grad_(True)
    (traj(theta, x0=x0, v0=v0) * torch.as_tensor(sx)).sum().backward()
    assert numpy.array_equal(x0.grad.numpy(), g_all["x"])
    assert numpy.array_equal(v0.grad.numpy(), g_all["v"])
    # (parameter sums are formed in a fixed order: the same bits)
    want = [g_all["k_n"] + g_all["k_t"], g_all["gamma_n"], g_all["mu_s"] + g_all["mu_d"]]
    assert numpy.array_equal(theta.grad.numpy(), want)


def _crit(s, k_n, frac):
    return frac * scenes.critical_damping(float(s.radius.min()), float(s.rho[0]), k_n)


def test_column_collapse_gradient_descends():
    s = scenes.collapse_column(2000, seed=3001, k_n=1.0e7, damping_frac=0.2,
                               open_stepper=lambda img: engine.Context(img))
    every, nsamples = 50, 10
    traj = autograd.Trajectory(s, every, nsamples, params=("k_n", "gamma_n", "mu"))
    star = torch.tensor([2.0e7, _crit(s, 2.0e7, 0.3), 0.5], dtype=torch.float64)
    with torch.no_grad():
        target = traj(star)
    theta0 = torch.tensor([1.0e7, _crit(s, 1.0e7, 0.2), 0.5], dtype=torch.float64)
    theta = theta0.clone().requires_grad_(True)
    v0 = torch.as_tensor(s.vel[:, :3].copy()).requires_grad_(True)
    loss = ((traj(theta, v0=v0) - target) ** 2).mean()
    loss.backward()
    g = theta.grad.numpy()
    print("column: loss %.6e, dL/dk_n %.3e, dL/dgamma_n %.3e, dL/dmu %.3e, |dL/dv0| %.3e" % (
        float(loss.detach()), g[0], g[1], g[2], float(v0.grad.norm())), flush=True)
    assert numpy.all(numpy.isfinite(g)) and numpy.all(numpy.isfinite(v0.grad.numpy()))
    assert float(loss.detach()) > 0.0 and g[0] != 0.0 and g[1] != 0.0
    # a 1% step in log-parameters against the gradient lowers the loss
    glog = g * theta0.numpy()
    step = 0.01 / numpy.abs(glog).max()
    theta1 = theta0 * torch.as_tensor(numpy.exp(-step * glog))
    with torch.no_grad():
        loss1 = ((traj(theta1) - target) ** 2).mean()
    assert float(loss1) < float(loss.detach()), (float(loss1), float(loss.detach()))
