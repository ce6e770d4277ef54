"""A DEM trajectory as a differentiable torch function.

`Trajectory(s, every, nsamples)` maps contact parameters (and optionally the
initial positions and velocities) to the particle positions sampled every
`every` steps, nsamples times, running the HIP engine forward; its backward
is the engine's adjoint of a trajectory loss (sph_adjoint_samples), so that
any torch loss on the sampled positions has gradients with respect to k_n,
gamma_n, mu, x0 and v0.  `BatchTrajectory(sims, every, nsamples)` does the
same for many simulations at once, each with its own parameters, in one
batched context and one batched adjoint.  No counterpart in the reference.

The forward pass opens a fresh context from the .bin image of the
parameters and initial state.  When a gradient is wanted it runs
sph_adjoint_record: the same steps, recording the tape as it goes, and the
samples; the backward pass then runs sph_adjoint_reverse on that context
with the loss gradient of every sample (each enters the reverse pass at
its step, so the cotangents of the accelerations and contact
displacements carry across sampling times) and closes it.  Without a
gradient the forward pass simply steps.  The engine records the run in segments from checkpoints of its own
when the whole tape does not fit (SPH_S_ADJ_CHECKPOINT sets the segment
length).  Lengths are differentiated as exact roots and the contact list and
branches of each step are frozen, as in sph_adjoint.
"""

import numpy
import torch

from . import engine

# How each named parameter enters the simulation: a value sets every listed
# field of the .bin, and its gradient is the sum of theirs.  The contact
# laws take k_t = k_n and mu_s = mu_d in the workloads here.
TIES = {
    "k_n": ("k_n", "k_t"),
    "k_t": ("k_t",),
    "gamma_n": ("gamma_n",),
    "gamma_t": ("gamma_t",),
    "mu": ("mu_s", "mu_d"),
    "mu_s": ("mu_s",),
    "mu_d": ("mu_d",),
}


class Trajectory:
    """Positions sampled along a run of the engine, differentiable.

    `s` is a sphere_amd.sim.sim holding the initial state and the other
    parameters; `params` names the differentiable parameters (keys of
    `ties`).  Calling the object with a 1-D tensor theta (one value per
    name) and optional x0, v0 (np x 3 tensors) returns a float64 tensor of
    shape (nsamples, np, 3): positions after every, 2 every, ... steps."""

    def __init__(self, s, every, nsamples, params=("k_n", "gamma_n", "mu"), ties=None,
                 device=-1):
        if every <= 0 or nsamples <= 0:
            raise ValueError("every and nsamples must be positive")
        self.template = s
        self.every = int(every)
        self.nsamples = int(nsamples)
        self.params = tuple(params)
        self.ties = dict(TIES if ties is None else ties)
        for p in self.params:
            if p not in self.ties:
                raise ValueError("unknown parameter %r" % p)
        self.device = device
        self.np = int(s.np[0])

    def image(self, theta, x0=None, v0=None):
        """The .bin image of the template with theta (and x0, v0) applied."""
        from .sim import sim
        s = sim().frombytes(self.template.tobytes())
        for name, value in zip(self.params, numpy.asarray(theta, dtype=numpy.float64)):
            for field in self.ties[name]:
                getattr(s, field)[0] = float(value)
        if x0 is not None:
            s.x[:, :3] = numpy.asarray(x0, dtype=numpy.float64).reshape(self.np, 3)
        if v0 is not None:
            s.vel[:, :3] = numpy.asarray(v0, dtype=numpy.float64).reshape(self.np, 3)
        return s.tobytes()

    def __call__(self, theta, x0=None, v0=None):
        if x0 is None:
            x0 = torch.as_tensor(self.template.x[:, :3], dtype=torch.float64)
        if v0 is None:
            v0 = torch.as_tensor(self.template.vel[:, :3], dtype=torch.float64)
        return _TrajectoryFn.apply(theta, x0, v0, self)

    # -- the two passes ---------------------------------------------------
    def run_forward(self, theta, x0, v0, record=False):
        """Samples (nsamples, np, 3), and with `record` the context holding
        the recorded run for run_backward (else None)."""
        ctx = engine.Context(self.image(theta, x0, v0), device=self.device)
        if record:
            try:
                return ctx.adjoint_record(self.every, self.nsamples), ctx
            except Exception:
                ctx.close()
                raise
        samples = numpy.zeros((self.nsamples, self.np, 3))
        try:
            for k in range(self.nsamples):
                ctx.step(self.every)
                samples[k] = ctx.get("x")[:, :3]
        finally:
            ctx.close()
        return samples, None

    def run_backward(self, ctx, grad_samples):
        """Gradients (theta, x0, v0) of sum_k grad_samples[k] . sample_k;
        closes the context run_forward recorded the run in."""
        try:
            g = ctx.adjoint_reverse(self.every, self.nsamples, grad_samples, None)
        finally:
            ctx.close()
        gtheta = numpy.array([sum(g[f] for f in self.ties[name]) for name in self.params])
        return gtheta, g["x"], g["v"]


class BatchTrajectory:
    """Many independent trajectories (one per simulation) in one batched
    context, differentiable: each simulation has its own parameters.

    `sims` is a list of sphere_amd.sim.sim.  Calling the object with theta
    of shape (nsims, len(params)) and optional x0, v0 of shape (rows, 3)
    (the simulations' particles concatenated in order) returns a float64
    tensor of shape (nsamples, rows, 3).  Its backward runs one batched
    adjoint (sph_adjoint_samples_batch): each simulation's parameter
    gradient comes from its own rows of the loss gradient."""

    def __init__(self, sims, every, nsamples, params=("k_n", "gamma_n", "mu"), ties=None,
                 device=-1):
        if every <= 0 or nsamples <= 0 or not sims:
            raise ValueError("every, nsamples and the batch must be non-empty")
        self.parts = [Trajectory(s, every, nsamples, params=params, ties=ties, device=device)
                      for s in sims]
        self.every = int(every)
        self.nsamples = int(nsamples)
        self.params = tuple(params)
        self.ties = self.parts[0].ties
        self.device = device
        self.sizes = [t.np for t in self.parts]
        self.offsets = numpy.concatenate([[0], numpy.cumsum(self.sizes)]).astype(int)
        self.rows = int(self.offsets[-1])

    def images(self, theta, x0=None, v0=None):
        theta = numpy.asarray(theta, dtype=numpy.float64).reshape(len(self.parts), -1)
        out = []
        for k, t in enumerate(self.parts):
            a, b = self.offsets[k], self.offsets[k + 1]
            out.append(t.image(theta[k], None if x0 is None else numpy.asarray(x0)[a:b],
                               None if v0 is None else numpy.asarray(v0)[a:b]))
        return out

    def __call__(self, theta, x0=None, v0=None):
        if x0 is None:
            x0 = torch.as_tensor(numpy.concatenate([t.template.x[:, :3] for t in self.parts]),
                                 dtype=torch.float64)
        if v0 is None:
            v0 = torch.as_tensor(numpy.concatenate([t.template.vel[:, :3] for t in self.parts]),
                                 dtype=torch.float64)
        return _TrajectoryFn.apply(theta, x0, v0, self)

    def run_forward(self, theta, x0, v0, record=False):
        ctx = engine.Context(images=self.images(theta, x0, v0), device=self.device)
        if record:
            try:
                return ctx.adjoint_record(self.every, self.nsamples), ctx
            except Exception:
                ctx.close()
                raise
        samples = numpy.zeros((self.nsamples, self.rows, 3))
        try:
            for k in range(self.nsamples):
                ctx.step(self.every)
                for j in range(ctx.nsims):
                    a, b = self.offsets[j], self.offsets[j + 1]
                    samples[k, a:b] = ctx.get("x", sim=j)[:, :3]
        finally:
            ctx.close()
        return samples, None

    def run_backward(self, ctx, grad_samples):
        try:
            g = ctx.adjoint_reverse(self.every, self.nsamples, grad_samples, None)
        finally:
            ctx.close()
        gtheta = numpy.stack([sum(numpy.asarray(g[f]) for f in self.ties[name])
                              for name in self.params], axis=1)
        return gtheta, g["x"], g["v"]


class _TrajectoryFn(torch.autograd.Function):
    @staticmethod
    def forward(fctx, theta, x0, v0, traj):
        record = any(fctx.needs_input_grad[:3])
        samples, saved = traj.run_forward(theta.detach().cpu().numpy(),
                                          x0.detach().cpu().numpy(), v0.detach().cpu().numpy(),
                                          record=record)
        fctx.traj = traj
        fctx.saved = saved  # the context holding the recorded run
        fctx.dev = theta.device
        return torch.from_numpy(samples).to(fctx.dev)

    @staticmethod
    def backward(fctx, grad):
        if fctx.saved is None:
            raise RuntimeError("the trajectory's forward pass recorded no run to reverse")
        saved, fctx.saved = fctx.saved, None
        gtheta, gx, gv = fctx.traj.run_backward(saved, grad.detach().cpu().numpy())
        dev = fctx.dev
        return (torch.from_numpy(numpy.ascontiguousarray(gtheta)).to(dev),
                torch.from_numpy(gx).to(dev), torch.from_numpy(gv).to(dev), None)
