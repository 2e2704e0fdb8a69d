"""Device-resident HMC over many chains and many patches.

Every chain's position lives on the device; one leapfrog step is
    z -> theta = lower + sigmoid(z) (upper - lower)        (BoundedTransform,
                                                            forcepho/model.py:596-637)
      -> Source records (fluxes, shapes, mixture amplitudes and their
         derivatives)                                      (Galaxy.proposal,
                                                            forcepho/sources.py:607-621)
         both in one fpx_build_sources launch,
      -> one batched chi2 + gradient launch for all chains  (fpx_evaluate_batch)
      -> dU/dz = 1/2 stack_grad(dchi2) dq/dz - dlndetjac/dz (FastPosterior.evaluate,
                                                            forcepho/model.py:302-387)
         and the momentum / position updates in one fpx_leapfrog_update launch,
all as device operations with no host round trip, so a whole trajectory can
be recorded once into a graph (four kernels per step) and replayed.  The mixture amplitudes and
their derivatives are evaluated on the device for every (chain, source) by
device_scene.DeviceRecords: forcepho's spline tables (one
SplineMixtureAmplitudes shared by all scenes) or the analytic
LogNormalMixtureAmplitudes model.  Momentum draws and the Metropolis test run
outside the graph.
"""

import numpy as np

from ._lib import lib, check
from .kernel_limits import MAXBANDS, MAXSOURCES, NPARAMS
from .batch import BatchEvaluator
from .device_scene import DeviceRecords

__all__ = ["HMCEnsemble"]


def _torch():
    import torch
    return torch


class HMCEnsemble:
    """HMC chains on device-resident patches.

    Parameters
    ----------
    patches : list of HIPPatchMixin patches (sent to the device)
    scenes : list of Scenes (their sources' amplitude models give the radii)
    chains : int or list, chains per patch
    bounds : list of (lower, upper) flat parameter bounds per patch
    """

    def __init__(self, patches, scenes, chains, bounds, config=None, persistent=True):
        torch = _torch()
        self.batch = BatchEvaluator(patches, chains=chains, config=config,
                                    persistent=persistent)
        b = self.batch
        dev = b.device
        self.device = dev
        self.n_slots = b.n_slots
        self.S = max(int(p.n_sources) for p in patches)
        self.NB = max(int(p.n_bands) for p in patches)
        self.K = self.NB + 6
        n = self.n_slots
        lo = np.zeros((n, self.S, self.K))
        hi = np.ones((n, self.S, self.K))
        mask = np.zeros((n, self.S, self.K), dtype=bool)
        for s, (pi, _) in enumerate(b.slot_map):
            patch = patches[pi]
            ns, nb = int(patch.n_sources), int(patch.n_bands)
            l, h = (np.asarray(x, dtype=np.float64).reshape(ns, nb + 6) for x in bounds[pi])
            lo[s, :ns, :nb], hi[s, :ns, :nb] = l[:, :nb], h[:, :nb]
            lo[s, :ns, self.NB:], hi[s, :ns, self.NB:] = l[:, nb:], h[:, nb:]
            mask[s, :ns, :nb] = True
            mask[s, :ns, self.NB:] = True
        f64 = dict(dtype=torch.float64, device=dev)
        self.lo = torch.tensor(lo, **f64)
        self.hi = torch.tensor(hi, **f64)
        self.mask = torch.tensor(mask, device=dev)
        self.fmask = self.mask.to(torch.float64)
        self.records = DeviceRecords([scenes[pi] for pi, _ in b.slot_map],
                                     [int(patches[pi].n_bands) for pi, _ in b.slot_map],
                                     b.proposals, self.S, self.NB)
        self.z = torch.zeros((n, self.S, self.K), **f64)
        self.p = torch.zeros_like(self.z)
        self.g = torch.zeros_like(self.z)
        self.U = torch.zeros(n, **f64)
        self.U_start = torch.zeros(n, **f64)    # U at a trajectory's start point
        self.eps = torch.full((n, 1, 1), 0.01, **f64)
        self.resp = b.resp.view(n, MAXBANDS, MAXSOURCES, NPARAMS)
        self.chi2 = b.chi2.view(n, MAXBANDS)
        self.graph = None
        self.graph_steps = None

    # ------------------------------------------------------------------
    def set_theta(self, thetas):
        """Initial positions from flat parameter vectors (one per slot, in
        slot order), mapped through the bounded transform."""
        torch = _torch()
        z = np.zeros((self.n_slots, self.S, self.K))
        lo, hi = self.lo.cpu().numpy(), self.hi.cpu().numpy()
        for s, (pi, _) in enumerate(self.batch.slot_map):
            patch = self.batch.patches[pi]
            ns, nb = int(patch.n_sources), int(patch.n_bands)
            t = np.asarray(thetas[s], dtype=np.float64).reshape(ns, nb + 6)
            full = np.zeros((self.S, self.K))
            full[:ns, :nb], full[:ns, self.NB:] = t[:, :nb], t[:, nb:]
            y = (full - lo[s]) / (hi[s] - lo[s])
            y = np.clip(y, 1e-12, 1 - 1e-12)
            z[s] = np.where(self.mask[s].cpu().numpy(), -np.log(1.0 / y - 1.0), 0.0)
        self.z.copy_(torch.from_numpy(z))

    def theta(self):
        """Current positions in the constrained space, flat per slot."""
        th = (self.lo + torch_sigmoid(self.z) * (self.hi - self.lo)).cpu().numpy()
        out = []
        for s, (pi, _) in enumerate(self.batch.slot_map):
            patch = self.batch.patches[pi]
            ns, nb = int(patch.n_sources), int(patch.n_bands)
            out.append(np.concatenate([th[s, :ns, :nb], th[s, :ns, self.NB:]], axis=1).reshape(-1))
        return out

    # ------------------------------------------------------------------
    def _evaluate_and_update(self, kick=0.0, drift=False, move=False, U=None):
        """Records of the current z (bounded transform on the device), one
        batched chi2 launch, then fpx_leapfrog_update: U(z) and dU/dz into
        ``U`` (default self.U) / self.g and, with ``move``, p -= kick eps g
        and (with ``drift``) z += eps p.  Three launches plus the batch's
        two."""
        U = self.U if U is None else U
        self.records.write(self.z, lower=self.lo, upper=self.hi)
        self.batch.launch()
        check(lib().fpx_leapfrog_update(
            self.lo.data_ptr(), self.hi.data_ptr(), self.fmask.data_ptr(), self.eps.data_ptr(),
            self.resp.data_ptr(), self.chi2.data_ptr(), self.n_slots, self.S, self.NB,
            float(kick), int(drift), self.z.data_ptr(), self.p.data_ptr() if move else None,
            U.data_ptr(), self.g.data_ptr(),
            _torch().cuda.current_stream(self.device).cuda_stream), "fpx_leapfrog_update")

    def _potential_and_grad(self):
        """U(z) and dU/dz for the current z into self.U / self.g (device)."""
        self._evaluate_and_update()

    def _trajectory(self, n_steps):
        # potential at the start (into U_start), half kick, drift; then per
        # step the potential and a full kick + drift, the last one a half
        # kick: U then holds the potential at the end point
        if n_steps < 1:
            raise ValueError("a trajectory has at least one leapfrog step")
        self._evaluate_and_update(0.5, drift=True, move=True, U=self.U_start)
        for i in range(n_steps):
            last = i == n_steps - 1
            self._evaluate_and_update(0.5 if last else 1.0, drift=not last, move=True)

    def capture(self, n_steps):
        """Record one n_steps leapfrog trajectory of every chain in a graph."""
        torch = _torch()
        s = torch.cuda.Stream(device=self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        z0, p0 = self.z.clone(), self.p.clone()
        with torch.cuda.stream(s):
            self._trajectory(n_steps)          # warm-up outside the graph
        torch.cuda.current_stream(self.device).wait_stream(s)
        self.z.copy_(z0)
        self.p.copy_(p0)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._trajectory(n_steps)
        self.graph_steps = n_steps
        self.z.copy_(z0)
        self.p.copy_(p0)

    def trajectory(self, n_steps, use_graph=True):
        """Advance every chain by one leapfrog trajectory (no Metropolis)."""
        if use_graph:
            if self.graph is None or self.graph_steps != n_steps:
                self.capture(n_steps)
            self.graph.replay()
        else:
            self._trajectory(n_steps)

    def potential(self):
        """U(z) of every chain at the current position (device launch)."""
        self._potential_and_grad()
        return self.U.clone(), self.g.clone()

    def hmc_step(self, n_steps, generator=None, use_graph=True):
        """One HMC transition of every chain: fresh momenta, a leapfrog
        trajectory, Metropolis accept/reject (fpx_hmc_metropolis: the
        energies at the trajectory's two end points, rejected chains back to
        their start).  Returns acceptance flags (bool device tensor)."""
        torch = _torch()
        p0 = torch.randn(self.p.shape, generator=generator, dtype=torch.float64,
                         device=self.device)
        p0.mul_(self.fmask)
        z0 = self.z.clone()
        self.p.copy_(p0)
        self.trajectory(n_steps, use_graph=use_graph)
        u = torch.rand(self.n_slots, generator=generator, dtype=torch.float64, device=self.device)
        accept = torch.empty(self.n_slots, dtype=torch.int32, device=self.device)
        check(lib().fpx_hmc_metropolis(
            self.U_start.data_ptr(), self.U.data_ptr(), p0.data_ptr(), self.p.data_ptr(),
            u.data_ptr(), z0.data_ptr(), self.n_slots, self.S * self.K, self.z.data_ptr(),
            accept.data_ptr(), torch.cuda.current_stream(self.device).cuda_stream),
            "fpx_hmc_metropolis")
        return accept.bool()


    # ------------------------------------------------------------------
    def _energy_change(self, z0, p0, n_steps):
        """H1 - H0 of an eager n_steps trajectory of every chain from (z0,
        p0), H = U + |p|^2 / 2 (a non-finite energy gives +inf); z and p are
        left at the trajectory's end."""
        torch = _torch()
        self.z.copy_(z0)
        self.p.copy_(p0)
        self.trajectory(n_steps, use_graph=False)
        h0 = self.U_start + 0.5 * (p0 * p0).sum(dim=(1, 2))
        h1 = self.U + 0.5 * (self.p * self.p).sum(dim=(1, 2))
        dh = h1 - h0
        return torch.where(torch.isfinite(dh), dh, torch.full_like(dh, float("inf")))

    def find_step_size(self, generator=None, max_iter=20):
        """A starting step size per chain by the heuristic of Hoffman & Gelman
        (2014, Algorithm 4), for every chain at once: from the current
        position and fresh momenta, double (or halve) eps until the
        acceptance probability of one leapfrog step crosses 1/2, each chain
        stopping at its own crossing.  The position is left unchanged.
        Returns eps (device, per chain)."""
        import math
        torch = _torch()
        z0 = self.z.clone()
        p0 = torch.randn(self.p.shape, generator=generator, dtype=torch.float64,
                         device=self.device) * self.fmask
        half = math.log(0.5)
        logacc = -self._energy_change(z0, p0, 1)
        direction = torch.where(logacc > half, 1.0, -1.0).to(torch.float64)
        active = torch.ones(self.n_slots, dtype=torch.bool, device=self.device)
        for _ in range(max_iter):
            factor = torch.where(active, 2.0 ** direction, torch.ones_like(direction))
            self.eps.mul_(factor[:, None, None])
            logacc = -self._energy_change(z0, p0, 1)
            active &= direction * logacc > direction * half
            if not bool(active.any()):
                break
        self.z.copy_(z0)
        self.p.zero_()
        return self.eps.view(-1).clone()

    def adapt_step_size(self, n_rounds, n_steps, generator=None, grow=1.15, shrink=0.6,
                        use_graph=False):
        """n_rounds HMC transitions of n_steps leapfrog steps, each chain's
        eps multiplied by ``grow`` after an accepted transition and by
        ``shrink`` after a rejected one.  Returns the last round's
        acceptance flags."""
        torch = _torch()
        g = torch.tensor(grow, dtype=torch.float64, device=self.device)
        k = torch.tensor(shrink, dtype=torch.float64, device=self.device)
        acc = None
        for _ in range(n_rounds):
            acc = self.hmc_step(n_steps, generator=generator, use_graph=use_graph)
            self.eps.mul_(torch.where(acc, g, k)[:, None, None])
        return acc


def torch_sigmoid(x):
    return _torch().sigmoid(x)
