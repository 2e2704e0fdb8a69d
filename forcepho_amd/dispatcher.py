"""Fitting many regions of a SuperScene at once on one MI355X.

forcepho's parent process (dispatcher.py:156-300) checks a region out of the
SuperScene for every idle MPI child; each child builds its patch, subtracts
the fixed sources and samples the active ones alone on its GPU, and the
parent checks the result back in.  A patch of up to 20 sources keeps a few
percent of an MI355X busy, so here one process per GPU plays both roles:
``PatchDispatcher`` checks out up to ``max_regions`` disjoint regions,
builds a device patch for each (``make_patch(region)``), subtracts its fixed
sources (``prepare_model``, device_patch.py:28-72), samples the active
sources of *all* the regions with one ``HMCEnsemble`` -- every leapfrog step
is one batched launch over every chain of every region -- and checks each
region back in with its posterior mean and the iteration count.  On a node
with several GPUs, one dispatcher process per GPU would share the SuperScene
through its parent (checkout/checkin are its only two calls), as the MPI
children of the reference do.

The sampler is fixed-length HMC with a per-chain step size adapted during
warm-up (grown after an accepted trajectory, shrunk after a rejected one),
not the reference's NUTS (littlemcmc, fitting.py:254-341).
"""

import numpy as np

from .superscene import bounds_vectors
from .sources import Galaxy

__all__ = ["PatchDispatcher", "RegionTask"]


class RegionTask:
    """A checked-out region and its device patch, scene and prior box."""

    def __init__(self, region, active, fixed, patch, scene, bounds, q, task_id):
        self.region, self.active, self.fixed = region, active, fixed
        self.patch, self.scene, self.bounds, self.q = patch, scene, bounds, q
        self.task_id = task_id


class PatchDispatcher:
    """Check out, sample and check in many regions at once on one GPU.

    Parameters
    ----------
    sceneDB : SuperScene or LinkedSuperScene
    make_patch : callable(region) -> device patch with pixels packed
        (e.g. a FITSDevicePatch after ``build_patch(region, None, bands)``)
    max_regions : regions sampled together
    chains : chains per region
    n_leapfrog, n_warmup, n_samples, step_size : HMC settings
    n_tries : checkout attempts per round before giving up on more regions
    maxactive : block size for subtracting fixed sources
    """

    def __init__(self, sceneDB, make_patch, max_regions=16, chains=2, n_leapfrog=16,
                 n_warmup=20, n_samples=20, step_size=0.05, n_tries=200, maxactive=15,
                 seed=None, use_graph=True):
        self.sceneDB = sceneDB
        self.make_patch = make_patch
        self.max_regions = int(max_regions)
        self.chains = int(chains)
        self.n_leapfrog, self.n_warmup, self.n_samples = n_leapfrog, n_warmup, n_samples
        self.step_size = step_size
        self.n_tries = n_tries
        self.maxactive = maxactive
        self.seed = seed
        self.use_graph = use_graph
        self.n_tasks = 0
        self.history = []

    def checkout(self):
        """Up to max_regions disjoint regions [(region, active, fixed)]."""
        out = []
        for _ in range(self.n_tries):
            if len(out) >= self.max_regions or not self.sceneDB.undone:
                break
            region, active, fixed = self.sceneDB.checkout_region()
            if active is not None:
                out.append((region, active, fixed))
        return out

    def prepare(self, regions):
        """Device patches with fixed sources subtracted, scenes, prior boxes
        (bounds catalog rows relative to each patch's reference point)."""
        tasks = []
        for region, active, fixed in regions:
            patch = self.make_patch(region)
            model, q = patch.prepare_model(active=active, fixed=fixed, maxactive=self.maxactive)
            rows = self.sceneDB.bounds_catalog[active["source_index"]]
            lo, hi = bounds_vectors(rows, patch.bandlist,
                                    reference_coordinates=patch.patch_reference_coordinates)
            self.n_tasks += 1
            tasks.append(RegionTask(region, active, fixed, patch, model.scene, (lo, hi), q,
                                    self.n_tasks))
        return tasks

    def sample(self, tasks):
        """HMC over every chain of every task together; returns per task the
        posterior mean (flat Scene parameters) and the acceptance rate."""
        import torch
        from .ensemble import HMCEnsemble
        ens = HMCEnsemble([t.patch for t in tasks], [t.scene for t in tasks], self.chains,
                          [t.bounds for t in tasks])
        ens.set_theta([tasks[pi].q for pi, _ in ens.batch.slot_map])
        ens.eps.fill_(self.step_size)
        gen = torch.Generator(device=ens.device)
        gen.manual_seed(int(self.seed if self.seed is not None else np.random.randint(2 ** 31)))
        ens.adapt_step_size(self.n_warmup, self.n_leapfrog, generator=gen,
                            use_graph=self.use_graph)
        sums = [np.zeros_like(np.asarray(t.q, dtype=np.float64)) for t in tasks]
        n_acc = np.zeros(len(tasks))
        for _ in range(self.n_samples):
            acc = ens.hmc_step(self.n_leapfrog, generator=gen, use_graph=self.use_graph)
            acc = acc.cpu().numpy()
            for s, (pi, _), th in zip(range(ens.n_slots), ens.batch.slot_map, ens.theta()):
                sums[pi] += th
                n_acc[pi] += acc[s]
        norm = max(1, self.n_samples) * self.chains
        return [s / norm for s in sums], n_acc / norm

    def checkin(self, tasks, means):
        """Write each region's posterior mean back to the SuperScene."""
        for t, mean in zip(tasks, means):
            nb = len(t.patch.bandlist)
            cols = list(t.patch.bandlist) + list(Galaxy.SHAPE_COLS)
            m = np.asarray(mean).reshape(len(t.active), nb + 6)
            new = t.active.copy()
            for j, c in enumerate(cols):
                new[c] = m[:, j]
            ref = t.patch.patch_reference_coordinates
            new["ra"] += ref[0]
            new["dec"] += ref[1]
            self.sceneDB.checkin_region(new, t.fixed, niter=self.n_samples, taskID=t.task_id)

    def round(self):
        """One checkout / sample / checkin round; the number of regions."""
        regions = self.checkout()
        if not regions:
            return 0
        try:
            tasks = self.prepare(regions)
            means, acc = self.sample(tasks)
        except BaseException:
            for region, active, fixed in regions:      # release on failure
                self.sceneDB.checkin_region(active, fixed, niter=0)
            raise
        self.checkin(tasks, means)
        self.history.append(dict(n_regions=len(tasks), acceptance=acc.tolist(),
                                 n_active=[len(t.active) for t in tasks]))
        return len(tasks)

    def run(self, max_rounds=100):
        """Rounds until every source has its target iterations (or
        max_rounds, or no region can be checked out)."""
        rounds = 0
        while self.sceneDB.undone and rounds < max_rounds:
            if self.round() == 0:
                break
            rounds += 1
        return rounds
