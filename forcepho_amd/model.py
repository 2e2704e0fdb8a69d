"""Posterior objects: ln-probability and its gradient of a parameter vector,
evaluated through a proposer.

Mirrors forcepho/model.py: ``Posterior`` caching and ``check_grad``
(:42-257), ``FastPosterior.evaluate`` / ``stack_grad`` / ``residuals``
(:260-405) and ``Transform`` / ``BoundedTransform`` (:572-637):
    ll      = -1/2 sum_b chi2_b                      (float64)
    ll_grad = -1/2 stack_grad(dchi2/dparam)
    lnp     = ll + lnprior (+ ln|dq/dz| when sampling)
    grad    = (ll_grad + lnprior_grad) * dq/dz (+ d ln|dq/dz| / dz)
"""

import numpy as np

from .kernel_limits import NPARAMS

__all__ = ["Posterior", "FastPosterior", "GPUPosterior", "Transform", "BoundedTransform",
           "sigmoid", "sigmoid_grad", "logit", "stack_grad"]


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def sigmoid_grad(z):
    return (0.5 / np.cosh(0.5 * z)) ** 2


def logit(y):
    return -np.log(1.0 / y - 1.0)


def stack_grad(chi2_derivs, n_sources, n_bands):
    """[n_bands, n_sources, NPARAMS] -> flat [n_sources * (n_bands + 6)] in
    parameter-vector order: per source the band fluxes (each from its own
    band) then ra, dec, q, pa, sersic, rh summed over bands."""
    grads = np.zeros([n_sources, n_bands + (NPARAMS - 1)], dtype=np.float64)
    for band, derivs in enumerate(chi2_derivs):
        grads[:, n_bands:] += derivs[:, 1:]
        grads[:, band] += derivs[:, 0]
    return grads.reshape(-1)


class Transform:
    """Identity transform z -> q."""

    def __init__(self, ndim):
        self.ndim = ndim

    def transform(self, z):
        return np.array(z)

    def jacobian(self, z):
        return np.ones(len(z))

    def lndetjac(self, z):
        return np.log(np.abs(np.prod(self.jacobian(z))))

    def lndetjac_grad(self, z):
        return np.zeros(len(z))

    def inverse_transform(self, Theta):
        return np.array(Theta)


class BoundedTransform(Transform):
    """q = lower + sigmoid(z) (upper - lower)."""

    def __init__(self, lower, upper):
        self.lower = np.atleast_1d(np.asarray(lower, dtype=np.float64))
        self.upper = np.atleast_1d(np.asarray(upper, dtype=np.float64))
        assert len(self.upper) == len(self.lower)
        self.ndim = len(self.upper)

    @property
    def range(self):
        return self.upper - self.lower

    def transform(self, z):
        return self.lower + sigmoid(z) * self.range

    def jacobian(self, z):
        return self.range * sigmoid_grad(z)

    def lndetjac(self, z):
        return np.sum(np.log(self.jacobian(z)))

    def lndetjac_grad(self, z):
        return 1.0 - 2.0 * sigmoid(z)

    def inverse_transform(self, Theta):
        return logit((np.asarray(Theta) - self.lower) / self.range)


class Posterior:
    """Caching of lnp and its gradient plus a finite-difference check."""

    def __init__(self, **kwargs):
        self.ncall = 0
        self.ndof = 1.0
        has_bounds = kwargs.get("lower") is not None and kwargs.get("upper") is not None
        if has_bounds == (kwargs.get("transform") is not None):
            raise ValueError('Must specify either "transform" or ("lower","upper")')
        transform = kwargs.pop("transform", None)
        if transform is not None:
            self.transform = transform
        else:
            self.transform = BoundedTransform(kwargs.pop("lower"), kwargs.pop("upper"))
        if kwargs:
            raise ValueError(f"Passed argument(s) {list(kwargs)} to Posterior that were not understood")
        self._z = None

    def _stale(self, z):
        return self._z is None or np.shape(z) != np.shape(self._z) or np.any(z != self._z)

    def evaluate(self, z):
        raise NotImplementedError

    def lnprob(self, z):
        if self._stale(z):
            self.evaluate(z)
        return self._lnp

    def lnprob_grad(self, z):
        if self._stale(z):
            self.evaluate(z)
        return self._lnp_grad

    def lnprob_and_grad(self, z):
        if self._stale(z):
            self.evaluate(z)
        return self._lnp, self._lnp_grad

    def nll(self, z):
        if self._stale(z):
            self.evaluate(z)
        return -self._lnp / self.ndof, -self._lnp_grad / self.ndof

    def nll_nograd(self, z):
        if self._stale(z):
            self.evaluate(z)
        return -self._lnp / self.ndof

    def lnprior(self, q):
        if getattr(self, "_lnpriorfn", None) is None:
            return 0.0, 0.0
        return self._lnpriorfn(q)

    def make_transform(self, z):
        if self.transform is not None:
            self._jacobian = self.transform.jacobian(z)
            self._lndetjac = self.transform.lndetjac(z)
            self._lndetjac_grad = self.transform.lndetjac_grad(z)
            return self.transform.transform(z)
        self._jacobian, self._lndetjac, self._lndetjac_grad = 1.0, 0.0, 0.0
        return np.array(z)

    def residuals(self, z):
        raise NotImplementedError

    def check_grad(self, z, delta=1e-5):
        """Analytic gradient and central finite differences at z
        (forcepho/model.py:225-257)."""
        z0 = np.array(z, dtype=np.float64)
        dlnp = np.array(self.lnprob_grad(z0), dtype=np.float64)
        delta = np.zeros_like(dlnp) + delta
        dlnp_num = np.zeros(len(z0), dtype=np.float64)
        for i, dp in enumerate(delta):
            theta = z0.copy()
            theta[i] -= dp
            lo = self.lnprob(theta.copy())
            theta[i] += 2 * dp
            hi = self.lnprob(theta.copy())
            dlnp_num[i] = (hi - lo) / (2 * dp)
        return dlnp, dlnp_num


class FastPosterior(Posterior):
    """Posterior evaluated by a device proposer on a packed patch.

    ``device_records``: build the Source records of each evaluation on the
    device (fpx_build_sources) instead of through ``Scene.get_proposal`` on
    the host, and stack the gradient there (fpx_stack_gradients); one copy
    of chi2 and the gradient comes back.  Default: on when the proposer is
    the HIP proposer and the scene's amplitude model can be evaluated on the
    device (log-normal, or one shared spline model).  Evaluations that return
    residual images (``patch.return_residual``) use the host records."""

    def __init__(self, proposer, patch, scene=None, lnprior=None, transform=None,
                 name="", print_interval=1000, verbose=False, debug=False,
                 logging=False, sampling=True, device_records=None, **kwargs):
        super().__init__(transform=transform, **kwargs)
        self.proposer = proposer
        self.patch = patch
        self.scene = scene if scene is not None else patch.scene
        self._lnpriorfn = lnprior
        self.sampling = sampling
        self.mhalf = np.array(-0.5, dtype=np.float64)
        self.debug, self.logging, self.verbose = debug, logging, verbose
        self.print_interval, self.name = print_interval, name
        if logging:
            self.pos_history, self.lnp_history, self.grad_history = [], [], []
        self.device_records = device_records
        self._dev = None

    def _use_device_records(self):
        if self.device_records is False or getattr(self.patch, "return_residual", False):
            return False
        from .proposal import HIPProposer
        if not isinstance(self.proposer, HIPProposer):
            return False
        if self.device_records is None:
            from .device_scene import device_amplitude_models_ok
            self.device_records = device_amplitude_models_ok(self.scene)
        return bool(self.device_records)

    def _device_chi2(self, q):
        """chi2 summed over bands and d chi2 / dq of scene parameters q from
        records built on the device: one H2D copy, record build, chi2
        launch, gradient stacking, one D2H copy."""
        import torch
        from .device_scene import DeviceRecords, stack_grad_device
        from .kernel_limits import MAXBANDS, MAXSOURCES
        p = self.proposer
        ns, nb = int(self.patch.n_sources), int(self.patch.n_bands)
        if self._dev is None:
            pin = dict(dtype=torch.float64, pin_memory=True)
            buf = torch.zeros(MAXSOURCES * 264, dtype=torch.uint8, device=p.device)
            self._dev = dict(
                records=DeviceRecords([self.scene], [nb], buf, ns, nb),
                theta_h=torch.zeros(ns * (nb + 6), **pin),
                theta_d=torch.zeros(ns * (nb + 6), dtype=torch.float64, device=p.device),
                out_d=torch.zeros(1 + ns * (nb + 6), dtype=torch.float64, device=p.device),
                out_h=torch.zeros(1 + ns * (nb + 6), **pin))
        d = self._dev
        stream = p._stream()
        d["theta_h"].numpy()[:] = q
        d["theta_d"].copy_(d["theta_h"], non_blocking=True)
        d["records"].write(d["theta_d"].view(1, ns, nb + 6))
        p.launch(d["records"].records, self.patch, stream=stream)
        g, c = stack_grad_device(p._resp.view(1, MAXBANDS, MAXSOURCES, NPARAMS), ns, nb,
                                 chi2=p._chi2.view(1, MAXBANDS))
        d["out_d"][:1].copy_(c)
        d["out_d"][1:].copy_(g.view(-1))
        d["out_h"].copy_(d["out_d"], non_blocking=True)
        stream.synchronize()
        out = d["out_h"].numpy()
        return float(out[0]), out[1:].copy()

    def evaluate(self, z):
        q = self.make_transform(z)
        self.scene.set_all_source_params(q)
        lpr, lpr_grad = self.lnprior(q)
        if self._use_device_records():
            chi2_sum, dchi2 = self._device_chi2(np.asarray(q, dtype=np.float64))
            ll = self.mhalf * np.array(chi2_sum, dtype=np.float64)
            ll_grad = self.mhalf * dchi2
        else:
            proposal = self.scene.get_proposal()
            ret = self.proposer.evaluate_proposal(proposal, patch=self.patch)
            if len(ret) == 3:
                chi2, chi2_derivs, self._residuals = ret
            else:
                chi2, chi2_derivs = ret
            ll = self.mhalf * np.array(chi2.astype(np.float64).sum(), dtype=np.float64)
            ll_grad = self.mhalf * self.stack_grad(chi2_derivs)
        self.ncall += 1
        if self.logging:
            self.pos_history.append(z)
            self.lnp_history.append(ll)
            self.grad_history.append(ll_grad)
        self._lnp = ll + lpr
        self._lnp_grad = (ll_grad + lpr_grad) * self._jacobian
        self._q = q
        self._z = np.array(z, copy=True)
        if self.sampling:
            self._lnp = self._lnp + self._lndetjac
            self._lnp_grad = self._lnp_grad + self._lndetjac_grad
        if self.verbose and self.ncall % self.print_interval == 0:
            print(f"-------\n {self.name} @ {self.ncall}\nq: {q}\nlnlike: {ll}")

    def stack_grad(self, chi2_derivs):
        return stack_grad(chi2_derivs, int(self.patch.n_sources), int(self.patch.n_bands))

    def residuals(self, q, unpack=True):
        """Residual images at scene parameters q (constrained space)."""
        orig = self.patch.return_residual
        self.patch.return_residual = True
        if "residual" not in getattr(self.patch, "device_ptrs", {}):
            import torch
            self.patch.device_ptrs["residual"] = torch.zeros(
                self.patch.xpix.nbytes, dtype=torch.uint8, device=self.patch._dev())
        self.patch.send_patchstruct_to_device()
        self.scene.set_all_source_params(q)
        proposal = self.scene.get_proposal()
        _, _, self._residuals = self.proposer.evaluate_proposal(proposal, patch=self.patch,
                                                                unpack=unpack)
        self.patch.return_residual = orig
        self.patch.send_patchstruct_to_device()
        return self._residuals


class GPUPosterior(FastPosterior):
    """forcepho's name for FastPosterior, kept for backwards compatibility
    (forcepho/model.py:408-411)."""
