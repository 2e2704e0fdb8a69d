"""torch.autograd access to the patch model.

``PatchResidual`` maps the scene parameter vector theta
[flux_1..flux_B, ra, dec, q, pa, sersic, rhalf] x n_sources (the order of
forcepho's Scene, forcepho/sources.py:46-97) to the residual image
data - model(theta) over the packed pixels.  Its backward pass is the
kernel's gradient pass alone (FpxSlot.pixel_weight mode): given dL/dresidual
it recomputes every ImageGaussian and its moments from the proposal and
returns dL/dtheta, never storing per-pixel intermediates.  The theano Ops
of forcepho/model.py:515-551 played the same role for pymc3.

``PatchModel.chi2(theta)`` is the reference's chi2 convention
sum (r ierr)^2 - (d ierr)^2 (compute_gaussians_kernel.cc:286-293), so
``-0.5 * chi2`` is forcepho's log-likelihood and its autograd gradient
matches FastPosterior.lnprob_grad in the untransformed space.  It takes the
value and the gradient from one chi2-mode launch, as the reference's
FastPosterior.evaluate does (model.py:302-340), and with device-built records replays
one recorded graph per call; ``chi2_from_residual`` builds the same chi2 in
torch on top of the residual image.

When the scene's amplitude model can be evaluated on the device (the
analytic LogNormalMixtureAmplitudes, or forcepho's spline tables shared by
every source: device_scene.DeviceRecords) the Source records are built on the
device from theta and the gradient is stacked there too: forward and backward
then make no host round trip.  Other amplitude models build the records on
the host (Galaxy.proposal).
"""

import numpy as np

from .model import stack_grad
from .device_scene import DeviceRecords, device_amplitude_models_ok, stack_grad_device
from .kernel_limits import MAXBANDS, MAXSOURCES

__all__ = ["PatchModel", "PatchResidual"]


def _torch():
    import torch
    return torch


class PatchModel:
    """A device-resident patch, its scene and a proposer, for autograd use.

    Parameters
    ----------
    patch : HIPPatchMixin patch (packed; sent to the device here with a
            residual buffer)
    scene : Scene whose active sources map onto the patch's sources
    proposer : HIPProposer (created if None)
    use_graph : with device-built records, each device call (chi2 and its
                gradient, the residual image, the backward pass) replays a
                graph recorded on its first use: record build, kernel launch
                and gradient stacking in one submission
    """

    def __init__(self, patch, scene, proposer=None, use_graph=True):
        torch = _torch()
        from .proposal import HIPProposer
        self.patch, self.scene = patch, scene
        patch.return_residual = True
        if getattr(patch, "device_patch", None) is None or "residual" not in getattr(
                patch, "device_ptrs", {}):
            patch.send_to_device()
        self.proposer = proposer or HIPProposer(patch=patch)
        dev = self.proposer.device
        self.data = torch.from_numpy(np.asarray(patch.data, np.float64)).to(dev)
        self.ierr = torch.from_numpy(np.asarray(patch.ierr, np.float64)).to(dev)
        self.n_sources, self.n_bands = int(patch.n_sources), int(patch.n_bands)
        self.ndim = self.n_sources * (self.n_bands + 6)
        self._data_chi2 = ((self.data * self.ierr) ** 2).sum()
        self.records = None
        if device_amplitude_models_ok(scene):
            # the launches read this buffer in place (their device slots are
            # then built once)
            buf = torch.zeros(MAXSOURCES * 264, dtype=torch.uint8, device=dev)
            self.records = DeviceRecords([scene], [self.n_bands], buf, self.n_sources,
                                         self.n_bands)
        self.use_graph = bool(use_graph) and self.records is not None
        self._graphs = {}

    def device_proposal(self, theta, copy=True):
        """Source records of theta (any device, float64 [ndim]) as a uint8
        device tensor: a copy owned by the caller, or (copy=False) the
        builder's own buffer, overwritten by the next call."""
        torch = _torch()
        t = theta.detach().to(self.records.records.device, dtype=torch.float64)
        self.records.write(t.reshape(1, self.n_sources, self.n_bands + 6).contiguous())
        return self.records.records.clone() if copy else self.records.records

    # -- device calls (records built on the device) ---------------------------

    def _stacked(self):
        """dchi2/dtheta [ndim] and the chi2 summed over bands (0-d) of the
        proposer's last launch (one fpx_stack_gradients launch)."""
        p = self.proposer
        g, c = stack_grad_device(p._resp.view(1, MAXBANDS, MAXSOURCES, 7), self.n_sources,
                                 self.n_bands, chi2=p._chi2.view(1, MAXBANDS))
        return g.reshape(-1), c[0]

    def _chi2_body(self, theta):
        self.proposer.launch(self.device_proposal(theta, copy=False), self.patch)
        g, c = self._stacked()
        return c, g

    def _residual_body(self, theta):
        self.proposer.launch(self.device_proposal(theta, copy=False), self.patch,
                             model_only=True)
        return (self.patch.device_ptrs["residual"].view(_torch().float32)[:self.patch.npix],)

    def _vjp_body(self, theta, weight):
        # residual = data - model: d(sum_pix w r)/dtheta = -sum_pix w dmodel/dtheta
        self.proposer.launch(self.device_proposal(theta, copy=False), self.patch,
                             pixel_weight=weight)
        return (self._stacked()[0].neg(),)

    def _graph_state(self):
        """What a recorded graph depends on: the device patch struct, the
        residual buffer (``swap_on_device`` exchanges it with the data), the
        tile table and the kernel configuration."""
        p = self.proposer
        tiles, _, split_buf = p.tile_table(self.patch)
        return (self.patch.device_patch.data_ptr(), self.patch.device_ptrs["residual"].data_ptr(),
                tiles.data_ptr(), 0 if split_buf is None else split_buf.data_ptr(), p.config,
                p.persistent)

    def _call(self, body, *inputs):
        """Run ``body`` on device copies of ``inputs``: directly, or by
        replaying its recorded graph.  Returns the body's output tensors
        (a graph's are overwritten by its next replay)."""
        torch = _torch()
        dev = self.proposer.device
        if not self.use_graph:
            return body(*(x.detach().to(dev) for x in inputs))
        key = self._graph_state()
        rec = self._graphs.get(body.__name__)
        if rec is not None and rec["key"] == key:
            for s, x in zip(rec["inputs"], inputs):
                s.copy_(x.detach())
        else:
            static = [x.detach().to(dev).clone() for x in inputs]
            side = torch.cuda.Stream(device=dev)
            side.wait_stream(torch.cuda.current_stream(dev))
            with torch.cuda.stream(side):
                body(*static)               # builds the device slot outside the graph
            torch.cuda.current_stream(dev).wait_stream(side)
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph):
                out = body(*static)
            # the buffers the graph reads by address stay referenced with it
            refs = (self.proposer._slot, *self.proposer.tile_table(self.patch))
            rec = dict(key=key, graph=graph, inputs=static, out=out, refs=refs)
            self._graphs[body.__name__] = rec
        rec["graph"].replay()
        return rec["out"]

    def chi2_and_grad(self, theta):
        """chi2 and dchi2/dtheta of theta (float64 tensor [ndim]) as new
        float64 device tensors, from one chi2-mode launch; no host
        synchronisation."""
        value, grad = self._call(self._chi2_body, theta.to(_torch().float64))
        return value.clone(), grad.clone()

    def residual_device(self, theta):
        """data - model(theta) over the packed pixels (float32 device tensor
        [npix], a copy) from one model-only launch."""
        return self._call(self._residual_body, theta.to(_torch().float64))[0].clone()

    def vjp_theta(self, theta, weight):
        """sum_pix weight[pix] * d residual[pix] / d theta (float64 device
        tensor [ndim]) from one gradient-pass launch; weight: float32
        [npix] device tensor."""
        torch = _torch()
        w = weight.to(torch.float32).contiguous()
        return self._call(self._vjp_body, theta.to(torch.float64), w)[0].clone()

    # -- host records (any amplitude model) -----------------------------------

    def proposal(self, theta):
        self.scene.set_all_source_params(np.asarray(theta, dtype=np.float64))
        return self.scene.get_proposal()

    def forward_residual(self, proposal, outputs=True):
        """Launch the chi2 kernel for ``proposal``; returns (residual float32
        device tensor [npix] (a copy), chi2 per band, dchi2 [n_bands, n_src, 7]),
        or, without ``outputs``, the residual alone from a model-only launch
        (no host synchronisation)."""
        torch = _torch()
        p = self.proposer
        p.launch(proposal, self.patch, model_only=not outputs)
        resid = self.patch.device_ptrs["residual"].view(torch.float32)[:self.patch.npix].clone()
        if not outputs:
            return resid
        nb = self.n_bands
        chi2 = p._chi2[:nb].cpu().numpy()
        resp = p._resp[:nb * 210].view(nb, 30, 7)[:, :self.n_sources].cpu().numpy()
        return resid, chi2, resp.astype(np.float64)

    # -- differentiable entry points ------------------------------------------

    def residual(self, theta):
        """Differentiable residual image; theta: float64 tensor [ndim]."""
        return PatchResidual.apply(theta, self)

    def chi2(self, theta):
        """Differentiable sum over pixels of (r ierr)^2 - (d ierr)^2 (float64).
        One launch of the chi2 kernel computes the value and its gradient
        together (as ``evaluate_proposal`` does); backward scales the saved
        gradient.  Equal to ``chi2_from_residual``."""
        return PatchChi2.apply(theta, self)

    def chi2_from_residual(self, theta):
        """The same chi2 built by torch from the differentiable residual
        image (forward: model pass; backward: the kernel's gradient pass)."""
        r = self.residual(theta).double()
        return ((r * self.ierr) ** 2).sum() - self._data_chi2


class PatchResidual:
    """Namespace holding the autograd.Function (created lazily so that this
    module imports without torch)."""

    _fn = None

    @classmethod
    def apply(cls, theta, model):
        if cls._fn is None:
            cls._fn = _make_function()
        return cls._fn.apply(theta, model)


class PatchChi2:
    """Namespace of the chi2 autograd.Function: value and gradient from one
    kernel launch."""

    _fn = None

    @classmethod
    def apply(cls, theta, model):
        if cls._fn is None:
            cls._fn = _make_chi2_function()
        return cls._fn.apply(theta, model)


def _make_chi2_function():
    torch = _torch()

    class _Chi2(torch.autograd.Function):

        @staticmethod
        def forward(ctx, theta, model):
            p = model.proposer
            nb, ns = model.n_bands, model.n_sources
            if model.records is not None:
                value, g = model.chi2_and_grad(theta)
            else:
                chi2, d = p.evaluate_proposal(model.proposal(theta.detach().cpu().numpy()),
                                              patch=model.patch)[:2]
                g = torch.from_numpy(stack_grad(d, ns, nb)).to(p.device)
                value = torch.tensor(float(np.sum(chi2)), dtype=torch.float64, device=p.device)
            ctx.save_for_backward(g.to(dtype=theta.dtype, device=theta.device))
            return value.to(dtype=torch.float64, device=theta.device)

        @staticmethod
        def backward(ctx, grad_out):
            (g,) = ctx.saved_tensors
            return grad_out.to(g.dtype) * g, None

    return _Chi2


def _make_function():
    torch = _torch()

    class _Residual(torch.autograd.Function):

        @staticmethod
        def forward(ctx, theta, model):
            ctx.model = model
            ctx.theta_meta = (theta.dtype, theta.device)
            if model.records is not None:
                ctx.save_for_backward(theta)
                return model.residual_device(theta)
            proposal = model.proposal(theta.detach().cpu().numpy())
            ctx.proposal = proposal
            return model.forward_residual(proposal, outputs=False)

        @staticmethod
        def backward(ctx, grad_resid):
            model = ctx.model
            dtype, device = ctx.theta_meta
            if model.records is not None:
                (theta,) = ctx.saved_tensors
                g = model.vjp_theta(theta, grad_resid)
                return g.to(dtype=dtype, device=device), None
            # residual = data - model, so dL/dtheta = -sum_pix g * dmodel/dtheta
            pw = (-grad_resid).to(torch.float32).contiguous()
            d = model.proposer.vjp(ctx.proposal, pw, model.patch)
            g = stack_grad(d, model.n_sources, model.n_bands)
            return torch.from_numpy(g).to(dtype=dtype, device=device), None

    return _Residual
