"""Many evaluations in one kernel launch: uneven patches and many chains.

A ``BatchEvaluator`` owns one FpxSlot per (patch, chain) pair, the slots'
proposals, outputs and accumulators, all in device memory, and evaluates all
of them with one ``fpx_evaluate_batch`` launch.  The reference has no
counterpart (forcepho evaluates one patch per call, forcepho/proposal.py:
96-152, and spreads patches over MPI ranks, forcepho/dispatcher.py:156-300);
here a resident grid pulls (slot, tile) work items from a queue counter.

Load balance: slots are ordered by decreasing estimated cost per tile
(ImageGaussians per exposure), so the queue hands out the heaviest tiles
first and the cheap ones fill the tail.  ``autotune`` times every kernel
configuration (tile shape) on the batch and keeps the fastest.
"""

import numpy as np

from .kernel_limits import MAXBANDS, MAXSOURCES, NPARAMS
from ._lib import lib, check, FpxError, SLOT_DTYPE
from .proposal import source_struct_dtype

__all__ = ["BatchEvaluator"]

_SRC_BYTES = source_struct_dtype.itemsize
_RESP = MAXSOURCES * NPARAMS


def _torch():
    import torch
    return torch


class BatchEvaluator:
    """Evaluate proposals on many device-resident patches in one launch.

    Parameters
    ----------
    patches : list of HIPPatchMixin patches, already ``send_to_device``-ed
    chains : int or list of int, proposals (chains) per patch
    config : kernel configuration; None = DEFAULT (use ``autotune`` to pick)
    persistent : resident grid with a work queue (True) or one workgroup per
        tile (False)
    """

    def __init__(self, patches, chains=1, config=None, persistent=True):
        from .proposal import DEFAULT_CONFIG
        torch = _torch()
        self.patches = list(patches)
        if np.isscalar(chains):
            chains = [int(chains)] * len(self.patches)
        self.chains = [int(c) for c in chains]
        # slot -> (patch index, chain index)
        self.slot_map = [(p, c) for p, n in enumerate(self.chains) for c in range(n)]
        self.n_slots = len(self.slot_map)
        self.persistent = bool(persistent)
        self.device = torch.device("cuda", torch.cuda.current_device())
        dev = self.device
        L = lib()
        self.n_acc = L.fpx_accum_doubles(MAXBANDS)
        n = self.n_slots
        self.accum = torch.zeros(n * self.n_acc, dtype=torch.float64, device=dev)
        # per-slot status words (FpxSlot.status): nonzero when the kernel left
        # a tile of the slot out (checked by results())
        self.status = torch.zeros(n, dtype=torch.int32, device=dev)
        self.queue = torch.zeros(2, dtype=torch.int32, device=dev)
        self.chi2 = torch.zeros(n * MAXBANDS, dtype=torch.float64, device=dev)
        self.resp = torch.zeros(n * MAXBANDS * _RESP, dtype=torch.float32, device=dev)
        self.proposals = torch.zeros(n * MAXSOURCES * _SRC_BYTES, dtype=torch.uint8, device=dev)
        for p in self.patches:
            p.check_limits()
        self.max_gauss = max(p.max_gauss() for p in self.patches)
        self.set_config(DEFAULT_CONFIG if config is None else config)

    # ------------------------------------------------------------------
    def set_config(self, config):
        """Select the kernel configuration and rebuild the slot table."""
        import ctypes
        torch = _torch()
        L = lib()
        ppt = ctypes.c_int32()
        check(L.fpx_config_info(int(config), None, None, ctypes.byref(ppt)), "fpx_config_info")
        self.config = int(config)
        self.pix_per_tile = ppt.value
        tiles = [p.device_tiles(self.pix_per_tile) for p in self.patches]
        # order slots by decreasing work per tile (heaviest tiles leave the queue first)
        cost = np.array([self.patches[p].max_gauss() for p, _ in self.slot_map], dtype=np.int64)
        self.order = np.argsort(-cost, kind="stable")
        slots = np.zeros(self.n_slots, dtype=SLOT_DTYPE)
        n_tiles = np.zeros(self.n_slots, dtype=np.int64)
        base = lambda t: t.data_ptr()
        for k, s in enumerate(self.order):
            p, c = self.slot_map[s]
            patch = self.patches[p]
            tens, nt = tiles[p]
            slots[k]["patch"] = patch.device_patch.data_ptr()
            slots[k]["proposal"] = base(self.proposals) + s * MAXSOURCES * _SRC_BYTES
            slots[k]["tiles"] = tens.data_ptr()
            slots[k]["pchi2"] = base(self.chi2) + s * MAXBANDS * 8
            slots[k]["pdchi2_dp"] = base(self.resp) + s * MAXBANDS * _RESP * 4
            slots[k]["accum"] = base(self.accum) + s * self.n_acc * 8
            slots[k]["status"] = base(self.status) + s * 4
            slots[k]["n_bands"] = int(patch.n_bands)
            slots[k]["n_tiles"] = nt
            n_tiles[k] = nt
        offsets = np.zeros(self.n_slots + 1, dtype=np.int64)
        offsets[1:] = np.cumsum(n_tiles)
        if offsets[-1] > np.iinfo(np.int32).max:
            raise ValueError("too many tiles in one launch")
        self.n_items = int(offsets[-1])
        self._tiles_keep = tiles
        self.slots_dev = torch.from_numpy(slots.view(np.uint8).reshape(-1).copy()).to(self.device)
        self.offsets_dev = torch.from_numpy(offsets.astype(np.int32)).to(self.device)

    def set_proposals(self, proposals):
        """proposals: list (one per slot, in slot order: patch-major, chain
        minor) of source_struct_dtype arrays, or a uint8 device tensor of
        n_slots * MAXSOURCES * 264 bytes."""
        torch = _torch()
        if isinstance(proposals, torch.Tensor):
            self.proposals.copy_(proposals, non_blocking=True)
            return
        host = np.zeros((self.n_slots, MAXSOURCES * _SRC_BYTES), dtype=np.uint8)
        for s, prop in enumerate(proposals):
            p, _ = self.slot_map[s]
            if len(prop) != int(self.patches[p].n_sources):
                raise ValueError(f"slot {s}: proposal has {len(prop)} sources")
            raw = np.ascontiguousarray(prop, dtype=source_struct_dtype).view(np.uint8)
            host[s, :raw.size] = raw.reshape(-1)
        self.proposals.copy_(torch.from_numpy(host.reshape(-1)), non_blocking=False)

    def launch(self, stream=None, outputs=True):
        """Queue the batched evaluation (asynchronous).  ``outputs=False``
        launches the tile kernel alone (fpx_evaluate_tiles): its sums stay in
        the slots' accumulators until ``finalize`` or a launch with
        outputs."""
        torch = _torch()
        s = stream or torch.cuda.current_stream(self.device)
        fn = lib().fpx_evaluate_batch if outputs else lib().fpx_evaluate_tiles
        check(fn(self.slots_dev.data_ptr(), self.n_slots, self.offsets_dev.data_ptr(),
                 self.n_items, self.max_gauss, self.config, int(self.persistent),
                 self.queue.data_ptr(), s.cuda_stream),
              "fpx_evaluate_batch" if outputs else "fpx_evaluate_tiles")

    def finalize(self, stream=None):
        """Write every slot's outputs from its accumulator and zero it
        (fpx_finalize)."""
        torch = _torch()
        s = stream or torch.cuda.current_stream(self.device)
        check(lib().fpx_finalize(self.slots_dev.data_ptr(), self.n_slots, s.cuda_stream),
              "fpx_finalize")

    def results(self):
        """Synchronise and return per-slot (chi2 [n_bands], derivs
        [n_bands, n_sources, 7]) in slot order.  Raises FpxError (after
        clearing the words) if the kernel left a tile of some slot out."""
        torch = _torch()
        torch.cuda.current_stream(self.device).synchronize()
        self.check_status()
        chi2 = self.chi2.view(self.n_slots, MAXBANDS).cpu().numpy()
        resp = self.resp.view(self.n_slots, MAXBANDS, MAXSOURCES, NPARAMS).cpu().numpy()
        out = []
        for s, (p, _) in enumerate(self.slot_map):
            nb, ns = int(self.patches[p].n_bands), int(self.patches[p].n_sources)
            out.append((chi2[s, :nb].copy(), resp[s, :nb, :ns].astype(np.float64)))
        return out

    def check_status(self):
        """Raise FpxError, naming the slots, if any slot's status word is set
        (the kernel left a tile of that slot out and wrote NaN chi2); the
        words are cleared first."""
        st = self.status.cpu().numpy()
        bad = np.flatnonzero(st)
        if len(bad):
            self.status.zero_()
            raise FpxError(f"the kernel left tiles out in slots {bad[:10].tolist()} "
                           f"(status bits {sorted(set(st[bad].tolist()))})")

    def count_pairs(self):
        """Per slot (slot order), the (pixel, ImageGaussian) pairs inside the
        cutoff dx^T F dx <= MAX_EXP_ARG for the current proposals
        (fpx_count_pairs)."""
        torch = _torch()
        out = torch.zeros(self.n_slots, dtype=torch.int64, device=self.device)
        s = torch.cuda.current_stream(self.device)
        check(lib().fpx_count_pairs(self.slots_dev.data_ptr(), self.n_slots,
                                    self.offsets_dev.data_ptr(), self.n_items, self.max_gauss,
                                    out.data_ptr(), None, s.cuda_stream), "fpx_count_pairs")
        counts = out.cpu().numpy()
        res = np.zeros(self.n_slots, dtype=np.int64)
        res[self.order] = counts     # slot table position k holds slot order[k]
        return res

    def n_evals(self):
        """Pixel x Gaussian pairs of one batched evaluation."""
        return sum(self.patches[p].n_evals() for p, _ in self.slot_map)

    # ------------------------------------------------------------------
    def autotune(self, repeats=3, configs=None, persistent=(True, False)):
        """Time each (config, persistent) pair and keep the fastest.
        Returns {(config, persistent): seconds}."""
        torch = _torch()
        L = lib()
        configs = range(L.fpx_n_configs()) if configs is None else configs
        timings = {}
        for cfg in configs:
            self.set_config(cfg)
            for pers in persistent:
                self.persistent = pers
                self.launch()
                torch.cuda.synchronize()
                a = torch.cuda.Event(enable_timing=True)
                b = torch.cuda.Event(enable_timing=True)
                a.record()
                for _ in range(repeats):
                    self.launch()
                b.record()
                torch.cuda.synchronize()
                timings[(cfg, pers)] = a.elapsed_time(b) * 1e-3 / repeats
        best = min(timings, key=timings.get)
        self.set_config(best[0])
        self.persistent = best[1]
        return timings
