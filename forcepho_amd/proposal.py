"""Host interface that evaluates a proposal on the MI355X.

``HIPProposer.evaluate_proposal`` keeps the contract of forcepho's
GPUProposer/CPUProposer.evaluate_proposal (forcepho/proposal.py:96-152,
180-226): given an array of ``source_struct_dtype`` records it returns

    chi2        float64 [n_bands]
    chi2_derivs [n_bands, n_sources, NPARAMS]
    residuals   only when ``patch.return_residual``: list of per-exposure
                padded arrays (``unpack=True``) or one flat array

A numpy proposal is evaluated by one native call on the current torch
stream, fpx_evaluate_host of one of the patch's bound evaluators: over the
split table ``balance`` planned for the patch (persistent launches), else on
a patch of at most HOST_PATH_TILES tiles over tiles the library plans (the
tiles read the proposal from, and leave their sums in, mapped host memory),
else over the patch's own (balanced) tile table.  One kernel in each case;
the call returns when the launch signals through mapped memory that the
results are there.  A device proposal is launched over the proposer's own
slot.
"""

import sys

import numpy as np

from .kernel_limits import MAXBANDS, MAXRADII, MAXSOURCES, NPARAMS
from ._lib import lib, check, FpxError, SLOT_DTYPE, SLOT_MODEL_ONLY, STATUS_SPLIT_WAIT
from . import tile_split

# device output block of a proposer: the slot's status word (uint32, padded
# to 8 bytes), pchi2 (double[MAXBANDS]), then the Responses
# (float[MAXBANDS][MAXSOURCES][NPARAMS]); one copy brings all three back
_STATUS_BYTES = 8
_CHI2_END = _STATUS_BYTES + MAXBANDS * 8
_OUT_BYTES = _CHI2_END + MAXBANDS * MAXSOURCES * NPARAMS * 4

__all__ = ["ProposerBase", "HIPProposer", "GPUProposer", "Proposer",
           "source_struct_dtype", "response_struct_dtype", "DEFAULT_CONFIG"]

# forcepho/proposal.py:229-244
source_float_dt = np.float32
source_struct_dtype = np.dtype([("ra", source_float_dt),
                                ("dec", source_float_dt),
                                ("q", source_float_dt),
                                ("pa", source_float_dt),
                                ("nsersic", source_float_dt),
                                ("rh", source_float_dt),
                                ("fluxes", source_float_dt, MAXBANDS),
                                ("mixture_amplitudes", source_float_dt, MAXRADII),
                                ("damplitude_drh", source_float_dt, MAXRADII),
                                ("damplitude_dnsersic", source_float_dt, MAXRADII)],
                               align=True)
response_float_dt = source_float_dt
response_struct_dtype = np.dtype([("dchi2_dparam", response_float_dt, NPARAMS * MAXSOURCES)],
                                 align=True)

# kernel configuration used when none is given: 8 wavefronts x 4 pixels/lane
DEFAULT_CONFIG = 5

# launch flags (FPX_LAUNCH_PERSISTENT, FPX_LAUNCH_SPLIT)
LAUNCH_PERSISTENT, LAUNCH_SPLIT = 1, 2

# fixed cost of one piece of a split tile, in pairs inside the cutoff: its
# pixel loads, cull, partial-model exchange and wait.  Measured on the MI355X
# with 8 x 4 tiles: splitting pays on patches with fewer tiles than
# workgroup slots (16-144 tiles of 540-1080 Gaussians: 1.6-2.1x faster) and
# this overhead keeps the planner from splitting patches of many tiles,
# where the pieces cost more than they balance (384 and 1024 tiles)
PIECE_OVERHEAD = 30000

# largest tile count for which evaluate_proposal takes the bound evaluator's
# host path (fpx_evaluate_host: one kernel, no copies); the library sums the
# tiles on the host up to this count
HOST_PATH_TILES = 64


class ProposerBase:

    def evaluate_proposal(self, proposal, **kwargs):
        raise NotImplementedError

    def unpack_residuals(self, residuals_flat, patch, reshape=False):
        """Split a flat padded residual array into per-exposure arrays
        (forcepho/proposal.py:38-52); ``reshape`` returns images."""
        residuals = np.split(residuals_flat, np.cumsum(patch.exposure_N)[:-1])
        if reshape:
            residuals = [patch.unpack_image(residuals_flat, e) for e in range(len(residuals))]
        return residuals


class HIPProposer(ProposerBase):
    """Evaluates proposals with the gfx950 kernel.

    Parameters
    ----------
    patch : HIPPatchMixin patch, optional default patch
    config : int, kernel configuration (``fpx_config_info``); default
        DEFAULT_CONFIG
    persistent : launch a resident grid that pulls tiles from a queue
        (default) or one workgroup per tile
    """

    def __init__(self, patch=None, config=None, persistent=True, chi_dtype=np.float64):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("HIPProposer needs a ROCm device")
        self.patch = patch
        self.chi_dtype = chi_dtype
        self.persistent = bool(persistent)
        self.set_config(DEFAULT_CONFIG if config is None else config)
        self.device = torch.device("cuda", torch.cuda.current_device())
        dev = self.device
        n_acc = lib().fpx_accum_doubles(MAXBANDS)
        self._accum = torch.zeros(n_acc, dtype=torch.float64, device=dev)
        # status word and outputs back to back, so that one copy brings them
        # to the host
        self._out = torch.zeros(_OUT_BYTES, dtype=torch.uint8, device=dev)
        self._status = self._out[:4].view(torch.int32)
        self._chi2 = self._out[_STATUS_BYTES:_CHI2_END].view(torch.float64)
        self._resp = self._out[_CHI2_END:].view(torch.float32)
        self._prop = torch.zeros(MAXSOURCES * source_struct_dtype.itemsize, dtype=torch.uint8,
                                 device=dev)
        self._slot = None       # device FpxSlot of the last launch (built by _slot_for)
        self._slots = {}        # device FpxSlot per (patch struct, tiles, weights)
        self._queue = torch.zeros(2, dtype=torch.int32, device=dev)
        pin = dict(pin_memory=True)
        self._prop_h = torch.zeros(self._prop.numel(), dtype=torch.uint8, **pin)
        self._prop_np = self._prop_h.numpy()
        self._out_h = torch.zeros(_OUT_BYTES, dtype=torch.uint8, **pin)
        self._status_np = self._out_h.numpy()[:4].view(np.uint32)
        self._chi2_np = self._out_h.numpy()[_STATUS_BYTES:_CHI2_END].view(np.float64)
        self._resp_np = self._out_h.numpy()[_CHI2_END:].view(np.float32)
        self._slot_key = None
        self._launch_flags = int(self.persistent)
        self._split_bufs = {}   # split buffer (FpxSlot.split) per split tile table
        self._split_buf = None  # the one the current slot uses, or None
        self._h2d_done = None   # event of the last copy out of the pinned proposal buffer
        # proposal and outputs of the host path (fpx_evaluate_host), with their
        # addresses taken once (numpy's .ctypes.data costs microseconds)
        self._host_prop = np.zeros(MAXSOURCES, dtype=source_struct_dtype)
        self._host_chi2 = np.zeros(MAXBANDS, dtype=np.float64)
        self._host_resp = np.zeros(MAXBANDS * MAXSOURCES * NPARAMS, dtype=np.float32)
        self._host_ptrs = (self._host_prop.ctypes.data, self._host_chi2.ctypes.data,
                           self._host_resp.ctypes.data)

    def set_config(self, config):
        """Select the kernel configuration (tile shape, ``fpx_config_info``)."""
        import ctypes
        ppt = ctypes.c_int32()
        check(lib().fpx_config_info(int(config), None, None, ctypes.byref(ppt)), "fpx_config_info")
        self.config = int(config)
        self.pix_per_tile = ppt.value
        self._slot_key = None

    def autotune(self, proposal, patch=None, repeats=5, configs=None, persistent=(True, False),
                 balance=True, split=(False, True)):
        """Time every (configuration, persistent) pair on ``proposal`` and keep
        the fastest; with ``balance`` the tiles are first reordered for each
        pair (``balance``), and persistent launches are timed with each
        ``split`` choice of ``balance`` (the faster is kept).  Returns
        {(config, persistent): seconds per evaluation}; the times of every
        split choice are in ``split_timings`` {(config, persistent, split):
        seconds}."""
        import torch
        if patch is None:
            patch = self.patch
        configs = range(lib().fpx_n_configs()) if configs is None else configs
        timings, chosen = {}, {}
        self.split_timings = {}
        for cfg in configs:
            self.set_config(cfg)
            for pers in persistent:
                self.persistent = pers
                for sp in (split if (pers and balance) else (False,)):
                    if balance:
                        self.balance(proposal, patch, split=sp)
                    self.launch(proposal, patch)
                    torch.cuda.synchronize()
                    a = torch.cuda.Event(enable_timing=True)
                    b = torch.cuda.Event(enable_timing=True)
                    a.record(self._stream())
                    for _ in range(repeats):
                        self.launch(proposal, patch)
                    b.record(self._stream())
                    torch.cuda.synchronize()
                    t = a.elapsed_time(b) * 1e-3 / repeats
                    self.split_timings[(cfg, pers, sp)] = t
                    if (cfg, pers) not in timings or t < timings[(cfg, pers)]:
                        timings[(cfg, pers)], chosen[(cfg, pers)] = t, sp
        best = min(timings, key=timings.get)
        self.set_config(best[0])
        self.persistent = best[1]
        if balance:
            self.balance(proposal, patch, split=chosen[best])
        return timings

    def _stream(self):
        import torch
        return torch.cuda.current_stream(self.device)

    def _raw_stream(self):
        """Handle of the current stream of the proposer's device, without
        building a torch Stream object (which costs ~2 us per call)."""
        import torch
        raw = getattr(torch._C, "_cuda_getCurrentRawStream", None)
        if raw is None:
            return self._stream().cuda_stream
        return raw(self.device.index)

    def tile_table(self, patch, split=True):
        """(device tile table, entries, split buffer or None) that this
        proposer's launches on ``patch`` use: the patch's split table
        (``balance``) when the launch is persistent and one is set, else its
        plain table."""
        tiles, n = patch.device_tiles(self.pix_per_tile)
        if split and self.persistent:
            table = patch.split_tiles(self.pix_per_tile)
            if table is not None:
                tiles, n = table
                return tiles, n, self._split_buffer(tiles, n)
        return tiles, n, None

    def _split_buffer(self, tiles, n):
        """Zeroed split buffer (FpxSlot.split) for a split table, one per
        table and configuration, kept with the proposer."""
        import torch
        key = (tiles.data_ptr(), n, self.config)
        buf = self._split_bufs.get(key)
        if buf is None:
            if len(self._split_bufs) >= 64:
                self._split_bufs.clear()
            nbytes = int(lib().fpx_split_bytes(n, self.config))
            buf = torch.zeros(max(nbytes, 8), dtype=torch.uint8, device=self.device)
            self._split_bufs[key] = buf
        return buf

    def _slot_for(self, patch, pixel_weight=None, prop_ptr=None, flags=0, split=True):
        """(Re)build the device FpxSlot when the patch struct, tiles, weights,
        proposal buffer or flags change."""
        tiles, n_tiles, split_buf = self.tile_table(patch, split)
        self._split_buf = split_buf
        sp = 0 if split_buf is None else split_buf.data_ptr()
        self._launch_flags = ((LAUNCH_PERSISTENT if self.persistent else 0)
                              | (LAUNCH_SPLIT if split_buf is not None else 0))
        pw = 0 if pixel_weight is None else pixel_weight.data_ptr()
        prop_ptr = self._prop.data_ptr() if prop_ptr is None else prop_ptr
        key = (patch.device_patch.data_ptr(), tiles.data_ptr(), n_tiles, int(patch.n_bands), pw,
               prop_ptr, flags, sp)
        if key != self._slot_key and key in self._slots:
            self._slot, self._slot_key = self._slots[key], key
        elif key != self._slot_key:
            import torch
            if len(self._slots) >= 4096:
                self._slots.clear()
            self._slot = torch.empty(SLOT_DTYPE.itemsize, dtype=torch.uint8, device=self.device)
            self._slots[key] = self._slot
            slot = np.zeros(1, dtype=SLOT_DTYPE)
            slot["patch"] = patch.device_patch.data_ptr()
            slot["proposal"] = prop_ptr
            slot["tiles"] = tiles.data_ptr()
            slot["pixel_weight"] = pw
            slot["pchi2"] = self._chi2.data_ptr()
            slot["pdchi2_dp"] = self._resp.data_ptr()
            slot["accum"] = self._accum.data_ptr()
            slot["status"] = self._out.data_ptr()
            slot["n_bands"] = int(patch.n_bands)
            slot["n_tiles"] = n_tiles
            slot["flags"] = flags
            slot["split"] = sp
            self._slot.copy_(torch.from_numpy(slot.view(np.uint8).reshape(-1).copy()))
            self._slot_key = key
        return n_tiles

    def launch(self, proposal, patch, pixel_weight=None, outputs=True, stream=None,
               model_only=False):
        """Queue one evaluation on the current stream (no synchronisation).
        ``proposal`` is a numpy source_struct_dtype array or a device uint8
        tensor already holding the records.  ``outputs=False`` launches only
        the tile kernel (fpx_evaluate_tiles): its sums stay in the
        accumulator until a launch with outputs (or ``finalize``).
        ``model_only`` (chi2 mode) runs the model pass alone: the residual
        and chi2 are written, the gradient is zero."""
        if model_only and pixel_weight is not None:
            raise ValueError("model_only is a chi2-mode option (no pixel_weight)")
        flags = SLOT_MODEL_ONLY if model_only else 0
        if stream is None and not isinstance(proposal, np.ndarray):
            # a device proposal needs only the stream's handle
            handle = self._raw_stream()
            n_tiles = self._load(proposal, patch, pixel_weight, None, flags)
        else:
            stream = stream or self._stream()
            handle = stream.cuda_stream
            n_tiles = self._load(proposal, patch, pixel_weight, stream, flags)
        fn = lib().fpx_evaluate_batch if outputs else lib().fpx_evaluate_tiles
        check(fn(self._slot.data_ptr(), 1, None, n_tiles, patch.max_gauss(), self.config,
                 self._launch_flags, self._queue.data_ptr(), handle),
              "fpx_evaluate_batch" if outputs else "fpx_evaluate_tiles")
        return n_tiles

    def finalize(self):
        """Write the outputs of the accumulated tile launches and zero the
        accumulator (fpx_finalize).  Needs a launch before it: the output
        pass reads the slot that launch built."""
        if self._slot is None:
            raise RuntimeError("finalize needs a launch first (no device slot has been built)")
        check(lib().fpx_finalize(self._slot.data_ptr(), 1, self._stream().cuda_stream),
              "fpx_finalize")

    def count_pairs(self, proposal, patch=None, per_tile=False):
        """Number of (pixel, ImageGaussian) pairs of the patch inside the
        cutoff dx^T F dx <= MAX_EXP_ARG for this proposal (fpx_count_pairs):
        the pairs that carry the full per-pair work of common_kernel.cc.
        With per_tile, also the counts of the tiles of the current tile table
        (numpy, tile order)."""
        import torch
        if patch is None:
            patch = self.patch
        n_tiles = self._load(proposal, patch, None, split=False)
        out = torch.zeros(1, dtype=torch.int64, device=self.device)
        tiles = torch.zeros(max(n_tiles, 1), dtype=torch.int32, device=self.device)
        check(lib().fpx_count_pairs(self._slot.data_ptr(), 1, None, n_tiles, patch.max_gauss(),
                                    out.data_ptr(), tiles.data_ptr() if per_tile else None,
                                    self._stream().cuda_stream),
              "fpx_count_pairs")
        if per_tile:
            return int(out.item()), tiles[:n_tiles].cpu().numpy().view(np.uint32).astype(np.int64)
        return int(out.item())

    def balance(self, proposal, patch=None, tile_overhead=5000, split=True,
                piece_overhead=PIECE_OVERHEAD):
        """Reorder the patch's tiles for the current configuration so that the
        compute units get equal work, the work of a tile being its pairs
        inside the cutoff for ``proposal`` (fpx_count_pairs) plus
        ``tile_overhead`` pair-equivalents of fixed cost.

        When the whole table is resident at once (fpx_resident_groups) and the
        launch is not persistent, tile i runs on compute unit i mod n_cu: the
        tiles are dealt heaviest first to the least-loaded unit that still
        has a free slot.  Otherwise (a queue) they are sorted heaviest first,
        and with ``split`` the tiles that would hold the launch back are split
        into pieces (``tile_split.plan_parts``; a piece costs its share plus
        ``piece_overhead`` pair-equivalents): the patch then keeps a
        split table beside the plain one, which persistent launches of this
        proposer use.  Returns the permutation applied to the plain table
        (new position -> old position)."""
        import ctypes
        import heapq
        if patch is None:
            patch = self.patch
        _, cost = self.count_pairs(proposal, patch, per_tile=True)
        cost = cost.astype(np.float64) + float(tile_overhead)
        n = len(cost)
        L = lib()
        n_cu, resident = ctypes.c_int32(), ctypes.c_int32()
        check(L.fpx_device_info(ctypes.byref(n_cu), None, None), "fpx_device_info")
        check(L.fpx_resident_groups(self.config, patch.max_gauss(), ctypes.byref(resident)),
              "fpx_resident_groups")
        order = np.argsort(-cost, kind="stable")
        if self.persistent or n > resident.value or n_cu.value < 1:
            perm = order
        else:
            ncu = n_cu.value
            rounds = -(-n // ncu)
            # position r * ncu + u runs on unit u; the last round is partial,
            # so units u >= n - (rounds - 1) * ncu hold one tile fewer
            cap = np.where(np.arange(ncu) < n - (rounds - 1) * ncu, rounds, rounds - 1)
            filled = np.zeros(ncu, dtype=np.int64)
            load = [(0.0, u) for u in range(ncu) if cap[u] > 0]
            heapq.heapify(load)
            perm = np.full(n, -1, dtype=np.int64)
            for t in order:
                w, u = heapq.heappop(load)
                perm[filled[u] * ncu + u] = t
                filled[u] += 1
                if filled[u] < cap[u]:
                    heapq.heappush(load, (w + cost[t], u))
        patch.reorder_tiles(self.pix_per_tile, perm)   # drops an earlier split table
        if self.persistent and split and n > 0:
            over = float(piece_overhead)
            parts = tile_split.plan_parts(cost[perm], resident.value, over)
            if np.any(parts > 1):
                table = tile_split.split_table(patch.tile_table_host(self.pix_per_tile),
                                               cost[perm], parts, over)
                patch.set_split_tiles(self.pix_per_tile, table)
        return perm

    def _load(self, proposal, patch, pixel_weight, stream=None, flags=0, split=True):
        """Stage the proposal and the slot for ``patch``; returns the tile
        count.  A numpy proposal is copied through pinned memory into the
        proposer's buffer on ``stream`` (the launch stream); a device tensor
        is read in place by the launch (it must stay alive and unchanged
        until the launch has run)."""
        import torch
        patch.check_limits()
        n_src = int(patch.n_sources)
        if isinstance(proposal, np.ndarray):
            stream = stream or self._stream()
            if len(proposal) != n_src:
                raise ValueError(f"proposal has {len(proposal)} sources, patch has {n_src}")
            raw = np.ascontiguousarray(proposal, dtype=source_struct_dtype).view(np.uint8)
            if self._h2d_done is not None:
                self._h2d_done.synchronize()   # the pinned buffer may still be in flight
            self._prop_np[:raw.size] = raw.reshape(-1)
            with torch.cuda.stream(stream):
                self._prop.copy_(self._prop_h, non_blocking=True)
            self._h2d_done = torch.cuda.Event()
            self._h2d_done.record(stream)
            return self._slot_for(patch, pixel_weight, flags=flags, split=split)
        if proposal.device != self.device or not proposal.is_contiguous():
            raise ValueError("device proposals must be contiguous tensors on the proposer's device")
        if proposal.numel() * proposal.element_size() < n_src * source_struct_dtype.itemsize:
            raise ValueError("device proposal holds fewer records than the patch has sources")
        return self._slot_for(patch, pixel_weight, prop_ptr=proposal.data_ptr(), flags=flags,
                              split=split)

    def _check_status(self):
        """Raise FpxError if the status word copied back with the outputs is
        set (the kernel left tiles out); the device word is cleared first."""
        st = int(self._status_np[0])
        if st:
            self._status_np[0] = 0
            self._status.zero_()
            if st & STATUS_SPLIT_WAIT and self._split_buf is not None:
                self._split_buf.zero_()   # counters left behind by the timed-out pieces
            raise FpxError(f"the kernel left tiles out (status {st:#x}): a size limit or a "
                           "radius bin of the packed patch is out of range")

    def evaluate_proposal(self, proposal, patch=None, verbose=False, unpack=True):
        """Evaluate chi2 and dchi2/dparam of ``proposal`` on ``patch``."""
        if patch is None:
            patch = self.patch
        import torch
        n_bands, n_sources = int(patch.n_bands), int(patch.n_sources)
        if verbose:
            print(f"fpx launch: config {self.config}, {n_bands} bands, {n_sources} sources",
                  file=sys.stderr, flush=True)
        chi2_np, resp_np = self._chi2_np, self._resp_np
        if isinstance(proposal, np.ndarray):
            if len(proposal) != n_sources:
                raise ValueError(f"proposal has {len(proposal)} sources, patch has {n_sources}")
            patch.check_limits()
            # one native call through a bound evaluator (fpx_evaluate_host):
            # the split table ``balance`` planned for the patch (persistent
            # launches), else, on a small patch, tiles the library plans
            # (read the proposal from and leave their sums in mapped host
            # memory), else the patch's own (balanced) table
            bound = patch.bound_evaluator(self.config, "split") if self.persistent else None
            if bound is None:
                if patch.device_tiles(self.pix_per_tile)[1] <= HOST_PATH_TILES:
                    bound = patch.bound_evaluator(self.config)
                else:
                    bound = patch.bound_evaluator(self.config, "plain", self.persistent)
            if proposal.dtype == source_struct_dtype and proposal.flags.c_contiguous:
                src = proposal.ctypes.data    # read in place (the call copies it)
            else:
                self._host_prop[:n_sources] = proposal
                src = self._host_ptrs[0]
            check(lib().fpx_evaluate_host(bound[0], src, n_sources, self._host_ptrs[1],
                                          self._host_ptrs[2], self._raw_stream()),
                  "fpx_evaluate_host")
            return self._outputs(self._host_chi2, self._host_resp, patch, unpack)
        stream = self._stream()
        self.launch(proposal, patch, stream=stream)
        with torch.cuda.stream(stream):
            self._out_h.copy_(self._out, non_blocking=True)
        stream.synchronize()
        self._check_status()
        return self._outputs(chi2_np, resp_np, patch, unpack)

    def _outputs(self, chi2_np, resp_np, patch, unpack):
        """evaluate_proposal's return value from the output arrays."""
        n_bands, n_sources = int(patch.n_bands), int(patch.n_sources)
        chi_out = chi2_np[:n_bands].astype(self.chi_dtype, copy=True)
        derivs = resp_np[:n_bands * MAXSOURCES * NPARAMS].reshape(n_bands, MAXSOURCES, NPARAMS)
        chi2_derivs = derivs[:, :n_sources].astype(np.float64)
        if patch.return_residual:
            residuals = patch.retrieve_array("residual")
            if unpack:
                residuals = self.unpack_residuals(residuals, patch)
            return chi_out, chi2_derivs, residuals
        return chi_out, chi2_derivs


    def vjp(self, proposal, pixel_weight, patch=None):
        """sum_pix pixel_weight[pix] * d model[pix] / d param for every band
        and source: float64 [n_bands, n_sources, NPARAMS].  ``pixel_weight``
        is a float32 device tensor over the packed pixels.  The kernel runs
        only its gradient pass, recomputing the Gaussians from the proposal.
        """
        if patch is None:
            patch = self.patch
        if pixel_weight.dtype.itemsize != 4 or pixel_weight.numel() < patch.npix:
            raise ValueError("pixel_weight must be float32 over all packed pixels")
        n_bands, n_sources = int(patch.n_bands), int(patch.n_sources)
        stream = self._stream()
        self.launch(proposal, patch, pixel_weight=pixel_weight, stream=stream)
        nr = n_bands * MAXSOURCES * NPARAMS
        import torch
        with torch.cuda.stream(stream):
            self._out_h.copy_(self._out, non_blocking=True)
        stream.synchronize()
        self._check_status()
        d = self._resp_np[:nr].reshape(n_bands, MAXSOURCES, NPARAMS)
        return d[:, :n_sources].astype(np.float64)


# forcepho's names for the device proposer (forcepho/proposal.py:55,155)
GPUProposer = HIPProposer
Proposer = HIPProposer
