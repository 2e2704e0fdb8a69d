"""Device residency of a patch on an MI355X (HIP, device memory via torch).

``HIPPatchMixin`` keeps the method surface of forcepho's GPUPatchMixin
(forcepho/patches/device_patch.py:257-435): ``send_to_device``,
``replace_device_meta_ptrs``, ``send_patchstruct_to_device``,
``swap_on_device``, ``retrieve_array``, ``free`` and ``get_proposer``, plus
the attributes ``device_ptrs`` (name -> device buffer), ``device_patch`` (the
FpxPatch struct in device memory) and ``patch_struct`` (its host copy).
``prepare_model``/``subtract_fixed``/``design_matrix`` follow
forcepho/patches/device_patch.py:28-238.

torch is used only to own device buffers and for copies; all computation
happens in libfpx.so.
"""

import ctypes
import weakref

import numpy as np

from .._lib import lib, check, TILE_DTYPE

__all__ = ["DevicePatchMixin", "HIPPatchMixin", "device_bytes"]


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("no ROCm device visible: the HIP path needs an MI355X")
    return torch


def device_bytes(arr, device):
    """Copy a numpy array's bytes to a new uint8 device tensor."""
    torch = _torch()
    raw = np.ascontiguousarray(arr).view(np.uint8).reshape(-1)
    if raw.size == 0:
        raw = np.zeros(8, dtype=np.uint8)
    return torch.from_numpy(raw.copy()).to(device)


# evaluators whose release was put off because a graph was being captured
_DEFERRED = []


def _capturing():
    try:
        import torch
        return torch.cuda.is_available() and torch.cuda.is_current_stream_capturing()
    except Exception:
        return False


def _release_handles(handles):
    """fpx_release each handle, or keep them for later while the current
    stream is capturing a graph (a release synchronises the device and frees
    memory, neither of which a capture allows; a collected patch can trigger
    one at any time)."""
    if _capturing():
        _DEFERRED.extend(handles)
        return
    pending = handles + _DEFERRED
    del _DEFERRED[:]
    for ev in pending:
        try:
            lib().fpx_release(ev)
        except Exception:
            pass


_PIXELS_PER_TILE = {}


def _config_pixels(config):
    """Pixels per tile of a kernel configuration (fpx_config_info, cached)."""
    ppt = _PIXELS_PER_TILE.get(config)
    if ppt is None:
        v = ctypes.c_int32()
        check(lib().fpx_config_info(int(config), None, None, ctypes.byref(v)), "fpx_config_info")
        ppt = _PIXELS_PER_TILE[config] = v.value
    return ppt


class _BoundEvaluators:
    """FpxEvaluator handles of one upload of a patch's device struct: per
    kernel configuration (fpx_bind), config -> (handle, n_tiles), and per
    tile table of the patch (fpx_bind_tiles), (kind, config, persistent,
    table generation) -> (handle, n_entries), kind "split" or "plain".
    Released when the struct is uploaded again, when the patch is freed and
    when it is collected; a table's evaluators also when the table is
    replaced."""

    def __init__(self):
        self.handles = {}
        self.generation = None

    def release(self):
        handles, self.handles = self.handles, {}
        _release_handles([ev for ev, _ in handles.values()])

    def release_kind(self, kind):
        keys = [k for k in self.handles if isinstance(k, tuple) and k[0] == kind]
        _release_handles([self.handles.pop(k)[0] for k in keys])


class DevicePatchMixin:
    """Device-independent logic shared by device mixins."""

    def prepare_model(self, active=None, fixed=None, big=None, bounds=None, maxactive=15,
                      shapes=None, big_scene_kwargs={}, model_kwargs={}):
        """Subtract big and fixed sources, load the active scene and wrap the
        patch in a FastPosterior (forcepho/patches/device_patch.py:28-72).

        Parameters
        ----------
        active : Scene, or catalog (structured array) of active sources
        fixed, big : Scene or catalog of sources to subtract, or None
            (catalogs are subtracted in blocks of ``maxactive`` sources)
        bounds : (lower, upper) arrays for a BoundedTransform, or None
        shapes : accepted for interface compatibility (the Galaxy SHAPE_COLS)
        """
        from ..model import FastPosterior, Transform
        proposer, scene = self.subtract_fixed(fixed, active=active, big=big,
                                              maxactive=maxactive,
                                              big_scene_kwargs=big_scene_kwargs)
        self.return_residual = False
        self.send_patchstruct_to_device()
        q = scene.get_all_source_params().copy()
        if bounds is None:
            model = FastPosterior(proposer, scene=scene, patch=self,
                                  transform=Transform(len(q)), **model_kwargs)
        else:
            lo, hi = bounds
            model = FastPosterior(proposer, scene=scene, patch=self, lower=lo, upper=hi,
                                  **model_kwargs)
        return model, q

    def _scenes(self, sources, maxactive, **scene_kwargs):
        """Scenes of a Scene (itself) or of a catalog (blocks of maxactive)."""
        from ..sources import Scene
        if sources is None:
            return []
        if isinstance(sources, Scene):
            return [sources]
        cat = np.atleast_1d(sources)
        blocks = np.array_split(cat, np.arange(maxactive, len(cat), maxactive))
        return [self.set_scene(block, **scene_kwargs) for block in blocks]

    def subtract_fixed(self, fixed, active=None, big=None, maxactive=15,
                       big_scene_kwargs={}, **scene_kwargs):
        """Subtract big and fixed sources from the device-side data, then load
        the active scene's metadata (forcepho/patches/device_patch.py:113-183).
        After this the device ``data`` holds data - model(big + fixed)."""
        scenes = self._scenes(big, maxactive, **big_scene_kwargs)
        scenes += self._scenes(fixed, maxactive, **scene_kwargs)
        assert active is not None, "an active scene is required"
        scenes += self._scenes(active, max(1, len(np.atleast_1d(getattr(active, "sources", active)))),
                               **scene_kwargs)
        proposer = self.get_proposer()
        self.return_residual = True
        self.pack_meta(scenes[0])
        self.send_to_device()
        for i, scene in enumerate(scenes[1:]):
            proposal = scenes[i].get_proposal()
            proposer.evaluate_proposal(proposal, patch=self)
            self.pack_meta(scene)
            self.swap_on_device()
        return proposer, scenes[-1]

    def design_matrix(self, active, proposer=None):
        """Per-band design matrices of unit-flux source models
        (forcepho/patches/device_patch.py:185-238).

        Returns
        -------
        Xes : list of (n_source, n_pix_band) arrays, model of each source per
              unit flux in each band
        ys : list of (n_pix_band,) arrays, the device-side data split by band
        """
        y = self.retrieve_array("data")
        ys = np.split(y, np.cumsum(self.band_N_pix[:-1]))
        n_src = len(active.sources)
        Xes = [np.zeros((n_src, n)) for n in self.band_N_pix]
        if proposer is None:
            proposer = self.get_proposer()
        self.return_residual = True
        for i, source in enumerate(active.sources):
            single = active.subscene([i])
            self.pack_meta(single)
            self.replace_device_meta_ptrs()
            self.send_patchstruct_to_device()
            proposal = single.get_proposal()
            _, _, residual = proposer.evaluate_proposal(proposal, patch=self, unpack=False)
            model = y - residual
            split = np.split(model, np.cumsum(self.band_N_pix[:-1]))
            for j in range(self.n_bands):
                Xes[j][i, :] = split[j] / source.flux[j]
        return Xes, ys


class HIPPatchMixin(DevicePatchMixin):
    """Mix-in that puts a packed patch in MI355X device memory."""

    device = None   # torch device; default cuda:0

    def _dev(self):
        torch = _torch()
        if self.device is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        return self.device

    def get_proposer(self):
        from ..proposal import HIPProposer
        return HIPProposer(patch=self)

    def send_to_device(self):
        """Copy every array named in the Patch struct to the device and build
        the device-side struct; returns ``device_patch``."""
        dev = self._dev()
        self.device_ptrs = {}
        for name in self.patch_struct_dtype.names:
            if self.patch_struct_dtype[name] != self.ptr_dtype or name == "residual":
                continue
            self.device_ptrs[name] = device_bytes(getattr(self, name), dev)
        if self.return_residual:
            torch = _torch()
            self.device_ptrs["residual"] = torch.zeros(self.xpix.nbytes, dtype=torch.uint8,
                                                       device=dev)
        self._tiles = {}
        self._tiles_host = {}
        self._split_tables = {}
        self.send_patchstruct_to_device()
        self._dirty_data = False
        return self.device_patch

    def replace_device_meta_ptrs(self):
        """Upload the currently packed metadata arrays, replacing (and
        releasing) the previous device copies."""
        dev = self._dev()
        for name in self.meta_names:
            if self.patch_struct_dtype[name] == self.ptr_dtype:
                self.device_ptrs[name] = device_bytes(getattr(self, name), dev)

    def send_patchstruct_to_device(self):
        """Fill the FpxPatch struct with device pointers and scalar values and
        copy it to the device."""
        ps = np.zeros(1, dtype=self.patch_struct_dtype)[0]
        for name in self.patch_struct_dtype.names:
            if self.patch_struct_dtype[name] == self.ptr_dtype:
                buf = self.device_ptrs.get(name)
                if name == "residual" and not self.return_residual:
                    buf = None
                ps[name] = 0 if buf is None else buf.data_ptr()
            else:
                ps[name] = getattr(self, name)
        self.patch_struct = ps
        self.device_patch = device_bytes(np.array([ps]), self._dev())
        self._struct_generation = getattr(self, "_struct_generation", 0) + 1
        return self.device_patch

    def swap_on_device(self):
        """Upload the packed metadata, swap the device data and residual
        buffers and rebuild the struct: afterwards the device uses the former
        residual as its data (forcepho/patches/device_patch.py:343-364)."""
        assert "residual" in self.device_ptrs, "Must instantiate the Patch with `return_residual=True`"
        self.replace_device_meta_ptrs()
        d = self.device_ptrs
        d["data"], d["residual"] = d["residual"], d["data"]
        self.send_patchstruct_to_device()
        self._dirty_data = True
        return self.device_patch

    def retrieve_array(self, arrname="residual"):
        """Copy a device pixel array back to the host (float32 [npix])."""
        buf = self.device_ptrs[arrname]
        return buf.cpu().numpy().view(self.pix_dtype)[:self.npix].copy()

    def tile_order(self, tiles):
        """Permutation of a tile table putting tiles near a source first: the
        work queue then hands out the expensive tiles early.  Distance from
        the tile's pixel centroid to the nearest source centre of the packed
        scene (crpix), in pixels."""
        if len(tiles) == 0 or getattr(self, "crpix", None) is None:
            return np.arange(len(tiles))
        n = tiles["pix_count"].astype(np.int64)
        idx = np.repeat(np.arange(len(tiles)), n)
        first = np.cumsum(n) - n
        pix = np.arange(n.sum()) - first[idx] + tiles["pix_start"].astype(np.int64)[idx]
        count = np.maximum(n, 1)
        cx = np.bincount(idx, self.xpix[pix].astype(np.float64), len(tiles)) / count
        cy = np.bincount(idx, self.ypix[pix].astype(np.float64), len(tiles)) / count
        c = np.asarray(self.crpix, dtype=np.float64).reshape(int(self.n_exp), -1, 2)
        c = c[tiles["exposure"].astype(np.int64)]                 # (n_tiles, n_src, 2)
        d = np.hypot(c[..., 0] - cx[:, None], c[..., 1] - cy[:, None]).min(axis=1)
        return np.argsort(d, kind="stable")

    def device_tiles(self, pix_per_tile):
        """Device FpxTile table of this patch for a tile size (cached), ordered
        by ``tile_order``; returns (tensor, n_tiles)."""
        if pix_per_tile not in self._tiles:
            L = lib()
            args = [np.ascontiguousarray(self.exposure_start, dtype=np.int32),
                    np.ascontiguousarray(self.exposure_N, dtype=np.int32),
                    np.ascontiguousarray(self.band_start, dtype=np.int16),
                    np.ascontiguousarray(self.band_N, dtype=np.int16)]
            ptrs = [a.ctypes.data for a in args]
            n = check(L.fpx_plan_tiles(*ptrs, int(self.n_bands), int(pix_per_tile), None, 0),
                      "fpx_plan_tiles")
            tiles = np.zeros(max(n, 1), dtype=TILE_DTYPE)
            check(L.fpx_plan_tiles(*ptrs, int(self.n_bands), int(pix_per_tile),
                                   tiles.ctypes.data, n), "fpx_plan_tiles")
            tiles[:n] = tiles[:n][self.tile_order(tiles[:n])]
            self._tiles[pix_per_tile] = (device_bytes(tiles, self._dev()), n)
            self._tiles_host[pix_per_tile] = tiles
        return self._tiles[pix_per_tile]

    def reorder_tiles(self, pix_per_tile, perm):
        """Replace the cached tile table for pix_per_tile by its permutation
        ``perm`` (new position k holds old tile perm[k]); returns (tensor,
        n_tiles)."""
        self.device_tiles(pix_per_tile)
        tiles = self._tiles_host[pix_per_tile]
        n = self._tiles[pix_per_tile][1]
        perm = np.asarray(perm, dtype=np.int64)
        self.set_split_tiles(pix_per_tile, None)
        if sorted(perm.tolist()) != list(range(n)):
            raise ValueError("perm is not a permutation of the tile table")
        tiles = tiles.copy()
        tiles[:n] = tiles[:n][perm]
        self._tiles[pix_per_tile] = (device_bytes(tiles, self._dev()), n)
        self._tiles_host[pix_per_tile] = tiles
        self._tiles_generation = self.__dict__.get("_tiles_generation", 0) + 1
        cache = self.__dict__.get("_bound")
        if cache is not None:
            cache.release_kind("plain")
        return self._tiles[pix_per_tile]

    def tile_table_host(self, pix_per_tile):
        """Host copy (numpy TILE_DTYPE) of the tile table ``device_tiles``
        returns."""
        _, n = self.device_tiles(pix_per_tile)
        return self._tiles_host[pix_per_tile][:n].copy()

    def set_split_tiles(self, pix_per_tile, tiles):
        """Keep a tile table with split tiles (FpxTile parts >= 2, each split
        tile's pieces in a row) for persistent launches of this tile size,
        beside the plain table of ``device_tiles``; ``tiles=None`` drops it.
        Dropped too when the plain table is reordered or the patch is sent
        to the device again."""
        tables = self.__dict__.setdefault("_split_tables", {})
        self._split_generation = self.__dict__.get("_split_generation", 0) + 1
        cache = self.__dict__.get("_bound")
        if cache is not None:
            cache.release_kind("split")
        if tiles is None:
            tables.pop(pix_per_tile, None)
            return None
        from ..tile_split import check_table
        tiles = check_table(tiles)
        tables[pix_per_tile] = (device_bytes(tiles, self._dev()), len(tiles), tiles.copy())
        return tables[pix_per_tile][:2]

    def split_tiles(self, pix_per_tile, host=False):
        """(tensor, n_items) of the split table set for this tile size (with
        ``host``, its numpy copy), or None."""
        table = self.__dict__.get("_split_tables", {}).get(pix_per_tile)
        if table is None:
            return None
        return table[2] if host else table[:2]

    def max_gauss(self):
        """n_sources x the largest n_psf_per_source of the packed patch (kept
        per upload of the device struct: launches ask for it every call)."""
        gen = getattr(self, "_struct_generation", None)
        cached = self.__dict__.get("_max_gauss")
        if cached is not None and gen is not None and cached[0] == gen:
            return cached[1]
        value = int(self.n_sources) * int(np.max(self.n_psf_per_source))
        self._max_gauss = (gen, value)
        return value

    def bound_evaluator(self, config, table=None, persistent=False):
        """An FpxEvaluator bound to the current device struct for kernel
        configuration ``config``, kept until the struct is uploaded again;
        returns (handle, n_entries).  table=None: the library plans the
        tiles (fpx_bind, one workgroup per tile).  table="plain": the patch's
        own tile table for the configuration's tile size (``device_tiles``,
        balanced by ``HIPProposer.balance``), launched persistent or not;
        table="split": the split table set for it (persistent), or None if
        there is none (both fpx_bind_tiles, kept until the table changes)."""
        cache = self.__dict__.get("_bound")
        if cache is None:
            cache = self._bound = _BoundEvaluators()
            weakref.finalize(self, cache.release)
        gen = getattr(self, "_struct_generation", None)
        if cache.generation != gen:
            cache.release()
            cache.generation = gen
        if table is not None:
            # the key first (dictionary lookups: this runs on every
            # evaluation), the table only when binding
            ppt = _config_pixels(config)
            if table == "split":
                if ppt not in self.__dict__.get("_split_tables", {}):
                    return None
                persistent = True
                gen = self.__dict__.get("_split_generation", 0)
            elif table == "plain":
                gen = self.__dict__.get("_tiles_generation", 0)
            else:
                raise ValueError("table must be None, 'plain' or 'split'")
            key = (table, config, bool(persistent), gen)
            hit = cache.handles.get(key)
            if hit is not None:
                return hit
            if key not in cache.handles:
                tiles = (self.split_tiles(ppt, host=True) if table == "split"
                         else self.tile_table_host(ppt))
                tiles = np.ascontiguousarray(tiles, dtype=TILE_DTYPE)
                ev = ctypes.c_void_p()
                check(lib().fpx_bind_tiles(self.device_patch.data_ptr(), int(self.n_bands),
                                           int(config), int(bool(persistent)),
                                           tiles.ctypes.data, len(tiles), ctypes.byref(ev)),
                      "fpx_bind_tiles")
                cache.handles[key] = (ev, len(tiles))
            return cache.handles[key]
        if config not in cache.handles:
            ev = ctypes.c_void_p()
            check(lib().fpx_bind(self.device_patch.data_ptr(), int(self.n_bands), int(config), 0,
                                 ctypes.byref(ev)), "fpx_bind")
            n = ctypes.c_int32()
            check(lib().fpx_evaluator_info(ev, ctypes.byref(n), None, None), "fpx_evaluator_info")
            cache.handles[config] = (ev, n.value)
        return cache.handles[config]

    def check_limits(self):
        """Raise ValueError unless the packed patch is within the limits the
        kernel relies on: n_sources, n_bands and n_radii within MAXSOURCES,
        MAXBANDS and MAXRADII, n_psf_per_source within FPX_MAX_PSF_PER_SOURCE,
        and every radius bin the kernel reads (the first n_psf_per_source
        entries from psfgauss_start of each exposure) within [0, n_radii).
        The reference checks none of these (exceeding them is undefined,
        forcepho/src/kernel_limits.h); the kernel itself leaves out a tile
        that breaks one and reports it through the slot's status word.
        Checked once per upload of the patch struct."""
        gen = getattr(self, "_struct_generation", None)
        if gen is not None and getattr(self, "_limits_checked", None) == gen:
            return
        from ..kernel_limits import MAXBANDS, MAXRADII, MAXSOURCES, FPX_MAX_PSF_PER_SOURCE
        ns, nb, nr = int(self.n_sources), int(self.n_bands), int(self.n_radii)
        if not 1 <= ns <= MAXSOURCES:
            raise ValueError(f"n_sources {ns} outside [1, MAXSOURCES={MAXSOURCES}]")
        if not 1 <= nb <= MAXBANDS:
            raise ValueError(f"n_bands {nb} outside [1, MAXBANDS={MAXBANDS}]")
        if not 1 <= nr <= MAXRADII:
            raise ValueError(f"n_radii {nr} outside [1, MAXRADII={MAXRADII}]")
        n_psf = np.asarray(self.n_psf_per_source[:nb], dtype=np.int64)
        if n_psf.min() < 1 or n_psf.max() > FPX_MAX_PSF_PER_SOURCE:
            raise ValueError(f"n_psf_per_source outside [1, {FPX_MAX_PSF_PER_SOURCE}]")
        bins = np.asarray(self.psfgauss["sersic_bin"])
        for b in range(nb):
            for e in range(int(self.band_start[b]), int(self.band_start[b]) + int(self.band_N[b])):
                s = int(self.psfgauss_start[e])
                block = bins[s:s + int(n_psf[b])]
                if len(block) < n_psf[b] or block.min() < 0 or block.max() >= nr:
                    raise ValueError(f"exposure {e}: PSF radius bins outside [0, n_radii={nr})")
        self._limits_checked = gen

    def free(self):
        cache = self.__dict__.get("_bound")
        if cache is not None:
            cache.release()
        self.device_ptrs = {}
        self.device_patch = None
        self._tiles = {}
        self._tiles_host = {}
        self._split_tables = {}
