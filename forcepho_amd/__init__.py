"""forcepho_amd: forcepho's chi2 + gradient evaluation on AMD MI355X (gfx950).

The hot path (``Proposer.evaluate_proposal``) runs hand-written HIP kernels in
``libfpx.so`` through a C ABI (include/forcepho_amd.h).  The modules mirror
the reference's host API and add the batched and device-resident forms:

* ``patches`` (packing: stamps, FITS images, PSF stores; device residency),
  ``proposal`` (the proposer, struct dtypes), ``tile_split`` (tiles shared by
  several workgroups: the planner), ``sources`` (Scene, Galaxy and
  the mixture-amplitude models), ``model`` (Posterior, FastPosterior,
  BoundedTransform), ``kernel_limits``;
* ``batch`` (many patches and chains in one persistent launch),
  ``device_scene`` (Source records and gradient stacking on the device),
  ``autograd`` (torch.autograd access to a patch), ``hmc`` and ``ensemble``
  (leapfrog / HMC, and many chains with their state on the device);
* ``formats`` (FITS, HDF5 and TAN WCS readers), ``region``, ``superscene``
  and ``dispatcher`` (checking regions out of a catalog and sampling many at
  once), ``fitting`` (linear flux solves and optimisation), ``scenes``
  (seeded synthetic workloads).

Nothing here is imported eagerly; the modules that need torch import it on
use.
"""

__version__ = "0.1.0"
