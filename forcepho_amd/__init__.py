"""forcepho_amd: forcepho's chi2 + gradient evaluation on AMD MI355X (gfx950).

The hot path (``Proposer.evaluate_proposal``) runs hand-written HIP kernels in
``libfpx.so`` through a C ABI (include/forcepho_amd.h); the Python modules
mirror the reference's host API: ``patches`` (packing, device residency),
``proposal`` (proposer, struct dtypes), ``sources`` (Scene/Galaxy),
``model`` (Posterior), ``kernel_limits``.
"""

__version__ = "0.1.0"
