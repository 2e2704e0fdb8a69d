"""The device record builder (fpx_build_sources) and gradient stacking
(fpx_stack_gradients) against the host:

* spline amplitudes and their first derivatives against scipy's own
  evaluation of the same splines (forcepho/sources.py:657-681: spline(n, rh),
  dx=1, dy=1) for the RectBivariateSpline and SmoothBivariateSpline forms of
  the reference's table, for other degrees and per-radius knots, inside and
  outside the knot range (FITPACK clamps the arguments);
* whole records against Galaxy.proposal (sources.py:607-621) for spline and
  log-normal scenes, with proposals of fewer sources and bands than the
  parameter tensor (zero records, zero fluxes);
* stacked gradients and chi2 sums against model.stack_grad
  (forcepho/model.py:370-388).

The records hold float32: a double value v is stored within 2^-24 |v|, so
the amplitude bars are that rounding plus 1e-12 of the largest value for the
double evaluation."""

import numpy as np
import pytest

from fpx_testutil import verification_amplitudes
from forcepho_amd.sources import Galaxy, Scene, SplineMixtureAmplitudes, source_struct_dtype
from forcepho_amd.kernel_limits import MAXBANDS, MAXSOURCES

pytestmark = pytest.mark.gpu

F32_REL = 2.0 ** -24


def _device_amplitudes(model, n, r):
    """Records of points (n_i, r_i), 30 sources per proposal, built by
    fpx_build_sources; returns A, dA/dn, dA/drh [P, R] (float32 values)."""
    import torch
    from forcepho_amd.device_scene import DeviceRecords
    P, S = len(n), MAXSOURCES
    n_prop = (P + S - 1) // S
    scene = Scene([Galaxy(filters=["b"], amplitude_model=model) for _ in range(S)])
    buf = torch.zeros(n_prop * MAXSOURCES * 264, dtype=torch.uint8, device="cuda")
    recs = DeviceRecords([scene] * n_prop, [1] * n_prop, buf, S, 1)
    theta = np.zeros((n_prop * S, 7))
    theta[:, 0], theta[:, 3], theta[:, 5], theta[:, 6] = 1.0, 0.7, 2.0, 0.1
    theta[:P, 5], theta[:P, 6] = n, r
    recs.write(torch.tensor(theta.reshape(n_prop, S, 7), device="cuda"))
    rec = buf.cpu().numpy().view(source_struct_dtype)[:P]
    R = len(model.splines)
    return (rec["mixture_amplitudes"][:, :R], rec["damplitude_dnsersic"][:, :R],
            rec["damplitude_drh"][:, :R], rec)


def _scipy(model, n, r):
    A = np.array([[s(x, y)[0, 0] for s in model.splines] for x, y in zip(n, r)])
    dn = np.array([[s(x, y, dx=1)[0, 0] for s in model.splines] for x, y in zip(n, r)])
    dr = np.array([[s(x, y, dy=1)[0, 0] for s in model.splines] for x, y in zip(n, r)])
    return A, dn, dr


def _points(rng, nrange, rrange, k=150):
    n = np.concatenate([rng.uniform(*nrange, k), nrange, [2.2, 3.0]])
    r = np.concatenate([rng.uniform(*rrange, k), rrange, [0.1, 0.09]])
    return n, r


def _close(got, ref, floor):
    tol = F32_REL * np.abs(ref) + floor * np.abs(ref).max()
    assert np.all(np.abs(got - ref) <= tol), np.max(np.abs(got - ref) - tol)


@pytest.mark.parametrize("spline_type", ["rect", "smooth"])
def test_reference_table_splines(spline_type):
    model = verification_amplitudes(spline_type)
    rng = np.random.default_rng(1)
    # inside the table (1..5, 0.03..0.3) and beyond it on every side
    n, r = _points(rng, (0.5, 5.5), (0.01, 0.4))
    A, dn, dr, _ = _device_amplitudes(model, n, r)
    A0, dn0, dr0 = _scipy(model, n, r)
    _close(A, A0, 1e-12)
    _close(dn, dn0, 1e-11)
    _close(dr, dr0, 1e-11)


@pytest.mark.parametrize("spline_type", ["rect", "smooth"])
def test_nan_shape_gives_nan_amplitudes(spline_type):
    """A NaN sersic index or half-light radius gives NaN amplitudes and
    derivatives, as scipy's splines do (the reference's Galaxy.amplitudes,
    forcepho/sources.py:655-681): the knot-range clamp must not turn NaN
    into the first knot.  Finite points beyond the table stay clamped."""
    model = verification_amplitudes(spline_type)
    n = np.array([np.nan, 2.2, np.nan, 0.5, 2.2])
    r = np.array([0.1, np.nan, np.nan, 0.1, 0.5])
    A, dn, dr, _ = _device_amplitudes(model, n, r)
    A0, dn0, dr0 = _scipy(model, n, r)
    for got, ref in ((A, A0), (dn, dn0), (dr, dr0)):
        assert np.all(np.isnan(got[:3])) and np.all(np.isnan(ref[:3]))
        _close(got[3:], ref[3:], 1e-11)


@pytest.mark.parametrize("kx,ky", [(2, 2), (2, 3), (3, 2), (5, 4)])
def test_other_degrees_and_uneven_knots(kx, ky):
    """Degrees other than cubic and per-radius knots of different lengths
    (a smoothing spline per radius with its own knots)."""
    from scipy.interpolate import RectBivariateSpline, SmoothBivariateSpline
    rng = np.random.default_rng(kx * 10 + ky)
    xg = np.sort(np.concatenate([[0.0, 4.0], rng.uniform(0, 4, 9)]))
    yg = np.sort(np.concatenate([[0.0, 1.0], rng.uniform(0, 1, 7)]))
    model = SplineMixtureAmplitudes.__new__(SplineMixtureAmplitudes)
    model.radii = np.arange(3.0)
    Z = [np.sin(xg[:, None] * (i + 1)) * np.cos(3 * yg[None, :]) + i for i in range(2)]
    model.splines = [RectBivariateSpline(xg, yg, z, kx=kx, ky=ky) for z in Z]
    xs, ys = rng.uniform(0, 4, 80), rng.uniform(0, 1, 80)
    zs = np.exp(-xs / 3) * ys ** 2 + 0.01 * rng.normal(size=80)
    model.splines.append(SmoothBivariateSpline(xs, ys, zs, kx=kx, ky=ky, s=0.05))
    assert len({len(s.tck[0]) for s in model.splines}) > 1
    n, r = _points(rng, (-0.5, 4.5), (-0.1, 1.1))
    A, dn, dr, _ = _device_amplitudes(model, n, r)
    A0, dn0, dr0 = _scipy(model, n, r)
    _close(A, A0, 1e-12)
    _close(dn, dn0, 1e-11)
    _close(dr, dr0, 1e-11)


def _scenes(model_for, rng, counts):
    scenes = []
    for ns, nb in counts:
        srcs = []
        for k in range(ns):
            g = Galaxy(filters=[f"b{i}" for i in range(nb)], amplitude_model=model_for())
            g.set_params(np.concatenate([rng.uniform(0.5, 2, nb), rng.normal(0, 1e-5, 2),
                                         [rng.uniform(0.4, 1), rng.uniform(-1, 1),
                                          rng.uniform(1.2, 4.8), rng.uniform(0.04, 0.28)]]))
            srcs.append(g)
        scenes.append(Scene(srcs))
    return scenes


@pytest.mark.parametrize("kind", ["spline", "lognormal"])
def test_records_match_galaxy_proposal(kind):
    """Three proposals of (sources, bands) (3, 2), (1, 4), (5, 1) in one
    launch with S = 5, NB = 4: each record equals Galaxy.proposal's (shape and
    flux fields exactly), fluxes of a proposal's missing bands are zero and
    the records past its sources are all zero."""
    import torch
    from forcepho_amd.device_scene import DeviceRecords
    from forcepho_amd.sources import LogNormalMixtureAmplitudes
    rng = np.random.default_rng(3)
    spline = verification_amplitudes("rect")
    model_for = (lambda: spline) if kind == "spline" else LogNormalMixtureAmplitudes
    counts = [(3, 2), (1, 4), (5, 1)]
    scenes = _scenes(model_for, rng, counts)
    S, NB = 5, 4
    buf = torch.full((3 * MAXSOURCES * 264,), 7, dtype=torch.uint8, device="cuda")
    recs = DeviceRecords(scenes, [nb for _, nb in counts], buf, S, NB)
    theta = np.full((3, S, NB + 6), 0.5)
    for k, ((ns, nb), scene) in enumerate(zip(counts, scenes)):
        p = scene.get_all_source_params().reshape(ns, nb + 6)
        theta[k, :ns, :nb], theta[k, :ns, NB:] = p[:, :nb], p[:, nb:]
    recs.write(torch.tensor(theta, device="cuda"))
    got = buf.cpu().numpy().view(source_struct_dtype).reshape(3, MAXSOURCES)
    for k, ((ns, nb), scene) in enumerate(zip(counts, scenes)):
        host = scene.get_proposal()
        for f in ("ra", "dec", "q", "pa", "nsersic", "rh"):
            np.testing.assert_array_equal(got[k, :ns][f], host[f])
        np.testing.assert_array_equal(got[k, :ns]["fluxes"][:, :nb], host["fluxes"][:, :nb])
        assert not got[k, :ns]["fluxes"][:, nb:].any()
        for f in ("mixture_amplitudes", "damplitude_drh", "damplitude_dnsersic"):
            ref = host[f].astype(np.float64)
            tol = 4 * F32_REL * np.abs(ref) + 1e-7 * np.abs(ref).max()
            assert np.all(np.abs(got[k, :ns][f] - ref) <= tol), (k, f)
        assert not got[k, ns:S].view(np.uint8).any()
        assert np.all(got[k, S:].view(np.uint8) == 0)     # zeroed once at set-up


def test_stack_gradients_match_host():
    """fpx_stack_gradients against model.stack_grad on random responses, the
    chi2 summed over bands b < NB; bands at or past NB do not enter."""
    import torch
    from forcepho_amd.device_scene import stack_grad_device
    from forcepho_amd.model import stack_grad
    rng = np.random.default_rng(9)
    n, S, NB = 5, 7, 3
    resp = rng.normal(size=(n, MAXBANDS, MAXSOURCES, 7)).astype(np.float32)
    chi2 = rng.normal(size=(n, MAXBANDS))
    g, c = stack_grad_device(torch.tensor(resp, device="cuda"), S, NB,
                             chi2=torch.tensor(chi2, device="cuda"))
    g = g.cpu().numpy()
    for k in range(n):
        ref = stack_grad(resp[k, :NB, :S].astype(np.float64), S, NB).reshape(S, NB + 6)
        np.testing.assert_allclose(g[k], ref, rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(c.cpu().numpy(), chi2[:, :NB].sum(axis=1), rtol=1e-14)
    g2 = stack_grad_device(torch.tensor(resp, device="cuda"), S, NB).cpu().numpy()
    np.testing.assert_array_equal(g2, g)


def test_bounded_records_match_transformed_theta():
    """Records from the unconstrained z with lower / upper bounds equal the
    records of theta = lower + sigmoid(z) (upper - lower) computed by torch
    (shape and flux fields to one float32 rounding, amplitudes to a few)."""
    import torch
    from forcepho_amd.device_scene import DeviceRecords
    from forcepho_amd.sources import LogNormalMixtureAmplitudes
    rng = np.random.default_rng(8)
    counts = [(3, 2), (2, 2)]
    scenes = _scenes(LogNormalMixtureAmplitudes, rng, counts)
    S, NB = 3, 2
    theta = np.zeros((2, S, NB + 6))
    for k, ((ns, nb), scene) in enumerate(zip(counts, scenes)):
        theta[k, :ns] = scene.get_all_source_params().reshape(ns, nb + 6)
    lo, hi = theta * 0.8 - 1e-3, theta * 1.2 + 1e-3
    lo[..., NB + 2], hi[..., NB + 2] = 0.2, 1.0
    z = rng.normal(0, 0.5, theta.shape)
    dev = dict(dtype=torch.float64, device="cuda")
    zt, lt, ht = (torch.tensor(a, **dev) for a in (z, lo, hi))
    out = []
    for bounded in (True, False):
        buf = torch.zeros(2 * MAXSOURCES * 264, dtype=torch.uint8, device="cuda")
        recs = DeviceRecords(scenes, [nb for _, nb in counts], buf, S, NB)
        if bounded:
            recs.write(zt, lower=lt, upper=ht)
        else:
            recs.write((lt + torch.sigmoid(zt) * (ht - lt)).contiguous())
        out.append(buf.cpu().numpy().view(source_struct_dtype).reshape(2, MAXSOURCES))
    a, b = out
    for k, (ns, nb) in enumerate(counts):
        for f in ("ra", "dec", "q", "pa", "nsersic", "rh", "fluxes"):
            x, y = a[k, :ns][f].astype(np.float64), b[k, :ns][f].astype(np.float64)
            assert np.all(np.abs(x - y) <= F32_REL * np.abs(y)), f
        for f in ("mixture_amplitudes", "damplitude_drh", "damplitude_dnsersic"):
            x, y = a[k, :ns][f].astype(np.float64), b[k, :ns][f].astype(np.float64)
            assert np.all(np.abs(x - y) <= 4 * F32_REL * np.abs(y) + 1e-7 * np.abs(y).max()), f
