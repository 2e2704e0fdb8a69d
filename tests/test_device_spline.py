"""Device evaluation of forcepho's spline mixture amplitudes
(forcepho/sources.py:623-681) against scipy's own evaluation of the same
splines: values and first derivatives of RectBivariateSpline and
SmoothBivariateSpline forms of the reference's table, of splines of other
degrees, inside and outside the knot range (FITPACK clamps the arguments),
and the Source records DeviceRecords writes from them against
Galaxy.proposal.  torch runs on the CPU here; the GPU tests run the same
code on the device."""

import numpy as np
import pytest

from fpx_testutil import verification_amplitudes
from forcepho_amd.sources import Galaxy, Scene, SplineMixtureAmplitudes
from forcepho_amd.device_scene import DeviceSplineAmplitudes, DeviceRecords

torch = pytest.importorskip("torch")


def _scipy(model, n, r):
    A = np.array([[s(x, y)[0, 0] for s in model.splines] for x, y in zip(n, r)])
    dn = np.array([[s(x, y, dx=1)[0, 0] for s in model.splines] for x, y in zip(n, r)])
    dr = np.array([[s(x, y, dy=1)[0, 0] for s in model.splines] for x, y in zip(n, r)])
    return A, dn, dr


def _points(rng, nrange, rrange, k=150):
    n = np.concatenate([rng.uniform(*nrange, k), nrange, [2.2, 3.0]])
    r = np.concatenate([rng.uniform(*rrange, k), rrange, [0.1, 0.09]])
    return n, r


@pytest.mark.parametrize("spline_type", ["rect", "smooth"])
def test_reference_table_splines(spline_type):
    model = verification_amplitudes(spline_type)
    dev = DeviceSplineAmplitudes(model, torch.device("cpu"))
    rng = np.random.default_rng(1)
    # inside the table (1..5, 0.03..0.3) and beyond it on every side
    n, r = _points(rng, (0.5, 5.5), (0.01, 0.4))
    A, dn, dr = (t.numpy() for t in dev(torch.tensor(n), torch.tensor(r)))
    A0, dn0, dr0 = _scipy(model, n, r)
    np.testing.assert_allclose(A, A0, rtol=0, atol=1e-13)
    np.testing.assert_allclose(dn, dn0, rtol=0, atol=1e-12 * np.abs(dn0).max())
    np.testing.assert_allclose(dr, dr0, rtol=0, atol=1e-12 * np.abs(dr0).max())
    # batch shape [n, S] -> [n, S, R]
    A2, _, _ = dev(torch.tensor(n[:12]).reshape(3, 4), torch.tensor(r[:12]).reshape(3, 4))
    assert tuple(A2.shape) == (3, 4, len(model.radii))
    np.testing.assert_allclose(A2.reshape(12, -1).numpy(), A0[:12], atol=1e-13)


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
    dev = DeviceSplineAmplitudes(model, torch.device("cpu"))
    n, r = _points(rng, (-0.5, 4.5), (-0.1, 1.1))
    A, dn, dr = (t.numpy() for t in dev(torch.tensor(n), torch.tensor(r)))
    A0, dn0, dr0 = _scipy(model, n, r)
    np.testing.assert_allclose(A, A0, rtol=0, atol=1e-12 * max(1, np.abs(A0).max()))
    np.testing.assert_allclose(dn, dn0, rtol=0, atol=1e-11 * np.abs(dn0).max())
    np.testing.assert_allclose(dr, dr0, rtol=0, atol=1e-11 * np.abs(dr0).max())


def test_device_records_from_spline_scene():
    """Records written on the device from theta equal Galaxy.proposal's."""
    from forcepho_amd.kernel_limits import MAXSOURCES
    from forcepho_amd.proposal import source_struct_dtype
    model = verification_amplitudes("rect")
    rng = np.random.default_rng(3)
    nb, ns = 2, 3
    srcs = []
    for k in range(ns):
        g = Galaxy(filters=["a", "b"], amplitude_model=model)
        g.set_params(np.concatenate([rng.uniform(0.5, 2, nb), rng.normal(0, 1e-5, 2),
                                     [rng.uniform(0.4, 1), rng.uniform(-1, 1),
                                      rng.uniform(1.2, 4.8), rng.uniform(0.04, 0.28)]]))
        srcs.append(g)
    scene = Scene(srcs)
    buf = torch.zeros(2 * MAXSOURCES * 264, dtype=torch.uint8)
    recs = DeviceRecords([scene, scene], [nb, nb], buf, ns, nb)
    theta = torch.tensor(np.stack([scene.get_all_source_params().reshape(ns, nb + 6)] * 2))
    recs.write(theta)
    got = buf.numpy().view(source_struct_dtype).reshape(2, MAXSOURCES)
    host = scene.get_proposal()
    for f in ("ra", "dec", "q", "pa", "nsersic", "rh", "fluxes"):
        np.testing.assert_array_equal(got[0, :ns][f], host[f])
        np.testing.assert_array_equal(got[1, :ns][f], host[f])
    for f in ("mixture_amplitudes", "damplitude_drh", "damplitude_dnsersic"):
        np.testing.assert_allclose(got[0, :ns][f], host[f], rtol=1e-6,
                                   atol=1e-7 * np.abs(host[f]).max())


def test_device_records_reject_mixed_models():
    from forcepho_amd.device_scene import device_amplitude_models_ok
    a = Galaxy(amplitude_model=verification_amplitudes("rect"))
    b = Galaxy()
    assert device_amplitude_models_ok(Scene([a]))
    assert device_amplitude_models_ok(Scene([b]))
    assert not device_amplitude_models_ok(Scene([a, b]))
    with pytest.raises(TypeError):
        DeviceRecords([Scene([a]), Scene([b])], [1, 1], torch.zeros(2 * 30 * 264, dtype=torch.uint8), 1, 1)
