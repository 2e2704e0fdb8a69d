"""Host side of the device records (forcepho/sources.py:607-681 on the
device): the spline tables DeviceSplineAmplitudes lays out for
fpx_build_sources, the checks of DeviceRecords, and the argument checks of
fpx_build_sources / fpx_stack_gradients, which return before any launch (no
GPU needed).  The kernels' values are tested against scipy and
Galaxy.proposal in test_gpu_records.py."""

import ctypes

import numpy as np
import pytest

from fpx_testutil import verification_amplitudes
from forcepho_amd import _lib
from forcepho_amd.sources import Galaxy, Scene, SplineMixtureAmplitudes
from forcepho_amd.device_scene import DeviceSplineAmplitudes, DeviceRecords

torch = pytest.importorskip("torch")


@pytest.mark.parametrize("spline_type", ["rect", "smooth"])
def test_spline_tables_layout(spline_type):
    """Knots padded with their last value, lengths per radius, coefficient
    c_i[a, b] at a * Cy + b, and the FpxAmplitudeModel pointing at them."""
    model = verification_amplitudes(spline_type)
    dev = DeviceSplineAmplitudes(model, torch.device("cpu"))
    assert dev.R == len(model.splines) and (dev.kx, dev.ky) == (3, 3)
    for i, s in enumerate(model.splines):
        tx, ty, c = (np.asarray(v) for v in s.tck)
        n = len(tx)
        assert int(dev.tx_len[i]) == n and int(dev.ty_len[i]) == len(ty)
        np.testing.assert_array_equal(dev.tx[i, :n].numpy(), tx)
        assert (dev.tx[i, n:].numpy() == tx[-1]).all()
        nx, ny = n - 4, len(ty) - 4
        cc = dev.c[i].numpy().reshape(dev.Cx, dev.Cy)
        np.testing.assert_array_equal(cc[:nx, :ny], c.reshape(nx, ny))
        assert not cc[nx:].any() and not cc[:, ny:].any()
    st = dev.struct()
    assert st.kind == _lib.AMP_SPLINE and st.n_radii == dev.R
    assert st.tx_stride == dev.tx.shape[1] and st.cy == dev.Cy
    assert st.coef == dev.c.data_ptr() and st.ty_len == dev.ty_len.data_ptr()


def test_spline_tables_reject_high_degree():
    from scipy.interpolate import RectBivariateSpline
    x = np.linspace(0, 1, 12)
    model = SplineMixtureAmplitudes.__new__(SplineMixtureAmplitudes)
    model.radii = np.arange(1.0)
    model.splines = [RectBivariateSpline(x, x, np.outer(x, x), kx=5, ky=5)]
    DeviceSplineAmplitudes(model, torch.device("cpu"))
    model.splines[0].degrees = (6, 5)          # beyond FPX_MAX_SPLINE_DEGREE
    with pytest.raises(ValueError):
        DeviceSplineAmplitudes(model, torch.device("cpu"))


def test_device_records_reject_mixed_models():
    from forcepho_amd.device_scene import device_amplitude_models_ok
    a = Galaxy(amplitude_model=verification_amplitudes("rect"))
    b = Galaxy()
    assert device_amplitude_models_ok(Scene([a]))
    assert device_amplitude_models_ok(Scene([b]))
    assert not device_amplitude_models_ok(Scene([a, b]))
    with pytest.raises(TypeError):
        DeviceRecords([Scene([a]), Scene([b])], [1, 1], torch.zeros(2 * 30 * 264, dtype=torch.uint8), 1, 1)


def test_device_records_check_theta():
    """write() takes only a contiguous float64 [n, S, K] tensor on the
    records' device (checked before any launch)."""
    scene = Scene([Galaxy(filters=["a", "b"]) for _ in range(3)])
    recs = DeviceRecords([scene], [2], torch.zeros(30 * 264, dtype=torch.uint8), 3, 2)
    assert recs.model.kind == _lib.AMP_LOGNORMAL
    assert recs.n_sources.tolist() == [3] and recs.n_bands.tolist() == [2]
    for bad in (torch.zeros(1, 3, 7), torch.zeros(1, 3, 8, dtype=torch.float32),
                torch.zeros(1, 8, 3, dtype=torch.float64).transpose(1, 2)):
        with pytest.raises(ValueError):
            recs.write(bad)
    with pytest.raises(ValueError):
        DeviceRecords([scene], [2], torch.zeros(30 * 264, dtype=torch.uint8), 2, 2)
    with pytest.raises(ValueError):
        DeviceRecords([scene], [2], torch.zeros(10 * 264, dtype=torch.uint8), 3, 2)


def _rc(fn, *args):
    rc = fn(*args)
    return rc, _lib.lib().fpx_last_error().decode()


def test_build_sources_argument_checks():
    L = _lib.lib()
    good = _lib.AmplitudeModel(kind=_lib.AMP_LOGNORMAL, lnr=8, lnr_count=8)
    ptr = ctypes.c_void_p(8)

    def call(model, n=1, S=1, NB=1, theta=ptr, recs=ptr, lower=None, upper=None):
        return _rc(L.fpx_build_sources, theta, lower, upper, n, S, NB, None, None,
                   None if model is None else ctypes.byref(model), recs, None)

    assert call(good, S=0)[0] == -1 and "n_src_max" in call(good, S=0)[1]
    assert call(good, S=31)[0] == -1
    assert call(good, NB=31)[0] == -1 and "n_bands_max" in call(good, NB=31)[1]
    assert call(good, n=-1)[0] == -1
    assert call(None)[0] == -1
    assert call(_lib.AmplitudeModel(kind=7))[1] == "unknown amplitude model kind"
    assert "without radii" in call(_lib.AmplitudeModel(kind=_lib.AMP_LOGNORMAL))[1]
    spl = dict(kind=_lib.AMP_SPLINE, n_radii=9, kx=3, ky=3, tx_stride=11, ty_stride=11,
               cx=7, cy=7, tx=8, ty=8, tx_len=8, ty_len=8, coef=8)
    assert call(_lib.AmplitudeModel(**dict(spl, kx=6)))[0] == -1
    assert call(_lib.AmplitudeModel(**dict(spl, n_radii=11)))[0] == -1
    assert "NULL table" in call(_lib.AmplitudeModel(**dict(spl, coef=None)))[1]
    assert "smaller than" in call(_lib.AmplitudeModel(**dict(spl, tx_stride=7)))[1]
    assert "come together" in call(good, lower=ptr)[1]
    assert call(good, n=0)[0] == 0          # nothing to write: no launch
    assert call(good, theta=None)[0] == -1


def test_stack_gradients_argument_checks():
    L = _lib.lib()
    ptr = ctypes.c_void_p(8)
    f = L.fpx_stack_gradients
    assert _rc(f, ptr, ptr, 1, 0, 1, ptr, ptr, None)[0] == -1
    assert _rc(f, ptr, ptr, 1, 1, 31, ptr, ptr, None)[0] == -1
    assert "without its input" in _rc(f, None, ptr, 1, 1, 1, ptr, None, None)[1]
    assert "without its input" in _rc(f, ptr, None, 1, 1, 1, None, ptr, None)[1]
    assert _rc(f, ptr, ptr, 1, 1, 1, None, None, None)[0] == 0     # no outputs
    assert _rc(f, ptr, ptr, 0, 1, 1, ptr, ptr, None)[0] == 0       # no proposals


def test_leapfrog_update_argument_checks():
    L = _lib.lib()
    p = ctypes.c_void_p(8)

    def call(n=1, S=1, NB=1, drift=0, mom=p, z=p):
        return _rc(L.fpx_leapfrog_update, p, p, p, p, p, p, n, S, NB, 1.0, drift, z, mom, p, p,
                   None)

    assert call(S=0)[0] == -1 and call(S=31)[0] == -1 and call(NB=31)[0] == -1
    assert call(n=-1)[0] == -1
    assert call(z=None) == (-1, "NULL array")
    assert call(drift=1, mom=None) == (-1, "drift needs the momenta")
    assert call(n=0)[0] == 0


def test_hmc_metropolis_argument_checks():
    L = _lib.lib()
    p = ctypes.c_void_p(8)
    f = L.fpx_hmc_metropolis
    assert _rc(f, p, p, p, p, p, p, -1, 4, p, p, None)[0] == -1
    assert _rc(f, p, p, p, p, None, p, 1, 4, p, p, None) == (-1, "NULL array")
    assert _rc(f, p, p, p, p, p, p, 0, 4, p, p, None)[0] == 0
