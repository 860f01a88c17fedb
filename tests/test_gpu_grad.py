"""GPU tests of the hand-written backward (mesh2_spectrum_poles): gradients of a spectrum loss
with respect to positions and weights against central finite differences of the float64
oracle, and the forward against the oracle.

Tolerance: each checked gradient component g agrees with the finite difference g_fd to
|g - g_fd| <= 1e-4 max|g_fd| over the checked components (the forward runs in float32 with
float64 bin sums); the forward num_raw follows |dP| <= 1e-5 max(|P_l|, |P_0|).
"""
import numpy as np
import pytest
import torch

from oracle import power_oracle as po

from test_gpu_parity import DEV, assert_poles_close, host, say, to_dev

pytestmark = pytest.mark.gpu


def _mock(attrs, size, seed):
    rng = np.random.default_rng(seed)
    lo = attrs.boxcenter - attrs.boxsize / 2.
    pos = rng.uniform(0., 1., (size, 3)) * attrs.boxsize + lo
    n = size // 2
    centers = rng.uniform(0., 1., (4, 3)) * attrs.boxsize + lo
    pos[:n] = centers[rng.integers(0, 4, n)] + rng.normal(0., 0.05, (n, 3)) * attrs.boxsize
    pos[:n] = lo + np.mod(pos[:n] - lo, attrs.boxsize)
    w = rng.uniform(0.5, 1.5, size)
    return pos, w


def _oracle_num(oattrs, obin, pos, w, resampler, interlacing, compensate, los):
    mesh = po.paint(oattrs, pos, w, resampler=resampler, interlacing=interlacing, compensate=compensate, out='complex')
    mesh[0, 0, 0] = 0.
    return np.asarray(po.spectrum_poles(oattrs, obin, mesh, los=los)['num_raw'])


def _oracle_values(oattrs, obin, pos, w, resampler, interlacing, compensate, los):
    """P_ell = (num_raw - SN [ell = 0]) / norm with the box normalization and shot noise."""
    num = _oracle_num(oattrs, obin, pos, w, resampler, interlacing, compensate, los)
    norm = po.box_normalization(oattrs, w)
    sn = po.fkp_shotnoise(w)
    return np.array([po.pole_value(num[i], norm, sn if ell == 0 else 0.) for i, ell in enumerate(obin.ells)])


def _chi2(values, target):
    """sum over ell and k of (P_ell(k) - T_ell(k))^2 / T_ell(k)^2 over the bins that hold modes."""
    ok = target != 0.
    return np.sum(((values - target)[ok] / target[ok])**2)


@pytest.mark.parametrize('resampler,interlacing,compensate,los', [('tsc', 3, True, 'z'), ('cic', 1, False, 'x'),
                                                                  ('pcs', 2, True, (0.6, 0., 0.8))])
def test_gradient_matches_finite_differences(resampler, interlacing, compensate, los):
    """The loss of the differentiated-estimator workload, sum over ell = 0, 2 and k of
    (P_ell - T_ell)^2 / T_ell^2 with P the normalized, shot-noise-subtracted multipoles and T
    those of a second mock, differentiated by the hand-written backward (box normalization and
    shot noise through autograd) against central finite differences of the float64 oracle, with
    position steps of 1e-3 Mpc/h."""
    from jaxpower_amd import MeshAttrs, BinMesh2SpectrumPoles, mesh2_spectrum_poles, box_spectrum_values
    say('grad', resampler, interlacing, compensate, los)
    attrs = MeshAttrs(meshsize=16, boxsize=400., boxcenter=0.)
    ells = (0, 2)
    step = np.min(attrs.kfun) * 8 / 6.
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': step}, ells=ells)
    oattrs = po.Attrs(16, 400., 0.)
    obin = po.BinPoles(oattrs, edges={'step': step}, ells=ells)
    pos, w = _mock(attrs, 400, 1)
    tpos, tw = _mock(attrs, 400, 2)
    target = _oracle_values(oattrs, obin, tpos, tw, resampler, interlacing, compensate, los)
    w64 = w.astype('f4').astype('f8')

    def loss_np(p, ww):
        return float(_chi2(_oracle_values(oattrs, obin, p, ww, resampler, interlacing, compensate, los), target))

    pos_t = to_dev(pos).requires_grad_(True)
    w_t = to_dev(w.astype('f4')).requires_grad_(True)
    num = mesh2_spectrum_poles(pos_t, w_t, attrs=attrs, bin=bin, resampler=resampler, interlacing=interlacing,
                               compensate=compensate, los=los)
    assert_poles_close(host(num), _oracle_num(oattrs, obin, pos, w64, resampler, interlacing, compensate, los), label='forward')
    values = box_spectrum_values(num, w_t, attrs, ells)
    assert np.allclose(host(values), _oracle_values(oattrs, obin, pos, w64, resampler, interlacing, compensate, los), rtol=1e-5,
                       atol=1e-5 * np.abs(target).max())
    tt = to_dev(target)
    ok = tt != 0.
    loss = torch.sum(((values - tt)[ok] / tt[ok])**2)
    assert np.isclose(float(loss), loss_np(pos, w64), rtol=1e-4)
    loss.backward()
    gpos, gw = host(pos_t.grad), host(w_t.grad)
    h, hw = 1e-3, 1e-4
    got, fd = [], []
    for i in (0, 1, 7, 250, 399):
        for a in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[i, a] += h
            pm[i, a] -= h
            fd.append((loss_np(pp, w64) - loss_np(pm, w64)) / (2 * h))
            got.append(gpos[i, a])
        wp, wm = w64.copy(), w64.copy()
        wp[i] += hw
        wm[i] -= hw
        fd.append((loss_np(pos, wp) - loss_np(pos, wm)) / (2 * hw))
        got.append(gw[i])
    got, fd = np.array(got), np.array(fd)
    err = np.abs(got - fd).max() / np.abs(fd).max()
    print('max |g - g_fd| / max |g_fd|', err)
    assert err < 1e-4, (err, got, fd)


def test_gradient_directional_derivative_64():
    """At 64^3 with 5e4 lognormal-like particles, TSC x3: the gradient contracted with a random
    displacement field v (all particles at once) and a random weight perturbation against the
    central difference of the float64 oracle loss along them."""
    from jaxpower_amd import MeshAttrs, BinMesh2SpectrumPoles, mesh2_spectrum_poles, box_spectrum_values
    say('grad directional 64^3')
    attrs = MeshAttrs(meshsize=64, boxsize=500., boxcenter=250.)
    ells = (0, 2)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.02}, ells=ells)
    oattrs = po.Attrs(64, 500., 250.)
    obin = po.BinPoles(oattrs, edges={'step': 0.02}, ells=ells)
    pos, w = _mock(attrs, 50000, 11)
    tpos, tw = _mock(attrs, 50000, 12)
    w64 = w.astype('f4').astype('f8')
    target = _oracle_values(oattrs, obin, tpos, tw, 'tsc', 3, True, 'z')
    rng = np.random.default_rng(13)
    v = rng.normal(size=pos.shape)
    u = rng.normal(size=w.shape)
    pos_t = to_dev(pos).requires_grad_(True)
    w_t = to_dev(w.astype('f4')).requires_grad_(True)
    num = mesh2_spectrum_poles(pos_t, w_t, attrs=attrs, bin=bin, resampler='tsc', interlacing=3, compensate=True, los='z')
    values = box_spectrum_values(num, w_t, attrs, ells)
    tt = to_dev(target)
    ok = tt != 0.
    torch.sum(((values - tt)[ok] / tt[ok])**2).backward()
    dpos = float(np.sum(host(pos_t.grad).astype('f8') * v))
    dw = float(np.sum(host(w_t.grad).astype('f8') * u))

    def loss_np(p, ww):
        return float(_chi2(_oracle_values(oattrs, obin, p, ww, 'tsc', 3, True, 'z'), target))

    h, hw = 1e-3, 1e-5
    fdpos = (loss_np(pos + h * v, w64) - loss_np(pos - h * v, w64)) / (2 * h)
    fdw = (loss_np(pos, w64 + hw * u) - loss_np(pos, w64 - hw * u)) / (2 * hw)
    print('directional', dpos, fdpos, dw, fdw)
    assert abs(dpos - fdpos) < 1e-4 * abs(fdpos)
    assert abs(dw - fdw) < 1e-4 * abs(fdw)


def test_gradient_linear_in_upstream_and_zero_for_odd_ells():
    """The backward is linear in d loss / d num; odd multipoles of an auto spectrum are zero,
    so their upstream gradient does not reach the particles."""
    from jaxpower_amd import MeshAttrs, BinMesh2SpectrumPoles, mesh2_spectrum_poles
    say('grad linearity')
    attrs = MeshAttrs(meshsize=(32, 16, 16), boxsize=(640., 320., 320.))
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.03}, ells=(0, 1, 2))
    pos, w = _mock(attrs, 3000, 5)
    grads = []
    for up in ([1., 0., 0.], [0., 0., 1.], [2., 0., -3.], [0., 1., 0.]):
        pos_t = to_dev(pos.astype('f4')).requires_grad_(True)
        num = mesh2_spectrum_poles(pos_t, to_dev(w.astype('f4')), attrs=attrs, bin=bin, resampler='tsc', interlacing=2,
                                   compensate=True, los='z')
        u = to_dev(np.array(up)[:, None] * np.ones((1, bin.nbins)))
        (num * u).sum().backward()
        grads.append(host(pos_t.grad).astype('f8'))
    assert np.allclose(grads[2], 2 * grads[0] - 3 * grads[1], rtol=1e-4, atol=1e-6 * np.abs(grads[2]).max())
    assert np.all(grads[3] == 0.)
