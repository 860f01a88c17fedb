"""GPU parity of the power-spectrum window matrix (compute_mesh2_spectrum_window, global line
of sight, theory defined on the mesh; jaxpower/mesh2.py:921-1118, 1453-1482) against the
float64 oracle.

Tolerance: |dW| <= 1e-5 |W| + 1e-5 max|W| entry by entry.  The periodic window involves no FFT
(float32 Legendre factors, FP64 bin sums); with a selection every column goes through a float32
c2r and r2c, whose rounding is ~1e-7 of the largest mode of the convolved shell, spread over all
modes.  Either way the terms summed are O(max|W|) and carry float32 rounding, so blocks that
vanish by symmetry (e.g. ell = 0 from ellin = 2 on full shells) hold residues ~1e-8 max|W|:
hence the floor relative to the whole matrix.
"""
import numpy as np
import pytest
import torch

from oracle import power_oracle as po

from test_gpu_parity import DEV, host, say, to_dev

pytestmark = pytest.mark.gpu


def assert_window_close(got, want, nout, nin, rtol=1e-5, ftol=1e-5, label=''):
    got, want = np.asarray(got), np.asarray(want)
    assert got.shape == want.shape, (got.shape, want.shape)
    nell, nellin = want.shape[0] // nout, want.shape[1] // nin
    floor = ftol * np.nanmax(np.abs(want))
    for i in range(nell):
        for j in range(nellin):
            g = got[i * nout:(i + 1) * nout, j * nin:(j + 1) * nin]
            w = want[i * nout:(i + 1) * nout, j * nin:(j + 1) * nin]
            nan = np.isnan(w)
            assert np.array_equal(nan, np.isnan(g)), f'{label} block ({i}, {j}): NaN pattern differs'
            err = np.abs(g - w)[~nan]
            bad = err > rtol * np.abs(w[~nan]) + floor
            assert not np.any(bad), f'{label} block ({i}, {j}): max err {err.max():.3e}, floor {floor:.3e}'


def _survey_selection(attrs, size, seed, scale=0.2):
    """Gaussian blob of points, squashed along y and z and cut to the box (the selection of
    the reference's tests/test_mesh2.py:528-560)."""
    rng = np.random.default_rng(seed)
    pos = np.array([1., 0.2, 0.2]) * scale * rng.normal(size=(size, 3))
    mask = np.all((pos > -scale) & (pos < scale), axis=-1)
    pos = pos * attrs.boxsize + attrs.boxcenter
    return pos, mask.astype('f4')


@pytest.mark.parametrize('meshsize,boxsize,edgesin,los', [
    (32, 320., 'common', 'z'),
    (32, 320., np.arange(0.01, 0.3, 0.013), 'x'),
    ((16, 32, 64), (320., 400., 640.), np.arange(0., 0.25, 0.02), [0.3, -0.4, np.sqrt(0.75)]),
])
def test_window_periodic(meshsize, boxsize, edgesin, los):
    from jaxpower_amd import MeshAttrs, BinMesh2SpectrumPoles, compute_mesh2_spectrum_window
    say('window periodic', meshsize, boxsize, los)
    attrs = MeshAttrs(meshsize=meshsize, boxsize=boxsize, boxcenter=0.)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.02}, ells=(0, 2, 4))
    if isinstance(edgesin, str):
        edgesin = bin.edges
    ellsin = (0, 2, 4)
    wm = compute_mesh2_spectrum_window(attrs, edgesin=edgesin, ellsin=ellsin, bin=bin, los=los)
    oattrs = po.Attrs(meshsize, boxsize, 0.)
    obin = po.BinPoles(oattrs, edges={'step': 0.02}, ells=(0, 2, 4))
    want = po.spectrum_window(oattrs, obin, edgesin, ellsin=ellsin, los=los)
    nin = len(wm.theory.get(0).coords())
    assert_window_close(wm.value(), want, len(obin.nmodes), nin, label='periodic')
    assert wm.observable.ells == (0, 2, 4) and wm.theory.ells == ellsin


@pytest.mark.parametrize('cross,los,norm', [(False, 'x', None), (True, [0.6, 0.8, 0.], 2.5e-3), (False, 'z', [1e-3, 2e-3, 4e-3])])
def test_window_selection(cross, los, norm):
    from jaxpower_amd import MeshAttrs, ParticleField, BinMesh2SpectrumPoles, compute_mesh2_spectrum_window
    say('window selection', cross, los, norm)
    attrs = MeshAttrs(meshsize=32, boxsize=500., boxcenter=1000.)
    pos, w = _survey_selection(attrs, 200000, 42)
    sel1 = ParticleField(to_dev(pos), to_dev(w), attrs=attrs).paint(resampler='cic', interlacing=1, compensate=False)
    meshes = [sel1]
    if cross:
        pos2, w2 = _survey_selection(attrs, 150000, 7, scale=0.25)
        meshes.append(ParticleField(to_dev(pos2), to_dev(w2), attrs=attrs).paint(resampler='tsc', interlacing=1, compensate=False))
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.02}, ells=(0, 2, 4))
    edgesin = np.arange(0.01, 0.2, 0.02)
    ellsin = (0, 2, 4)
    wm = compute_mesh2_spectrum_window(*meshes, edgesin=edgesin, ellsin=ellsin, bin=bin, los=los, norm=norm)
    oattrs = po.Attrs(32, 500., 1000.)
    obin = po.BinPoles(oattrs, edges={'step': 0.02}, ells=(0, 2, 4))
    omeshes = [host(m.value).astype('f8') for m in meshes]
    want = po.spectrum_window(oattrs, obin, edgesin, ellsin=ellsin, meshes=omeshes, los=los, norm=norm)
    nout, nin = len(obin.nmodes), len(edgesin) - 1
    assert_window_close(wm.value(), want, nout, nin, label='selection')
    # W . theory for a theory piecewise constant on the input shells
    rng = np.random.default_rng(1)
    theta = rng.uniform(0.5, 2., want.shape[1])
    obs = wm.dot(theta)
    got = np.concatenate([p.value() for p in obs])
    ref = want @ theta
    assert np.allclose(got, ref, rtol=1e-5, atol=1e-5 * np.abs(ref).max())


def test_window_constant_selection_is_periodic():
    """A constant selection (Q = 1) must give the periodic window: the convolution path against
    the direct one on the GPU itself."""
    from jaxpower_amd import MeshAttrs, RealMeshField, BinMesh2SpectrumPoles, compute_mesh2_spectrum_window
    say('window constant selection')
    attrs = MeshAttrs(meshsize=32, boxsize=400., boxcenter=0.)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.025}, ells=(0, 2, 4))
    ones = RealMeshField(torch.ones((32, 32, 32), dtype=torch.float32, device=DEV), attrs=attrs)
    edgesin = np.arange(0., 0.3, 0.03)
    w1 = compute_mesh2_spectrum_window(ones, edgesin=edgesin, ellsin=(0, 2), bin=bin, los='y').value()
    w0 = compute_mesh2_spectrum_window(attrs, edgesin=edgesin, ellsin=(0, 2), bin=bin, los='y').value()
    assert_window_close(w1, w0, bin.nbins, len(edgesin) - 1, label='constant selection')


def test_window_unsupported_raise():
    from jaxpower_amd import MeshAttrs, BinMesh2SpectrumPoles, compute_mesh2_spectrum_window
    attrs = MeshAttrs(meshsize=16, boxsize=160., boxcenter=0.)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.05}, ells=(0, 2))
    with pytest.raises(NotImplementedError):
        compute_mesh2_spectrum_window(attrs, edgesin=np.arange(0., 0.2, 0.05), bin=bin, los='firstpoint', ellsin=([(0, 1), (2, 0)], 'local'),
                                      flags=('infinite',))
    with pytest.raises(NotImplementedError):
        compute_mesh2_spectrum_window(attrs, edgesin=np.arange(0., 0.2, 0.05), bin=bin, los='z', ellsin=0, flags=('fftlog',))


def test_window_fused_column_matches_written_spectrum():
    """mp_window_bin (shell generated in the inverse x pass, binned in the forward x pass) against
    mp_window_shell (the convolved spectrum written out) followed by mp_bin_poles, through the
    C-ABI, on the same selection and shell."""
    from jaxpower_amd import MeshAttrs, ParticleField, BinMesh2SpectrumPoles, _lib
    from jaxpower_amd.mesh import _empty_spectrum, _spectrum_pitch, _c2r
    from jaxpower_amd.mesh2 import _k2_thresholds
    say('window fused column vs written spectrum')
    attrs = MeshAttrs(meshsize=(32, 64, 32), boxsize=(400., 800., 400.), boxcenter=0.)
    pos, w = _survey_selection(attrs, 100000, 5)
    sel = ParticleField(to_dev(pos), to_dev(w), attrs=attrs).paint(resampler='cic', interlacing=1, compensate=False)
    A = sel.r2c().value
    Q = _c2r((A * A.conj()).to(torch.complex64), attrs.meshsize)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.02}, ells=(0, 2, 4))
    t = bin._tables
    n = _lib.i32x3(attrs.meshsize)
    los = _lib.f64x3([0., 0.6, 0.8])
    t_lo, t_hi = _k2_thresholds(np.array([0.03, 0.07]))
    ells = (_lib.ctypes.c_int32 * 3)(0, 2, 4)
    ws = _lib.workspace(_lib.load().mp_window_bin_workspace_bytes(n, 3, len(bin._edges1d)), Q.device)
    outs = torch.empty((2, 3, bin.nbins, 2), dtype=torch.float64, device=Q.device)
    _lib.call('mp_window_bin', _lib.stream_handle(Q.device), _lib.ptr(Q), n, _lib.ptr(t['kx']), _lib.ptr(t['ky']),
              _lib.ptr(t['kz']), _lib.ptr(t['thr']), len(bin._edges1d), 2, (_lib.ctypes.c_double * 2)(t_lo, t_lo),
              (_lib.ctypes.c_double * 2)(t_hi, t_hi), (_lib.ctypes.c_int32 * 2)(0, 2), los, 1., 3, ells, _lib.ptr(outs),
              _lib.ptr(ws), ws.numel())
    for ic, ellin in enumerate((0, 2)):
        out = outs[ic]
        B = _empty_spectrum(attrs.meshsize, Q.device)
        _lib.call('mp_window_shell', _lib.stream_handle(Q.device), _lib.ptr(Q), n, _lib.ptr(t['kx']), _lib.ptr(t['ky']),
                  _lib.ptr(t['kz']), float(t_lo), float(t_hi), ellin, los, 1., _lib.ptr(B), _spectrum_pitch(B))
        ref = bin.products([B], [(-1, 0, ell, True) for ell in (0, 2, 4)], los=[0., 0.6, 0.8])
        got = host(out)[..., 0]
        floor = 1e-6 * np.abs(ref.real).max()
        assert np.all(np.abs(got - ref.real) <= 1e-6 * np.abs(ref.real) + floor), (ellin, np.abs(got - ref.real).max(), floor)
        assert np.all(host(out)[..., 1] == 0.)


@pytest.mark.parametrize('periodic,los', [(True, 'z'), (False, 'x'), (False, [0.36, 0.48, 0.8])])
def test_window_infinite(periodic, los):
    """flags=('infinite',) (mesh2.py:1078-1092): configuration-space shells from the closed-form
    Bessel integrals, times the selection, r2c and binned, against the oracle."""
    from jaxpower_amd import MeshAttrs, ParticleField, BinMesh2SpectrumPoles, compute_mesh2_spectrum_window
    say('window infinite', periodic, los)
    attrs = MeshAttrs(meshsize=32, boxsize=500., boxcenter=1000.)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.02}, ells=(0, 2, 4))
    edgesin = np.arange(0.01, 0.2, 0.03)
    ellsin = (0, 2, 4)
    oattrs = po.Attrs(32, 500., 1000.)
    obin = po.BinPoles(oattrs, edges={'step': 0.02}, ells=(0, 2, 4))
    if periodic:
        wm = compute_mesh2_spectrum_window(attrs, edgesin=edgesin, ellsin=ellsin, bin=bin, los=los, flags=('infinite',))
        want = po.spectrum_window_infinite(oattrs, obin, edgesin, ellsin=ellsin, los=los)
    else:
        pos, w = _survey_selection(attrs, 200000, 11)
        sel = ParticleField(to_dev(pos), to_dev(w), attrs=attrs).paint(resampler='cic', interlacing=1, compensate=False)
        wm = compute_mesh2_spectrum_window(sel, edgesin=edgesin, ellsin=ellsin, bin=bin, los=los, flags=('infinite',))
        want = po.spectrum_window_infinite(oattrs, obin, edgesin, ellsin=ellsin, meshes=[host(sel.value).astype('f8')], los=los)
    assert_window_close(wm.value(), want, len(obin.nmodes), len(edgesin) - 1, label='infinite')


@pytest.mark.parametrize('periodic,los', [(True, 'z'), (False, 'x'), (False, [0.36, 0.48, 0.8])])
def test_window_smooth(periodic, los):
    """flags=('smooth',) (mesh2.py:1045-1076): the selection times P_l P_lin binned in
    separation shells (one GPU pass over the Legendre series of the products), Hankel-summed
    against each input shell, against the oracle."""
    from jaxpower_amd import MeshAttrs, ParticleField, BinMesh2SpectrumPoles, compute_mesh2_spectrum_window
    say('window smooth', periodic, los)
    attrs = MeshAttrs(meshsize=32, boxsize=500., boxcenter=1000.)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.02}, ells=(0, 2, 4))
    edgesin = np.arange(0.01, 0.2, 0.03)
    ellsin = (0, 2, 4)
    oattrs = po.Attrs(32, 500., 1000.)
    obin = po.BinPoles(oattrs, edges={'step': 0.02}, ells=(0, 2, 4))
    if periodic:
        wm = compute_mesh2_spectrum_window(attrs, edgesin=edgesin, ellsin=ellsin, bin=bin, los=los, flags=('smooth',))
        want = po.spectrum_window_smooth(oattrs, obin, edgesin, ellsin=ellsin, los=los)
    else:
        pos, w = _survey_selection(attrs, 200000, 13)
        sel = ParticleField(to_dev(pos), to_dev(w), attrs=attrs).paint(resampler='cic', interlacing=1, compensate=False)
        wm = compute_mesh2_spectrum_window(sel, edgesin=edgesin, ellsin=ellsin, bin=bin, los=los, flags=('smooth',))
        want = po.spectrum_window_smooth(oattrs, obin, edgesin, ellsin=ellsin, meshes=[host(sel.value).astype('f8')], los=los)
    assert_window_close(wm.value(), want, len(obin.nmodes), len(edgesin) - 1, label='smooth')


@pytest.mark.parametrize('periodic,los,ellsin', [(True, 'firstpoint', (0, 2)), (False, 'firstpoint', (0, 2, 4)),
                                                 (False, 'endpoint', ((0, 2), 'firstpoint'))])
def test_window_local_infinite(periodic, los, ellsin):
    """flags=('infinite',) with a local line of sight (mesh2.py:1235-1260): per input multipole and
    output (l, m) the selection term sum_m1 Y_lin,m1(shat) c2r(conj(r2c(rho Y_lm Y_lin,m1)) A0),
    then per input shell xi_lin(|s|) times it, r2c, Y_lm(khat)-weighted binning, summed over m;
    against the oracle (whose Y_lm sign conventions may differ: only the sum over m is compared)."""
    from jaxpower_amd import MeshAttrs, ParticleField, BinMesh2SpectrumPoles, compute_mesh2_spectrum_window
    say('window local infinite', periodic, los, ellsin)
    attrs = MeshAttrs(meshsize=16, boxsize=400., boxcenter=(0., 0., 0.) if periodic else (900., 300., 500.))
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.03}, ells=(0, 2, 4))
    edgesin = np.arange(0.01, 0.12, 0.025)
    oattrs = po.Attrs(16, 400., (0., 0., 0.) if periodic else (900., 300., 500.))
    obin = po.BinPoles(oattrs, edges={'step': 0.03}, ells=(0, 2, 4))
    oellsin = ellsin[0] if isinstance(ellsin[-1], str) else ellsin
    if periodic:
        wm = compute_mesh2_spectrum_window(attrs, edgesin=edgesin, ellsin=ellsin, bin=bin, los=los, flags=('infinite',))
        want = po.spectrum_window_local_infinite(oattrs, obin, edgesin, ellsin=oellsin)
    else:
        pos, w = _survey_selection(attrs, 100000, 17, scale=0.3)
        sel = ParticleField(to_dev(pos), to_dev(w), attrs=attrs).paint(resampler='cic', interlacing=1, compensate=False)
        wm = compute_mesh2_spectrum_window(sel, edgesin=edgesin, ellsin=ellsin, bin=bin, los=los, flags=('infinite',))
        want = po.spectrum_window_local_infinite(oattrs, obin, edgesin, ellsin=oellsin, mesh=host(sel.value).astype('f8'))
    assert_window_close(wm.value(), want, len(obin.nmodes), len(edgesin) - 1, label='local infinite')


@pytest.mark.parametrize('periodic,los,ellsin', [(True, 'firstpoint', (0, 2)), (False, 'firstpoint', (0, 2)),
                                                 (False, 'endpoint', (2,))])
def test_window_local_mesh_theory(periodic, los, ellsin):
    """No flags with a local line of sight (mesh2.py:1262-1305): xi_m1 = c2r(closed shell
    Y_lin,m1(khat)), A_lm = sum_m1 xi_m1 c2r(conj(r2c(rho Y_lm Y_lin,m1)) A0), binned with
    Y_lm(khat) and summed over m, against the oracle."""
    from jaxpower_amd import MeshAttrs, ParticleField, BinMesh2SpectrumPoles, compute_mesh2_spectrum_window
    say('window local mesh theory', periodic, los, ellsin)
    center = (0., 0., 0.) if periodic else (900., 300., 500.)
    attrs = MeshAttrs(meshsize=16, boxsize=400., boxcenter=center)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.03}, ells=(0, 2, 4))
    kf = 2. * np.pi / 400.
    edgesin = np.array([[kf, 3 * kf], [3 * kf, 5 * kf], [5 * kf, 8 * kf]])  # closed right edges on lattice shells
    oattrs = po.Attrs(16, 400., center)
    obin = po.BinPoles(oattrs, edges={'step': 0.03}, ells=(0, 2, 4))
    if periodic:
        wm = compute_mesh2_spectrum_window(attrs, edgesin=edgesin, ellsin=ellsin, bin=bin, los=los)
        want = po.spectrum_window_local(oattrs, obin, edgesin, ellsin=ellsin)
    else:
        pos, w = _survey_selection(attrs, 100000, 19, scale=0.3)
        sel = ParticleField(to_dev(pos), to_dev(w), attrs=attrs).paint(resampler='cic', interlacing=1, compensate=False)
        wm = compute_mesh2_spectrum_window(sel, edgesin=edgesin, ellsin=ellsin, bin=bin, los=los)
        want = po.spectrum_window_local(oattrs, obin, edgesin, ellsin=ellsin, mesh=host(sel.value).astype('f8'))
    assert_window_close(wm.value(), want, len(obin.nmodes), len(edgesin), label='local mesh theory')


@pytest.mark.parametrize('periodic,los,ellsin', [(True, 'firstpoint', (0, 2)), (False, 'firstpoint', (0, 2, 4))])
def test_window_local_smooth(periodic, los, ellsin):
    """flags=('smooth',) with a local line of sight (mesh2.py:1195-1233): the selection terms times
    Y_lm(shat) Y_lin,m1(shat) summed over (m, m1), binned in separation shells, Hankel-summed
    against each input shell, against the oracle."""
    from jaxpower_amd import MeshAttrs, ParticleField, BinMesh2SpectrumPoles, compute_mesh2_spectrum_window
    say('window local smooth', periodic, los, ellsin)
    center = (0., 0., 0.) if periodic else (900., 300., 500.)
    attrs = MeshAttrs(meshsize=16, boxsize=400., boxcenter=center)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.03}, ells=(0, 2, 4))
    edgesin = np.arange(0.01, 0.12, 0.025)
    oattrs = po.Attrs(16, 400., center)
    obin = po.BinPoles(oattrs, edges={'step': 0.03}, ells=(0, 2, 4))
    if periodic:
        wm = compute_mesh2_spectrum_window(attrs, edgesin=edgesin, ellsin=ellsin, bin=bin, los=los, flags=('smooth',))
        want = po.spectrum_window_local_smooth(oattrs, obin, edgesin, ellsin=ellsin)
    else:
        pos, w = _survey_selection(attrs, 100000, 23, scale=0.3)
        sel = ParticleField(to_dev(pos), to_dev(w), attrs=attrs).paint(resampler='cic', interlacing=1, compensate=False)
        wm = compute_mesh2_spectrum_window(sel, edgesin=edgesin, ellsin=ellsin, bin=bin, los=los, flags=('smooth',))
        want = po.spectrum_window_local_smooth(oattrs, obin, edgesin, ellsin=ellsin, mesh=host(sel.value).astype('f8'))
    assert_window_close(wm.value(), want, len(obin.nmodes), len(edgesin) - 1, label='local smooth')


@pytest.mark.parametrize('periodic,los', [(True, 'z'), (False, 'x'), (False, 'firstpoint'), (False, 'endpoint')])
def test_spectrum_mean(periodic, los):
    """compute_mesh2_spectrum_mean (mesh2.py:1485-1680) for a theory tabulated on k edges against
    the oracle: |d num_l| <= 1e-5 |num_l| + 1e-5 max|num_0|."""
    from jaxpower_amd import MeshAttrs, ParticleField, BinMesh2SpectrumPoles, compute_mesh2_spectrum_mean
    say('spectrum mean', periodic, los)
    center = (0., 0., 0.) if periodic else (900., 300., 500.)
    attrs = MeshAttrs(meshsize=16, boxsize=400., boxcenter=center)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.02}, ells=(0, 2, 4))
    rng = np.random.default_rng(3)
    kedges = np.arange(0., 0.25, 0.015)
    poles = {0: rng.uniform(1e3, 2e3, len(kedges) - 1), 2: rng.uniform(-5e2, 5e2, len(kedges) - 1)}
    oattrs = po.Attrs(16, 400., center)
    obin = po.BinPoles(oattrs, edges={'step': 0.02}, ells=(0, 2, 4))
    if periodic:
        got = compute_mesh2_spectrum_mean(attrs, theory=(kedges, poles), bin=bin, los=los)
        want = po.spectrum_mean(oattrs, obin, kedges, poles, edges=True, los=los)
    else:
        pos, w = _survey_selection(attrs, 100000, 29, scale=0.3)
        sel = ParticleField(to_dev(pos), to_dev(w), attrs=attrs).paint(resampler='cic', interlacing=1, compensate=False)
        got = compute_mesh2_spectrum_mean(sel, theory=(kedges, poles), bin=bin, los=los)
        want = po.spectrum_mean(oattrs, obin, kedges, poles, edges=True, mesh=host(sel.value).astype('f8'), los=los)
    floor = 1e-5 * np.nanmax(np.abs(want[0]))
    for ill, ell in enumerate((0, 2, 4)):
        g = got.get(ell).values('num_raw')
        assert np.all(np.abs(g - want[ill]) <= 1e-5 * np.abs(want[ill]) + floor), (ell, np.max(np.abs(g - want[ill])), floor)
    assert np.allclose(got.get(0).values('norm'), np.prod(oattrs.meshsize) / np.prod(oattrs.cellsize))


def test_spectrum_mean_callable_theory():
    """A theory given as callables of |k| (looked up at the mesh's distinct k^2) equals the same
    theory tabulated on k edges when the callables are that step function."""
    from jaxpower_amd import MeshAttrs, BinMesh2SpectrumPoles, compute_mesh2_spectrum_mean
    say('spectrum mean callable')
    attrs = MeshAttrs(meshsize=16, boxsize=(400., 300., 500.), boxcenter=0.)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.02}, ells=(0, 2))
    kedges = np.arange(0.005, 0.3, 0.015)
    rng = np.random.default_rng(4)
    vals = {0: rng.uniform(1e3, 2e3, len(kedges) - 1), 2: rng.uniform(-5e2, 5e2, len(kedges) - 1)}

    def step(values):
        def f(k):
            idx = np.digitize(k, kedges, right=False) - 1
            ok = (idx >= 0) & (idx <= len(kedges) - 2)
            return np.where(ok, values[np.where(ok, idx, 0)], 0.)
        return f

    a = compute_mesh2_spectrum_mean(attrs, theory=(kedges, vals), bin=bin, los='y')
    b = compute_mesh2_spectrum_mean(attrs, theory={ell: step(v) for ell, v in vals.items()}, bin=bin, los='y')
    for ell in (0, 2):
        x, y = a.get(ell).values('num_raw'), b.get(ell).values('num_raw')
        assert np.allclose(x, y, rtol=1e-6, atol=1e-6 * np.abs(a.get(0).values('num_raw')).max()), ell


def test_separation_binning_many_shells():
    """Separation binning with more shells than an LDS histogram holds (the unique shells of a
    128^3 mesh, the default of the smooth window) goes through FP64 global atomics: against the
    oracle's bincount; and the smooth window at the reference's test size (64^3,
    tests/test_mesh2.py:528-560)."""
    from jaxpower_amd import MeshAttrs, ParticleField, BinMesh2CorrelationPoles, BinMesh2SpectrumPoles, compute_mesh2_spectrum_window
    say('separation binning, unique shells at 128^3')
    big = MeshAttrs(meshsize=128, boxsize=500., boxcenter=1000.)
    sbin = BinMesh2CorrelationPoles(big, edges=None)
    osbin = po.BinCorrelationPoles(po.Attrs(128, 500., 1000.), edges=None)
    assert sbin.nbins > 8000 and np.array_equal(sbin.nmodes, osbin.nmodes)
    rng = np.random.default_rng(2)
    mesh = rng.normal(size=(128, 128, 128)).astype('f4')
    got = sbin(to_dev(mesh))
    want = osbin(mesh.astype('f8'))
    ok = osbin.nmodes > 0
    assert np.allclose(got[ok], want[ok], rtol=1e-6, atol=1e-6 * np.abs(want[ok]).max())
    attrs = MeshAttrs(meshsize=64, boxsize=500., boxcenter=1000.)
    oattrs = po.Attrs(64, 500., 1000.)
    pos, w = _survey_selection(attrs, 300000, 31)
    sel = ParticleField(to_dev(pos), to_dev(w), attrs=attrs).paint(resampler='cic', interlacing=1, compensate=False)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.01}, ells=(0, 2, 4))
    edgesin = np.array([[0.1, 0.11]])
    wm = compute_mesh2_spectrum_window(sel, edgesin=edgesin, ellsin=(0, 2, 4), bin=bin, los='x', flags=('smooth',))
    obin = po.BinPoles(oattrs, edges={'step': 0.01}, ells=(0, 2, 4))
    want = po.spectrum_window_smooth(oattrs, obin, edgesin, ellsin=(0, 2, 4), meshes=[host(sel.value).astype('f8')], los='x')
    assert_window_close(wm.value(), want, len(obin.nmodes), 1, label='smooth 64^3')


@pytest.mark.parametrize('periodic', [True, False])
def test_window_local_theory_infinite(periodic):
    """flags=('infinite',) with a local line of sight and theory line of sight 'local'
    (mesh2.py:1382-1413, 1446-1457): Gaunt-weighted selection terms, Bessel shells of order p,
    Y_lm(khat) binning, combined into the input multipoles; against the oracle, whose Y_lm and
    Gaunt coefficients follow another sign convention (only the convention-free result is compared)."""
    from jaxpower_amd import MeshAttrs, ParticleField, BinMesh2SpectrumPoles, compute_mesh2_spectrum_window
    say('window local theory infinite', periodic)
    center = (0., 0., 0.) if periodic else (900., 300., 500.)
    attrs = MeshAttrs(meshsize=16, boxsize=400., boxcenter=center)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.03}, ells=(0, 2, 4))
    edgesin = np.arange(0.01, 0.12, 0.025)
    oattrs = po.Attrs(16, 400., center)
    obin = po.BinPoles(oattrs, edges={'step': 0.03}, ells=(0, 2, 4))
    if periodic:
        wm = compute_mesh2_spectrum_window(attrs, edgesin=edgesin, ellsin=((0, 2, 4), 'local'), bin=bin, los='firstpoint',
                                           flags=('infinite',))
        want = po.spectrum_window_local_theory_infinite(oattrs, obin, edgesin, ellsin=(0, 2, 4))
    else:
        pos, w = _survey_selection(attrs, 100000, 37, scale=0.3)
        sel = ParticleField(to_dev(pos), to_dev(w), attrs=attrs).paint(resampler='cic', interlacing=1, compensate=False)
        wm = compute_mesh2_spectrum_window(sel, edgesin=edgesin, ellsin=((0, 2, 4), 'local'), bin=bin, los='firstpoint',
                                           flags=('infinite',))
        want = po.spectrum_window_local_theory_infinite(oattrs, obin, edgesin, ellsin=(0, 2, 4), mesh=host(sel.value).astype('f8'))
    assert_window_close(wm.value(), want, len(obin.nmodes), len(edgesin) - 1, label='local theory infinite')


@pytest.mark.parametrize('periodic', [True, False])
def test_window_local_theory_smooth(periodic):
    """flags=('smooth',) with the 'local' theory line of sight (mesh2.py:1346-1380, 1446-1457)
    against the oracle (another Y_lm and Gaunt sign convention)."""
    from jaxpower_amd import MeshAttrs, ParticleField, BinMesh2SpectrumPoles, compute_mesh2_spectrum_window
    say('window local theory smooth', periodic)
    center = (0., 0., 0.) if periodic else (900., 300., 500.)
    attrs = MeshAttrs(meshsize=16, boxsize=400., boxcenter=center)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.03}, ells=(0, 2, 4))
    edgesin = np.arange(0.01, 0.12, 0.025)
    oattrs = po.Attrs(16, 400., center)
    obin = po.BinPoles(oattrs, edges={'step': 0.03}, ells=(0, 2, 4))
    mesh = None
    inputs = [attrs]
    if not periodic:
        pos, w = _survey_selection(attrs, 100000, 41, scale=0.3)
        sel = ParticleField(to_dev(pos), to_dev(w), attrs=attrs).paint(resampler='cic', interlacing=1, compensate=False)
        inputs, mesh = [sel], host(sel.value).astype('f8')
    wm = compute_mesh2_spectrum_window(*inputs, edgesin=edgesin, ellsin=((0, 2, 4), 'local'), bin=bin, los='firstpoint', flags=('smooth',))
    want = po.spectrum_window_local_theory_smooth(oattrs, obin, edgesin, ellsin=(0, 2, 4), mesh=mesh)
    assert_window_close(wm.value(), want, len(obin.nmodes), len(edgesin) - 1, label='local theory smooth')


@pytest.mark.parametrize('periodic', [True, False])
def test_window_local_theory_mesh(periodic):
    """No flags with the 'local' theory line of sight (mesh2.py:1415-1457): shells weighted by
    Y_l1,m1(khat) Y_l2,m2(-khat), the stored selection terms, Y_lm(khat) binning, against the oracle."""
    from jaxpower_amd import MeshAttrs, ParticleField, BinMesh2SpectrumPoles, compute_mesh2_spectrum_window
    say('window local theory mesh', periodic)
    center = (0., 0., 0.) if periodic else (900., 300., 500.)
    attrs = MeshAttrs(meshsize=16, boxsize=400., boxcenter=center)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.03}, ells=(0, 2, 4))
    kf = 2. * np.pi / 400.
    edgesin = np.array([[kf, 3 * kf], [3 * kf, 5 * kf]])
    oattrs = po.Attrs(16, 400., center)
    obin = po.BinPoles(oattrs, edges={'step': 0.03}, ells=(0, 2, 4))
    mesh = None
    inputs = [attrs]
    if not periodic:
        pos, w = _survey_selection(attrs, 100000, 43, scale=0.3)
        sel = ParticleField(to_dev(pos), to_dev(w), attrs=attrs).paint(resampler='cic', interlacing=1, compensate=False)
        inputs, mesh = [sel], host(sel.value).astype('f8')
    wm = compute_mesh2_spectrum_window(*inputs, edgesin=edgesin, ellsin=((0, 2, 4), 'local'), bin=bin, los='firstpoint')
    want = po.spectrum_window_local_theory(oattrs, obin, edgesin, ellsin=(0, 2, 4), mesh=mesh)
    assert_window_close(wm.value(), want, len(obin.nmodes), len(edgesin), label='local theory mesh')


@pytest.mark.parametrize('flag', ['infinite', 'smooth', None])
def test_window_local_wide_angle(flag):
    """Wide-angle orders of the firstpoint theory (|x|^-wa and |s|^wa, mesh2.py:1244, 1269,
    1207-1209) for each theory flag, against the oracle; selection kept off the observer."""
    from jaxpower_amd import MeshAttrs, ParticleField, BinMesh2SpectrumPoles, compute_mesh2_spectrum_window
    say('window local wide angle', flag)
    center = (900., 300., 500.)
    attrs = MeshAttrs(meshsize=16, boxsize=400., boxcenter=center)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.03}, ells=(0, 2))
    kf = 2. * np.pi / 400.
    edgesin = np.array([[kf, 3 * kf], [3 * kf, 5 * kf]])
    ellsin = [(0, 0), (0, 1), (2, 1)]
    oattrs = po.Attrs(16, 400., center)
    obin = po.BinPoles(oattrs, edges={'step': 0.03}, ells=(0, 2))
    pos, w = _survey_selection(attrs, 100000, 47, scale=0.3)
    sel = ParticleField(to_dev(pos), to_dev(w), attrs=attrs).paint(resampler='cic', interlacing=1, compensate=False)
    mesh = host(sel.value).astype('f8')
    flags = (flag,) if flag else ()
    wm = compute_mesh2_spectrum_window(sel, edgesin=edgesin, ellsin=ellsin, bin=bin, los='firstpoint', flags=flags)
    fn = {'infinite': po.spectrum_window_local_infinite, 'smooth': po.spectrum_window_local_smooth, None: po.spectrum_window_local}[flag]
    want = fn(oattrs, obin, edgesin, ellsin=ellsin, mesh=mesh)
    assert_window_close(wm.value(), want, len(obin.nmodes), len(edgesin), label=f'wide angle {flag}')
    assert [p.attrs.get('wa_order') for p in wm.theory] == [0, 1, 1]


def test_spectrum_mean_wide_angle():
    """compute_mesh2_spectrum_mean with wide-angle theory terms (mesh2.py:1640-1645) against the oracle."""
    from jaxpower_amd import MeshAttrs, ParticleField, BinMesh2SpectrumPoles, compute_mesh2_spectrum_mean
    say('spectrum mean wide angle')
    center = (900., 300., 500.)
    attrs = MeshAttrs(meshsize=16, boxsize=400., boxcenter=center)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.02}, ells=(0, 2))
    rng = np.random.default_rng(5)
    kedges = np.arange(0., 0.25, 0.015)
    poles = {(0, 0): rng.uniform(1e3, 2e3, len(kedges) - 1), (1, 1): rng.uniform(-5e2, 5e2, len(kedges) - 1)}
    oattrs = po.Attrs(16, 400., center)
    obin = po.BinPoles(oattrs, edges={'step': 0.02}, ells=(0, 2))
    pos, w = _survey_selection(attrs, 100000, 53, scale=0.3)
    sel = ParticleField(to_dev(pos), to_dev(w), attrs=attrs).paint(resampler='cic', interlacing=1, compensate=False)
    got = compute_mesh2_spectrum_mean(sel, theory=(kedges, poles), bin=bin, los='firstpoint')
    want = po.spectrum_mean(oattrs, obin, kedges, poles, edges=True, mesh=host(sel.value).astype('f8'), los='firstpoint')
    floor = 1e-5 * np.nanmax(np.abs(want[0]))
    for ill, ell in enumerate((0, 2)):
        g = got.get(ell).values('num_raw')
        assert np.all(np.abs(g - want[ill]) <= 1e-5 * np.abs(want[ill]) + floor), (ell, np.max(np.abs(g - want[ill])), floor)


def test_spectrum_mean_local_theory():
    """compute_mesh2_spectrum_mean with the 'local' theory line of sight (mesh2.py:1647-1662)
    against the oracle (another Y_lm sign convention)."""
    from jaxpower_amd import MeshAttrs, ParticleField, BinMesh2SpectrumPoles, compute_mesh2_spectrum_mean
    say('spectrum mean local theory')
    center = (900., 300., 500.)
    attrs = MeshAttrs(meshsize=16, boxsize=400., boxcenter=center)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.02}, ells=(0, 2, 4))
    rng = np.random.default_rng(6)
    kedges = np.arange(0., 0.25, 0.015)
    poles = {0: rng.uniform(1e3, 2e3, len(kedges) - 1), 2: rng.uniform(-5e2, 5e2, len(kedges) - 1), 4: rng.uniform(-1e2, 1e2, len(kedges) - 1)}
    oattrs = po.Attrs(16, 400., center)
    obin = po.BinPoles(oattrs, edges={'step': 0.02}, ells=(0, 2, 4))
    pos, w = _survey_selection(attrs, 100000, 59, scale=0.3)
    sel = ParticleField(to_dev(pos), to_dev(w), attrs=attrs).paint(resampler='cic', interlacing=1, compensate=False)
    got = compute_mesh2_spectrum_mean(sel, theory=((kedges, poles), 'local'), bin=bin, los='firstpoint')
    want = po.spectrum_mean(oattrs, obin, kedges, poles, edges=True, mesh=host(sel.value).astype('f8'), los='firstpoint',
                            theory_los='local')
    floor = 1e-5 * np.nanmax(np.abs(want[0]))
    for ill, ell in enumerate((0, 2, 4)):
        g = got.get(ell).values('num_raw')
        assert np.all(np.abs(g - want[ill]) <= 1e-5 * np.abs(want[ill]) + floor), (ell, np.max(np.abs(g - want[ill])), floor)
