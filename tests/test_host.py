"""CPU tests: the C-ABI library loads and exports every declared symbol, and the host-side
logic of the package (geometry, bin thresholds, containers, argument checks) matches the
oracle.  No GPU compute is called here."""
import itertools
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, 'include', 'jaxpower_amd.h')).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(mp_[a-z0-9_]+)\s*\(', text)))


def test_library_exports_every_declared_symbol():
    from jaxpower_amd import _lib
    lib = _lib.load()
    declared = _declared_symbols()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, f'{name} has no ctypes signature'
    assert lib.mp_version() == 1


def test_host_queries_without_device():
    from jaxpower_amd import _lib
    lib = _lib.load()
    assert lib.mp_spectrum_pitch(512) == 272 and lib.mp_spectrum_pitch(64) == 48 and lib.mp_spectrum_pitch(8) == 16
    n = _lib.i32x3([512, 512, 512])
    rows = lib.mp_paint_workspace_bytes(n, 16_000_000, 3, 3)
    assert rows >= 16_000_000 * 34  # row sort: two 16-byte records and a 2-byte key per particle
    assert lib.mp_paint_workspace_bytes(_lib.i32x3([6, 6, 6]), 1000, 3, 3) == 256  # atomic path needs none
    assert lib.mp_bin_workspace_bytes(3, 161) >= 1024 * 3 * 160 * 16
    # argument errors are reported without touching a device
    rc = lib.mp_bin_poles(None, 0, None, None, n, None, None, None, None, 1, None, None, 0, None, None, None, None, None, None, 0)
    assert rc == -1 and b'null' in lib.mp_last_error() or b'edges' in lib.mp_last_error()


def test_mesh_attrs_matches_oracle():
    from jaxpower_amd import MeshAttrs, get_mesh_attrs, next_fft_size
    from oracle import power_oracle as po
    a = MeshAttrs(meshsize=64, boxsize=1000., boxcenter=500.)
    b = po.Attrs(64, 1000., 500.)
    for name in ['meshsize', 'boxsize', 'boxcenter', 'cellsize', 'kfun', 'knyq']:
        assert np.allclose(getattr(a, name), getattr(b, name), rtol=0, atol=0)
    assert dict(**a)['meshsize'].tolist() == [64] * 3
    assert a.clone(meshsize=32).cellsize[0] == 1000. / 32
    with pytest.raises(NotImplementedError):
        MeshAttrs(meshsize=8, dtype='f8')
    with pytest.raises(NotImplementedError):
        MeshAttrs(meshsize=8, fft_backend='jaxdecomp')
    # tests/test_mesh.py:24-48 KATs
    assert np.all(get_mesh_attrs(boxsize=1000., boxcenter=0., cellsize=10.).meshsize == 100)
    rng = np.random.default_rng(84)
    pos = rng.uniform(0., 100., (100, 3))
    attrs = get_mesh_attrs(pos, cellsize=12., primes=(2, 3, 5), divisors=(2, 3))
    assert np.all(attrs.meshsize == 18) and np.allclose(attrs.cellsize, 12.)
    ref = po.get_mesh_attrs(pos, meshsize=128)
    attrs = get_mesh_attrs(pos, meshsize=128)
    assert np.allclose(attrs.boxsize, ref.boxsize) and np.allclose(attrs.boxcenter, ref.boxcenter)
    for n in range(1, 300):
        assert next_fft_size(n) == po.next_fft_size(n)


def test_regrid_matches_oracle():
    from jaxpower_amd import MeshAttrs
    from jaxpower_amd.mesh import get_halo_for_new_mattrs
    from oracle import power_oracle as po
    for L in (995., 1000., 1005., 1234.5):
        a = MeshAttrs(meshsize=64, boxsize=L, boxcenter=500.)
        new = get_halo_for_new_mattrs(a, cellsize=10.)[1]
        ref = po.regrid_attrs(po.Attrs(64, L, 500.), 10.)
        assert np.array_equal(new.meshsize, ref.meshsize) and np.allclose(new.boxsize, ref.boxsize)


def test_k2_thresholds_reproduce_float64_digitize():
    from jaxpower_amd.mesh2 import _k2_thresholds, _kvectors
    from jaxpower_amd import MeshAttrs
    from oracle import power_oracle as po
    for meshsize, boxsize, step in [(64, 1000., 0.01), (128, 2000., 0.005), ((32, 16, 64), (300., 150., 500.), 0.013)]:
        a = MeshAttrs(meshsize=meshsize, boxsize=boxsize)
        o = po.Attrs(meshsize, boxsize)
        edges = np.arange(0., a.kfun.min() * np.min(a.meshsize) / 2., step)
        thr = _k2_thresholds(edges)
        kx, ky, kz = _kvectors(a)
        for got, want in zip((kx, ky, kz), o.kcoords('wavenumber')):
            assert np.array_equal(got, want)
        k2 = (kx[:, None, None]**2 + ky[None, :, None]**2) + kz[None, None, :]**2
        ib = np.searchsorted(thr, k2.ravel(), side='right')
        assert np.array_equal(ib, np.digitize(np.sqrt(k2).ravel(), edges, right=False))
    # thresholds sit exactly at the float64 boundary
    e = np.array([0., 0.355, 0.71, 1e-3])
    t = _k2_thresholds(e)
    for ee, tt in zip(e[1:], t[1:]):
        assert np.sqrt(tt) >= ee and np.sqrt(np.nextafter(tt, 0.)) < ee


def test_spectrum_container():
    from jaxpower_amd import Mesh2SpectrumPole, Mesh2SpectrumPoles, read_observable
    k = np.array([0.1, 0.2, 0.3])
    poles = Mesh2SpectrumPoles([Mesh2SpectrumPole(k=k, k_edges=np.column_stack([k - 0.05, k + 0.05]), nmodes=np.array([1, 2, 3]),
                                                  num_raw=np.array([10., 20., 30.]) * (ell + 1), norm=2., ell=ell,
                                                  attrs={'los': 'z'}) for ell in (0, 2)])
    assert poles.ells == (0, 2)
    poles = poles.clone(norm=[4., 4.], num_shotnoise=[np.full(3, 2.), np.zeros(3)])
    assert np.allclose(poles.get(0).value(), (np.array([10., 20., 30.]) - 2.) / 4.)
    assert np.allclose(poles.get(ells=2).value(), np.array([30., 60., 90.]) / 4.)
    assert np.allclose(poles.get(0).values('shotnoise'), 0.5)
    assert poles.get(0).values('nmodes').dtype.kind == 'i'
    sel = poles.select(k=(0.15, 0.35))
    assert np.allclose(sel.get(0).coords('k'), [0.2, 0.3])


def test_spectrum_container_io(tmp_path):
    from jaxpower_amd import Mesh2SpectrumPole, Mesh2SpectrumPoles, read_observable
    k = np.array([0.1, 0.2])
    poles = Mesh2SpectrumPoles([Mesh2SpectrumPole(k=k, k_edges=np.column_stack([k - 0.05, k + 0.05]), nmodes=np.array([3, 4]),
                                                  num_raw=np.array([1., 2.]), norm=1., ell=0, attrs={'los': 'x'})])
    fn = tmp_path / 'p.npz'
    poles.write(fn)
    back = read_observable(fn)
    assert back.ells == (0,) and np.allclose(back.get(0).value(), [1., 2.]) and back.attrs['los'] == 'x'


def test_format_meshes():
    from jaxpower_amd.mesh import format_meshes
    m, f = format_meshes('a', nmeshes=2)
    assert m == ['a', 'a'] and f == (0, 0)
    m, f = format_meshes('a', 'b', nmeshes=2)
    assert m == ['a', 'b'] and f == (0, 1)


def test_deferred_values_in_containers():
    # device results (here CPU tensors) and Deferred scalars are read back only on access
    import torch
    from jaxpower_amd import Mesh2SpectrumPole
    from jaxpower_amd.types import Deferred
    k = np.array([0.1, 0.2, 0.3])
    norm = Deferred(torch.tensor(4., dtype=torch.float64))
    sn = Deferred(torch.full((3,), 2., dtype=torch.float64))
    pole = Mesh2SpectrumPole(k=torch.tensor(k), k_edges=np.column_stack([k - 0.05, k + 0.05]), nmodes=np.array([1, 2, 3]),
                             num_raw=torch.tensor([10., 20., 30.], dtype=torch.float64), norm=1., ell=0)
    assert pole._cache is None
    pole = pole.clone(norm=norm, num_shotnoise=sn)
    assert pole._cache is None and norm._host is None
    assert np.allclose(pole.value(), (np.array([10., 20., 30.]) - 2.) / 4.)
    assert float(norm) == 4. and np.isclose(norm, 4.) and norm * 2 == 8. and 1. / norm == 0.25 and np.allclose(sn + 1., 3.)
    assert np.asarray(sn).shape == (3,) and len(sn) == 3 and sn[1] == 2.


def test_mesh_and_particle_npz_round_trip(tmp_path):
    """save/load (mesh.py:1352-1450, 1783-1797, 2646-2670) on the host: values, attrs and extra
    fields come back; attrs are JSON, so loading needs no pickle."""
    import torch
    from jaxpower_amd import MeshAttrs, RealMeshField, ComplexMeshField, ParticleField
    attrs = MeshAttrs(meshsize=(4, 6, 8), boxsize=(40., 60., 80.), boxcenter=(1., 2., 3.))
    rng = np.random.default_rng(0)
    value = rng.normal(size=(4, 6, 8)).astype('f4')
    mesh = RealMeshField(torch.from_numpy(value), attrs=attrs, device='cpu')
    mesh.save(tmp_path / 'sub' / 'mesh.npz')
    back = RealMeshField.load(tmp_path / 'sub' / 'mesh.npz', device='cpu')
    assert np.array_equal(np.asarray(back), value)
    assert back.attrs == attrs
    cvalue = (rng.normal(size=(4, 6, 5)) + 1j * rng.normal(size=(4, 6, 5))).astype('c8')
    cmesh = ComplexMeshField(torch.from_numpy(cvalue), attrs=attrs, device='cpu')
    cmesh.save(tmp_path / 'cmesh.npz')
    assert np.array_equal(np.asarray(ComplexMeshField.load(tmp_path / 'cmesh.npz', device='cpu')), cvalue)
    pos = rng.uniform(0., 40., (100, 3)).astype('f4')
    p = ParticleField(torch.from_numpy(pos), torch.ones(100), attrs=attrs, extra={'id': torch.arange(100)}, device='cpu')
    p.save(tmp_path / 'particles.npz')
    q = ParticleField.load(tmp_path / 'particles.npz', device='cpu')
    assert np.array_equal(q.positions.numpy(), pos) and np.array_equal(q.extra['id'].numpy(), np.arange(100))
    assert q.attrs == attrs
    with pytest.raises(ValueError):
        mesh.save(tmp_path / 'mesh.h5')


def test_load_reference_style_file_needs_allow_pickle(tmp_path):
    """A file written the way jax-power writes it (attrs as a dict, pickled by numpy) loads only
    with an explicit allow_pickle=True."""
    from jaxpower_amd import RealMeshField
    value = np.arange(64, dtype='f4').reshape(4, 4, 4)
    state = dict(meshsize=(4, 4, 4), boxsize=(10., 10., 10.), boxcenter=(0., 0., 0.), dtype='<f4', fft_backend='auto')
    fn = tmp_path / 'ref.npz'
    np.savez(fn, value=value, attrs=state)
    with pytest.raises(ValueError, match='allow_pickle'):
        RealMeshField.load(fn, device='cpu')
    back = RealMeshField.load(fn, device='cpu', allow_pickle=True)
    assert np.array_equal(np.asarray(back), value)


def test_tophat_shell_xi_host_matches_oracle():
    """The library's float64 configuration-space shell (closed-form Bessel integrals, its own
    sine integral) against the oracle's (scipy's sici), across both sides of x = 0.1 and the
    sine-integral series/asymptotic switch at x = 20."""
    from jaxpower_amd import _lib
    from oracle import power_oracle as po
    lib = _lib.load()
    k0, k1 = 0.05, 0.08
    for ell in range(6):
        for s in np.concatenate([[0., 0.5, 1.3], np.geomspace(2., 2000., 57), [250., 251., 400.]]):
            got = lib.mp_tophat_shell_xi(ell, k0, k1, float(s))
            want = float(po.tophat_shell_xi(ell, k0, k1, s))
            scale = k1**3 / (2. * np.pi)**3 * 4. * np.pi * 1e-9  # ~1e-9 of the shell's peak value
            assert abs(got - want) <= 1e-9 * abs(want) + scale * (1. if ell <= 2 else 1e3), (ell, s, got, want)


def test_spherical_jn_and_shell_arrays_host():
    """The library's host arrays of j_ell and of the shell xi_ell (used by the smooth window)
    against the oracle's restatement of the reference's closed forms."""
    from jaxpower_amd import _lib
    from jaxpower_amd.mesh2 import _host_array_fn
    from oracle import power_oracle as po
    _lib.load()
    x = np.concatenate([np.linspace(0., 0.0999, 5), np.geomspace(0.2, 300., 40)])
    for ell in range(6):
        got = _host_array_fn('mp_spherical_jn_array', ell, x)
        want = po.spherical_jn(ell, x)
        # above x = 0.1 the closed forms sum terms up to ~(2l+1)!! x^-(l+1): rounding scales with them
        dfact = np.prod(np.arange(2 * ell + 1, 0, -2, dtype='f8'))
        atol = np.where(x > 0.1, 1e-14 * dfact / np.maximum(x, 0.1)**(ell + 1), 1e-15)
        assert np.all(np.abs(got - want) <= 1e-10 * np.abs(want) + atol), ell
        got = _host_array_fn('mp_tophat_shell_xi_array', ell, x * 10., 0.02, 0.05)
        want = po.tophat_shell_xi(ell, 0.02, 0.05, x * 10.)
        assert np.allclose(got, want, rtol=1e-8, atol=1e-9 * np.abs(want).max()), ell


def test_real_gaunt_host_invariants():
    """The library's real Gaunt coefficients (its own Y_lm convention): the same convention-free
    invariant as the oracle's, the m-independent G(0 0; l m; l m) = 1/sqrt(4 pi), and |G| equal to
    the oracle's coefficient by coefficient (conventions differ by signs only)."""
    from jaxpower_amd import _lib
    from oracle import power_oracle as po
    lib = _lib.load()
    for l in (2, 4):
        for m in range(-l, l + 1):
            assert abs(lib.mp_real_gaunt(0, 0, l, m, l, m) - 0.5 / np.sqrt(np.pi)) < 1e-6
    for l1, l2, l3 in [(0, 2, 2), (2, 2, 2), (4, 2, 2), (4, 4, 2)]:
        tot_lib = tot_or = 0.
        for a in range(-l1, l1 + 1):
            for b in range(-l2, l2 + 1):
                for c in range(-l3, l3 + 1):
                    g = lib.mp_real_gaunt(l1, a, l2, b, l3, c)
                    tot_lib += g * g
                    tot_or += po.real_gaunt(l1, a, l2, b, l3, c)**2
        assert abs(tot_lib - tot_or) < 1e-6 * max(tot_or, 1e-3), (l1, l2, l3, tot_lib, tot_or)


def test_shell_cap_positions_distribution():
    """generate_shell_cap_positions (the survey geometry of the FKP workload) against the law it
    states, on the CPU: r^3 uniform in [rmin^3, rmax^3], cos(theta) uniform in [cos 60deg, 1],
    phi uniform; shell_cap_volume against a Monte Carlo count in the bounding box."""
    from scipy import stats
    from jaxpower_amd.mock import generate_shell_cap_positions, shell_cap_volume
    pos = generate_shell_cap_positions(200_000, seed=3, device='cpu').numpy().astype('f8')
    r = np.linalg.norm(pos, axis=1)
    assert np.all((r > 1000. * (1. - 1e-6)) & (r < 2000. * (1. + 1e-6)))
    assert stats.kstest((r**3 - 1000.**3) / (2000.**3 - 1000.**3), 'uniform').pvalue > 1e-4
    assert stats.kstest((pos[:, 2] / r - 0.5) / 0.5, 'uniform').pvalue > 1e-4
    assert stats.kstest((np.arctan2(pos[:, 1], pos[:, 0]) + np.pi) / (2. * np.pi), 'uniform').pvalue > 1e-4
    rng = np.random.default_rng(4)
    u = rng.uniform(-2000., 2000., (2_000_000, 3))
    u[:, 2] = np.abs(u[:, 2])  # the cap is in z > 0
    ru = np.linalg.norm(u, axis=1)
    inside = (ru >= 1000.) & (ru < 2000.) & (u[:, 2] >= 0.5 * ru)
    frac, cube = inside.mean(), 4000.**2 * 2000.
    assert abs(shell_cap_volume() - frac * cube) < 5. * np.sqrt(frac * (1. - frac) / len(u)) * cube


def test_sugiyama_host_coefficients():
    """The package's Wigner 3j, triposh rows, complex-harmonic components and complex harmonic
    at a line of sight against the oracle's (pinned to sympy in test_oracle.py)."""
    from jaxpower_amd import mesh3
    from oracle import power_oracle as po
    for j1, j2, j3 in itertools.product(range(5), repeat=3):
        for m1 in range(-j1, j1 + 1):
            for m2 in range(-j2, j2 + 1):
                assert np.isclose(mesh3._wigner_3j(j1, j2, j3, m1, m2, -m1 - m2), po.wigner_3j(j1, j2, j3, m1, m2, -m1 - m2),
                                  atol=1e-14)
    for ell in [(0, 0, 0), (2, 0, 2), (1, 1, 2), (2, 2, 2), (4, 2, 2)]:
        for los_z in (False, True):
            got, want = mesh3._triposh_rows(*ell, los_z=los_z), po.triposh_rows(*ell, los_z=los_z)
            assert [r[:3] for r in got] == [r[:3] for r in want]
            assert np.allclose([r[3] for r in got], [r[3] for r in want], rtol=1e-14)
    rng = np.random.default_rng(3)
    u = rng.normal(size=(20, 3))
    for ell in range(5):
        for m in range(-ell, ell + 1):
            comp = mesh3._ylm_components(ell, m)
            got = sum(c * po.ref_real_ylm(ell, mu, *u.T) for mu, c in comp.items())
            assert np.allclose(got, po.complex_ylm_conj(ell, m, *u.T), atol=1e-13), (ell, m)
            for v in u[:3]:
                assert np.isclose(mesh3._complex_ylm_conj_at(ell, m, v), po.complex_ylm_conj(ell, m, *v), atol=1e-13)


def test_tuning_table_json_cache(tmp_path):
    """The autotuner's table imports a JSON cache entry by entry, exports it again, skips a
    cache measured on another device type and rejects malformed entries (no device needed)."""
    import json
    from jaxpower_amd import tuning
    tuning.clear()
    try:
        doc = {'format': tuning._FORMAT, 'device': 'gfx950', 'entries': {
            'xpass_cols L=512 ny=512 ncols=257 interlacing=3 products=3 catalogues=1': {'knob': 'xpass_cols', 'choice': 2},
            'sort_chunk nx=512 ny=512 positions=f32': {'knob': 'sort_chunk', 'choice': 1}}}
        path = tmp_path / 'cache.json'
        path.write_text(json.dumps(doc))
        assert tuning.load(str(path), device='gfx942') == 0 and tuning.table() == {}
        assert tuning.load(str(path), device='gfx950') == 2
        table = tuning.table()
        assert {k: (e['knob'], e['choice'], e['measured']) for k, e in table.items()} == {
            k: (e['knob'], e['choice'], False) for k, e in doc['entries'].items()}
        out = tmp_path / 'out.json'
        assert tuning.save(str(out)) == 2
        assert json.loads(out.read_text())['entries'] == table
        bad = dict(doc, entries={'x': {'knob': 'nope', 'choice': 0}})
        path.write_text(json.dumps(bad))
        with pytest.raises(ValueError):
            tuning.load(str(path), device='gfx950')
        with pytest.raises(ValueError):
            path.write_text(json.dumps(dict(doc, format=99)))
            tuning.load(str(path), device='gfx950')
    finally:
        tuning.clear()
