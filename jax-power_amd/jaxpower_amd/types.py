"""Power-spectrum and correlation-function multipole containers with the API jax-power uses
from lsstypes (``Mesh2SpectrumPole(s)``, constructed at jaxpower/mesh2.py:340-344, 370-374;
``Mesh2CorrelationPole(s)``, mesh2.py:467-468, 502-503; registered at jaxpower/types.py:10-51,
93-96).

A pole holds float64 arrays: k (mode-averaged |k|), k_edges (nbins, 2), nmodes (int64),
num_raw, num_shotnoise, norm, volume, and ``ell``.  ``value()`` is
(num_raw - num_shotnoise) / norm.  Fields computed on the GPU are read back to the host on
first access, so building and cloning a spectrum never waits for the device.
"""
import numpy as np


class Deferred:
    """A float64 scalar or array computed on the GPU and read back on first use, so that the
    estimator returns without waiting for the device (JAX arrays play this role in the
    reference).  Converts like an array (``np.asarray``, ``float``) and supports arithmetic
    through its host value."""

    __array_ufunc__ = None  # numpy binary operators defer to the reflected methods below

    def __init__(self, tensor):
        self._tensor, self._host = tensor, None

    def numpy(self):
        if self._host is None:
            self._host = self._tensor.detach().cpu().numpy()
            self._tensor = None
        return self._host

    def __array__(self, dtype=None, copy=None):
        a = self.numpy()
        return a.astype(dtype) if dtype is not None else a

    def __float__(self):
        return float(self.numpy())

    def __len__(self):
        return len(self.numpy())

    def __getitem__(self, item):
        return self.numpy()[item]

    @property
    def shape(self):
        return self.numpy().shape

    def __repr__(self):
        return f'Deferred({self.numpy()!r})'


def _binop(op, reflected=False):
    def fn(self, other):
        other = other.numpy() if isinstance(other, Deferred) else other
        return op(other, self.numpy()) if reflected else op(self.numpy(), other)
    return fn


import operator as _operator  # noqa: E402

for _name, _op in [('add', _operator.add), ('sub', _operator.sub), ('mul', _operator.mul), ('truediv', _operator.truediv),
                   ('pow', _operator.pow), ('lt', _operator.lt), ('le', _operator.le), ('gt', _operator.gt),
                   ('ge', _operator.ge), ('eq', _operator.eq), ('ne', _operator.ne)]:
    setattr(Deferred, f'__{_name}__', _binop(_op))
    if _name in ('add', 'sub', 'mul', 'truediv', 'pow'):
        setattr(Deferred, f'__r{_name}__', _binop(_op, reflected=True))
Deferred.__neg__ = lambda self: -self.numpy()
Deferred.__abs__ = lambda self: np.abs(self.numpy())
Deferred.__hash__ = None


def _host(x):
    if isinstance(x, np.ndarray):
        return x
    if isinstance(x, Deferred):
        return x.numpy()
    try:
        import torch
        if isinstance(x, torch.Tensor):
            return x.detach().cpu().numpy()
    except ImportError:  # pragma: no cover
        pass
    if isinstance(x, (list, tuple)) and any(isinstance(v, Deferred) for v in x):
        return np.asarray([_host(v) for v in x])
    return np.asarray(x)


def _format_ell(ell):
    """An int multipole, or a tuple of ints for the Sugiyama bispectrum's (l1, l2, L)."""
    if np.ndim(ell):
        return tuple(int(v) for v in np.ravel(ell))
    return int(ell)


def _ell_key(ell):
    return '_'.join(str(v) for v in ell) if isinstance(ell, tuple) else str(ell)


class _MeshPole:
    """One multipole of a binned two-point measurement; ``_x`` names its coordinate ('k' for
    spectra, 's' for correlation functions)."""

    _x = 'k'
    _values_names = ('num_raw', 'num_shotnoise', 'norm', 'nmodes', 'volume')

    def __init__(self, ell=0, attrs=None, **fields):
        names = self._field_names()
        unknown = set(fields) - set(names)
        if unknown:
            raise TypeError(f'{type(self).__name__}: unexpected fields {sorted(unknown)}')
        # fields may be device tensors or Deferred values: they are read back on first access
        self._raw = {name: fields.get(name) for name in names}
        self._cache = None
        self.ell = _format_ell(ell)
        self.attrs = dict(attrs or {})

    @classmethod
    def _field_names(cls):
        return (cls._x, cls._x + '_edges', 'nmodes', 'num_raw', 'num_shotnoise', 'norm', 'volume')

    @property
    def _data(self):
        if self._cache is None:
            r, x = self._raw, self._x
            xv = _host(r[x]).astype('f8')
            nbins = len(xv)

            def bcast(v, dtype='f8'):  # a copy of shape (nbins,)
                a = _host(v)
                if a.shape == (nbins,):
                    return a.astype(dtype)
                if a.ndim == 0:
                    return np.full(nbins, a, dtype=dtype)
                return np.broadcast_to(a.astype(dtype), (nbins,)).copy()

            nmodes = _host(r['nmodes'])
            self._cache = {x: xv, x + '_edges': _host(r[x + '_edges']).astype('f8'),
                           'nmodes': nmodes.astype('i8') if np.issubdtype(nmodes.dtype, np.integer) else bcast(nmodes),
                           'num_raw': bcast(r['num_raw']),
                           'num_shotnoise': bcast(0. if r['num_shotnoise'] is None else r['num_shotnoise']),
                           'norm': bcast(1. if r['norm'] is None else r['norm']),
                           'volume': bcast(0. if r['volume'] is None else r['volume'])}
        return self._cache

    def clone(self, **kwargs):
        state = dict(self._raw if self._cache is None else self._cache)
        state.update(ell=self.ell, attrs=self.attrs)
        state.update(kwargs)
        return type(self)(**state)

    def value(self):
        d = self._data
        return (d['num_raw'] - d['num_shotnoise']) / d['norm']

    def coords(self, name=None):
        return self._data[name or self._x]

    def edges(self, name=None):
        return self._data[(name or self._x) + '_edges']

    def values(self, name='value'):
        if name == 'value':
            return self.value()
        if name == 'shotnoise':
            return self._data['num_shotnoise'] / self._data['norm']
        return self._data[name]

    @property
    def shape(self):
        return self._data[self._x].shape

    def select(self, **kwargs):
        """Keep bins with xmin <= x < xmax (x given as (xmin, xmax) under the coordinate's
        name, e.g. k=(0.01, 0.2)) or a slice of bins."""
        sel = kwargs.pop(self._x, None)
        if kwargs:
            raise TypeError(f'select: unexpected arguments {sorted(kwargs)}')
        if sel is None:
            return self
        if isinstance(sel, slice):
            mask = np.zeros(self.shape, dtype=bool)
            mask[sel] = True
        else:
            xx = self._data[self._x]
            mask = (xx >= sel[0]) & (xx < sel[1])
        data = {name: v[mask] for name, v in self._data.items()}
        return type(self)(**data, ell=self.ell, attrs=self.attrs)

    def __repr__(self):
        return f'{type(self).__name__}(ell={self.ell}, nbins={len(self._data[self._x])})'


class Mesh2SpectrumPole(_MeshPole):
    """Power-spectrum multipole: k (mode-averaged |k|), k_edges, nmodes, num_raw,
    num_shotnoise, norm, volume, ell (mesh2.py:340-344, 370-374)."""

    _x = 'k'


class Mesh2CorrelationPole(_MeshPole):
    """Correlation-function multipole: s (cell-averaged |s|), s_edges, nmodes, num_raw,
    num_shotnoise, norm, volume, ell (mesh2.py:467-468, 502-503)."""

    _x = 's'


class Mesh3SpectrumPole(_MeshPole):
    """Bispectrum multipole (mesh3.py:625-627): k (ntri, 3) shell-averaged |k| of the three
    sides, k_edges (ntri, 3, 2), nmodes (closed-triangle counts, float), num_raw,
    num_shotnoise, norm, volume, ell, and the binning ``basis``."""

    _x = 'k'

    def __init__(self, ell=0, attrs=None, basis='scoccimarro', **fields):
        super().__init__(ell=ell, attrs=attrs, **fields)
        self.basis = basis

    def clone(self, **kwargs):
        kwargs.setdefault('basis', self.basis)
        return super().clone(**kwargs)

    def select(self, **kwargs):
        """Keep triangles whose three sides all satisfy kmin <= k < kmax (k=(kmin, kmax)),
        or a slice of triangles."""
        sel = kwargs.pop('k', None)
        if kwargs:
            raise TypeError(f'select: unexpected arguments {sorted(kwargs)}')
        if sel is None:
            return self
        if isinstance(sel, slice):
            mask = np.zeros(len(self._data['k']), dtype=bool)
            mask[sel] = True
        else:
            kk = self._data['k']
            mask = np.all((kk >= sel[0]) & (kk < sel[1]), axis=-1)
        data = {name: v[mask] for name, v in self._data.items()}
        return type(self)(**data, ell=self.ell, attrs=self.attrs, basis=self.basis)


class Mesh3CorrelationPole(Mesh3SpectrumPole):
    """Three-point correlation multipole (mesh3.py:723-725): s (npairs, 2) bin-averaged |s| of
    the two separations, s_edges (npairs, 2, 2), nmodes (products of the two cell counts),
    num_raw, num_shotnoise, norm, volume, ell (l1, l2, L) and the binning ``basis``."""

    _x = 's'

    def __init__(self, ell=0, attrs=None, basis='sugiyama', **fields):
        super().__init__(ell=ell, attrs=attrs, basis=basis, **fields)

    def select(self, **kwargs):
        """Keep pairs whose separations all satisfy smin <= s < smax (s=(smin, smax)), or a
        slice of pairs."""
        sel = kwargs.pop('s', None)
        if kwargs:
            raise TypeError(f'select: unexpected arguments {sorted(kwargs)}')
        if sel is None:
            return self
        if isinstance(sel, slice):
            mask = np.zeros(len(self._data['s']), dtype=bool)
            mask[sel] = True
        else:
            ss = self._data['s']
            mask = np.all((ss >= sel[0]) & (ss < sel[1]), axis=-1)
        data = {name: v[mask] for name, v in self._data.items()}
        return type(self)(**data, ell=self.ell, attrs=self.attrs, basis=self.basis)


class _MeshPoles:

    _pole = Mesh2SpectrumPole

    def __init__(self, poles, attrs=None):
        self._branches = list(poles)
        self.attrs = dict(attrs or (self._branches[0].attrs if self._branches else {}))

    @property
    def ells(self):
        return tuple(p.ell for p in self._branches)

    def get(self, ells=None, **kwargs):
        if ells is None:
            ells = kwargs.get('ell')
        if isinstance(ells, (list, tuple)) and not any(p.ell == ells for p in self._branches):
            return type(self)([self.get(ell) for ell in ells])  # a list of multipoles (a Sugiyama (l1, l2, L) matches first)
        for p in self._branches:
            if p.ell == ells:
                return p
        raise KeyError(f'ell = {ells} not in {self.ells}')

    def __iter__(self):
        return iter(self._branches)

    def __len__(self):
        return len(self._branches)

    def clone(self, **kwargs):
        """Per-pole fields given as lists (one entry per ell) or as one value for all."""
        poles = []
        for i, p in enumerate(self._branches):
            kw = {}
            for name, v in kwargs.items():
                kw[name] = v[i] if isinstance(v, (list, tuple)) else v
            poles.append(p.clone(**kw))
        return type(self)(poles, attrs=self.attrs)

    def value(self):
        return np.concatenate([p.value() for p in self._branches])

    def values(self, name='value'):
        return [p.values(name) for p in self._branches]

    def coords(self, name=None):
        return self._branches[0].coords(name)

    def edges(self, name=None):
        return self._branches[0].edges(name)

    def select(self, **kwargs):
        return type(self)([p.select(**kwargs) for p in self._branches], attrs=self.attrs)

    def write(self, fn, extra=None):
        """npz with one group of arrays per ell."""
        arrays = dict(extra or {})
        for p in self._branches:
            for name, v in p._data.items():
                arrays[f'ell{_ell_key(p.ell)}_{name}'] = v
        arrays['ells'] = np.array(self.ells)
        arrays['kind'] = np.array(type(self).__name__)
        los = self.attrs.get('los')
        if los is not None:
            arrays['los'] = np.asarray(los)
        np.savez(fn, **arrays)

    def __repr__(self):
        return f'{type(self).__name__}(ells={self.ells})'


class Mesh2SpectrumPoles(_MeshPoles):
    _pole = Mesh2SpectrumPole


class Mesh2CorrelationPoles(_MeshPoles):
    _pole = Mesh2CorrelationPole


class Mesh3SpectrumPoles(_MeshPoles):
    _pole = Mesh3SpectrumPole

    def write(self, fn):
        super().write(fn, extra={'basis': np.array(self._branches[0].basis if self._branches else 'scoccimarro')})


class Mesh3CorrelationPoles(Mesh3SpectrumPoles):
    _pole = Mesh3CorrelationPole


class WindowMatrix:
    """Window matrix W mapping a theory (input multipoles on input k bins) to the binned
    observable: observable = W . theory (mesh2.py:1453-1482).  ``observable`` holds the output
    poles (k, edges, nmodes, norm; num_raw zero), ``theory`` the input poles (k = mid-points of
    the input edges), ``value()`` the (n_ells n_bins, n_ellsin n_in) matrix, ordered by ell then
    bin on both sides."""

    def __init__(self, observable, theory, value):
        self.observable = observable
        self.theory = theory
        self._value = np.asarray(value)

    def value(self):
        return self._value

    @property
    def shape(self):
        return self._value.shape

    def dot(self, theory, return_type=None):
        """W . theory: ``theory`` is a poles container (its concatenated ``value()``) or a flat
        array; returns the observable poles with that value (num_raw = value x norm), or the
        flat array when return_type is 'array'."""
        vec = np.asarray(theory.value() if hasattr(theory, 'value') else theory, dtype='f8').ravel()
        out = self._value @ vec
        if return_type == 'array':
            return out
        poles, start = [], 0
        for p in self.observable:
            size = p.shape[0]
            norm = p.values('norm')
            poles.append(p.clone(num_raw=out[start:start + size] * norm, num_shotnoise=np.zeros(size)))
            start += size
        return type(self.observable)(poles, attrs=self.observable.attrs)

    def __repr__(self):
        return f'WindowMatrix(shape={self._value.shape}, ells={self.observable.ells}, ellsin={self.theory.ells})'


def read(fn):
    """Read a Mesh2SpectrumPoles, Mesh2CorrelationPoles, Mesh3SpectrumPoles or Mesh3CorrelationPoles
    written by ``write``."""
    with np.load(fn, allow_pickle=False) as data:
        kind = str(data['kind']) if 'kind' in data.files else 'Mesh2SpectrumPoles'
        cls = {'Mesh2SpectrumPoles': Mesh2SpectrumPoles, 'Mesh2CorrelationPoles': Mesh2CorrelationPoles,
               'Mesh3SpectrumPoles': Mesh3SpectrumPoles, 'Mesh3CorrelationPoles': Mesh3CorrelationPoles}[kind]
        attrs = {'los': data['los'].tolist()} if 'los' in data.files else {}
        extra = {'basis': str(data['basis'])} if 'basis' in data.files else {}
        poles = []
        for ell in data['ells']:
            ell = _format_ell(ell)
            kw = {name: data[f'ell{_ell_key(ell)}_{name}'] for name in cls._pole._field_names()}
            poles.append(cls._pole(**kw, ell=ell, attrs=attrs, **extra))
    return cls(poles, attrs=attrs)


def write(fn, observable):
    observable.write(fn)
