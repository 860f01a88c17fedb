"""Power-spectrum multipole containers with the API jax-power uses from lsstypes
(``Mesh2SpectrumPole(s)``, constructed at jaxpower/mesh2.py:340-344, 370-374; registered at
jaxpower/types.py:10-51, 93-94).

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


class Mesh2SpectrumPole:

    _coords_names = ('k',)
    _values_names = ('num_raw', 'num_shotnoise', 'norm', 'nmodes', 'volume')

    def __init__(self, k=None, k_edges=None, nmodes=None, num_raw=None, num_shotnoise=None, norm=None, volume=None, ell=0,
                 attrs=None):
        # fields may be device tensors or Deferred values: they are read back on first access
        self._raw = dict(k=k, k_edges=k_edges, nmodes=nmodes, num_raw=num_raw, num_shotnoise=num_shotnoise, norm=norm,
                         volume=volume)
        self._cache = None
        self.ell = int(ell)
        self.attrs = dict(attrs or {})

    @property
    def _data(self):
        if self._cache is None:
            r = self._raw
            k = _host(r['k']).astype('f8')
            nbins = len(k)
            bcast = lambda v, dtype='f8': np.broadcast_to(_host(v).astype(dtype), (nbins,)).copy()
            nmodes = _host(r['nmodes'])
            self._cache = dict(k=k, k_edges=_host(r['k_edges']).astype('f8'),
                               nmodes=nmodes.astype('i8') if np.issubdtype(nmodes.dtype, np.integer) else bcast(nmodes),
                               num_raw=bcast(r['num_raw']),
                               num_shotnoise=bcast(0. if r['num_shotnoise'] is None else r['num_shotnoise']),
                               norm=bcast(1. if r['norm'] is None else r['norm']),
                               volume=bcast(0. if r['volume'] is None else r['volume']))
        return self._cache

    def clone(self, **kwargs):
        state = dict(self._raw if self._cache is None else self._cache)
        state.update(ell=self.ell, attrs=self.attrs)
        state.update(kwargs)
        return type(self)(**state)

    def value(self):
        d = self._data
        return (d['num_raw'] - d['num_shotnoise']) / d['norm']

    def coords(self, name='k'):
        return self._data[name]

    def edges(self, name='k'):
        return self._data[name + '_edges']

    def values(self, name='value'):
        if name == 'value':
            return self.value()
        if name == 'shotnoise':
            return self._data['num_shotnoise'] / self._data['norm']
        return self._data[name]

    @property
    def shape(self):
        return self._data['k'].shape

    def select(self, k=None):
        """Keep bins with kmin <= k < kmax (k given as (kmin, kmax)) or a slice of bins."""
        if k is None:
            return self
        if isinstance(k, slice):
            mask = np.zeros(self.shape, dtype=bool)
            mask[k] = True
        else:
            kk = self._data['k']
            mask = (kk >= k[0]) & (kk < k[1])
        data = {name: v[mask] for name, v in self._data.items()}
        return type(self)(**data, ell=self.ell, attrs=self.attrs)

    def __repr__(self):
        return f'Mesh2SpectrumPole(ell={self.ell}, nbins={len(self._data["k"])})'


class Mesh2SpectrumPoles:

    def __init__(self, poles, attrs=None):
        self._branches = list(poles)
        self.attrs = dict(attrs or (self._branches[0].attrs if self._branches else {}))

    @property
    def ells(self):
        return tuple(p.ell for p in self._branches)

    def get(self, ells=None, **kwargs):
        if ells is None:
            ells = kwargs.get('ell')
        if isinstance(ells, (list, tuple)):
            return Mesh2SpectrumPoles([self.get(ell) for ell in ells])
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

    def coords(self, name='k'):
        return self._branches[0].coords(name)

    def edges(self, name='k'):
        return self._branches[0].edges(name)

    def select(self, **kwargs):
        return type(self)([p.select(**kwargs) for p in self._branches], attrs=self.attrs)

    def write(self, fn):
        """npz with one group of arrays per ell."""
        arrays = {}
        for p in self._branches:
            for name, v in p._data.items():
                arrays[f'ell{p.ell}_{name}'] = v
        arrays['ells'] = np.array(self.ells)
        los = self.attrs.get('los')
        if los is not None:
            arrays['los'] = np.asarray(los)
        np.savez(fn, **arrays)

    def __repr__(self):
        return f'Mesh2SpectrumPoles(ells={self.ells})'


def read(fn):
    """Read a Mesh2SpectrumPoles written by ``write``."""
    with np.load(fn, allow_pickle=False) as data:
        attrs = {'los': data['los'].tolist()} if 'los' in data.files else {}
        poles = []
        for ell in data['ells']:
            kw = {name: data[f'ell{ell}_{name}'] for name in ['k', 'k_edges', 'nmodes', 'num_raw', 'num_shotnoise', 'norm', 'volume']}
            poles.append(Mesh2SpectrumPole(**kw, ell=int(ell), attrs=attrs))
    return Mesh2SpectrumPoles(poles, attrs=attrs)


def write(fn, observable):
    observable.write(fn)
