"""Power-spectrum multipole containers with the API jax-power uses from lsstypes
(``Mesh2SpectrumPole(s)``, constructed at jaxpower/mesh2.py:340-344, 370-374; registered at
jaxpower/types.py:10-51, 93-94).

A pole holds host float64 arrays: k (mode-averaged |k|), k_edges (nbins, 2), nmodes (int64),
num_raw, num_shotnoise, norm, volume, and ``ell``.  ``value()`` is
(num_raw - num_shotnoise) / norm.
"""
import numpy as np


def _host(x):
    try:
        import torch
        if isinstance(x, torch.Tensor):
            return x.detach().cpu().numpy()
    except ImportError:  # pragma: no cover
        pass
    return np.asarray(x)


class Mesh2SpectrumPole:

    _coords_names = ('k',)
    _values_names = ('num_raw', 'num_shotnoise', 'norm', 'nmodes', 'volume')

    def __init__(self, k=None, k_edges=None, nmodes=None, num_raw=None, num_shotnoise=None, norm=None, volume=None, ell=0,
                 attrs=None):
        nbins = len(_host(k))
        bcast = lambda v, dtype='f8': np.broadcast_to(_host(v).astype(dtype), (nbins,)).copy()
        self._data = dict(k=_host(k).astype('f8'), k_edges=_host(k_edges).astype('f8'),
                          nmodes=_host(nmodes).astype('i8') if np.issubdtype(_host(nmodes).dtype, np.integer) else bcast(nmodes),
                          num_raw=bcast(num_raw), num_shotnoise=bcast(0. if num_shotnoise is None else num_shotnoise),
                          norm=bcast(1. if norm is None else norm), volume=bcast(0. if volume is None else volume))
        self.ell = int(ell)
        self.attrs = dict(attrs or {})

    def clone(self, **kwargs):
        state = dict(self._data)
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
