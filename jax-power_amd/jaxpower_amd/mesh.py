"""Mesh geometry, mesh fields, particles and painting on one MI355X.

Mirrors the single-device part of jaxpower/mesh.py: ``MeshAttrs`` (1453-1720),
``RealMeshField`` / ``ComplexMeshField`` (1723-2196), ``get_mesh_attrs`` and
``next_fft_size`` (2199-2381), ``ParticleField`` (2397-2670), ``_paint`` (2673-2754),
``FKPField`` (2913-2999) and the normalisation helpers (3002-3129).

Field values are torch tensors on the GPU (torch provides the device memory and the stream);
painting, FFTs, readout and FP64 reductions run in libjaxpower_amd.so.  Half spectra are
(N0, N1, N2//2+1) complex64 views of buffers whose rows are padded to a multiple of 16
elements, so every row starts on a 128-byte boundary.
"""
import json
import math
import operator
import os
from functools import lru_cache

import numpy as np
import torch

from . import _lib
from . import resamplers


def _save_npz(fn, state, attrs):
    """npz with the arrays of ``state`` and the attributes (mesh.py:1352-1390): the attributes
    are stored as a JSON string, so reading them back executes nothing.  h5py is not part of
    this build, so only '.npz' is written."""
    fn = str(fn)
    if not fn.endswith('.npz'):
        raise ValueError(f'extension of {fn!r} not supported here (npz only)')
    dirname = os.path.dirname(fn)
    if dirname:
        os.makedirs(dirname, exist_ok=True)
    arrays = {}
    for name, value in state.items():
        arrays[name] = value.detach().cpu().numpy() if isinstance(value, torch.Tensor) else np.asarray(value)
    arrays['attrs'] = np.array(json.dumps(attrs))
    np.savez(fn, **arrays)


def _load_npz(fn, allow_pickle=False):
    """(arrays, attrs) of a file written by ``_save_npz``.  Files written by the reference keep
    their attributes as a pickled dict (mesh.py:1413-1416): those are read only with an explicit
    allow_pickle=True, for files from a trusted source."""
    fn = str(fn)
    if not fn.endswith('.npz'):
        raise ValueError(f'extension of {fn!r} not supported here (npz only)')
    try:
        with np.load(fn, allow_pickle=allow_pickle) as data:
            raw = data['attrs']
            attrs = json.loads(str(raw)) if raw.dtype.kind == 'U' else raw[()]
            arrays = {name: data[name] for name in data.files if name != 'attrs'}
    except ValueError as exc:
        if not allow_pickle and 'allow_pickle' in str(exc):
            raise ValueError(f'{fn}: attributes are pickled (a file written by jax-power); pass allow_pickle=True '
                             'only for a trusted file') from exc
        raise
    return arrays, dict(attrs)


def _attrs_from_state(state):
    keys = ('meshsize', 'boxsize', 'boxcenter', 'dtype', 'fft_backend')
    return MeshAttrs.from_state({k: state[k] for k in keys if k in state})


def _default_device():
    return torch.device('cuda', torch.cuda.current_device()) if torch.cuda.is_available() else torch.device('cpu')


def _fill3(value, dtype='f8'):
    return np.broadcast_to(np.asarray(value, dtype=dtype), (3,)).copy()


def _as_device_tensor(value, dtype=None, device=None):
    if isinstance(value, torch.Tensor):
        t = value
        if device is not None and t.device != device:
            t = t.to(device)
        if dtype is not None and t.dtype != dtype:
            t = t.to(dtype)
        return t
    if isinstance(value, BaseMeshField):
        return _as_device_tensor(value.value, dtype=dtype, device=device)
    arr = np.asarray(value)
    t = torch.from_numpy(np.ascontiguousarray(arr)).to(device or _default_device())
    if dtype is not None:
        t = t.to(dtype)
    return t


# ----------------------------------------------------------------------------- geometry

class MeshAttrs:
    """Mesh attributes (jaxpower/mesh.py:1453-1720).

    meshsize, boxsize, boxcenter are host float64 / int arrays of length 3.  ``dtype`` is the
    real mesh dtype; only float32 is supported (the reference's default without x64).
    ``fft_backend`` 'auto', 'jax' and 'hip' all select this library's hermitian r2c FFT.
    """

    __annotations__ = {'meshsize': np.ndarray, 'boxsize': np.ndarray, 'boxcenter': np.ndarray, 'dtype': object,
                       'fft_backend': str}

    def __init__(self, meshsize=None, boxsize=None, boxcenter=0., dtype=float, fft_backend='auto'):
        if meshsize is not None:
            meshsize = _fill3(meshsize, 'i4')
            if boxsize is None:
                boxsize = 1. * meshsize
        dt = np.dtype(dtype)
        if dt == np.float64 and dtype is float:
            dt = np.dtype('f4')  # python float -> float32 (no x64 on this path)
        if dt.kind == 'c':
            dt = np.dtype('f4') if dt.itemsize == 8 else dt
        if dt != np.dtype('f4'):
            raise NotImplementedError(f'mesh dtype {dt} not supported: meshes are float32 (bin sums, normalisation '
                                      'and shot noise are accumulated in float64)')
        if fft_backend not in ('auto', 'jax', 'hip'):
            raise NotImplementedError(f"fft_backend {fft_backend!r}: only the hermitian r2c layout ('auto', 'jax', 'hip')")
        self.__dict__.update(meshsize=meshsize, boxsize=None if boxsize is None else _fill3(boxsize),
                             boxcenter=None if boxcenter is None else _fill3(boxcenter), dtype=dt, fft_backend='hip')

    def __setattr__(self, name, value):
        raise AttributeError('MeshAttrs is immutable; use clone()')

    # mapping protocol, so that dict(**attrs) works (mesh.py:1520-1527)
    def keys(self):
        return ['meshsize', 'boxsize', 'boxcenter', 'dtype', 'fft_backend']

    def __getitem__(self, key):
        return getattr(self, key)

    def __eq__(self, other):
        return isinstance(other, MeshAttrs) and all(np.array_equal(self[k], other[k]) for k in ['meshsize', 'boxsize', 'boxcenter']) \
            and self.dtype == other.dtype

    def __hash__(self):
        return hash((tuple(self.meshsize), tuple(self.boxsize), tuple(self.boxcenter)))

    def __repr__(self):
        return (f'MeshAttrs(meshsize={self.meshsize.tolist()}, boxsize={self.boxsize.tolist()}, '
                f'boxcenter={self.boxcenter.tolist()}, dtype={self.dtype})')

    def clone(self, **kwargs):
        state = {k: self[k] for k in self.keys()}
        state.update(kwargs)
        return MeshAttrs(**state)

    def __getstate__(self):
        return dict(meshsize=tuple(self.meshsize.tolist()), boxsize=tuple(self.boxsize.tolist()),
                    boxcenter=tuple(self.boxcenter.tolist()), dtype=self.dtype.str, fft_backend=self.fft_backend)

    @classmethod
    def from_state(cls, state):
        return cls(**state)

    @property
    def rdtype(self):
        return self.dtype

    @property
    def cdtype(self):
        return np.dtype('c8')

    @property
    def is_real(self):
        return True

    @property
    def is_hermitian(self):
        return True

    @property
    def ndim(self):
        return 3

    @property
    def cellsize(self):
        return self.boxsize / self.meshsize

    @property
    def kfun(self):
        return 2. * np.pi / self.boxsize

    @property
    def knyq(self):
        return np.pi * self.meshsize / self.boxsize

    @property
    def sharding_mesh(self):
        return None

    def rcoords(self, kind='position', sparse=None, device=None):
        """Node coordinates (mesh.py:1573-1603): 'position' or 'index'."""
        device = device or _default_device()
        if kind == 'index':
            toret = [torch.arange(int(n), dtype=torch.int32, device=device) for n in self.meshsize]
        else:
            offset = self.boxcenter - self.boxsize / 2.
            toret = [(torch.arange(int(n), dtype=torch.float64, device=device) * float(d) + float(o)).to(torch.float32)
                     for n, d, o in zip(self.meshsize, self.cellsize, offset)]
        return _sparsify(toret, sparse)

    xcoords = rcoords

    def kcoords(self, kind='wavenumber', sparse=None, device=None):
        """Half-spectrum Fourier coordinates (mesh.py:1605-1636): 'wavenumber', 'circular' or
        'index'; the last axis uses rfftfreq."""
        device = device or _default_device()
        toret = []
        for axis, n in enumerate(self.meshsize):
            n = int(n)
            f = np.fft.rfftfreq(n) if axis == 2 else np.fft.fftfreq(n)
            if kind in ('wavenumber', 'separation'):
                k = f * self.kfun[axis] * n
            elif kind == 'circular':
                k = f * (2. * np.pi / n) * n
            elif kind == 'index':
                k = np.round(f * n)
            else:
                raise ValueError(f'unknown kind {kind}')
            toret.append(torch.from_numpy(k.astype('f4' if kind != 'index' else 'i4')).to(device))
        return _sparsify(toret, sparse)

    ccoords = kcoords

    def create(self, kind='real', fill=None, device=None):
        """New mesh (mesh.py:1638-1677); complex meshes are padded-pitch half spectra."""
        device = device or _default_device()
        name = kind if isinstance(kind, str) else kind.__name__.lower()
        if 'complex' in name:
            value = _empty_spectrum(self.meshsize, device)
            cls = ComplexMeshField
        else:
            value = torch.empty(tuple(int(n) for n in self.meshsize), dtype=torch.float32, device=device)
            cls = RealMeshField
        if fill is not None:
            if callable(fill):
                value.copy_(_as_device_tensor(fill(tuple(value.shape)), device=device))
            elif isinstance(fill, (torch.Tensor, np.ndarray)) and tuple(fill.shape) == tuple(value.shape):
                value.copy_(_as_device_tensor(fill, device=device))
            else:
                value.fill_(fill)
        return cls(value, attrs=self)

    def r2c(self, value):
        return _r2c(value, self.meshsize)

    def c2r(self, value):
        return _c2r(value, self.meshsize)


def _sparsify(coords, sparse):
    if sparse is None:
        return tuple(coords)
    shapes = [[-1 if i == a else 1 for i in range(3)] for a in range(3)]
    toret = [c.reshape(s) for c, s in zip(coords, shapes)]
    if sparse is False:
        toret = list(torch.broadcast_tensors(*toret))
    return tuple(toret)


def _empty_spectrum(meshsize, device):
    n0, n1, n2 = (int(n) for n in meshsize)
    pitch = int(_lib.load().mp_spectrum_pitch(n2))
    buf = torch.empty((n0, n1, pitch), dtype=torch.complex64, device=device)
    return buf[:, :, :n2 // 2 + 1]


def _spectrum_pitch(value):
    """Row pitch (elements) of a (N0, N1, N2//2+1) complex64 tensor usable by the kernels."""
    if value.dtype != torch.complex64:
        raise ValueError(f'half spectra must be complex64, got {value.dtype}')
    s0, s1, s2 = value.stride()
    if s2 != 1 or s1 < value.shape[2] or s0 != s1 * value.shape[1]:
        return None
    return s1


def _as_spectrum(value):
    """Return (tensor, pitch) with a layout the kernels accept (copying if needed)."""
    pitch = _spectrum_pitch(value)
    if pitch is None:
        value = value.contiguous()
        pitch = value.shape[2]
    return value, pitch


def _r2c(value, meshsize):
    _lib.require_device(value, 'mesh')
    value = value.contiguous()
    if value.dtype != torch.float32:
        value = value.to(torch.float32)
    out = _empty_spectrum(meshsize, value.device)
    _lib.call('mp_rfft3d', _lib.stream_handle(value.device), _lib.ptr(value), _lib.ptr(out), _lib.i32x3(meshsize),
              _spectrum_pitch(out))
    return out


def _c2r(value, meshsize):
    _lib.require_device(value, 'mesh')
    value, pitch = _as_spectrum(value)
    out = torch.empty(tuple(int(n) for n in meshsize), dtype=torch.float32, device=value.device)
    n = _lib.i32x3(meshsize)
    ws = _lib.workspace(_lib.load().mp_irfft3d_workspace_bytes(n), value.device)
    _lib.call('mp_irfft3d', _lib.stream_handle(value.device), _lib.ptr(value), pitch, _lib.ptr(out), n, _lib.ptr(ws),
              ws.numel())
    return out


@lru_cache(maxsize=32, typed=False)
def next_fft_size(n, primes=(2, 3, 5, 7), divisors=()):
    """Smallest m >= n with prime factors in primes, divisible by divisors (mesh.py:2226-2262)."""
    n = int(n)
    primes = tuple(p for p in primes if p > 1)
    divisors = tuple(int(d) for d in divisors if d > 1)
    D = 1
    for d in divisors:
        D = D * d // math.gcd(D, d)
    m = ((n + D - 1) // D) * D
    if not primes:
        return m
    while True:
        x = m
        for p in primes:
            while x % p == 0:
                x //= p
        if x == 1:
            return m
        m += D


def next_fft_shape(shape, primes=(2, 3, 5, 7)):
    return tuple(next_fft_size(n, primes) for n in shape)


def fftfreq(shape, kind='separation', sparse=None, hermitian=False, spacing=1., dtype=None, device=None, **kwargs):
    """Mesh frequencies (mesh.py:360-403)."""
    device = device or _default_device()
    spacing = _fill3(spacing) if len(shape) == 3 else np.broadcast_to(spacing, (len(shape),))
    toret = []
    for axis, s in enumerate(shape):
        if kind == 'position':
            k = np.arange(s) * spacing[axis]
        else:
            f = np.fft.rfftfreq(s) if (axis == len(shape) - 1 and hermitian) else np.fft.fftfreq(s)
            k = f * spacing[axis] * s
        toret.append(torch.from_numpy(np.asarray(k, dtype=dtype or 'f4')).to(device))
    if len(shape) == 3:
        return _sparsify(toret, sparse)
    return tuple(toret)


def _extent(*positions):
    nonempty = [p for p in positions if p.shape[0]]
    if not nonempty:
        raise ValueError('<= 1 particles found; cannot infer boxsize')
    pmin = np.min([_to_numpy(p.min(dim=0).values if isinstance(p, torch.Tensor) else p.min(axis=0)) for p in nonempty], axis=0)
    pmax = np.max([_to_numpy(p.max(dim=0).values if isinstance(p, torch.Tensor) else p.max(axis=0)) for p in nonempty], axis=0)
    return pmin.astype('f8'), pmax.astype('f8')


def _to_numpy(x):
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    return np.asarray(x)


def get_mesh_attrs(*positions, meshsize=None, boxsize=None, boxcenter=None, cellsize=None, boxpad=2., check=False,
                   approximate=False, dtype=None, primes=None, divisors=None, **kwargs):
    """Enclosing box and mesh size (mesh.py:2270-2381)."""
    if cellsize is not None and meshsize is not None:
        if boxsize is not None:
            raise ValueError('cannot specify boxsize, cellsize and meshsize simultaneously')
        meshsize = _fill3(meshsize, 'i4')
        boxsize = meshsize * _fill3(cellsize)
    if boxsize is None or boxcenter is None or (positions and check):
        if not positions:
            raise ValueError('positions must be provided if boxsize and boxcenter are not specified, or check is True')
        pos_min, pos_max = _extent(*positions)
        delta = np.abs(pos_max - pos_min)
        if boxcenter is None:
            boxcenter = 0.5 * (pos_min + pos_max)
        if boxsize is None:
            boxsize = delta.max() * boxpad
        if check and (np.asarray(boxsize) < delta).any():
            raise ValueError(f'boxsize {boxsize} too small to contain all data (max {delta})')
    boxsize = _fill3(boxsize)
    toret = {}
    if meshsize is None:
        if cellsize is None:
            raise ValueError('meshsize (or cellsize) must be specified')
        cellsize = _fill3(cellsize)
        meshsize = np.ceil(boxsize / cellsize).astype('i4')
        primes = tuple(primes or ())
        divisors = divisors or ()
        if not divisors or np.ndim(divisors[0]) == 0:
            divisors = [tuple(divisors)] * 3
        meshsize = np.array([next_fft_size(m, primes=primes, divisors=tuple(d)) for m, d in zip(meshsize, divisors)])
        toret['boxsize'] = meshsize * cellsize
        if not positions and not approximate and not np.allclose(toret['boxsize'], boxsize):
            raise ValueError(f"cannot enforce cellsize = {cellsize} with meshsize = {meshsize}, as it would lead to "
                             f"boxsize = {toret['boxsize']} != input boxsize = {boxsize}")
    else:
        meshsize = _fill3(meshsize, 'i4')
    toret.setdefault('boxsize', boxsize)
    kw = dict(kwargs)
    kw.pop('sharding_mesh', None)
    return MeshAttrs(meshsize=meshsize, boxsize=toret['boxsize'], boxcenter=_fill3(boxcenter), dtype=dtype or float, **kw)


# ----------------------------------------------------------------------------- mesh fields

class _UpdateRef:

    def __init__(self, field, item):
        self.field, self.item = field, item

    def _op(self, name, other):
        field = self.field
        if (name == 'set' and isinstance(field, ComplexMeshField) and field._pending_combine() is not None
                and _is_origin(self.item) and _is_zero_scalar(other)):
            # mesh.at[0, 0, 0].set(0.) on a mesh whose interlacing combine is still pending:
            # the combine (or its fused binning) zeroes the mode itself
            return field._with_pending(field._pending_combine().with_zero_dc())
        value = field.value.clone()
        other = other.value if isinstance(other, BaseMeshField) else other
        view = value[self.item]
        if name == 'set':
            view = torch.as_tensor(other, device=value.device, dtype=value.dtype) + torch.zeros_like(view)
        elif name == 'add':
            view = view + other
        elif name == 'multiply':
            view = view * other
        elif name == 'divide':
            view = view / other
        elif name == 'power':
            view = view**other
        elif name == 'min':
            view = torch.minimum(view, torch.as_tensor(other, device=value.device, dtype=value.dtype))
        elif name == 'max':
            view = torch.maximum(view, torch.as_tensor(other, device=value.device, dtype=value.dtype))
        value[self.item] = view
        return self.field.clone(value=value)

    def set(self, other):
        return self._op('set', other)

    def add(self, other):
        return self._op('add', other)

    def multiply(self, other):
        return self._op('multiply', other)

    def divide(self, other):
        return self._op('divide', other)

    def power(self, other):
        return self._op('power', other)

    def min(self, other):
        return self._op('min', other)

    def max(self, other):
        return self._op('max', other)


def _is_origin(item):
    if not isinstance(item, tuple) or len(item) != 3:
        return False
    try:
        return all(int(i) == 0 and not isinstance(i, (slice, bool)) for i in item)
    except TypeError:
        return False


def _is_zero_scalar(other):
    if isinstance(other, torch.Tensor):
        return other.numel() == 1 and not other.is_cuda and complex(other.item()) == 0
    try:
        return complex(other) == 0
    except TypeError:
        return False


class _UpdateHelper:

    def __init__(self, field):
        self.field = field

    def __getitem__(self, item):
        return _UpdateRef(self.field, item)


class BaseMeshField:
    """A field as a 3-D array with its MeshAttrs (mesh.py:1723-1797)."""

    def __init__(self, value, *args, **kwargs):
        if 'attrs' in kwargs:
            args = args + (kwargs.pop('attrs'),)
        is_complex = 'complex' in type(self).__name__.lower()
        value = _as_device_tensor(value, device=kwargs.pop('device', None))
        if not is_complex and value.dtype == torch.float64:
            value = value.to(torch.float32)
        if is_complex and value.dtype == torch.complex128:
            value = value.to(torch.complex64)
        if args:
            attrs = args[0].clone(**kwargs) if kwargs else args[0]
        else:
            meshsize = kwargs.pop('meshsize', None)
            if meshsize is None:
                meshsize = tuple(value.shape)
                if is_complex:
                    raise ValueError('give attrs (or meshsize) for a complex mesh')
            attrs = MeshAttrs(meshsize=meshsize, **kwargs)
        self.__dict__.update(value=value, attrs=attrs)

    def __setattr__(self, name, value):
        raise AttributeError(f'{type(self).__name__} is immutable; use clone()')

    def clone(self, **kwargs):
        state = dict(value=self.value, attrs=self.attrs)
        if 'value' in kwargs:
            state['value'] = _as_device_tensor(kwargs.pop('value'), device=self.value.device)
        if 'attrs' in kwargs:
            state['attrs'] = kwargs.pop('attrs')
        if kwargs:
            state['attrs'] = state['attrs'].clone(**kwargs)
        new = type(self).__new__(type(self))
        new.__dict__.update(state)
        return new

    def __getitem__(self, item):
        return self.value[item]

    def save(self, fn):
        """Save the mesh (value and attrs) to an npz file (mesh.py:1783-1787)."""
        _save_npz(fn, {'value': self.value}, self.attrs.__getstate__())

    @classmethod
    def load(cls, fn, device=None, allow_pickle=False):
        """Mesh saved by ``save`` (mesh.py:1789-1797)."""
        arrays, attrs = _load_npz(fn, allow_pickle=allow_pickle)
        return cls(arrays['value'], attrs=_attrs_from_state(attrs), device=device)

    @property
    def at(self):
        return _UpdateHelper(self)

    def __array__(self, dtype=None, copy=None):
        arr = self.value.detach().cpu().numpy()
        return arr.astype(dtype) if dtype is not None else arr

    shape = property(lambda self: tuple(self.value.shape))
    ndim = property(lambda self: self.value.ndim)
    size = property(lambda self: self.value.numel())
    dtype = property(lambda self: self.value.dtype)
    real = property(lambda self: self.value.real)
    imag = property(lambda self: self.value.imag if self.value.is_complex() else torch.zeros_like(self.value))
    meshsize = property(lambda self: self.attrs.meshsize)
    boxsize = property(lambda self: self.attrs.boxsize)
    boxcenter = property(lambda self: self.attrs.boxcenter)
    cellsize = property(lambda self: self.attrs.cellsize)

    def conj(self):
        return self.clone(value=torch.conj(self.value).resolve_conj())

    def sum(self, *args, **kwargs):
        return self.value.sum(*args, **kwargs)

    def mean(self, *args, **kwargs):
        return self.value.mean(*args, **kwargs)

    def var(self, *args, **kwargs):
        return self.value.var(*args, unbiased=False, **kwargs) if not args and not kwargs else self.value.var(*args, **kwargs)

    def std(self, *args, **kwargs):
        return self.value.std(*args, unbiased=False, **kwargs) if not args and not kwargs else self.value.std(*args, **kwargs)

    def min(self, *args, **kwargs):
        return self.value.min(*args, **kwargs)

    def max(self, *args, **kwargs):
        return self.value.max(*args, **kwargs)

    def argmin(self, *args, **kwargs):
        return self.value.argmin(*args, **kwargs)

    def argmax(self, *args, **kwargs):
        return self.value.argmax(*args, **kwargs)

    def ravel(self):
        return self.value.reshape(-1)


def _binary(name, op):
    def fn(self, other):
        value = op(self.value, other.value if isinstance(other, BaseMeshField) else other)
        return self.clone(value=value)
    fn.__name__ = name
    return fn


def _rev(op):
    return lambda x, y: op(y, x)


for _name, _op in [('__add__', operator.add), ('__sub__', operator.sub), ('__mul__', operator.mul),
                   ('__truediv__', operator.truediv), ('__floordiv__', operator.floordiv), ('__mod__', operator.mod),
                   ('__pow__', operator.pow), ('__radd__', _rev(operator.add)), ('__rsub__', _rev(operator.sub)),
                   ('__rmul__', _rev(operator.mul)), ('__rtruediv__', _rev(operator.truediv)),
                   ('__rpow__', _rev(operator.pow)), ('__lt__', operator.lt), ('__le__', operator.le),
                   ('__gt__', operator.gt), ('__ge__', operator.ge)]:
    setattr(BaseMeshField, _name, _binary(_name, _op))

BaseMeshField.__neg__ = lambda self: self.clone(value=-self.value)
BaseMeshField.__pos__ = lambda self: self
BaseMeshField.__abs__ = lambda self: self.clone(value=self.value.abs())


class RealMeshField(BaseMeshField):
    """Real-space field (mesh.py:1900-2062)."""

    @property
    def spacing(self):
        return self.cellsize

    def coords(self, kind='position', sparse=None):
        return self.attrs.xcoords(kind=kind, sparse=sparse, device=self.value.device)

    def r2c(self):
        """FFT to the half spectrum (HIP r2c)."""
        return ComplexMeshField(_r2c(self.value, self.attrs.meshsize), attrs=self.attrs)

    def apply(self, fn, kind=Ellipsis, sparse=False):
        args = [self.value]
        if kind is Ellipsis:
            kind = getattr(fn, 'kind', None)
        if kind in ('index', 'position'):
            args.append(self.coords(kind=kind, sparse=sparse))
        elif kind is not None:
            raise ValueError(f'do not recognize kind = {kind}')
        value = fn(*args)
        assert tuple(value.shape) == self.shape, f'kernel does not return correct shape: {value.shape} vs {self.shape}'
        return self.clone(value=value)

    def read(self, positions, resampler='cic', compensate=False, exchange=False, **kwargs):
        """Read the mesh at physical positions (mesh.py:2030-2115)."""
        if isinstance(positions, ParticleField):
            positions = positions.positions
        resampler = resamplers.get_resampler(resampler)
        mesh = self
        if compensate:
            kernel = resampler.compensate if isinstance(compensate, bool) else compensate
            mesh = mesh.r2c().apply(kernel).c2r()
        if not resamplers.is_builtin(resampler):  # the object's own gather (mesh.py:2065-2115)
            pos = _as_device_tensor(positions, device=mesh.value.device)
            return resampler.read(mesh.value, _grid_positions(mesh.attrs, pos))
        return _read(mesh, positions, resampler.order)


def _read(mesh, positions, order):
    pos = _as_device_tensor(positions, device=mesh.value.device)
    if pos.dtype not in (torch.float32, torch.float64):
        pos = pos.to(torch.float32)
    pos = pos.contiguous()
    out = torch.empty(pos.shape[0], dtype=torch.float32, device=pos.device)
    a = mesh.attrs
    _lib.call('mp_read', _lib.stream_handle(pos.device), _lib.ptr(mesh.value.contiguous()), _lib.i32x3(a.meshsize),
              _lib.f64x3(a.boxsize), _lib.f64x3(a.boxcenter), _lib.ptr(pos), int(pos.dtype == torch.float64), pos.shape[0],
              order, _lib.ptr(out))
    return out


class _PendingCombine:
    """The n_i interlaced meshes of _paint after their z and y FFT passes, waiting for the x
    pass that combines them (mesh.py:2739-2751).  Holding them lets a global-line-of-sight
    auto spectrum bin the combined mesh without ever writing it."""

    def __init__(self, partial, pitch, order, interlacing, compensate, zero_dc=False):
        self.partial, self.pitch = partial, int(pitch)
        self.order, self.interlacing, self.compensate, self.zero_dc = int(order), int(interlacing), bool(compensate), bool(zero_dc)

    def with_zero_dc(self):
        return _PendingCombine(self.partial, self.pitch, self.order, self.interlacing, self.compensate, zero_dc=True)

    def materialize(self, attrs):
        device = self.partial.device
        value = _empty_spectrum(attrs.meshsize, device)
        n = _lib.i32x3(attrs.meshsize)
        ws = _lib.workspace(_lib.load().mp_combine_partial_workspace_bytes(n, self.interlacing), device)
        _lib.call('mp_combine_partial', _lib.stream_handle(device), _lib.ptr(self.partial), self.pitch, n, self.order,
                  self.interlacing, int(self.compensate), int(self.zero_dc), _lib.ptr(value), _spectrum_pitch(value),
                  _lib.ptr(ws), ws.numel())
        return value


class ComplexMeshField(BaseMeshField):
    """Half-spectrum field (mesh.py:2118-2196).

    A mesh returned by an interlaced or compensated paint(out='complex') may hold its
    interlaced meshes before their combine; `value` runs the combine on first access."""

    def __getattr__(self, name):
        if name == 'value':
            pending = self.__dict__.get('_pending')
            if pending is not None:
                value = pending.materialize(self.__dict__['attrs'])
                self.__dict__['value'] = value
                del self.__dict__['_pending']
                return value
        raise AttributeError(f'{type(self).__name__!r} object has no attribute {name!r}')

    @classmethod
    def _from_pending(cls, pending, attrs):
        new = cls.__new__(cls)
        new.__dict__.update(attrs=attrs, _pending=pending)
        return new

    def _pending_combine(self):
        return self.__dict__.get('_pending')

    def _with_pending(self, pending):
        return type(self)._from_pending(pending, self.attrs)

    def clone(self, **kwargs):
        pending = self._pending_combine()
        if pending is not None and 'value' not in kwargs and 'attrs' not in kwargs and not kwargs:
            return self._with_pending(pending)
        return super().clone(**kwargs)

    @property
    def spacing(self):
        return self.kfun

    @property
    def kfun(self):
        return self.attrs.kfun

    @property
    def knyq(self):
        return self.attrs.knyq

    def coords(self, kind='wavenumber', sparse=None):
        return self.attrs.kcoords(kind=kind, sparse=sparse, device=self.value.device)

    def c2r(self):
        """Inverse FFT to real space (HIP c2r, 1/N^3)."""
        return RealMeshField(_c2r(self.value, self.attrs.meshsize), attrs=self.attrs)

    def apply(self, fn, sparse=False, kind=Ellipsis):
        args = [self.value]
        if kind is Ellipsis:
            kind = getattr(fn, 'kind', None)
        if kind in ('circular', 'wavenumber'):
            args.append(self.coords(kind=kind, sparse=sparse))
        elif kind is not None:
            raise ValueError(f'do not recognize kind = {kind}')
        value = fn(*args)
        assert tuple(value.shape) == self.shape, f'kernel does not return correct shape: {value.shape} vs {self.shape}'
        return self.clone(value=value)


# ----------------------------------------------------------------------------- particles

def _sum_f64(a, b=None, square=False):
    """FP64 sum of a float32 GPU vector (or of a*b, a^2), as a 0-d float64 GPU tensor."""
    a = a.contiguous().to(torch.float32)
    if b is not None:
        b = b.contiguous().to(torch.float32)
    out = torch.empty((), dtype=torch.float64, device=a.device)
    ws = _lib.workspace(_lib.load().mp_reduce_workspace_bytes(), a.device)
    _lib.call('mp_sum_f64', _lib.stream_handle(a.device), _lib.ptr(a), _lib.ptr(b), a.numel(), int(square), _lib.ptr(out),
              _lib.ptr(ws), ws.numel())
    return out


class ParticleField:
    """Particles with weights on a mesh geometry (mesh.py:2397-2670), one device."""

    def __init__(self, positions, weights=None, attrs=None, extra=None, exchange=False, backend='auto', **kwargs):
        if attrs is None:
            attrs = kwargs.pop('mattrs', None)
        if attrs is None:
            raise ValueError('attrs must be provided')
        if not isinstance(attrs, MeshAttrs):
            attrs = MeshAttrs(**attrs)
        device = kwargs.pop('device', None)
        positions = _as_device_tensor(positions, device=device)
        if positions.dtype not in (torch.float32, torch.float64):
            positions = positions.to(torch.float32)
        positions = positions.contiguous()
        if weights is None:
            weights = torch.ones(positions.shape[0], dtype=torch.float32, device=positions.device)
        else:
            weights = _as_device_tensor(weights, device=positions.device)
            weights = torch.broadcast_to(weights, positions.shape[:1]).to(torch.float32).contiguous()
        extra = {k: _as_device_tensor(v, device=positions.device) for k, v in (extra or {}).items()}
        self.__dict__.update(positions=positions, weights=weights, attrs=attrs, extra=extra, exchange_direct=None,
                             exchange_inverse=None)

    def __repr__(self):
        return f'{type(self).__name__}(size={self.size}, extra={list(self.extra)})'

    def clone(self, **kwargs):
        state = dict(positions=self.positions, weights=self.weights, attrs=self.attrs, extra=self.extra)
        kwargs.pop('exchange', None)
        state.update(kwargs)
        return type(self)(**state)

    def exchange(self, **kwargs):
        """Single device: particles already sit with their mesh."""
        return self.clone(**kwargs)

    def save(self, fn):
        """Save positions, weights, extra fields and attrs to an npz file (mesh.py:2646-2658)."""
        state = {'positions': self.positions, 'weights': self.weights}
        for name, value in self.extra.items():
            if name in state:
                raise ValueError(f'extra field name {name} conflicts with reserved field names')
            state[name] = value
        attrs = self.attrs.__getstate__()
        if self.extra:
            attrs['extra'] = list(self.extra)
        _save_npz(fn, state, attrs)

    @classmethod
    def load(cls, fn, device=None, allow_pickle=False):
        """Particles saved by ``save`` (mesh.py:2660-2670)."""
        arrays, attrs = _load_npz(fn, allow_pickle=allow_pickle)
        extra = {name: arrays.pop(name) for name in attrs.pop('extra', [])}
        return cls(arrays['positions'], arrays['weights'], attrs=_attrs_from_state(attrs), extra=extra, device=device)

    def __getitem__(self, name):
        return self.clone(positions=self.positions[name], weights=self.weights[name],
                          extra={k: v[name] for k, v in self.extra.items()})

    @property
    def size(self):
        return self.positions.shape[0]

    def sum(self):
        """Sum of weights, accumulated in float64."""
        return _sum_f64(self.weights)

    @classmethod
    def concatenate(cls, others, weights=None, local=True, extra_weights=None, **kwargs):
        """mesh.py:2543-2578."""
        if weights is None:
            weights = [1] * len(others)
        extra_names = [n for n in others[0].extra if all(n in o.extra for o in others[1:])]
        extra_weights = extra_weights or {}
        if not isinstance(extra_weights, dict):
            extra_weights = {n: extra_weights for n in extra_names}
        pos = torch.cat([o.positions for o in others], dim=0)
        w = torch.cat([_scale_weights(o.weights, f) for o, f in zip(others, weights)], dim=0)
        extra = {n: torch.cat([extra_weights[n] * o.extra[n] if n in extra_weights else o.extra[n] for o in others], dim=0)
                 for n in extra_names}
        return cls(positions=pos, weights=w, attrs=others[0].attrs, extra=extra, **kwargs)

    def __add__(self, other):
        if isinstance(other, ParticleField):
            return self.concatenate([self, other], [1, 1])
        raise RuntimeError(f'Type of `other` {type(other)} not understood.')

    def __sub__(self, other):
        if isinstance(other, ParticleField):
            return self.concatenate([self, other], [1, -1])
        raise RuntimeError(f'Type of `other` {type(other)} not understood.')

    def __mul__(self, other):
        return self.concatenate([self], [other])

    __rmul__ = __mul__

    def __truediv__(self, other):
        return self.concatenate([self], [1. / other])

    def paint(self, resampler='cic', interlacing=0, compensate=False, out='real', dtype=None, **kwargs):
        """Paint particles to a mesh (mesh.py:2604-2640): HIP paint, and for interlacing or
        compensation the fused paint -> FFT -> phase/compensation pipeline."""
        attrs = self.attrs if dtype is None else self.attrs.clone(dtype=dtype)
        return _paint(attrs, self.positions, self.weights, resampler=resampler, interlacing=interlacing,
                      compensate=compensate, out=out)


def _scale_weights(w, factor):
    if isinstance(factor, (int, float)) and factor == 1:
        return w
    if isinstance(factor, torch.Tensor):
        return (w.to(torch.float64) * factor.to(w.device, torch.float64)).to(torch.float32)
    return (w.to(torch.float64) * float(factor)).to(torch.float32)


for _name in ['meshsize', 'boxsize', 'boxcenter', 'dtype', 'fft_backend', 'cellsize']:
    setattr(ParticleField, _name, property(lambda self, _n=_name: getattr(self.attrs, _n)))


def _paint_partial(attrs, positions, weights, order, ni, compensate):
    """The n_i interlaced meshes after their z and y FFT passes (mp_paint_partial)."""
    device = positions.device
    pos = positions.contiguous()
    n = _lib.i32x3(attrs.meshsize)
    np_ = pos.shape[0]
    pitch = int(_lib.load().mp_spectrum_pitch(int(attrs.meshsize[2])))
    partial = torch.empty((ni, int(attrs.meshsize[0]), int(attrs.meshsize[1]), pitch), dtype=torch.complex64, device=device)
    ws = _lib.workspace(_lib.load().mp_paint_partial_workspace_bytes(n, np_, order, ni), device)
    _lib.call('mp_paint_partial', _lib.stream_handle(device), _lib.ptr(pos), int(pos.dtype == torch.float64), _lib.ptr(weights),
              np_, n, _lib.f64x3(attrs.boxsize), _lib.f64x3(attrs.boxcenter), order, ni, _lib.ptr(partial), pitch, _lib.ptr(ws),
              ws.numel())
    return _PendingCombine(partial, pitch, order, ni, compensate)


def _grid_positions(attrs, positions):
    """(x + L/2 - c) / cellsize in the positions' dtype (mesh.py:2684)."""
    dtype = positions.dtype if positions.dtype in (torch.float32, torch.float64) else torch.float32
    shift = torch.as_tensor(attrs.boxsize / 2. - attrs.boxcenter, dtype=torch.float64, device=positions.device)
    cell = torch.as_tensor(attrs.cellsize, dtype=torch.float64, device=positions.device)
    return ((positions.to(torch.float64) + shift) / cell).to(dtype)


def _paint_custom(attrs, positions, weights, resampler, interlacing, compensate, out):
    """_paint (mesh.py:2673-2754) with a resampler object other than the four built-in ones
    (the plugin point of resamplers.get_resampler, resamplers.py:134-144): its ``.paint(mesh,
    positions, weights)`` scatters each shifted copy of the grid-unit positions (device tensors)
    into a zero mesh; the r2c, the interlacing phases exp(+i s sum k_circ) / n_i, the sum, the
    compensation (its ``.compensate`` kernel, or the given callable) and the c2r follow the
    reference's order."""
    device = positions.device
    shape = tuple(int(n) for n in attrs.meshsize)
    xg = _grid_positions(attrs, positions)
    ni = max(int(interlacing), 1)
    if isinstance(compensate, bool):
        kernel_compensate = resampler.compensate if compensate else None
    else:
        kernel_compensate = compensate

    def painted(shift):
        value = resampler.paint(torch.zeros(shape, dtype=torch.float32, device=device), xg + shift if shift else xg, weights)
        return RealMeshField(torch.as_tensor(value, dtype=torch.float32, device=device), attrs=attrs)

    if ni <= 1:
        mesh = painted(0.)
        if kernel_compensate is not None:
            mesh = mesh.r2c().apply(kernel_compensate)
            return mesh.c2r() if out == 'real' else mesh
        return mesh if out == 'real' else mesh.r2c()
    acc = None
    for j in range(ni):
        s = j / ni

        def phase(value, kvec, s=s):
            return value * torch.exp(1j * s * sum(kvec)).to(value.dtype) / ni

        term = painted(s).r2c().apply(phase, kind='circular', sparse=True)
        acc = term if acc is None else acc + term
    if kernel_compensate is not None:
        acc = acc.apply(kernel_compensate)
    return acc.c2r() if out == 'real' else acc


def _paint(attrs, positions, weights=None, resampler='cic', interlacing=0, compensate=False, out='real'):
    """_paint, mesh.py:2673-2754, single device."""
    resampler = resamplers.get_resampler(resampler)
    if not resamplers.is_builtin(resampler):
        _lib.require_device(positions, 'positions')
        return _paint_custom(attrs, positions, weights, resampler, interlacing, compensate, out)
    order = int(resampler.order)
    ni = max(int(interlacing), 1)
    device = positions.device
    _lib.require_device(positions, 'positions')
    pos = positions.contiguous()
    w = weights.contiguous().to(torch.float32) if weights is not None else None
    np_ = pos.shape[0]
    pos_f64 = int(pos.dtype == torch.float64)
    n = _lib.i32x3(attrs.meshsize)
    box, center = _lib.f64x3(attrs.boxsize), _lib.f64x3(attrs.boxcenter)
    stream = _lib.stream_handle(device)
    kernel_compensate = None
    if not isinstance(compensate, bool):
        kernel_compensate, compensate = compensate, False
    if ni <= 1 and not compensate:
        value = torch.empty(tuple(int(s) for s in attrs.meshsize), dtype=torch.float32, device=device)
        ws = _lib.workspace(_lib.load().mp_paint_workspace_bytes(n, np_, order, 1), device)
        _lib.call('mp_paint', stream, _lib.ptr(pos), pos_f64, _lib.ptr(w), np_, n, box, center, order, 1, _lib.ptr(value),
                  _lib.ptr(ws), ws.numel())
        mesh = RealMeshField(value, attrs=attrs)
        if kernel_compensate is not None:
            mesh = mesh.r2c().apply(kernel_compensate)
            return mesh.c2r() if out == 'real' else mesh
        return mesh if out == 'real' else mesh.r2c()
    mesh = ComplexMeshField._from_pending(_paint_partial(attrs, pos, w, order, ni, compensate), attrs)
    if kernel_compensate is not None:
        mesh = mesh.apply(kernel_compensate)
    return mesh.c2r() if out == 'real' else mesh


# Particles per mesh cell (on average over the mesh) above which FKPField.paint paints data
# and randoms apart: measured at 0.5 per cell the single paint stays at 3.5e-6 of the float64
# oracle, at 2 per cell it reached 1.5e-5.
_SPLIT_DENSITY = 1.


def _paint_sum(attrs, parts, resampler, interlacing, compensate, out):
    """Sum of the paints of [(positions, weights), ...] with the built-in kernels: the real
    meshes added, or the n_i partial spectra added before their pending combine (the z and y
    passes are linear).  None for custom resamplers or compensation callables."""
    rs = resamplers.get_resampler(resampler)
    if not resamplers.is_builtin(rs) or not isinstance(compensate, bool):
        return None
    order, ni = int(rs.order), max(int(interlacing), 1)
    if ni <= 1 and not compensate:
        meshes = [_paint(attrs, pos, w, resampler=resampler, interlacing=interlacing, compensate=False, out='real')
                  for pos, w in parts]
        value = meshes[0].value
        for m in meshes[1:]:
            value.add_(m.value)
        mesh = RealMeshField(value, attrs=attrs)
        return mesh if out == 'real' else mesh.r2c()
    pending = []
    for pos, w in parts:
        _lib.require_device(pos, 'positions')
        pending.append(_paint_partial(attrs, pos.contiguous(), w.contiguous().to(torch.float32), order, ni, compensate))
    for q in pending[1:]:
        pending[0].partial.add_(q.partial)
    mesh = ComplexMeshField._from_pending(pending[0], attrs)
    return mesh.c2r() if out == 'real' else mesh


# functional interface (mesh.py:2760-2785)

def c2r(mesh):
    return mesh.c2r()


def r2c(mesh):
    return mesh.r2c()


def apply(mesh, fn, sparse=False, **kwargs):
    return mesh.apply(fn, sparse=sparse, **kwargs)


def read(mesh, positions, *args, **kwargs):
    return mesh.read(positions, *args, **kwargs)


def paint(particles, *args, **kwargs):
    return particles.paint(*args, **kwargs)


class FKPField:
    """Data minus alpha * randoms (mesh.py:2913-2999), alpha = sum w_d / sum w_r."""

    def __init__(self, data, randoms, attrs=None):
        if attrs is not None:
            data = data.clone(attrs=attrs)
            randoms = randoms.clone(attrs=attrs)
        self.__dict__.update(data=data, randoms=randoms)

    def clone(self, **kwargs):
        state = dict(data=self.data, randoms=self.randoms)
        state.update(kwargs)
        return type(self)(**state)

    def exchange(self, **kwargs):
        return self

    @property
    def attrs(self):
        return self.data.attrs

    @property
    def alpha(self):
        return self.data.sum() / self.randoms.sum()

    @property
    def particles(self):
        particles = self.__dict__.get('_particles')
        if particles is None:
            particles = ParticleField.concatenate([self.data, self.randoms], [1, -self.alpha]).clone(attrs=self.data.attrs)
            self.__dict__['_particles'] = particles
        return particles

    def paint(self, resampler='cic', interlacing=1, compensate=False, dtype=None, out='real', **kwargs):
        """Paint data - alpha * randoms (mesh.py:2973-2980).  Catalogues denser than
        _SPLIT_DENSITY particles per mesh cell are painted one sign at a time and the two meshes
        added in float32: the fixed-point paint scales each block by a bound on the sum of |w|
        that reaches a cell, and where data and randoms cancel cell by cell that bound is far
        above the cells, which then lose precision (7 data and 130 randoms per cell: 1.5e-5 on
        P_2 at 128^3, 2 particles per mesh cell on average; painted apart, each catalogue's
        bound is of the order of its cells).  Sparser catalogues (the 1024^3 survey
        configuration has 0.1 per cell) keep one paint of the concatenation."""
        attrs = self.attrs if dtype is None else self.attrs.clone(dtype=dtype)
        if self.data.size + self.randoms.size > _SPLIT_DENSITY * float(np.prod(attrs.meshsize)):
            parts = [(self.data.positions, self.data.weights),
                     (self.randoms.positions, _scale_weights(self.randoms.weights, -self.alpha))]
            mesh = _paint_sum(attrs, parts, resampler, interlacing, compensate, out)
            if mesh is not None:
                return mesh
        return self.particles.paint(resampler=resampler, interlacing=interlacing, compensate=compensate, dtype=dtype, out=out)


# ----------------------------------------------------------------------------- normalisation

def get_halo_for_new_mattrs(mattrs, **kwargs):
    """New MeshAttrs for normalisation meshes (mesh.py:3002-3039); halo is 0 on one device."""
    new = mattrs
    if kwargs:
        old = {k: mattrs[k] for k in mattrs.keys()}
        update = dict(kwargs)
        if 'cellsize' in update:
            old.pop('meshsize')
            old.setdefault('approximate', True)
        kw = {**old, **update}
        new_attrs = get_mesh_attrs(**kw)
        new = mattrs.clone(**{k: new_attrs[k] for k in new_attrs.keys()})
        if np.any((new.meshsize - mattrs.meshsize) % 2):  # keep a divisor of 2 (no half-cell shift)
            meshsize = new.meshsize + 1
            new = new.clone(meshsize=meshsize, boxsize=new.cellsize * meshsize)
    return 0, new


def _iter_meshes(*inputs, resampler='cic', interlacing=0, compensate=False, **kwargs):
    """mesh.py:3042-3089."""
    meshes, particles = [], []
    attrs = {k: v for k, v in kwargs.items() if v is not None}
    for inp in inputs:
        if isinstance(inp, RealMeshField):
            meshes.append(inp)
            attrs = {name: getattr(inp, name) for name in ['boxsize', 'boxcenter', 'meshsize']}
        else:
            particles.append(inp)
    if particles and attrs:
        _, new = get_halo_for_new_mattrs(particles[0].attrs, **attrs)
        clones = {}
        for p in particles:
            if id(p) not in clones:
                clones[id(p)] = p.clone(attrs=new)
        particles = [clones[id(p)] for p in particles]
    for mesh in meshes:
        yield mesh
    painted = {}  # the same particles passed several times (auto normalizations) are painted once
    for p in particles:
        if id(p) not in painted:
            painted[id(p)] = p.paint(resampler=resampler, interlacing=interlacing, compensate=compensate)
        yield painted[id(p)]


def compute_normalization(*inputs, resampler='cic', start=1., **kwargs):
    """Sum of the product of the painted inputs times cellvolume^(1 - n) (mesh.py:3092-3118),
    accumulated in float64."""
    meshes = list(_iter_meshes(*inputs, resampler=resampler, **kwargs))
    cellvol = float(np.prod(meshes[0].cellsize))
    if len(meshes) == 2 and isinstance(start, (int, float)):
        total = _sum_f64(meshes[0].value.reshape(-1), meshes[1].value.reshape(-1)) * float(start)
    else:
        prod = meshes[0].value.to(torch.float32) * start if not isinstance(start, (int, float)) else meshes[0].value * float(start)
        for m in meshes[1:]:
            prod = prod * m.value
        total = _sum_f64(prod.reshape(-1))
    return total * cellvol**(1 - len(meshes))


def compute_box_normalization(*inputs):
    """Constant-density normalisation (mesh.py:3121-3129)."""
    attrs = inputs[0].attrs
    size = float(np.prod(attrs.meshsize))
    norm = 1.
    sums = {}  # an auto normalisation passes the same field twice: sum it once
    for inp in inputs:
        if id(inp) not in sums:
            sums[id(inp)] = inp.sum() if isinstance(inp, ParticleField) else _sum_f64(inp.value.reshape(-1))
        norm = norm * sums[id(inp)] / size
    return norm * size * float(np.prod(attrs.cellsize))**(1 - len(inputs))


def _make_input_tuple(*inputs):
    if len(inputs) == 1 and isinstance(inputs[0], (tuple, list)):
        inputs = inputs[0]
    return tuple(inputs)


def format_meshes(*meshes, fields=None, nmeshes=None):
    """Pad the list of inputs for auto/cross correlations (mesh.py:3138-3156)."""
    meshes = list(_make_input_tuple(*meshes))
    assert len(meshes) >= 1
    if nmeshes is not None:
        assert len(meshes) <= nmeshes
        meshes = meshes + [None] * (nmeshes - len(meshes))
    fields = [0] if fields is None else list(fields)
    for mesh in meshes[len(fields):]:
        fields.append(fields[-1] if mesh is None else fields[-1] + 1)
    for i, (mesh, field) in enumerate(zip(meshes, fields)):
        if mesh is None:
            meshes[i] = meshes[fields.index(field)]
        assert meshes[i] is not None, f'first mesh of field {field} must not be None'
    return meshes, tuple(fields)
