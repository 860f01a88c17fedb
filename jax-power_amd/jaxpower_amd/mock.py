"""Seeded synthetic inputs: uniform particles, Gaussian meshes and lognormal Poisson mocks.

jax-power draws its mocks from jax.random streams (jaxpower/mock.py:15-47, 180-207); those
streams are not reproduced here: the generators below are seeded with torch.Generator on the
device, so one seed gives the same catalogue on every run of this package.  The lognormal
mock (a Gaussian field with the given P(k), exponentiated, Poisson-sampled per cell, with
positions uniform within each cell and the catalogue shuffled) has no counterpart in the
reference.
"""
import numpy as np
import torch

from .mesh import MeshAttrs, ParticleField, RealMeshField, _default_device


def _generator(seed, device):
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    return g


def generate_uniform_particles(mattrs, size, seed=42, dtype=torch.float32, device=None, **kwargs):
    """Uniform positions in the box (mock.py:180-207), unit weights."""
    device = device or _default_device()
    g = _generator(seed, device)
    u = torch.rand((int(size), 3), generator=g, device=device, dtype=torch.float64)
    box = torch.as_tensor(mattrs.boxsize, device=device)
    lo = torch.as_tensor(mattrs.boxcenter - mattrs.boxsize / 2., device=device)
    return ParticleField((u * box + lo).to(dtype), attrs=mattrs, **kwargs)


def generate_gaussian_mesh(mattrs, power=lambda kvec: 1., seed=42, unitary_amplitude=False, device=None):
    """Gaussian random field with power spectrum ``power(kvec)`` (mock.py:15-47)."""
    device = device or _default_device()
    g = _generator(seed, device)
    white = torch.randn(tuple(int(n) for n in mattrs.meshsize), generator=g, device=device, dtype=torch.float32)
    mesh = RealMeshField(white, attrs=mattrs).r2c()
    cellvol = float(np.prod(mattrs.cellsize))
    nmesh = float(np.prod(mattrs.meshsize))

    def kernel(value, kvec):
        ker = torch.sqrt(power(kvec) / cellvol)
        if unitary_amplitude:
            ker = ker * np.sqrt(nmesh) / torch.abs(value)
        return value * ker

    return mesh.apply(kernel, kind='wavenumber').c2r()


def power_law_cutoff(amplitude=2e4, k0=0.1, index=-1.5, kcut=0.5):
    """P(k) = amplitude (k/k0)^index exp(-(k/kcut)^2), zero at k = 0."""

    def power(kvec):
        k = torch.sqrt(sum(kk.to(torch.float64)**2 for kk in kvec))
        pk = amplitude * torch.where(k > 0, k / k0, torch.ones_like(k))**index * torch.exp(-(k / kcut)**2)
        return torch.where(k > 0, pk, torch.zeros_like(pk)).to(torch.float32)

    return power


def generate_lognormal_particles(mattrs, nbar=None, size=None, power=None, seed=42, dtype=torch.float32, device=None,
                                 shuffle=True):
    """Lognormal Poisson mock: delta_G with P(k) = ``power``, delta_LN = exp(G - var/2) - 1,
    counts ~ Poisson(nbar Vcell (1 + delta_LN)), positions uniform in their cell."""
    device = device or _default_device()
    power = power or power_law_cutoff()
    cellvol = float(np.prod(mattrs.cellsize))
    ncell = int(np.prod(mattrs.meshsize))
    if nbar is None:
        nbar = float(size) / (ncell * cellvol)
    G = generate_gaussian_mesh(mattrs, power=power, seed=seed, device=device).value
    var = G.double().var().item()
    lam = (nbar * cellvol) * torch.exp(G - 0.5 * var)
    del G
    g = _generator(seed + 1_000_003, device)
    counts = torch.poisson(lam.reshape(-1), generator=g).to(torch.int64)
    del lam
    cells = torch.repeat_interleave(torch.arange(ncell, device=device), counts)
    del counts
    n = torch.as_tensor(mattrs.meshsize.astype('i8'), device=device)
    ix = cells // (n[1] * n[2])
    iy = (cells // n[2]) % n[1]
    iz = cells % n[2]
    del cells
    idx = torch.stack([ix, iy, iz], dim=1).to(torch.float64)
    del ix, iy, iz
    u = torch.rand(idx.shape, generator=g, device=device, dtype=torch.float64)
    cell = torch.as_tensor(mattrs.cellsize, device=device)
    lo = torch.as_tensor(mattrs.boxcenter - mattrs.boxsize / 2., device=device)
    pos = ((idx + u) * cell + lo).to(dtype)
    del idx, u
    if shuffle:
        perm = torch.randperm(pos.shape[0], generator=g, device=device)
        pos = pos[perm]
    return ParticleField(pos, attrs=mattrs)


def generate_shell_cap_positions(size, rmin=1000., rmax=2000., cap_degrees=60., seed=42, dtype=torch.float32, device=None):
    """(size, 3) positions uniform in volume in the shell rmin <= r < rmax, within cap_degrees of
    the +z axis, with the observer at the origin (the survey geometry of an FKP measurement;
    no counterpart in the reference)."""
    device = device or _default_device()
    g = _generator(seed, device)
    u = torch.rand((int(size), 3), generator=g, device=device, dtype=torch.float64)
    r = (rmin**3 + u[:, 0] * (rmax**3 - rmin**3))**(1. / 3.)
    cmin = np.cos(np.radians(cap_degrees))
    cost = cmin + u[:, 1] * (1. - cmin)
    sint = torch.sqrt(torch.clamp(1. - cost**2, min=0.))
    phi = 2. * np.pi * u[:, 2]
    pos = torch.stack([r * sint * torch.cos(phi), r * sint * torch.sin(phi), r * cost], dim=1)
    return pos.to(dtype)


def shell_cap_volume(rmin=1000., rmax=2000., cap_degrees=60.):
    """Volume of the shell-cap region of generate_shell_cap_positions."""
    return 2. * np.pi * (1. - np.cos(np.radians(cap_degrees))) * (rmax**3 - rmin**3) / 3.
