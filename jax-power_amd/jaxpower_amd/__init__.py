"""MI355X-native mesh power-spectrum path with jax-power's interface (jaxpower/__init__.py:1-12):
paint particles (NGP/CIC/TSC/PCS, interlacing, compensation), real-to-complex FFT, and
binning of |delta(k)|^2 into power-spectrum multipoles (global or local line of sight), with
FKP normalisation and shot noise; correlation-function multipoles from the same meshes, and
bispectrum multipoles (Scoccimarro basis, global line of sight) and power-spectrum window
matrices (global line of sight).  Compute runs in
libjaxpower_amd.so (hand-written HIP for gfx950); torch provides device memory and streams.
"""
from .mesh import (MeshAttrs, BaseMeshField, RealMeshField, ComplexMeshField, ParticleField, FKPField, r2c, c2r, apply,
                   read, paint, fftfreq, get_mesh_attrs, next_fft_size, compute_normalization, compute_box_normalization,
                   get_halo_for_new_mattrs)
from .mesh2 import (BinMesh2SpectrumPoles, BinMesh2CorrelationPoles, compute_mesh2, compute_mesh2_spectrum,
                    compute_mesh2_correlation, compute_mesh2_spectrum_window, compute_mesh2_spectrum_mean,
                    compute_fkp2_normalization,
                    compute_box2_normalization, compute_fkp2_shotnoise)
from .mesh3 import (BinMesh3SpectrumPoles, BinMesh3CorrelationPoles, compute_mesh3, compute_mesh3_spectrum, compute_mesh3_correlation,
                    compute_box3_normalization, compute_fkp3_normalization, compute_fkp3_shotnoise)
from .mock import generate_uniform_particles, generate_gaussian_mesh, generate_lognormal_particles
from .types import (Mesh2SpectrumPole, Mesh2SpectrumPoles, Mesh2CorrelationPole, Mesh2CorrelationPoles, Mesh3SpectrumPole,
                    Mesh3SpectrumPoles, Mesh3CorrelationPole, Mesh3CorrelationPoles, WindowMatrix, read as read_observable, write as write_observable)
from .autodiff import mesh2_spectrum_poles, box_spectrum_values
from .batch import Mesh2SpectrumBatch
from . import resamplers
from . import tuning

tuning._autoload()  # only when JAXPOWER_AMD_TUNING_CACHE names a cache file

__all__ = ['MeshAttrs', 'BaseMeshField', 'RealMeshField', 'ComplexMeshField', 'ParticleField', 'FKPField', 'r2c', 'c2r',
           'apply', 'read', 'paint', 'fftfreq', 'get_mesh_attrs', 'next_fft_size', 'compute_normalization',
           'compute_box_normalization', 'BinMesh2SpectrumPoles', 'BinMesh2CorrelationPoles', 'compute_mesh2',
           'compute_mesh2_spectrum', 'compute_mesh2_correlation', 'Mesh2CorrelationPole', 'Mesh2CorrelationPoles',
           'compute_fkp2_normalization', 'compute_box2_normalization', 'compute_fkp2_shotnoise',
           'generate_uniform_particles', 'generate_gaussian_mesh', 'generate_lognormal_particles', 'Mesh2SpectrumPole',
           'Mesh2SpectrumPoles', 'mesh2_spectrum_poles', 'box_spectrum_values', 'Mesh2SpectrumBatch', 'resamplers', 'tuning', 'BinMesh3SpectrumPoles',
           'compute_mesh3', 'compute_mesh3_spectrum', 'compute_box3_normalization', 'compute_fkp3_normalization', 'Mesh3SpectrumPole', 'Mesh3SpectrumPoles',
           'compute_mesh2_spectrum_window', 'compute_mesh2_spectrum_mean', 'WindowMatrix', 'compute_fkp3_shotnoise',
           'BinMesh3CorrelationPoles', 'compute_mesh3_correlation', 'Mesh3CorrelationPole', 'Mesh3CorrelationPoles']
