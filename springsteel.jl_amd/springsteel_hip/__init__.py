"""springsteel_hip -- MI355X (gfx950) transform path for Springsteel.jl RL / RZ / RLZ grids.

The package is the host-side mirror of the reference's grid API over the
C-ABI library ``libspringsteel_hip.so`` (include/springsteel_hip.h).  The
compute runs only in that library's HIP kernels.
"""
from .bcs import (AntisymmetricBC, BoundaryConditions, CauchyBC, Chebyshev, CubicBSpline, DirichletBC,
                  ExponentialBC, FixedBC, Fourier, NaturalBC, NeumannBC, PeriodicBC, RobinBC, SecondDerivativeBC,
                  SymmetricBC, ZerosBC, bc_rank)
from .capi import SSG_FP32_REFINE, SSG_FP64, SSG_MEM_DEVICE, SSG_MEM_HOST, Plan
from .errors import (ArgumentError, BoundsError, DomainError, ExtensionMissingError, HIPError, NotSupportedError,
                     PosDefException, SpringsteelError)
from .filters import GaussianFilter, SpectralFilter
from .grid import (SpringsteelGrid, applyFilter, applyFilter_, applyFilter_device, createGrid, evaluate_unstructured, evaluate_unstructured_device, getGridpoints,
                   getRegularGridpoints, gridTransform, gridTransform_, gridTransform_device, regularGridTransform,
                   regularGridTransform_device, set_ahat, spectralTransform, spectralTransform_,
                   spectralTransform_device)
from .params import SpringsteelGridParameters, compute_derived_params, resolve
from .tiles import (calcTileSizes, setSpectralTile_, tileParameters, splineTransform_, splineTransform_device, sumSharedSpectral,
                    sumSpectralTile_, tileTransform_, tileTransform_device)

__all__ = [n for n in dir() if not n.startswith("_")]
