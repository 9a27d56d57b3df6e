"""dlux_amd -- an MI355X-native (gfx950) implementation of dLux's diffraction hot path.

Hand-written HIP kernels (``dlux_amd/csrc``) behind a C ABI
(``include/dlux_amd.h``), driven from a Python layer that keeps dLux's
optical-system / layer / propagator API:

    import dlux_amd as dl
    import dlux_amd.utils as dlu
"""

from . import runtime, utils
from . import parametric, polynomials
from .parametric import Parametric, ParametricBasis, ExplicitBasis, ImplicitBasis, CoordBasis
from .polynomials import DynamicZernike, ZernikeBasis, DynamicZernikeBasis
from .runtime import set_precision, get_precision, precision
from .base import Base
from .coordinates import (Spec, PadSpec, CoordSpec, BaseCoordTransform, CoordTransform,
                          DistortedCoords)
from .psfs import PSF
from .wavefronts import Wavefront
from . import layers
from .layers import (OpticalLayer, TransmissiveLayer, AberratedLayer, BasisLayer, Tilt,
                     Normalise, Optic, BasisOptic, Propagator, FFT, MFT, ASMPropagator)
from .optical_systems import (BaseOpticalSystem, OpticalSystem, ParametricOpticalSystem,
                              LayeredOpticalSystem, AngularOpticalSystem,
                              CartesianOpticalSystem)
from .spectra import BaseSpectrum, SimpleSpectrum, Spectrum, PolySpectrum
from .sources import (BaseSource, Source, PointSource, PointSources, BinarySource, ResolvedSource,
                      PointResolvedSource, Scene)
from .detectors import BaseDetector, LayeredDetector
from .instruments import Instrument, Telescope, Dither

__version__ = "0.1.0"
