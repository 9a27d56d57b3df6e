"""Optical and detector layers (dLux.layers subset on the hot path)."""

from .optical_layers import (OpticalLayer, Interpolate, TransmissiveLayer, AberratedLayer,
                             BasisLayer, Tilt, Normalise)
from .optics import Optic, BasisOptic
from .propagators import (Propagator, FFT, MFT, ASMPropagator, ABCDElement, ABCDFreeSpace,
                          ABCDLens, ABCDMirror, ABCDConjugatePlane, ABCDPropagator,
                          MFTPropagator, FFTPropagator, Fraunhofer, Fresnel)
from .detector_layers import (DetectorLayer, ApplyInterpolation, ApplyPixelResponse, ApplyJitter, ApplySaturation,
                              AddConstant, Downsample)
from .apertures import (ApertureLayer, BaseDynamicAperture, DynamicAperture, CircularAperture,
                        SquareAperture, RectangularAperture, RegPolyAperture, Spider,
                        AberratedAperture, CompositeAperture, CompoundAperture, MultiAperture)
