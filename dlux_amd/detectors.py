"""Detectors (detectors.py:1-159): a keyed list of detector layers.

``LayeredDetector.__call__`` lowers its whole layer list to one HIP detector
chain (dlx_detector_apply): elementwise layers fuse into the convolution and
downsampling passes, so the chain reads and writes the image once per
structural layer.  Layers without a chain form run one by one.
"""

from __future__ import annotations

from collections import OrderedDict

from . import runtime
from .base import Base
from .layers.detector_layers import DetectorLayer
from .psfs import PSF

__all__ = ["BaseDetector", "LayeredDetector"]


class BaseDetector(Base):
    """Abstract detector (detectors.py:18-48)."""

    def __call__(self, psf: PSF, return_psf: bool = False):  # pragma: no cover
        raise NotImplementedError

    def model(self, psf: PSF, return_psf: bool = False):
        """Alias of __call__ (detectors.py:33-48)."""
        return self(psf, return_psf)


class LayeredDetector(BaseDetector):
    """Applies its layers in order (detectors.py:51-159)."""

    def __init__(self, layers):
        from .optical_systems import list2dictionary
        self.layers = list2dictionary(layers, True, DetectorLayer)

    def __getattr__(self, key):
        layers = self.__dict__.get("layers")
        if layers is None:
            raise AttributeError(key)
        if key in layers:
            return layers[key]
        for layer in layers.values():
            if hasattr(layer, key):
                return getattr(layer, key)
        raise AttributeError(
            f"{type(self).__name__} has no attribute {key!r} (layers: {list(layers)})")

    def _chain(self):
        ops, scale = [], 1
        for layer in self.layers.values():
            op = layer._det_op()
            if op is None:
                return None, 1
            ops.append(op)
            scale *= layer._scale_factor()
        return ops, scale

    def __call__(self, psf: PSF, return_psf: bool = False):
        ops, scale = self._chain()
        if ops is not None:
            psf = PSF(runtime.detector(psf.data, ops), psf.pixel_scale * scale)
        else:
            for layer in self.layers.values():
                psf = layer(psf)
        return psf if return_psf else psf.data

    def insert_layer(self, layer, index: int):
        """Insert a layer (or a (key, layer) tuple) and re-key duplicates
        (detectors.py:130-155, utils/helpers.py:165-196)."""
        from .optical_systems import list2dictionary
        items = list(self.layers.items())
        items.insert(index, layer)
        return self.set("layers", list2dictionary(items, True, DetectorLayer))

    def remove_layer(self, key: str):
        layers = OrderedDict(self.layers)
        layers.pop(key)
        return self.set("layers", layers)

