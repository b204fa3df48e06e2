"""Staggered 2D channel grid (mac2d/grid.py:7-43).

Same fields, validation, spacing, face shapes, coordinates, zero fields and
velocity checks as ``ChannelGrid``. ``periodic=True`` (build-own) makes the
channel periodic in x: u then has nx faces (face x at x*hx), v keeps its
ny+1 wall-bounded faces.
"""
from dataclasses import dataclass
from math import isfinite, prod
import torch

from .. import _abi


@dataclass(frozen=True)
class ChannelGrid:
    shape: tuple = (256, 64)
    lengths: tuple = (8., 1.61)
    periodic: bool = False

    def __post_init__(self):
        if len(self.shape) != 2 or any(type(n) is not int or n < 4 for n in self.shape):
            raise ValueError('two integer cell counts >=4 required')
        if len(self.lengths) != 2 or any(not isfinite(v) or v <= 0 for v in self.lengths):
            raise ValueError('two finite positive channel lengths required')

    @property
    def spacing(self):
        return tuple(L / n for L, n in zip(self.lengths, self.shape))

    @property
    def volume(self):
        return prod(self.spacing)

    def face_shape(self, c):
        if self.periodic and c == 0:
            return tuple(self.shape)
        return tuple(n + (axis == c) for axis, n in enumerate(self.shape))

    def coordinates(self, c=None, *, device='cpu', dtype=torch.float64):
        """Positions, shape (n0, n1, 2): cell centres (c None) or the faces
        of component c; the channel starts at the origin."""
        counts = self.shape if c is None else self.face_shape(c)
        out = torch.empty((*counts, 2), device=device, dtype=dtype)
        for a, (n, h) in enumerate(zip(counts, self.spacing)):
            index = torch.arange(n, device=device, dtype=dtype) + (0 if c == a else .5)
            out[..., a] = (h * index).view([n if b == a else 1 for b in range(2)])
        return out

    def zeros(self, *, device='cpu', dtype=torch.float64):
        return tuple(torch.zeros(self.face_shape(c), device=device, dtype=dtype) for c in range(2))

    def check_velocity(self, velocity):
        """Two components on their face shapes, one device and one dtype."""
        if len(velocity) != 2:
            raise ValueError('two velocity components required')
        lead = velocity[0]
        if any(tuple(u.shape) != self.face_shape(c) or u.dtype != lead.dtype or u.device != lead.device
               for c, u in enumerate(velocity)):
            raise ValueError('channel velocity shape/device/dtype mismatch')

    def abi(self):
        g = _abi.Grid2()
        for a in range(2):
            g.n[a] = self.shape[a]
            g.h[a] = self.spacing[a]
            g.lengths[a] = self.lengths[a]
        g.periodic = int(self.periodic)
        return g
