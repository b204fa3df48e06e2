"""Three-dimensional MAC grid (x, y, z), closed Neumann box or periodic.

Mirrors ``MACGrid`` of mac/grid.py:13-59 (same fields, spacing, volume,
face_shape, face_origin, zeros, coordinates, check_velocity). ``periodic``
is a build-own extension: every face array then has n entries per axis and
face i of component c sits at origin_c + i*h_c.
"""
from dataclasses import dataclass
from math import isfinite, prod
import torch

from .._abi import Grid3


@dataclass(frozen=True)
class MACGrid:
    shape: tuple
    lengths: tuple
    origin: tuple = (0., 0., 0.)
    periodic: bool = False

    def __post_init__(self):
        if len(self.shape) != 3 or any(type(n) is not int or n < 4 for n in self.shape):
            raise ValueError('MAC requires three integer cell counts >= 4')
        if (len(self.lengths) != 3 or len(self.origin) != 3 or
                any(not isfinite(v) or v <= 0 for v in self.lengths) or
                any(not isfinite(v) for v in self.origin)):
            raise ValueError('finite origin and positive box lengths required')

    @property
    def spacing(self):
        return tuple(L / n for L, n in zip(self.lengths, self.shape))

    @property
    def volume(self):
        return prod(self.spacing)

    @property
    def cells(self):
        return prod(self.shape)

    def face_shape(self, component):
        if self.periodic:
            return tuple(self.shape)
        return tuple(n + (axis == component) for axis, n in enumerate(self.shape))

    def face_origin(self, component):
        return tuple(o + (0. if axis == component else .5 * h)
                     for axis, (o, h) in enumerate(zip(self.origin, self.spacing)))

    def zeros(self, *, device='cpu', dtype=torch.float64):
        return tuple(torch.zeros(self.face_shape(c), device=device, dtype=dtype) for c in range(3))

    def coordinates(self, component=None, *, device='cpu', dtype=torch.float64):
        """Positions, shape (n0, n1, n2, 3): cell centres (component None)
        or the faces of velocity component ``component``; entry [i, j, k, a]
        is start_a + h_a * index_a."""
        if component is None:
            counts = self.shape
            start = tuple(o + .5 * h for o, h in zip(self.origin, self.spacing))
        else:
            counts, start = self.face_shape(component), self.face_origin(component)
        out = torch.empty((*counts, 3), device=device, dtype=dtype)
        for a, (n, h, s) in enumerate(zip(counts, self.spacing, start)):
            line = s + h * torch.arange(n, device=device, dtype=dtype)
            out[..., a] = line.view([n if b == a else 1 for b in range(3)])
        return out

    def check_velocity(self, velocity):
        """Three components on their face shapes, on one device, in one
        floating dtype (FP32 or FP64)."""
        if len(velocity) != 3:
            raise ValueError('three MAC velocity components required')
        lead = velocity[0]
        same = all(u.device == lead.device and u.dtype == lead.dtype for u in velocity)
        shaped = all(tuple(u.shape) == self.face_shape(c) for c, u in enumerate(velocity))
        if not (same and shaped and lead.dtype in (torch.float32, torch.float64)):
            raise ValueError('MAC field shape/device/dtype mismatch')

    def abi(self):
        """The afsi_grid3 struct; centre origins use face_origin's own FP64
        expression so integer stencil bases match the reference bit for bit."""
        g = Grid3()
        h = self.spacing
        for a in range(3):
            g.n[a] = self.shape[a]
            g.h[a] = h[a]
            g.origin[a] = self.origin[a]
            g.center_origin[a] = self.origin[a] + .5 * h[a]
        g.periodic = 1 if self.periodic else 0
        return g
