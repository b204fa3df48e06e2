"""2D staggered channel on the GPU (mirror of afsi_torch.mac2d)."""
from .grid import ChannelGrid
from .multigrid import ChannelMultigrid
from .flow import ChannelFlow, ChannelResult, divergence, correct
from .transfer import Stencil, prepare, spread, gather
from .step import ChannelIBStep

__all__ = ['ChannelGrid', 'ChannelMultigrid', 'ChannelFlow', 'ChannelResult', 'divergence',
           'correct', 'Stencil', 'prepare', 'spread', 'gather', 'ChannelIBStep']
