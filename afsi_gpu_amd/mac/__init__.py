"""3D MAC grid: IB transfer backend, multigrid pressure solve, projection."""
from .grid import MACGrid
from .multigrid import MGOptions, GeometricMultigrid, MixedMultigrid
from .flow import MACFlow, MACFlowResult

__all__ = ['MACGrid', 'MGOptions', 'GeometricMultigrid', 'MixedMultigrid', 'MACFlow', 'MACFlowResult']
