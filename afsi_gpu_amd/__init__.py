"""afsi_gpu_amd: MI355X-native MAC/IB per-timestep hot path.

Drop-in for the afsi_torch IB backend (``mac.cuda_ib``) and MAC pressure
solver (``mac.multigrid``, ``mac.flow``) of loveIroha/AFSI_GPU, executed by
hand-written gfx950 HIP kernels in ``libafsi_hip.so`` through a C ABI
(``include/afsi_hip.h``). There is no CPU fallback: every GPU entry point
raises if the library is missing.
"""
__version__ = '0.1.0'
