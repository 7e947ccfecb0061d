"""ropesim for MI355X: the batched Cosserat-rod solver behind the reference's Python surface.

- RopeSim:   single rope, host-driven root (validation API).
- RopeBatch: many ropes (uneven lengths allowed) in one solver; roots driven from GPU tensors,
             positions / velocities / frames exposed zero-copy through __cuda_array_interface__.
- CudaView:  the device-array view type.
- ropesim.sdf: signed-distance grids (closed forms, triangle meshes baked on the GPU, trilinear host sampler).
- ropesim.diff: differentiable rollouts (torch.autograd.Function over the hand-written adjoint).

The solver is librope_hip.so (hand-written HIP for gfx950) behind a C ABI; _ropesim is its pybind11
boundary. Importing fails loudly when the built extension is missing: there is no CPU fallback.
Units: density = rho [kg/m], k_stretch = EA [N], k_bend = EI [N m^2] (> 0).
"""
try:  # let torch (when present) load its HIP runtime first so one runtime serves both
    import torch as _torch  # noqa: F401
except ImportError:  # pragma: no cover
    _torch = None

from . import sdf  # noqa: F401,E402

try:
    from ._ropesim import CudaView, RopeBatch, RopeSim, cell_variant_info, num_cell_variants  # noqa: F401,E402
except ImportError as e:  # pragma: no cover
    raise ImportError("ropesim: the HIP extension is not built (python -m ropesim._build): " + str(e)) from e

if _torch is not None:
    from . import diff  # noqa: F401,E402

__all__ = ["RopeSim", "RopeBatch", "CudaView", "sdf", "diff"]
