"""qsdvr_amd -- MI355X-native (gfx950) QSDVR render hot path.

Modules:
  qsdvr    drop-in for the reference extension module (14 ops + CameraArgs)
  render   autograd layer (models/render.py caller contract)
  camera   per-camera ray lists (datasets/utils.py caller contract)
  fused    fused plain renderer with the recompute adjoint
  proposal proposal MLP + inverse-CDF fine sampler
  scenes   synthetic workloads (cameras, quadric and appearance grids)
  build    hipcc build of libqsdvr_hip.so
"""
__version__ = "0.1.0"
