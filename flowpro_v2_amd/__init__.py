"""MI355X (gfx950) implementation of FlowPro's implicit-DG Newton step.

Modules:
  cases      synthetic meshes in FlowPro's format and workload parameters
  equation   model constants of the supported `Equation` plugins
  quadrature QuadratureCentral rule selection over the bundled gaussRules data
  native     ctypes binding of libflowpro_hip.so (include/flowpro_hip.h)
  solver     DeviceCase (one fp_ctx) and GpuImplicitSolver (LocalImplicitSolver loop)
  build      in-tree hipcc build of libflowpro_hip.so
"""
__all__ = ["cases", "equation", "quadrature", "native", "solver", "build"]
