"""HFPx3D hot path on MI355X: displacement-discontinuity BEM assembly, dense
LU / mixed-precision solve and stress postprocessing as HIP kernels behind a
C ABI (include/hfpx3d.h)."""
__all__ = ["hfp3d", "meshgen"]
