"""parallax_amd -- MI355X-native Gaussian-splat training step.

The engine's splat trainer (render, photometric loss, adjoint, Adam, densify)
as hand-written HIP kernels for gfx950 behind a C ABI (include/splat_hip.h),
built into libparallax_splat.so. `parallax_amd.splat` binds that ABI for
Python; `parallax_amd.synthetic` holds the seeded scenes and cameras.
"""

__all__ = ["splat", "synthetic"]
