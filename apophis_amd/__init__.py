"""apophis_amd -- MI355X (gfx950) hydro flux path of AlexHls/apophis.

PLM/PPM reconstruction, HLLE/HLLC Riemann solvers and the flux-divergence RK2 update
as hand-written HIP kernels behind a C ABI (include/apophis_hydro.h), with the
reference's Hydro package / task interface mirrored in :mod:`apophis_amd.hydro`.
Importing this package does not load the native library; :mod:`apophis_amd.hydro`
does, and fails loudly if it has not been built.
"""
__all__ = ["mesh", "problems"]
