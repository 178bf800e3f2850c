"""MI355X-native drop-in for the reference's `xpbd_gpu` package.

Modules keep the reference's names: `neohookean_solver` (XPBDNeoHookeanSolver,
calc_first_lame, calc_second_lame), `xpbd_base` (XPBDSolverBase,
calc_tet_volume), `sdf` (PlaneSdf, SphereSdf, CubeSdf, HangSdfModel) and `constants` (EPSILON).
`mesh` replaces the MeshTaichi patcher's loader.  The compute path is
libxpbd_hip.so (hand-written gfx950 kernels behind a C ABI).
"""
