"""plume_amd: an MI355X (gfx950) implementation of Plume's simulation substep and volume ray march.

`plume_amd.sim` and `plume_amd.render` keep the call surface of the reference's `plume.sim` and
`plume.render` (same names, arguments and behaviour), over float32 torch tensors on the GPU; every
operator is a hand-written HIP kernel in libplume_hip.so, reached through the C ABI of
include/plume_hip.h. There is no CPU compute path.
"""

__version__ = "0.1.0"
