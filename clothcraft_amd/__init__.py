"""clothcraft_amd — MI355X-native Vertex Block Descent cloth step with
Offset Geometric Contact, behind the C ABI of include/clothvbd.h.

Host-side mirror of the reference stepping API: `VbdSolver`, `GpuVbd`,
`VbdSettings`, `OgcSettings`, `Grab` (src/vbd.rs, src/gpu_vbd.rs), plus scene
generators (`scenes`).  The HIP library libclothvbd.so is built in-tree by
`clothcraft_amd.build.build()`.
"""
from .vbd import GpuVbd, Grab, OgcSettings, VbdError, VbdSettings, VbdSolver, version  # noqa: F401
