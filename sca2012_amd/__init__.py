"""sca2012_amd: the adaptive SPH step of finallyjustice/sca2012 on MI355X (gfx950).

The product is the HIP shared library ``libsph_amd.so`` with the C ABI of
``include/sph_amd.h``; this package holds its sources (``csrc/``), the build recipe
(``build.py``), a ctypes binding (``_abi.py``) and the scene seeding (``scenes.py``).
"""
from ._abi import (V2D, V3D2, V3D4, STEP_SPLIT, STEP_MERGE, STEP_MC, SphError, SPHContext, SPHBatch, SPHAdjoint,
                   add_extern_force, load_library, param_defaults, param_derive, particle_dtype, param_dtype,
                   PARTICLE2D_DTYPE, PARTICLE3D_DTYPE, SYSPARAM2D_DTYPE, SYSPARAM3D_DTYPE)

__all__ = ["V2D", "V3D2", "V3D4", "STEP_SPLIT", "STEP_MERGE", "STEP_MC", "SphError", "SPHContext", "SPHBatch", "SPHAdjoint",
           "add_extern_force", "load_library", "param_defaults", "param_derive", "particle_dtype", "param_dtype",
           "PARTICLE2D_DTYPE", "PARTICLE3D_DTYPE", "SYSPARAM2D_DTYPE", "SYSPARAM3D_DTYPE"]
