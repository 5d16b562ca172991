"""MI355X implementation of MSTM's solution hot path.

The product is the C-ABI library ``libmstm_hip.so`` (declared in
``include/mstm_hip.h``).  This package holds its HIP sources (``csrc/``),
the Fortran ISO_C_BINDING shim (``fortran/``), the build script, seeded
geometry generators and a thin ctypes view of the C-ABI that mirrors the
reference's Fortran procedure names (``mstm_amd.solver``).
"""
from .lib import MstmHipError, load_library  # noqa: F401
