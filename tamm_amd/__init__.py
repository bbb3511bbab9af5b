"""MI355X-native block-sparse tensor-contraction engine with TAMM's contraction semantics.

The engine is the C ABI in include/tamm_mi355x.h (libtamm_mi355x.so, HIP for gfx950).
`tamm_amd.engine` binds it with ctypes (the library is loaded on first use);
`tamm_amd.configs` holds the workload definitions.
"""
from . import configs, engine  # noqa: F401

__all__ = ["configs", "engine"]
