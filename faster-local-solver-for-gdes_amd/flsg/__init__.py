"""flsg -- MI355X-native local graph-diffusion push solvers.

Host package: device-resident graphs, solve calls into libflsg.so (HIP
kernels for gfx950), graph generators and the reference's harness
conventions.  The reference-compatible entry points are the drop-in modules
``ppr_solver``, ``katz_solver``, ``hk_solver`` and ``InstantGNN`` next to this
package.
"""
from . import _abi
from ._abi import (FLSG_CHEB, FLSG_FP32, FLSG_FP64, FLSG_GD, FLSG_HK, FLSG_KATZ, FLSG_PPR)

__all__ = ["_abi", "FLSG_PPR", "FLSG_KATZ", "FLSG_HK", "FLSG_GD", "FLSG_CHEB", "FLSG_FP64", "FLSG_FP32"]
