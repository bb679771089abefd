"""pywave_amd -- MI355X-native acoustic finite-difference engine with pywave's Python API.

Drop-ins for the reference's hot path (github.com/chenyk1990/pywave):
  afd2d, afd3d   2D/3D acoustic FD modelling and time-reversal imaging (pywave/afd2d.py, afd3d.py)
  aps2d, aps3d   the same with the pseudo-spectral Laplacian (pywave/aps2d.py, aps3d.py)
  pfwi, fillpar  FWI forward modelling (mode 1), L-BFGS FWI (mode 2), passive modelling, LS-TRI
                 source inversion and passive velocity FWI (mode 4) (pywave/pfwi.py)
  afd2dc, afd3dc, aps2dc, aps3dc, forward_modeling_ac, fwic, lstric, pfwic
                 the extension-module functions pywave's namespace re-exports (and
                 forward_modelingc, which pywave/pfwi.py:72 calls but the reference lacks)
Extensions beyond the reference's Python surface:
  AfdPlan        resident modelling problem (set-up once, run many times)
  tri3d_image    batched time-reversal imaging of receiver groups with a fused image
  FwiProblem     FWI gradients (reference imaging condition or exact adjoint), batched shots
  FwiForward     torch.autograd.Function wrapper of the FWI forward map and its adjoint
  LstriProblem   resident passive modelling / time-reversal image / LS-TRI inversion
All numerics run in hand-written HIP kernels (pywave_amd/csrc); there is no CPU fallback.
"""
from .afd import afd2d, afd3d, AfdPlan, count_snaps  # noqa: F401
from .aps import aps2d, aps3d  # noqa: F401
from .pfwi import pfwi, fillpar  # noqa: F401
from .fwi import FwiProblem, FwiForward, fwi_traces, make_pararray  # noqa: F401
from .tri import tri3d_image, group_masks  # noqa: F401
from .batch import Afd2dBatch, afd2d_batch  # noqa: F401
from .lstri import LstriProblem  # noqa: F401
from .io import binread, binwrite  # noqa: F401
# the extension functions pywave's namespace re-exports (afdcfun, apscfun, pfwicfun)
from .cfun import (afd2dc, afd3dc, aps2dc, aps3dc, forward_modeling_ac, forward_modelingc, fwic,  # noqa: F401
                   lstric, pfwic)

__version__ = "0.1.0"
