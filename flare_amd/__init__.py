"""flare_amd: FLARE's sparse-GP force evaluation on MI355X (gfx950).

The prediction path of FLARE (neighbour list -> B2 descriptor ->
NormalizedDotProduct against the sparse set -> SparseGP mean and local
uncertainty) as hand-written HIP kernels behind a C ABI
(include/flare_mi.h), with ``flare_amd._C_flare`` mirroring the reference's
pybind11 module for that path.
"""
from . import _lib  # noqa: F401
from ._lib import FORM_CHOSEN, FORM_FP64  # noqa: F401

__all__ = ["FORM_FP64", "FORM_CHOSEN"]
