"""MI355X reverse k-ranks (US sampled-score-table path) -- host-side mirror.

The compute path is librkr.so (HIP kernels for gfx950 behind the C ABI in
include/rkr.h); this package only binds it.
"""
from ._lib import RkrError, load  # noqa: F401
from .index import Index, QueryResult, UserRankElement, build_index  # noqa: F401
from . import formats, serve, synth, tune  # noqa: F401
from .serve import Server  # noqa: F401
