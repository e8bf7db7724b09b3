"""argus -- Python binding of libargus_hip.so (MI355X per-signal extraction).

Add the package directory (argus-sdr_amd/) to sys.path and `import argus`.
"""
from ._lib import ErrUnavailable, GpudError, EXPORTS, LIB_PATH, lib  # noqa: F401
from . import gpudemod  # noqa: F401
from . import jobs  # noqa: F401
from . import rds_runtime  # noqa: F401
from . import serve  # noqa: F401
from . import surveillance  # noqa: F401
from . import wfm  # noqa: F401
