"""MI355X-native Fourier-domain projection gridder.

A drop-in for the gridder API of tagarelab/MultiGPUGridder (setVolume,
forwardProject, backProject, resetVolume, getVol, reconstructVol) over the C
ABI of ``libMultiGPUGridder.so`` (include/mgg.h), whose kernels are
hand-written HIP for gfx950.
"""
from ._lib import LibraryNotBuilt, MggError, lib  # noqa: F401
from .gridder import MultiGPUGridder  # noqa: F401
from .server import Server, pinned_empty  # noqa: F401
from .autograd import ForwardProjection, project, project_transpose  # noqa: F401
from .mrc import read_mrc, write_mrc  # noqa: F401
from .hostmath import kb_table, precomp, uniform_axes  # noqa: F401
from .synth import blob_volume, random_directions  # noqa: F401

__version__ = "0.1.0"
