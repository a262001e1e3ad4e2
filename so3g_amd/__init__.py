"""so3g_amd: an MI355X-native implementation of so3g's projection engine.

The ProjEng_* classes (one per coordinate system x spin x tiling x
interpolation, plus the HEALPix and Precomp engines) and the
so3g_amd.proj package (Projectionist, ProjectionistHealpix, FocalPlane,
CelestialSightLine, Assembly, quat, Ranges, RangesMatrix) are drop-in
replacements for so3g's; compute runs in the hand-written HIP kernels of
so3g_amd/lib/libso3g_hip.so (build with ``python -m so3g_amd.build``).
"""
import os

from .engine import make_engine_classes as _make_engine_classes
from .engine import PrecompPlan  # noqa: F401
from .batch import ObservationBatch  # noqa: F401
from .array_ops import nmat_detvecs_apply, process_cuts, translate_cuts  # noqa: F401
from .g3super import G3SuperTimestream  # noqa: F401
from .proj.ranges import Ranges, RangesInt32, RangesMatrix  # noqa: F401

globals().update(_make_engine_classes())

from . import proj  # noqa: E402,F401

__version__ = '0.1.0'


def useful_info():
    """Counterpart of so3g.useful_info() (src/main.cxx:26-39).  The thread
    count is the number of sample domains the thread-assignment helpers
    produce by default: OMP_NUM_THREADS if set, else the CPU count."""
    n = os.environ.get('OMP_NUM_THREADS')
    try:
        n = int(n) if n else (os.cpu_count() or 1)
    except ValueError:
        n = os.cpu_count() or 1
    return {'omp_num_threads': n, 'backend': 'hip', 'arch': 'gfx950'}
