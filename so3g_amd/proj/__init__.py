"""Drop-in counterpart of so3g.proj (python/proj/__init__.py)."""
from . import quat  # noqa: F401
from . import util  # noqa: F401
from .quat import DEG  # noqa: F401
from .coords import CelestialSightLine, EarthlySite, Assembly, FocalPlane, SITES  # noqa: F401
from .ranges import Ranges, RangesMatrix  # noqa: F401
from . import wcs  # noqa: F401
from .wcs import Projectionist, ProjectionistHealpix  # noqa: F401
from . import mapthreads  # noqa: F401
from . import wcslite  # noqa: F401
