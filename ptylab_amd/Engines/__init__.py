"""Engines of the hot path (PtyLab Engines/__init__.py names)."""

from .BaseEngine import BaseEngine  # noqa: F401
from .Batched import Batched  # noqa: F401
from .e3PIE import e3PIE  # noqa: F401
from .ePIE import ePIE  # noqa: F401
from .mPIE import mPIE  # noqa: F401
from .mqNewton import mqNewton  # noqa: F401
from .pcPIE import pcPIE  # noqa: F401
from .multiPIE import multiPIE  # noqa: F401
from .qNewton import qNewton  # noqa: F401
