"""dlpx: MI355X-native force evaluation for DL_POLY-style MD (host side)."""
from . import systems, fileio  # noqa: F401
