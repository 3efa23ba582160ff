"""phasecon_amd -- MI355X-native MBIR engine (drop-in for the reference's
reconstruct() C library).  Native code lives in csrc/ and is built into
libmbir4d.so by ``python -m phasecon_amd.build``."""
from .lib import load, LIB_PATH  # noqa: F401

__all__ = ["load", "LIB_PATH"]
