"""Host-side Python mirror of the gfx950 SMS library (ctypes over include/sms.h)."""
from . import abi, scenes  # noqa: F401
