"""ctypes signatures of the step-level C ABI (include/pgdg.h)."""
import ctypes as C


def declare(L: C.CDLL) -> None:
    """Attach argtypes/restype for every step-level entry point present."""
    return None
