"""MI355X-native ray-tracing path of AviatorEngine2: HIP kernels for gfx950
behind a C ABI (include/avrt.h), with a Python mirror of the reference's
Renderer interface for tests and benchmarks."""
from . import abi  # noqa: F401

__all__ = ["abi"]
