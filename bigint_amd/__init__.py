"""MI355X-native three-prime NTT multiply engine (gfx950), a drop-in for the reference's
BigInt/BigFloat multiply path.  The compute lives in libbigint_hip.so (HIP kernels + C ABI,
include/bigint_hip.h); this package holds its Python binding and limb utilities."""
from . import limbs  # noqa: F401

__all__ = ["limbs", "engine"]
