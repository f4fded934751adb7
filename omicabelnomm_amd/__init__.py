"""MI355X (gfx950) implementation of OmicABELnoMM's blocked
normal-equation OLS scan.  The compute lives in liboagpu.so (HIP, C ABI in
include/oagpu.h); this package holds its Python binding, the host-side
mirror of the reference's Settings/Algorithm interface and the synthetic
workload generator."""
from .lib import OagpuError  # noqa: F401
from .scan import Algorithm, Outputs, Scanner, Settings, normalize_thresholds, tvalue  # noqa: F401

__all__ = ["Algorithm", "Outputs", "Scanner", "Settings", "normalize_thresholds", "tvalue", "OagpuError"]
