"""feellgood_amd -- an MI355X (gfx950) engine for FeeLLGood's per-timestep LLG path.

The hot path lives in the HIP library libfgmi.so behind the C-ABI of include/fgmi.h; the host
side (mesh input, preprocessing, .sol files, the time-loop driver) in libfgmi_host.so and the
fgmi_run executable. This package holds their Python bindings (ctypes) and the build recipe.
"""
from . import native  # noqa: F401
