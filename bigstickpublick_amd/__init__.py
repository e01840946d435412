"""MI355X engine for BIGSTICK's M-scheme Hamiltonian application and Lanczos
driver.  The compute lives in libbigstick_amd.so (HIP, gfx950) behind the C
ABI declared in include/bigstick_amd.h; this module is a thin ctypes binding
plus torch plumbing for device memory and streams."""
from .engine import Batch, BatchResult, LanczosResult, StrengthResult, System, load_library  # noqa: F401
