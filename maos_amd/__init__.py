"""MI355X-native tomographic reconstruction solve (MAOS FDPCG-PCG).

Host FP64 setup in C++ and the GPU solve in HIP for gfx950, both in
libmaos_tomo.so behind the C ABI of include/maos_tomo.h.
"""
from . import configs  # noqa: F401
