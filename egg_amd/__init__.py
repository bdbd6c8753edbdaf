"""MI355X-native structured TMOP mesh smoother / untangler.

Compute core: hand-written HIP kernels for gfx950 behind the C ABI in
``include/eggmi.h`` (``egg_amd/_lib/libeggmi.so``), bound to Python as
``egg_amd._cpp.cpp_core`` with the reference's ``egg._cpp.cpp_core`` surface.
"""

__version__ = "0.1.0"
