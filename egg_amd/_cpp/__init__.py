"""Native extension: ``cpp_core`` (pybind11 over the eggmi C ABI, HIP on gfx950)."""
