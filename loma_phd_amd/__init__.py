"""MI355X-native LoMa RK3 time advance (HIP kernels for gfx950 behind a C ABI).

Modules:
  capi     -- ctypes binding of libloma_hip.so (include/loma_amd.h)
  grid     -- mesh generator and workload configurations
  ic       -- seeded initial condition (host set-up)
  restart  -- reference restart-file reader/writer
  serve    -- continuous-batching request loop
"""
__all__ = ["capi", "grid", "ic", "restart"]
