"""MI355X-native split-form DGSEM right-hand-side engine (FLUXO hot path).

The compute path is the HIP library libfluxo_gpu.so (C ABI in include/fluxo_gpu.h);
fluxo_amd.native binds it for Python callers.
"""
