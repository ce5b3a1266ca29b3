"""MI355X-native big-integer multiply hot path of the Piculator pi calculator.

The GPU engine is libpicugpu.so (HIP kernels behind the C ABI of
include/picugpu.h); picugpu.py binds it for Python callers.
"""
