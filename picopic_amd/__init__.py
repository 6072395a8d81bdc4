"""picopic_amd: MI355X-native per-step particle loop of PiCoPiC.

The HIP library (csrc/, built into _lib/libpicopic_hip.so) is the product;
``capi`` binds its C ABI (include/picopic.h) and ``config`` reads the
reference's JSON configuration schema.
"""
__all__ = ["capi", "build"]
