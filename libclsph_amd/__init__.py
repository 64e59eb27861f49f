"""libclsph_amd: MI355X-native (gfx950 HIP) implementation of libclsph's
per-substep SPH pipeline behind a C ABI (include/clsph_hip.h).

`native` binds the C ABI; `scenes` builds the synthetic configurations;
`layout` holds numpy views of the reference's particle / parameter structs.
"""
__version__ = "0.1.0"
