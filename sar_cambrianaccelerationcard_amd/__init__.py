"""MI355X (gfx950) time-domain back-projection SAR imager.

The hot path is libsarbp.so (hand-written HIP kernels behind the C ABI in
include/sarbp.h).  Python modules here are host glue: `native` binds the ABI,
`scene` simulates seeded stripmap scenes, `device` moves them into HBM,
`metrics` scores images.
"""
__version__ = "0.1.0"
