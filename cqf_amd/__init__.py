"""cqf_amd -- counting quotient filter for AMD MI355X (gfx950).

libgqf.so (HIP kernels + C ABI, built by cqf_amd.build) is the product;
this package is its Python mirror (cqf_amd.gqf) plus the many-filter
batching and serving layers built on the same C ABI.
"""
from .gqf import GQF, load_library, to_device_u64, to_numpy_u64  # noqa: F401
