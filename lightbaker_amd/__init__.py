"""lightbaker_amd: jer-nc/lightbaker's ray-cast bake on one AMD MI355X (gfx950).

The HIP kernels and their C-ABI live in liblightbaker_amd.so (built from csrc/ by
`python -m lightbaker_amd.build`); this package binds them with ctypes and restates the
reference's host API around them.
"""
from ._native import FORM_PLAIN, FORM_TUNED, NODES_QUANT, NODES_WIDE  # noqa: F401
