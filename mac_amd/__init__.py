"""mac_amd -- MI355X-native MAC-grid free-surface projection (HIP, gfx950).

The product is the C-ABI shared library ``mac_amd/libmac_amd.so`` declared in
``include/mac_amd.h``.  This package only binds it with ctypes for Python
callers (tests, benchmarks) and provides the seeded workload generators.
Nothing here computes a projection on the CPU: when the library is missing
the import of :mod:`mac_amd.lib` fails loudly.
"""
