"""spike_pstrsv for AMD MI355X: Spike sparse triangular solve on gfx950.

The product is the C ABI in include/spike_pstrsv.h, implemented by
libspike_pstrsv.so (host C++ + HIP kernels).  `native` binds it through ctypes;
`gen` builds the synthetic workloads.
"""
__all__ = ["native", "gen"]
