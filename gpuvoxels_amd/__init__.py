"""MI355X-native Voxelyze3 physics engine (host-side Python mirror).

The engine itself is C++/HIP behind the C ABI in include/vx3h.h; this package
holds the ctypes bindings, model descriptions and synthetic workloads.
"""
