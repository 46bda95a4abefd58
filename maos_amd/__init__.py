"""MI355X-native Shack-Hartmann physical-optics WFS path of MAOS.

The product is the C ABI in include/maos_mi355x.h (libmaos_mi355x.so, HIP
kernels for gfx950).  This package holds its sources (csrc/), the host-side
table builders (host/, include/maos_setup.h) and thin Python front ends:
  powfs    host setup of a powfs (geometry, DTF, ETF, LLT, screens, MF)
  engine   the GPU engine through the C ABI
  configs  the workloads in MAOS configuration keys
"""
