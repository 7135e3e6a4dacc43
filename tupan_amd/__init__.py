"""tupan_amd -- MI355X (gfx950) build of Tupan's direct-summation kernels.

Layout:
  csrc/          hand-written HIP kernels and the C ABI (include/tupan_hip.h)
  build.py       in-tree build of lib/libtupan_hip.so and the libtupan.h
                 drop-in libraries lib/libtupan_f{64,32}.so
  _lib.py        ctypes binding; no CPU fallback
  spec.py        kernel table (fields, outputs, flop conventions)
  kernels.py     host-array calls (numpy in, numpy out)
  device.py      device-tensor calls (torch HIP tensors, zero copy)
  extensions.py  drop-in for tupan.lib.extensions (classes + singletons)
  particles.py   minimal SoA ParticleSystem with the reference caller contract
  ics.py         Plummer sphere / mass functions (reference RNG order)
"""
__version__ = "0.1.0"
