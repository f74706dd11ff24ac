"""MI355X-native regular-force engine for NBODY6++GPU.

The product is the C-ABI library ``lib/libgpunb.so`` (HIP, gfx950) built from
``csrc/``; ``gpunb`` is its Python front end with the reference's function
names, ``ics`` generates the synthetic star-cluster inputs.
"""
__all__ = ["gpunb", "ics"]
