"""ct2 -- MI355X (gfx950) implementation of ct-2's reconstruction hot path.

Modules named after the reference's f2py extension modules
(``view_proj3D``, ``view_proj3D_tomo``, ``ray_proj3D``, ``totalvar``) and
its ``FDK`` functions call hand-written HIP kernels in ``libct2.so`` through
ctypes; ``tomo3D`` carries the object model (sinogram3D / image3D /
linogram3D) whose methods are the reference's call sites.  Geometry and
phantom helpers are pure Python and import without a GPU.
"""
__version__ = "0.1.0"
