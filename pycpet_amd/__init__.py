"""MI355X (gfx950) implementation of PyCPET's Coulomb-grid and streamline
hot paths behind the reference's calculator / options-file interface.

Native code: pycpet_amd/csrc (HIP kernels + C-ABI, built into
pycpet_amd/libcpet_hip.so by pycpet_amd.build).  Host side: _native (ctypes
binding), ops (array API), c_ops (Math_ops drop-in), and the calculator /
CPET / io mirrors of the reference's interface.
"""
__version__ = "0.1.0"
