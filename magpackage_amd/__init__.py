"""MI355X-native element-assembly engine for MagPackage (Nedelec curl-curl K,
sigma-mass M, B-H Newton Jacobian/residual, COO->CSC).

The compute path is libmpk_assembly.so (hand-written gfx950 HIP kernels
behind the C ABI in include/mpk_assembly.h).  Submodules:

- ``assembler``   Mesh / DofHandler / Assembler mirroring src/Assembler.m
- ``mex``         the reference's MEX entry points (same signatures)
- ``material``    MaterialLib tables (PCHIP, Marrocco) and their packing
- ``fem_tables``  quadrature and reference shape-function tables
- ``meshgen``     seeded jittered Kuhn-box meshes
"""
__version__ = "0.1.0"
