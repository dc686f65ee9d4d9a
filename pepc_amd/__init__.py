"""MI355X-native Coulomb Barnes-Hut tree solve (PEPC's grow + traverse path).

* ``pepc_amd.solver.Solver``  -- Python front end over the C-ABI
* ``pepc_amd.workloads``       -- seeded synthetic particle systems
* ``include/pepc_gpu.h``       -- the C-ABI (libpepc_gpu.so)
* ``pepc_amd/fortran``         -- iso_c_binding shim with PEPC's Fortran API
"""
__all__ = ["solver", "workloads"]
