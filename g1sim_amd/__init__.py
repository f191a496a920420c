"""g1sim_amd -- MI355X-native step engine for the Unitree G1 (29-DoF, floating base).

The product is the HIP library ``libg1sim_amd.so`` (C-ABI in
include/g1sim_amd.h) and the torch drop-in ``G1Sim`` (g1sim_amd.sim).
Submodules are imported lazily so that the numpy-only parts (model data,
workload generators) load without torch.
"""
NB, NQ, NV, NU, NC = 30, 36, 35, 29, 8


def __getattr__(name):
    if name == "G1Sim":
        from .sim import G1Sim
        return G1Sim
    raise AttributeError(name)
