"""stormm_amd: MI355X-native batched molecular-mechanics energy / force and minimization path.

The product is the HIP library libstormm_amd.so (C-ABI in include/stormm_amd.h).  This package is
its host-side mirror: topology intake (AMBER prmtop / inpcrd), synthetic workloads, and the
synthesis objects and launch functions of the reference's API.
"""
from .topology import Topology, read_inpcrd, read_prmtop, write_inpcrd, write_prmtop  # noqa: F401

__all__ = ["Topology", "read_prmtop", "read_inpcrd", "write_prmtop", "write_inpcrd"]
