"""Host-side Python binding of the rpcf C ABI (include/rpcf.h, include/rpcf_group.h)."""
from .binding import (FIELD_OX, FIELD_OY, FIELD_OZ, FIELD_U, FIELD_V, FIELD_W, HIP_LIB, ORACLE_LIB, Group, Params,
                      RpcfError, Simulation, load, make_params, seed_state, sizes, slot_bytes)

__all__ = ["FIELD_U", "FIELD_V", "FIELD_W", "FIELD_OX", "FIELD_OY", "FIELD_OZ", "HIP_LIB", "ORACLE_LIB",
           "Group", "Params", "RpcfError", "Simulation", "load", "make_params", "seed_state", "sizes", "slot_bytes"]
