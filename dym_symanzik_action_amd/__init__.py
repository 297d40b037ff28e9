"""MI355X-native discrete-group Symanzik Monte Carlo (dym-mod-metro path).

Python mirror of the reference's interface for this path, over the C ABI of
``build/libdsa.so`` (``include/dsa.h``): ``load_group`` (group.c:13-81),
``Chain`` with ``update`` (dym-mod-metro.cpp:146-197) and ``measure``
(dym-mod-metro.cpp:281-318), the gauge field ``a`` (lattice.h:9) and the loop
functions of lattice.c.  All compute runs in the HIP kernels: the first call
into the library raises ImportError when ``build/libdsa.so`` is missing.
``cfgio`` holds the configuration-file formats and ``analysis`` the
Gamma-method errors and LLR density-of-states reconstruction (pure numpy).
"""
from .dsa import (  # noqa: F401
    GROUP_DIR,
    Batch,
    Chain,
    DSAError,
    Group,
    group_path,
    lib,
    load_group,
    schedule,
)

__all__ = ["Batch", "Chain", "Group", "DSAError", "load_group", "group_path", "lib", "GROUP_DIR", "schedule"]
