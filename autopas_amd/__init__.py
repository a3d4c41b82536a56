"""autopas_amd -- MI355X-native engine for AutoPas's Lennard-Jones short-range path.

The compute lives in ``libljgpu.so`` (HIP kernels for gfx950 behind the C ABI in include/ljgpu.h);
this package only binds it.
"""
from .engine import (  # noqa: F401
    KERNEL_AUTO,
    KERNEL_CLUSTER_FP64,
    KERNEL_CLUSTER_FP64_N3,
    KERNEL_CLUSTER_JROWS_MIXED,
    KERNEL_CLUSTER_MIXED,
    KERNEL_CLUSTER_MIXED_N3,
    KERNEL_LC_FP64,
    KERNEL_VERLET_MIXED,
    KERNEL_NAMES,
    Engine,
    EngineError,
    generate_lattice,
    generate_slab,
    load_library,
    read_checkpoint,
    write_checkpoint,
)
