"""MI355X Barnes-Hut engine for ultracold-plasma MD (HIP kernels for gfx950)."""
from .engine import (  # noqa: F401
    EPS, NDIST, NOALLOC_MAX_TARGETS, PT_DTYPE, TPS, XM_E, XM_I, BuildLocalTree, Engine, Ensemble, Force,
    LocalCoulombSum, TreeError, VVerletNotemp1, VVerletNotemp2, gen_plasma, kelbg_length, load_library, make_config,
)
