"""MI355X-native FTLM hot path for the 2D extended Hubbard model.

Mirrors the reference package `ftlm` (python/ftlm/__init__.py): HubbardParams, nsites and
version_string, plus the hot-path API (FockBasis, apply_extended_hubbard, Lanczos, FTLM) backed by
the gfx950 library libftlm_hip.so.  Unlike the reference there is no silent fallback: importing a
hot-path symbol without the built library raises.
"""

try:  # torch first: its bundled HIP runtime must be the one every library of this package binds to
    import torch  # noqa: F401
except ImportError:
    pass

from . import _lib
from .api import (
    FockBasis,
    FtlmParams,
    HubbardOperator,
    HubbardParams,
    KBlockOperator,
    KStackOperator,
    KGramOperator,
    KrylovSession,
    LanczosExtrema,
    MomentumSector,
    OrbitKRepEntries,
    RectLattice,
    RitzData,
    apply_extended_hubbard,
    build_hubbard_geometry,
    canonical_states,
    device_count,
    derivative_operator,
    exact_log_partition,
    ftlm_gradient,
    ftlm_log_partition_complex,
    ftlm_log_partition_gradient,
    ftlm_log_partition_vjp,
    ftlm_vjp,
    FtlmTape,
    ftlm_observables,
    ftlm_log_partition_observables,
    geometry_derivative,
    ftlm_log_partition_real,
    ftlm_ritz,
    ftlm_thermo,
    lanczos_extrema,
    lanczos_tridiagonal,
    nsites,
    pack_raw_state,
    quadrature_ritz,
    translate_states,
    translation_bloch_phase,
    tridiagonal_extrema,
    version_string,
)

__all__ = [
    "HubbardParams", "nsites", "version_string", "RectLattice", "build_hubbard_geometry", "FockBasis",
    "HubbardOperator", "apply_extended_hubbard", "LanczosExtrema", "lanczos_tridiagonal", "lanczos_extrema",
    "tridiagonal_extrema", "FtlmParams", "ftlm_log_partition_complex", "ftlm_log_partition_real", "RitzData",
    "ftlm_ritz", "ftlm_thermo", "quadrature_ritz", "device_count", "KrylovSession", "MomentumSector", "KBlockOperator", "KStackOperator", "KGramOperator", "OrbitKRepEntries", "pack_raw_state",
    "translate_states", "canonical_states", "translation_bloch_phase", "exact_log_partition", "geometry_derivative", "derivative_operator",
    "ftlm_gradient", "ftlm_log_partition_gradient", "ftlm_vjp", "ftlm_log_partition_vjp", "FtlmTape", "ftlm_observables", "ftlm_log_partition_observables", "_lib",
]
