"""Parameter containers for the half-sarcomere hot path.

Mirrors the reference's two-tier split (multifil_jax/core/params.py):

* ``StaticParams`` (params.py:563-737) -- structural, shape-defining fields.
  Frozen dataclass; ``replace()`` derives a modified copy.
* ``DynamicParams`` (params.py:739-857) -- the 49 sweepable physics values,
  in the reference's field order (``_DYNAMIC_DEFAULTS``, params.py:73-553).
  Values are plain numpy float32 scalars or (B,) float32 arrays; the device
  path packs them into one float32 row per instance (``DynamicParams.row``).

Units follow the reference: ms, nm, pN, kT, pCa, degrees Celsius.
"""
from __future__ import annotations

import warnings
from dataclasses import dataclass, asdict, fields as dc_fields
from typing import Dict, Tuple

import numpy as np

# Field order and skeletal defaults (params.py:73-553).  The order is part of
# the device ABI: a parameter row is laid out in exactly this order.
DYNAMIC_DEFAULTS: Dict[str, float] = {
    'thick_k': 7500.0,
    'thin_k': 5500.0,
    'xb_c_rest_weak': 0.82309,
    'xb_c_rest_strong': 1.27758,
    'xb_c_k_weak': 8.0,
    'xb_c_k_strong': 40.0,
    'xb_g_rest_weak': 19.93,
    'xb_g_rest_strong': 16.47,
    'xb_g_k_weak': 0.4,
    'xb_g_k_strong': 5.0,
    'titin_a': 43.0,
    'titin_b': 0.004,
    'titin_rest': 215.0,
    'tm_k_01': 100000.0,
    'tm_k_12': 1.0,
    'tm_k_23': 0.1,
    'tm_k_30': 0.2,
    'tm_Keq_01': 500000.0,
    'tm_Keq_12': 130.0,
    'tm_Keq_23': 0.1,
    'tm_J_C': 0.0,
    'tm_J_M': 2.70,
    'xb_r01_coeff': 305.99,
    'xb_r12_coeff': 0.6,
    'xb_r23_coeff': 0.15,
    'xb_r34_coeff': 0.6,
    'xb_delta_23': 1.0,
    'xb_delta_34': 0.5,
    'xb_r40': 0.1,
    'xb_r04': 0.01,
    'xb_r05': 0.007,
    'xb_U_DRX': -2.3,
    'xb_U_loose': -4.3,
    'xb_U_tight_1': -15.0,
    'xb_U_tight_2': -21.0,
    'xb_srx_k0': 0.007,
    'xb_srx_kmax': 0.4,
    'xb_srx_b': 5.0,
    'xb_srx_ca50': 1e-6,
    'temp_celsius': 26.15,
    'solver_tol': 0.3,
    'pCa': 4.5,
    'z_line': 900.0,
    'lattice_spacing': 14.0,
}
DYNAMIC_FIELDS: Tuple[str, ...] = tuple(DYNAMIC_DEFAULTS)
N_PARAM_FIELDS = len(DYNAMIC_FIELDS)
PARAM_INDEX = {name: i for i, name in enumerate(DYNAMIC_FIELDS)}

# Shape-defining fields (params.py:61).
STATIC_FIELDS = frozenset({
    'n_crowns', 'n_polymers_per_thin', 'solver_max_iter', 'actin_geometry',
    'n_newton_steps', 'n_cg_steps', 'solver_residual_tol', 'n_xb_bins',
    'xb_bin_lo', 'xb_bin_hi', 'thick_bare_zone', 'thick_crown_spacing',
    'actin_half_pitch', 'mono_per_poly', 'polymer_base_turns',
    'target_zone_wiggle', 'n_xb_per_crown'})


@dataclass(frozen=True)
class StaticParams:
    """Structural configuration (params.py:563-737); same fields and defaults."""
    n_crowns: int = 52
    n_polymers_per_thin: int = 15
    solver_max_iter: int = 50
    actin_geometry: str = "vertebrate"
    n_newton_steps: int = 4
    n_cg_steps: int = 6
    solver_residual_tol: float = 1.5
    n_xb_bins: int = 200
    xb_bin_lo: float = -8.0
    xb_bin_hi: float = 35.0
    thick_bare_zone: float = 58.0
    thick_crown_spacing: float = 14.3
    n_superlattice_classes: int = 1
    actin_half_pitch: float = 36.0
    mono_per_poly: int = 26
    polymer_base_turns: float = 12.0
    target_zone_wiggle: float = float(np.float32(2 * np.pi) / np.float32(24))
    n_xb_per_crown: int = 3
    crown_rotation_deg: float = 60.0
    crown_face_wiggle_deg: float = 15.0
    legacy_crown_geometry: bool = False

    def __post_init__(self):
        # Same validation as params.py:710-737.
        if self.legacy_crown_geometry and self.n_xb_per_crown != 3:
            raise ValueError(
                "legacy_crown_geometry=True requires n_xb_per_crown == 3 "
                f"(the legacy face table has no n={self.n_xb_per_crown} analogue)")
        if self.legacy_crown_geometry and (
                self.crown_rotation_deg != 60.0 or self.crown_face_wiggle_deg != 15.0):
            warnings.warn(
                "legacy_crown_geometry=True ignores crown_rotation_deg/"
                "crown_face_wiggle_deg; non-default values here are a silent no-op.",
                stacklevel=2)
        if self.n_superlattice_classes not in (1, 3):
            raise ValueError(
                "n_superlattice_classes must be 1 (no superlattice) or 3 "
                f"(3-class axial superlattice), got {self.n_superlattice_classes}")
        if self.actin_geometry not in ("vertebrate", "invertebrate"):
            raise ValueError(f"unknown actin_geometry {self.actin_geometry!r}")

    def replace(self, **kwargs) -> 'StaticParams':
        return StaticParams(**{**asdict(self), **kwargs})

    def __repr__(self) -> str:
        return (f"StaticParams(n_crowns={self.n_crowns}, "
                f"n_polymers_per_thin={self.n_polymers_per_thin}, "
                f"actin_geometry='{self.actin_geometry}')")


class DynamicParams:
    """The sweepable physics values (params.py:739-857).

    Each field holds a float32 scalar or a (B,) float32 array.  ``copy`` checks
    field names like the reference (params.py:803-820); ``with_drivers``
    replaces only pCa / z_line / lattice_spacing (params.py:825-854).
    """

    __slots__ = DYNAMIC_FIELDS

    def __init__(self, **kwargs):
        unknown = set(kwargs) - set(DYNAMIC_FIELDS)
        if unknown:
            raise ValueError(f"Unknown parameter: {sorted(unknown)}. "
                             f"Valid: {list(DYNAMIC_FIELDS)}")
        for name, default in DYNAMIC_DEFAULTS.items():
            object.__setattr__(self, name,
                               np.asarray(kwargs.get(name, default), dtype=np.float32))

    def __setattr__(self, key, value):
        raise AttributeError("DynamicParams is immutable; use copy(**updates)")

    def to_dict(self) -> Dict[str, float]:
        return {n: (float(getattr(self, n)) if np.ndim(getattr(self, n)) == 0
                    else np.asarray(getattr(self, n)).tolist())
                for n in DYNAMIC_FIELDS}

    def copy(self, **updates) -> 'DynamicParams':
        invalid = set(updates) - set(DYNAMIC_FIELDS)
        if invalid:
            raise ValueError(f"Unknown parameter: {sorted(invalid)}. "
                             f"Valid: {list(DYNAMIC_FIELDS)}")
        kw = {n: updates.get(n, getattr(self, n)) for n in DYNAMIC_FIELDS}
        return DynamicParams(**kw)

    def with_drivers(self, pCa, z_line, lattice_spacing) -> 'DynamicParams':
        return self.copy(pCa=pCa, z_line=z_line, lattice_spacing=lattice_spacing)

    @property
    def batch_size(self) -> int:
        sizes = {np.asarray(getattr(self, n)).shape[0]
                 for n in DYNAMIC_FIELDS if np.ndim(getattr(self, n)) > 0}
        if len(sizes) > 1:
            raise ValueError(f"inconsistent batch sizes across fields: {sorted(sizes)}")
        return sizes.pop() if sizes else 1

    def row(self) -> np.ndarray:
        """Pack a scalar DynamicParams into one float32 row in field order."""
        return np.array([float(getattr(self, n)) for n in DYNAMIC_FIELDS], dtype=np.float32)

    def rows(self, batch: int) -> np.ndarray:
        """(batch, N_PARAM_FIELDS) float32, broadcasting scalar fields."""
        out = np.empty((batch, N_PARAM_FIELDS), dtype=np.float32)
        for i, n in enumerate(DYNAMIC_FIELDS):
            out[:, i] = np.broadcast_to(np.asarray(getattr(self, n), np.float32), (batch,))
        return out

    @staticmethod
    def from_row(row) -> 'DynamicParams':
        row = np.asarray(row, dtype=np.float32)
        return DynamicParams(**{n: row[i] for i, n in enumerate(DYNAMIC_FIELDS)})

    def __repr__(self) -> str:
        return (f"DynamicParams(thick_k={float(np.ravel(self.thick_k)[0]):.1f}, "
                f"thin_k={float(np.ravel(self.thin_k)[0]):.1f}, ...)")


Constants = DynamicParams


def get_skeletal_params() -> Tuple[StaticParams, DynamicParams]:
    """Fast-twitch skeletal defaults (params.py:860-897)."""
    return StaticParams(), DynamicParams()


def get_cardiac_params() -> Tuple[StaticParams, DynamicParams]:
    """Cardiac overrides on the skeletal defaults (params.py:900-979)."""
    overrides = {
        'tm_k_12': 0.5, 'tm_k_01': 80000.0, 'tm_Keq_01': 750000.0, 'tm_k_30': 0.04,
        'xb_r12_coeff': 0.175, 'xb_r23_coeff': 0.065, 'xb_r34_coeff': 0.065,
        'xb_r05': 0.2, 'xb_srx_k0': 0.005,
        'titin_a': 55.0, 'titin_b': 0.008, 'titin_rest': 140.0,
    }
    return StaticParams(), DynamicParams(**overrides)


_IFM_STATIC = dict(
    actin_geometry='invertebrate', thick_crown_spacing=14.5, actin_half_pitch=38.7,
    mono_per_poly=28, polymer_base_turns=13, n_xb_per_crown=4,
    crown_rotation_deg=33.75, target_zone_wiggle=float(np.radians(26.0)),
    n_crowns=100, n_polymers_per_thin=20)
_IFM_DYNAMIC = dict(titin_a=40.0, titin_b=0.025, titin_rest=50.0)


def get_lethocerus_params() -> Tuple[StaticParams, DynamicParams]:
    """Insect flight muscle preset (params.py:982-1203)."""
    return StaticParams(**_IFM_STATIC), DynamicParams(**_IFM_DYNAMIC)


def get_drosophila_params() -> Tuple[StaticParams, DynamicParams]:
    """Drosophila preset: IFM plus the 3-class superlattice (params.py:1206-1281)."""
    return (StaticParams(**_IFM_STATIC, n_superlattice_classes=3),
            DynamicParams(**_IFM_DYNAMIC))


def static_field_names():
    return [f.name for f in dc_fields(StaticParams)]
