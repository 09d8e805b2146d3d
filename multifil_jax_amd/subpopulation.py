"""Mixed populations of motors with different kinetics (core/subpopulation.py).

A Subpopulation holds K populations, index 0 the wild type. Each population
is a dict of multiplicative scales on xb_* / tm_* DynamicParams fields;
mechanics always stay at the wild-type values. Three modes, as in the
reference:

* ``mean_field``: every head and site uses one generator, the
  fraction-weighted sum of the K population generators (summed before the
  matrix exponential).
* ``random``: each head and site independently belongs to population 1 with
  probability ``fraction`` and uses its own population's transition rows.
  Labels are drawn per simulation from the mask seed and the simulation
  index.
* ``c_zone``: heads on crowns whose axial offset lies in
  [c_zone_min_nm, c_zone_max_nm] belong to population 1. Thin sites in the
  same band do too, but only when a tm_* field is scaled.

The engine receives these as an ``engine.Populations`` (run() converts):
population parameter rows, mean-field fractions, or per-head / per-site
labels.

Random masks are drawn with numpy's PCG64. The reference draws them from
JAX's threefry stream (subpopulation.py:210-235), which does not run here.
The distribution is the same: independent Bernoulli(fraction) per head and
per site. The individual draws differ.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from .params import DYNAMIC_FIELDS, DynamicParams

# cMyBP-C band on the thick filament, nm from the M-line (subpopulation.py:75-76)
C_ZONE_MIN_NM = 350.0
C_ZONE_MAX_NM = 650.0


def _validate_scales(scales: Dict[str, float]) -> Dict[str, float]:
    """Unknown field names are rejected before any device work
    (subpopulation.py:79-87)."""
    invalid = set(scales) - set(DYNAMIC_FIELDS)
    if invalid:
        raise ValueError(f"Unknown subpopulation scale field(s): {sorted(invalid)}. "
                         f"Valid DynamicParams fields: {list(DYNAMIC_FIELDS)}")
    return dict(scales)


@dataclass
class Subpopulation:
    """K populations (index 0 = wild type) (subpopulation.py:90-192).

    scales: K dicts field -> factor; scales[0] is {}.
    fractions: (K,) weights summing to 1: the mean-field blend weights and
        the label probabilities of random masks.
    mode: 'mean_field' | 'random' | 'c_zone'.
    xb_mask / tm_mask: (total_xbs,) / (n_thin * n_sites,) int labels
        (c_zone), or None.
    seed: base seed of random masks.
    """
    scales: List[Dict[str, float]]
    fractions: np.ndarray
    mode: str
    xb_mask: Optional[np.ndarray] = None
    tm_mask: Optional[np.ndarray] = None
    seed: Optional[int] = None

    def __post_init__(self):
        self.fractions = np.asarray(self.fractions, dtype=np.float32)

    @property
    def K(self) -> int:
        return len(self.scales)

    def scaled_fields(self) -> tuple:
        """Sorted union of the fields any population scales."""
        fields = set()
        for s in self.scales:
            fields |= set(s)
        return tuple(sorted(fields))

    def scale_array(self, field_names: tuple) -> np.ndarray:
        """(K, F) float32 factors in ``field_names`` order (wild-type rows 1)."""
        return np.asarray([[float(s.get(f, 1.0)) for f in field_names] for s in self.scales],
                          dtype=np.float32)

    def apply_scales(self, constants: DynamicParams) -> List[DynamicParams]:
        """Each population's constants: the scales multiplied onto
        ``constants`` (entry 0 is ``constants`` itself)."""
        out = []
        for s in self.scales:
            if not s:
                out.append(constants)
                continue
            out.append(constants.copy(**{f: np.float32(getattr(constants, f)) * np.float32(v)
                                         for f, v in s.items()}))
        return out

    @classmethod
    def mean_field(cls, fraction: float, **scales) -> "Subpopulation":
        """One blended generator: (1 - fraction) Q_wt + fraction Q_mut."""
        _validate_scales(scales)
        return cls(scales=[{}, dict(scales)], fractions=[1.0 - fraction, fraction], mode="mean_field")

    @classmethod
    def random(cls, fraction: float, seed: int = 0, **scales) -> "Subpopulation":
        """Each head / site is mutant with probability ``fraction``; labels are
        drawn per simulation from ``seed`` and the simulation index."""
        _validate_scales(scales)
        return cls(scales=[{}, dict(scales)], fractions=[1.0 - fraction, fraction], mode="random",
                   seed=seed)

    @classmethod
    def c_zone(cls, topology, c_zone_min_nm: float = C_ZONE_MIN_NM,
               c_zone_max_nm: float = C_ZONE_MAX_NM, **scales) -> "Subpopulation":
        """Heads on crowns in the axial band are mutant; sites in the band too
        when a tm_* field is scaled."""
        _validate_scales(scales)
        xb_mask, tm_mask = c_zone_masks(topology, c_zone_min_nm, c_zone_max_nm, scales)
        frac = float(np.mean(xb_mask))
        return cls(scales=[{}, dict(scales)], fractions=[1.0 - frac, frac], mode="c_zone",
                   xb_mask=xb_mask, tm_mask=tm_mask)


def c_zone_masks(topology, lo_nm: float, hi_nm: float, scales: Dict[str, float]):
    """Per-head and per-site labels of an axial band (subpopulation.py:
    196-226): a head is labelled by its crown's offset (per filament, so a
    superlattice's classes differ); sites by the cumulative rest position of
    the site, and only when tm_* fields are scaled."""
    off = np.asarray(topology.crown_offsets)
    crown_in = ((off >= lo_nm) & (off <= hi_nm)).astype(np.int32)
    xb_mask = np.broadcast_to(crown_in[:, :, None],
                              (topology.n_thick, topology.n_crowns, topology.n_xb_per_crown)).reshape(-1)
    if any(f.startswith("tm_") for f in scales):
        site_pos = np.cumsum(np.asarray(topology.binding_rests), axis=1)
        tm_mask = ((site_pos >= lo_nm) & (site_pos <= hi_nm)).astype(np.int32).reshape(-1)
    else:
        tm_mask = np.zeros(topology.n_thin * topology.n_sites, dtype=np.int32)
    return xb_mask.astype(np.int32), tm_mask


def random_masks(seed: int, p_mut, topology, total_batch: int):
    """(total_batch, total_xbs) and (total_batch, n_sites) Bernoulli(p_mut)
    labels, one independent draw per simulation keyed by (seed, index)
    (the reference's generate_random_masks_batch, subpopulation.py:229-253,
    with numpy's generator in place of JAX's)."""
    p = np.broadcast_to(np.asarray(p_mut, np.float64), (total_batch,))
    X = topology.n_thick * topology.n_crowns * topology.n_xb_per_crown
    S = topology.n_thin * topology.n_sites
    xb = np.empty((total_batch, X), np.int32)
    tm = np.empty((total_batch, S), np.int32)
    for i in range(total_batch):
        rng = np.random.default_rng(np.random.SeedSequence([int(seed) & 0xffffffff, i]))
        xb[i] = rng.random(X) < p[i]
        tm[i] = rng.random(S) < p[i]
    return xb, tm


@dataclass
class ResolvedPopulations:
    """A Subpopulation (or a swept list of them) resolved per simulation
    (simulation.py:514-577)."""
    mode: str                      # 'mean_field' | 'explicit'
    K: int
    field_names: tuple
    has_xb: bool
    has_tm: bool
    scale_matrix: np.ndarray       # (B, K, F)
    fractions: np.ndarray          # (B, K)
    xb_mask: Optional[np.ndarray] = None   # (B, X)
    tm_mask: Optional[np.ndarray] = None   # (B, S)
    extra: dict = field(default_factory=dict)


def resolve(subpopulation, topology, total_batch: int, idx=None) -> ResolvedPopulations:
    """Validate and broadcast one Subpopulation, or gather a swept list by
    ``idx`` (the per-simulation list index)."""
    subpops = subpopulation if isinstance(subpopulation, list) else [subpopulation]
    modes = {sp.mode for sp in subpops}
    if len(modes) > 1:
        raise ValueError(f"All swept subpopulations must share one mode; got {sorted(modes)}")
    ks = {sp.K for sp in subpops}
    if len(ks) > 1:
        raise ValueError(f"All swept subpopulations must have the same K; got {sorted(ks)}")
    mode, K = subpops[0].mode, subpops[0].K
    if mode not in ("mean_field", "random", "c_zone"):
        raise ValueError(f"Unknown subpopulation mode '{mode}'")
    fields = set()
    for sp in subpops:
        fields |= set(sp.scaled_fields())
    bad = [f for f in fields if not (f.startswith("xb_") or f.startswith("tm_"))]
    if bad:
        raise ValueError(f"Subpopulation scales must be xb_* or tm_* fields; got {sorted(bad)}")
    names = tuple(sorted(fields))
    has_xb = any(f.startswith("xb_") for f in names)
    has_tm = any(f.startswith("tm_") for f in names)
    scales = np.stack([sp.scale_array(names) for sp in subpops])     # (V, K, F)
    fracs = np.stack([sp.fractions for sp in subpops])                # (V, K)
    if idx is None:
        idx = np.zeros(total_batch, np.int64)
    idx = np.asarray(idx)
    out = ResolvedPopulations("mean_field" if mode == "mean_field" else "explicit", K, names,
                              has_xb, has_tm, scales[idx], fracs[idx])
    if mode == "c_zone":
        out.xb_mask = np.stack([np.asarray(sp.xb_mask, np.int32) for sp in subpops])[idx]
        out.tm_mask = np.stack([np.asarray(sp.tm_mask, np.int32) for sp in subpops])[idx]
    elif mode == "random":
        seeds = {int(sp.seed or 0) for sp in subpops}
        if len(seeds) > 1:
            raise ValueError(f"All swept 'random' subpopulations must share one seed; got {sorted(seeds)}")
        out.xb_mask, out.tm_mask = random_masks(seeds.pop(), out.fractions[:, 1], topology, total_batch)
    return out


def population_params(params: np.ndarray, rp: ResolvedPopulations) -> np.ndarray:
    """(B, K, 44) population rows: each simulation's parameter row with the
    scaled fields multiplied in float32 (simulation.py:660-667)."""
    from .params import DYNAMIC_FIELDS as F
    params = np.asarray(params, np.float32)
    out = np.repeat(params[:, None, :], rp.K, axis=1).copy()
    for j, name in enumerate(rp.field_names):
        c = F.index(name)
        out[:, :, c] = params[:, None, c] * rp.scale_matrix[:, :, j]
    return out
