"""Lattice topology: which filament is where and what can reach what.

Host-side NumPy restatement of the reference's ``SarcTopology.create``
(multifil_jax/core/sarc_geometry.py:379-645 and helpers :725-1492).  The
arrays it produces are the static input of the device kernels, so every one
of them must be element-for-element what the reference builds; the float32
arithmetic of the reference's geometry helpers is reproduced where it decides
a discrete outcome (site selection, face matching).

The fields keep the reference's names and shapes so callers can switch over;
``pack()`` (in ``engine.py``) turns them into the flat device tables.
"""
from __future__ import annotations

from typing import List, Optional, Tuple

import numpy as np

from .params import StaticParams, DynamicParams

SQRT3 = np.float32(np.sqrt(3))
THICK_THIN_DISTANCE = np.float32(1.0 / np.sqrt(3))
THICK_THIN_DISTANCE_INVERTEBRATE = np.float32(0.5)
N_ACTIN_REGISTRATION_CLASSES = 3

# Six hexagonal face directions (sarc_geometry.py:88-95).
ORIENTATION_VECTORS = np.array([
    [0.866, -0.5], [0.0, -1.0], [-0.866, -0.5],
    [-0.866, 0.5], [0.0, 1.0], [0.866, 0.5]], dtype=np.float32)
THICK_FACE_ANGLES = np.array([150, 90, 30, 330, 270, 210], dtype=np.float32) * np.pi / 180
THICK_FACE_ANGLES_INVERTEBRATE = np.array([0, 60, 120, 180, 240, 300], dtype=np.float32) * np.pi / 180
THIN_ORIENTATION_UPWARD = (4, 0, 2)
THIN_ORIENTATION_DOWNWARD = (3, 5, 1)
_LEGACY_FACE_PATTERN = np.array([[0, 2, 4], [1, 3, 5], [0, 2, 4]])


# --------------------------------------------------------------------------
# lattice placement
# --------------------------------------------------------------------------

def _hex_thick_positions(n_rows: int, n_cols: int):
    """Row-major hexagonal thick grid (sarc_geometry.py:725-745)."""
    one = np.float32(1.0)
    pts = []
    for row in range(n_rows):
        for col in range(n_cols):
            x_off = np.float32(0.5) * one if row % 2 == 1 else np.float32(0.0)
            pts.append([col * one + x_off, -row * SQRT3 / np.float32(2.0) * one])
    box_x = np.float32(n_cols) * one
    box_y = np.float32(n_rows) * SQRT3 / np.float32(2.0) * one
    return np.array(pts, dtype=np.float32), box_x, box_y


def _wrap_y(y, box_y):
    while y < -box_y / 2:
        y += box_y
    while y > box_y / 2:
        y -= box_y
    return np.float32(y)


def _wrap_x(x, box_x, tol=1e-6):
    x = x % box_x
    if abs(x - box_x) < tol or abs(x) < tol:
        x = 0.0
    return x


def _min_image(dx, dy, box_x, box_y):
    return dx - box_x * np.round(dx / box_x), dy - box_y * np.round(dy / box_y)


def _periodic_distance(p, q, box_x, box_y):
    dx, dy = _min_image(p[0] - q[0], p[1] - q[1], box_x, box_y)
    return np.float32(np.sqrt(dx * dx + dy * dy))


def _thin_index_at(thin_pos, x, y, box_x, box_y, periodic, tol=0.01):
    for idx, pos in enumerate(thin_pos):
        d = (_periodic_distance([x, y], pos, box_x, box_y) if periodic
             else np.sqrt((x - pos[0]) ** 2 + (y - pos[1]) ** 2))
        if d < tol:
            return idx
    return None


def _matching_thin_face(thick_pos, thin_pos, orientations, box_x, box_y, periodic):
    """Thin face whose direction is closest to the thick filament (:786-808)."""
    dx = thick_pos[0] - thin_pos[0]
    dy = thick_pos[1] - thin_pos[1]
    if periodic:
        dx, dy = _min_image(dx, dy, box_x, box_y)
    ang = np.arctan2(dy, dx)
    best, best_diff = 0, np.inf
    for face, o in enumerate(orientations):
        v = ORIENTATION_VECTORS[o]
        diff = abs(ang - np.arctan2(v[1], v[0]))
        if diff > np.pi:
            diff = 2 * np.pi - diff
        if diff < best_diff:
            best_diff, best = diff, face
    return best


def _thins_at_interstices(thick_pos, box_x, box_y, periodic):
    """Vertebrate thin placement at trigonal interstices (:811-912).

    Thins are enumerated row by row (top row first), each row's thick
    filaments left to right, upper faces (0,1,2) of the whole row before its
    lower faces (5,4,3); orientation is 'upward' when at least as many of a
    thin's thick neighbours lie below it as above.
    """
    tol = 0.05

    def site_of(tp, face):
        ang = THICK_FACE_ANGLES[face]
        tx = tp[0] + THICK_THIN_DISTANCE * np.cos(ang)
        ty = tp[1] + THICK_THIN_DISTANCE * np.sin(ang)
        if periodic:
            tx = _wrap_x(tx, box_x)
            ty = _wrap_y(ty, box_y)
        return tx, ty, (round(tx / tol), round(ty / tol))

    members = {}
    for ti, tp in enumerate(thick_pos):
        for face in range(6):
            tx, ty, key = site_of(tp, face)
            if key not in members:
                members[key] = {'pos': (np.float32(tx), np.float32(ty)), 'faces': []}
            members[key]['faces'].append((ti, face))

    rows = {}
    for ti, tp in enumerate(thick_pos):
        rows.setdefault(round(tp[1] / 0.1), []).append((ti, tp))

    order, seen = [], set()
    for rk in sorted(rows, reverse=True):
        in_row = sorted(rows[rk], key=lambda t: t[1][0])
        for faces in ([0, 1, 2], [5, 4, 3]):
            for _, tp in in_row:
                for face in faces:
                    _, _, key = site_of(tp, face)
                    if key not in seen:
                        seen.add(key)
                        order.append(members[key])

    positions, orientations = [], []
    for m in order:
        tx, ty = m['pos']
        above = below = 0
        for ti, _ in m['faces']:
            thy = thick_pos[ti][1]
            if periodic:
                dy = thy - ty
                dy = dy - box_y * np.round(dy / box_y)
                above += dy > tol
                below += dy < -tol
            else:
                above += thy > ty + tol
                below += thy < ty - tol
        positions.append([tx, ty])
        orientations.append(THIN_ORIENTATION_UPWARD if above <= below
                            else THIN_ORIENTATION_DOWNWARD)
    return np.array(positions, dtype=np.float32), orientations


def _thins_at_edges(thick_pos, box_x, box_y, periodic):
    """Invertebrate thin placement at thick-thick edge midpoints (:915-1005)."""
    tol = 0.1
    nb_angles = np.array([0, 60, 120, 180, 240, 300], dtype=np.float32) * np.pi / 180
    o_angles = np.array([np.arctan2(v[1], v[0]) for v in ORIENTATION_VECTORS])

    def best_orientation(a):
        d = np.abs(o_angles - a)
        d = np.minimum(d, 2 * np.pi - d)
        return int(np.argmin(d))

    def thick_at(x, y):
        for idx, pos in enumerate(thick_pos):
            if periodic:
                dx, dy = _min_image(x - pos[0], y - pos[1], box_x, box_y)
                d = np.sqrt(dx ** 2 + dy ** 2)
            else:
                d = np.sqrt((x - pos[0]) ** 2 + (y - pos[1]) ** 2)
            if d < tol:
                return idx
        return None

    found, seen = [], set()
    for i, pi in enumerate(thick_pos):
        for ang in nb_angles:
            nx, ny = pi[0] + np.cos(ang), pi[1] + np.sin(ang)
            if periodic:
                nx, ny = _wrap_x(nx, box_x), _wrap_y(ny, box_y)
            j = thick_at(nx, ny)
            if j is None:
                continue
            tx, ty = pi[0] + 0.5 * np.cos(ang), pi[1] + 0.5 * np.sin(ang)
            if periodic:
                tx, ty = _wrap_x(tx, box_x), _wrap_y(ty, box_y)
            ek = (round(round(tx, 4) % box_x, 3), round(round(ty, 4) % box_y, 3))
            if ek in seen:
                continue
            seen.add(ek)
            a_i = ang + np.pi
            if a_i > np.pi:
                a_i -= 2 * np.pi
            oi, oj = best_orientation(a_i), best_orientation(ang)
            if oi == oj:
                oj = (oi + 3) % 6
            found.append((np.float32(tx), np.float32(ty), (oi, oj), (i, j)))
    found.sort(key=lambda t: (-t[1], t[0]))
    if not found:
        return np.zeros((0, 2), np.float32), [], []
    return (np.array([[t[0], t[1]] for t in found], dtype=np.float32),
            [t[2] for t in found], [t[3] for t in found])


def _connectivity(thick_pos, thin_pos, thin_orient, box_x, box_y, periodic,
                  n_faces_per_thin, dist, thin_pairs):
    """thick face <-> thin face map (:1008-1088)."""
    n_thick, n_thin = len(thick_pos), len(thin_pos)
    t2n = [[None] * 6 for _ in range(n_thick)]
    n2t = [[None] * n_faces_per_thin for _ in range(n_thin)]
    if n_faces_per_thin == 2 and thin_pairs is not None:
        for ni in range(n_thin):
            for face, ti in enumerate(thin_pairs[ni]):
                dx = thin_pos[ni][0] - thick_pos[ti][0]
                dy = thin_pos[ni][1] - thick_pos[ti][1]
                if periodic:
                    dx, dy = _min_image(dx, dy, box_x, box_y)
                a = np.arctan2(dy, dx)
                best, best_diff = 0, np.inf
                for tf in range(6):
                    diff = abs((a - THICK_FACE_ANGLES_INVERTEBRATE[tf] + np.pi) % (2 * np.pi) - np.pi)
                    if diff < best_diff:
                        best_diff, best = diff, tf
                n2t[ni][face] = (ti, best)
                t2n[ti][best] = (ni, face)
    else:
        for ti in range(n_thick):
            tp = thick_pos[ti]
            for tf in range(6):
                a = THICK_FACE_ANGLES[tf]
                ex = tp[0] + dist * np.cos(a)
                ey = tp[1] + dist * np.sin(a)
                if periodic:
                    ex = ex % box_x
                    ey = _wrap_y(ey, box_y)
                ni = _thin_index_at(thin_pos, ex, ey, box_x, box_y, periodic, tol=0.1)
                if ni is None:
                    continue
                nf = _matching_thin_face(tp, thin_pos[ni], thin_orient[ni], box_x, box_y, periodic)
                t2n[ti][tf] = (ni, nf)
                n2t[ni][nf] = (ti, tf)
    return t2n, n2t


def _spread_step(nv):
    """Step coprime to nv nearest nv/golden ratio (:1095-1106)."""
    target = max(1, round(nv / 1.6180339887))
    for d in range(nv):
        for s in (target - d, target + d):
            if 1 <= s < nv and np.gcd(s, nv) == 1:
                return s
    return 1


def _spread_starts(n, low, high):
    nv = high - low
    return (low + (np.arange(n) * _spread_step(nv)) % nv).astype(np.int32)


def _superlattice_classes(n_thick, ncols, n_classes):
    """Axial 3-colouring of thick filaments (:1118-1131)."""
    idx = np.arange(n_thick)
    row, col = idx // ncols, idx % ncols
    q = col - (row - (row & 1)) // 2
    return ((q - row) % n_classes).astype(np.int32)


def _thin_registration_classes(thick_pos, thin_pairs, box_x, box_y, periodic):
    """Edge-direction class of each invertebrate thin (:1134-1170)."""
    if thin_pairs is None:
        return None
    out = np.empty(len(thin_pairs), dtype=np.int32)
    for k, (i, j) in enumerate(thin_pairs):
        dx = thick_pos[j][0] - thick_pos[i][0]
        dy = thick_pos[j][1] - thick_pos[i][1]
        if periodic:
            dx, dy = _min_image(dx, dy, box_x, box_y)
        ang = np.degrees(np.arctan2(dy, dx)) % 180.0
        out[k] = int(np.round(ang / 60.0)) % N_ACTIN_REGISTRATION_CLASSES
    return out


def _crown_offsets(n_crowns, bare_zone, spacing):
    """Crown positions from the M-line and per-segment rests (:1173-1191)."""
    bare_zone = np.asarray(bare_zone, dtype=np.float32)
    spacing = np.float32(spacing)
    offs = bare_zone[:, None] + np.arange(n_crowns, dtype=np.float32)[None, :] * spacing
    rests = np.full((bare_zone.shape[0], n_crowns), spacing, dtype=np.float32)
    rests[:, 0] = bare_zone
    return offs.astype(np.float32), rests


def _binding_sites(orientations, thin_starts, n_thin, n_poly, half_pitch, mono_per_poly,
                   base_turns, wiggle, thin_class=None):
    """Actin helix -> binding sites per thin (:1194-1360).

    A monomer is a site of a face when its azimuth is within ``wiggle`` of
    the face direction.  Sites are ordered by decreasing offset from the
    Z-line (increasing axial position); the TM chain of a site is its
    monomer's parity; prev/next are same-chain neighbours in site order
    (self-references at chain ends).
    """
    if thin_class is None:
        thin_class = np.zeros(n_thin, dtype=np.int32)
    ang_step = int(round(mono_per_poly / N_ACTIN_REGISTRATION_CLASSES))
    ax_step = np.float32(half_pitch / N_ACTIN_REGISTRATION_CLASSES)
    base_len = np.float32(2 * half_pitch)
    turns = np.float32(base_turns)
    rev = np.float32(2 * np.pi)
    pitch = turns * rev / mono_per_poly
    rise = base_len / mono_per_poly
    n_mono = mono_per_poly * n_poly
    mono_offsets = (n_mono - np.arange(n_mono, dtype=np.float32)) * rise
    wig = np.float32(wiggle)

    out = []
    for ti in range(n_thin):
        cls = int(thin_class[ti])
        start = thin_starts[ti] + cls * ang_step
        offs_this = mono_offsets + cls * ax_step
        angles = np.array([((m + start + 1) % mono_per_poly) * pitch % rev
                           for m in range(n_mono)], dtype=np.float32)
        vecs = ORIENTATION_VECTORS[list(orientations[ti])]
        face_ang = np.arctan2(vecs[:, 1], vecs[:, 0])
        face_ang = np.where(face_ang < 0, face_ang + rev, face_ang)
        mono_in_face = [np.where(np.abs(angles - fa) < wig)[0] for fa in face_ang]
        offs_by_face = [offs_this[mi] for mi in mono_in_face]
        if offs_by_face and any(len(f) > 0 for f in offs_by_face):
            flat = np.sort(np.hstack(offs_by_face))[::-1]
        else:
            flat = offs_this[::-1]
        chain = []
        for off in flat:
            hit = np.where(np.abs(offs_this - off) < 1e-6)[0]
            chain.append((hit[0] if len(hit) else 0) % 2)
        chain = np.array(chain, dtype=np.int32)
        n_s = len(flat)
        prev = np.arange(n_s, dtype=np.int32)
        nxt = np.arange(n_s, dtype=np.int32)
        for c in (0, 1):
            idx = np.where(chain == c)[0]
            prev[idx[1:]] = idx[:-1]
            nxt[idx[:-1]] = idx[1:]
        faces = []
        for fo in offs_by_face:
            sites = []
            for off in fo:
                hit = np.where(np.abs(flat - off) < 1e-6)[0]
                if len(hit):
                    sites.append(hit[0])
            faces.append(np.array(sites, dtype=np.int32))
        if len(flat) > 1:
            rests = np.append(flat[:-1] - flat[1:], flat[-1])
        else:
            rests = np.array([flat[0] if len(flat) else rise], dtype=np.float32)
        out.append(dict(offsets=flat.astype(np.float32), rests=rests.astype(np.float32),
                        faces=faces, chain=chain, prev=prev, next=nxt))
    return out


def _flat_xb_maps(t2n, face_sites, n_per_face, thick_starts, n_thick, n_crowns, width,
                  nxpc, rotation_deg, face_wiggle_deg, legacy):
    """Per-XB target thin/face, fixed-width candidate list, validity (:1370-1492)."""
    total = n_thick * n_crowns * nxpc
    thin_id = np.zeros(total, np.int32)
    thin_face = np.zeros(total, np.int32)
    cand = np.zeros((total, width), np.int32)
    valid = np.ones(total, bool)
    for xb in range(total):
        thick = xb // (n_crowns * nxpc)
        local = xb % (n_crowns * nxpc)
        crown, h = local // nxpc, local % nxpc
        if legacy:
            level = (crown + thick_starts[thick] - 1) % 3 + 1
            face, matched = _LEGACY_FACE_PATTERN[level - 1, h], True
        else:
            eff = crown + (thick_starts[thick] - 1)
            az = (h * (360.0 / nxpc) + eff * rotation_deg) % 360.0
            face = int(round(az / 60.0)) % 6
            res = abs(az - face * 60.0)
            res = min(res, 360.0 - res)
            matched = res < face_wiggle_deg
        if matched:
            ni, nf = t2n[thick, face, 0], t2n[thick, face, 1]
        else:
            ni, nf = -1, 0
        valid[xb] = ni >= 0
        if ni < 0:
            ni, nf = 0, 0
        thin_id[xb], thin_face[xb] = ni, nf
        nv = n_per_face[ni, nf]
        row = face_sites[ni, nf, :]
        first = row[0] if nv > 0 else 0
        cand[xb] = np.where(np.arange(width) < nv, row, first)
    return thin_id, thin_face, cand, valid


def xb_bin_edges(lo: float, hi: float, n_bins: int) -> np.ndarray:
    """Float32 bin edges as a float32 linspace computes them.

    The reference calls ``jnp.linspace(lo, hi, n+1, dtype=float32)``
    (sarc_geometry.py:598-603), which evaluates ``lo*(1-s) + hi*s`` with
    ``s = i/n`` in float32 and appends ``hi`` exactly.
    """
    lo32, hi32 = np.float32(lo), np.float32(hi)
    s = (np.arange(n_bins, dtype=np.float32) / np.float32(n_bins)).astype(np.float32)
    body = (lo32 * (np.float32(1.0) - s)).astype(np.float32) + (hi32 * s).astype(np.float32)
    return np.concatenate([body.astype(np.float32), np.array([hi32], np.float32)])


class SarcTopology:
    """Static lattice description (sarc_geometry.py:112-377), NumPy-resident.

    Integer attributes: n_thick, n_crowns, n_thin, n_sites, n_titin,
    n_xb_per_crown, n_faces_per_thin, max_sites_per_face, total_xbs, n_xb_bins.
    Array attributes carry the reference's names, shapes and dtypes.
    """

    ARRAY_FIELDS = (
        'xb_to_thin_id', 'xb_to_thin_face', 'xb_to_site_indices', 'xb_valid',
        'thick_to_thin', 'thin_to_thick', 'face_to_sites', 'n_sites_per_face',
        'titin_connections', 'crown_offsets', 'crown_rests', 'binding_offsets',
        'binding_rests', 'tm_chains', 'thick_starts', 'thin_starts', 'eye_4', 'eye_6',
        'xb_bin_edges', 'xb_bin_centers', 'tm_prev_neighbor', 'tm_next_neighbor')
    INT_FIELDS = ('n_thick', 'n_crowns', 'n_thin', 'n_sites', 'n_titin', 'n_xb_per_crown',
                  'n_faces_per_thin', 'max_sites_per_face', 'total_xbs', 'n_xb_bins')

    def __init__(self, **kw):
        for k in self.INT_FIELDS:
            setattr(self, k, int(kw.pop(k)))
        for k in self.ARRAY_FIELDS:
            setattr(self, k, kw.pop(k))
        self.nrows = kw.pop('nrows', None)
        self.ncols = kw.pop('ncols', None)
        self.periodic = kw.pop('periodic', True)
        if kw:
            raise TypeError(f"unexpected fields {sorted(kw)}")

    @property
    def n_dof(self) -> int:
        return self.n_thick * self.n_crowns + self.n_thin * self.n_sites

    @classmethod
    def create(cls, nrows: int, ncols: int, static_params: StaticParams = None,
               dynamic_params: DynamicParams = None, periodic: bool = True,
               lattice_spacing: float = 14.0, thin_starts: List[int] = None,
               thick_starts: List[int] = None) -> 'SarcTopology':
        """Build the lattice (sarc_geometry.py:379-645), same arguments."""
        sp = static_params if static_params is not None else StaticParams()
        vertebrate = sp.actin_geometry == "vertebrate"
        n_faces = 3 if vertebrate else 2

        thick_pos, box_x, box_y = _hex_thick_positions(nrows, ncols)
        n_thick = len(thick_pos)
        if thick_starts is None:
            thick_starts_arr = _spread_starts(n_thick, 1, sp.n_xb_per_crown + 1)
        else:
            if len(thick_starts) != n_thick:
                raise ValueError(f"thick_starts must have length n_thick={n_thick}, "
                                 f"got {len(thick_starts)}")
            thick_starts_arr = np.array(thick_starts, dtype=np.int32)

        pairs = None
        if vertebrate:
            thin_pos, orient = _thins_at_interstices(thick_pos, box_x, box_y, periodic)
        else:
            thin_pos, orient, pairs = _thins_at_edges(thick_pos, box_x, box_y, periodic)
        n_thin = len(thin_pos)
        thin_class = _thin_registration_classes(thick_pos, pairs, box_x, box_y, periodic)
        if thin_starts is None:
            thin_starts_arr = (_spread_starts(n_thin, 0, 26) if vertebrate
                               else np.zeros(n_thin, dtype=np.int32))
        else:
            if len(thin_starts) != n_thin:
                raise ValueError(f"thin_starts must have length n_thin={n_thin}, "
                                 f"got {len(thin_starts)}")
            thin_starts_arr = np.array(thin_starts, dtype=np.int32)

        t2n_list, n2t_list = _connectivity(
            thick_pos, thin_pos, orient, box_x, box_y, periodic, n_faces,
            THICK_THIN_DISTANCE if vertebrate else THICK_THIN_DISTANCE_INVERTEBRATE, pairs)
        t2n = np.full((n_thick, 6, 2), -1, np.int32)
        for ti, fl in enumerate(t2n_list):
            for f, c in enumerate(fl):
                if c is not None:
                    t2n[ti, f] = c
        n2t = np.full((n_thin, n_faces, 2), -1, np.int32)
        for ni, fl in enumerate(n2t_list):
            for f, c in enumerate(fl):
                if c is not None:
                    n2t[ni, f] = c

        if periodic:
            missing = np.argwhere(t2n[:, :, 0] < 0)
            if len(missing):
                t0, f0 = missing[0]
                raise ValueError(
                    f"thick_to_thin has {len(missing)} unconnected (thick, face) pair(s) "
                    f"under periodic=True (e.g. thick={t0}, face={f0}) -- malformed "
                    f"lattice for nrows={nrows}, ncols={ncols}, "
                    f"actin_geometry='{sp.actin_geometry}'. Invertebrate geometry "
                    "requires an even nrows.")
        if periodic and sp.n_superlattice_classes == 3 and (nrows % 2 or ncols % 3):
            raise ValueError(
                "n_superlattice_classes=3 requires nrows even AND ncols a multiple of 3 "
                f"under periodic=True (got nrows={nrows}, ncols={ncols})")

        titin = [(ti, f, c[0], c[1]) for ti in range(n_thick)
                 for f, c in enumerate(t2n_list[ti]) if c is not None]
        titin_arr = np.array(titin, np.int32) if titin else np.zeros((0, 4), np.int32)

        sl = _superlattice_classes(n_thick, ncols, sp.n_superlattice_classes)
        bare = (sp.thick_bare_zone + sl.astype(np.float32)
                * (sp.thick_crown_spacing / sp.n_superlattice_classes))
        crown_offsets, crown_rests = _crown_offsets(sp.n_crowns, bare, sp.thick_crown_spacing)

        sites = _binding_sites(orient, thin_starts_arr, n_thin, sp.n_polymers_per_thin,
                               sp.actin_half_pitch, sp.mono_per_poly, sp.polymer_base_turns,
                               sp.target_zone_wiggle, thin_class=thin_class)
        max_sites = max(len(s['offsets']) for s in sites)
        b_off = np.zeros((n_thin, max_sites), np.float32)
        b_rest = np.full((n_thin, max_sites), 2.77, np.float32)
        chains = np.zeros((n_thin, max_sites), np.int32)
        prev = np.tile(np.arange(max_sites, dtype=np.int32), (n_thin, 1))
        nxt = np.tile(np.arange(max_sites, dtype=np.int32), (n_thin, 1))
        for i, s in enumerate(sites):
            n = len(s['offsets'])
            b_off[i, :n] = s['offsets']
            b_rest[i, :n] = s['rests']
            chains[i, :n] = s['chain']
            prev[i, :n] = s['prev']
            nxt[i, :n] = s['next']
        width = max(len(f) for s in sites for f in s['faces'])
        f2s = np.full((n_thin, n_faces, width), -1, np.int32)
        npf = np.zeros((n_thin, n_faces), np.int32)
        for i, s in enumerate(sites):
            for j, f in enumerate(s['faces']):
                f2s[i, j, :len(f)] = f
                npf[i, j] = len(f)

        total = n_thick * sp.n_crowns * sp.n_xb_per_crown
        tid, tface, cand, valid = _flat_xb_maps(
            t2n, f2s, npf, thick_starts_arr, n_thick, sp.n_crowns, width,
            sp.n_xb_per_crown, sp.crown_rotation_deg, sp.crown_face_wiggle_deg,
            sp.legacy_crown_geometry)

        edges = xb_bin_edges(sp.xb_bin_lo, sp.xb_bin_hi, sp.n_xb_bins)
        centers = (np.float32(0.5) * (edges[:-1] + edges[1:])).astype(np.float32)

        return cls(
            n_thick=n_thick, n_crowns=sp.n_crowns, n_thin=n_thin, n_sites=max_sites,
            n_titin=len(titin), n_xb_per_crown=sp.n_xb_per_crown, n_faces_per_thin=n_faces,
            max_sites_per_face=width, total_xbs=total, n_xb_bins=sp.n_xb_bins,
            xb_to_thin_id=tid, xb_to_thin_face=tface, xb_to_site_indices=cand, xb_valid=valid,
            thick_to_thin=t2n, thin_to_thick=n2t, face_to_sites=f2s, n_sites_per_face=npf,
            titin_connections=titin_arr, crown_offsets=crown_offsets, crown_rests=crown_rests,
            binding_offsets=b_off, binding_rests=b_rest, tm_chains=chains,
            thick_starts=thick_starts_arr, thin_starts=thin_starts_arr,
            eye_4=np.eye(4, dtype=np.float32), eye_6=np.eye(6, dtype=np.float32),
            xb_bin_edges=edges, xb_bin_centers=centers,
            tm_prev_neighbor=prev, tm_next_neighbor=nxt,
            nrows=nrows, ncols=ncols, periodic=periodic)

    # ------------------------------------------------------------------
    # on-disk format: one .npz of the SoA arrays under the reference's field
    # names, plus the integer counts and a format tag / version
    FORMAT = "multifil-topology"
    FORMAT_VERSION = 1

    def save(self, path) -> None:
        """Write the topology to ``path`` (.npz, no pickled objects)."""
        meta = {k: np.int64(getattr(self, k)) for k in self.INT_FIELDS}
        arrays = {k: np.asarray(getattr(self, k)) for k in self.ARRAY_FIELDS}
        np.savez(path, format=np.array(self.FORMAT), format_version=np.int64(self.FORMAT_VERSION),
                 nrows=np.int64(-1 if self.nrows is None else self.nrows),
                 ncols=np.int64(-1 if self.ncols is None else self.ncols),
                 periodic=np.bool_(self.periodic), **meta, **arrays)

    @classmethod
    def load(cls, path) -> 'SarcTopology':
        """Read a topology written by ``save``; ValueError on another format,
        a newer version, a missing field or inconsistent counts."""
        with np.load(path, allow_pickle=False) as z:
            if "format" not in z.files or str(z["format"]) != cls.FORMAT:
                raise ValueError(f"{path}: not a {cls.FORMAT} file")
            ver = int(z["format_version"])
            if ver > cls.FORMAT_VERSION:
                raise ValueError(f"{path}: format version {ver} is newer than {cls.FORMAT_VERSION}")
            missing = [k for k in cls.INT_FIELDS + cls.ARRAY_FIELDS if k not in z.files]
            if missing:
                raise ValueError(f"{path}: missing fields {missing}")
            kw = {k: int(z[k]) for k in cls.INT_FIELDS}
            kw.update({k: np.array(z[k]) for k in cls.ARRAY_FIELDS})
            nr, nc = int(z["nrows"]), int(z["ncols"])
            kw.update(nrows=None if nr < 0 else nr, ncols=None if nc < 0 else nc,
                      periodic=bool(z["periodic"]))
        t = cls(**kw)
        t._check()
        return t

    def _check(self):
        """Shape and count consistency of the arrays the engine packs."""
        X = self.n_thick * self.n_crowns * self.n_xb_per_crown
        want = {
            'crown_offsets': (self.n_thick, self.n_crowns), 'crown_rests': (self.n_thick, self.n_crowns),
            'binding_offsets': (self.n_thin, self.n_sites), 'binding_rests': (self.n_thin, self.n_sites),
            'tm_prev_neighbor': (self.n_thin, self.n_sites), 'tm_next_neighbor': (self.n_thin, self.n_sites),
            'xb_to_thin_id': (X,), 'xb_to_thin_face': (X,), 'xb_valid': (X,),
            'xb_to_site_indices': (X, self.max_sites_per_face),
            'face_to_sites': (self.n_thin, self.n_faces_per_thin, self.max_sites_per_face),
            'n_sites_per_face': (self.n_thin, self.n_faces_per_thin),
            'xb_bin_edges': (self.n_xb_bins + 1,), 'xb_bin_centers': (self.n_xb_bins,),
        }
        if self.total_xbs != X:
            raise ValueError(f"total_xbs {self.total_xbs} != n_thick * n_crowns * n_xb_per_crown = {X}")
        for k, shape in want.items():
            if tuple(np.shape(getattr(self, k))) != shape:
                raise ValueError(f"{k}: shape {np.shape(getattr(self, k))}, expected {shape}")

    def valid_xb_targets(self) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        """Indices/targets of XBs with a real partner (sarc_geometry.py:285-316)."""
        idx = np.where(self.xb_valid)[0]
        return idx, self.xb_to_thin_id[idx], self.xb_to_thin_face[idx]

    def __repr__(self) -> str:
        return (f"SarcTopology(n_thick={self.n_thick}, n_crowns={self.n_crowns}, "
                f"n_thin={self.n_thin}, n_sites={self.n_sites}, "
                f"total_xbs={self.total_xbs}, n_titin={self.n_titin})")
