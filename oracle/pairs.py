"""Minimum image and neighbour pairs of the oracle (float64).

* `min_image`: sequential reduction of a reduced lower-triangular box, along c, then b, then a,
  with nearest-integer multiples (pgm_jax/md/box.py:130-153).
* `pairs`: every unordered atom pair i < j with |min_image(r_i - r_j)| < rc, intra- and
  intermolecular (the rows of pgm_jax/md/forcefield.py:1140-1186 hold each of them twice).
  Intramolecular displacements are differences of offsets from the molecule's first atom
  (forcefield.py:1063-1073).  Brute force for small systems, scipy cKDTree for large
  orthorhombic boxes.

Parity: `min_image` is pinned against the brute-force minimum over the 27 neighbouring images in a
truncated octahedron (the reference's test_md.py:137-157 check); `pairs` against brute force.
"""

from __future__ import annotations

import numpy as np


def min_image(dx: np.ndarray, H: np.ndarray) -> np.ndarray:
    """Return the minimum-image displacements of a reduced lower-triangular box (box.py:130-153)."""
    dx = np.array(dx, float)
    dx = dx - np.round(dx[..., 2:3] / H[2, 2]) * H[2]
    dx = dx - np.round(dx[..., 1:2] / H[1, 1]) * H[1]
    dx = dx - np.round(dx[..., 0:1] / H[0, 0]) * H[0]
    return dx


def _candidates(pos: np.ndarray, H: np.ndarray, rc: float) -> tuple[np.ndarray, np.ndarray]:
    """Return candidate pairs i < j (a superset of the pairs within rc)."""
    n = len(pos)
    if n <= 6000:
        return np.triu_indices(n, k=1)
    if np.allclose(H, np.diag(np.diag(H))):
        from scipy.spatial import cKDTree

        L = np.diag(H)
        w = np.mod(pos, L)
        w = np.where(w >= L, 0.0, w)
        ij = cKDTree(w, boxsize=L).query_pairs(rc * (1 + 1e-9) + 1e-9, output_type="ndarray")
        i, j = np.minimum(ij[:, 0], ij[:, 1]), np.maximum(ij[:, 0], ij[:, 1])
        return i, j
    raise NotImplementedError("oracle pairs: large triclinic boxes are not supported")


def pairs(pos: np.ndarray, H: np.ndarray, rc: float, mol: np.ndarray, first: np.ndarray) -> dict:
    """Return the pairs within rc: {"i", "j", "x" (r_i - r_j, minimum image), "r", "intra"}."""
    pos = np.asarray(pos, float)
    i, j = _candidates(pos, H, rc)
    x = min_image(pos[i] - pos[j], H)
    r2 = np.sum(x * x, -1)
    keep = r2 < rc * rc
    i, j, x = i[keep], j[keep], x[keep]
    intra = mol[i] == mol[j]
    if np.any(intra):  # exact within-molecule displacements (forcefield.py:1063-1073)
        off = pos - pos[first][mol]
        x[intra] = min_image(off[i[intra]] - off[j[intra]], H)
    return {"i": i, "j": j, "x": x, "r": np.sqrt(np.sum(x * x, -1)), "intra": intra}
