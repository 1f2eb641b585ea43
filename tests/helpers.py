"""Small numpy helpers shared by the tests (no oracle logic here)."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def barycentric_of_ref(pts):
    """Barycentric coordinates of points on the [-1,1] reference simplex."""
    pts = np.atleast_2d(pts)
    lam_rest = (1.0 + pts) / 2.0
    lam0 = 1.0 - lam_rest.sum(axis=1, keepdims=True)
    return np.hstack([lam0, lam_rest])


def dof_positions(coords, elems, elem2dof, ref_nodes, ndof, ncomp=1):
    """Physical position of every (scalar) global dof; checks consistency.

    Returns (pos[ndof//ncomp...], max_mismatch) where max_mismatch is the largest
    distance between two placements of the same global dof seen from different
    elements (0 for a conforming numbering)."""
    lam = barycentric_of_ref(ref_nodes)              # nloc x (d+1)
    X = coords[elems]                                # ne x (d+1) x d
    P = np.einsum("lv,evd->eld", lam, X)             # ne x nloc x d
    nl = lam.shape[0]
    dofs = elem2dof[:, :nl] if ncomp > 1 else elem2dof
    pos = np.full((ndof, coords.shape[1]), np.nan)
    flat_d = dofs.reshape(-1)
    flat_p = P.reshape(-1, coords.shape[1])
    pos[flat_d] = flat_p
    mism = np.abs(pos[flat_d] - flat_p).max()
    return pos, mism


def rel_fro(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)
