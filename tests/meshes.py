"""Corner meshes of the unit square / cube whose cells are not axis-aligned
rectangles, for the parity tests of the element geometry (J, det J and
J^-T grad N on cells where J is not diagonal)."""
import numpy as np

# the shear of the parallelogram meshes: x -> S x (constant, non-symmetric J)
SHEAR = {2: np.array([[1.0, 0.35], [0.0, 1.0]]),
         3: np.array([[1.0, 0.35, 0.2], [0.0, 1.0, -0.25], [0.0, 0.0, 1.0]])}


def corner_mesh(orc, dim, n, p, kind, seed=0):
    """(nodes, cells) with n cells per axis, corner order as the DoF numbering
    expects for p (lexicographic for p = 1, counter-clockwise for p >= 2).
    kind 'jitter': every interior corner node moved by up to 0.3 h per axis
    (seeded), boundary nodes stay on the boundary, so the domain is still the
    unit square / cube; 'shear': every node mapped by SHEAR[dim], so every cell
    is the same parallelogram / parallelepiped."""
    nodes, cells = orc.structured_mesh(dim, n, n, n if dim == 3 else 0, 0 if p == 1 else 1)
    nodes = nodes.copy()
    if kind == "jitter":
        rng = np.random.default_rng(seed)
        x = nodes[:, :dim]
        inner = np.all((x > 1e-9) & (x < 1 - 1e-9), axis=1)
        x[inner] += rng.uniform(-0.3, 0.3, (inner.sum(), dim)) / n
        nodes[:, :dim] = x
    elif kind == "shear":
        nodes[:, :dim] = nodes[:, :dim] @ SHEAR[dim].T
    else:
        raise ValueError(kind)
    return nodes, cells


def interior(dn, dim, kind):
    """Dof nodes off the domain boundary (the unit cube, or its sheared image)."""
    x = dn[:, :dim]
    if kind == "shear":
        x = np.linalg.solve(SHEAR[dim], x.T).T
    return np.all((x > 1e-9) & (x < 1 - 1e-9), axis=1)
