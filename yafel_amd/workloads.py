"""Synthetic workloads of the benchmark, built entirely on the device.

Each workload is a structured unit-square/cube mesh numbered by the device
DoFManager, a sparsity pattern + gather plan, the physics descriptor, the
Dirichlet mask and the right-hand side.  `step()` runs the hot path once:
assembly (values into the fixed pattern) -> Dirichlet -> Krylov solve.

    c1  2-D Poisson -lap u = 1, Q1 128x128, homogeneous Dirichlet on all edges,
        CSR, Jacobi-PCG FP64 to ||b-Ax||/||b|| <= 1e-10
    c2  3-D linear elasticity, Q1 hex 64^3, E = 1, nu = 0.3, body force (0,0,-1),
        clamped x = 0 face, 3x3-block BCSR, CG to 1e-8 (FP64 or FP32-inner/FP64-refined)
    c3  2-D Poisson with k = exp(g), g ~ N(0, 0.5^2) per element, Q4 256x256,
        5x5 Gauss, rhs ~ N(0,1) per node, homogeneous Dirichlet on all edges, CSR CG to 1e-8
    c5  2-D SUPG advection-diffusion, Q2 512x512, U = (cos 30, sin 30), cell Peclet 100,
        homogeneous Dirichlet on all edges, 64 N(0,1) right-hand sides, Jacobi-BiCGSTAB
"""
import math

import numpy as np

from . import hip

E_YOUNG, NU = 1.0, 0.3
LAMBDA = E_YOUNG * NU / ((1 + NU) * (1 - 2 * NU))
MU = E_YOUNG / (2 * (1 + NU))

SEED_C3_COEF, SEED_C3_RHS, SEED_C5_RHS = 1003, 2003, 5000


def supg_params(n):
    h = 1.0 / n
    U = (math.cos(math.radians(30)), math.sin(math.radians(30)), 0.0)
    Pe = 100.0
    kdiff = 1.0 * h / (2 * Pe)
    alpha = math.cosh(Pe) / math.sinh(Pe) - 1.0 / Pe
    return dict(U=U, kdiff=kdiff, h=h, alpha=alpha)


class Workload:
    def __init__(self, ctx, name, dim, p, dpn, n, fmt, ph, all_edges=True, clamp_x0=False,
                 rhs=None, precisions=hip.FP64 | hip.FP32):
        self.ctx, self.name = ctx, name
        self.mesh = hip.Mesh.structured(ctx, dim, p, dpn, n)
        self.A = self.mesh.pattern(fmt, precisions)
        self.n = self.mesh.ndofs
        self.ph = ph
        self.mask = ctx.alloc(self.n, np.uint8)
        self.mask.upload(np.zeros(self.n, np.uint8))
        if all_edges:
            for axis in range(dim):
                self.mesh.bc_mask_plane(self.mask, axis, 0.0)
                self.mesh.bc_mask_plane(self.mask, axis, 1.0)
        if clamp_x0:
            self.mesh.bc_mask_plane(self.mask, 0, 0.0)
        self.rhs_given = rhs
        self.b = ctx.alloc(self.n)
        self.x = ctx.alloc(self.n)
        self.req = hip.TANGENT | (hip.RESIDUAL if rhs is None else 0)
        self.extra_req = 0  # e.g. hip.FP32_ELEMENTS

    def assemble(self):
        """CGAssembly + DirichletBC::apply; leaves A and b ready to solve."""
        if self.rhs_given is None:
            self.b.zero()
        else:
            self.b.copy_from(self.rhs_given)
        self.mesh.assemble(self.ph, self.A, self.b if self.req & hip.RESIDUAL else None, self.req | self.extra_req)
        self.A.dirichlet(self.mask, self.b)

    def solve(self, **kw):
        self.x.zero()
        return self.A.solve(self.b, self.x, **kw)

    def step(self, **kw):
        self.assemble()
        return self.solve(**kw)


def c1(ctx, n=128):
    ph = hip.physics(hip.POISSON, 2, f=1.0)
    return Workload(ctx, "c1", 2, 1, 1, n, hip.CSR, ph)


def c2(ctx, n=64, fmt=hip.BCSR3):
    ph = hip.physics(hip.ELASTICITY, 3, lam=LAMBDA, mu=MU, body=(0.0, 0.0, -1.0))
    return Workload(ctx, "c2", 3, 1, 3, n, fmt, ph, all_edges=False, clamp_x0=True)


def c3(ctx, n=256, p=4):
    nel = n * n
    coef = np.exp(hip.normal(SEED_C3_COEF, nel, 0.0, 0.5))
    ph = hip.physics(hip.POISSON, 2, coef=coef, f=0.0)
    nn = (p * n + 1) ** 2
    rhs = ctx.upload(hip.normal(SEED_C3_RHS, nn))
    w = Workload(ctx, "c3", 2, p, 1, n, hip.CSR, ph, rhs=rhs)
    w.coef = coef
    return w


def c5(ctx, n=512, nrhs=64):
    prm = supg_params(n)
    ph = hip.physics(hip.SUPG, 2, **prm)
    w = Workload(ctx, "c5", 2, 2, 1, n, hip.CSR, ph, rhs=None)
    w.req = hip.TANGENT
    w.nrhs = nrhs
    B = np.stack([hip.normal(SEED_C5_RHS + j, w.n) for j in range(nrhs)], axis=1)
    w.B_host = np.ascontiguousarray(B)
    return w


# ---------------------------------------------------------------- serving
def c4_requests(n_req, seed=4000, lo2=16, hi2=256, lo3=8, hi3=48):
    """The served stream's requests: 2-D Poisson (Q1 or Q2, equal odds) or, one
    in four, 3-D Q1 elasticity; cells per axis log-uniform in [16,256] (2-D) or
    [8,48] (3-D); lognormal(0, 0.5) per-element coefficients; N(0,1) loads per
    dof; homogeneous Dirichlet on every face (Poisson) or a clamped x = 0 face
    (elasticity).  Deterministic in `seed`."""
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n_req):
        if rng.random() < 0.25:
            dim, p, dpn = 3, 1, 3
            n = int(round(math.exp(rng.uniform(math.log(lo3), math.log(hi3)))))
            nel, ndofs = n ** 3, (n + 1) ** 3 * 3
        else:
            dim, dpn = 2, 1
            p = 1 if rng.random() < 0.5 else 2
            n = int(round(math.exp(rng.uniform(math.log(lo2), math.log(hi2)))))
            nel, ndofs = n * n, (p * n + 1) ** 2
        reqs.append(dict(dim=dim, p=p, dpn=dpn, n=n, seed=seed * 7919 + i, nel=nel, ndofs=ndofs))
    return reqs


def c4_materialise(r):
    """Host arrays of one request (mesh, coefficients, load, physics)."""
    dim, p, n = r["dim"], r["p"], r["n"]
    nodes, cells = hip.structured_mesh(dim, n, n, n if dim == 3 else 0, p)
    coef = np.exp(hip.normal(r["seed"], r["nel"], 0.0, 0.5))
    load = hip.normal(r["seed"] + 1, r["ndofs"])
    if dim == 3:
        ph = hip.physics(hip.ELASTICITY, 3, coef=coef, lam=LAMBDA, mu=MU)
        bc = hip.BC_XMIN
    else:
        ph = hip.physics(hip.POISSON, 2, coef=coef)
        bc = hip.BC_BOX
    return dict(dim=dim, p=p, dpn=r["dpn"], nodes=nodes, cells=cells, ph=ph, load=load, bc=bc)


def poisson_arrivals(n, rate, seed):
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))
