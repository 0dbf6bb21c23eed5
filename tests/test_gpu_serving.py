"""Serving: continuous batching, admission under an HBM budget, and the
same result for a request alone and in a mix."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MIX = [dict(dim=2, p=1, dpn=1, n=20), dict(dim=2, p=2, dpn=1, n=16), dict(dim=3, p=1, dpn=3, n=8),
       dict(dim=2, p=1, dpn=1, n=40), dict(dim=2, p=2, dpn=1, n=24), dict(dim=3, p=1, dpn=3, n=6)]


def materialised(i, r):
    from yafel_amd import workloads
    r = dict(r)
    r["nel"] = r["n"] ** r["dim"]
    r["ndofs"] = (r["n"] + 1) ** 3 * 3 if r["dim"] == 3 else (r["p"] * r["n"] + 1) ** 2
    r["seed"] = 900 + i
    return workloads.c4_materialise(r)


def run_all(H, ctx, reqs, budget=8 << 30, slots=16, stagger=False, graphs=False, passes=None, limit=0):
    srv = H.Server(ctx, budget, slots)
    if graphs:
        srv.set_graphs(True)
    if passes is not None:
        srv.set_cg_passes(passes, limit)
    ids, out = [], {}
    for i, m in enumerate(reqs):
        ids.append(srv.submit(m["dim"], m["p"], m["dpn"], m["nodes"], m["cells"], m["ph"], load=m["load"],
                              bc=m["bc"], tol=1e-8))
        if stagger:
            srv.progress()       # later requests arrive while earlier ones run
    pending = set(ids)
    guard = 0
    while pending:
        for rid in srv.progress():
            pending.discard(rid)
        guard += 1
        assert guard < 200000
    for rid in ids:
        out[rid] = srv.result(rid)
    st = srv.stats()
    srv.close()
    return [out[r] for r in ids], st


def test_alone_and_in_the_mix_are_identical(ctx):
    from yafel_amd import hip as H
    reqs = [materialised(i, r) for i, r in enumerate(MIX)]
    mix, st = run_all(H, ctx, reqs, stagger=True)
    assert st["completed"] == len(MIX) and st["failed"] == 0
    for i, m in enumerate(reqs):
        alone, _ = run_all(H, ctx, [m])
        x_alone, info_alone = alone[0]
        x_mix, info_mix = mix[i]
        assert info_mix.status == 0 and info_mix.rel_residual < 1e-8
        assert np.array_equal(x_alone, x_mix)          # bit-identical alone vs in the mix
        assert info_alone.iterations == info_mix.iterations


def test_graph_replay_matches_direct_submission(ctx):
    """Chunks replayed from graphs whose kernels read the slot table and its
    totals from device memory (uploaded when a request joins or retires; a
    new capture only when the totals outgrow the captured grids) give the
    same bits as directly submitted launches, with up to four chunks in
    flight, in the three-pass form and with the form chosen per chunk."""
    from yafel_amd import hip as H
    reqs = [materialised(i, r) for i, r in enumerate(MIX)]
    direct, st0 = run_all(H, ctx, reqs, stagger=True)
    assert st0["graph_captures"] == 0 and st0["graph_replays"] == 0
    for passes, limit in ((3, 0), (0, 100_000)):
        replay, st1 = run_all(H, ctx, reqs, stagger=True, graphs=True, passes=passes, limit=limit)
        assert st1["graph_captures"] >= 1 and st1["graph_replays"] > 2 * st1["graph_captures"]
        assert st1["completed"] == len(reqs) and st1["failed"] == 0
        if passes == 0:
            assert st1["two_pass_chunks"] > 0 and st1["three_pass_chunks"] > 0
        for (x0, i0), (x1, i1) in zip(direct, replay):
            assert i1.status == 0 and i1.rel_residual < 1e-8
            assert np.array_equal(x0, x1)
            assert i0.iterations == i1.iterations


def test_two_and_three_pass_iterations_are_identical(ctx):
    """The batched CG iteration as two passes (the SpMV forms p = z + beta p
    in its gather and applies x += alpha p; a closing pass per chunk) gives
    the same bits as three, also when the form changes from chunk to chunk
    (a limit of 100 k stored scalars: the 3-D requests switch it as they join
    and retire)."""
    from yafel_amd import hip as H
    reqs = [materialised(i, r) for i, r in enumerate(MIX)]
    three, st3 = run_all(H, ctx, reqs, stagger=True, passes=3)
    assert st3["completed"] == len(reqs) and st3["failed"] == 0 and st3["two_pass_chunks"] == 0
    for passes, limit in ((2, 0), (0, 100_000)):
        other, st = run_all(H, ctx, reqs, stagger=True, passes=passes, limit=limit)
        assert st["completed"] == len(reqs) and st["failed"] == 0
        for (x2, i2), (x3, i3) in zip(other, three):
            assert i2.status == 0 and i2.rel_residual < 1e-8
            assert np.array_equal(x2, x3)
            assert i2.iterations == i3.iterations and i2.refinements == i3.refinements
        assert st["batched_launches"] < st3["batched_launches"]
        if passes == 2:
            assert st["three_pass_chunks"] == 0
        else:
            assert st["two_pass_chunks"] > 0 and st["three_pass_chunks"] > 0
    srv = H.Server(ctx, 1 << 30, 4)
    try:
        with pytest.raises(H.YafelHipError):
            srv.set_cg_passes(4)
    finally:
        srv.close()


def test_served_result_matches_direct_solve(ctx, orc):
    from yafel_amd import hip as H
    reqs = [materialised(i, r) for i, r in enumerate(MIX[:3])]
    served, _ = run_all(H, ctx, reqs)
    for m, (x, info) in zip(reqs, served):
        mesh = H.Mesh(ctx, m["dim"], m["p"], m["dpn"], m["nodes"], m["cells"])
        A = mesh.pattern(H.BCSR3 if m["dpn"] == 3 else H.CSR, H.FP64 | H.FP32)
        mesh.assemble(m["ph"], A, None, H.TANGENT)
        b = ctx.upload(m["load"])
        mask = ctx.upload(np.zeros(mesh.ndofs, np.uint8))
        if m["bc"] == H.BC_BOX:
            for d in range(m["dim"]):
                mesh.bc_mask_plane(mask, d, 0.0, 1e-9)
                mesh.bc_mask_plane(mask, d, 1.0, 1e-9)
        else:
            mesh.bc_mask_plane(mask, 0, 0.0, 1e-9)
        A.dirichlet(mask, b)
        bh = b.download()
        ptr, col, val = A.arrays()
        O = orc.wrap(3 if m["dpn"] == 3 else 1, mesh.ndofs // m["dpn"] if m["dpn"] == 3 else mesh.ndofs,
                     mesh.ndofs // m["dpn"] if m["dpn"] == 3 else mesh.ndofs, ptr, col, val)
        r = bh - O.spmv(x)
        assert np.linalg.norm(r) / np.linalg.norm(bh) < 1e-8
        x2, i2 = A.solve_host(bh, method=H.PCG_JACOBI, precision=H.MIXED, stop=H.STOP_RELRES, tol=1e-8)
        assert np.abs(x - x2).max() <= 1e-6 * np.abs(x2).max()


def test_admission_under_budget_and_rejection(ctx):
    from yafel_amd import hip as H
    reqs = [materialised(i, r) for i, r in enumerate(MIX)]
    # a budget that holds about two requests at a time
    sizes = []
    srv = H.Server(ctx, 1 << 40, 4)
    budget = 60 << 20
    srv.close()
    res, st = run_all(H, ctx, reqs, budget=budget, slots=2)
    assert st["completed"] == len(reqs) and st["hbm_peak"] <= budget
    for x, info in res:
        assert info.status == 0 and info.rel_residual < 1e-8
    srv = H.Server(ctx, 1 << 20, 4)  # 1 MiB: nothing fits
    m = reqs[0]
    with pytest.raises(H.YafelHipError) as e:
        srv.submit(m["dim"], m["p"], m["dpn"], m["nodes"], m["cells"], m["ph"], load=m["load"], bc=m["bc"])
    assert e.value.code == H.ENOMEM
    srv.close()


def test_cancel_queued_request(ctx):
    from yafel_amd import hip as H
    reqs = [materialised(i, r) for i, r in enumerate(MIX[:3])]
    srv = H.Server(ctx, 8 << 30, 1)
    ids = [srv.submit(m["dim"], m["p"], m["dpn"], m["nodes"], m["cells"], m["ph"], load=m["load"], bc=m["bc"])
           for m in reqs]
    srv.cancel(ids[2])
    pending = set(ids[:2])
    while pending:
        for rid in srv.progress():
            pending.discard(rid)
    assert srv.query(ids[2])[0] == H.REQ_CANCELLED
    assert srv.stats()["cancelled"] == 1
    assert srv.stats()["queued"] == 0
    # result reports the cancellation once, then the request is forgotten
    with pytest.raises(H.YafelHipError) as e:
        lib_x = np.zeros(1)
        H.check(H.lib.yafel_hip_server_result(srv.h, ids[2], H._ptr(lib_x), 1, None), ctx.h)
    assert e.value.code == H.ESTATE and "cancelled" in str(e.value)
    with pytest.raises(H.YafelHipError):
        srv.query(ids[2])
    srv.close()
