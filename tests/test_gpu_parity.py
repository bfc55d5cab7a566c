"""GPU parity: the HIP path (through the C ABI) against the CPU FP64 oracle on the same inputs.

Tolerances: Verlet pairs and the cell list bit-exact; FP64 mobility apply / saddle apply / RHS
1e-12 relative; saddle solves and the Lanczos square root 1e-10 relative (both sides solve to
rtol 1e-12); one full step (positions, velocities, RFD drift) 1e-10 relative.
"""
import numpy as np
import pytest

from polystokes_brownian_amd import configs
from polystokes_brownian_amd.polystokes import Context

pytestmark = pytest.mark.gpu


def _oracle():
    from oracle.pyoracle import Oracle
    return Oracle


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


def pair(w, **over):
    pd = configs.params_dict(w)
    pd.update(over)
    x0 = configs.positions(w)
    g = Context(pd)
    g.initial_configuration(x0)
    o = _oracle()(pd)
    o.initial_configuration(x0)
    return g, o, x0


CASES = {
    "c1": lambda: configs.workload("c1"),
    "small_chainHI": lambda: configs.small(3, 4, 10, 6.0, seed=3, chain_HI=1),
    "small_dense": lambda: configs.small(4, 6, 10, 4.5, seed=5),
    "nobox": lambda: dict(configs.small(2, 3, 8, 6.0, seed=7), L=None),
}


def _w(name):
    w = CASES[name]()
    if w.get("L") is None:
        # unbounded: generate inside a box, then switch PBC off
        w2 = dict(w, L=6.0)
        x0 = configs.positions(w2)
        w["_x0"] = x0
    return w


def build(name, **over):
    w = _w(name)
    pd = configs.params_dict(w)
    pd.update(over)
    x0 = w.get("_x0")
    if x0 is None:
        x0 = configs.positions(w)
    g = Context(pd)
    g.initial_configuration(x0)
    o = _oracle()(pd)
    o.initial_configuration(x0)
    return g, o, x0


@pytest.mark.parametrize("name", list(CASES))
def test_vlist_bit_exact(name):
    g, o, x0 = build(name)
    print("case", name, "Np", g.Np, flush=True)
    ng = g.check_dist()
    no = o.check_dist()
    rg, pg, ig = g.vlist()
    ro, po, io = o.vlist()
    assert ng == no
    assert np.array_equal(rg, ro) and np.array_equal(pg, po) and np.array_equal(ig, io)


def test_vlist_bit_exact_after_perturbation():
    # positions spread over many cells, including exact cell boundaries and wrapped images
    g, o, x0 = build("small_dense")
    rng = np.random.default_rng(11)
    for trial in range(3):
        x = x0 + rng.normal(scale=0.3, size=x0.shape)
        x[:6] = np.round(x[:6] * 4) / 4   # values on cell faces
        g.set_positions(x)
        o.set_positions(x)
        assert g.check_dist() == o.check_dist()
        for a, b in zip(g.vlist(), o.vlist()):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("name", list(CASES))
def test_mobility_apply(name):
    g, o, _ = build(name)
    print("case", name, flush=True)
    g.check_dist(); o.check_dist()
    g.mob(); o.mob()
    np.testing.assert_allclose(g.stale_tile(), o.stale_tile(), rtol=1e-13, atol=1e-14)
    rng = np.random.default_rng(1)
    for _ in range(2):
        f = rng.normal(size=g.n)
        ug, uo = g.mob_mult(f), o.mob_mult(f)
        assert rel(ug, uo) < 1e-12
        # bit-reproducible across calls (deterministic reductions)
        assert np.array_equal(ug, g.mob_mult(f))
    x = rng.normal(size=g.ns)
    assert rel(g.saddle_mult(x), o.saddle_mult(x)) < 1e-12
    np.testing.assert_array_equal(g.saddle_diag(), o.saddle_diag())


SYM_CASES = {
    # one partial tile (diagonal block only)
    "c1": lambda: configs.workload("c1"),
    # 3 tiles, the last one partial (off-diagonal blocks with and without ghost beads)
    "box_700": lambda: configs.small(3, 35, 20, 14.0, seed=12),
    # exactly 4 full tiles, no ghost beads
    "box_1024": lambda: configs.small(2, 64, 16, 16.0, seed=13),
    "nobox_700": lambda: dict(configs.small(3, 35, 20, 14.0, seed=14), L=None),
}


def _sym_build(name, **over):
    w = SYM_CASES[name]()
    if w.get("L") is None:
        x0 = configs.positions(dict(w, L=14.0))
    else:
        x0 = configs.positions(w)
    pd = configs.params_dict(w)
    pd.update(over)
    return pd, x0


@pytest.mark.parametrize("name", list(SYM_CASES))
def test_symmetric_aa_kernel(name):
    """The symmetric monomer-monomer kernel (one evaluation per unordered pair, applied both ways)
    against the oracle and against the ordered kernel; FP64 to 1e-12, FP32 tiles to 1e-5."""
    pd, x0 = _sym_build(name)
    o = _oracle()(pd)
    o.initial_configuration(x0)
    o.check_dist()
    o.mob()
    rng = np.random.default_rng(21)
    f = rng.normal(size=o.n)
    uo = o.mob_mult(f)
    us = {}
    for kern in (1, 2):
        for prec in (0, 1):
            g = Context(dict(pd, aa_kernel=kern, precision=prec))
            g.initial_configuration(x0)
            g.check_dist()
            g.mob()
            u = g.mob_mult(f)
            print("case", name, "kernel", kern, "precision", prec, "rel", rel(u, uo), flush=True)
            assert rel(u, uo) < (1e-12 if prec == 0 else 1e-5)
            assert np.array_equal(u, g.mob_mult(f))   # deterministic
            us[(kern, prec)] = u
            # chain_HI systems run the ordered kernel whatever is asked
            want = "symmetric" if kern == 2 and not pd.get("chain_HI") else "ordered"
            assert g.kernel_choice()["fp64"] == want
            g.close()
    assert rel(us[(2, 0)], us[(1, 0)]) < 1e-13
    # aa_kernel "auto": one of the two kernels, timed at create, and the same operator as that kernel
    for prec in (0, 1):
        g = Context(dict(pd, aa_kernel=0, precision=prec))
        kc = g.kernel_choice()
        print("case", name, "auto precision", prec, kc, flush=True)
        if pd.get("chain_HI"):   # only the ordered kernel applies: its source splits are timed
            assert kc["fp64"] == "ordered" and kc["ms64"] > 0
        else:
            assert kc["fp64"] in ("ordered", "symmetric") and kc["ms64"] > 0
            assert (kc["fp32"] in ("ordered", "symmetric") and kc["ms32"] > 0) if prec == 1 else kc["fp32"] is None
        # the tile shape: the ordered kernel's source split is a whole number of 256-bead tiles
        for key, sk in (("fp64", "split64"), ("fp32", "split32")):
            if kc[key] == "ordered":
                assert kc[sk] > 0 and kc[sk] % 256 == 0, kc
            else:
                assert kc[sk] == 0, kc
        g.initial_configuration(x0)
        g.check_dist()
        g.mob()
        u = g.mob_mult(f)
        k = {"ordered": 1, "symmetric": 2}[kc["fp64" if prec == 0 else "fp32"]]
        if prec == 0:
            assert rel(u, us[(k, 0)]) < 1e-13
        elif k == 2:
            assert np.array_equal(u, us[(k, 1)])
        else:   # the tuned source split may differ from the default one: the FP32 partial sums differ
            assert rel(u, us[(k, 1)]) < 1e-6
        # the second context of the same size class takes the cached choice
        g2 = Context(dict(pd, aa_kernel=0, precision=prec))
        assert g2.kernel_choice() == kc
        g2.close()
        g.close()


@pytest.mark.parametrize("name", ["c1", "small_chainHI"])
def test_rhs(name):
    g, o, _ = build(name)
    g.check_dist(); o.check_dist()
    rg, ro = g.rhs(), o.rhs()
    assert rel(rg, ro) < 1e-13
    np.testing.assert_allclose(rg, ro, rtol=1e-12, atol=1e-9 * max(1.0, np.abs(ro).max()))


@pytest.mark.parametrize("name,tol,same_its", [("c1", 1e-10, True), ("small_chainHI", 1e-10, True),
                                               ("small_dense", 1e-7, False)])
def test_saddle_solve(name, tol, same_its):
    """Both MINRES solutions against the direct dense solve of the saddle system.  The dense
    4.5^3 box case is ill-conditioned (thousands of iterations), so only its solution is pinned."""
    g, o, _ = build(name, ksp_rtol=1e-12)
    g.check_dist(); o.check_dist()
    g.mob(); o.mob()
    b = o.rhs()
    b[: o.n] += np.random.default_rng(2).normal(size=o.n)
    A = np.zeros((o.ns, o.ns))
    A[: o.n, : o.n] = o.mob_dense()
    A[: o.nu, o.n:] = -np.eye(o.nu)
    A[o.n:, : o.nu] = -np.eye(o.nu)
    xd = np.linalg.solve(A, b)
    xg, ig, _ = g.saddle_solve(b)
    xo, io, _ = o.saddle_solve(b)
    print("case", name, "its gpu", ig.iterations, "oracle", io.iterations,
          "err gpu", rel(xg, xd), "err oracle", rel(xo, xd), flush=True)
    assert ig.reason > 0 and io.reason > 0
    if same_its:
        assert abs(ig.iterations - io.iterations) <= 2
        assert rel(xg, xo) < tol
    assert rel(xg, xd) < tol and rel(xo, xd) < tol
    # warm start from the solution: converges at once (well-conditioned) to the same answer
    xw, iw, _ = g.saddle_solve(b, x0=xg)
    assert rel(xw, xd) < tol
    if same_its:
        assert iw.iterations <= 2


@pytest.mark.parametrize("name", ["c1", "small_dense"])
def test_sqrt(name):
    g, o, _ = build(name, sqrt_rtol=1e-10, sqrt_kmax=64)
    g.check_dist(); o.check_dist()
    g.mob(); o.mob()
    xi = np.random.default_rng(4).normal(size=o.n)
    wg, ig = g.sqrt_apply(xi)
    wo, io = o.sqrt_apply(xi)
    print("case", name, "lanczos gpu", ig.iterations, "oracle", io.iterations, flush=True)
    assert ig.iterations == io.iterations and ig.nneg == io.nneg
    assert rel(wg, wo) < 1e-10


@pytest.mark.parametrize("name", ["c1", "small_chainHI", "nobox"])
def test_one_step(name):
    g, o, _ = build(name, ksp_rtol=1e-12, sqrt_rtol=1e-10)
    g.run(1)
    o.run(1)
    xg, qg, tg = g.state()
    xo, qo, to = o.state()
    vg, vo = g.vectors(), o.vectors()
    print("case", name, "pos rel", rel(xg, xo), flush=True)
    assert tg == to
    assert rel(xg, xo) < 1e-12
    # displacement over the step, the quantity the tolerance is stated for
    x0 = build(name)[2]
    assert rel(xg - x0, xo - x0) < 1e-10
    for k in ("up", "v_det_n", "v_det_c", "v_brown", "fext"):
        assert rel(vg[k], vo[k]) < 1e-10, k
    # udiff and sb are differences of two probe solves 2*eps = 2e-3 apart (src/drift.cpp:57-58): each
    # solve is pinned at rtol 1e-12, so their difference is pinned only to ~1e-12 / eps
    assert rel(vg["udiff"], vo["udiff"]) < 1e-8
    assert rel(vg["sb"], vo["sb"]) < 1e-8
    assert rel(qg, qo) < 1e-12


def test_five_steps_c1():
    g, o, x0 = build("c1", ksp_rtol=1e-12, sqrt_rtol=1e-10)
    g.run(5)
    o.run(5)
    xg, _, _ = g.state()
    xo, _, _ = o.state()
    assert rel(xg - x0, xo - x0) < 1e-9
    assert g.check_dist() == o.check_dist()


@pytest.mark.parametrize("name", ["c1", "small_chainHI", "small_dense"])
def test_block_solver_matches_direct_solve(name):
    """PS_SOLVER_BLOCK solves the same saddle system exactly (two applies + the M_ss solve)."""
    g, o, _ = build(name, solver=1)
    g.check_dist(); o.check_dist()
    g.mob(); o.mob()
    b = o.rhs()
    b[: o.n] += np.random.default_rng(3).normal(size=o.n)
    A = np.zeros((o.ns, o.ns))
    A[: o.n, : o.n] = o.mob_dense()
    A[: o.nu, o.n:] = -np.eye(o.nu)
    A[o.n:, : o.nu] = -np.eye(o.nu)
    xd = np.linalg.solve(A, b)
    xg, info, _ = g.saddle_solve(b)
    print("case", name, "block M_ss its", info.iterations, "err", rel(xg, xd), flush=True)
    assert rel(xg, xd) < 1e-10
    assert rel(A @ xg, b) < 1e-11


def test_block_solver_step_matches_minres_step():
    g, o, x0 = build("c1", ksp_rtol=1e-12, sqrt_rtol=1e-10, solver=1)
    g.run(1)
    o.run(1)
    xg, _, _ = g.state()
    xo, _, _ = o.state()
    assert rel(xg - x0, xo - x0) < 1e-9
    vg, vo = g.vectors(), o.vectors()
    assert rel(vg["up"], vo["up"]) < 1e-9


def _direct(o, b):
    A = np.zeros((o.ns, o.ns))
    A[: o.n, : o.n] = o.mob_dense()
    A[: o.nu, o.n:] = -np.eye(o.nu)
    A[o.n:, : o.nu] = -np.eye(o.nu)
    return np.linalg.solve(A, b)


def _mixed_case(**over):
    w = configs.small(4, 10, 20, 8.0, seed=9)
    pd = configs.params_dict(w)
    x0 = configs.positions(w)
    g = Context(dict(pd, precision=1, **over))
    g.initial_configuration(x0)
    g.check_dist()
    g.mob()
    o = _oracle()(pd)
    o.initial_configuration(x0)
    o.check_dist()
    o.mob()
    b = o.rhs()
    b[: o.n] += np.random.default_rng(5).normal(size=o.n)
    return g, o, b, _direct(o, b)


def test_mixed_block_velocity_error():
    """Mixed form with block elimination: without refinement the solve's one apply runs on FP32 tiles
    and the velocity carries their rounding (below 1e-6 against the dense FP64 direct solve); with
    FP64 residual refinement it recovers FP64 accuracy.  The FP32 tiles also serve the square root."""
    g, o, b, xd = _mixed_case(solver=1, ir_maxit=0)
    x, info, _ = g.saddle_solve(b)
    err = rel(x[o.n:], xd[o.n:])
    print("mixed block velocity rel err, no refinement", err, flush=True)
    assert err < 1e-6
    g2 = _mixed_case(solver=1, ir_maxit=10, ir_rtol=1e-13)[0]
    x2, info2, _ = g2.saddle_solve(b)
    err2 = rel(x2[o.n:], xd[o.n:])
    print("mixed block velocity rel err, refined", err2, "sweeps", info2.refinements, flush=True)
    assert err2 < 1e-10 and info2.refinements >= 1
    xi = np.random.default_rng(6).normal(size=o.n)
    w32, i32 = g.sqrt_apply(xi)
    w64, i64 = o.sqrt_apply(xi)
    print("mixed sqrt rel err", rel(w32, w64), i32.iterations, i64.iterations, flush=True)
    assert rel(w32, w64) < 1e-5


def test_mixed_minres_velocity_error():
    """FP32-apply MINRES corrections + FP64 refinement against the dense FP64 direct solve."""
    g, o, b, xd = _mixed_case(solver=0, ksp_rtol=1e-6, ir_rtol=1e-9, ir_maxit=30)
    x, info, _ = g.saddle_solve(b)
    err = rel(x[o.n:], xd[o.n:])
    print("mixed minres velocity rel err", err, "sweeps", info.refinements, "its", info.iterations, flush=True)
    assert err < 1e-6


def _bf16_round(v):
    """Round float64 -> float32 -> bfloat16 (round to nearest even), returned as float64."""
    f = np.asarray(v, dtype=np.float32).copy()
    b = f.view(np.uint32).astype(np.uint64)
    b = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def _h_round(v, storage):
    if storage == "bf16":
        return _bf16_round(v)
    return np.asarray(v, dtype=np.float32).astype(np.float16).astype(np.float64)


@pytest.mark.parametrize("name", ["c1", "box_700"])
@pytest.mark.parametrize("storage", ["bf16", "f16"])
def test_dense_mfma_apply(name, storage):
    """Stored-dense M on MFMA: columns read back through the apply (exact for unit vectors) equal the
    oracle's FP64 entries rounded to the storage type (to one 16-bit ulp where the two FP64 values
    straddle a rounding boundary), and M X for 16 random vectors matches the FP64 product."""
    pd, x0 = _sym_build(name)
    o = _oracle()(pd)
    o.initial_configuration(x0)
    o.check_dist()
    o.mob()
    M = o.mob_dense()
    g = Context(pd)
    g.initial_configuration(x0)
    g.check_dist()
    g.mob()
    with pytest.raises(Exception):
        g.dense_mult(np.zeros((1, g.n)))          # before dense_build
    g.dense_build(storage)
    n = g.n
    # columns: the first monomer ones, every colloid force/torque/stresslet column of colloid 0, the last
    nu = o.nu
    cols =[0, 1, 2, 3, 3 * pd["Nm"], 3 * pd["Nm"] + 1, 3 * pd["Nm"] + 3 * pd["Nc"],
            3 * pd["Nm"] + 3 * pd["Nc"] + 2, nu, nu + 1, nu + 4, n - 1, n // 2, n // 3, 7, n - 5]
    E = np.zeros((16, n))
    for k, cidx in enumerate(cols):
        E[k, cidx] = 1.0
    got = g.dense_mult(E)
    want = _h_round(M[:, cols].T, storage)
    ulp = 2.0 ** (-7 if storage == "bf16" else -10)
    diff = np.abs(got - want)
    tol = ulp * np.abs(want) + 1e-7 * np.abs(M).max()
    print("case", name, storage, "exact fraction", np.mean(diff == 0), "max rel", (diff / (np.abs(want) + 1e-30)).max(),
          flush=True)
    assert np.all(diff <= tol)
    assert np.mean(diff == 0) > 0.995
    X = np.random.default_rng(31).normal(size=(16, n))
    Y = g.dense_mult(X)
    err = np.linalg.norm(Y - X @ M.T) / np.linalg.norm(X @ M.T)
    print("case", name, storage, "M X rel err", err, flush=True)
    assert err < (2e-2 if storage == "bf16" else 3e-3)
    # a smaller batch gives the same vectors
    Y3 = g.dense_mult(X[:3])
    np.testing.assert_array_equal(Y3, Y[:3])
    with pytest.raises(Exception):
        g.dense_mult(np.zeros((17, n)))
    g.close()


def test_frame_outputs_and_restart(tmp_path):
    """ps_run_output writes frames after loop indices 0, samplerate, 2 samplerate, ... with the
    positions, quaternions and stresslets of those steps (against the oracle stepped to the same
    times), and a restart from the last written frame (restart = its time: seed offset, trap kept at
    the origin, quaternions restored) continues like the oracle restarted from the same files."""
    from polystokes_brownian_amd.frames import format_time, read_frames, read_quaternions, read_trajectory
    w = configs.workload("c1")
    pd = configs.params_dict(w)
    pd.update(samplerate=2, ksp_rtol=1e-12, sqrt_rtol=1e-10)
    x0 = configs.positions(w)
    out = tmp_path / "frames"
    out.mkdir()
    g = Context(pd)
    g.initial_configuration(x0)
    g.run_output(5, str(out))
    o = _oracle()(pd)
    o.initial_configuration(x0)
    times, frames = read_trajectory(str(out))
    qt, quats = read_quaternions(str(out))
    st, stress = read_frames(str(out), "brownian_stresslet_")
    ft, forces = read_frames(str(out), "force_")
    assert len(times) == 3 and len(qt) == 3 and len(st) == 3 and len(ft) == 3
    names = set(p.name for p in out.iterdir())
    done = 0
    want = set()
    for k, steps in enumerate((1, 3, 5)):   # steps completed after loop indices 0, 2, 4
        o.run(steps - done)
        done = steps
        xo, qo, to = o.state()
        tt = format_time(to)
        want |= {f"{pre}{tt}_.txt" for pre in ("config_", "quats_", "force_", "brownian_stresslet_")}
        assert times[k] == float(tt)
        np.testing.assert_allclose(frames[k], xo.reshape(-1, 3), rtol=0, atol=1e-9)
        np.testing.assert_allclose(quats[k], qo.reshape(-1, 4), rtol=0, atol=1e-9)
        vo = o.vectors()
        np.testing.assert_allclose(stress[k].ravel(), vo["sb"], rtol=1e-7, atol=1e-8 * max(1.0, np.abs(vo["sb"]).max()))
        np.testing.assert_allclose(forces[k].ravel(), vo["fext"], rtol=1e-7, atol=1e-8 * max(1.0, np.abs(vo["fext"]).max()))
    assert names == want
    # restart from the last frame
    t_last = float(times[-1])
    pd2 = dict(pd, restart=t_last)
    xl, ql = frames[-1].ravel(), quats[-1].ravel()
    g2 = Context(pd2)
    g2.initial_configuration(xl, ql)
    g2.run(2)
    o2 = _oracle()(pd2)
    o2.initial_configuration(xl, ql)
    o2.run(2)
    xg, qg, tg = g2.state()
    xo, qo, to = o2.state()
    assert tg == to
    assert rel(xg - xl, xo - xl) < 1e-9
    assert rel(qg, qo) < 1e-12


@pytest.mark.parametrize("name,over", [("c1", dict()), ("box_1024", dict(aa_kernel=2, precision=1, ir_maxit=0)),
                                       ("box_1024", dict(aa_kernel=1))])
def test_mmhi_step_graph_replays_the_eager_step(name, over):
    """Full monomer HI with block-eliminated solves: ps_run replays each step as a HIP graph (the
    stresslet MINRES and the Lanczos stopping test run on the device).  Four replayed steps give the
    same bits as four steps launched kernel by kernel, and the same device-side counts."""
    if name in SYM_CASES:
        pd, x0 = _sym_build(name)
    else:
        w = configs.workload(name)
        pd, x0 = configs.params_dict(w), configs.positions(w)
    pd = dict(pd, solver=1, ksp_rtol=1e-12, sqrt_rtol=1e-10, **over)
    out = {}
    for graph in (1, 0):
        g = Context(dict(pd, graph=graph))
        g.initial_configuration(x0)
        g.run(4)
        out[graph] = (g.state()[0], g.vectors(), g.counters())
        g.close()
    (x1, v1, c1), (x0_, v0, c0) = out[1], out[0]
    print("case", name, over, "counters graph", c1, flush=True)
    assert np.array_equal(x1, x0_)
    for k in ("up", "udiff", "v_brown", "sb"):
        assert np.array_equal(v1[k], v0[k]), k
    for k in ("applies", "mob_calls", "saddle_solves", "sqrt_calls", "lanczos_iterations",
              "stresslet_iterations", "lanczos_nneg_count", "nvlist"):
        assert c1[k] == c0[k], k
    assert c1["stresslet_iterations"] > 0 and c1["lanczos_iterations"] > 0


def test_stresslet_solve_reports_iterations_and_counts():
    """The persistent stresslet MINRES: the solve information read back equals what the device
    counters accumulated, and the solution still matches the direct solve."""
    g, o, _ = build("small_dense", solver=1)
    g.check_dist(); o.check_dist()
    g.mob(); o.mob()
    b = o.rhs()
    b[: o.n] += np.random.default_rng(8).normal(size=o.n)
    c0 = g.counters()["stresslet_iterations"]
    x, info, _ = g.saddle_solve(b)
    c1 = g.counters()["stresslet_iterations"]
    assert info.iterations > 0 and c1 - c0 == info.iterations
    assert rel(x, _direct(o, b)) < 1e-10
