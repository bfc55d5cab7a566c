"""The joint step against the sequential one.

With block-eliminated solves and thermal noise the library runs each step with the stresslet solves
of the predictor, the slip solve and both RFD probes in one launch and the probes' mob() on a side
stream (one_step_joint).  It performs the same operations as the sequential step (src/run.cpp:70-162
in the reference's order), only reordered, and every kernel is deterministic, so several steps must
give the same bits as the sequential order (PS_JOINT_STEP=0), launched kernel by kernel or replayed
as step graphs, in FP64 and in the mixed form.
"""
import os

import numpy as np
import pytest

from polystokes_brownian_amd import configs
from polystokes_brownian_amd.polystokes import Context

pytestmark = pytest.mark.gpu

CASES = {
    "c1": lambda: configs.workload("c1"),
    "box_700": lambda: configs.small(3, 35, 20, 14.0, seed=12),
    "dense_probes": lambda: configs.small(12, 20, 10, 9.0, seed=31),
}


def _run(pd, x0, steps, joint, graph):
    old = os.environ.get("PS_JOINT_STEP")
    if joint:
        os.environ.pop("PS_JOINT_STEP", None)
    else:
        os.environ["PS_JOINT_STEP"] = "0"
    try:
        g = Context(dict(pd, graph=graph))
    finally:
        if old is None:
            os.environ.pop("PS_JOINT_STEP", None)
        else:
            os.environ["PS_JOINT_STEP"] = old
    g.initial_configuration(x0)
    g.run(steps)
    out = (g.state()[0], g.vectors(), g.counters())
    g.close()
    return out


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name", list(CASES))
def test_joint_step_equals_sequential_step(name, precision):
    w = CASES[name]()
    pd = dict(configs.params_dict(w), solver=1, precision=precision, ir_maxit=0, ksp_rtol=1e-12,
              sqrt_rtol=1e-10)
    x0 = configs.positions(w)
    print("case", name, "precision", precision, "Np", pd["Np"], flush=True)
    ref = _run(pd, x0, 3, joint=False, graph=0)
    for graph in (0, 1):
        got = _run(pd, x0, 3, joint=True, graph=graph)
        assert np.array_equal(got[0], ref[0]), graph
        for k in ("up", "udiff", "v_brown", "sb"):
            assert np.array_equal(got[1][k], ref[1][k]), (graph, k)
        for k in ("applies", "mob_calls", "saddle_solves", "sqrt_calls", "lanczos_iterations",
                  "stresslet_iterations", "nvlist"):
            assert got[2][k] == ref[2][k], (graph, k)
    assert ref[2]["stresslet_iterations"] > 0
