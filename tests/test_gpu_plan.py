"""The factorization's launch plan (iegpu_set_launch_plan): after a full
factorization the ID-stage kernels whose lists were empty are not launched
again for the same inputs; k_root checks every skipped list and a miss
repeats the factorization with every kernel.  The result must not depend
on the plan: every case below is compared bit for bit with a context that
launches every kernel."""
import numpy as np
import pytest
import torch

import geometry as G
import iesolver

pytestmark = pytest.mark.gpu

MODEL = iesolver.__file__.replace("iesolver.py", "rank_model_default.npy")


def _solver(g, form, plan):
    s = iesolver.RSSolver(0)
    s.set_geometry(g)
    s.build_tree(64)
    if form == "learned":
        s.set_rank_model(np.load(MODEL), margin=3)
        s.set_id_method(iesolver.ID_RANDOMIZED, 0)
    elif form == "randomized":
        s.set_id_method(iesolver.ID_RANDOMIZED, 32)
    s.set_launch_plan(plan)
    return s


@pytest.mark.parametrize("form", ["learned", "randomized", "plain"])
def test_plan_result_is_bitwise_the_full_launch_result(form):
    g = G.wavy(65536, seed=4)
    ys, q = G.exterior_sources(4)
    f = G.harmonic_data(g["points"], ys, q)
    ref = _solver(g, form, 0)
    ref.factor(1e-10)
    m_ref, _ = ref.solve(f)
    assert ref.launch_plan_info()[0] == 0

    s = _solver(g, form, 1)
    s.factor(1e-10)
    used, misses, skipped = s.launch_plan_info()
    assert used == 0 and misses == 0 and skipped > 0  # recorded, not used yet
    m1, _ = s.solve(f)
    s.factor(1e-10)
    used, misses, _ = s.launch_plan_info()
    print(f"{form}: {skipped} kernel groups skipped, used {used}", flush=True)
    assert used == 1 and misses == 0
    m2, _ = s.solve(f)
    assert np.array_equal(m1, m_ref) and np.array_equal(m2, m_ref)

    # a tolerance change is a new key: full launches, then the new plan
    s.factor(1e-6)
    assert s.launch_plan_info()[0] == 0
    s.factor(1e-6)
    assert s.launch_plan_info()[0] == 1
    ref.factor(1e-6)
    assert np.array_equal(s.solve(f)[0], ref.solve(f)[0])


def test_plan_miss_repeats_in_full():
    g = G.star(32768)
    ys, q = G.exterior_sources(1)
    f = G.harmonic_data(g["points"], ys, q)
    ref = _solver(g, "plain", 0)
    ref.factor(1e-10)
    m_ref, _ = ref.solve(f)
    s = _solver(g, "plain", 1)
    s.set_launch_plan(2)  # every list claimed empty: k_root must catch it
    s.factor(1e-10)
    used, misses, skipped = s.launch_plan_info()
    assert misses == 1 and used == 0 and skipped > 0
    assert np.array_equal(s.solve(f)[0], m_ref)
    s.factor(1e-10)  # the plan recorded by the repeat
    assert s.launch_plan_info()[:2] == (1, 1)
    assert np.array_equal(s.solve(f)[0], m_ref)


def test_plan_with_fused_factor_solve_and_updates():
    g = G.wavy(32768, seed=5)
    ys, q = G.exterior_sources(5)
    f = G.harmonic_data(g["points"], ys, q)
    dev = torch.device("cuda:0")
    fd = torch.from_numpy(f).to(dev)
    ref = _solver(g, "learned", 0)
    x0 = torch.empty_like(fd)
    ref.factor_solve_device(1e-10, iesolver.FP64, fd.data_ptr(), x0.data_ptr())
    s = _solver(g, "learned", 1)
    for k in range(3):
        x = torch.empty_like(fd)
        s.factor_solve_device(1e-10, iesolver.FP64, fd.data_ptr(), x.data_ptr())
        torch.cuda.synchronize()
        assert s.launch_plan_info()[0] == (1 if k > 0 else 0)
        assert torch.equal(x, x0)
    # a geometry update invalidates the plan (and the incremental run takes none)
    g2 = G.bump_wavy(32768, 0.01, 1.0, seed=5)
    s.update_geometry(g2)
    s.factor(1e-10)
    assert s.launch_plan_info()[0] == 0
