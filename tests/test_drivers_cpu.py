"""Host-side logic of the batched-CG and learned-preconditioner drivers (no GPU)."""
import numpy as np

from tbsla_amd import batched, learned


def test_batched_instance_sizes():
    n, c = batched.instance_sizes(512, seed=0)
    assert len(n) == len(c) == 512
    assert n.min() >= 2 ** 12 and n.max() < 2 ** 18
    assert c.min() >= 8 and c.max() <= 32
    # log-uniform: about half the systems below 2^15
    assert 0.4 < np.mean(n < 2 ** 15) < 0.6
    n2, c2 = batched.instance_sizes(512, seed=0)
    assert np.array_equal(n, n2) and np.array_equal(c, c2)
    # the whole batch stays within int32 indexing after symmetrisation
    assert 2 * int((n * c).sum()) + int(n.sum()) < 2 ** 31 - 1


def test_learned_summary():
    rows = [{"method": "jacobi_fp64", "ms": 2.0, "iterations": 100, "true_rel_res": 9e-9,
             "fallback": False, "seed": 1},
            {"method": "jacobi_fp64", "ms": 4.0, "iterations": 120, "true_rel_res": 8e-9,
             "fallback": False, "seed": 2},
            {"method": "learned_mixed", "ms": 1.0, "iterations": 50, "true_rel_res": 9.9e-9,
             "fallback": True, "seed": 1}]
    s = learned.summarise(rows)
    assert s["jacobi_fp64"]["ms_median"] == 3.0
    assert s["jacobi_fp64"]["iterations_median"] == 110
    assert s["jacobi_fp64"]["max_true_rel_res"] == 9e-9
    assert s["learned_mixed"]["fallbacks"] == 1
    assert "plain_fp64" not in s


def test_held_out_seeds_disjoint_from_training():
    # training uses problem seeds 0..1999, evaluation starts at 100000
    assert learned.RHS_SEED_OFFSET > 2000
    assert 100_000 > 2000
