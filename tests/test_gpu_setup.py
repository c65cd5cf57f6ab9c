"""create_cache's derived fields against the oracle's restatement, exactly:
the pipe index lists in the reference's order (create_advection_index_lists,
src/cache.jl:9-44), countz and the list lengths, and the eigenvalue estimate
(eigen_estimator_pre_calculation, src/solver.jl:474-488).  The library's
context set-up runs on the GPU box only, hence the gpu mark."""
import functools

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


def W():
    from geothermalwells_jl_amd import workloads
    return workloads


@functools.lru_cache(maxsize=None)
def _c2():
    return W().c2()


CASES = {
    "single_borehole": lambda: W().c0(),
    "array_3x3": lambda: W().c1(),
    "layered_adjoint": lambda: W().c3(),
    "array_5x5": lambda: W().c4(),
    "two_bh_depths": lambda: H.small_case(nbh=2, inlet="perwell", hs=(60.0, 41.0)),
    "two_bh_touching": lambda: H.small_case(nbh=2, inlet="hx", xs=[(0.0, 0.0), (0.2, 0.0)], hs=(60.0, 50.0)),
    "golden": lambda: W().golden(),
}
for _s in (0, 17, 40, 63):
    CASES[f"ragged_{_s}"] = (lambda s=_s: _c2()[s])


@pytest.mark.parametrize("case", list(CASES))
def test_index_lists_counts_and_eigen_estimate(case):
    g = H.gw()
    w = CASES[case]()
    cache = g.create_cache(**W().scenario_kwargs(w))
    orc = H.oracle_for(w)
    assert (cache.countz, cache.countxy_inner, cache.countxy_outer) == (orc.countz, orc.n_inner, orc.n_outer)
    assert cache.eigen_estimate == orc.eigen_estimate
    inner, outer = cache.index_lists(0)
    o_in, o_out = orc.index_lists()
    assert np.array_equal(np.array(inner, dtype=np.int32).reshape(-1, 3) - 1, o_in)
    assert np.array_equal(np.array(outer, dtype=np.int32).reshape(-1, 3) - 1, o_out)
    assert len(inner) > 0 and len(outer) > 0
