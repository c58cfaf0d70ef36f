"""tb_direct_sample_noalloc: the direct sum staged in the library's fixed
per-GPU array, with no device allocation, against tb_direct_sample on the same
handle (the same kernel, so bit for bit), on two calls in a row, the second
with a smaller target list; sizes beyond its maximum are refused."""
import numpy as np
import pytest

import tree_amd as ta
from tests.conftest import announce

pytestmark = pytest.mark.gpu


def test_direct_sample_noalloc_matches_allocating_entry():
    announce("direct_sample_noalloc")
    rng = np.random.default_rng(5)
    qel, qion = ta.gen_plasma(31, 5000, 4000, 100.0)
    with ta.Engine(rs=ta.kelbg_length(10.0), theta=0.5, walk="mixed", theta_mixed=0.4) as eng:
        eng.set_particles(qel, qion)
        eng.build()  # key order differs from the caller's: the mapping is exercised
        for nt in (3000, 700):
            tsp = rng.integers(0, 2, nt).astype(np.int32)
            tix = np.where(tsp == 0, rng.integers(0, 5000, nt), rng.integers(0, 4000, nt)).astype(np.int32)
            tix[:5] = tix[5]  # repeated targets
            tsp[:5] = tsp[5]
            ref = eng.direct_sample(tsp, tix)
            got = eng.direct_sample_noalloc(tsp, tix)
            assert np.array_equal(got, ref), nt
        with pytest.raises(ta.TreeError) as ex:
            n = ta.NOALLOC_MAX_TARGETS + 1
            eng.direct_sample_noalloc(np.zeros(n, np.int32), np.zeros(n, np.int32))
        assert ex.value.code == 1
        with pytest.raises(ta.TreeError):
            eng.direct_sample_noalloc(np.array([0], np.int32), np.array([5000], np.int32))
