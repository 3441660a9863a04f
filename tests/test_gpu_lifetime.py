"""Handle lifetimes through the C ABI: a context released before the
structures and models made on it (the order a garbage collector may pick
for a reference cycle) keeps working until the last of them is destroyed,
and results on it equal those on another context."""
import gc

import numpy as np
import pytest

from helpers import random_cell

pytestmark = pytest.mark.gpu


def test_context_released_first():
    from flare_amd import _C_flare as C
    from flare_amd import _lib
    print("context released before its objects", flush=True)
    cell, species, pos = random_cell(20, 5, nos=2, scale=6.5, skew=0.2)
    calc = C.B2("chebyshev", "quadratic", [0.0, 4.0], [], [2, 4, 3])
    ref = C.Structure(cell, species, pos, 4.0, [calc])
    sgp = C.SparseGP([C.NormalizedDotProduct(1.5, 2.0)], 0.1, 0.1, 0.1)
    sgp.add_specific_environments(ref, [0, 3, 7, 11])
    sgp.update_L_inverse()
    sgp.set_alpha(np.linspace(-0.2, 0.3, sgp.n_sparse))
    sgp.predict_local_uncertainties(ref)
    ctx = _lib.Context(0)
    s = C.Structure(cell, species, pos, 4.0, [calc], context=ctx)
    sgp.predict_mean(s)  # makes the model on ctx
    ctx.close()  # the structure and the model keep it alive
    sgp.predict_local_uncertainties(s)
    np.testing.assert_array_equal(s.mean_efs, ref.mean_efs)
    np.testing.assert_array_equal(s.local_uncertainties[0], ref.local_uncertainties[0])
    del s
    sgp._invalidate()
    gc.collect()
    # A context reachable only from a structure <-> descriptor cycle.
    for _ in range(3):
        c2 = _lib.Context(0)
        t = C.Structure(cell, species, pos, 4.0, [calc], context=c2)
        sgp.predict_mean(t)
        del c2, t
        gc.collect()
    sgp.predict_mean(ref)
