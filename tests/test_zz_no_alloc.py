"""komd_rdf_noalloc (the RDF sample without any device allocation: molecule centres and
counts in a static device array per GPU) against komd_rdf through the C ABI.  Integer
counts: bit-exact.  Two calls in a row share the GPU's array, the second with fewer
molecules and fewer bins, so stale centres or counts of the first would show."""
import numpy as np
import pytest

import komd_amd as K
from komd_amd import systems as S

pytestmark = pytest.mark.gpu


def test_rdf_noalloc_matches_rdf_on_two_calls_in_a_row():
    first = S.chains((8, 8, 8), 4, 0.7, seed=501)   # 2,048 four-site molecules
    second = S.fcc(6, 0.8, 0.8, 502)                # 864 single-site molecules
    for s, max_r, dr in ((first, 4.0, 0.05), (second, 3.0, 0.1)):
        print(f"rdf_noalloc n={s.n} max_r={max_r} dr={dr}", flush=True)
        with K.Context(s) as c:
            got = c.rdf_noalloc(max_r, dr)
            want = c.rdf(max_r, dr)
            again = c.rdf_noalloc(max_r, dr)
        assert got.sum() > 0
        assert np.array_equal(got, want)
        assert np.array_equal(again, want)


def test_rdf_noalloc_rejects_more_molecules_than_its_scratch():
    s = S.fcc(41, 0.8, 0.8, 503)                    # 275,684 > 262,144 molecules
    print(f"rdf_noalloc over capacity n={s.n}", flush=True)
    with K.Context(s) as c:
        with pytest.raises(K.KomdError, match="capacity of the static scratch") as ei:
            c.rdf_noalloc(3.0, 0.1)
        assert ei.value.status == 1
        c.rdf(3.0, 0.1)                             # the allocating entry point still serves it
