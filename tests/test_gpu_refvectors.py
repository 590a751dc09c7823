"""The HIP path against vectors the reference itself printed
(tests/golden/ref_vectors.json; see tests/test_oracle_refvectors.py): every
batch's global delta and every relation's final full version, intern ids
included, bit-exact for each program, rank count (= virtual buckets) and batch
size, through the C ABI -- once with the dense membership of dictionary-ranked
outputs (the default) and once with the general hash table for every output
(PRA_DENSE_MEMBERSHIP=0, read when a context is created)."""
import os

import pytest

from test_oracle_refvectors import VECTORS, check_against_reference, run_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dense", ["1", "0"])
@pytest.mark.parametrize("key", sorted(VECTORS))
def test_gpu_matches_reference_vectors(key, dense, monkeypatch):
    import pra_amd
    monkeypatch.setenv("PRA_DENSE_MEMBERSHIP", dense)
    k = int(key.split("/")[1][1:])
    print("case", key, "dense", dense, flush=True)
    g = pra_amd.Context(nbuckets=k, device=0)
    try:
        hb, bs = run_case(g, key)
        g.set_batch_size(bs)
        g.execute()
        check_against_reference(g, (hb.rels, hb.sccs), VECTORS[key], bs)
    finally:
        g.close()
