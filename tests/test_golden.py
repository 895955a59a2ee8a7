"""The oracle and the scene generator against the committed golden fixtures
(tests/golden/*.npz, written by tests/golden/make_golden.py): regenerating the
inputs reproduces the stored frames and poses bit for bit, and the oracle's
outputs on the stored inputs reproduce the stored block keys, visible counts,
voxel state (SHA-256 plus raw samples), vertex and normal maps and mesh."""
import hashlib
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden as G  # noqa: E402

CASES = sorted(G.CASES)


def load(name):
    with np.load(os.path.join(HERE, "golden", name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", CASES)
def test_generator_reproduces_inputs(name):
    gold = load(name)
    inp = G.case_inputs(G.CASES[name])
    assert set(inp) <= set(gold)
    for k, v in inp.items():
        assert np.array_equal(v, gold[k]), f"{name}: input {k} differs from the fixture"


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_outputs(name, oracle_threads):
    gold = load(name)
    out = G.oracle_outputs(gold)
    for k, v in out.items():
        assert np.array_equal(v, gold[k], equal_nan=v.dtype.kind == "f"), f"{name}: {k} differs from the fixture"


@pytest.mark.parametrize("name", CASES)
def test_fixture_self_consistency(name):
    """Stored outputs agree with each other: the sampled blocks are rows of
    the key list, the mesh head is the start of the hashed soup's order, the
    counts are positive, the raycast hits most pixels."""
    gold = load(name)
    assert len(gold["keys"]) == len(np.unique(gold["keys"], axis=0)) > 1000
    assert gold["sample_sdf"].shape == (8, 512) and gold["sample_rows"].max() < len(gold["keys"])
    assert np.all(gold["counts"] > 0)
    hit = ~np.isnan(gold["vmap"][..., 0])
    assert hit.mean() > 0.9
    head = gold["mesh_head"]
    assert np.all(np.lexsort(head.T[::-1]) == np.arange(len(head)))
    assert len(hashlib.sha256(b"").hexdigest()) == len(str(gold["mesh_sha"]))
