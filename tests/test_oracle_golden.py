"""The oracle's random numbers against known-answer vectors worked out from
the reference's definitions (dense.rs:273-279, 374, 446-448) with Python
integers (tests/golden/make_rng_kat.py, committed as tests/golden/rng_kat.json)."""
import json
import os

import numpy as np

import oracle as orc

KAT = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                  "rng_kat.json")))


def test_initial_seeds():
    for i, want in KAT["seed_init"]:
        assert format(orc.seed_init(i), "016x") == want, i


def test_refinement_seeds():
    for i, it, parity, want in KAT["seed_refine"]:
        assert format(orc.seed_refine(i, it, parity), "016x") == want, (i, it, parity)


def test_draws():
    for case in KAT["draws"]:
        seed = int(case["seed"], 16)
        draws, _ = orc.rnd_stream(seed, len(case["draws"]))
        for k, (state, value) in enumerate(case["draws"]):
            assert np.float32(value) == draws[k], (case["seed"], k)
            _, s_k = orc.rnd_stream(seed, k + 1)
            assert format(s_k, "016x") == state, (case["seed"], k)
