"""The device form of the seeded generators (hpcsharp_amd.datagen's torch
functions, used by the full-size GPU tests) produces the numpy generators'
bits exactly; checked here on torch's CPU device."""
from __future__ import annotations

import numpy as np
import pytest

from hpcsharp_amd import datagen


@pytest.mark.parametrize("seed", [0, 1, 0x5EED0002, 2 ** 64 - 1, 0x8000000000000000])
def test_splitmix64_torch_matches_numpy(seed):
    a = datagen.splitmix64(seed, 70001)
    b = datagen.splitmix64_torch(seed, 70001, "cpu", chunk=30000).numpy().view(np.uint64)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("dist", ["uniform", "low8", "sorted", "constant"])
def test_keys_u32_torch_matches_numpy(dist):
    a = datagen.keys_u32(40009, 0x5EED0003, dist)
    b = datagen.keys_u32_torch(40009, 0x5EED0003, dist, "cpu").numpy().view(np.uint32)
    assert np.array_equal(a, b)


def test_splitmix64_first_outputs():
    # published reference outputs of splitmix64 seeded with 1234567
    x = datagen.splitmix64_torch(1234567, 3, "cpu").numpy().view(np.uint64)
    assert x.tolist() == [6457827717110365317, 3203168211198807973, 9817491932198370423]
