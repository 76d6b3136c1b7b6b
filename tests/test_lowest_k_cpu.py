"""The long-vector selection behind the pspace choice (direct_spin1.lowest_k)
against torch.topk on the CPU: the same values in the same order, the same
indices where the values are distinct, including vectors with a tail past
the last whole chunk and minima concentrated in one chunk."""

import numpy
import pytest
import torch

from conftest import package

pkg = package()


@pytest.mark.parametrize("n,k,seed", [(3_000_000, 400, 0), (2_000_003, 64, 1), (70_000, 16, 2),
                                      (1_000_000, 1, 3), (500_000, 400, 4)])
def test_lowest_k_matches_topk(n, k, seed):
    rng = numpy.random.default_rng(seed)
    x = rng.normal(size=n)
    if seed == 1:
        x[n - 5:] -= 10.0           # the lowest values in the tail
    if seed == 4:
        x[16384 * 7:16384 * 7 + 1000] -= 10.0   # all in one chunk
    t = torch.from_numpy(x)
    want = torch.topk(t, k, largest=False).indices
    got = pkg.direct_spin1.lowest_k(t, k)
    numpy.testing.assert_array_equal(x[got.numpy()], x[want.numpy()])
    numpy.testing.assert_array_equal(got.numpy(), want.numpy())


def test_lowest_k_ties_give_a_lowest_set():
    x = numpy.zeros(1_000_000)
    x[::7] = -1.0
    t = torch.from_numpy(x)
    got = pkg.direct_spin1.lowest_k(t, 300).numpy()
    assert len(set(got.tolist())) == 300 and (x[got] == -1.0).all()
