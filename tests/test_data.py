"""The package's device-side data generators against the oracle's restatement
(CPU tensors here; the same torch code runs on the GPU)."""
import numpy as np
import pytest
import torch

from oracle import sbp_oracle as O
from sc_2022_amd import data as D


@pytest.mark.parametrize("seed", [0, 7, 2022, 2**63 + 5])
def test_fourier_modes_match_oracle(seed):
    assert D.fourier_modes(seed) == O.fourier_modes(seed)


def test_fourier_modes_ranges():
    src, bnd = D.fourier_modes(11)
    assert len(src) == 64 and len(bnd) == 4 and all(len(m) == 8 for m in bnd)
    for kx, ky, a, ph in src:
        assert 1 <= kx <= 32 and 1 <= ky <= 32 and 0 <= ph < 2 * np.pi
        assert a == pytest.approx(1.0 / (kx * kx + ky * ky + 1.0), rel=0, abs=0)


@pytest.mark.parametrize("level", [3, 5])
def test_fourier_fields_match_oracle(level):
    N = 2 ** level + 1
    got = D.fourier_fields(N, 7, device="cpu")
    ref = O.fourier_data(level, 7)
    for g, r in zip(got, ref):
        np.testing.assert_allclose(g.numpy(), r, rtol=0, atol=1e-13)


def test_manufactured_fields_match_oracle():
    level = 4
    got = D.manufactured_fields(2 ** level + 1, device="cpu")
    ref = O.manufactured_data(level)
    for g, r in zip(got, ref):
        np.testing.assert_allclose(g.numpy(), r, rtol=0, atol=1e-13)


def test_reference_fields_match_oracle():
    """u = sin(pi (x + y)) and its data (level_2_multigrid_new.jl:36-38,
    171-177) against the oracle's reference_data."""
    level = 4
    got = D.reference_fields(2 ** level + 1, device="cpu")
    ref = O.reference_data(level)
    for g, r in zip(got, ref):
        np.testing.assert_allclose(g.numpy(), r, rtol=0, atol=1e-13)


def test_batch_specs():
    s1 = D.batch_specs(2022)
    s2 = D.batch_specs(2022)
    assert s1 == s2 and len(s1) == 256
    Ns = {N for N, _, _ in s1}
    assert Ns == {2 ** k + 1 for k in range(6, 11)}
    assert {p for _, p, _ in s1} == {2, 4}
    # roughly uniform: every size class holds a fair share of 256
    for N in Ns:
        c = sum(1 for n, _, _ in s1 if n == N)
        assert 25 <= c <= 80, (N, c)
    assert D.batch_specs(2023) != s1
