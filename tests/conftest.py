"""Shared test set-up: import paths, the `gpu` marker, oracle fixtures."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "fast-matmul_amd"), os.path.join(ROOT, "oracle"),
          os.path.join(ROOT, "tests", "golden"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oracle_mod():
    """The CPU oracle with the image's MKL as its single-threaded leaf BLAS
    (the reference's own leaf library; threaded MKL 2021.4 dgemm returns
    wrong results for m = 1 products in this image)."""
    import oracle
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        pytest.skip("oracle not built (make -C oracle)")
    oracle.use_mkl(1)
    return oracle


@pytest.fixture(scope="session")
def golden():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "golden.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def fm():
    """The product package on a GPU (skips the test where no GPU is visible)."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import fastmm
    return fastmm
