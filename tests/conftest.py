import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "peridynamics.jl_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libpdx kernels)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(ROOT, "tests", "golden", "reference_values.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def oracle():
    import oracle as orc
    orc.build()
    return orc


def gpu_available():
    try:
        import torch  # noqa: F401  (only to count devices; the product path does not use it)
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def need_gpu():
    if not gpu_available():
        pytest.skip("no GPU visible")


def bb_params(E=70e9, delta=3.015e-3, rho=2700.0, Gc=100.0):
    """Point parameters of a bond-based material (material! derivation, host side)."""
    import peridynamics as pd
    kw = dict(horizon=delta, rho=rho, E=E)
    if Gc is not None:
        kw["Gc"] = Gc
    p = pd.materials.get_point_params(pd.BBMaterial(), kw)
    return p


def as_oracle_params(p, **extra):
    d = dict(delta=p.delta, rho=p.rho, E=p.E, nu=p.nu, G=p.G, K=p.K, bc=p.bc, epsc=p.epsc,
             Gc=p.Gc, mu=p.mu)
    d["lambda"] = p.lam
    d.update(extra)
    return d


class lds_budget:
    """PDX_LDS_BUDGET (bytes of LDS a block's neighbour table may take) for the contexts
    created inside the block; pdx_create reads it.  A small budget sends the kernels down
    their paths for tables longer than LDS.  0 keeps the default."""

    def __init__(self, nbytes):
        self.nbytes = nbytes

    def __enter__(self):
        self.old = os.environ.get("PDX_LDS_BUDGET")
        if self.nbytes:
            os.environ["PDX_LDS_BUDGET"] = str(self.nbytes)
        return self

    def __exit__(self, *exc):
        if self.old is None:
            os.environ.pop("PDX_LDS_BUDGET", None)
        else:
            os.environ["PDX_LDS_BUDGET"] = self.old


class bond_layout:
    """The device bond layout for the contexts created inside the block: "lds" (default),
    "lds12k" (the LDS-staged layout with a 12 KB table budget, so the tables are longer
    than what a block stages at once) or "global" (PDX_NO_LDS: the global-gather kernels
    that serve tables the 16-bit bond words cannot address)."""

    def __init__(self, layout):
        self.layout = layout

    def __enter__(self):
        self.old = {k: os.environ.get(k) for k in ("PDX_LDS_BUDGET", "PDX_NO_LDS")}
        if self.layout == "lds12k":
            os.environ["PDX_LDS_BUDGET"] = "12288"
        elif self.layout == "global":
            os.environ["PDX_NO_LDS"] = "1"
        return self

    def __exit__(self, *exc):
        for k, v in self.old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
