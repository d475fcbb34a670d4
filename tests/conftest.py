import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (os.path.join(ROOT, "oracle"), os.path.join(ROOT, "calypso-gap_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def frames():
    return np.load(os.path.join(GOLDEN, "c64_diamond", "frames.npz"))


@pytest.fixture(scope="session")
def example_model():
    import oracle
    return oracle.parse_model(os.path.join(GOLDEN, "c64_diamond", "gap_parameters"))


@pytest.fixture(scope="session")
def ref_vectors():
    return np.load(os.path.join(GOLDEN, "ref_vectors.npz"))


def golden_model(ref, name):
    import oracle
    g = lambda k: ref["model_%s_%s" % (name, k)]  # noqa: E731
    z = g("sf_zeta")
    return oracle.Model(z=g("z"), w=g("w"), sf_type=g("sf_type"), sf_alpha=g("sf_alpha"),
                        sf_cut=g("sf_cut"), theta=g("theta"), mm=g("mm"), coeff=g("coeff"),
                        sf_zeta=None if np.all(z == 1) else z)


def golden_cases(ref):
    out = []
    for cname, mname in zip(ref["case_names"], ref["case_models"]):
        p = "case_%s_" % cname
        out.append(dict(name=str(cname), model=str(mname), species=ref[p + "species"], lat=ref[p + "lat"],
                        pos=ref[p + "pos"], rcut=float(ref[p + "rcut"]), lgrad=bool(ref[p + "lgrad"]),
                        energy=float(ref[p + "energy"]), forces=ref[p + "forces"], stress=ref[p + "stress"]))
    return out


GOLDEN_CASE_NAMES = ["b8c56", "bc24_triclinic", "bc6_small_cell", "bc6_small_cell_nograd", "t16", "t40",
                     "t40_nograd", "t8_small_cell", "c64_single_species", "bc8_dense_920", "t8_dense_920"]


CAR_CASE_NAMES = ["bc6_small_cell", "t8_small_cell", "bc8_dense_920"]


def car_case(ref, name):
    """A CAR2ACSF golden case: FGAP_CALC's neighbour table of a structure and
    the reference CAR2ACSF's xx, dxdy, strs on it (Fortran layouts)."""
    p = "car_%s_" % name
    return dict(name=name, model=str(ref[p + "model"]), pos=ref[p + "pos"], neighbor=ref[p + "neighbor"],
                count=ref[p + "count"], xx=ref[p + "xx"], dxdy=ref[p + "dxdy"], strs=ref[p + "strs"])
