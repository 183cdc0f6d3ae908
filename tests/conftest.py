import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libstaqdic.so)")


@pytest.fixture(scope="session")
def ctx():
    from staq_dic_amd import Context
    c = Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def plain_cfg():
    """512^2 affine pair on the uniform 900-node grid (seed 1001)."""
    from staq_dic_amd import synth
    return synth.config_plain()


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    pyoracle.build()
    return pyoracle


@pytest.fixture(scope="session")
def ref_frames():
    """The reference's own frames 000 / 001 of the circular-hole sequence, its
    1434-node quadtree mesh and its stored post-ADMM result for that pair
    (tests/golden/circular_hole_000_001.npz), prepared as the reference's
    driver prepares them: each image normalized over the ROI
    (func/funNormalizeImg.m:34-38, std with N-1) and multiplied by its mask
    (main_ALDIC_Quadtree_v2.m:57-64)."""
    import numpy as np
    z = np.load(os.path.join(ROOT, "tests", "golden", "circular_hole_000_001.npz"))
    gx, gy = z["gridx"], z["gridy"]

    def norm(img):
        t = img[gx[0] - 1:gx[1], gy[0] - 1:gy[1]].astype(np.float64)
        return (img.astype(np.float64) - t.mean()) / t.std(ddof=1)

    m0 = z["mask0"].astype(np.float64)
    m1 = z["mask1"].astype(np.float64)
    return dict(f=norm(z["img0"]) * m0, g=norm(z["img1"]) * m1, mask=m0,
                coords=z["coordinatesFEM"], winsize=int(z["winsize"]),
                winsize_List=z["winsize_List"].astype(np.int64), U1=z["U1"], F1=z["F1"],
                mesh=dict(coordinatesFEM=z["coordinatesFEM"], elementsFEM=z["elementsFEM"],
                          dirichlet=z["dirichlet"], markCoordHoleEdge=z["markCoordHoleEdge"]),
                winstepsize=float(z["winstepsize"]))
