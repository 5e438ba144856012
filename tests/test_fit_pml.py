"""FIT perfectly matched layers (D:4414-4556 grid and solid extension,
D:5519-5557 / D:5609-5647 coefficients): the grid grows by
floor(width scale / L) cells on each +x, +y, +z side with a pml condition,
the solids grow into the layer, and every edge (face) whose barycentre lies
beyond the inner box on axis a gets
    upml (fpml) = prod_a exp(-((c_a - max_a) / (true_max_a - max_a))^3 * 4/2/150/pi/L_a),
1 elsewhere."""
import numpy as np
import pytest

from tetfit_amd import Model, configs
from tetfit_amd.host import TetfitError


def pml_case(tmp_path, pml, n=10, step=0.1, name="pml"):
    p = configs.fit_cavity(str(tmp_path), n=n, step=step, n_steps=20, n_probes=3, name=name, pml=pml)
    return p, Model(p).build(1)


def test_pml_grid_and_coefficients(tmp_path):
    n, step = 10, 0.1
    pml = {"+x": 0.35, "+z": 0.25}
    _, b = pml_case(tmp_path, pml, n, step)
    f = b.fit_ops()
    nx, ny, nz = f["cells"]
    assert (nx, ny, nz) == (n + int(0.35 / step), n, n + int(0.25 / step))
    t = b.fit_tables()
    pts, g = t["pts"], t["g_nodes"]
    eb = 0.5 * (pts[g[:, 0]] + pts[g[:, 1]])
    inner = n * step + 0.25 * step  # the grid's upper bound (xgrid {0, step, inner}), not a cell face
    true_max = np.array([inner + int(0.35 / step) * step, inner, inner + int(0.25 / step) * step])
    ref = np.ones(b.n_edges, dtype=np.longdouble)
    for a in (0, 2):
        d = np.abs(inner - eb[:, a]) / abs(inner - true_max[a])
        fac = np.exp(-(d.astype(np.longdouble) ** 3) * 4 / 2 / 150 / np.longdouble(np.pi) / step)
        ref = np.where(eb[:, a] > inner, ref * fac, ref)
    np.testing.assert_allclose(f["upml"], ref.astype(float), rtol=1e-15, atol=0)
    # 1 inside the inner box, below 1 and decreasing outwards in the layer
    inside = (eb[:, 0] <= inner) & (eb[:, 2] <= inner)
    assert np.all(f["upml"][inside] == 1) and np.all(f["upml"][~inside] < 1)
    xs = eb[:, 0] > inner
    order = np.argsort(eb[xs, 0])
    assert np.all(np.diff(f["upml"][xs][order][(eb[xs, 2][order] <= inner)]) <= 0)
    assert np.all(f["fpml"] <= 1) and np.any(f["fpml"] < 1)


def test_pml_layer_takes_the_extended_solid_material(tmp_path):
    """The box solid grows into the layer (Solid::ConstructPMLExtension), so
    the layer's edges carry the box's eps (M_q = eps A / L as inside)."""
    p = configs.fit_cavity(str(tmp_path), n=8, step=0.1, n_steps=10, n_probes=2, name="pmlm", eps_r=3.0,
                           pml={"+y": 0.2})
    b = Model(p).build(1)
    f = b.fit_ops()
    common = f["common_edges"]
    mq = f["M_q"][common]
    assert np.allclose(mq, mq.max(), rtol=1e-12)  # interior edges: 4 cells of eps_r 3 each


@pytest.mark.parametrize("side", ["-x", "-y", "-z", "rho"])
def test_pml_on_lower_sides_is_rejected(tmp_path, side):
    with pytest.raises(TetfitError, match="pml direction"):
        pml_case(tmp_path, {side: 0.2}, name="pmlneg")
