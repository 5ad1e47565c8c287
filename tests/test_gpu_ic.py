"""Device generators of the piecewise initial states against their host
restatement (htx.ic.piecewise_state)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scheme", ["teno5", "hteno"])
def test_quadrant_generator(scheme):
    import htx
    from htx.ic import piecewise_state, quadrant_workload

    n = 96
    w = quadrant_workload(n)
    print(f"[case] quadrant generator {scheme} n={n}", flush=True)
    s = htx.System2D(w["xl"], w["xr"], n, w["yb"], w["yt"], n, w["cfl"], w["tmax"], 0.0, w["gamma"], scheme=scheme,
                     bc="zero_grad")
    s.init_workload(w)
    np.testing.assert_array_equal(s.get_state(), piecewise_state(w, n, n, hermite=scheme == "hteno"))


def test_shock_blob_generator():
    import htx
    from htx.ic import piecewise_state, shock_blob_workload

    nx, ny = 256, 64
    w = shock_blob_workload(nx)
    print(f"[case] shock-blob generator {nx}x{ny}", flush=True)
    s = htx.System2D(w["xl"], w["xr"], nx, w["yb"], w["yt"], ny, w["cfl"], w["tmax"], 0.0, w["gamma"],
                     bc="zero_grad")
    s.init_workload(w)
    np.testing.assert_allclose(s.get_state(), piecewise_state(w, nx, ny), rtol=2e-15, atol=0)
