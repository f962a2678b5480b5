"""Parity at the benchmark's full size: Kuhn 128^3 (12,582,912 cells) with the
benchmark's Gaussian field, VV + classes on the device against the OpenMP
oracle (restatement of src/cpu_extraction.cpp:259-289 and
src/alg/critPoints.cu:49-338), bit-exact, plus the Kuhn closed form of the
VV size (2E, E = 7n^3 + 9n^2 + 3n)."""
import os

import numpy as np
import pytest

from oracle import binding as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", [(2, 0), (3, 0), (4, 0), (1, 0)],
                         ids=["bucketed-auto", "owner-auto", "lean-auto", "vertex-lane"])
def test_kuhn128_gaussians_identical_to_oracle(gpu, form):
    n = 128
    m = gpu.kuhn(n, seed=1234, scalar="gaussians")
    print(f"kuhn{n} gaussians V={m.n_points} T={m.n_cells} form={form}", flush=True)
    with gpu.Mesh(m.cells, m.n_points, m.scalars) as d:
        d.set_config(variant=form[0])
        d.vv_critpoints()
        off, vals = d.relation("VV")
        cls = d.classes()
    E = 7 * n**3 + 9 * n**2 + 3 * n
    assert off[-1] == 2 * E
    threads = min(16, os.cpu_count() or 1)
    VV, ref, _ = O.vv_cp_omp(m.cells, m.n_points, m.scalars, threads)
    assert np.array_equal(off, VV.off)
    assert np.array_equal(vals[:, 0], VV.val[:, 0])
    bad = np.nonzero((cls != ref).any(axis=1))[0]
    assert bad.size == 0, f"{bad.size} vertices differ, first {bad[:5]}"
