"""The reference's own meshes (mwe, tiny, big of artificial_dataset.sh) on
the GPU: VV, classes and every relation against the oracle, with the
float64 scalars the reference reads from the .vtu files and with those
scalars rounded to float32."""
import numpy as np
import pytest

from oracle import binding as O
from test_gpu_relations import check_all
from test_ref_meshes import load_mesh

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["mwe", "tiny", "big"])
@pytest.mark.parametrize("f64", [True, False])
def test_vv_classes(gpu, name, f64):
    cells, V, sc = load_mesh(name)
    if not f64:
        sc = sc.astype(np.float32)
    print(f"ref mesh {name} f64={f64}", flush=True)
    with gpu.Mesh(cells, V, sc) as d:
        d.vv_critpoints()
        off, vals = d.relation("VV")
        cls = d.classes()
    VV = O.vv(cells, V)
    S = VV.sorted_rows() if V <= 1000 else O.vv_cp_omp(cells, V, np.zeros(V, np.float32))[0]
    assert np.array_equal(off, S.off) and np.array_equal(vals[:, 0], S.val[:, 0])
    assert np.array_equal(cls, O.critpoints(VV, sc.astype(np.float64), V))


@pytest.mark.parametrize("name", ["mwe", "tiny", "big"])
def test_all_relations(gpu, name):
    cells, V, _ = load_mesh(name)
    print(f"ref mesh {name} relations", flush=True)
    ref = check_all(gpu, cells, V)
    assert ref["E"] == 3 * len(cells) + 3 and ref["F"] == 3 * len(cells) + 1
