"""tg_vv_critpoints_noalloc against tg_vv_critpoints on the same meshes.

The no-allocation entry point takes caller-owned device buffers and keeps its
intermediates in a static device array; its VV CSR and classes must equal the
mesh path's exactly, also on a second call with a smaller mesh (the array
keeps the first call's contents).  Device buffers come from torch here.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def mesh_path(t, m):
    with t.Mesh(m.cells, m.n_points, m.scalars) as d:
        d.vv_critpoints()
        off, vals = d.relation("VV")
        return off, vals[:, 0], d.classes()


def noalloc(t, torch, m, f64=False):
    from tetragpu_amd import _native

    T, V = m.n_cells, m.n_points
    dev = "cuda"
    cells = torch.from_numpy(np.ascontiguousarray(m.cells, np.int32)).to(dev)
    sc = torch.from_numpy(np.ascontiguousarray(m.scalars, np.float64 if f64 else np.float32)).to(dev)
    cap = min(12 * T, V * (V - 1))
    off = torch.full((V + 1,), -7, dtype=torch.int32, device=dev)
    cols = torch.full((cap,), -7, dtype=torch.int32, device=dev)
    cls = torch.full((V * 3,), 7, dtype=torch.int32, device=dev)
    st = torch.full((1,), -7, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    torch.cuda.synchronize()
    print(f"noalloc T={T} V={V} f64={f64}", flush=True)
    _native.check(_native.hip().tg_vv_critpoints_noalloc(
        cells.data_ptr(), T, V, sc.data_ptr(), int(f64), off.data_ptr(), cols.data_ptr(), cap,
        cls.data_ptr(), st.data_ptr(), C.c_void_p(stream)))
    torch.cuda.synchronize()
    o = off.cpu().numpy().astype(np.int64)
    return (int(st.item()), o, cols.cpu().numpy()[: o[-1]],
            cls.cpu().numpy().view(np.uint32).reshape(V, 3))


def test_noalloc_matches_mesh_path(gpu):
    import torch

    for n, seed in ((24, 7), (9, 8)):  # second call smaller than the first
        m = gpu.kuhn(n, seed=seed)
        off_ref, cols_ref, cls_ref = mesh_path(gpu, m)
        st, off, cols, cls = noalloc(gpu, torch, m)
        assert st == 0
        assert np.array_equal(off, off_ref), f"VV offsets differ at n={n}"
        assert np.array_equal(cols, cols_ref), f"VV columns differ at n={n}"
        assert np.array_equal(cls, cls_ref), f"classes differ at n={n}"


def test_noalloc_hubs_and_f64(gpu):
    """Vertices over 32 neighbours (heavy kernel) and float64 scalars."""
    import torch

    m = gpu.kuhn(8, seed=11, split_fraction=0.2, n_hubs=3, hub_splits=40)
    m.scalars = m.scalars.astype(np.float64)
    off_ref, cols_ref, cls_ref = mesh_path(gpu, m)
    st, off, cols, cls = noalloc(gpu, torch, m, f64=True)
    assert st == 0
    assert np.array_equal(off, off_ref) and np.array_equal(cols, cols_ref)
    assert np.array_equal(cls, cls_ref)


def test_noalloc_limits(gpu):
    """Sizes beyond the static array and a short vv_cols are refused before
    any work is enqueued."""
    from tetragpu_amd import _native

    lib = _native.hip()
    need, mx = C.c_int64(), C.c_int64()
    _native.check(lib.tg_noalloc_limits(2_000_000, 340_000, C.byref(need), C.byref(mx)))
    assert need.value > mx.value == 128 << 20
    rc = lib.tg_vv_critpoints_noalloc(C.c_void_p(16), 2_000_000, 340_000, C.c_void_p(16), 0,
                                      C.c_void_p(16), C.c_void_p(16), 24_000_000,
                                      C.c_void_p(16), None, None)
    assert rc == -5 and b"scratch" in lib.tg_last_error()
    rc = lib.tg_vv_critpoints_noalloc(C.c_void_p(16), 10, 20, C.c_void_p(16), 0,
                                      C.c_void_p(16), C.c_void_p(16), 119,
                                      C.c_void_p(16), None, None)
    assert rc == -1
