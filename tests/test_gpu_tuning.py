"""The tuning knobs change work, never results: the neighbor search with 2, 4 and 8 waves per workgroup gives the same
rows and counts (plain and slot-major rows, open and periodic boxes), and the batched clouds with the autotuned
workgroup shape still equal their plain runs (test_gpu_batched.py runs at the default shape)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("periodic", [False, True])
def test_neighbor_rows_do_not_depend_on_the_workgroup_shape(hip, periodic):
    from hipcstone import inputs

    n, ngmax, T = 150_000, 96, np.float32
    x, y, z, r = inputs.plummer(n, 0.1, seed=61, dtype=T)
    h = inputs.plummer_h(r, n, 0.1, 70, dtype=T)
    bnd = (1, 1, 1) if periodic else (0, 0, 0)
    box = hip.make_box(-1.0, 1.0, T, bnd)
    pipe = hip.Pipeline(curve=hip.HILBERT, key_bits=64, real_bits=32, bucket_size=32, ngmax=ngmax)
    d = [hip.DeviceArray.from_numpy(a) for a in (x, y, z, h)]
    pipe.sync(*d, n, box, full_build=True)
    v = pipe.view()
    before = hip.get_neighbor_waves()
    out = {}
    try:
        for w in (2, 4, 8):
            hip.set_neighbor_waves(w)
            print(f"neighbor search, {w} waves per workgroup, periodic={periodic}", flush=True)
            nb = hip.DeviceArray(n * ngmax, np.uint32)
            nc = hip.DeviceArray(n, np.uint32)
            nb.fill(0)
            hip.find_neighbors(v.x, v.y, v.z, v.h, 0, n, box, v.child_offsets, v.parents, v.internal_to_leaf,
                               v.layout, v.centers, v.sizes, v.num_nodes, ngmax, nb, nc)
            out[w] = (nb.download(), nc.download())
    finally:
        hip.set_neighbor_waves(before)
    for w in (2, 8):
        assert np.array_equal(out[w][1], out[4][1])
        m = np.minimum(out[4][1], ngmax)
        mask = np.arange(ngmax)[None, :] < m[:, None]
        assert np.array_equal(np.where(mask, out[w][0].reshape(n, ngmax), 0),
                              np.where(mask, out[4][0].reshape(n, ngmax), 0))
    assert out[4][1].mean() > 30


def test_neighbor_waves_rejects_other_shapes(hip):
    with pytest.raises(hip.HipcstoneError):
        hip.set_neighbor_waves(3)
    assert hip.get_neighbor_waves() in (2, 4, 8)
