"""The reference's on-disk cache format (RW_Cache, 2D/vortex/lfm_midpoint_simulator.py:46-184)
on the CPU: chunk file names and contents, reads across chunks, and the
scalar2vec / vec2mat packing round trips."""
import os

import numpy as np
import pytest

from dfm import rw_cache as W


def test_chunked_series_files_and_reads(tmp_path):
    c = W.RWCache()
    d = str(tmp_path)
    data = {j: np.full((5, 4, 1), j, np.float32) for j in range(250)}
    for j in range(250):
        c.write(data[j], "s1", d, j)
    c.force_write("s1", d)
    assert sorted(os.listdir(d)) == ["s1_0_99.npy", "s1_100_199.npy", "s1_200_249.npy"]
    last = np.load(os.path.join(d, "s1_200_249.npy"))
    assert last.shape == (100, 5, 4, 1)  # the whole buffer is saved, as the reference does
    assert last[49, 0, 0, 0] == 249 and last[50, 0, 0, 0] == 150  # slots past the last entry keep the previous chunk
    r = W.RWCache()
    for j in (0, 99, 100, 173, 249, 5):
        assert np.array_equal(r.read("s1", d, j), data[j])
    with pytest.raises(ValueError):
        r.read("s1", d, 250)
    W.write_disk(7, "final_ind", d, 0)
    assert W.read_disk("final_ind", d, 0).tolist() == [7]


@pytest.mark.parametrize("n", [(6, 5), (4, 3, 5)])
def test_packing_round_trips(n):
    rng = np.random.default_rng(0)
    dim = len(n)
    comps = [rng.standard_normal(W.face_shape(n, a)).astype(np.float32) for a in range(dim)]
    buf = W.pack_mac(comps, n)
    assert buf.shape == tuple(s + 1 for s in n) + (dim,)
    # scalar2vec: component a sits in channel a over its face grid, zero elsewhere
    assert buf[(n[0],) + (0,) * (dim - 1) + (0,)] == comps[0][(n[0],) + (0,) * (dim - 1)]
    assert np.count_nonzero(buf[..., 0]) == comps[0].size
    for a, b in zip(W.unpack_mac(buf, n), comps):
        assert np.array_equal(a, b)
    vecs = [rng.standard_normal(W.face_shape(n, a) + (dim,)).astype(np.float32) for a in range(dim)]
    mb = W.pack_face_vectors(vecs, n)
    assert mb.shape == tuple(s + 1 for s in n) + (dim * dim,)
    # vec2mat row a = face grid a's vector: channels [a*dim, (a+1)*dim)
    assert np.array_equal(mb[(0,) * dim][dim:2 * dim], vecs[1][(0,) * dim])
    for a, b in zip(W.unpack_face_vectors(mb, n), vecs):
        assert np.array_equal(a, b)
    assert W.as_channels(np.zeros(n), dim).shape == tuple(n) + (1,)
