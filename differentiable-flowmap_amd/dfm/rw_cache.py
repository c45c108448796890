"""The reference's on-disk storage format (RW_Cache and write_disk / read_disk,
2D/vortex/lfm_midpoint_simulator.py:46-184,453-470; 3D/lfm_midpoint_simulator.py:46-165),
for debugging and for resuming an adjoint sweep from files.

The simulators keep this data in HBM; ``LFM_Simulator.export_cache(log_dir)``
writes it in this layout and ``LFM_Diff_Simulator.import_cache(log_dir)``
reads it back, so runs can be exchanged with the reference's file layout:

* chunked series ``<dir>/<name>_<first>_<last>.npy``: a (100, X, Y[, Z], C)
  FP32 buffer whose slot i holds entry first + i (written when 100 entries
  have accumulated, or by force_write; the whole 100-slot buffer is saved);
* single arrays ``<dir>/<name>_<j>.npy`` (write_disk / read_disk).

Packed shapes (scalar2vec, vec2mat, 2D/vortex/taichi_utils.py:21-54):
MAC velocity -> (n_0 + 1, ..., dim), component a in channel a over its face
grid, zero elsewhere; face maps / Jacobian columns -> (n_0 + 1, ..., dim*dim),
face grid a's vector in channels [a*dim, (a+1)*dim); cell fields ->
(n..., C).
"""
from __future__ import annotations

import os
import re

import numpy as np

CHUNK = 100
_SERIES = re.compile(r"^(.*)_(\d+)_(\d+)\.npy$")


def face_shape(n, a):
    return tuple(n[b] + (1 if b == a else 0) for b in range(len(n)))


def pack_mac(comps, n):
    """scalar2vec: MAC components -> (n+1..., dim)."""
    dim = len(n)
    out = np.zeros(tuple(s + 1 for s in n) + (dim,), np.float32)
    for a in range(dim):
        out[tuple(slice(0, s) for s in face_shape(n, a)) + (a,)] = comps[a]
    return out


def unpack_mac(buf, n):
    """vec2scalar: (n+1..., dim) -> MAC components."""
    return [np.ascontiguousarray(buf[tuple(slice(0, s) for s in face_shape(n, a)) + (a,)], dtype=np.float32)
            for a in range(len(n))]


def pack_face_vectors(vecs, n):
    """vec2mat: per face grid a a (*face_a, dim) vector field -> (n+1..., dim*dim)."""
    dim = len(n)
    out = np.zeros(tuple(s + 1 for s in n) + (dim * dim,), np.float32)
    for a in range(dim):
        out[tuple(slice(0, s) for s in face_shape(n, a)) + (slice(a * dim, (a + 1) * dim),)] = vecs[a]
    return out


def unpack_face_vectors(buf, n):
    dim = len(n)
    return [np.ascontiguousarray(buf[tuple(slice(0, s) for s in face_shape(n, a)) + (slice(a * dim, (a + 1) * dim),)],
                                 dtype=np.float32) for a in range(dim)]


def as_channels(arr, spatial_dim):
    """write_disk_with_cache's reshape: a scalar field gains a channel axis, a
    matrix field's last two axes merge."""
    arr = np.asarray(arr, np.float32)
    if arr.ndim == spatial_dim:
        return arr[..., None]
    if arr.ndim == spatial_dim + 2:
        return arr.reshape(arr.shape[:spatial_dim] + (arr.shape[-2] * arr.shape[-1],))
    return arr


class RWCache:
    """Chunked series writer/reader with the reference's file names and
    buffer semantics; buffers are allocated per name on first use."""

    def __init__(self, n_steps=CHUNK):
        self.n_steps = n_steps
        self.buffer = {}

    def _entry(self, name, shape):
        e = self.buffer.get(name)
        if e is None or e["buffer"].shape[1:] != tuple(shape):
            e = {"buffer": np.zeros((self.n_steps,) + tuple(shape), np.float32), "info": [], "file_info": []}
            self.buffer[name] = e
        return e

    def init(self):
        for e in self.buffer.values():
            e["info"].clear()
            e["file_info"].clear()

    def write(self, arr, name, dir, j, shape=None):
        """write_disk_with_cache: append entry j (arr in channel layout,
        placed at the origin of a buffer of `shape`, default arr's)."""
        arr = np.asarray(arr, np.float32)
        e = self._entry(name, shape or arr.shape)
        if len(e["info"]) >= self.n_steps:
            self.force_write(name, dir)
        e["buffer"][(len(e["info"]),) + tuple(slice(0, s) for s in arr.shape)] = arr
        e["info"].append(int(j))

    def force_write(self, name, dir):
        e = self.buffer.get(name)
        if e is not None and e["info"]:
            info = e["info"]
            np.save(os.path.join(dir, f"{name}_{info[0]}_{info[-1]}.npy"), e["buffer"])
            info.clear()

    def read(self, name, dir, j):
        """read_disk_with_cache: the buffer slot of entry j (channel layout)."""
        e = self.buffer.get(name)
        if e is None:
            e = {"buffer": None, "info": [], "file_info": []}
            self.buffer[name] = e
        if not e["file_info"]:
            for fname in os.listdir(dir):
                m = _SERIES.match(fname)
                if m and m.group(1) == name:
                    e["file_info"].append((int(m.group(2)), int(m.group(3)), fname))
            e["file_info"].sort()
        info = e["info"]
        if info and info[0] <= j <= info[-1] and j in info:
            return e["buffer"][info.index(j)]
        for lo, hi, fname in e["file_info"]:
            if lo <= j <= hi:
                data = np.load(os.path.join(dir, fname))
                e["buffer"] = data
                info.clear()
                info.extend(range(lo, hi + 1))
                return data[info.index(j)]
        raise ValueError(f"Frame {j} not found in any available file.")


def write_disk(data, name, dir, j):
    """write_disk: <dir>/<name>_<j>.npy (a Python number becomes a 1-array)."""
    arr = np.array([data]) if isinstance(data, (int, float)) else np.asarray(data)
    np.save(os.path.join(dir, f"{name}_{j}.npy"), arr)


def read_disk(name, dir, j):
    path = os.path.join(dir, f"{name}_{j}.npy")
    if not os.path.exists(path):
        raise FileNotFoundError(f"File not found: {path}")
    return np.load(path)


def layout(dim, n_passive=1):
    """Directory and series names the reference uses (2D/vortex/lfm_midpoint_simulator.py:371-400,
    3D/lfm_midpoint_simulator.py:395-412): {key: (directory, series name)}; the
    passive entry holds one name per passive scalar."""
    xy = "phi_xy" if dim == 2 else "phi_xyz"
    if dim == 2:
        passive = [f"s{k + 1}" for k in range(n_passive)]
    else:
        passive = ["s"] + [f"s{k + 1}" for k in range(1, n_passive)]
    return {"midpoint": ("midpoint", "u"), "phi": ("flowmap_phi", "phi"), "phi_xy": ("flowmap_" + xy, xy),
            "F": ("flowmap_F", "F"), "passive": ("passive", passive), "final": ("final", None)}
