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


class _Series:
    """One named series: its slot buffer, the entry index each filled slot
    holds (in slot order), and the chunk files of the series found on disk as
    (first entry, last entry, file name), listed on the first read."""

    __slots__ = ("slots", "held", "chunks")

    def __init__(self, slots=None):
        self.slots = slots
        self.held = []
        self.chunks = None


class RWCache:
    """Chunked series writer/reader with the reference's file names and
    chunk semantics (a whole CHUNK-slot buffer per file, slot k holding
    entry first + k); one slot buffer per series name, allocated on its
    first write."""

    def __init__(self, n_steps=CHUNK):
        self.n_steps = n_steps
        self._series = {}

    def _for_write(self, name, shape):
        s = self._series.get(name)
        if s is None or s.slots is None or s.slots.shape[1:] != tuple(shape):
            s = _Series(np.zeros((self.n_steps,) + tuple(shape), np.float32))
            self._series[name] = s
        return s

    def init(self):
        for s in self._series.values():
            s.held.clear()
            s.chunks = None

    def write(self, arr, name, dir, j, shape=None):
        """write_disk_with_cache: append entry j (arr in channel layout,
        placed at the origin of a slot of `shape`, default arr's); a full
        buffer is saved first."""
        arr = np.asarray(arr, np.float32)
        s = self._for_write(name, shape or arr.shape)
        if len(s.held) >= self.n_steps:
            self.force_write(name, dir)
        s.slots[(len(s.held),) + tuple(slice(0, k) for k in arr.shape)] = arr
        s.held.append(int(j))

    def force_write(self, name, dir):
        """Save the series' buffer as <name>_<first>_<last>.npy (all slots, as
        the reference saves them) and start a new chunk."""
        s = self._series.get(name)
        if s is not None and s.held:
            np.save(os.path.join(dir, f"{name}_{s.held[0]}_{s.held[-1]}.npy"), s.slots)
            s.held.clear()

    def read(self, name, dir, j):
        """read_disk_with_cache: entry j of the series (channel layout), from
        the chunk in memory or else from the chunk file that covers j."""
        s = self._series.setdefault(name, _Series())
        if s.chunks is None:
            s.chunks = sorted((int(m.group(2)), int(m.group(3)), f) for f in os.listdir(dir)
                              for m in [_SERIES.match(f)] if m and m.group(1) == name)
        if j in s.held:
            return s.slots[s.held.index(j)]
        for first, last, fname in s.chunks:
            if first <= j <= last:
                s.slots = np.load(os.path.join(dir, fname))
                s.held = list(range(first, last + 1))
                return s.slots[j - first]
        raise ValueError(f"series {name!r}: entry {j} lies in no chunk file under {dir}")


def write_disk(data, name, dir, j):
    """write_disk: <dir>/<name>_<j>.npy (a Python number becomes a 1-array)."""
    arr = np.array([data]) if isinstance(data, (int, float)) else np.asarray(data)
    np.save(os.path.join(dir, f"{name}_{j}.npy"), arr)


def read_disk(name, dir, j):
    path = os.path.join(dir, f"{name}_{j}.npy")
    if not os.path.exists(path):
        raise FileNotFoundError(f"read_disk: no file {path}")
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
