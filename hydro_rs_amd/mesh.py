"""Host tessellation (replaces meshless_voronoi::Voronoi::build, src/space.rs:238-262, :574-587)."""
import ctypes as C

import numpy as np

from . import _lib


class VoronoiMesh:
    """Face-indexed Voronoi mesh with a cell->face CSR, built on the host.

    Arrays (numpy, owned copies): area[F], normal[F,3], centroid[F,3], shift[F,3], left[F], right[F]
    (-1 = wall), volume[N], cell_centroid[N,3], search_radius[N], offset[N], count[N], conn[*].
    """

    def __init__(self, dim, box, periodic, loc, n_threads=0, mask=None):
        """mask (optional, [N] bool): Voronoi::build_partial, only the masked generators' cells are built."""
        loc = np.ascontiguousarray(loc, dtype=np.float64).reshape(-1, 3)
        box = np.ascontiguousarray(box, dtype=np.float64)
        h = _lib.host()
        ptr = C.c_void_p()
        if mask is None:
            code = h.hfv_voronoi_build(int(dim), box.ctypes.data_as(_lib.DP), int(bool(periodic)), loc.shape[0],
                                       loc.ctypes.data_as(_lib.DP), int(n_threads), C.byref(ptr))
        else:
            m8 = np.ascontiguousarray(mask, dtype=np.uint8)
            assert m8.size == loc.shape[0]
            code = h.hfv_voronoi_build_partial(int(dim), box.ctypes.data_as(_lib.DP), int(bool(periodic)),
                                               loc.shape[0], loc.ctypes.data_as(_lib.DP),
                                               m8.ctypes.data_as(_lib.U8P), int(n_threads), C.byref(ptr))
        if code != 0:
            raise _lib.HfvError(code, "tessellation failed")
        try:
            v = _lib.MeshArrays()
            h.hfv_mesh_view(ptr, C.byref(v))
            n, nf, nc = v.n_cells, v.n_faces, v.n_conn

            def cp(p, k, dt=np.float64):
                return np.ctypeslib.as_array(p, shape=(k,)).astype(dt, copy=True) if k else np.zeros(0, dt)

            self.area = cp(v.area, nf)
            self.normal = cp(v.normal, 3 * nf).reshape(nf, 3)
            self.centroid = cp(v.centroid, 3 * nf).reshape(nf, 3)
            self.shift = cp(v.shift, 3 * nf).reshape(nf, 3)
            self.left = cp(v.left, nf, np.int64)
            self.right = cp(v.right, nf, np.int64)
            self.volume = cp(v.volume, n)
            self.cell_centroid = cp(v.cell_centroid, 3 * n).reshape(n, 3)
            self.search_radius = cp(v.search_radius, n)
            self.offset = cp(v.offset, n, np.int64)
            self.count = cp(v.count, n, np.int64)
            self.conn = cp(v.conn, nc, np.int64)
        finally:
            h.hfv_mesh_free(ptr)
        self.dim = dim
        self.box = box.copy()
        self.periodic = bool(periodic)
        self.n_cells = loc.shape[0]
        self.n_faces = self.area.size

    def as_dict(self):
        """Flat arrays in the layout of hfv_mesh_arrays (also what the oracle consumes)."""
        return {
            "n_faces": self.n_faces, "area": self.area, "normal": self.normal.ravel(),
            "centroid": self.centroid.ravel(), "shift": self.shift.ravel(), "left": self.left, "right": self.right,
            "conn": self.conn, "offset": self.offset, "count": self.count, "volume": self.volume,
            "cell_centroid": self.cell_centroid.ravel(), "search_radius": self.search_radius,
        }

    def arrays_struct(self):
        """hfv_mesh_arrays pointing into this object's arrays (keep the object alive while in use)."""
        self._flat = {k: np.ascontiguousarray(v) for k, v in self.as_dict().items() if k != "n_faces"}
        f = self._flat
        m = _lib.MeshArrays()
        m.n_cells = self.n_cells
        m.n_faces = self.n_faces
        m.n_conn = f["conn"].size
        for k in ("area", "normal", "centroid", "shift", "volume", "search_radius"):
            setattr(m, k, f[k].ctypes.data_as(_lib.DP))
        m.cell_centroid = f["cell_centroid"].ctypes.data_as(_lib.DP)
        for k in ("left", "right", "offset", "count", "conn"):
            setattr(m, k, f[k].ctypes.data_as(_lib.I64P))
        return m
