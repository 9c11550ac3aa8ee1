"""Reader of AxiSEM3D's Exodus meshes (NetCDF4 files, read through the
HDF5 C library with ctypes; ExodusMesh.cpp:205-619 lists the variables).

ExodusFile(path) exposes the datasets the mesher writes: the quad
connectivity (connect1, 1-based in the file, 0-based here), the nodal
coordinates (coordx, coordy), side sets (ss_names, elem_ss<i>, side_ss<i>,
0-based), elemental variables by name (name_elem_var and
vals_elem_var<i>eb1), global variables (name_glo_var, vals_glo_var) and the
info records.  ExodusMesh wraps them in the element interface the preloop
builds domains from (nodal (s, z), axial elements, the bilinear mapping of
linear elements, GLL points shared through the connectivity)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import spectral

NPED, NPEM = 5, 25

_HDF5_PATHS = ("/opt/conda/lib/libhdf5.so", "libhdf5.so")
_hid = C.c_int64


def _hdf5():
    """The HDF5 C library with the handful of calls the reader needs."""
    for p in _HDF5_PATHS:
        try:
            h = C.CDLL(p)
            break
        except OSError:
            continue
    else:
        raise RuntimeError("exodus: the HDF5 C library (libhdf5.so) is not "
                           "available")
    h.H5open()
    h.H5Eset_auto2.argtypes = [_hid, C.c_void_p, C.c_void_p]
    h.H5Eset_auto2(0, None, None)
    h.H5Fopen.restype = _hid
    h.H5Fopen.argtypes = [C.c_char_p, C.c_uint, _hid]
    h.H5Fclose.argtypes = [_hid]
    h.H5Lexists.argtypes = [_hid, C.c_char_p, _hid]
    h.H5Dopen2.restype = _hid
    h.H5Dopen2.argtypes = [_hid, C.c_char_p, _hid]
    h.H5Dclose.argtypes = [_hid]
    h.H5Dget_space.restype = _hid
    h.H5Dget_space.argtypes = [_hid]
    h.H5Dget_type.restype = _hid
    h.H5Dget_type.argtypes = [_hid]
    h.H5Sget_simple_extent_ndims.argtypes = [_hid]
    h.H5Sget_simple_extent_dims.argtypes = [_hid, C.POINTER(C.c_uint64),
                                            C.c_void_p]
    h.H5Sclose.argtypes = [_hid]
    h.H5Tget_class.argtypes = [_hid]
    h.H5Tget_size.restype = C.c_size_t
    h.H5Tget_size.argtypes = [_hid]
    h.H5Tclose.argtypes = [_hid]
    h.H5Dread.argtypes = [_hid, _hid, _hid, _hid, _hid, C.c_void_p]
    return h


_H5T_INTEGER, _H5T_FLOAT, _H5T_STRING = 0, 1, 3


class ExodusFile:
    """Datasets of one Exodus (NetCDF4/HDF5) file, read whole on demand."""

    def __init__(self, path):
        self.path = path
        self._h = _hdf5()
        self._f = self._h.H5Fopen(os.fsencode(path), 0, 0)
        if self._f < 0:
            raise OSError(f"exodus: cannot open {path}")
        # the native types are library globals, set by H5open
        self._t_double = _hid.in_dll(self._h, "H5T_NATIVE_DOUBLE_g").value
        self._t_int32 = _hid.in_dll(self._h, "H5T_NATIVE_INT_g").value

    def close(self):
        if self._f >= 0:
            self._h.H5Fclose(self._f)
            self._f = -1

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def has(self, name):
        return self._h.H5Lexists(self._f, name.encode(), 0) > 0

    def read(self, name):
        """The dataset `name` as a numpy array (float64, int32, or for
        character arrays the list of its rows as str)."""
        h = self._h
        d = h.H5Dopen2(self._f, name.encode(), 0)
        if d < 0:
            raise KeyError(f"exodus: no dataset {name} in {self.path}")
        try:
            sp = h.H5Dget_space(d)
            nd = h.H5Sget_simple_extent_ndims(sp)
            dims = (C.c_uint64 * max(nd, 1))()
            h.H5Sget_simple_extent_dims(sp, dims, None)
            h.H5Sclose(sp)
            shape = tuple(int(x) for x in dims[:nd])
            t = h.H5Dget_type(d)
            cls, size = h.H5Tget_class(t), h.H5Tget_size(t)
            try:
                if cls == _H5T_FLOAT:
                    out = np.empty(shape, dtype=np.float64)
                    mt = self._t_double
                elif cls == _H5T_INTEGER:
                    out = np.empty(shape, dtype=np.int32)
                    mt = self._t_int32
                elif cls == _H5T_STRING and size == 1:
                    out = np.empty(shape, dtype=np.uint8)
                    mt = t
                else:
                    raise TypeError(f"exodus: dataset {name} has an "
                                    f"unsupported type (class {cls})")
                if out.size and h.H5Dread(d, mt, 0, 0, 0,
                                          out.ctypes.data) < 0:
                    raise OSError(f"exodus: cannot read {name}")
            finally:
                h.H5Tclose(t)
        finally:
            h.H5Dclose(d)
        if out.dtype == np.uint8:
            rows = out.reshape(-1, shape[-1]) if out.ndim > 1 else out[None]
            return [bytes(r).split(b"\0", 1)[0].decode().strip()
                    for r in rows]
        return out


class ExodusMesh:
    """An AxiSEM3D Exodus mesh (ExodusMesh::ExodusMesh, ExodusMesh.cpp):
    quads with 4 corner nodes, their geometry type (0 spherical, 1 linear,
    2 semi-spherical) and medium (solid or fluid), side sets, global
    variables and records, and the radial model variables stored per
    element node (VAR_0..VAR_3) or per element."""

    def __init__(self, path=None, arrays=None):
        """From an Exodus file, or from the arrays to_arrays() gave."""
        if arrays is not None:
            self._from_arrays(arrays)
        else:
            self._read(path)
        self.geometry_type = np.round(
            self.elem_vars["element_type"]).astype(np.int64)
        self.fluid = np.round(self.elem_vars["fluid"]).astype(np.int64) != 0
        self.cartesian = self.records.get("crdsys", "").lower() == \
            "cartesian"
        left = "x0" if self.cartesian else "t0"
        self._orient_axial(left)
        self.left = np.full(self.nelem, -1, dtype=np.int64)
        for e, s in self.side_sets.get(left, {}).items():
            self.left[e] = s

    def _orient_axial(self, left):
        """Cycles the nodes of every axial element so that its side on the
        axis becomes side 3 (xi = -1), with its element-node variables and
        side-set sides (ExodusMesh.cpp:585-610); a no-op on an oriented
        mesh."""
        node_vars = sorted({k[:-2] for k in self.elem_vars
                            if k[-2:] == "_0"})
        for e, a in list(self.side_sets.get(left, {}).items()):
            if a == 3:
                continue
            src = [(k + a - 3) % 4 for k in range(4)]
            self.connectivity[e] = self.connectivity[e][src]
            for v in node_vars:
                vals = [self.elem_vars[f"{v}_{k}"][e] for k in range(4)]
                for k in range(4):
                    self.elem_vars[f"{v}_{k}"][e] = vals[src[k]]
            for m in self.side_sets.values():
                if e in m:
                    m[e] = (m[e] - a + 3) % 4

    def _read(self, path):
        f = ExodusFile(path)
        self.path = path
        self.connectivity = f.read("connect1").astype(np.int64) - 1
        self.coords = np.stack([f.read("coordx"), f.read("coordy")], axis=1)
        names = f.read("name_glo_var") if f.has("name_glo_var") else []
        vals = (np.asarray(f.read("vals_glo_var")).reshape(-1)
                if f.has("vals_glo_var") else np.zeros(0))
        self.global_variables = dict(zip(names, vals.tolist()))
        self.records = {}
        for r in (f.read("info_records") if f.has("info_records") else []):
            if "=" in r:
                k, v = r.split("=", 1)
                self.records[k.strip()] = v.strip()
        self.side_sets = {}
        for i, n in enumerate(f.read("ss_names")):
            el = f.read(f"elem_ss{i + 1}").astype(np.int64) - 1
            sd = f.read(f"side_ss{i + 1}").astype(np.int64) - 1
            self.side_sets[n] = dict(zip(el.tolist(), sd.tolist()))
        evn = f.read("name_elem_var")
        self.elem_vars = {n: np.asarray(f.read(f"vals_elem_var{i + 1}eb1"),
                                        dtype=np.float64).reshape(-1)
                          for i, n in enumerate(evn)}
        f.close()

    def to_arrays(self, elem_vars=None):
        """The mesh as plain arrays (numpy.savez): connectivity, coords,
        side sets, the named elemental variables (all by default), global
        variables and the crdsys record."""
        keep = list(self.elem_vars) if elem_vars is None else elem_vars
        out = {"connectivity": self.connectivity.astype(np.int32),
               "coords": self.coords,
               "crdsys": np.array(self.records.get("crdsys", "")),
               "glo_names": np.array(list(self.global_variables)),
               "glo_values": np.array(list(self.global_variables.values()),
                                      dtype=np.float64),
               "ss_names": np.array(list(self.side_sets))}
        for i, (n, m) in enumerate(self.side_sets.items()):
            out[f"ss{i}"] = np.array(sorted(m.items()),
                                     dtype=np.int32).reshape(-1, 2)
        out["ev_names"] = np.array(keep)
        for i, n in enumerate(keep):
            out[f"ev{i}"] = self.elem_vars[n]
        return out

    def _from_arrays(self, a):
        self.path = None
        self.connectivity = np.asarray(a["connectivity"], dtype=np.int64)
        self.coords = np.asarray(a["coords"], dtype=np.float64)
        self.records = {"crdsys": str(a["crdsys"])}
        self.global_variables = dict(zip(
            [str(x) for x in a["glo_names"]],
            np.asarray(a["glo_values"], dtype=np.float64).tolist()))
        self.side_sets = {str(n): {int(e): int(s) for e, s in a[f"ss{i}"]}
                          for i, n in enumerate(a["ss_names"])}
        self.elem_vars = {str(n): np.asarray(a[f"ev{i}"], dtype=np.float64)
                          for i, n in enumerate(a["ev_names"])}

    @property
    def nelem(self):
        return len(self.connectivity)

    def nodal_variable(self, name):
        """(E, 4) values of a model variable at the element corners: the
        element-node storage VAR_0..VAR_3 or one value per element."""
        if f"{name}_0" in self.elem_vars:
            return np.stack([self.elem_vars[f"{name}_{k}"]
                             for k in range(4)], axis=1)
        return np.repeat(self.elem_vars[name][:, None], 4, axis=1)
