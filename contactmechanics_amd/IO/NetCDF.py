"""Frames of displacement and force fields in a NetCDF file.

Mirrors ``ContactMechanics/IO/NetCDF.py:45-403`` (``NetCDFContainer``,
``NetCDFContainerFrame``, ``open``).  The reference writes through netCDF4
when it is installed and otherwise through ``scipy.io.netcdf_file`` (NetCDF 3
classic); netCDF4 is not installed here, so this container always takes the
scipy path, which is the reference's own fallback.  File layout (the same as
the reference writes, so either side reads the other's files):

* dimensions ``frame`` (unlimited), ``nx``, ``ny`` (and ``ndof`` for vector
  fields); a rigid surface of another shape gets ``rigid_nx``, ``rigid_ny``;
* global attributes ``program``, ``programVersion``;
* per-frame variables created on first assignment: Python integers as
  ``i4 (frame)``, reals and 0-d arrays as ``f8 (frame)``, grid-shaped arrays
  as ``f4`` (``f8`` with ``double=True``) ``(frame, nx, ny)`` or
  ``(frame, ndof, nx, ny)``;
* static fields ``h`` (rigid surface) and ``elastic_surface`` as ``f8``.

Files with an ``atom`` dimension (AMBER convention) are read with ``u`` and
``f`` mapped to ``coordinates`` and ``forces`` and reshaped to
``(nx, nx, 3)``.  ``format`` is accepted for interface compatibility and
ignored (NetCDF 3 64-bit offset files are written).
"""

import numbers
from math import isqrt

import numpy as np
from scipy.io import netcdf_file


class NetCDFContainerFrame:
    """One frame of a container; fields are attributes or items."""

    def __init__(self, parent, i):
        self._parent = parent
        self._i = i + len(parent) if i < 0 else i

    @property
    def index(self):
        return self._i

    def get_index(self):
        return self._i

    @property
    def traj(self):
        return self._parent

    def get_traj(self):
        return self._parent

    def _var(self, name):
        return self._parent._data.variables[self._parent._mangle(name)]

    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        value = np.array(self._var(name)[self._i])
        if self._parent._is_amber:
            nx, ny = self._parent._shape
            value = value.reshape(nx, ny, 3)
        return value

    def __setattr__(self, name, value):
        if name.startswith("_"):
            object.__setattr__(self, name, value)
            return
        key = self._parent._mangle(name)
        self._parent._declare_frame_variable(key, value)
        self._parent._data.variables[key][self._i] = value

    def __getitem__(self, name):
        return self.__getattr__(name)

    def __setitem__(self, name, value):
        self.__setattr__(name, value)

    def set_grid(self, name, value, x0=0, y0=0):
        """Write ``value`` into the grid variable ``name`` at offset (x0, y0);
        the rest of the frame's grid keeps the file's fill value."""
        key = self._parent._mangle(name)
        self._parent._declare_frame_variable(key, value, shape=self._parent._shape)
        nx, ny = np.shape(value)
        self._parent._data.variables[key][self._i, x0:x0 + nx, y0:y0 + ny] = value

    def sync(self):
        self._parent.sync()


class NetCDFContainer:
    """A NetCDF file of frames on one grid (see module docstring)."""

    def __init__(self, fn, frame=0, double=False, store_force=False, mode="r", format="NETCDF4"):
        self._fn = fn
        self._data = None
        self._float_str = "f8" if double else "f4"
        self._store_force = store_force
        self._is_amber = False
        self._shape = self._shape2 = None
        m = "w" if mode == "ws" else mode
        try:
            self._data = netcdf_file(fn, m, version=2, mmap=False)
        except (OSError, RuntimeError, TypeError, ValueError) as err:
            raise RuntimeError('Error opening file "{0}": {1}'.format(fn, err)) from err
        if m == "w":
            self._data.program = "PyCo"
            self._data.programVersion = "N/A"
            self._is_defined = False
        else:
            dims = self._data.dimensions
            if "nx" in dims and "ny" in dims:
                self._shape = self._shape2 = (int(dims["nx"]), int(dims["ny"]))
            elif "atom" in dims:
                n = int(dims["atom"])
                nx = isqrt(n)
                if nx * nx != n:
                    raise RuntimeError("AMBER file {0}: {1} atoms is not a square grid".format(fn, n))
                self._shape = self._shape2 = (nx, nx)
                self._is_amber = True
            else:
                raise RuntimeError("Unknown NetCDF convention used for file %s." % fn)
            self._is_defined = True
        self._cur_frame = len(self) + frame if frame < 0 else frame

    # ---- structure
    def _mangle(self, name):
        if self._is_amber:
            return {"u": "coordinates", "f": "forces"}.get(name, name)
        return name

    def _define_file_structure(self, shape):
        shape = tuple(int(n) for n in shape)
        ndof, (nx, ny) = (shape[0], shape[1:]) if len(shape) == 3 else (1, shape)
        self._shape, self._shape2 = shape, (nx, ny)
        dims = self._data.dimensions
        if "frame" not in dims:
            self._data.createDimension("frame", None)
        if ndof > 1 and "ndof" not in dims:
            self._data.createDimension("ndof", ndof)
        for name, n in (("nx", nx), ("ny", ny)):
            if name not in dims:
                self._data.createDimension(name, n)
        self._is_defined = True

    def set_shape(self, x, ndof=None):
        shape = tuple(getattr(x, "shape", x))
        if ndof is not None and ndof != 1:
            shape = (ndof,) + shape
        if not self._is_defined:
            self._define_file_structure(shape)
        elif tuple(shape) != tuple(self._shape):
            raise RuntimeError("Shape mismatch: NetCDF file has shape {0}, but someone is trying to override this "
                               "with shape {1}.".format(self._shape, shape))

    def _declare_frame_variable(self, name, value, shape=None):
        if name in self._data.variables:
            return
        if "frame" not in self._data.dimensions:
            raise RuntimeError("set_shape() must define the grid before frames are written")
        if isinstance(value, (bool, np.bool_)) or isinstance(value, numbers.Integral):
            self._data.createVariable(name, "i4", ("frame",))
            return
        if isinstance(value, numbers.Real):
            self._data.createVariable(name, "f8", ("frame",))
            return
        if isinstance(value, np.ndarray):
            shape = tuple(value.shape if shape is None else shape)
            if shape == ():
                self._data.createVariable(name, "f8", ("frame",))
            elif shape == tuple(self._shape):
                dims = ("frame", "ndof", "nx", "ny") if len(self._shape) == 3 else ("frame", "nx", "ny")
                self._data.createVariable(name, self._float_str, dims)
            elif len(shape) == 2 and shape == tuple(self._shape2):
                self._data.createVariable(name, self._float_str, ("frame", "nx", "ny"))
            else:
                raise RuntimeError('Not sure how to guess NetCDF type for field "{0}" which is a numpy ndarray with '
                                   "type {1} and shape {2}.".format(name, value.dtype, value.shape))
            return
        raise RuntimeError('Not sure how to guess NetCDF type for field "{0}" with type {1}.'.format(name, type(value)))

    # ---- frames
    def __len__(self):
        if "frame" not in self._data.dimensions:
            return 0
        n = self._data.dimensions["frame"]
        if n is None:  # unlimited: the record count
            return int(self._data._recs)
        return int(n)

    def __getitem__(self, i):
        if isinstance(i, str):
            return self.__getattr__(i)
        if isinstance(i, slice):
            return [NetCDFContainerFrame(self, j) for j in range(*i.indices(len(self)))]
        return NetCDFContainerFrame(self, i)

    def __setitem__(self, i, value):
        if isinstance(i, str):
            return self.__setattr__(i, value)
        raise RuntimeError("Cannot set full frame.")

    def __iter__(self):
        for i in range(len(self)):
            yield NetCDFContainerFrame(self, i)

    def get_next_frame(self):
        frame = NetCDFContainerFrame(self, self._cur_frame)
        self._cur_frame += 1
        return frame

    def set_cursor(self, cur_frame):
        self._cur_frame = cur_frame

    def get_cursor(self):
        return self._cur_frame

    # ---- static fields
    def has_h(self):
        return "h" in self._data.variables

    def set_rigid_surface(self, h, ndof=None):
        self.set_shape(h, ndof=ndof)
        hnx, hny = np.shape(h)
        if "h" not in self._data.variables:
            if (hnx, hny) != tuple(self._shape2):
                for name, n in (("rigid_nx", hnx), ("rigid_ny", hny)):
                    if name not in self._data.dimensions:
                        self._data.createDimension(name, n)
                self._data.createVariable("h", "f8", ("rigid_nx", "rigid_ny"))
            else:
                self._data.createVariable("h", "f8", ("nx", "ny"))
        self._data.variables["h"][:, :] = h

    set_h = set_rigid_surface

    def get_rigid_surface(self):
        return np.array(self._data.variables["h"][:, :])

    get_h = get_rigid_surface

    def set_elastic_surface(self, h):
        if "elastic_surface" not in self._data.variables:
            self._data.createVariable("elastic_surface", "f8", ("nx", "ny"))
        self._data.variables["elastic_surface"][:, :] = h

    def get_elastic_surface(self):
        return np.array(self._data.variables["elastic_surface"][:, :])

    # ---- whole-file attributes and fields
    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        data = self.__dict__.get("_data")
        if data is None:
            raise AttributeError(name)
        if name in data.variables:
            return np.array(data.variables[name][...])
        if name in data._attributes:
            value = data._attributes[name]
            return value.decode() if isinstance(value, bytes) else value
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name.startswith("_"):
            object.__setattr__(self, name, value)
            return
        if isinstance(value, np.ndarray) and value.shape != ():
            if name not in self._data.variables:
                if self._shape is not None and tuple(value.shape) == tuple(self._shape):
                    self._data.createVariable(name, "f8", ("nx", "ny"))
                else:
                    raise RuntimeError('Not sure how to guess NetCDF type for field "{0}" which is a numpy ndarray '
                                       "with type {1} and shape {2}.".format(name, value.dtype, value.shape))
            self._data.variables[name][:, :] = value
            return
        setattr(self._data, name, value)

    def get_size(self):
        return self._shape

    def get_filename(self):
        return self._fn

    def sync(self):
        self._data.flush()

    def close(self):
        if self._data is not None:
            self._data.close()
            self._data = None
            self._is_defined = False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __str__(self):
        return self._fn


def open(fn, mode="r", frame=None, **kwargs):  # noqa: A001  (the reference's module-level name)
    """A container, or one of its frames for ``"file.nc@N"`` or ``frame=N``."""
    if isinstance(fn, NetCDFContainer):
        return fn
    i = fn.find("@")
    if i > 0:
        return NetCDFContainer(fn[:i], mode=mode, **kwargs)[int(fn[i + 1:])]
    if frame is not None:
        return NetCDFContainer(fn, mode=mode, **kwargs)[frame]
    return NetCDFContainer(fn, mode=mode, **kwargs)

