"""MPS files, read and written by the library (the C++ reader behind
anuk_mps_read, which accepts what read_mps accepts and builds the same model:
src/anukulan/mps_reader.hpp:12-50, mps_reader.cpp:711-728).

    m = read_mps("plan.mps")          # MpsModel: the file's model, names, warnings
    lp = m.to_lp()                    # standard form for solve_pdhg
    write_mps("copy.mps", m)          # reads back to the same model, bit for bit

A file the reader rejects raises AnukError with code ANUK_ERR_PARSE (6) and
the reader's "source:line: what" message; a file that cannot be opened raises
with ANUK_ERR_IO (7).
"""
from __future__ import annotations

from ctypes import POINTER, byref, c_char_p, c_void_p
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib
from ._lib import DP, IP, lib
from .api import LP, AnukError, _dp, _f64, _i32, _ip

FORMATS = {"auto": 0, "free": 1, "fixed": 2}
NEGATIVE_UP = {"highs": 0, "cplex": 1}
COL_TYPES = ("continuous", "integer", "semi-continuous", "semi-integer")
PARSE_ERROR, IO_ERROR = 6, 7


@dataclass
class MpsModel:
    """The model as the file states it (model.hpp:22-45): A as CSR, the
    objective c'x + offset (+ x'Qx/2 when `hessian` is given), row and column
    bounds, column types (0 continuous, 1 integer, 2 semi-continuous,
    3 semi-integer) and names, plus the reader's warnings and counters."""
    name: str
    rows: int
    cols: int
    start: np.ndarray
    index: np.ndarray
    value: np.ndarray
    objective: np.ndarray
    row_lower: np.ndarray
    row_upper: np.ndarray
    col_lower: np.ndarray
    col_upper: np.ndarray
    col_type: np.ndarray
    row_names: List[str]
    col_names: List[str]
    maximize: bool = False
    objective_offset: float = 0.0
    hessian: Optional[tuple] = None
    warnings: List[str] = field(default_factory=list)
    free_rows_dropped: int = 0
    objective_nonzeros: int = 0
    fixed_format: bool = False

    @property
    def nnz(self) -> int:
        return int(self.value.size)

    def row_index(self) -> np.ndarray:
        return np.repeat(np.arange(self.rows, dtype=np.int32), np.diff(self.start))

    def to_natural(self):
        """The same model as a generators.NaturalModel (triplets in CSR order)."""
        from .generators import NaturalModel
        return NaturalModel(rows=self.rows, cols=self.cols, row_index=self.row_index(),
                            col_index=self.index.copy(), value=self.value.copy(),
                            objective=self.objective.copy(), row_lower=self.row_lower.copy(),
                            row_upper=self.row_upper.copy(), col_lower=self.col_lower.copy(),
                            col_upper=self.col_upper.copy(), maximize=self.maximize,
                            objective_offset=self.objective_offset, name=self.name)

    def to_lp(self) -> LP:
        """Standard form of the LP (to_standard_form): the continuous
        relaxation. Raises for a quadratic objective, which PDHG does not take."""
        if self.hessian is not None:
            raise ValueError(f"{self.name}: the file has a quadratic objective; PDHG solves LPs")
        return LP.from_model(self.rows, self.cols, self.row_index(), self.index, self.value,
                             self.objective, self.row_lower, self.row_upper, self.col_lower,
                             self.col_upper, self.maximize, self.objective_offset)


def options(format="auto", negative_up_bound="highs", infinity_threshold=1e30) -> _lib.MpsOptions:
    o = _lib.MpsOptions()
    lib.anuk_mps_default_options(byref(o))
    o.format = FORMATS[format]
    o.negative_up_bound = NEGATIVE_UP[negative_up_bound]
    o.infinity_threshold = float(infinity_threshold)
    return o


def _copy(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


def _names_of(get, h, count):
    """The names, or [] for a model that has none (one built in memory)."""
    if count == 0 or get(h, 0) is None:
        return []
    return [get(h, k).decode() for k in range(count)]


def _take(h, owned=True) -> MpsModel:
    """Copies the model out of an anuk_mps handle, freeing the handle when
    this side owns it."""
    try:
        info = _lib.MpsInfo()
        lib.anuk_mps_info_get(h, byref(info))
        m = _lib.Model()
        lib.anuk_mps_model(h, byref(m))
        rows, cols, nnz = info.rows, info.cols, int(info.nnz)
        s, i, v = IP(), IP(), DP()
        lib.anuk_mps_matrix(h, byref(s), byref(i), byref(v))
        start = _copy(s, rows + 1, np.int32)
        hess = None
        if info.hessian_nnz > 0:
            hs, hi, hv = IP(), IP(), DP()
            lib.anuk_mps_hessian(h, byref(hs), byref(hi), byref(hv))
            hn = int(info.hessian_nnz)
            hess = (_copy(hs, cols + 1, np.int32), _copy(hi, hn, np.int32), _copy(hv, hn, np.float64))
        return MpsModel(
            name=lib.anuk_mps_name(h).decode(), rows=rows, cols=cols, start=start,
            index=_copy(i, nnz, np.int32), value=_copy(v, nnz, np.float64),
            objective=_copy(m.objective, cols, np.float64),
            row_lower=_copy(m.row_lower, rows, np.float64),
            row_upper=_copy(m.row_upper, rows, np.float64),
            col_lower=_copy(m.col_lower, cols, np.float64),
            col_upper=_copy(m.col_upper, cols, np.float64),
            col_type=_copy(lib.anuk_mps_col_types(h), cols, np.int32),
            row_names=_names_of(lib.anuk_mps_row_name, h, rows),
            col_names=_names_of(lib.anuk_mps_col_name, h, cols),
            maximize=bool(info.maximize), objective_offset=info.objective_offset, hessian=hess,
            warnings=[lib.anuk_mps_warning(h, k).decode() for k in range(info.warnings)],
            free_rows_dropped=info.free_rows_dropped,
            objective_nonzeros=int(info.objective_nonzeros),
            fixed_format=bool(info.fixed_format))
    finally:
        if owned:
            lib.anuk_mps_free(h)


def read_mps(path: str, **opts) -> MpsModel:
    h = c_void_p()
    rc = lib.anuk_mps_read(str(path).encode(), byref(options(**opts)), byref(h))
    if rc != 0:
        raise AnukError(rc, lib.anuk_last_error(None).decode())
    return _take(h)


def read_mps_text(text, source: str = "<text>", **opts) -> MpsModel:
    data = text.encode() if isinstance(text, str) else bytes(text)
    h = c_void_p()
    rc = lib.anuk_mps_read_text(data, len(data), source.encode(), byref(options(**opts)), byref(h))
    if rc != 0:
        raise AnukError(rc, lib.anuk_last_error(None).decode())
    return _take(h)


def _names(names, count):
    if names is None:
        return None, None
    if len(names) != count:
        raise ValueError("one name per row (column) is needed")
    encoded = [str(n).encode() for n in names]
    arr = (c_char_p * max(count, 1))(*encoded)
    return arr, encoded


def write_mps(path: str, model, name: Optional[str] = None, row_names=None, col_names=None,
              col_type=None) -> None:
    """Writes an MpsModel or a generators.NaturalModel as free-format MPS.
    Names and column types default to the model's own (MpsModel) or to
    generated names and continuous columns."""
    if isinstance(model, MpsModel):
        ri, ci, v = model.row_index(), model.index, model.value
        row_names = (model.row_names or None) if row_names is None else row_names
        col_names = (model.col_names or None) if col_names is None else col_names
        col_type = model.col_type if col_type is None else col_type
    else:
        ri, ci, v = model.row_index, model.col_index, model.value
    ri, ci, v = _i32(ri), _i32(ci), _f64(v)
    arrs = [_f64(a) for a in (model.objective, model.row_lower, model.row_upper,
                              model.col_lower, model.col_upper)]
    m = _lib.Model(int(model.rows), int(model.cols), int(v.size), _ip(ri), _ip(ci), _dp(v),
                   _dp(arrs[0]), _dp(arrs[1]), _dp(arrs[2]), _dp(arrs[3]), _dp(arrs[4]),
                   1 if model.maximize else 0, float(model.objective_offset))
    rn, _keep_r = _names(row_names, model.rows)
    cn, _keep_c = _names(col_names, model.cols)
    ct = _i32(col_type) if col_type is not None else None
    title = (name if name is not None else getattr(model, "name", "")) or ""
    rc = lib.anuk_mps_write(str(path).encode(), byref(m), title.encode(),
                            rn if rn is not None else None, cn if cn is not None else None,
                            _ip(ct) if ct is not None else None)
    if rc != 0:
        raise AnukError(rc, lib.anuk_last_error(None).decode())
