"""GAMS scalar-format models (MINLPLib's GAMS Convert output) read into engine
problems, and the solver's RESULT line.

Mirrors the reference frontend's interface (include/cuminlp/gams.hpp:1-136):
``parse`` / ``parse_file`` return a ``ParsedModel`` whose ``problem`` is ready
for ``engine.Tape`` / ``engine.solve``; failures raise ``ParseError`` with the
line and ``kind`` the reference reports. Parsing runs in the engine library
(host C++), so no GPU is needed.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

from . import capi as _c
from .engine import EngineProblem

# gams::ErrorKind (reference gams.hpp:59-70)
SYNTAX, NON_SCALAR, DISCRETE, UNSUPPORTED_FUNCTION, UNREPRESENTABLE, IO = range(6)
KIND_NAMES = ["syntax", "non-scalar", "discrete", "unsupported-function", "unrepresentable", "io"]

# gams::Sense
MINIMISE, MAXIMISE = range(2)


class ParseError(Exception):
    """gams::ParseError: ``line`` (0 = the model as a whole) and ``kind``."""

    def __init__(self, message: str, line: int, kind: int):
        super().__init__(message)
        self.line = line
        self.kind = kind


@dataclass
class ParseOptions:
    """gams::ParseOptions (reference gams.hpp:21-50)."""

    default_bound: float = 1e6
    fold_fixed_to_const: bool = True
    reject_discrete: bool = False

    def _c(self) -> _c.GamsOptions:
        o = _c.GamsOptions()
        o.default_bound = float(self.default_bound)
        o.fold_fixed_to_const = int(bool(self.fold_fixed_to_const))
        o.reject_discrete = int(bool(self.reject_discrete))
        return o


@dataclass
class Diagnostic:
    line: int
    message: str


@dataclass
class ParsedModel:
    """gams::ParsedModel (reference gams.hpp:85-116). Under MAXIMISE the
    problem's objective is already negated."""

    problem: EngineProblem
    sense: int = MINIMISE
    var_names: list = field(default_factory=list)
    constraint_names: list = field(default_factory=list)
    initial_point: np.ndarray = field(default_factory=lambda: np.zeros(0))
    objvar_kept: bool = False
    warnings: list = field(default_factory=list)


def _collect(problem: C.c_void_p, meta: C.c_void_p) -> ParsedModel:
    lib = _c.lib()
    try:
        info = _c.GamsInfo()
        _c.check(lib.cmi_gams_info_get(meta, C.byref(info)))
        warnings = []
        for i in range(info.n_warnings):
            line = C.c_int(0)
            msg = lib.cmi_gams_warning(meta, i, C.byref(line))
            warnings.append(Diagnostic(int(line.value), msg.decode()))
        point = np.zeros(0)
        if info.has_initial_point:
            point = np.zeros(info.n_vars)
            _c.check(lib.cmi_gams_initial_point(meta, _c.ptr(point), info.n_vars))
        return ParsedModel(
            problem=EngineProblem(problem),
            sense=int(info.sense),
            var_names=[lib.cmi_gams_var_name(meta, i).decode() for i in range(info.n_vars)],
            constraint_names=[lib.cmi_gams_constraint_name(meta, i).decode() for i in range(info.n_constraints)],
            initial_point=point,
            objvar_kept=bool(info.objvar_kept),
            warnings=warnings,
        )
    finally:
        lib.cmi_gams_destroy(meta)


def _finish(status: int, err: _c.GamsError, problem: C.c_void_p, meta: C.c_void_p) -> ParsedModel:
    if status == _c.PARSE_ERROR:
        raise ParseError(_c.lib().cmi_last_error().decode(), int(err.line), int(err.kind))
    _c.check(status)
    return _collect(problem, meta)


def parse(source: str | bytes, options: ParseOptions | None = None) -> ParsedModel:
    """gams::parse (reference gams.hpp:118-121, frontend.cpp:1592-1667)."""
    text = source.encode() if isinstance(source, str) else bytes(source)
    opts = (options or ParseOptions())._c()
    problem, meta, err = C.c_void_p(), C.c_void_p(), _c.GamsError()
    status = _c.lib().cmi_gams_parse(text, len(text), C.byref(opts), C.byref(problem), C.byref(meta),
                                     C.byref(err))
    return _finish(status, err, problem, meta)


def parse_file(path, options: ParseOptions | None = None) -> ParsedModel:
    """gams::parse_file (reference gams.hpp:123-126, frontend.cpp:1669-1681)."""
    opts = (options or ParseOptions())._c()
    problem, meta, err = C.c_void_p(), C.c_void_p(), _c.GamsError()
    status = _c.lib().cmi_gams_parse_file(os.fsencode(path), C.byref(opts), C.byref(problem), C.byref(meta),
                                          C.byref(err))
    return _finish(status, err, problem, meta)


def result_line(sense: int, lower: float, upper: float) -> str:
    """The RESULT line for a solve's minimisation bounds (reference
    source/gams/solve.cu:836-867)."""
    n = C.c_size_t(0)
    _c.check(_c.lib().cmi_gams_result_line(int(sense), float(lower), float(upper), None, 0, C.byref(n)))
    buf = C.create_string_buffer(n.value + 1)
    _c.check(_c.lib().cmi_gams_result_line(int(sense), float(lower), float(upper), buf, len(buf), None))
    return buf.value.decode()


__all__ = ["ParseError", "ParseOptions", "Diagnostic", "ParsedModel", "parse", "parse_file", "result_line",
           "MINIMISE", "MAXIMISE", "SYNTAX", "NON_SCALAR", "DISCRETE", "UNSUPPORTED_FUNCTION", "UNREPRESENTABLE",
           "IO", "KIND_NAMES"]
