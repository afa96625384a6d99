"""Shared backend enums and parsers.

Mirrors ``pygrgl_spmv/backends/types.py:8-113`` of the reference. The reference
uses ``enum.StrEnum`` (Python >= 3.11); these are ``(str, Enum)`` so they behave
the same (string equality, ``.value``) on Python 3.10.
"""

from __future__ import annotations

from enum import Enum


class _StrEnum(str, Enum):
    def __str__(self) -> str:  # StrEnum formats as its value
        return str(self.value)


class Direction(_StrEnum):
    """Traversal direction: UP = samples -> mutations, DOWN = mutations -> samples."""

    UP = "up"
    DOWN = "down"


class InitMode(_StrEnum):
    """Node initialization modes for a single traversal."""

    NONE = "none"
    XTX = "xtx"
    VECTOR = "vector"
    MATRIX = "matrix"


class StoredMatrix(_StrEnum):
    """Whether a backend stores ``A`` or ``A.T``."""

    N = "N"
    T = "T"


class SparseFormat(_StrEnum):
    """Sparse storage formats shared by backends."""

    CSR = "CSR"
    CSC = "CSC"
    COO = "COO"


_TRANSPOSE_FMT_MAP = {
    SparseFormat.CSR: SparseFormat.CSC,
    SparseFormat.CSC: SparseFormat.CSR,
    SparseFormat.COO: SparseFormat.COO,
}


def parse_direction(direction: str | Direction) -> Direction:
    """Parse a traversal direction string/enum (reference ``types.py:46-57``)."""
    if isinstance(direction, Direction):
        return direction
    key = str(direction).lower()
    if key == "up":
        return Direction.UP
    if key == "down":
        return Direction.DOWN
    raise ValueError(f"Unknown direction {direction!r}; expected up or down")


def parse_init_mode(init_mode: str | InitMode) -> InitMode:
    """Parse an init mode string/enum (reference ``types.py:60-78``)."""
    if isinstance(init_mode, InitMode):
        return init_mode
    key = str(init_mode).lower()
    for mode in InitMode:
        if mode.value == key:
            return mode
    raise ValueError(
        "Unknown init_mode "
        f"{init_mode!r}; expected one of (none, xtx, vector, matrix)"
    )


def parse_store(value: str | StoredMatrix) -> StoredMatrix:
    """Parse whether a backend stores ``A`` or ``A.T`` (reference ``types.py:81-92``)."""
    if isinstance(value, StoredMatrix):
        return value
    key = str(value).strip().upper()
    if key == "N":
        return StoredMatrix.N
    if key == "T":
        return StoredMatrix.T
    raise ValueError(f"Unknown store {value!r}; expected N or T")


def parse_sparse_format(value: str | SparseFormat) -> SparseFormat:
    """Parse a sparse storage format (reference ``types.py:95-108``)."""
    if isinstance(value, SparseFormat):
        return value
    key = str(value).strip().upper()
    for fmt in SparseFormat:
        if fmt.value == key:
            return fmt
    raise ValueError(f"Unknown sparse format {value!r}; expected CSR, CSC, or COO")


def transpose_compatible_format(fmt: SparseFormat) -> SparseFormat:
    """Return the transpose-compatible sparse format."""
    return _TRANSPOSE_FMT_MAP[parse_sparse_format(fmt)]


__all__ = [
    "Direction",
    "InitMode",
    "StoredMatrix",
    "SparseFormat",
    "parse_direction",
    "parse_init_mode",
    "parse_sparse_format",
    "parse_store",
    "transpose_compatible_format",
]
