"""Error types, mirroring the reference's SolveError and its sources.

fasterhenry/src/solve.rs:567-637 (SolveError), inductance/mod.rs:93-115
(KernelError), mesh.rs:99-148 (MeshError), geometry.rs (GeometryError).
Each native status code of include/fasterhenry_gpu.h maps to one class.
"""
from __future__ import annotations


class SolveError(Exception):
    """Why an impedance extraction failed (solve.rs:567-637)."""

    def __init__(self, message: str, **detail) -> None:
        super().__init__(message)
        self.detail = detail
        for key, value in detail.items():
            setattr(self, key, value)


class GeometryError(SolveError):
    """A node or segment of the geometry is invalid (geometry.rs GeometryError)."""


class DiscretizeError(SolveError):
    """A segment could not be cut into filaments (SolveError::Discretize)."""


class KernelError(SolveError):
    """A partial inductance could not be evaluated (inductance/mod.rs:93-115)."""


class MeshError(SolveError):
    """Ports or per-segment counts do not fit the geometry (mesh.rs:99-148)."""


class InvalidGrading(SolveError):
    """A grading parameter is out of range (SolveError::InvalidGrading)."""


class InvalidFrequency(SolveError):
    """A frequency is negative or not finite (SolveError::InvalidFrequency)."""


class Singular(SolveError):
    """Zero pivot or non-finite impedance at a frequency (SolveError::Singular)."""


class NotConverged(SolveError):
    """An iterative refinement or Krylov solve missed its tolerance."""


class HipError(SolveError):
    """The HIP runtime reported an error."""


class InvalidArgument(SolveError):
    """A null pointer, a size out of range or an unsupported option."""


class PfftError(SolveError):
    """The pFFT operator cannot be built (pfft/mod.rs:236-251): code 1 is
    InvalidParameter, code 2 GridTooLarge (index0 = padded points, index1 =
    the limit)."""


class InvalidExcitation(SolveError):
    """The excitation does not have one finite entry per port
    (SolveError::InvalidExcitation): index0 = ports, index1 = entries given."""


class DeckError(SolveError):
    """A deck line the reader rejects (fasterhenry-cli/src/inp.rs:682-702
    ParseError): index0 = the line (0 for the deck as a whole)."""
