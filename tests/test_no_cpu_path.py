"""There is no CPU path: every compute entry point refuses host tensors with a
RuntimeError before any arithmetic, non-floating input raises the reference's
TypeError (core/dtype.py:581-590), and a missing libpfm.so raises instead of
falling back.  No GPU is needed: the checks run before any device call."""

import pytest
import torch

from pyopenfoam_amd import _lib
from pyopenfoam_amd.core.ldu_matrix import LduMatrix
from pyopenfoam_amd.solvers import GAMGSolver, PBiCGSTABSolver, PCGSolver
from pyopenfoam_amd.solvers.diagonal import DiagonalSolver
from pyopenfoam_amd.solvers.pbicg import PBiCGSolver
from pyopenfoam_amd.solvers.preconditioners import DICPreconditioner, DILUPreconditioner
from pyopenfoam_amd.solvers.smooth_solver import SmoothSolver

F64 = torch.float64


def _host_matrix():
    m = LduMatrix(3, torch.tensor([0, 1]), torch.tensor([1, 2]), device="cpu")
    m.diag = torch.tensor([4.0, 4.0, 4.0], dtype=F64)
    m.lower = torch.tensor([-1.0, -1.0], dtype=F64)
    m.upper = torch.tensor([-1.0, -2.0], dtype=F64)
    return m


@pytest.mark.parametrize("call", [
    "Ax", "Ax_T", "PCG", "PCG-mixed", "PBiCGStab", "PBiCG", "GAMG", "smooth", "diagonal",
    "DIC", "DILU"])
def test_host_tensors_are_refused(call):
    m = _host_matrix()
    b = torch.tensor([1.0, 2.0, 3.0], dtype=F64)
    x0 = torch.zeros(3, dtype=F64)
    run = {
        "Ax": lambda: m.Ax(b),
        "Ax_T": lambda: m.Ax_T(b),
        "PCG": lambda: PCGSolver()(m, b, x0),
        "PCG-mixed": lambda: PCGSolver(precision="mixed")(m, b, x0),
        "PBiCGStab": lambda: PBiCGSTABSolver()(m, b, x0),
        "PBiCG": lambda: PBiCGSolver()(m, b, x0),
        "GAMG": lambda: GAMGSolver()(m, b, x0),
        "smooth": lambda: SmoothSolver()(m, b, x0),
        "diagonal": lambda: DiagonalSolver()(m, b, x0),
        "DIC": lambda: DICPreconditioner(m),
        "DILU": lambda: DILUPreconditioner(m),
    }[call]
    with pytest.raises(RuntimeError, match="GPU only"):
        run()
    assert torch.equal(x0, torch.zeros(3, dtype=F64))  # nothing was written


def test_non_floating_input_is_a_type_error():
    m = _host_matrix()
    with pytest.raises(TypeError, match="floating-point"):
        m.Ax(torch.ones(3, dtype=torch.int64))


def test_missing_library_raises(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "library_path", tmp_path / "libpfm.so")
    with pytest.raises(_lib.PfmError, match="missing"):
        _lib.lib()
