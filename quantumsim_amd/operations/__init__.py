from .operation import PTMOp, from_kraus, from_ptm
from . import qubits

__all__ = ['PTMOp', 'from_kraus', 'from_ptm', 'qubits']
