from .pauli_basis import PauliBasis
from .library import general, gell_mann

__all__ = ['PauliBasis', 'general', 'gell_mann']
