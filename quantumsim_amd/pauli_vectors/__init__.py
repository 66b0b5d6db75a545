"""Pauli-vector backends: ``PauliVectorHIP`` (gfx950 kernels) is the only
one, and ``Default`` (quantumsim/pauli_vectors/__init__.py:5-14) names it."""
from .pauli_vector import PauliVectorBase
from .hip import PauliVectorHIP

Default = PauliVectorHIP

__all__ = ['Default', 'PauliVectorBase', 'PauliVectorHIP']
