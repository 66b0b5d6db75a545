"""quantumsim_amd: an MI355X-native Pauli-vector backend with quantumsim's
PauliVector API (quantumsim/__init__.py:9 exposes ``PauliVector``)."""
from . import bases, algebra
from .pauli_vectors import PauliVectorBase, PauliVectorHIP

PauliVector = PauliVectorHIP

__all__ = ['bases', 'algebra', 'PauliVector', 'PauliVectorBase', 'PauliVectorHIP']
