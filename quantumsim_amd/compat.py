"""Drop-in hooks for an installed quantumsim.

``register_with_quantumsim()`` makes ``PauliVectorHIP`` a virtual subclass
of quantumsim's ``PauliVectorBase`` (``State`` checks
``issubclass(pauli_vector_class, PauliVectorBase)``,
quantumsim/states/state.py:31-35); ``install_as_default()`` additionally
points ``quantumsim.pauli_vectors.Default`` and ``quantumsim.PauliVector``
at it (quantumsim/pauli_vectors/__init__.py:5-14, quantumsim/__init__.py:9).
Both raise ImportError when quantumsim is not importable.
"""
from .pauli_vectors.hip import PauliVectorHIP

__all__ = ['register_with_quantumsim', 'install_as_default']


def register_with_quantumsim():
    from quantumsim.pauli_vectors.pauli_vector import PauliVectorBase as _Base
    _Base.register(PauliVectorHIP)
    return PauliVectorHIP


def install_as_default():
    register_with_quantumsim()
    import quantumsim
    import quantumsim.pauli_vectors as pvs
    pvs.Default = PauliVectorHIP
    quantumsim.PauliVector = PauliVectorHIP
    return PauliVectorHIP
