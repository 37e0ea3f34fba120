"""fast_pauli on MI355X: Pauli strings, operators and summed operators applied on gfx950.

Same import surface as the reference package (fast_pauli/__init__.py:18-24):
Pauli, PauliString, PauliOp, SummedPauliOp and helpers, backed by the C++ host
API over the HIP C ABI (libfastpauli_hip.so).  There is no CPU fallback: the
classes come from the compiled extension ``_fast_pauli`` and accessing them
raises ImportError when it has not been built.  Submodules that talk to the C
ABI directly (``_abi``, ``batch``, ``serving``) import without it.
"""

_PUBLIC = ("Pauli", "PauliOp", "PauliString", "SummedPauliOp", "helpers")
__all__ = list(_PUBLIC)


def __getattr__(name):
    if name in _PUBLIC:
        from . import _fast_pauli  # raises ImportError when the extension is missing

        value = getattr(_fast_pauli, name)
        globals()[name] = value
        return value
    raise AttributeError(f"module 'fast_pauli' has no attribute {name!r}")
