"""landschaft_amd: the circuit-connectivity hot path of landschaft on AMD MI355X (gfx950).

The package mirrors ``landschaft.circuit`` (Circuitscape's advanced and pairwise
modes, Omniscape's moving windows) on a HIP library reached through a C-ABI
(``include/landschaft_circuit.h``); see ``landschaft_amd.circuit``.
"""

from . import circuit
from .circuit import omniscape

__all__ = ["circuit", "omniscape"]
