"""Circuit-theory connectivity on rasters (Circuitscape, Omniscape) on the GPU.

Mirrors ``landschaft.circuit`` (the reference's src/landschaft/circuit/__init__.py).
"""

from ._backend import core, settings
from .omniscape import omniscape, read_reclass_table, reclassify, source_from_resistance

__all__ = ["core", "omniscape", "read_reclass_table", "reclassify", "settings", "source_from_resistance"]
