"""Circuit-theory connectivity on rasters (Circuitscape, Omniscape) on the GPU.

Mirrors ``landschaft.circuit`` (the reference's src/landschaft/circuit/__init__.py).
"""

from ._backend import core, settings
from .advanced import AdvancedResult, advanced
from .circuitscape import (CircuitscapeResult, FocalPoints, IncludedPairs, circuitscape_advanced, circuitscape_ini,
                           circuitscape_one_to_all, circuitscape_pairwise, habitat_conductance)
from .ini import omniscape_ini, parse_ini
from .network import NetworkResult, circuitscape_network_advanced, circuitscape_network_pairwise, read_graph
from .omniscape import omniscape, read_reclass_table, reclassify, source_from_resistance

__all__ = ["AdvancedResult", "CircuitscapeResult", "FocalPoints", "IncludedPairs", "NetworkResult", "advanced",
           "circuitscape_advanced", "circuitscape_ini", "circuitscape_network_advanced",
           "circuitscape_network_pairwise", "circuitscape_one_to_all", "circuitscape_pairwise", "core",
           "habitat_conductance", "omniscape", "omniscape_ini", "parse_ini", "read_graph", "read_reclass_table",
           "reclassify", "settings", "source_from_resistance"]
