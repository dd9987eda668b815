"""MI355X-native RoadRunner transit model (PyTransit's RoadRunnerModel / TSModel path).

Public API mirrors ``pytransit`` (``pytransit/__init__.py:64-66``):
``RoadRunnerModel`` (alias ``RRModel``), ``TransmissionSpectroscopyModel`` (alias
``TSModel``) and the ``LDModel`` plug-in protocol.
"""
from .ldmodel import LDModel, TabulatedLDModel
from .rrmodel import RoadRunnerModel, RRModel
from .tsmodel import TransmissionSpectroscopyModel, TSModel

__version__ = '0.1.0'

__all__ = ['RoadRunnerModel', 'RRModel', 'TransmissionSpectroscopyModel', 'TSModel', 'LDModel', 'TabulatedLDModel']
