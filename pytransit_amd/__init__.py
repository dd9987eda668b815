"""MI355X-native RoadRunner transit model (PyTransit's RoadRunnerModel / TSModel path).

Public API mirrors ``pytransit`` (``pytransit/__init__.py:64-66``):
``RoadRunnerModel`` (alias ``RRModel``), ``TransmissionSpectroscopyModel`` (alias
``TSModel``), ``QuadraticModel`` (Mandel & Agol), the eclipse variants
``EclipseModel`` and ``EclipseSpectroscopyModel``, ``OblatePlanetModel`` and the
``LDModel`` plug-in protocol; ``RoadRunnerLPFModel`` takes the layouts of
PyTransit's LPFs.
"""
from .eclipse import EclipseModel, EclipseSpectroscopyModel
from .ldmodel import LDModel, LDTkLDModel, TabulatedLDModel
from .lpf import RoadRunnerLPFModel
from .opmodel import OblatePlanetModel
from .qmodel import QuadraticModel
from .rrmodel import RoadRunnerModel, RRModel
from .tsmodel import TransmissionSpectroscopyModel, TSModel

__version__ = '0.1.0'

__all__ = ['RoadRunnerModel', 'RRModel', 'TransmissionSpectroscopyModel', 'TSModel', 'QuadraticModel',
           'RoadRunnerLPFModel', 'LDModel', 'TabulatedLDModel', 'LDTkLDModel', 'EclipseModel',
           'EclipseSpectroscopyModel', 'OblatePlanetModel']
