"""robast_amd: MI355X-native drop-in for ROBAST's non-sequential tracer
(AOpticsManager::TraceNonSequential), built on librbgpu (HIP kernels for
gfx950 behind the C ABI in include/rbgpu.h)."""
from . import _abi
from .geometry import (ABorderSurfaceCondition, AFocalSurface, AGeoAsphericDisk, AGeoWinstonCone2D,
                       AGeoWinstonConePoly, ALens, AMirror,
                       AObscuration, AOpticalComponent, TGeoBBox,
                       TGeoCombiTrans, TGeoCompositeShape, TGeoHMatrix, TGeoParaboloid, TGeoPcon, TGeoPgon,
                       TGeoRotation, TGeoSphere, TGeoTranslation, TGeoTube, TGeoVolume, cm, deg, m, mm, nm,
                       rad, um)
from .manager import AOpticsManager
from .materials import ARefractiveIndex, ASchottFormula, ASellmeierFormula, TGraph
from .rays import ARay, ARayArray, ARayShooter
