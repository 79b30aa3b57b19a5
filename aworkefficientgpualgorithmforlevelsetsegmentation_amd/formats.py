"""The reference application's volume and parameter files (Python mirror of
``include/lsseg/formats.hpp``).

  - raw volumes: VolumeLoader::LoadVolumeRaw (src/VolumeLoader.cpp:275-400);
  - parameter files (*.par): the "Segmenter" system of the ParameterManager XML
    (bin/parameters, bin/defaults; Segmenter.cpp:34-40);
  - default parameters by the volume's maximum (Engine.cpp:196-216);
  - projects (*.gproj): MainWindow::openProject (src/widgets/MainWindow.cpp:692-725).

Reference paths are relative to code/gpuseg/code/ (sources) and code/gpuseg/bin/
(data) of the reference repository.  DICOM series are out of scope.
"""
from __future__ import annotations

import os
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field

import numpy as np

from .segmenter import Parameters

_RAW_TYPES = {(1, False): np.uint8, (1, True): np.int8, (2, False): np.uint16, (2, True): np.int16,
              (4, False): np.uint32, (4, True): np.int32}


@dataclass
class VolumeFileDesc:
    """VolumeFileDesc (code/gpuseg/code/VolumeFileDesc.hpp)."""
    fileName: str
    numVoxelsX: int
    numVoxelsY: int
    numVoxelsZ: int
    numBytesPerVoxel: int = 1
    isSigned: bool = False
    zAnisotropy: float = 1.0
    upDirection: tuple = (0, 0, 1)


@dataclass
class Volume:
    data: np.ndarray          # (z, y, x)
    minValue: float
    maxValue: float
    zAnisotropy: float = 1.0


@dataclass
class Project:
    volume: VolumeFileDesc
    parametersFile: str = ""  # empty: defaults by type
    extra: dict = field(default_factory=dict)


def load_volume_raw(desc: VolumeFileDesc) -> Volume:
    """nx*ny*nz voxels, x fastest, with the minimum and maximum as floats."""
    key = (int(desc.numBytesPerVoxel), bool(desc.isSigned))
    if key not in _RAW_TYPES:
        raise ValueError("raw volumes have 1, 2 or 4 bytes per voxel")
    dt = np.dtype(_RAW_TYPES[key])
    count = desc.numVoxelsX * desc.numVoxelsY * desc.numVoxelsZ
    data = np.fromfile(desc.fileName, dtype=dt, count=count)
    if data.size != count:
        raise ValueError(f"cannot read {count * dt.itemsize} bytes from {desc.fileName}")
    vals = data.astype(np.float32)
    # VolumeLoader.cpp:305-384 starts from 65535 / -1
    lo = min(65535.0, float(vals.min()))
    hi = max(-1.0, float(vals.max()))
    return Volume(data.reshape(desc.numVoxelsZ, desc.numVoxelsY, desc.numVoxelsX), lo, hi, desc.zAnisotropy)


def load_parameters(par_file: str, base: Parameters | None = None) -> Parameters:
    """The Segmenter system of a parameter file; unnamed parameters keep `base`'s values."""
    p = Parameters(**vars(base)) if base is not None else Parameters()
    root = ET.parse(par_file).getroot()
    system = next((s for s in root.iter("System") if s.get("name") == "Segmenter"), None)
    if system is None:
        raise ValueError(f"{par_file}: no Segmenter system")
    for prm in system.iter("Parameter"):
        v = prm.find("ValueFloat")
        if v is None:
            continue
        value = float(np.float32(float(v.get("value"))))
        name = prm.get("name")
        if name == "numSubSteps":
            p.numSubSteps = int(value)
        elif name in ("growLevel", "growWindow", "timeStep", "curvatureInfluence", "convergenceThreshold"):
            setattr(p, name, value)
    return p


def default_parameters(max_value: float) -> Parameters:
    """Engine.cpp:196-216 with the values of bin/defaults/*.par."""
    common = dict(timeStep=0.2, curvatureInfluence=0.3, numSubSteps=10, convergenceThreshold=0.0)
    if max_value <= 255.0:
        return Parameters(growLevel=150.0, growWindow=50.0, **common)
    if max_value <= 4095.0:
        return Parameters(growLevel=790.0, growWindow=350.0, **common)
    if max_value <= 65535.0:
        return Parameters(growLevel=15000.0, growWindow=5000.0, **common)
    raise ValueError("no default parameters for a maximum above 65535")


def load_project(gproj_file: str, root: str | None = None) -> Project:
    """A .gproj; dataFile and parametersFile are relative to the application's
    working directory, by default the parent of the project file's directory."""
    el = ET.parse(gproj_file).getroot()
    proj = el if el.tag == "GPUSegProject" else el.find("GPUSegProject")
    if proj is None:
        raise ValueError(f"{gproj_file}: no GPUSegProject element")
    a = proj.attrib
    base = root if root is not None else os.path.dirname(os.path.dirname(os.path.abspath(gproj_file)))

    def resolve(rel):
        rel = rel.replace("\\", "/")
        return rel if not rel or os.path.isabs(rel) else os.path.join(base, rel)

    vol = VolumeFileDesc(
        fileName=resolve(a.get("dataFile", "")),
        numVoxelsX=int(a.get("numVoxelsX", 0)), numVoxelsY=int(a.get("numVoxelsY", 0)),
        numVoxelsZ=int(a.get("numVoxelsZ", 0)),
        numBytesPerVoxel=int(a.get("numBytesPerVoxel", 1)), isSigned=int(a.get("isSigned", 0)) != 0,
        zAnisotropy=float(np.float32(float(a.get("zAnisotropy", 1.0)))),
        upDirection=(int(a.get("upDirectionX", 0)), int(a.get("upDirectionY", 0)), int(a.get("upDirectionZ", 1))))
    return Project(vol, resolve(a.get("parametersFile", "")))


def project_parameters(project: Project, volume: Volume) -> Parameters:
    """Engine.cpp:196-226: the project's parameter file, else the default for the type."""
    if project.parametersFile:
        return load_parameters(project.parametersFile)
    return default_parameters(volume.maxValue)


def write_project(directory: str, name: str, volume_zyx: np.ndarray, parameters: Parameters,
                  z_anisotropy: float = 1.0) -> str:
    """Writes <directory>/data/<name>.raw, parameters/<name>.par and
    projects/<name>.gproj in the reference's layout; returns the project path."""
    vol = np.ascontiguousarray(volume_zyx)
    nz, ny, nx = vol.shape
    key = next((k for k, v in _RAW_TYPES.items() if np.dtype(v) == vol.dtype), None)
    if key is None:
        raise ValueError(f"unsupported raw type {vol.dtype}")
    for sub in ("data", "parameters", "projects"):
        os.makedirs(os.path.join(directory, sub), exist_ok=True)
    vol.tofile(os.path.join(directory, "data", f"{name}.raw"))
    p = parameters
    entries = [("growLevel", p.growLevel), ("growWindow", p.growWindow), ("timeStep", p.timeStep),
               ("curvatureInfluence", p.curvatureInfluence), ("numSubSteps", p.numSubSteps),
               ("convergenceThreshold", p.convergenceThreshold)]
    params_xml = "".join(f'                    <Parameter type="ParameterType_Float" name="{k}" >\n'
                         f'                        <ValueFloat value="{v!r}" />\n'
                         f'                    </Parameter>\n' for k, v in entries)
    with open(os.path.join(directory, "parameters", f"{name}.par"), "w") as f:
        f.write("<Root>\n    <ParameterManager>\n        <Systems>\n            <System name=\"Segmenter\" >\n"
                f"                <Parameters>\n{params_xml}                </Parameters>\n"
                "            </System>\n        </Systems>\n    </ParameterManager>\n</Root>\n")
    path = os.path.join(directory, "projects", f"{name}.gproj")
    with open(path, "w") as f:
        f.write(f'<Root>\n    <GPUSegProject isSigned="{int(key[1])}" numBytesPerVoxel="{key[0]}" upDirectionX="0" '
                f'parametersFile="parameters/{name}.par" upDirectionY="0" upDirectionZ="1" numVoxelsX="{nx}" '
                f'dataFile="data/{name}.raw" numVoxelsY="{ny}" zAnisotropy="{z_anisotropy}" numVoxelsZ="{nz}" />\n'
                '</Root>\n')
    return path
