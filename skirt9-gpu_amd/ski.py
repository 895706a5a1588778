"""Ski-file front end and FITS output for the engine.

`load(path)` reads the subset of SKIRT's ski configuration (SMILE XML) that the engine runs:
an oligochromatic `MonteCarloSimulation` with one `PointSource` (`BlackBodySED` with a
`SpecificLuminosityNormalization`), one `GeometricMedium` (`ExpDiskGeometry`, optionally
inside a `SpiralStructureGeometryDecorator`) with a `MeanListDustMix` and an
`OpticalDepthMaterialNormalization`, a `CartesianSpatialGrid` (`LinMesh` axes) or a
`PolicyTreeSpatialGrid` with a `DensityTreePolicy`, a `VoronoiMeshSpatialGrid` or a
`TetraMeshSpatialGrid` with sites / vertices from the dust density or a column file -- or one
`VoronoiMeshMedium` (sites and number or mass densities imported from a text file, with a
`MeanListDustMix`) on a `VoronoiMeshSpatialGrid` that employs its mesh -- `PhotonPacketOptions`,
`FrameInstrument`s, `SEDInstrument`s and `PerspectiveInstrument`s.  Anything else raises `UnsupportedSki`, so a caller
can fall back to the reference.  `run(sim, outdir, prefix)` sets the engine up, runs the
photon packets and writes `<prefix>_<instrument>_total.fits` per frame as
`FrameInstrument`/`FluxRecorder` do (`FluxRecorder.cpp:876-960`): surface brightness in
MJy/sr (ExtragalacticUnits with `fluxOutputStyle="Frequency"`), float32, one plane per
wavelength; `<prefix>_<probe>_J.dat` per `RadiationFieldProbe` (`PerCellForm`) when the
radiation field is stored (mean intensity J_nu per cell); and
`<prefix>_<instrument>_sed.dat` per SED instrument, the text columns of
`FluxRecorder.cpp:1093-1115` (wavelength in micron, total flux density F_nu in Jy =
recorded luminosity / (4 pi d^2) / (bin width) * lambda^2 / c * 1e26).

Calibration (FluxRecorder::calibrateAndWrite): value = recorded luminosity
/ (4 pi d^2) / omega / (bin width) * lambda^2 / c * 1e20, with omega = 4 atan(dx / 2d)
atan(dy / 2d) per pixel and the oligochromatic bin width 2 * 1e-3 * lambda_min
(OligoWavelengthGrid.cpp:20-25).  The recorded luminosity of a bin is its share of the
source luminosity, L_lambda(lambda) * (bin width), with L_lambda the black body scaled to
the normalization's specific luminosity (BlackBodySED, SpecificLuminosityNormalization).
"""

from __future__ import annotations

import math
import os
import time
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field
from typing import List

import numpy as np

# SKIRT's constants (Constants.hpp) and the units the supported items use (SkirtUnitDef.cpp)
C_LIGHT = 2.99792458e8
H_PLANCK = 6.62606957e-34
K_BOLTZ = 1.3806488e-23
PC = 3.08567758e16
MPC = 1e6 * PC
LSUN = 3.839e26
AU = 1.495978707e11

_UNITS = {
    "m": 1.0, "cm": 1e-2, "mm": 1e-3, "km": 1e3, "AU": AU, "pc": PC, "kpc": 1e3 * PC, "Mpc": 1e6 * PC,
    "micron": 1e-6, "nm": 1e-9, "Angstrom": 1e-10,
    "deg": math.pi / 180.0, "rad": 1.0, "arcsec": math.pi / 180.0 / 3600.0,
    "K": 1.0, "km/s": 1e3, "m/s": 1.0, "uG": 1e-10, "nT": 1e-9,
    "Lsun/micron": LSUN * 1e6, "W/m": 1.0, "W/micron": 1e6, "erg/s/Angstrom": 1e-7 * 1e10,
    "m2/kg": 1.0, "cm2/g": 0.1, "pc2/Msun": PC * PC / 1.9891e30,
}


class UnsupportedSki(ValueError):
    """The ski file asks for something the engine does not run."""


def quantity(text: str) -> float:
    """'10 pc' -> metres, '60 deg' -> radians, '1' -> 1.0 (SI values, as SKIRT stores them)."""
    parts = text.strip().split()
    if len(parts) == 1:
        return float(parts[0])
    if len(parts) != 2 or parts[1] not in _UNITS:
        raise UnsupportedSki(f"unsupported quantity {text!r}")
    return float(parts[0]) * _UNITS[parts[1]]


def quantities(text: str) -> List[float]:
    return [quantity(t) for t in text.split(",")]


@dataclass
class Frame:
    name: str
    distance: float
    inclination: float
    azimuth: float
    roll: float
    fovx: float
    fovy: float
    nx: int
    ny: int
    cx: float
    cy: float


@dataclass
class Sed:
    """SEDInstrument: distance, direction and aperture radius (0 = none)."""
    name: str
    distance: float
    inclination: float
    azimuth: float
    roll: float
    radius: float


@dataclass
class Perspective:
    """PerspectiveInstrument: viewport origin, crosshair, upwards direction, focal length,
    viewport width and pixels."""
    name: str
    view: tuple
    cross: tuple
    up: tuple
    focal: float
    width: float
    nx: int
    ny: int


@dataclass
class Simulation:
    num_packets: int
    seed: int
    wavelengths: np.ndarray          # m, sorted
    source_position: tuple
    temperature: float               # black body temperature (K)
    norm_wavelength: float           # m
    norm_specific_luminosity: float  # W/m
    spiral: bool
    scale_length: float
    scale_height: float
    spiral_params: dict
    dust_table: np.ndarray           # (n, 4): lambda (m), kappa_ext (m2/kg), albedo, g
    tau_v: float
    tau_wavelength: float
    grid: dict
    box: tuple
    frames: List[Frame] = field(default_factory=list)
    rf_probes: List[str] = field(default_factory=list)  # RadiationFieldProbe names (PerCellForm)
    force_scattering: bool = True
    min_weight_reduction: float = 1e4
    min_scatt_events: int = 0
    path_length_bias: float = 0.5
    num_density_samples: int = 100
    cell_volumes: object = None  # set by setup(): volume per cell (m3)
    # ParticleSource instead of the point source: xyzh (n, 4) m, radius (n,) m, temperature (n,) K,
    # kernel 0 uniform / 1 cubic spline
    particles: object = None
    # VoronoiMeshMedium instead of the GeometricMedium: sites (n, 3) m and number density (n,) m^-3
    imported: object = None


def _one(root, tag):
    found = root.findall(".//" + tag)
    if len(found) != 1:
        raise UnsupportedSki(f"expected exactly one {tag}, found {len(found)}")
    return found[0]


def load(path: str) -> Simulation:
    root = ET.parse(path).getroot()
    sim = _one(root, "MonteCarloSimulation")
    if sim.get("simulationMode") != "OligoExtinctionOnly":
        raise UnsupportedSki("only simulationMode OligoExtinctionOnly is supported")
    media = root.findall(".//media/*")
    if not media or media[0].tag not in ("GeometricMedium", "VoronoiMeshMedium"):
        raise UnsupportedSki("a GeometricMedium or a VoronoiMeshMedium is supported")
    sources = root.findall(".//sources/*")
    if len(media) != 1 or len(sources) != 1:
        raise UnsupportedSki("exactly one medium and one source are supported")
    instruments = root.findall(".//instruments/*")
    if any(i.tag not in ("FrameInstrument", "SEDInstrument", "PerspectiveInstrument") for i in instruments):
        raise UnsupportedSki("only FrameInstrument, SEDInstrument and PerspectiveInstrument instruments are supported")
    seed = int(_one(root, "Random").get("seed", "0"))
    ss = _one(root, "SourceSystem")
    lam = np.array(sorted(quantities(ss.get("wavelengths"))))
    particles = None
    if sources[0].tag == "ParticleSource":
        if float(ss.get("sourceBias", "0.5")) != 0.5:
            raise UnsupportedSki("a ParticleSource runs with sourceBias 0.5")
        particles = _particles(path, sources[0])
        pos, temperature, norm_wl, norm_l = (0.0, 0.0, 0.0), 0.0, 0.0, 0.0
    elif sources[0].tag == "PointSource":
        ps = sources[0]
        pos = tuple(quantity(ps.get(k, "0 pc")) for k in ("positionX", "positionY", "positionZ"))
        bb = _one(root, "BlackBodySED")
        norm = _one(root, "SpecificLuminosityNormalization")
        if norm.get("unitStyle", "wavelengthmonluminosity") != "wavelengthmonluminosity":
            raise UnsupportedSki("only wavelengthmonluminosity normalizations are supported")
        temperature, norm_wl = quantity(bb.get("temperature")), quantity(norm.get("wavelength"))
        norm_l = quantity(norm.get("specificLuminosity"))
    else:
        raise UnsupportedSki("only a PointSource or a ParticleSource is supported")
    opts = _one(root, "PhotonPacketOptions")
    imported = None
    if media[0].tag == "VoronoiMeshMedium":
        imported = _voronoi_medium(path, media[0])
        disk = ET.Element("ExpDiskGeometry", scaleLength="0 pc", scaleHeight="0 pc")
    else:
        disk = _one(root, "ExpDiskGeometry")
    for k in ("minRadius", "maxRadius", "maxZ"):
        if quantity(disk.get(k, "0 pc")) != 0:
            raise UnsupportedSki("truncated exponential disks are not supported")
    spiral_el = root.findall(".//SpiralStructureGeometryDecorator")
    spiral = dict()
    if spiral_el:
        s = spiral_el[0]
        spiral = dict(arms=int(s.get("numArms")), pitch=quantity(s.get("pitchAngle")),
                      radius_zero=quantity(s.get("radiusZeroPoint")), phase_zero=quantity(s.get("phaseZeroPoint")),
                      weight=float(s.get("perturbationWeight")), index=int(s.get("index")))
    mix = _one(root, "MeanListDustMix")
    tab = np.column_stack([quantities(mix.get("wavelengths")), quantities(mix.get("extinctionCoefficients")),
                           [float(x) for x in mix.get("albedos").split(",")],
                           [float(x) for x in mix.get("asymmetryParameters").split(",")]])
    if imported is None:
        mn = _one(root, "OpticalDepthMaterialNormalization")
        if mn.get("axis", "Z") != "Z":
            raise UnsupportedSki("only optical depth normalizations along Z are supported")
    else:
        mn = ET.Element("OpticalDepthMaterialNormalization", opticalDepth="0", wavelength="0.55 micron")
    samp = root.findall(".//SamplingOptions")
    nds = int(samp[0].get("numDensitySamples", "100")) if samp else 100
    grid_el = root.find(".//grid/*")
    if grid_el is None:
        raise UnsupportedSki("no spatial grid")
    box = tuple(quantity(grid_el.get(k)) for k in ("minX", "minY", "minZ", "maxX", "maxY", "maxZ"))
    if grid_el.tag == "CartesianSpatialGrid":
        ns = []
        for ax in ("meshX", "meshY", "meshZ"):
            m = grid_el.find(f"{ax}/LinMesh")
            if m is None:
                raise UnsupportedSki("only LinMesh axes are supported")
            ns.append(int(m.get("numBins")))
        grid = dict(kind="cartesian", n=ns)
    elif grid_el.tag == "PolicyTreeSpatialGrid":
        if grid_el.get("treeType", "OctTree") != "OctTree":
            raise UnsupportedSki("only octrees are supported")
        pol = grid_el.find(".//DensityTreePolicy")
        if pol is None:
            raise UnsupportedSki("only DensityTreePolicy is supported")
        if float(pol.get("maxDustOpticalDepth", "0")) or float(pol.get("maxDustDensityDispersion", "0")):
            raise UnsupportedSki("only maxDustFraction refinement is supported")
        grid = dict(kind="octree", min_level=int(pol.get("minLevel")), max_level=int(pol.get("maxLevel")),
                    max_fraction=float(pol.get("maxDustFraction")))
    elif imported is not None:
        # VoronoiMeshSpatialGrid employing the imported medium's mesh (VoronoiMeshSpatialGrid.cpp:133-141)
        if grid_el.tag != "VoronoiMeshSpatialGrid" or grid_el.get("policy") != "ImportedMesh":
            raise UnsupportedSki("a VoronoiMeshMedium runs on a VoronoiMeshSpatialGrid with policy ImportedMesh")
        if tuple(imported["box"]) != box:
            raise UnsupportedSki("the grid and the imported medium must have the same domain")
        grid = dict(kind="voronoi_imported")
    elif grid_el.tag in ("VoronoiMeshSpatialGrid", "TetraMeshSpatialGrid"):
        # vertices / sites sampled from the medium's density (policy DustDensity) or read from a
        # column file (policy File: VoronoiMeshSnapshot.cpp's file constructor,
        # TetraMeshSpatialGrid.cpp:303-312)
        policy = grid_el.get("policy", "DustDensity")
        if policy not in ("DustDensity", "File"):
            raise UnsupportedSki("only the DustDensity and File site policies are supported")
        if grid_el.get("relaxSites", "false") != "false" or grid_el.get("refine", "false") != "false":
            raise UnsupportedSki("site relaxation and tetra refinement are not supported")
        kind = "voronoi" if grid_el.tag == "VoronoiMeshSpatialGrid" else "tetra"
        if policy == "File":
            pts = _position_file(path, grid_el.get("filename"))
            grid = dict(kind=kind, n=len(pts), points=pts)
        else:
            key = "numSites" if kind == "voronoi" else "numSamples"
            grid = dict(kind=kind, n=int(grid_el.get(key)))
    else:
        raise UnsupportedSki(f"grid {grid_el.tag} is not supported by the ski front end")
    if not instruments:
        raise UnsupportedSki("no instruments")
    # radiation field: RadiationFieldOptions storeRadiationField and per-cell probes of it
    rfo = root.findall(".//RadiationFieldOptions")
    store_rf = bool(rfo) and rfo[0].get("storeRadiationField", "false") == "true"
    rf_probes = []
    for pr in root.findall(".//probes/*"):
        if pr.tag != "RadiationFieldProbe" or pr.find(".//PerCellForm") is None or not store_rf:
            raise UnsupportedSki("only RadiationFieldProbe with PerCellForm (and a stored radiation field) is supported")
        rf_probes.append(pr.get("probeName"))
    frames = []
    for fi in instruments:
        if fi.tag == "PerspectiveInstrument":
            q3 = lambda n: tuple(quantity(fi.get(n + c)) for c in "XYZ")  # noqa: E731
            frames.append(Perspective(fi.get("instrumentName"), q3("view"), q3("cross"), q3("up"),
                                      quantity(fi.get("focal")), quantity(fi.get("width")),
                                      int(fi.get("numPixelsX")), int(fi.get("numPixelsY"))))
            continue
        if fi.tag == "SEDInstrument":
            frames.append(Sed(fi.get("instrumentName"), quantity(fi.get("distance")), quantity(fi.get("inclination")),
                              quantity(fi.get("azimuth")), quantity(fi.get("roll", "0 deg")),
                              quantity(fi.get("radius", "0 pc"))))
            continue
        frames.append(Frame(fi.get("instrumentName"), quantity(fi.get("distance")), quantity(fi.get("inclination")),
                            quantity(fi.get("azimuth")), quantity(fi.get("roll")),
                            quantity(fi.get("fieldOfViewX")), quantity(fi.get("fieldOfViewY")),
                            int(fi.get("numPixelsX")), int(fi.get("numPixelsY")),
                            quantity(fi.get("centerX")), quantity(fi.get("centerY"))))
    return Simulation(
        num_packets=int(float(sim.get("numPackets"))), seed=seed, wavelengths=lam, source_position=pos,
        temperature=temperature, norm_wavelength=norm_wl, norm_specific_luminosity=norm_l, spiral=bool(spiral_el),
        scale_length=quantity(disk.get("scaleLength")), scale_height=quantity(disk.get("scaleHeight")),
        spiral_params=spiral, dust_table=tab, tau_v=float(mn.get("opticalDepth")),
        tau_wavelength=quantity(mn.get("wavelength")), grid=grid, box=box, frames=frames, rf_probes=rf_probes,
        force_scattering=opts.get("forceScattering", "true") == "true",
        min_weight_reduction=float(opts.get("minWeightReduction", "1e4")),
        min_scatt_events=int(opts.get("minScattEvents", "0")),
        path_length_bias=float(opts.get("pathLengthBias", "0.5")), num_density_samples=nds, particles=particles,
        imported=imported)


# VoronoiMeshMedium columns (VoronoiMeshMedium.hpp:20-40) and the units a header line may give;
# without one, x, y, z are in pc, a number density in 1/cm3, a mass density in Msun/pc3
_DENSITY_UNITS = {"1/m3": 1.0, "/m3": 1.0, "1/cm3": 1e6, "/cm3": 1e6, "1/pc3": PC ** -3,
                  "kg/m3": 1.0, "g/cm3": 1e3, "Msun/pc3": 1.9891e30 / PC ** 3, "Msun/AU3": 1.9891e30 / AU ** 3}


def _voronoi_medium(path, el):
    """VoronoiMeshMedium (MeshMedium.hpp, ImportedMedium.hpp): the text columns x, y, z and a
    number density (massType NumberDensity) or a mass density (MassDensity, converted with the
    dust mix's mass per hydrogen atom, MeanListDustMix.cpp:25), times massFraction, in the
    units of the '# column i: name (unit)' header lines or the defaults."""
    from .workloads import MU_DUST
    for k in ("importMetallicity", "importTemperature", "importVelocity", "importMagneticField",
              "importVariableMixParams"):
        if el.get(k, "false") != "false":
            raise UnsupportedSki(f"VoronoiMeshMedium {k} is not supported")
    if el.get("useColumns", ""):
        raise UnsupportedSki("VoronoiMeshMedium useColumns is not supported")
    mtype = el.get("massType", "MassDensity")
    if mtype not in ("NumberDensity", "MassDensity"):
        raise UnsupportedSki("VoronoiMeshMedium massType NumberDensity or MassDensity is supported")
    mix = el.findall("./materialMix/*")
    if not mix or mix[0].tag != "MeanListDustMix":
        raise UnsupportedSki("a VoronoiMeshMedium with a MeanListDustMix is supported")
    units = ["pc", "pc", "pc", "1/cm3" if mtype == "NumberDensity" else "Msun/pc3"]
    fname = os.path.join(os.path.dirname(os.path.abspath(path)), el.get("filename"))
    rows = []
    with open(fname) as f:
        for line in f:
            t = line.strip()
            if not t:
                continue
            if t.startswith("#"):
                low = t[1:].strip().lower()
                if low.startswith("column") and "(" in t and ":" in t:
                    i = int(low.split(":")[0].split()[1]) - 1
                    if 0 <= i < len(units):
                        units[i] = t[t.rindex("(") + 1:t.rindex(")")].strip()
                continue
            rows.append([float(v) for v in t.split()[:4]])
    a = np.array(rows, dtype=np.float64).reshape(-1, 4)
    if units[3] not in _DENSITY_UNITS:
        raise UnsupportedSki(f"density unit {units[3]} is not supported")
    sites = a[:, :3] * np.array([quantity("1 " + u) for u in units[:3]])
    dens = a[:, 3] * _DENSITY_UNITS[units[3]] * float(el.get("massFraction", "1"))
    if mtype == "MassDensity":
        dens = dens / MU_DUST
    box = tuple(quantity(el.get(k)) for k in ("minX", "minY", "minZ", "maxX", "maxY", "maxZ"))
    return dict(sites=sites, density=dens, box=box)


def _position_file(path, name):
    """Mesh vertices or sites from a text column file (policy File): the columns position x, y,
    z in the units of the '# column i: name (unit)' header lines, or pc (TextInFile's
    defaults for these columns)."""
    units = ["pc", "pc", "pc"]
    rows = []
    with open(os.path.join(os.path.dirname(os.path.abspath(path)), name)) as f:
        for line in f:
            t = line.strip()
            if not t:
                continue
            if t.startswith("#"):
                low = t[1:].strip().lower()
                if low.startswith("column") and "(" in t and ":" in t:
                    i = int(low.split(":")[0].split()[1]) - 1
                    if 0 <= i < 3:
                        units[i] = t[t.rindex("(") + 1:t.rindex(")")].strip()
                continue
            rows.append([float(v) for v in t.split()[:3]])
    a = np.array(rows, dtype=np.float64).reshape(-1, 3)
    return a * np.array([quantity("1 " + u) for u in units])


_PARTICLE_COLUMNS = (("position x", "pc"), ("position y", "pc"), ("position z", "pc"), ("size h", "pc"),
                     ("radius", "km"), ("temperature", "K"))


def _particles(path, el):
    """ParticleSource with BlackBodySEDFamily (ParticleSource.hpp, ImportedSource.hpp,
    BlackBodySEDFamily.cpp:11-14): the text columns x, y, z, h, radius, temperature in the
    units their '# column i: name (unit)' header lines give, or the defaults pc, pc, pc, pc,
    km, K (Snapshot.cpp:74-104)."""
    for k in ("importVelocity", "importVelocityDispersion", "importCurrentMass", "importBias"):
        if el.get(k, "false") != "false":
            raise UnsupportedSki(f"ParticleSource {k} is not supported")
    if el.get("useColumns", ""):
        raise UnsupportedSki("ParticleSource useColumns is not supported")
    kern = el.findall("./smoothingKernel/*")
    kname = kern[0].tag if kern else "CubicSplineSmoothingKernel"
    kernels = {"UniformSmoothingKernel": 0, "CubicSplineSmoothingKernel": 1}
    if kname not in kernels:
        raise UnsupportedSki(f"smoothing kernel {kname} is not supported")
    fam = el.findall("./sedFamily/*")
    if fam and fam[0].tag != "BlackBodySEDFamily":
        raise UnsupportedSki("only the BlackBodySEDFamily is supported")
    fname = os.path.join(os.path.dirname(os.path.abspath(path)), el.get("filename"))
    units = [u for _, u in _PARTICLE_COLUMNS]
    rows = []
    with open(fname) as f:
        for line in f:
            t = line.strip()
            if not t:
                continue
            if t.startswith("#"):
                low = t[1:].strip().lower()
                if low.startswith("column") and "(" in t and ":" in t:
                    i = int(low.split(":")[0].split()[1]) - 1
                    if 0 <= i < len(units):
                        units[i] = t[t.rindex("(") + 1:t.rindex(")")].strip()
                continue
            rows.append([float(v) for v in t.split()[:6]])
    a = np.array(rows, dtype=np.float64).reshape(-1, 6)
    fac = [1.0 if u == "K" else quantity("1 " + u) for u in units]
    a = a * np.array(fac)
    return dict(xyzh=a[:, :4].copy(), radius=a[:, 4].copy(), temperature=a[:, 5].copy(),
                kernel=kernels[kname])


def particle_bin_luminosities(sim: Simulation):
    """Luminosity (W) of each particle in each wavelength bin: BlackBodySEDFamily's
    L_lambda = 4 pi^2 R^2 B_lambda(T) (BlackBodySEDFamily.cpp:25-32) times the bin width."""
    p = sim.particles
    lam = sim.wavelengths
    R, T = p["radius"][:, None], p["temperature"][:, None]
    return 4 * math.pi * math.pi * R * R * planck(lam[None, :], T) * bin_widths(lam)[None, :]


def planck(lam, T):
    """Black-body specific intensity shape B_lambda(T) (PlanckFunction)."""
    x = H_PLANCK * C_LIGHT / (lam * K_BOLTZ * T)
    return 2 * H_PLANCK * C_LIGHT ** 2 / lam ** 5 / np.expm1(x)


def bin_widths(lam):
    """Oligochromatic bins: constant half width 1e-3 * lambda_min (OligoWavelengthGrid.cpp:24)."""
    delta = lam[0] * 1e-3
    return (lam + delta) - (lam - delta)


def bin_luminosities(sim: Simulation):
    """Source luminosity (W) per wavelength bin."""
    lam = sim.wavelengths
    Llam = sim.norm_specific_luminosity * planck(lam, sim.temperature) / planck(sim.norm_wavelength, sim.temperature)
    return Llam * bin_widths(lam)


def calibration(sim: Simulation, frame: Frame):
    """Recorded luminosity -> MJy/sr, per wavelength bin (FluxRecorder::calibrateAndWrite)."""
    lam = sim.wavelengths
    d = frame.distance
    dx, dy = frame.fovx / frame.nx, frame.fovy / frame.ny
    omega = 4.0 * math.atan(0.5 * dx / d) * math.atan(0.5 * dy / d)
    return 1.0 / (4.0 * math.pi * d * d) / omega / bin_widths(lam) * (lam * lam / C_LIGHT) * 1e20


def perspective_setup(view, cross, up, focal, width, nx, ny):
    """PerspectiveInstrument::setupSelfBefore (PerspectiveInstrument.cpp:22-80) with
    HomogeneousTransform's row-vector convention (p' = p M, each step post-multiplied,
    HomogeneousTransform.cpp): the 4x4 transform from world coordinates to pixel coordinates
    and depth, the eye position and the pixel size."""
    Vx, Vy, Vz = view
    Cx, Cy, Cz = cross
    Ux, Uy, Uz = up
    Fe = focal
    G = math.sqrt((Vx - Cx) ** 2 + (Vy - Cy) ** 2 + (Vz - Cz) ** 2)
    if G < 1e-20 or (Ux == 0 and Uy == 0 and Uz == 0):
        raise UnsupportedSki("degenerate perspective instrument")
    a, b, c = (Vx - Cx) / G, (Vy - Cy) / G, (Vz - Cz) / G
    s = width / nx
    E = (Vx + Fe * a, Vy + Fe * b, Vz + Fe * c)
    M = [[1.0 if i == j else 0.0 for j in range(4)] for i in range(4)]

    def cat(T):
        nonlocal M
        T = [list(map(float, row)) for row in T]
        M = [[M[i][0] * T[0][j] + M[i][1] * T[1][j] + M[i][2] * T[2][j] + M[i][3] * T[3][j] for j in range(4)]
             for i in range(4)]

    def translate(x, y, z):
        T = np.identity(4)
        T[3, :3] = (x, y, z)
        cat(T)

    def scale(x, y, z):
        cat(np.diag([x, y, z, 1.0]))

    def rot(i, j, cs, sn):
        T = np.identity(4)
        T[i, i] = cs
        T[j, j] = cs
        T[i, j] = -sn
        T[j, i] = sn
        cat(T)

    translate(-E[0], -E[1], -E[2])
    v = math.sqrt(b * b + c * c)
    if v > 0.3:
        rot(1, 2, c / v, -b / v)   # rotateX
        rot(0, 2, v, -a)           # rotateY
        k = (b * b + c * c) * Ux - a * b * Uy - a * c * Uz
        l = c * Uy - b * Uz
        u = math.sqrt(k * k + l * l)
        rot(0, 1, l / u, -k / u)   # rotateZ
    else:
        v = math.sqrt(a * a + c * c)
        rot(0, 2, c / v, -a / v)   # rotateY
        rot(1, 2, v, -b)           # rotateX
        k = c * Ux - a * Uz
        l = (a * a + c * c) * Uy - a * b * Ux - b * c * Uz
        u = math.sqrt(k * k + l * l)
        rot(0, 1, l / u, -k / u)   # rotateZ
    scale(1.0, 1.0, -1.0)
    T = np.identity(4)             # perspectiveZ
    T[2, 3] = 1.0 / Fe
    T[3, 2] = -Fe
    T[3, 3] = 0.0
    cat(T)
    scale(1.0 / s, 1.0 / s, 1.0)
    translate(nx / 2.0, ny / 2.0, 0.0)
    return np.array(M), E, s


def perspective_calibration(sim: Simulation, cam: Perspective):
    """Recorded luminosity (already divided by the depth squared) -> MJy/sr for a local
    instrument (FluxRecorder::calibrateAndWrite with 4 pi instead of 4 pi d^2 and the pixel
    solid angle alpha^2, alpha = 2 atan(s / 2 focal), PerspectiveInstrument.cpp:75-78)."""
    lam = sim.wavelengths
    s = cam.width / cam.nx
    alpha = 2.0 * math.atan(0.5 * s / cam.focal)
    return 1.0 / (4.0 * math.pi) / (alpha * alpha) / bin_widths(lam) * (lam * lam / C_LIGHT) * 1e20


def sed_calibration(sim: Simulation, sed: Sed):
    """Recorded luminosity -> F_nu in Jy, per wavelength bin (FluxRecorder::calibrateAndWrite)."""
    lam = sim.wavelengths
    d = sed.distance
    return 1.0 / (4.0 * math.pi * d * d) / bin_widths(lam) * (lam * lam / C_LIGHT) * 1e26


def _smart(v):
    """StringUtils::toString(double): %1.10g without exponent zero padding or '+'."""
    t = f"{v:1.10g}"
    return t.replace("e-0", "e-").replace("e+0", "e").replace("e+", "e")


def write_sed(path, values, sed: Sed, lam):
    """The SED text file of FluxRecorder (header, column lines, '%.9e' values)."""
    with open(path, "w") as f:
        f.write(f"# SED at inclination {_smart(math.degrees(sed.inclination))} deg, azimuth "
                f"{_smart(math.degrees(sed.azimuth))} deg, distance {_smart(sed.distance / MPC)} Mpc\n")
        f.write("# column 1: wavelength; lambda (micron)\n")
        f.write("# column 2: total flux; F_nu (Jy)\n")
        for l, v in zip(lam, values):
            f.write(f"{l * 1e6:.9e} {v:.9e}\n")


def write_mean_intensity(path, J, lam):
    """RadiationFieldProbe with PerCellForm: J_nu per cell and wavelength (W/m2/Hz/sr)."""
    with open(path, "w") as f:
        f.write("# Mean intensity per spatial cell\n")
        f.write("# column 1: spatial cell index (1)\n")
        for c, l in enumerate(lam):
            f.write(f"# column {c + 2}: J_nu at lambda = {_smart(l * 1e6)} micron (W/m2/Hz/sr)\n")
        for m in range(J.shape[0]):
            f.write(str(m) + "".join(f" {v:.9e}" for v in J[m]) + "\n")


def mean_intensity(sim: Simulation, rf):
    """MediumSystem::meanIntensity (MediumSystem.cpp:2670-2680): the stored field / (4 pi V
    Δλ), as J_nu = J_lambda * lambda^2 / c."""
    lam = sim.wavelengths
    return rf / (4.0 * math.pi * sim.cell_volumes[:, None]) / bin_widths(lam)[None, :] * (lam * lam / C_LIGHT)[None, :]


def read_sed(path):
    """(wavelength in micron, F_nu in Jy) columns of an SED text file."""
    rows = [list(map(float, ln.split())) for ln in open(path) if ln.strip() and not ln.startswith("#")]
    return np.array(rows).reshape(-1, 2)


def write_fits(path, cube, frame: Frame, lam):
    """Float32 FITS cube (x fastest, then y, then wavelength) with the reference's main cards."""
    cube = np.asarray(cube, dtype=np.float64)
    nz, ny, nx = cube.shape
    cards = [("SIMPLE", "T"), ("BITPIX", -32), ("NAXIS", 3), ("NAXIS1", nx), ("NAXIS2", ny), ("NAXIS3", nz),
             ("BSCALE", 1.0), ("BZERO", 0.0), ("ORIGIN", "'skirt9-gpu_amd'"), ("BUNIT", "'MJy/sr'"),
             ("CRPIX1", (nx + 1) / 2.0), ("CRVAL1", frame.cx / PC), ("CDELT1", frame.fovx / nx / PC),
             ("CUNIT1", "'pc'"), ("CRPIX2", (ny + 1) / 2.0), ("CRVAL2", frame.cy / PC),
             ("CDELT2", frame.fovy / ny / PC), ("CUNIT2", "'pc'"), ("CUNIT3", "'micron'"),
             ("CROTA1", math.degrees(frame.inclination)), ("CROTA2", math.degrees(frame.azimuth)),
             ("CROTA3", math.degrees(frame.roll)), ("DISTLUMI", frame.distance / (1e6 * PC)),
             ("DISTUNIT", "'Mpc'")]
    head = ""
    for k, v in cards:
        if isinstance(v, bool) or v == "T":
            s = "T"
        elif isinstance(v, str):
            s = v
        elif isinstance(v, int):
            s = str(v)
        else:
            s = repr(float(v)).upper()
        head += f"{k:<8}= {s:>20}".ljust(80)
    head += "END".ljust(80)
    head += " " * ((-len(head)) % 2880)
    data = cube.astype(">f4").tobytes()
    data += b"\0" * ((-len(data)) % 2880)
    with open(path, "wb") as f:
        f.write(head.encode("ascii"))
        f.write(data)


def read_fits(path):
    """Reads the primary image of a float32 FITS file as (nz, ny, nx) or (ny, nx) float64."""
    raw = open(path, "rb").read()
    hdr, pos = {}, 0
    while True:
        card = raw[pos:pos + 80].decode("ascii")
        pos += 80
        if card[:8].strip() == "END":
            break
        if card[8:10] == "= ":
            hdr[card[:8].strip()] = card[10:].split("/")[0].strip()
    pos = (pos + 2879) // 2880 * 2880
    dims = [int(hdr[f"NAXIS{i + 1}"]) for i in range(int(hdr["NAXIS"]))]
    n = int(np.prod(dims))
    return np.frombuffer(raw, dtype=">f4", count=n, offset=pos).astype(np.float64).reshape(dims[::-1])


def imported_cell_densities(sites, density, cell_sites):
    """The density of each cell of a mesh built from imported sites: the mesh's cells come in
    the builder's order (sorted by x), so the file's rows are matched to them by coordinates."""
    if len(cell_sites) != len(sites):
        raise UnsupportedSki("imported sites outside the domain or closer than eps are not supported")
    order_in = np.lexsort(np.asarray(sites).T[::-1])
    order_cells = np.lexsort(np.asarray(cell_sites).T[::-1])
    if not np.array_equal(np.asarray(sites)[order_in], np.asarray(cell_sites)[order_cells]):
        raise UnsupportedSki("the mesh's sites differ from the imported ones")
    dens = np.empty(len(order_cells))
    dens[order_cells] = np.asarray(density)[order_in]
    return dens


def setup(sim: Simulation, engine, threads=0, npk_norm=None):
    """Builds the grid and medium on the host, uploads them and configures the engine."""
    from . import workloads as W
    pkg = W._pkg()
    if not sim.force_scattering:
        raise UnsupportedSki("the engine runs the forced-scattering life cycle only")
    if sim.source_position != (0.0, 0.0, 0.0) or any(abs(b) != abs(sim.box[0]) for b in sim.box):
        raise UnsupportedSki("a point source at the centre of a cubic domain centred on the origin is supported")
    mu = W.MU_DUST
    tab = sim.dust_table
    sext, ssca, g = pkg.dust_meanlist(tab[:, 0], tab[:, 1], tab[:, 2], tab[:, 3], sim.wavelengths, mu)
    sv, _, _ = pkg.dust_meanlist(tab[:, 0], tab[:, 1], tab[:, 2], tab[:, 3], [sim.tau_wavelength], mu)
    sp = sim.spiral_params
    model = None if sim.imported is not None else pkg.disk_model(
        tau_v=sim.tau_v, sext_v=float(sv[0]), spiral=sim.spiral, scale_length=sim.scale_length,
        scale_height=sim.scale_height, **({} if not sp else dict(
            arms=sp["arms"], pitch=sp["pitch"], radius_zero=sp["radius_zero"], phase_zero=sp["phase_zero"],
            weight=sp["weight"], index=sp["index"])))
    h = sim.box[3]
    if sim.grid["kind"] == "cartesian":
        borders = []
        for n in sim.grid["n"]:
            t = np.array([0.0 + i * ((1.0 - 0.0) / n) for i in range(n + 1)])
            borders.append(t * (h - (-h)) + (-h))
        engine.grid_cartesian(*borders)
        dens = pkg.sample_density_cartesian(*borders, model, sim.num_density_samples, sim.seed + 1, threads)
        # cell volumes, cell index k + nz*j + nz*ny*i (CartesianSpatialGrid.cpp:200-203)
        dx, dy, dz = (np.diff(bb) for bb in borders)
        sim.cell_volumes = (dx[:, None, None] * dy[None, :, None] * dz[None, None, :]).ravel()
    elif sim.grid["kind"] == "voronoi_imported":
        # the imported sites' mesh; each cell keeps its site's density (cells in the builder's
        # order, sorted by x, matched to the file's rows by their coordinates)
        imp = sim.imported
        mesh = pkg.Voronoi.build(imp["sites"], sim.box, threads)
        ex = mesh.export()
        dens = imported_cell_densities(imp["sites"], imp["density"], ex["sites"])
        mesh.upload(engine)
        sim.cell_volumes = ex["volume"]
        mesh.close()
    elif sim.grid["kind"] in ("voronoi", "tetra"):
        # TetraMeshSpatialGrid / VoronoiMeshSpatialGrid with the DustDensity policy: points
        # distributed as the medium (TetraMeshSpatialGrid.cpp:23-40), then the mesh
        if "points" in sim.grid:
            pts = sim.grid["points"]
        else:
            pts = W.disk_points(sim.grid["n"], sim.box, sim.scale_length, sim.scale_height, seed=sim.seed + 3,
                                spiral=sim.spiral_params if sim.spiral else None)
        if sim.grid["kind"] == "voronoi":
            mesh = pkg.Voronoi.build(pts, sim.box, threads)
            mesh.upload(engine)
            dens = mesh.sample_density_model(model, sim.num_density_samples, sim.seed + 1, threads)
            sim.cell_volumes = mesh.export()["volume"]
        else:
            mesh = pkg.Tetra.build(pts, sim.box)
            mesh.upload(engine)
            dens = engine.sample_density(model, sim.num_density_samples, sim.seed + 1)
            ex = mesh.export()
            V, T = ex["verts"], ex["tet_verts"]
            a, b_, c, d = (V[T[:, q]] for q in range(4))
            sim.cell_volumes = np.abs(np.einsum("ij,ij->i", np.cross(b_ - a, c - a), d - a)) / 6
        mesh.close()
    else:
        gr = sim.grid
        tree = pkg.Octree.build(list(sim.box), model, gr["min_level"], gr["max_level"], gr["max_fraction"],
                                sim.num_density_samples, sim.seed + 1, threads)
        tree.upload(engine)
        dens = tree.sample_density(model, sim.num_density_samples, sim.seed + 2, threads)
        ex = tree.export()
        leaf = ex["cell_index"] >= 0
        bnd = ex["bounds"].reshape(-1, 6)[leaf]
        vol = np.empty(int(leaf.sum()))
        vol[ex["cell_index"][leaf]] = np.prod(bnd[:, 3:] - bnd[:, :3], axis=1)
        sim.cell_volumes = vol
    engine.medium(dens)
    engine.dust(sim.wavelengths, sext, ssca, g)
    engine.frame_remove_all()
    for f in sim.frames:
        if isinstance(f, Perspective):
            M, E, ps = perspective_setup(f.view, f.cross, f.up, f.focal, f.width, f.nx, f.ny)
            engine.perspective_add(M, E, ps, f.nx, f.ny)
        elif isinstance(f, Sed):
            engine.sed_add(f.inclination, f.azimuth, f.radius)
        else:
            engine.frame_add(f.inclination, f.azimuth, f.roll, f.nx, f.ny, f.fovx, f.fovy, f.cx, f.cy)
    # the bin is chosen uniformly, so each packet of bin ell carries nlambda * L_ell / N
    npk = npk_norm or sim.num_packets
    if sim.particles is not None:
        # the particles' launch ranges split the history indices [0, npk); the launch weights use
        # each particle's luminosity summed over the bins
        lbin = particle_bin_luminosities(sim)
        engine.source_point((0.0, 0.0, 0.0), np.zeros(len(sim.wavelengths)))
        engine.source_particles(sim.particles["xyzh"], lbin, lbin.sum(axis=1), npk, 0.5, sim.particles["kernel"])
    else:
        engine.source_point(sim.source_position, bin_luminosities(sim) * len(sim.wavelengths) / npk)
    engine.options(sim.path_length_bias, sim.min_weight_reduction, sim.min_scatt_events, 0, sim.seed)
    # the persistent engine: the faster one here, and the one that stores the radiation field
    engine.tune_pipeline(pkg.SKG_MODE_PERSISTENT, 1 << 21, 0, 8)


def run(sim: Simulation, outdir: str, prefix: str, engine=None, chunk=1 << 22, threads=0, log=print):
    """Runs the simulation on the GPU and writes one FITS cube per instrument; returns the
    calibrated cubes {instrument name: (nlambda, ny, nx)}."""
    from . import Engine
    own = engine is None
    eng = engine or Engine(0)
    t0 = time.time()
    setup(sim, eng, threads)
    log(f"set-up {time.time() - t0:.1f} s")
    eng.frame_clear()
    if sim.rf_probes:
        eng.rf_enable(True)
    t1 = time.time()
    done = 0
    while done < sim.num_packets:
        n = min(chunk, sim.num_packets - done)
        eng.run_packets(done, n)
        done += n
    log(f"{sim.num_packets} photon packets in {time.time() - t1:.2f} s")
    os.makedirs(outdir, exist_ok=True)
    out = {}
    for i, f in enumerate(sim.frames):
        if isinstance(f, Perspective):
            cube = eng.frame_read(i) * perspective_calibration(sim, f)[:, None, None]
            ps = f.width / f.nx
            hdr = Frame(f.name, 0.0, 0.0, 0.0, 0.0, ps * f.nx, ps * f.ny, f.nx, f.ny, 0.0, 0.0)
            write_fits(os.path.join(outdir, f"{prefix}_{f.name}_total.fits"), cube, hdr, sim.wavelengths)
            out[f.name] = cube
            continue
        if isinstance(f, Sed):
            vals = eng.frame_read(i)[:, 0, 0] * sed_calibration(sim, f)
            write_sed(os.path.join(outdir, f"{prefix}_{f.name}_sed.dat"), vals, f, sim.wavelengths)
            out[f.name] = vals
            continue
        cube = eng.frame_read(i) * calibration(sim, f)[:, None, None]
        write_fits(os.path.join(outdir, f"{prefix}_{f.name}_total.fits"), cube, f, sim.wavelengths)
        out[f.name] = cube
    if sim.rf_probes:
        J = mean_intensity(sim, eng.rf_read())
        eng.rf_enable(False)
        for name in sim.rf_probes:
            write_mean_intensity(os.path.join(outdir, f"{prefix}_{name}_J.dat"), J, sim.wavelengths)
        out["J"] = J
    if own:
        eng.close()
    return out
