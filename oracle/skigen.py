"""Render SKIRT ski files for the workloads used by the tests and the CPU baseline.

Test infrastructure: the ski files drive the reference build in oracle/_ref (fixture
generation and the "reference" CPU timing).  The item names and attributes are the
reference's configuration vocabulary (ExpDiskGeometry, SpiralStructureGeometryDecorator,
MeanListDustMix, CartesianSpatialGrid, PolicyTreeSpatialGrid/DensityTreePolicy,
FrameInstrument, PhotonPacketOptions); a working oligochromatic point-source setup needs a
BlackBodySED with a SpecificLuminosityNormalization (SingleWavelengthSED is rejected in
oligochromatic mode).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Sequence, Tuple


@dataclass
class SkiSpec:
    num_packets: float = 1e5
    wavelengths_um: Sequence[float] = (0.55,)
    # dust table: (lambda_um, kappa_ext m2/kg, albedo, g); log-log resampled by the reference
    dust: Sequence[Tuple[float, float, float, float]] = ((0.55, 1.0, 0.6, 0.6),)
    spiral: bool = False
    tau_v: float = 1.0
    grid: Tuple = ("cartesian", 64)  # or ("octree", min_level, max_level, max_dust_fraction)
    half_size_pc: float = 10.0
    # instruments: (name, inclination deg, azimuth deg)
    instruments: Sequence[Tuple[str, float, float]] = (("i60", 60.0, 0.0),)
    # SED instruments: (name, inclination deg, azimuth deg, aperture radius pc; 0 = none)
    sed_instruments: Sequence[Tuple[str, float, float, float]] = ()
    # perspective instruments: (name, view (x,y,z) pc, crosshair (x,y,z) pc, up (x,y,z), focal pc,
    # viewport width pc, pixels per side)
    persp_instruments: Sequence = ()
    # store the radiation field and probe its mean intensity per cell (RadiationFieldProbe)
    store_rf: bool = False
    # a ParticleSource (black-body SED family) read from this file, next to the ski file,
    # instead of the point source; and its smoothing kernel
    particle_file: str = ""
    particle_kernel: str = "CubicSplineSmoothingKernel"
    num_pixels: int = 200
    fov_pc: float = 20.0
    seed: int = 0
    force_scattering: bool = True
    min_weight_reduction: float = 1e4
    min_scatt_events: int = 0
    path_length_bias: float = 0.5
    num_density_samples: int = 100
    # an imported Voronoi medium (VoronoiMeshMedium: columns x, y, z in pc and the number density
    # in 1/m3, written by write_voronoi_medium) instead of the disk; pair it with the grid
    # ("voronoi_imported",), which employs that medium's mesh (policy ImportedMesh)
    medium_file: str = ""


def _fmt_list(vals, unit):
    return ", ".join(f"{v:.9g} {unit}" for v in vals)


def _source(spec: SkiSpec) -> str:
    if spec.particle_file:
        return (f'<ParticleSource filename="{spec.particle_file}" importVelocity="false" '
                'importVelocityDispersion="false" importCurrentMass="false" importBias="false" useColumns="" '
                'sourceWeight="1" wavelengthBias="0.5">'
                f'<smoothingKernel type="SmoothingKernel"><{spec.particle_kernel}/></smoothingKernel>'
                '<sedFamily type="SEDFamily"><BlackBodySEDFamily/></sedFamily></ParticleSource>')
    return ('<PointSource positionX="0 pc" positionY="0 pc" positionZ="0 pc" velocityX="0 km/s" velocityY="0 km/s" '
            'velocityZ="0 km/s" sourceWeight="1" wavelengthBias="0.5">'
            '<sed type="SED"><BlackBodySED temperature="5800 K"/></sed>'
            '<normalization type="LuminosityNormalization"><SpecificLuminosityNormalization wavelength="0.55 micron" '
            'unitStyle="wavelengthmonluminosity" specificLuminosity="1 Lsun/micron"/></normalization></PointSource>')


def disk_particles(n=30, seed=12):
    """A fixed set of stellar particles in the disk plane: x, y, z, h (pc), radius (km),
    temperature (K) per row (radii 1-7 pc, smoothing lengths 0.5-2 pc, 3500-12000 K)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    R = rng.uniform(1, 7, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    return np.column_stack([R * np.cos(ph), R * np.sin(ph), rng.normal(0, 0.2, n), rng.uniform(0.5, 2.0, n),
                            rng.uniform(3e5, 3e6, n), rng.uniform(3500, 12000, n)])


def write_particles(path, rows):
    """A particle file for ParticleSource with BlackBodySEDFamily: x, y, z, h (pc), radius (km),
    temperature (K) per row."""
    with open(path, "w") as f:
        for i, (name, unit) in enumerate((("position x", "pc"), ("position y", "pc"), ("position z", "pc"),
                                          ("size h", "pc"), ("radius", "km"), ("temperature", "K"))):
            f.write(f"# column {i + 1}: {name} ({unit})\n")
        for r in rows:
            f.write(" ".join(f"{v:.10g}" for v in r) + "\n")


def render(spec: SkiSpec) -> str:
    wl = _fmt_list(spec.wavelengths_um, "micron")
    dl = _fmt_list([d[0] for d in spec.dust], "micron")
    kappa = ", ".join(f"{d[1]:.9g} m2/kg" for d in spec.dust)
    alb = ", ".join(f"{d[2]:.9g}" for d in spec.dust)
    gg = ", ".join(f"{d[3]:.9g}" for d in spec.dust)
    disk = ('<ExpDiskGeometry scaleLength="3 pc" scaleHeight="0.3 pc" minRadius="0 pc" '
            'maxRadius="0 pc" maxZ="0 pc"/>')
    if spec.spiral:
        geom = ('<SpiralStructureGeometryDecorator numArms="2" pitchAngle="15 deg" radiusZeroPoint="3 pc" '
                'phaseZeroPoint="0 deg" perturbationWeight="0.5" index="1"><geometry type="AxGeometry">'
                + disk + '</geometry></SpiralStructureGeometryDecorator>')
    else:
        geom = disk
    h = spec.half_size_pc
    box = f'minX="-{h} pc" maxX="{h} pc" minY="-{h} pc" maxY="{h} pc" minZ="-{h} pc" maxZ="{h} pc"'
    if spec.grid[0] == "cartesian":
        n = spec.grid[1]
        grid = (f'<CartesianSpatialGrid {box}>'
                f'<meshX type="Mesh"><LinMesh numBins="{n}"/></meshX>'
                f'<meshY type="Mesh"><LinMesh numBins="{n}"/></meshY>'
                f'<meshZ type="Mesh"><LinMesh numBins="{n}"/></meshZ></CartesianSpatialGrid>')
    elif spec.grid[0] == "octree":
        _, lmin, lmax, frac = spec.grid
        grid = (f'<PolicyTreeSpatialGrid {box} treeType="OctTree"><policy type="TreePolicy">'
                f'<DensityTreePolicy minLevel="{lmin}" maxLevel="{lmax}" maxDustFraction="{frac:.9g}" '
                'maxDustOpticalDepth="0" wavelength="0.55 micron" maxDustDensityDispersion="0" '
                'maxElectronFraction="1e-6" maxGasFraction="1e-6"/></policy></PolicyTreeSpatialGrid>')
    elif spec.grid[0] == "voronoi":
        # ("voronoi", num_sites, policy): sites drawn by the grid's policy (VoronoiMeshSpatialGrid.cpp:51-140)
        _, nsites, policy = spec.grid
        grid = (f'<VoronoiMeshSpatialGrid {box} policy="{policy}" numSites="{nsites}" relaxSites="false"/>')
    elif spec.grid[0] == "voronoi_imported":
        # the imported medium's own mesh (VoronoiMeshSpatialGrid.cpp:133-141)
        grid = f'<VoronoiMeshSpatialGrid {box} policy="ImportedMesh"/>'
    elif spec.grid[0] == "tetra":
        # ("tetra", num_sites, policy): Delaunay tetrahedralisation of the sampled vertices
        _, nsites, policy = spec.grid
        grid = f'<TetraMeshSpatialGrid {box} policy="{policy}" numSamples="{nsites}" refine="false"/>'
    elif spec.grid[0] == "voronoi_file":
        # ("voronoi_file", filename): the Voronoi tessellation of given sites (policy File,
        # VoronoiMeshSpatialGrid.cpp:122-126), written by write_points
        grid = f'<VoronoiMeshSpatialGrid {box} policy="File" filename="{spec.grid[1]}" relaxSites="false"/>'
    elif spec.grid[0] == "tetra_file":
        # ("tetra_file", filename): Delaunay tetrahedralisation of given vertices (policy File,
        # TetraMeshSpatialGrid.cpp:303-312), written by write_points
        grid = f'<TetraMeshSpatialGrid {box} policy="File" filename="{spec.grid[1]}" refine="false"/>'
    else:
        raise ValueError(spec.grid)
    instr = "".join(
        f'<FrameInstrument instrumentName="{name}" distance="10 Mpc" inclination="{inc:.9g} deg" '
        f'azimuth="{az:.9g} deg" roll="0 deg" fieldOfViewX="{spec.fov_pc:.9g} pc" numPixelsX="{spec.num_pixels}" '
        f'centerX="0 pc" fieldOfViewY="{spec.fov_pc:.9g} pc" numPixelsY="{spec.num_pixels}" centerY="0 pc" '
        'recordComponents="false" numScatteringLevels="0" recordPolarization="false" recordStatistics="false"/>'
        for (name, inc, az) in spec.instruments) + "".join(
        f'<SEDInstrument instrumentName="{name}" distance="10 Mpc" inclination="{inc:.9g} deg" '
        f'azimuth="{az:.9g} deg" roll="0 deg" radius="{rad:.9g} pc" recordComponents="false" '
        'numScatteringLevels="0" recordPolarization="false" recordStatistics="false"/>'
        for (name, inc, az, rad) in spec.sed_instruments) + "".join(
        f'<PerspectiveInstrument instrumentName="{name}" numPixelsX="{n}" numPixelsY="{n}" width="{w:.9g} pc" '
        f'viewX="{v[0]:.9g} pc" viewY="{v[1]:.9g} pc" viewZ="{v[2]:.9g} pc" crossX="{c[0]:.9g} pc" '
        f'crossY="{c[1]:.9g} pc" crossZ="{c[2]:.9g} pc" upX="{u[0]:.9g} pc" upY="{u[1]:.9g} pc" upZ="{u[2]:.9g} pc" '
        f'focal="{fe:.9g} pc" recordComponents="false" numScatteringLevels="0" recordPolarization="false" '
        'recordStatistics="false"/>'
        for (name, v, c, u, fe, w, n) in spec.persp_instruments)
    mix = (f'<materialMix type="MaterialMix"><MeanListDustMix wavelengths="{dl}" extinctionCoefficients="{kappa}" '
           f'albedos="{alb}" asymmetryParameters="{gg}"/></materialMix>')
    if spec.medium_file:
        medium = (f'<VoronoiMeshMedium filename="{spec.medium_file}" minX="-{h} pc" maxX="{h} pc" minY="-{h} pc" '
                  f'maxY="{h} pc" minZ="-{h} pc" maxZ="{h} pc" massType="NumberDensity" massFraction="1" '
                  'importMetallicity="false" importTemperature="false" maxTemperature="0 K" importVelocity="false" '
                  f'importMagneticField="false" importVariableMixParams="false" useColumns="">{mix}</VoronoiMeshMedium>')
    else:
        medium = (f'<GeometricMedium velocityMagnitude="0 km/s" magneticFieldStrength="0 uG">'
                  f'<geometry type="Geometry">{geom}</geometry>{mix}'
                  '<normalization type="MaterialNormalization"><OpticalDepthMaterialNormalization axis="Z" '
                  f'wavelength="0.55 micron" opticalDepth="{spec.tau_v:.9g}"/></normalization></GeometricMedium>')
    force = "true" if spec.force_scattering else "false"
    rfopt = ('                <radiationFieldOptions type="RadiationFieldOptions">'
             '<RadiationFieldOptions storeRadiationField="true"/></radiationFieldOptions>\n') if spec.store_rf else ""
    probes = ('<ProbeSystem><probes type="Probe"><RadiationFieldProbe probeName="rf" writeWavelengthGrid="false">'
              '<form type="Form"><PerCellForm/></form></RadiationFieldProbe></probes></ProbeSystem>'
              if spec.store_rf else "<ProbeSystem/>")
    return f"""<?xml version="1.0" encoding="UTF-8"?>
<skirt-simulation-hierarchy type="MonteCarloSimulation" format="9" producer="skigen" time="2026-01-01T00:00:00">
    <MonteCarloSimulation userLevel="Expert" simulationMode="OligoExtinctionOnly" iteratePrimaryEmission="false" numPackets="{spec.num_packets:.9g}">
        <random type="Random"><Random seed="{spec.seed}"/></random>
        <units type="Units"><ExtragalacticUnits fluxOutputStyle="Frequency"/></units>
        <cosmology type="Cosmology"><LocalUniverseCosmology/></cosmology>
        <sourceSystem type="SourceSystem">
            <SourceSystem wavelengths="{wl}" sourceBias="0.5">
                <sources type="Source">
                    {_source(spec)}
                </sources>
            </SourceSystem>
        </sourceSystem>
        <mediumSystem type="MediumSystem">
            <MediumSystem>
                <photonPacketOptions type="PhotonPacketOptions">
                    <PhotonPacketOptions explicitAbsorption="false" forceScattering="{force}" minWeightReduction="{spec.min_weight_reduction:.9g}" minScattEvents="{spec.min_scatt_events}" pathLengthBias="{spec.path_length_bias:.9g}"/>
                </photonPacketOptions>
                <media type="Medium">
                    {medium}
                </media>
{rfopt}                <samplingOptions type="SamplingOptions">
                    <SamplingOptions numDensitySamples="{spec.num_density_samples}" numPropertySamples="1" aggregateVelocity="Average"/>
                </samplingOptions>
                <grid type="SpatialGrid">{grid}</grid>
            </MediumSystem>
        </mediumSystem>
        <instrumentSystem type="InstrumentSystem">
            <InstrumentSystem><instruments type="Instrument">{instr}</instruments></InstrumentSystem>
        </instrumentSystem>
        <probeSystem type="ProbeSystem">{probes}</probeSystem>
    </MonteCarloSimulation>
</skirt-simulation-hierarchy>
"""


def write_voronoi_medium(path, sites_m, density_m3):
    """A VoronoiMeshMedium file: site positions (pc) and the cell's number density (1/m3),
    one site per row."""
    import numpy as np
    pc = 3.08567758e16  # the reference's parsec (Constants::pc)
    with open(path, "w") as f:
        for i, (name, unit) in enumerate((("position x", "pc"), ("position y", "pc"), ("position z", "pc"),
                                          ("number density", "1/m3"))):
            f.write(f"# column {i + 1}: {name} ({unit})\n")
        np.savetxt(f, np.column_stack([np.asarray(sites_m) / pc, density_m3]), fmt="%.17g")


def write_points(path, points_m):
    """A site / vertex file of the mesh grids' File policy: positions x, y, z (pc), one per row."""
    import numpy as np
    pc = 3.08567758e16  # the reference's parsec (Constants::pc)
    with open(path, "w") as f:
        for i, name in enumerate(("position x", "position y", "position z")):
            f.write(f"# column {i + 1}: {name} (pc)\n")
        np.savetxt(f, np.asarray(points_m) / pc, fmt="%.17g")


def log_wavelengths(lo: float, hi: float, n: int) -> List[float]:
    import math
    return [lo * (hi / lo) ** (i / (n - 1)) for i in range(n)] if n > 1 else [lo]


def config0(num_packets=1e6, seed=0) -> SkiSpec:
    """Cartesian 64^3 exponential disk, tau_V=1, 0.55 um, albedo 0.6, g 0.6, frame at 60 deg."""
    return SkiSpec(num_packets=num_packets, seed=seed)


def config1(num_packets=1e8, seed=0, max_level=10, min_level=3, max_frac=1e-6, tau_v=1.0, nlambda=40,
            lam_range=(0.1, 10.0)) -> SkiSpec:
    """Spiral disk on a density-refined octree with a power-law opacity (kappa ~ lambda^-1.5)."""
    return SkiSpec(num_packets=num_packets, seed=seed, spiral=True, tau_v=tau_v,
                   wavelengths_um=log_wavelengths(lam_range[0], lam_range[1], nlambda),
                   dust=((0.1, 1000.0, 0.5, 0.6), (10.0, 1.0, 0.1, 0.6)),
                   grid=("octree", min_level, max_level, max_frac))
