"""Workload builders: the scenes, phantoms, sources and tallies of the
benchmark configurations, applied to anything with the Context call shapes
(set_domain / add_grid / set_source / set_tallies / enable_voxel_dose).

Material IDs follow the registry in data/materials.csv: 3 "Air, Dry (near sea
level)", 48 "Water, Liquid", 155 "Lung - C5", 168 "Breast - C5" (adipose),
169 "Cortical Bone - C5".
"""
import numpy as np

AIR, WATER, LUNG, ADIPOSE, BONE = 3, 48, 155, 168, 169
TALLY_AACUBOID, TALLY_DISC, TALLY_RECT, TALLY_IMAGE = 0, 1, 2, 3
DIR_BEAM, DIR_ISOTROPIC, DIR_RECT_ISOTROPIC, DIR_DISC_ISOTROPIC = 0, 1, 2, 3


def kramers(kvp_kev, e_lo_kev=15.0, step_kev=1.0):
    """Kramers spectrum weights (kVp - E) / E on 1 keV bins, energies in eV."""
    e = np.arange(e_lo_kev, kvp_kev, step_kev)
    return e * 1e3, (kvp_kev - e) / e


def ellipsoid_phantom(shape, seed, background=WATER, specs=((BONE, (3, 8), (0.5, 3.0)), (ADIPOSE, (2, 5), (2.0, 6.0))),
                      voxel_cm=0.125):
    """uint8 material grid [x, y, z] with seeded random axis-aligned ellipsoids.

    specs: (material, (min count, max count), (min semi-axis cm, max semi-axis cm)).
    """
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    ids = np.full(shape, background, dtype=np.uint8)
    x = (np.arange(nx) + 0.5) * voxel_cm
    y = (np.arange(ny) + 0.5) * voxel_cm
    z = (np.arange(nz) + 0.5) * voxel_cm
    for mat, (cmin, cmax), (rmin, rmax) in specs:
        for _ in range(int(rng.integers(cmin, cmax + 1))):
            c = rng.uniform([0, 0, 0], [nx * voxel_cm, ny * voxel_cm, nz * voxel_cm])
            r = rng.uniform(rmin, rmax, 3)
            ix = np.nonzero(np.abs(x - c[0]) <= r[0])[0]
            iy = np.nonzero(np.abs(y - c[1]) <= r[1])[0]
            iz = np.nonzero(np.abs(z - c[2]) <= r[2])[0]
            if not len(ix) or not len(iy) or not len(iz):
                continue
            X, Y, Z = np.meshgrid(x[ix], y[iy], z[iz], indexing="ij")
            inside = ((X - c[0]) / r[0]) ** 2 + ((Y - c[1]) / r[1]) ** 2 + ((Z - c[2]) / r[2]) ** 2 <= 1.0
            sub = ids[ix[0]:ix[-1] + 1, iy[0]:iy[-1] + 1, iz[0]:iz[-1] + 1]
            sub[inside] = mat
    return ids


# --------------------------------------------------------------------------- C1
C1_MATERIALS = (AIR, WATER)


def setup_slab(sim, energy=60e3, voxel_dose=True, hist_bins=61):
    """Plain method: 4x4x100 cm air domain, 40x40x10 water grid (1 mm) at
    z = 10 cm, 60 keV pencil beam along +z from (2, 2, 0); an AABB volume tally
    over the slab and a 4 cm disc at z = 100 with a 1 keV IncidentEnergy
    histogram."""
    sim.set_domain((4.0, 4.0, 100.0), AIR)
    sim.add_grid(np.full((40, 40, 10), WATER, np.uint8), (0.1, 0.1, 0.1), (0.0, 0.0, 10.0))
    sim.set_source(direction=DIR_BEAM, position=(2.0, 2.0, 0.0), a=(2.0, 2.0, 100.0), energy=energy)
    sim.set_tallies([
        dict(kind=TALLY_AACUBOID, p0=(0.0, 0.0, 10.0), p1=(4.0, 4.0, 11.0)),
        dict(kind=TALLY_DISC, p0=(2.0, 2.0, 100.0), normal=(0.0, 0.0, 1.0), radius=2.0, hist_bins=hist_bins,
             hist_lo=0.0, hist_width=1000.0),
    ])
    if voxel_dose:
        sim.enable_voxel_dose(True)


# --------------------------------------------------------------------------- C2
C2_MATERIALS = (AIR, WATER, BONE, ADIPOSE)


def setup_radiography(sim, shape=(256, 256, 160), voxel_cm=0.125, seed=2, voxel_dose=True, image=(256, 256),
                      kvp=80.0):
    """Radiography: air domain 30 x 30 cm laterally, phantom of seeded
    ellipsoids (water background, 3-8 cortical-bone ellipsoids of semi-axes
    0.5-3 cm, 2-5 adipose of 2-6 cm), point source 60 cm above the phantom's
    top face, cone beam (RectangularIsotropicDirectionality) over that face,
    Kramers spectrum 15-80 keV in 1 keV bins, an energy-fluence image at the
    far face and per-voxel deposited energy.  The domain is extended in z so
    that it contains the source (a photon born outside the domain dies on its
    first step in the reference)."""
    nx, ny, nz = shape
    lx, ly, lz = nx * voxel_cm, ny * voxel_cm, nz * voxel_cm
    dim = (30.0, 30.0, lz + 60.0 + 1.0)
    z0 = 0.5  # phantom from z0 to z0 + lz; image plane just below it
    sim.set_domain(dim, AIR)
    ox, oy = (30.0 - lx) / 2, (30.0 - ly) / 2
    sim.add_grid(ellipsoid_phantom(shape, seed, voxel_cm=voxel_cm), (voxel_cm,) * 3, (ox, oy, z0))
    src = (15.0, 15.0, z0 + lz + 60.0)
    e, w = kramers(kvp)
    sim.set_source(direction=DIR_RECT_ISOTROPIC, position=src, a=(ox, oy, z0 + lz), b=(lx, 0.0, 0.0),
                   c=(0.0, ly, 0.0), spectrum=(e, w))
    tallies = []
    if image:
        # edge1 x edge2 points along -z, the direction of travel
        tallies.append(dict(kind=TALLY_IMAGE, p0=(ox, oy, z0 - 0.25), p1=(0.0, ly, 0.0), p2=(lx, 0.0, 0.0),
                            nx=image[1], ny=image[0]))
    sim.set_tallies(tallies)
    if voxel_dose:
        sim.enable_voxel_dose(True)
    return dict(shape=shape, voxel_cm=voxel_cm, origin=(ox, oy, z0), source=src)
