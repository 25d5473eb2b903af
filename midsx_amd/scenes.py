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
    ids = np.full(shape, background, dtype=np.uint8, order="F")  # x fastest in memory, as the library stores it
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
                      kvp=80.0, with_domain=True):
    """Radiography: air domain 30 x 30 cm laterally, phantom of seeded
    ellipsoids (water background, 3-8 cortical-bone ellipsoids of semi-axes
    0.5-3 cm, 2-5 adipose of 2-6 cm), point source 60 cm above the phantom's
    top face, cone beam (RectangularIsotropicDirectionality) over that face,
    Kramers spectrum 15-80 keV in 1 keV bins, an energy-fluence image at the
    far face and per-voxel deposited energy.  The domain is extended in z so
    that it contains the source (a photon born outside the domain dies on its
    first step in the reference).  with_domain=False sets only the source and
    tallies (for a job that uses another job's phantom)."""
    nx, ny, nz = shape
    lx, ly, lz = nx * voxel_cm, ny * voxel_cm, nz * voxel_cm
    dim = (30.0, 30.0, lz + 60.0 + 1.0)
    z0 = 0.5  # phantom from z0 to z0 + lz; image plane just below it
    ox, oy = (30.0 - lx) / 2, (30.0 - ly) / 2
    if with_domain:
        sim.set_domain(dim, AIR)
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
    if voxel_dose and with_domain:
        sim.enable_voxel_dose(True)
    return dict(shape=shape, voxel_cm=voxel_cm, origin=(ox, oy, z0), source=src)


# --------------------------------------------------------------------------- dose jobs
DOSE_MATERIALS = (AIR, WATER, LUNG, ADIPOSE, BONE)
DOSE_SHAPES = (64, 128, 256, 512)
DOSE_KVP = (60.0, 80.0, 100.0, 120.0)


def dose_phantom(shape, voxel_cm, seed):
    """Water phantom with seeded ellipsoids of lung (1-2, semi-axes 10-25 % of
    the phantom's edge), adipose (2-5, 8-20 %) and cortical bone (3-8, 2-10 %)."""
    edge = min(shape) * voxel_cm
    specs = ((LUNG, (1, 2), (0.10 * edge, 0.25 * edge)), (ADIPOSE, (2, 5), (0.08 * edge, 0.20 * edge)),
             (BONE, (3, 8), (0.02 * edge, 0.10 * edge)))
    return ellipsoid_phantom(shape, seed, specs=specs, voxel_cm=voxel_cm)


def draw_dose_jobs(n, seed, shapes=DOSE_SHAPES, h_lo=1e6, h_hi=1e8, max_edge_cm=35.0):
    """n dose-job descriptions: cubic grid edge from shapes, voxel 0.5-2 mm
    (capped so that the phantom's edge stays within max_edge_cm), source
    monoenergetic 20-120 keV or Kramers with kVp from DOSE_KVP (one each half
    of the time), histories log-uniform in [h_lo, h_hi], a 128 x 128 image."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        m = int(rng.choice(shapes))
        vox = min(float(rng.uniform(0.05, 0.2)), max_edge_cm / m)
        if rng.uniform() < 0.5:
            src = dict(energy=float(rng.uniform(20e3, 120e3)))
        else:
            src = dict(kvp=float(rng.choice(DOSE_KVP)))
        h = int(np.exp(rng.uniform(np.log(h_lo), np.log(h_hi))))
        out.append(dict(shape=(m, m, m), voxel_cm=vox, phantom_seed=int(rng.integers(1 << 30)), histories=h,
                        seed=int(rng.integers(1, 1 << 62)), image=(128, 128), **src))
    return out


def setup_dose_job(sim, job, ids=None):
    """A dose job: air domain 2 cm wider than the phantom on each side, point
    source 20 cm above the phantom's top face, cone beam over that face,
    energy-fluence image just below the phantom and per-voxel deposited
    energy.  ids: the phantom grid if already built."""
    nx, ny, nz = job["shape"]
    v = job["voxel_cm"]
    lx, ly, lz = nx * v, ny * v, nz * v
    ssd = 20.0
    z0 = 0.5
    sim.set_domain((lx + 4.0, ly + 4.0, lz + ssd + 1.5), AIR)
    if ids is None:
        ids = dose_phantom(job["shape"], v, job["phantom_seed"])
    sim.add_grid(ids, (v, v, v), (2.0, 2.0, z0))
    src = (2.0 + lx / 2, 2.0 + ly / 2, z0 + lz + ssd)
    kw = dict(direction=DIR_RECT_ISOTROPIC, position=src, a=(2.0, 2.0, z0 + lz), b=(lx, 0.0, 0.0), c=(0.0, ly, 0.0))
    if "kvp" in job:
        e, w = kramers(job["kvp"])
        kw["spectrum"] = (e, w)
    else:
        kw["energy"] = job["energy"]
    sim.set_source(**kw)
    ny_i, nx_i = job["image"]
    sim.set_tallies([dict(kind=TALLY_IMAGE, p0=(2.0, 2.0, z0 - 0.25), p1=(0.0, ly, 0.0), p2=(lx, 0.0, 0.0),
                          nx=nx_i, ny=ny_i)])
    sim.enable_voxel_dose(True)


ROI_EDGE_CM = 1.0


def dose_job_roi(job):
    """The job's region of interest: four dose voxels of 1 cm^3 (2 x 2
    laterally), centred on the beam axis (the phantom's central column),
    1-2 cm below the entrance face (the phantom's top: the beam travels
    along -z); each ROI dose voxel holds the phantom voxels whose centres
    lie in its cube.  Returns the four boxes as half-open voxel index ranges
    (i0, i1, j0, j1, k0, k1)."""
    nx, ny, nz = job["shape"]
    v = job["voxel_cm"]

    def idx(lo, hi, n):  # voxels with centres in [lo, hi) cm
        a = int(np.ceil(lo / v - 0.5))
        b = int(np.ceil(hi / v - 0.5))
        return max(0, min(a, n)), max(0, min(b, n))

    cx, cy, top = nx * v / 2, ny * v / 2, nz * v
    k0, k1 = idx(top - 2 * ROI_EDGE_CM, top - ROI_EDGE_CM, nz)
    boxes = []
    for a in (0, 1):
        i0, i1 = idx(cx - ROI_EDGE_CM + a * ROI_EDGE_CM, cx + a * ROI_EDGE_CM, nx)
        for b in (0, 1):
            j0, j1 = idx(cy - ROI_EDGE_CM + b * ROI_EDGE_CM, cy + b * ROI_EDGE_CM, ny)
            boxes.append((i0, i1, j0, j1, k0, k1))
    return boxes


def roi_rse(stats):
    """Relative standard error of each ROI dose voxel's summed deposits from
    its (n, sum, sum^2): sigma_sum / sum with sigma_sum = sqrt(n s^2) and s^2
    the sample variance of the deposits (quantity.cpp:37-45, 80-94); inf for
    fewer than two deposits."""
    st = np.atleast_2d(np.asarray(stats, dtype=np.float64))
    n, s, s2 = st[:, 0], st[:, 1], st[:, 2]
    with np.errstate(divide="ignore", invalid="ignore"):
        var = np.maximum(s2 - s * s / n, 0.0) / (n - 1)
        rse = np.sqrt(n * var) / s
    return np.where((n >= 2) & (s > 0), rse, np.inf)


def dose_job_bytes(job, dose="f32"):
    """Device bytes a dose job holds at most: material IDs (1 B per voxel),
    the FP64 voxel records (sum, sum^2, count, 32 B), the macro-cell table
    (2 B per 8^3 voxels), with dose "f32" the staged FP32 (mean, relative
    error) pairs read back at the end (8 B per voxel), the image's (sum,
    sum^2, count) per pixel, plus small records."""
    nx, ny, nz = job["shape"]
    nv = nx * ny * nz
    cells = -(-nx // 8) * -(-ny // 8) * -(-nz // 8)
    ny_i, nx_i = job["image"]
    return nv * (1 + 32 + (8 if dose == "f32" else 0)) + 2 * cells + 24 * nx_i * ny_i + (1 << 16)


# --------------------------------------------------------------------------- CT
CT_MATERIALS = (AIR, WATER, LUNG, ADIPOSE, BONE)


def ct_phantom(shape=(512, 512, 256), voxel_cm=0.08, seed=4):
    """Water block with seeded ellipsoids of lung (2-3, semi-axes 3-7 cm),
    adipose (2-5, 2-6 cm) and cortical bone (3-8, 0.5-3 cm)."""
    specs = ((LUNG, (2, 3), (3.0, 7.0)), (ADIPOSE, (2, 5), (2.0, 6.0)), (BONE, (3, 8), (0.5, 3.0)))
    return ellipsoid_phantom(shape, seed, specs=specs, voxel_cm=voxel_cm)


def ct_geometry(shape=(512, 512, 256), voxel_cm=0.08, sod=60.0, odd=40.0, det_w=110.0, det_h=16.0):
    """Axial CT: rotation axis = the phantom's z axis through its centre;
    source at distance sod from the axis, a flat detector odd beyond it
    (source-detector 100 cm), det_w x det_h cm.  Returns the domain size, the
    phantom origin and the axis point."""
    nx, ny, nz = shape
    lx, ly, lz = nx * voxel_cm, ny * voxel_cm, nz * voxel_cm
    half = max(sod, float(np.hypot(odd, det_w / 2))) + 2.0
    dim = (2 * half, 2 * half, max(lz, det_h) + 4.0)
    axis = (half, half, dim[2] / 2)
    origin = (half - lx / 2, half - ly / 2, axis[2] - lz / 2)
    return dict(dim=dim, origin=origin, axis=axis, sod=sod, odd=odd, det_w=det_w, det_h=det_h)


def setup_ct_view(sim, geo, angle_deg, kvp=120.0, pixels=(512, 64)):
    """Source and detector of one view at angle_deg about the axis: a point
    source, a cone beam aimed over the detector (RectangularIsotropic over
    its rectangle) and an energy-fluence image of pixels = (columns, rows);
    edge1 runs along z (rows), edge2 along the detector's columns, so that
    edge1 x edge2 points from the source to the detector."""
    th = np.radians(angle_deg)
    c, s = float(np.cos(th)), float(np.sin(th))
    ax = geo["axis"]
    src = (ax[0] + geo["sod"] * c, ax[1] + geo["sod"] * s, ax[2])
    dc = (ax[0] - geo["odd"] * c, ax[1] - geo["odd"] * s, ax[2])
    u = (-s, c, 0.0)
    W, H = geo["det_w"], geo["det_h"]
    p0 = (dc[0] - W / 2 * u[0], dc[1] - W / 2 * u[1], dc[2] - H / 2)
    e1 = (0.0, 0.0, H)
    e2 = (W * u[0], W * u[1], 0.0)
    e, w = kramers(kvp)
    sim.set_source(direction=DIR_RECT_ISOTROPIC, position=src, a=p0, b=e2, c=e1, spectrum=(e, w))
    cols, rows = pixels
    sim.set_tallies([dict(kind=TALLY_IMAGE, p0=p0, p1=e1, p2=e2, nx=rows, ny=cols)])


def setup_ct(ctx, n_views=360, shape=(512, 512, 256), voxel_cm=0.08, seed=4, kvp=120.0, pixels=(512, 64),
             ids=None):
    """The CT scan on a Context: the current job holds the phantom (and its
    FP64 voxel dose, accumulated over all views); one job per view aliases
    it and carries its own source and detector image.  Returns (geometry,
    view job ids, view angles)."""
    geo = ct_geometry(shape, voxel_cm)
    owner = ctx.job
    ctx.set_domain(geo["dim"], AIR)
    if ids is None:
        ids = ct_phantom(shape, voxel_cm, seed)
    ctx.add_grid(ids, (voxel_cm,) * 3, geo["origin"])
    ctx.enable_voxel_dose(True)
    ctx.set_tallies([])
    views, angles = [], []
    for v in range(n_views):
        a = 360.0 * v / n_views
        j = ctx.job_create()
        ctx.job_alias_domain(owner)
        setup_ct_view(ctx, geo, a, kvp, pixels)
        views.append(j)
        angles.append(a)
    ctx.job_select(owner)
    return geo, views, angles


def ct_histories(n_views, seed=4, lo=2e6, hi=2e7):
    """Per-view history counts, uniform in [lo, hi] (seeded)."""
    rng = np.random.default_rng(seed)
    return rng.integers(int(lo), int(hi) + 1, n_views)
