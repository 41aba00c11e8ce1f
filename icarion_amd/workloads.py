"""Descriptor builders and synthetic workloads of the batch ion-step path.

Domain builders evaluate the quantities the reference config loader derives:
number density n = P/(kB T) (EnvironmentConfig.h:150-153), gas mass / radius /
polarizability (EnvironmentConfig.h:229-278) and the E/N -> axial_V conversion
(DomainConfigLoader.cpp:413-440).  Initial ion states are drawn host-side with
numpy from a fixed seed and handed identically to every implementation that is
compared, so their generator does not enter any parity check.
"""
import math

import numpy as np

from . import _abi
from .constants import AMU_TO_KG, BOLTZMANN, ELEM_CHARGE_C, gas_properties
from .ensemble import Ensemble

TD_TO_VM2 = 1e-21  # 1 Td = 1e-21 V m^2


def number_density(pressure_Pa, temperature_K):
    return pressure_Pa / (BOLTZMANN * temperature_K)


def axial_V_from_EN(EN_Td, pressure_Pa, temperature_K, length_m):
    """DomainConfigLoader.cpp:419-427: V = (E/N) * N * L."""
    k_B = 1.380649e-23
    N = pressure_Pa / (k_B * temperature_K)
    E_Vm = EN_Td * 1e-21 * N
    return E_Vm * length_m


def _base_domain(instrument, gas, pressure_Pa, temperature_K):
    d = _abi.DomainDesc()
    d.instrument = instrument
    d.rot_g2l = _abi.IDENTITY9
    d.rot_l2g = _abi.IDENTITY9
    d.field_array_scale = 1.0
    m, r, a = gas_properties(gas)
    d.temperature_K = temperature_K
    d.pressure_Pa = pressure_Pa
    d.gas_density_m3 = number_density(pressure_Pa, temperature_K)
    d.gas_mass_kg = m
    d.gas_radius_m = r
    d.gas_polarizability_m3 = a
    d.detector_z_m = math.nan
    return d


def ims_domain(origin_z=-5e-3, length=20e-3, radius=10e-3, EN_Td=10.0, axial_V=None,
               gas="N2", pressure_Pa=101325.0, temperature_K=300.0, detector_z=math.nan):
    d = _base_domain(_abi.INSTRUMENT_IMS, gas, pressure_Pa, temperature_K)
    d.origin_m = (0.0, 0.0, origin_z)
    d.length_m = length
    d.radius_m = radius
    d.dc_axial_V = (axial_V if axial_V is not None
                    else axial_V_from_EN(EN_Td, pressure_Pa, temperature_K, length))
    d.detector_z_m = detector_z
    return d


def quadrupole_domain(r0=4e-3, V=500.0, f=1e6, U=0.0, length=100e-3, origin_z=-50e-3,
                      axial_V=0.0, gas="He", pressure_Pa=1.0, temperature_K=300.0):
    d = _base_domain(_abi.INSTRUMENT_QUADRUPOLE, gas, pressure_Pa, temperature_K)
    d.origin_m = (0.0, 0.0, origin_z)
    d.length_m = length
    d.radius_m = r0
    d.rf_voltage_V = V
    d.rf_frequency_Hz = f
    d.dc_quad_V = U
    d.dc_axial_V = axial_V
    return d


def orbitrap_domain(R1=6e-3, R2=15e-3, Rm=22e-3, V=3500.0, length=40e-3, gas="He",
                    pressure_Pa=1e-7, temperature_K=300.0):
    """examples/orbitrap/orbitrap_basic.json geometry."""
    d = _base_domain(_abi.INSTRUMENT_ORBITRAP, gas, pressure_Pa, temperature_K)
    d.origin_m = (0.0, 0.0, 0.0)
    d.length_m = length
    d.radius_in_m = R1
    d.radius_out_m = R2
    d.radius_char_m = Rm
    d.dc_radial_V = V
    return d


def orbitrap_k(d):
    """AnalyticalFieldModel.cpp:201-202."""
    r_in, r_out, rc = d.radius_in_m, d.radius_out_m, d.radius_char_m
    return 2.0 * d.dc_radial_V / (rc * rc * math.log(r_out / r_in) - 0.5 * (r_out ** 2 - r_in ** 2))


def physics(integrator=_abi.INTEGRATOR_RK4, collision=_abi.COLLISION_NONE, ou=0,
            precision=_abi.PRECISION_FP64, seed=123, ehss_mode=_abi.EHSS_GEOMETRY, substeps=0):
    """substeps: collision stages per engine step (0 or 1: one)."""
    p = _abi.PhysicsDesc()
    p.integrator = integrator
    p.collision_model = collision
    p.ou_thermal_kicks = ou
    p.precision = precision
    p.ehss_mode = ehss_mode
    p.collision_substeps = substeps
    p.rng_seed = seed
    return p


def species(sigma_hss=0.0, sigma_ehss=0.0, ccs=0.0, K0=0.0, atom_offset=0, atom_count=0,
            features=(0.0, 0.0, 0.0)):
    s = _abi.SpeciesDesc()
    s.sigma_hss_m2 = sigma_hss
    s.sigma_ehss_m2 = sigma_ehss
    s.ccs_m2 = ccs
    s.K0_cm2_Vs = K0
    s.atom_offset = atom_offset
    s.atom_count = atom_count
    for f in range(3):
        s.surrogate_features[f] = float(features[f])
    return s


def gaussian_cloud(n, mass_amu, sigma_pos, center=(0.0, 0.0, 0.0), temperature_K=300.0,
                   charge=1, seed=0, species_index=None):
    """Gaussian positions, Maxwellian velocities (IonLoader.cpp:300-312: sigma = sqrt(kT/m))."""
    rng = np.random.default_rng(seed)
    e = Ensemble(n)
    m = np.broadcast_to(np.asarray(mass_amu, dtype=np.float64) * AMU_TO_KG, (n,)).copy()
    sp = np.broadcast_to(np.asarray(sigma_pos, dtype=np.float64), (3,))
    e.x[:] = center[0] + sp[0] * rng.standard_normal(n)
    e.y[:] = center[1] + sp[1] * rng.standard_normal(n)
    e.z[:] = center[2] + sp[2] * rng.standard_normal(n)
    sv = np.sqrt(BOLTZMANN * temperature_K / m)
    e.vx[:] = sv * rng.standard_normal(n)
    e.vy[:] = sv * rng.standard_normal(n)
    e.vz[:] = sv * rng.standard_normal(n)
    e.mass_kg[:] = m
    e.charge_C[:] = charge * ELEM_CHARGE_C
    if species_index is not None:
        e.species[:] = species_index
    return e


# ---- the five workloads ------------------------------------------------------

def ims_hss_workload(n=10_000, n_steps=100_000, seed=123, precision=_abi.PRECISION_FP64):
    """IMS drift tube, HSS, RK4: 10k ions (m = 100 Da, CCS = 100 A^2, z = +1),
    Gaussian cloud sigma 0.5 mm at the origin, uniform axial E at E/N = 10 Td,
    N2 at 101325 Pa and 300 K, dt = 1e-10 s x 1e5 steps, rng_seed 123.
    Domain: IMS cylinder origin (0, 0, -5 mm), L = 20 mm, R = 10 mm; detector
    plane at z = +0.5 mm."""
    dom = ims_domain(detector_z=0.5e-3)
    phys = physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_HSS, precision=precision, seed=seed)
    sp = [species(sigma_hss=100e-20, sigma_ehss=100e-20, ccs=100e-20)]
    ens = gaussian_cloud(n, 100.0, 0.5e-3, seed=seed)
    return dict(name="ims_hss_rk4", domain=dom, physics=phys, species=sp, atoms=None,
                ensemble=ens, dt=1e-10, n_steps=n_steps)


def ims_mixture_workload(n=4096, seed=31, precision=_abi.PRECISION_FP64,
                         integrator=_abi.INTEGRATOR_RK4):
    """The IMS drift of ims_hss_workload in a gas mixture at 1 atm / 300 K:
    N2 78 %, He 20 %, Ar 2 % and a trace of CO2 that does not take part in
    collisions; two ion species (100 Da / 100 A^2 in N2, 250 Da / 180 A^2) with
    per-gas cross sections.  Returns the workload dict with "mixture":
    (components [(n_i, m_i, participates)], sigma[n_species, n_components])."""
    dom = ims_domain(detector_z=0.5e-3)
    n_tot = dom.gas_density_m3
    comps = [(0.78 * n_tot, 28.0134 * AMU_TO_KG, 1), (0.20 * n_tot, 4.0026 * AMU_TO_KG, 1),
             (0.02 * n_tot, 39.948 * AMU_TO_KG, 1), (4e-4 * n_tot, 44.01 * AMU_TO_KG, 0)]
    sigma = np.array([[100e-20, 62e-20, 105e-20, 120e-20],
                      [180e-20, 120e-20, 190e-20, 210e-20]])
    phys = physics(integrator, _abi.COLLISION_HSS, precision=precision, seed=seed)
    sp = [species(sigma_hss=100e-20, ccs=100e-20), species(sigma_hss=180e-20, ccs=180e-20)]
    rng = np.random.default_rng(seed + 1)
    sidx = rng.integers(0, 2, n).astype(np.uint32)
    ens = gaussian_cloud(n, np.array([100.0, 250.0])[sidx], 0.5e-3, seed=seed)
    ens.species[:] = sidx
    return dict(name="ims_hss_mixture", domain=dom, physics=phys, species=sp, atoms=None,
                ensemble=ens, dt=1e-10, n_steps=2000, mixture=(comps, sigma))


def quadrupole_workload(n=1_000_000, n_steps=100_000, U=0.0, seed=123,
                        precision=_abi.PRECISION_FP64, ou=1):
    """RF quadrupole, Boris, Langevin: 1M ions, masses log-uniform 100-2000 Da,
    z = +1, r0 = 4 mm, V = 500 V at 1 MHz, U = 0-50 V, Langevin damping (+ OU
    thermal kicks) in He at 1 Pa and 300 K, Gaussian positions sigma 0.2 mm,
    Maxwellian velocities at 300 K, dt = 1e-9 s x 1e5 steps.  Domain: cylinder
    of radius r0 from z = -50 mm to +50 mm."""
    rng = np.random.default_rng(seed + 1)
    masses = np.exp(rng.uniform(math.log(100.0), math.log(2000.0), n))
    dom = quadrupole_domain(U=U)
    phys = physics(_abi.INTEGRATOR_BORIS, _abi.COLLISION_LANGEVIN, ou=ou, precision=precision,
                   seed=seed)
    sp = [species()]
    ens = gaussian_cloud(n, masses, 0.2e-3, seed=seed)
    return dict(name="quadrupole_boris_langevin", domain=dom, physics=phys, species=sp,
                atoms=None, ensemble=ens, dt=1e-9, n_steps=n_steps)


def orbitrap_workload(n=1000, n_steps=10_000, seed=7, precision=_abi.PRECISION_FP64):
    """Collisionless Orbitrap (examples/orbitrap/orbitrap_basic.json geometry,
    R1 = 6 mm, R2 = 15 mm, Rm = 22 mm, 3.5 kV): ions start at r = 9 mm, z =
    6 mm with the tangential speed of a circular orbit, RK4 dt = 1e-9 s."""
    rng = np.random.default_rng(seed)
    dom = orbitrap_domain()
    k = orbitrap_k(dom)
    masses = rng.uniform(150.0, 1000.0, n)
    e = Ensemble(n)
    m = masses * AMU_TO_KG
    q = ELEM_CHARGE_C
    r = 9e-3 + 1e-4 * rng.standard_normal(n)
    phi = rng.uniform(0, 2 * math.pi, n)
    e.x[:] = r * np.cos(phi)
    e.y[:] = r * np.sin(phi)
    e.z[:] = 6e-3 + 1e-4 * rng.standard_normal(n)
    # circular orbit: m v^2 / r = q E_r with E_r = k/2 (Rm^2/r - r)
    Er = 0.5 * k * (dom.radius_char_m ** 2 / r - r)
    v = np.sqrt(q * Er * r / m)
    e.vx[:] = -v * np.sin(phi)
    e.vy[:] = v * np.cos(phi)
    e.vz[:] = 0.0
    e.mass_kg[:] = m
    e.charge_C[:] = q
    phys = physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_NONE, precision=precision, seed=seed)
    return dict(name="orbitrap_rk4", domain=dom, physics=phys, species=[species()], atoms=None,
                ensemble=e, dt=1e-9, n_steps=n_steps)


def gaussian_potential_field(n=(256, 256, 256), extent=(10e-3, 10e-3, 50e-3),
                             origin=(-5e-3, -5e-3, 0.0), n_gauss=32, seed=11,
                             dtype=np.float32):
    """Smooth random E field on a node-centred grid: phi = sum_g A_g exp(-|r-c_g|^2/(2 s_g^2)),
    A_g in U[1, 5] kV, s_g in U[1, 5] mm, centres uniform in the box; E = -grad phi
    evaluated analytically at the nodes.  Returns (E[nz, ny, nx, 3], origin, spacing)."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = n
    spacing = tuple(extent[i] / (n[i] - 1) for i in range(3))
    xs = origin[0] + spacing[0] * np.arange(nx)
    ys = origin[1] + spacing[1] * np.arange(ny)
    zs = origin[2] + spacing[2] * np.arange(nz)
    A = rng.uniform(1e3, 5e3, n_gauss) * rng.choice([-1.0, 1.0], n_gauss)
    s = rng.uniform(1e-3, 5e-3, n_gauss)
    c = np.stack([origin[i] + extent[i] * rng.uniform(0, 1, n_gauss) for i in range(3)], 1)
    E = np.zeros((nz, ny, nx, 3), dtype=np.float64)
    X = xs[None, None, :]
    Y = ys[None, :, None]
    for kz in range(0, nz, 16):
        Z = zs[kz:kz + 16, None, None]
        acc = np.zeros((Z.shape[0], ny, nx, 3))
        for g in range(n_gauss):
            dx, dy, dz = X - c[g, 0], Y - c[g, 1], Z - c[g, 2]
            ph = A[g] * np.exp(-(dx * dx + dy * dy + dz * dz) / (2 * s[g] ** 2))
            f = ph / (s[g] ** 2)  # E = -grad phi = phi * (r - c) / s^2
            acc[..., 0] += f * dx
            acc[..., 1] += f * dy
            acc[..., 2] += f * dz
        E[kz:kz + 16] = acc
    return E.astype(dtype), origin, spacing


def field_array_workload(n=4_000_000, n_steps=20_000, grid=(256, 256, 256), seed=11,
                         precision=_abi.PRECISION_FP64, pressure_Pa=300.0):
    """Field-array IMS: 4M ions, FP32 3-component E on a 256^3 grid spanning
    10 x 10 x 50 mm (sum of 32 Gaussians, 1-5 kV), trilinear interpolation, HSS
    in N2 at 300 Pa and 300 K, 3 species (150/400/900 Da; CCS 120/200/320 A^2),
    RK4 dt = 5e-11 s x 2e4 steps.  Domain: cylinder along z, R = 5 mm, L = 50 mm."""
    E, origin, spacing = gaussian_potential_field(grid, seed=seed)
    dom = ims_domain(origin_z=0.0, length=50e-3, radius=5e-3, axial_V=0.0, gas="N2",
                     pressure_Pa=pressure_Pa)
    dom.use_field_array = 1
    phys = physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_HSS, precision=precision, seed=seed)
    sp = [species(sigma_hss=c * 1e-20, ccs=c * 1e-20) for c in (120.0, 200.0, 320.0)]
    rng = np.random.default_rng(seed + 5)
    sidx = rng.integers(0, 3, n).astype(np.uint32)
    masses = np.array([150.0, 400.0, 900.0])[sidx]
    ens = gaussian_cloud(n, masses, (1e-3, 1e-3, 2e-3), center=(0.0, 0.0, 25e-3), seed=seed)
    ens.species[:] = sidx
    return dict(name="field_array_hss_rk4", domain=dom, physics=phys, species=sp, atoms=None,
                ensemble=ens, dt=5e-11, n_steps=n_steps, field=(E, origin, spacing))


def stretched_axis(lo, hi, n, amp=0.12):
    """A non-uniform monotone axis: nodes lo + (hi - lo) s(u), u uniform in [0, 1],
    s(u) = u + amp sin(2 pi u) / (2 pi) (|amp| < 1 keeps s increasing)."""
    u = np.linspace(0.0, 1.0, n)
    return lo + (hi - lo) * (u + amp * np.sin(2 * np.pi * u) / (2 * np.pi))


def gaussian_field_on_axes(xs, ys, zs, n_gauss=8, amp=(0.2, 1.0), seed=0):
    """E = -grad phi at the nodes of a rectilinear grid, phi = sum_g A_g
    exp(-|r - c_g|^2 / (2 s_g^2)) with A_g in U[amp] (random sign), s_g a fifth to
    a half of the smallest extent and centres uniform in the box; returns
    E[nz, ny, nx, 3] (float64)."""
    rng = np.random.default_rng(seed)
    lo = np.array([xs[0], ys[0], zs[0]])
    ext = np.array([xs[-1] - xs[0], ys[-1] - ys[0], zs[-1] - zs[0]])
    A = rng.uniform(*amp, n_gauss) * rng.choice([-1.0, 1.0], n_gauss)
    s = rng.uniform(0.2, 0.5, n_gauss) * ext.min()
    c = lo + ext * rng.uniform(0.0, 1.0, (n_gauss, 3))
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    E = np.zeros(Z.shape + (3,))
    for g in range(n_gauss):
        dx, dy, dz = X - c[g, 0], Y - c[g, 1], Z - c[g, 2]
        f = A[g] * np.exp(-(dx * dx + dy * dy + dz * dz) / (2 * s[g] ** 2)) / s[g] ** 2
        E[..., 0] += f * dx
        E[..., 1] += f * dy
        E[..., 2] += f * dz
    return E


def field_terms_workload(n=4096, grid=(20, 16, 36), seed=3, precision=_abi.PRECISION_FP64,
                         collision=_abi.COLLISION_NONE, integrator=_abi.INTEGRATOR_RK4,
                         uniform=False, n_terms=3, dt=5e-11):
    """Field-array IMS with a composite source (CompositeFieldProvider): unit-voltage
    Gaussian-sum fields on a (by default non-uniform) grid over 10 x 10 x 50 mm,
    scaled by the domain's DC axial voltage (200 V), an RF term (150 V at
    2 MHz, phase 0.3 rad) and a constant (-50), N2 at 300 Pa / 300 K; ions of
    150/400/900 Da near the axis."""
    nx, ny, nz = grid
    if uniform:
        xs, ys, zs = (np.linspace(-5e-3, 5e-3, nx), np.linspace(-5e-3, 5e-3, ny),
                      np.linspace(0.0, 50e-3, nz))
    else:
        xs, ys, zs = (stretched_axis(-5e-3, 5e-3, nx), stretched_axis(-5e-3, 5e-3, ny, -0.2),
                      stretched_axis(0.0, 50e-3, nz, 0.3))
    kinds = [(_abi.FIELD_SCALE_DC_AXIAL, 0.0, 0.0, 0.0), (_abi.FIELD_SCALE_RF, 0.0, 2e6, 0.3),
             (_abi.FIELD_SCALE_CONSTANT, -50.0, 0.0, 0.0)][:n_terms]
    terms = [(gaussian_field_on_axes(xs, ys, zs, amp=(1.0, 5.0), seed=seed * 10 + k),) + kinds[k]
             for k in range(n_terms)]
    dom = ims_domain(origin_z=0.0, length=50e-3, radius=5e-3, axial_V=200.0, gas="N2",
                     pressure_Pa=300.0)
    dom.use_field_array = 1
    dom.rf_voltage_V, dom.rf_frequency_Hz = 150.0, 1e6
    phys = physics(integrator, collision, precision=precision, seed=seed)
    sp = [species(sigma_hss=c * 1e-20, ccs=c * 1e-20) for c in (120.0, 200.0, 320.0)]
    rng = np.random.default_rng(seed + 5)
    sidx = rng.integers(0, 3, n).astype(np.uint32)
    ens = gaussian_cloud(n, np.array([150.0, 400.0, 900.0])[sidx], (1e-3, 1e-3, 2e-3),
                         center=(0.0, 0.0, 25e-3), seed=seed)
    ens.species[:] = sidx
    return dict(name="field_terms", domain=dom, physics=phys, species=sp, atoms=None,
                ensemble=ens, dt=dt, n_steps=1000, axes=(xs, ys, zs), terms=terms)


def multi_domain_workload(n=4096, seed=51, integrator=_abi.INTEGRATOR_RK4, collision=_abi.COLLISION_NONE,
                          dt=1e-9, n_steps=10_000):
    """A three-domain instrument along z: an IMS region (0-5 mm, radius 3 mm,
    20 V axial, N2 at 10 Pa) with a 1.5 mm exit aperture, an RF quadrupole
    (5-15 mm, r0 2 mm, 200 V at 3 MHz, He at 1 Pa) and a second IMS region
    (15-20 mm, radius 3 mm, 10 V, N2 at 10 Pa).  Ions of 100-500 Da start in
    the first region (Gaussian around z = 1 mm) moving along +z at 1.5 km/s on
    top of 300 K thermal velocities, so that they cross both boundaries, and
    many are lost where the quadrupole's radius is smaller than the drift
    region's."""
    d0 = ims_domain(origin_z=0.0, length=5e-3, radius=3e-3, axial_V=20.0, gas="N2", pressure_Pa=10.0)
    d1 = quadrupole_domain(r0=2e-3, V=200.0, f=3e6, length=10e-3, origin_z=5e-3, gas="He", pressure_Pa=1.0)
    d2 = ims_domain(origin_z=15e-3, length=5e-3, radius=3e-3, axial_V=10.0, gas="N2", pressure_Pa=10.0)
    rng = np.random.default_rng(seed + 3)
    masses = rng.uniform(100.0, 500.0, n)
    ens = gaussian_cloud(n, masses, (0.8e-3, 0.8e-3, 0.5e-3), center=(0.0, 0.0, 1e-3), seed=seed)
    ens.vz += 1500.0
    sp = [species(sigma_hss=150e-20, ccs=150e-20, K0=2.0)]
    return dict(name="multi_domain", domains=[d0, d1, d2], apertures=[1.5e-3, 0.0, 0.0],
                physics=physics(integrator, collision, seed=seed), species=sp, atoms=None,
                ensemble=ens, dt=dt, n_steps=n_steps)


def ehss_cluster_counts(n, seed=0, lo=20, hi=500):
    """Atom counts of the EHSS clusters: uniform integers in [lo, hi]."""
    return np.random.default_rng(seed).integers(lo, hi + 1, n).astype(np.uint32)


def ehss_surrogate_workload(n=1_000_000, n_steps=1000, seed=21, pressure_Pa=101325.0,
                            dt=2e-12, EN_Td=20.0, precision=_abi.PRECISION_FP32_KAHAN,
                            ehss_mode=_abi.EHSS_SURROGATE):
    """EHSS drift in He: n ions, each its own random rigid cluster of 20-500
    atoms (radii 1.5-2.0 A, FP32 coordinates), He at 1 atm and 300 K, E/N = 20 Td,
    RK4 dt = 2 ps.  The clusters are generated on the device
    (icarion_gpu_synth_clusters), so the species table is filled by the handle;
    this builder returns the domain, physics, the atom counts and the ensemble
    (species[i] = i, mass = 12 Da per atom)."""
    dom = ims_domain(origin_z=-5e-3, length=20e-3, radius=10e-3, EN_Td=EN_Td, gas="He",
                     pressure_Pa=pressure_Pa, temperature_K=300.0)
    phys = physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_EHSS, precision=precision, seed=seed,
                   ehss_mode=ehss_mode)
    counts = ehss_cluster_counts(n, seed)
    ens = gaussian_cloud(n, counts.astype(np.float64) * 12.0, 0.5e-3, seed=seed)
    ens.species[:] = np.arange(n, dtype=np.uint32)
    return dict(name="ehss_he_20td", domain=dom, physics=phys, counts=counts, ensemble=ens, dt=dt,
                n_steps=n_steps)


def ehss_surrogate_cloud(n, sigma, features, counts, seed=0, integrator=_abi.INTEGRATOR_RK4,
                         dt=2e-12, n_steps=1000):
    """One EHSS-surrogate cloud for the batch: the He 20 Td drift of
    ehss_surrogate_workload with a species table of given clusters (sigma_ehss =
    their OAPA areas, their surrogate features, mass 12 Da per atom); each ion
    takes a random cluster.  sigma, features, counts: per cluster, e.g. from
    gpu.Handle.synth_clusters on a handle of the same domain."""
    dom = ims_domain(origin_z=-5e-3, length=20e-3, radius=10e-3, EN_Td=20.0, gas="He",
                     pressure_Pa=101325.0, temperature_K=300.0)
    phys = physics(integrator, _abi.COLLISION_EHSS, precision=_abi.PRECISION_FP32_KAHAN, seed=seed,
                   ehss_mode=_abi.EHSS_SURROGATE)
    sp = [species(sigma_ehss=float(sigma[k]), features=features[k]) for k in range(len(sigma))]
    rng = np.random.default_rng(seed)
    which = rng.integers(0, len(sigma), n).astype(np.uint32)
    ens = gaussian_cloud(n, np.asarray(counts, np.float64)[which] * 12.0, 0.5e-3, seed=seed)
    ens.species[:] = which
    return dict(name="ehss_surrogate_cloud", domain=dom, physics=phys, species=sp, ensemble=ens,
                dt=dt, n_steps=n_steps)


def batched_workload(n_clouds=256, size_lo=1_000, size_hi=200_000, seed=5,
                     precision=_abi.PRECISION_FP32_KAHAN, steps=None):
    """Batched mixed instruments: n_clouds independent clouds with sizes
    log-uniform over [size_lo, size_hi], the kinds cycling IMS (uniform E, HSS,
    RK4, dt 1e-10 s), RF quadrupole (Boris, Langevin + OU, dt 1e-9 s) and
    Orbitrap (collisionless RK4, dt 1e-9 s).  steps: engine steps per kind
    {"ims": .., "quadrupole": .., "orbitrap": ..}; default 1e5 / 1e5 / 1e6.
    Every cloud gets its own RNG seed.  Returns a list of cloud dicts for
    gpu.Handle.batch_upload."""
    steps = steps or {"ims": 100_000, "quadrupole": 100_000, "orbitrap": 1_000_000}
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(math.log(size_lo), math.log(size_hi), n_clouds)).astype(int)
    kinds = ("ims", "quadrupole", "orbitrap")
    clouds = []
    for c in range(n_clouds):
        kind = kinds[c % 3]
        n = int(sizes[c])
        s = seed * 1000 + c
        if kind == "ims":
            w = ims_hss_workload(n=n, seed=s, precision=precision)
        elif kind == "quadrupole":
            w = quadrupole_workload(n=n, seed=s, precision=precision)
        else:
            w = orbitrap_workload(n=n, seed=s, precision=precision)
        clouds.append(dict(kind=kind, domain=w["domain"], physics=w["physics"],
                           species=w["species"], ensemble=w["ensemble"], t0=0.0, dt=w["dt"],
                           n_steps=int(steps[kind]), step_offset=0))
    return clouds
