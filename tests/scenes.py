"""Seeded synthetic scenes shared by tests, the smoke check and the benchmark.

PRNG: SplitMix64 (Steele, Lea & Flood 2014), float = (x >> 40) * 2^-24.
"""
import numpy as np

from fenix_amd import abi
from fenix_amd.emitters import IgniterEmitter, ScalarEmitter, VelocityEmitter, pack_emitter_sources
from fenix_amd.simulation import make_desc

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & MASK64

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next_u64() >> 40) * (1.0 / (1 << 24))

    def uniform_array(self, n, lo=0.0, hi=1.0):
        return np.array([lo + (hi - lo) * self.uniform() for _ in range(n)], dtype=np.float64)


def sim_params(time=0.25, dt=1.0 / 60.0, dx=1.0, world=64.0, wind=(0.0, 0.0, 0.0, 0.0), gravity=(0.0, -1.0, 0.0, 0.45),
               buoyancy=3.6, vort=2.15, mask=(1.0, 0.4, 0.8, 0.35)):
    p = abi.SimParams()
    p.time, p.delta_time, p.previous_time, p.scalar_set_index = time, dt, max(time - dt, 0.0), 0.0
    p.wind[:] = wind
    p.gravity[:] = gravity
    p.buoyancy, p.vorticity_strength, p.dx, p.world_size = buoyancy, vort, dx, world
    p.vorticity_mask[:] = mask
    return p


def c1_desc(n=64, config=False, **kw):
    """Plain method C1: dense n^3, unit voxels, vorticity eps = 0.3 via masks (1,0,0,0)."""
    runtime = {"worldSize": float(max(n, n)), "vorticityStrength": 0.3 / 1.45, "vorticityConstantMask": 1.0,
               "vorticityVelocityMask": 0.0, "vorticityHeatMask": 0.0, "vorticityDensityMask": 0.0}
    quality = {"finePreIterations": 2, "finePostIterations": 2, "midPreIterations": 2, "midPostIterations": 2,
               "coarseIterations": 10}
    quality.update(kw.pop("quality", {}))
    runtime.update(kw.pop("runtime", {}))
    extra = {}
    if config:
        extra = dict(solver=abi.FX_SOLVER_VCYCLE)
    extra.update(kw)
    return make_desc((n, n, n), runtime=runtime, quality=quality, **extra)


def c1_sources():
    """One spherical smoke+fuel source of radius 6 voxels at (32,8,32) of 64^3 (normalised)."""
    return [ScalarEmitter(position=(0.5, 0.125, 0.5), radius=0.09375, startTime=0.0, duration=9999.0,
                          densityRate=1.0, heatRate=2.0, fuelRate=1.5, noiseScale=0.0, noiseMix=0.0, seed=0.0)]


def c1_initial_velocity(n=64, seed=1):
    """u ~ U[-0.1, 0.1]^3 inside the source sphere, 0 elsewhere (row-major (D,H,W) arrays)."""
    rng = SplitMix64(seed)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    pos = np.stack([(x + 0.5) / n, (y + 0.5) / n, (z + 0.5) / n], axis=-1)
    inside = np.linalg.norm(pos - np.array([0.5, 0.125, 0.5]), axis=-1) <= 0.09375
    u = [np.zeros((n, n, n)) for _ in range(3)]
    idx = np.argwhere(inside)
    for (k, j, i) in idx:
        for c in range(3):
            u[c][k, j, i] = -0.1 + 0.2 * rng.uniform()
    return u


def c2_sources(n=256, seed=2, count=2, radius_vox=10.0):
    """C2: emitters of radius 10 voxels on the floor at seeded x,z positions."""
    rng = SplitMix64(seed)
    r = radius_vox / n
    out = []
    for _ in range(count):
        x = r + (1 - 2 * r) * rng.uniform()
        z = r + (1 - 2 * r) * rng.uniform()
        out.append(ScalarEmitter(position=(x, radius_vox / n, z), radius=r, startTime=0.0, duration=9999.0,
                                 densityRate=1.0, heatRate=2.0, fuelRate=1.5, noiseScale=0.0, noiseMix=0.0, seed=0.0))
    return out


def c2_desc(n=256, **kw):
    """C2 config mode: sparse, threshold activation, burn rate 0.5, ignition T = 0.6, 4-level V-cycle."""
    runtime = {"worldSize": float(n), "vorticityStrength": 0.3 / 1.45, "vorticityConstantMask": 1.0,
               "vorticityVelocityMask": 0.0, "vorticityHeatMask": 0.0, "vorticityDensityMask": 0.0}
    args = dict(domain_mode=abi.FX_DOMAIN_SPARSE, activation_rule=abi.FX_ACTIVATION_THRESHOLD,
                solver=abi.FX_SOLVER_VCYCLE, burn_rate=0.5, ignition_lo=0.6 - 0.225, ignition_hi=0.6 + 0.225,
                debug_fields=0)
    runtime.update(kw.pop("runtime", {}))
    args.update(kw)
    return make_desc((n, n, n), runtime=runtime, **args)


def mixed_sources():
    """A scalar, a directional velocity and an igniter source (no hash noise)."""
    return [
        ScalarEmitter(position=(0.45, 0.2, 0.55), radius=0.18, densityRate=1.2, heatRate=1.5, fuelRate=1.0),
        VelocityEmitter(position=(0.5, 0.25, 0.5), radius=0.2, mode="directional", speed=6.0, direction=(0.2, 1.0, 0.1),
                        falloff=0.5, tightness=0.3),
        VelocityEmitter(position=(0.55, 0.3, 0.45), radius=0.15, mode="radial", speed=3.0, falloff=1.0, tightness=0.0),
        IgniterEmitter(position=(0.5, 0.22, 0.5), radius=0.12, intensity=0.8, startTime=0.0, duration=9999.0),
    ]


def random_state(shape, seed):
    """Physically plausible random fields (scalars in [0,1], velocity +-3) for per-pass parity."""
    rng = np.random.default_rng(seed)
    st = {}
    base = rng.random(shape)
    for name in ("density", "temperature", "fuel", "reaction", "adv_density", "adv_temperature", "adv_fuel", "adv_reaction"):
        st[name] = np.clip(base * rng.random(shape) * 1.2, 0, 1)
    for name in ("velocity_x", "velocity_y", "velocity_z", "scratch_x", "scratch_y", "scratch_z"):
        st[name] = rng.uniform(-3, 3, shape)
    for name in ("vorticity_x", "vorticity_y", "vorticity_z"):
        st[name] = rng.uniform(-1, 1, shape)
    st["vorticity_magnitude"] = np.sqrt(st["vorticity_x"] ** 2 + st["vorticity_y"] ** 2 + st["vorticity_z"] ** 2)
    st["divergence"] = rng.uniform(-1, 1, shape)
    # quiet region so activation yields a mix of active and inactive bricks
    d, h, w = shape
    for name in ("density", "temperature", "reaction", "adv_density", "adv_temperature", "adv_reaction"):
        st[name][:, :, : w // 3] = 0.0
    return {k: v.astype(np.float32).astype(np.float64) for k, v in st.items()}


def packed(sources):
    return pack_emitter_sources(sources), len(sources)
