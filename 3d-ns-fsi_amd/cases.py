"""Named workloads of the explicit RK4 step, as ns_config structures.

The reference's coordinates start at 0 (src/SmallVectors.h:60-62), so a domain [-a, a]^3 maps to
[0, 2a]^3 with the body centre shifted by +a. The reference's Reynolds number is rho_0 c_0 L / mu_0
(ConfigFile:109-110), so a body Reynolds number Re_D at Mach M and diameter D is Re = Re_D / (M D).
"Slip side walls" are symmetry planes on y and z; x is inlet (min) / outlet (max). CFL 0.5 is
convStabilityLimit; viscStabilityLimit stays at the ConfigFile's -0.68; the periodic density filter is off.
"""
from __future__ import annotations

import numpy as np

import nsfsi

SIDE_WALLS = ["inlet", "outlet", "symmetry", "symmetry", "symmetry", "symmetry"]


def _base(N, L, Re_D, M, D, steps, **kw):
    return nsfsi.make_config(
        NI=N, NJ=N, NK=N, L_x=L, L_y=L, L_z=L,
        Re=Re_D / (M * D), M_0=M, gamma=1.4, Pr=0.72,
        conv_stability_limit=0.5, visc_stability_limit=-0.68,
        stop_criterion="timesteps", stop_time_level=steps,
        filter_intervals=[0], filter_change_times=[],
        conv_history="all", bc=SIDE_WALLS, **kw)


def sphere48(N=48, steps=200, precision="fp64"):
    """48^3 on [-2,2]^3, sphere r=0.5 at the origin, Re_D=100, M=0.3, Gaussian sigma 1e-3 (seed 1)."""
    return _base(N, 4.0, 100.0, 0.3, 1.0, steps, body_type="sphere", body_centre=(2.0, 2.0, 2.0), body_radius=0.5,
                 ic_type="freestream_gaussian", ic_amplitude=1e-3, ic_seed=1, precision=precision)


def cylinder256(N=256, steps=1000, precision="fp64"):
    """256^3 on [-4,4]^3, z-cylinder r=0.5 at the origin spanning z, Re_D=200, M=0.2, U[-1e-3,1e-3] (seed 2)."""
    return _base(N, 8.0, 200.0, 0.2, 1.0, steps, body_type="cylinder", body_axis="z", body_centre=(4.0, 4.0, 0.0),
                 body_radius=0.5, ic_type="freestream_uniform", ic_amplitude=1e-3, ic_seed=2, precision=precision)


def sphere512(N=512, steps=1000, precision="fp32"):
    """512^3 on [-4,4]^3, sphere r=0.6 at (-1,0,0), Re_D=300, M=0.3, Gaussian sigma 1e-3 (seed 3)."""
    return _base(N, 8.0, 300.0, 0.3, 1.2, steps, body_type="sphere", body_centre=(3.0, 4.0, 4.0), body_radius=0.6,
                 ic_type="freestream_gaussian", ic_amplitude=1e-3, ic_seed=3, precision=precision)


def sphere96(N=96, steps=400, precision="fp64"):
    """96^3 on [-2,2]^3, sphere r=0.5 at the origin, Re_D=100, M=0.3, Gaussian sigma 1e-3 (seed 5)."""
    return _base(N, 4.0, 100.0, 0.3, 1.0, steps, body_type="sphere", body_centre=(2.0, 2.0, 2.0), body_radius=0.5,
                 ic_type="freestream_gaussian", ic_amplitude=1e-3, ic_seed=5, precision=precision)


def reference_default():
    """The reference's shipped ConfigFile: 161x83x3 quasi-2-D cylinder channel, periodic z (ConfigFile:22-125)."""
    return nsfsi.make_config()


def sweep64(n=64, steps=500, seed=4, precision="fp32"):
    """Batched sweep: n independent cases; case c draws from a generator seeded with seed + c: grid edge
    N in {32, 48, 64, 96, 128}, sphere radius U[0.3, 0.8] at the origin of [-2,2]^3, Mach U[0.1, 0.4],
    Re_D U[50, 300] (Re = Re_D / (M * 2r)); Gaussian perturbation sigma 1e-3 seeded with seed + c."""
    out = []
    for c in range(n):
        rng = np.random.default_rng(seed + c)
        N = int(rng.choice([32, 48, 64, 96, 128]))
        r = float(rng.uniform(0.3, 0.8))
        M = float(rng.uniform(0.1, 0.4))
        Re_D = float(rng.uniform(50.0, 300.0))
        out.append(_base(N, 4.0, Re_D, M, 2 * r, steps, body_type="sphere", body_centre=(2.0, 2.0, 2.0),
                         body_radius=r, ic_type="freestream_gaussian", ic_amplitude=1e-3, ic_seed=seed + c,
                         precision=precision))
    return out


def two_bodies(steps=10, precision="fp64"):
    """A sphere (r 0.45) upstream of a z-axis cylinder (r 0.4), 48x30x26 nodes, slip side walls, Mach 0.3,
    Re 300, Gaussian perturbation sigma 1e-3: two immersed bodies, classified and applied in list order."""
    return nsfsi.make_config(
        NI=48, NJ=30, NK=26, L_x=4.8, L_y=3.0, L_z=2.6, bc=SIDE_WALLS, body_type="sphere",
        body_centre=(1.2, 1.5, 1.3), body_radius=0.45,
        extra_bodies=[dict(type="cylinder", axis="z", centre=(3.3, 1.4, 0.0), radius=0.4)],
        stop_criterion="timesteps", stop_time_level=steps, filter_intervals=[0], filter_change_times=[],
        ic_type="freestream_gaussian", ic_amplitude=1e-3, ic_seed=23, M_0=0.3, Re=300.0, conv_stability_limit=0.5,
        precision=precision)


def three_bodies(steps=10, precision="fp64"):
    """Two spheres side by side in y (r 0.35) and a z-axis cylinder downstream (r 0.4), 48x40x26 nodes."""
    return nsfsi.make_config(
        NI=48, NJ=40, NK=26, L_x=4.8, L_y=4.0, L_z=2.6, bc=SIDE_WALLS, body_type="sphere",
        body_centre=(1.0, 1.0, 1.3), body_radius=0.35,
        extra_bodies=[dict(type="sphere", centre=(1.0, 3.0, 1.3), radius=0.35),
                      dict(type="cylinder", axis="z", centre=(3.4, 2.0, 0.0), radius=0.4)],
        stop_criterion="timesteps", stop_time_level=steps, filter_intervals=[0], filter_change_times=[],
        ic_type="freestream_gaussian", ic_amplitude=1e-3, ic_seed=29, M_0=0.3, Re=300.0, conv_stability_limit=0.5,
        precision=precision)


WORKLOADS = {
    "sphere48": sphere48,
    "cylinder256": cylinder256,
    "sphere512": sphere512,
    "sphere96": sphere96,
}
