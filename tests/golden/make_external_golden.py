"""Writes the external-force fixtures of tests/golden/external/ with the
reference binary (run in the build container, after `make -f oracle/ref.mk`):

    python tests/golden/make_external_golden.py [case names]

Each case is a system, an external_forces_file in the reference's format and
the step counter the configuration carries (kept with restart_step_counter
= false, so that the forces' rate terms act).  The reference CPU backend
(DNA2_nomesh, steps = 0) computes the forces and torques of that state with
the external forces (ForceDump observable plugin, %.17g), and the .npz holds
inputs and outputs only.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import make_golden as mg  # noqa: E402
from oxdna_amd import io, systems  # noqa: E402

OUT = os.path.join(HERE, "external")

# every supported type, with particle lists, rates, nearest images and the
# clamp of a moving plane; the duplex sits at the centre of a box of 20
FORCES_A = """# forces on a 20-bp duplex along z around (10, 10, 10)
{
type = string
particle = 0-3
F0 = 0.5
rate = 1e-4
dir = 1, 0.5, 0
}
{
type = string
particle = 39
F0 = 0.3
rate = -1e-5
dir = 10, 10, 30
dir_as_centre = true
}
{
type = trap
particle = 10,12,last
stiff = 1.5
rate = 2e-4
dir = 0, 0, 1
pos0 = 10, 10, 10
}
{
type = mutual_trap
particle = 5
ref_particle = 34
stiff = 2.0
r0 = 1.2
PBC = 1
rate = 1e-5
stiff_rate = 1e-4
}
{
type = constant_trap
particle = 7
ref_particle = 30
stiff = 0.4
r0 = 3
}
{
type = repulsion_plane
particle = -1
stiff = 3
dir = 0, 0, -1
position = 12.5
v = -1e-3
end_position = 12.0
}
{
type = sphere
particle = all
stiff = 2.5
r0 = 2.5
rate = 1e-4
center = 10, 10, 10
}
"""

# a hairpin near a box corner: a mutual trap through the periodic boundary
# with and without the nearest image, a range walked along the strand
FORCES_B = """{
type = mutual_trap
particle = 0
ref_particle = 20
stiff = 1.0
r0 = 2.0
PBC = 0
}
{
type = mutual_trap
particle = 20
ref_particle = 0
stiff = 1.0
r0 = 2.0
PBC = 1
}
{
type = string
particle = 3-9, 15
F0 = 0.05
rate = 0
dir = 0, 1, 0
}
{
type = trap
particle = 12
stiff = 0.8
rate = 0
dir = 1, 0, 0
pos0 = 0.5, 0.5, 0.5
}
"""


# one file per further type, each on the 20-bp duplex (centre near
# (9.6, 9.7, 10), z from 6.1 to 13.9) at step 1000, where the rates act;
# the parameters put a good share of the particles in each force's range
FORCES_TYPES = {
    "lowdim_trap": """{
type = lowdim_trap
particle = 0-9
stiff = 1.2
rate = 1e-4
pos0 = 9.5, 9.5, 10
dir = 0, 1, 1
visibility = 1, 0, 1
}
""",
    "attraction_plane": """{
type = attraction_plane
particle = -1
stiff = 0.7
position = -10.0
dir = 0, 0.2, 1
}
""",
    "repulsion_plane_moving": """{
type = repulsion_plane_moving
particle = 20-39
ref_particle = 3-7
stiff = 0.5
dir = 0, 0, 1
}
""",
    "sphere_smooth": """{
type = sphere_smooth
particle = all
stiff = 1.5
r0 = 1.0
r_ext = 3.0
center = 9.6, 9.7, 10.0
}
""",
    "repulsive_sphere_moving": """{
type = repulsive_sphere_moving
particle = all
stiff = 1.0
r0 = 0.5
rate = 1e-4
origin = 11.9, 9.7, 7
target = 11.9, 9.7, 13
steps = 2000
}
""",
    "repulsive_kepler_poinsot": """{
type = repulsive_kepler_poinsot
particle = all
stiff = 0.003
rate = 1e-4
apex = 3.0
base = 1.0
base_radius = 0.3
kappa = 10
center = 9.6, 9.7, 10.0
}
{
type = repulsive_kepler_poinsot
particle = all
stiff = 3e-5
apex = 3.0
base_radius = 0.35
center = 9.6, 9.7, 8.0
}
""",
    "LJ_wall": """{
type = LJ_wall
particle = all
stiff = 0.5
position = -5.0
sigma = 1.0
n = 6
dir = 0, 0, 1
}
{
type = LJ_wall
particle = 0-19
position = 15.0
only_repulsive = true
dir = 0, 0, -1
}
""",
    "LJ_cone": """{
type = LJ_cone
particle = all
stiff = 0.2
alpha = 0.6
n = 6
dir = 0, 0, 1
pos0 = 9.6, 9.7, 2.0
}
""",
    "twist": """{
type = twist
particle = 5-8
stiff = 1.0
rate = 1e-3
base = 0.3
axis = 0, 0, 1
center = 9.6, 9.7, 10
pos0 = 11, 9.7, 10
mask = 1, 1, 0
}
""",
    "generic_central_force": """{
type = generic_central_force
particle = all
force_type = gravity
F0 = 0.25
inner_cut_off = 1.0
outer_cut_off = 3.5
center = 9.6, 9.7, 11.0
}
""",
    "ellipsoid": """{
type = ellipsoid
particle = all
stiff = 0.6
r_2 = 2.0, 2.0, 3.0
center = 9.6, 9.7, 10.0
}
""",
    "com": """{
type = com
com_list = 0-9
ref_list = 30-39
stiff = 1.0
r0 = 1.5
rate = 1e-4
}
{
type = com
com_list = 12, 14
ref_list = 25-27
stiff = 0.5
r0 = 0.5
}
""",
    "yukawa_sphere": """{
type = yukawa_sphere
particle = all
radius = 7.0
center = 9.6,9.7,10.0
debye_length = 1.0
debye_A = 2.0
WCA_epsilon = 0.5
WCA_sigma = 2.0
WCA_n = 3
}
""",
}


def hairpin_at_corner():
    s = systems.hairpin_set(3, n=4)[1]
    s.pos = s.pos - s.pos.mean(axis=0) + np.array([0.5, 0.5, 0.5])
    return s


def cases():
    d = mg.equilibrate(systems.duplex_20bp(8), "25C", 0.5, 2000)
    yield "ext_duplex_step0", d, "25C", 0.5, FORCES_A, 0
    yield "ext_duplex_step1000", d, "25C", 0.5, FORCES_A, 1000
    yield "ext_hairpin_pbc", hairpin_at_corner(), "40C", 0.5, FORCES_B, 250
    for t, text in FORCES_TYPES.items():
        yield f"ext_{t}_step1000", d, "25C", 0.5, text, 1000


def main(only=None):
    """Writes every case, or those named in `only`."""
    os.makedirs(OUT, exist_ok=True)
    for name, s, T_str, salt, text, step in cases():
        if only and name not in only:
            continue
        E, F, T, stress, hb = mg.run_reference(s, T_str, salt, {"external_forces": "1", "external_forces_file": "ext.txt"},
                                               step=step, files={"ext.txt": text})
        E0, F0, _, _, _ = mg.run_reference(s, T_str, salt, step=step)
        np.savez_compressed(os.path.join(OUT, f"{name}.npz"), pos=s.pos, a1=s.a1, a3=s.a3, btype=s.btype, n3=s.n3, n5=s.n5,
                            strand=s.strand, box=s.box, T=io.temperature(T_str), salt=salt, step=step,
                            forces_file=np.array(text), E8=E, F=F, Tbody=T, F_pairs=F0)
        dF = np.linalg.norm(F - F0, axis=1)
        print(f"{name}: N={s.N} step {step}: external forces on {int((dF > 0).sum())} particles, max {dF.max():.3g}")


if __name__ == "__main__":
    main(sys.argv[1:])
