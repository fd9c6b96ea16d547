#!/usr/bin/env python3
"""Writes tests/golden/*.npz: input/output vectors of the reference's own
hot-path kernels, run from oracle/_ref/libhvref.so (compiled from the
reference headers by oracle/Makefile, target `ref`), single-threaded so the
OpenMP reductions run in index order.

The fixtures are data only (small arrays); they let the oracle be pinned
against the reference where the reference checkout is absent.

    OMP_NUM_THREADS=1 python tests/golden/make_golden.py        (kernels, loop)
    python tests/golden/make_golden.py vtk                      (snapshot files)
"""
import os
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))

from inputs import SIGMA_R, particles, smooth_field  # noqa: E402
from oracle_lib import Reference  # noqa: E402

N = (12, 10, 14)
D = (5.0, 4.5, 6.0)
NP, NN = 150, 190


def main():
    if os.environ.get("OMP_NUM_THREADS") != "1":
        sys.exit("run with OMP_NUM_THREADS=1 (index-order reductions)")
    R = Reference()
    G = int(np.prod(N))
    st = particles(2024, NP, NN, 3.5, spill=0.02)
    counts = R.deposit(st["x"], st["y"], st["z"], NP, NN, N, D)
    rho = R.smear(counts, N, D, SIGMA_R, 7)
    U = R.skyrme(rho, N, D, -356.8, 303.9, 32.0, 7.0 / 6.0)
    phi_bc = R.coulomb_bc(np.zeros(G), N, D, 50)
    phi = R.sor(phi_bc, rho[:G], N, D)
    Um = R.merge(U, phi, N, D)
    F = R.fdm(Um, N, D)
    f = R.gather(st["x"], st["y"], st["z"], NP, NN, F, N, D)
    k = R.kick([st["kx"], st["ky"], st["kz"]], f, 0.5)
    x = R.drift([st["x"], st["y"], st["z"]], k, 0.5)
    kz_exc = R.excite(k[2], NP, NN, 0.01, 50, 82)
    stats = np.array([R.center_of_mass(x[2], NP, NN, 0), R.center_of_mass(x[2], NP, NN, 1),
                      R.mean_squared_radius(x[0], x[1], x[2], NP, NN, N, D, 0),
                      R.mean_squared_radius(x[0], x[1], x[2], NP, NN, N, D, 1)])
    np.savez_compressed(
        HERE / "kernels_12x10x14.npz",
        n=np.array(N), d_max=np.array(D), np_=NP, nn=NN, sigma_r=SIGMA_R, ppn=7,
        skyrme=np.array([-356.8, 303.9, 32.0, 7.0 / 6.0]), zprot=50, nnuc=82, dt=0.5, eta=0.01,
        **{f"in_{key}": st[key] for key in ("x", "y", "z", "kx", "ky", "kz")},
        counts=counts, rho=rho, U=U, phi_bc=phi_bc, phi=phi, U_merged=Um, Fx=F[0], Fy=F[1], Fz=F[2],
        fx=f[0], fy=f[1], fz=f[2], kx=k[0], ky=k[1], kz=k[2], x=x[0], y=x[1], z=x[2], kz_exc=kz_exc,
        stats=stats)

    # a short trajectory of the reference step loop (cpu_simulate's call sequence)
    st2 = particles(77, 120, 160, 3.2)
    init = {key: st2[key] for key in ("x", "y", "z", "kx", "ky", "kz")}
    steps = 8
    out, stats2, phi2 = R.simulate(init, 120, 160, N, D, -356.8, 303.9, 32.0, 7.0 / 6.0, 0.5 * steps, steps,
                                   1.0, 0.05, 20, 30, SIGMA_R, 7)
    np.savez_compressed(
        HERE / "loop_12x10x14_8steps.npz",
        n=np.array(N), d_max=np.array(D), np_=120, nn=160, sigma_r=SIGMA_R, ppn=7,
        skyrme=np.array([-356.8, 303.9, 32.0, 7.0 / 6.0]), zprot=20, nnuc=30, dt=0.5, steps=steps, t_exc=1.0,
        eta=0.05, **{f"in_{key}": v for key, v in init.items()},
        **{f"out_{key}": out[key] for key in ("x", "y", "z", "kx", "ky", "kz", "fx", "fy", "fz")},
        stats=stats2, phi=phi2)
    print("wrote", sorted(p.name for p in HERE.glob("*.npz")))


def snapshots():
    """The reference's VTK snapshot (output_vtk_header_start,
    output_vector_field, output_scalar_field) of seeded fields on a small
    uneven grid, from the double and the float instantiation."""
    R = Reference()
    n, d = (5, 4, 3), (1.5, 2.0, 2.7)
    G2 = 2 * int(np.prod(n))
    rng = np.random.default_rng(31)
    fields = {key: rng.normal(0.0, 3.0, G2) for key in ("fx", "fy", "fz", "potentials")}
    fields["density"] = rng.uniform(0.0, 0.2, G2)
    np.savez_compressed(HERE / "snapshot_5x4x3.npz", n=np.array(n), d_max=np.array(d), **fields)
    for single, tag in ((False, "f64"), (True, "f32")):
        R.write_vtk(HERE / f"snapshot_5x4x3_{tag}.vtk", n, d, fields["fx"], fields["fy"], fields["fz"],
                    fields["density"], fields["potentials"], single=single)
    print("wrote", sorted(p.name for p in HERE.glob("snapshot_*")))


if __name__ == "__main__":
    if sys.argv[1:] == ["vtk"]:
        snapshots()
    else:
        main()
