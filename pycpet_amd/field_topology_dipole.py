"""Streamline topology of point-dipole frames, the command of the
reference's scripts/field_topology_dipole.py (CPET/source/scripts/
field_topology_dipole.py:10-157) on the GPU streamline kernels:

    python -m pycpet_amd.field_topology_dipole -d <dir> [-c X Y Z] [-x ...]
        [-y ...] [-dim DX DY DZ] [-n N] [-s STEP] [-t THREADS]

Every <dir>/*.dip file (a header line, then index, x, y, z (Å), mx, my, mz
(a.u.) per dipole) is taken once, in random order.  Starts are the
interior points of an n^(1/3)-per-axis grid in the box (rounded up to a
cube), each with an unseeded random step limit below round(2 |dims| / step)
(initialize_box_points_uniform); positions and moments are rotated into the
box frame; the [distance, mean curvature] rows go to <file stem>.top in the
working directory, byte-identical to numpy.savetxt.  The arguments, their
defaults (the y axis defaults to the x axis, as in the reference) and the
7-column check are the reference's.  --threads is accepted and unused: the
streamlines of a frame run in one GPU launch.
"""
from __future__ import annotations

import argparse
import logging
import os
import random
from glob import glob

import numpy as np


def _parser() -> argparse.ArgumentParser:
    ap = argparse.ArgumentParser(
        description="Compute electric field topology from a set of input dipoles")
    ap.add_argument("--dipole_dir", "-d", type=str, required=True,
                    help="directory of .dip files: a header line, then index, x(A), y(A), "
                         "z(A), mx(a.u.), my(a.u.), mz(a.u.) per dipole")
    ap.add_argument("--center", "-c", type=float, default=[0.0, 0.0, 0.0], nargs=3,
                    help="box centre")
    ap.add_argument("-x", type=float, default=[1.0, 0.0, 0.0], nargs=3,
                    help="point defining the box x axis, relative to the centre")
    ap.add_argument("-y", type=float, default=[1.0, 0.0, 0.0], nargs=3,
                    help="point defining the box y axis, relative to the centre")
    ap.add_argument("--dimensions", "-dim", type=np.float32, default=[1.0, 1.0, 1.0], nargs=3,
                    help="box half-widths")
    ap.add_argument("--n-samples", "-n", type=int, default=1000,
                    help="number of streamlines (rounded up to a cube)")
    ap.add_argument("--step-size", "-s", type=float, default=0.1, help="streamline step")
    ap.add_argument("--threads", "-t", type=int, default=1,
                    help="accepted for compatibility; one GPU launch per frame")
    return ap


def frame_topology(dipoles, center, x_axis, y_axis, dimensions, n_samples, step_size,
                   threads=1):
    """One .dip table (D, 7) -> (L, 2) float32 rows, L the cube of points
    per axis that covers n_samples."""
    from .calculator import calculator
    from .field_ops import initialize_box_points_uniform

    max_steps = round(2 * np.linalg.norm(dimensions) / step_size)
    per_axis = round(n_samples ** (1 / 3))
    if per_axis ** 3 < n_samples:
        per_axis += 1
    starts, limits, T = initialize_box_points_uniform(
        center=center, x=x_axis, y=y_axis, dimensions=dimensions,
        N_cr=[per_axis, per_axis, per_axis], dtype="float32", max_steps=max_steps,
        ret_rand_max=True, inclusive=False, seed=None)
    starts = starts.reshape(-1, 3)
    Tinv = np.linalg.inv(T)
    c = calculator.__new__(calculator)
    c.n_samples = len(starts)
    c.step_size = step_size
    c.x = ((dipoles[:, 1:4] - center) @ Tinv).astype(np.float32)
    c.mu = (dipoles[:, 4:7] @ Tinv).astype(np.float32)
    c.dimensions = dimensions
    c.transformation_matrix = T
    c.random_start_points = starts
    c.random_max_samples = limits
    c.concur_slip = threads
    c.center = center
    c.log = logging.getLogger(__name__)
    return c.compute_topo_complete_c_shared_dipole()


def main(argv=None) -> int:
    from .io import write_rows

    args = _parser().parse_args(argv)
    center = np.array(args.center)
    x_axis = np.array(args.x)
    y_axis = np.array(args.y)
    dimensions = np.array(args.dimensions)
    n_samples = args.n_samples
    pending = glob(args.dipole_dir + "/*.dip")
    if not pending:
        return 0  # the reference's loop never runs, and prints nothing
    while pending:
        path = random.choice(pending)
        pending.remove(path)
        dipoles = np.loadtxt(path, skiprows=1, dtype=np.float32)
        if dipoles.shape[1] != 7:
            raise ValueError("Dipole file must have 7 columns: index, x(A), y(A), z(A), "
                             "mx(a.u.), my(a.u.), mz(a.u.)")
        hist = frame_topology(dipoles, center, x_axis, y_axis, dimensions, n_samples,
                              args.step_size, args.threads)
        n_samples = len(hist)  # the cube, as the reference carries it to the next file
        stem = os.path.basename(path).split(".")[0]
        write_rows(f"{stem}.top", hist, "%.18e")
    print("No more files to process!")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
